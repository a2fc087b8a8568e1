/*
 * scheme.c — encryption layer, gates, switching and the LWE->RLWE cascade.
 * Test infrastructure (see via_oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "vo_internal.h"

/* ---- wide arithmetic (algebra/wide.rs:1-153) ----------------------------- */

static void mul_u128_wide(u128 a, u128 b, u128 *lo, u128 *hi) {
    const u128 M = ((u128)1 << 64) - 1;
    u128 a0 = a & M, a1 = a >> 64, b0 = b & M, b1 = b >> 64;
    u128 p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
    u128 mid = (p00 >> 64) + (p01 & M) + (p10 & M);
    *lo = (p00 & M) | (mid << 64);
    *hi = p11 + (p01 >> 64) + (p10 >> 64) + (mid >> 64);
}

/* floor((hi*2^128 + lo) / d). Long division in 32-bit digits when d < 2^96
 * (every divisor on this path is a modulus below 2^76), otherwise the
 * bit-serial division of wide.rs:43-72. Requires hi < d. */
static u128 div_u256(u128 hi, u128 lo, u128 d) {
    if (hi == 0) return lo / d;
    if ((d >> 96) == 0) {
        uint32_t w[8];
        for (int i = 0; i < 4; i++) {
            w[i] = (uint32_t)(lo >> (32 * i));
            w[4 + i] = (uint32_t)(hi >> (32 * i));
        }
        u128 rem = 0, quo = 0;
        for (int i = 7; i >= 0; i--) {
            u128 cur = (rem << 32) | w[i];
            u128 qd = cur / d;
            rem = cur - qd * d;
            if (i < 4) quo |= qd << (32 * i);
        }
        return quo;
    }
    u128 rem = 0, quo = 0;
    for (int i = 255; i >= 0; i--) {
        u128 bit = i >= 128 ? (hi >> (i - 128)) & 1 : (lo >> i) & 1;
        int of = (int)(rem >> 127);
        rem = (rem << 1) | bit;
        if (of || rem >= d) {
            rem -= d;
            if (i < 128) quo |= (u128)1 << i;
        }
    }
    return quo;
}

/* round_mul_div_u128 (wide.rs:145-153): floor((a*s + floor(d/2)) / d). */
static u128 round_mul_div(u128 a, u128 s, u128 d) {
    u128 lo, hi;
    mul_u128_wide(a, s, &lo, &hi);
    u128 half = d / 2;
    u128 nlo = lo + half;
    if (nlo < lo) hi++;
    return div_u256(hi, nlo, d);
}

void vo_round_mul_div(uint64_t a_lo, uint64_t a_hi, uint64_t s_lo, uint64_t s_hi, uint64_t d_lo,
                      uint64_t d_hi, uint64_t *q_lo, uint64_t *q_hi) {
    u128 a = ((u128)a_hi << 64) | a_lo, s = ((u128)s_hi << 64) | s_lo, d = ((u128)d_hi << 64) | d_lo;
    u128 q = round_mul_div(a, s, d);
    *q_lo = (uint64_t)q;
    *q_hi = (uint64_t)(q >> 64);
}

/* ---- gadget (encryption/gadget.rs) --------------------------------------- */

/* gadget_vector_values (gadget.rs:73-91): g_i = floor((Q + B^(i+1)/2) / B^(i+1)). */
void vo_gadget_vector(const vo_ring *r, uint64_t base, uint32_t L, uint64_t *out) {
    u128 Q = vo_big_q(r), div = 1;
    for (uint32_t i = 0; i < L; i++) {
        div *= base;
        u128 g = (Q + div / 2) / div;
        out[2 * i] = (uint64_t)g;
        out[2 * i + 1] = (uint64_t)(g >> 64);
    }
}

static u128 gadget_g(const vo_ring *r, uint64_t base, uint32_t i) {
    u128 Q = vo_big_q(r), div = 1;
    for (uint32_t k = 0; k <= i; k++) div *= base;
    return (Q + div / 2) / div;
}

/* gadget_scale_into (gadget.rs:115-151) then L rounds of
 * gadget_extract_lsb_into (gadget.rs:168-190): round-half-away-from-zero of
 * c_centred * B^L / Q, then signed digits in (-B/2, B/2], LSB first. */
void vo_gadget_digits_lsb(const vo_ring *r, const uint64_t *in, uint64_t base, uint32_t L, int64_t *digits) {
    u128 Q = vo_big_q(r), bl = 1;
    for (uint32_t k = 0; k < L; k++) bl *= base;
    i128 B = (i128)base, half = B / 2;
    for (size_t i = 0; i < r->n; i++) {
        i128 c = vo_centered(r, in, i);
        u128 mag = c < 0 ? (u128)(-c) : (u128)c;
        u128 v = round_mul_div(mag, bl, Q);
        i128 val = c >= 0 ? (i128)v : -(i128)v;
        for (uint32_t k = 0; k < L; k++) {
            i128 d = val % B;
            if (d < 0) d += B;
            if (d > half) d -= B;
            digits[(size_t)k * r->n + i] = (int64_t)d;
            val = (val - d) / B;
        }
    }
}

/* gadget_decompose_into (gadget.rs:199-217): MSB-first output. */
void vo_gadget_decompose(const vo_ring *r, const uint64_t *in, uint64_t base, uint32_t L, int64_t *digits) {
    int64_t *lsb = (int64_t *)malloc(sizeof(int64_t) * L * r->n);
    vo_gadget_digits_lsb(r, in, base, L, lsb);
    for (uint32_t k = 0; k < L; k++) memcpy(digits + (size_t)(L - 1 - k) * r->n, lsb + (size_t)k * r->n, sizeof(int64_t) * r->n);
    free(lsb);
}

void vo_rlev_to_eval(const vo_ring *r, const uint64_t *rlev, uint32_t L, uint64_t *out) {
    size_t w = vo_pw(r);
    memcpy(out, rlev, sizeof(uint64_t) * 2 * L * w);
    for (size_t k = 0; k < 2 * (size_t)L; k++) vo_to_eval(r, out + k * w);
}

/* RLevEval::gadget_product (rlev.rs:259-283): sum over k of
 * NTT(digit_k) (.) sample[L-1-k], accumulated in eval form. */
void vo_gadget_product_eval(const vo_ring *r, const uint64_t *rlev_eval, uint32_t L, uint64_t base,
                            const uint64_t *in, uint64_t *acc_mask, uint64_t *acc_body, int acc) {
    size_t w = vo_pw(r);
    int64_t *dg = (int64_t *)malloc(sizeof(int64_t) * L * r->n);
    uint64_t *dp = vo_alloc(w);
    vo_gadget_digits_lsb(r, in, base, L, dg);
    for (uint32_t k = 0; k < L; k++) {
        vo_poly_from_centered(r, dg + (size_t)k * r->n, dp);
        vo_to_eval(r, dp);
        const uint64_t *s = rlev_eval + (size_t)(L - 1 - k) * 2 * w;
        int a = acc || k > 0;
        vo_eval_mul_acc(r, acc_mask, dp, s, a);
        vo_eval_mul_acc(r, acc_body, dp, s + w, a);
    }
    free(dp);
    free(dg);
}

/* RLevCiphertext::gadget_product (rlev.rs:193-222). */
void vo_gadget_product(const vo_ring *r, const uint64_t *rlev, uint32_t L, uint64_t base, const uint64_t *in,
                       uint64_t *ct) {
    size_t w = vo_pw(r);
    uint64_t *ev = vo_alloc(2 * L * w);
    vo_rlev_to_eval(r, rlev, L, ev);
    vo_gadget_product_eval(r, ev, L, base, in, ct, ct + w, 0);
    vo_from_eval(r, ct);
    vo_from_eval(r, ct + w);
    free(ev);
}

/* RGSWCiphertext::external_product (rgsw.rs:108-120). */
void vo_external_product(const vo_ring *r, const uint64_t *rgsw, uint32_t L1, uint32_t L2, uint64_t b1,
                         uint64_t b2, const uint64_t *ct, uint64_t *out) {
    size_t w = vo_pw(r);
    uint64_t *t1 = vo_alloc(2 * w), *t2 = vo_alloc(2 * w);
    vo_gadget_product(r, rgsw, L1, b1, ct, t1);
    vo_gadget_product(r, rgsw + 2 * (size_t)L1 * w, L2, b2, ct + w, t2);
    vo_poly_add(r, out, t1, t2);
    vo_poly_add(r, out + w, t1 + w, t2 + w);
    free(t1);
    free(t2);
}

/* RLevCiphertext::key_switch (keyswitch.rs:85-102): (-gp.mask, body - gp.body). */
void vo_key_switch(const vo_ring *r, const uint64_t *rlev, uint32_t L, uint64_t base, const uint64_t *ct,
                   uint64_t *out) {
    size_t w = vo_pw(r);
    uint64_t *gp = vo_alloc(2 * w);
    vo_gadget_product(r, rlev, L, base, ct, gp);
    vo_poly_neg(r, out, gp);
    vo_poly_sub(r, out + w, ct + w, gp + w);
    free(gp);
}

/* ---- RLWE (encryption/rlwe.rs) ------------------------------------------- */

/* SecretKey::keygen (rlwe.rs:40-46). */
void vo_keygen(const vo_ring *r, vo_dist d, vo_prg *g, uint64_t *sk) {
    int64_t *s = (int64_t *)malloc(sizeof(int64_t) * r->n);
    vo_sample(g, d, s, r->n);
    vo_poly_from_centered(r, s, sk);
    free(s);
}

/* SecretKey::encrypt (rlwe.rs:87-106): mask uniform, then error, then
 * body = mask*s + e + m. */
void vo_encrypt(const vo_ring *r, const uint64_t *sk, const uint64_t *m, vo_dist d, vo_prg *g, uint64_t *ct) {
    size_t w = vo_pw(r);
    int64_t *es = (int64_t *)malloc(sizeof(int64_t) * r->n);
    uint64_t *e = vo_alloc(w);
    vo_poly_random(r, g, ct);
    vo_sample(g, d, es, r->n);
    vo_poly_from_centered(r, es, e);
    vo_poly_mul(r, ct + w, ct, sk);
    vo_poly_add(r, ct + w, ct + w, e);
    vo_poly_add(r, ct + w, ct + w, m);
    free(e);
    free(es);
}

/* encode (rlwe.rs:247-270): (pt * ceil(Q/p)) mod Q. */
void vo_encode(const vo_ring *r, uint64_t p, const uint64_t *pt, uint64_t *out) {
    u128 Q = vo_big_q(r), delta = (Q + p - 1) / p;
    for (size_t i = 0; i < r->n; i++) vo_set_u128(r, out, i, ((u128)pt[i] * delta) % Q);
}

/* decode (rlwe.rs:296-330): div_euclid(p*c + floor(Q/2), Q) rem_euclid p. */
void vo_decode(const vo_ring *r, uint64_t p, const uint64_t *in, uint64_t *pt) {
    i128 Q = (i128)vo_big_q(r), P = (i128)p, hq = Q / 2;
    for (size_t i = 0; i < r->n; i++) {
        i128 c = vo_centered(r, in, i);
        i128 num = P * c + hq;
        i128 q = num / Q;
        if (num % Q < 0) q -= 1;
        i128 m = q % P;
        if (m < 0) m += P;
        pt[i] = (uint64_t)m;
    }
}

/* SecretKey::encrypt_rlev (rlev.rs:40-62): sample k encrypts m * g_k. */
void vo_encrypt_rlev(const vo_ring *r, const uint64_t *sk, const uint64_t *m, uint64_t base, uint32_t L, vo_dist d,
                     vo_prg *g, uint64_t *rlev) {
    size_t w = vo_pw(r);
    uint64_t *sc = vo_alloc(w);
    for (uint32_t k = 0; k < L; k++) {
        vo_poly_scale(r, sc, m, gadget_g(r, base, k));
        vo_encrypt(r, sk, sc, d, g, rlev + (size_t)k * 2 * w);
    }
    free(sc);
}

/* SecretKey::encrypt_rgsw (rgsw.rs:47-62): RLev(-s*m, B1) then RLev(m, B2). */
void vo_encrypt_rgsw(const vo_ring *r, const uint64_t *sk, const uint64_t *m, uint64_t b1, uint64_t b2,
                     uint32_t L1, uint32_t L2, vo_dist d, vo_prg *g, uint64_t *rgsw) {
    size_t w = vo_pw(r);
    uint64_t *nsm = vo_alloc(w);
    vo_poly_mul(r, nsm, sk, m);
    vo_poly_neg(r, nsm, nsm);
    vo_encrypt_rlev(r, sk, nsm, b1, L1, d, g, rgsw);
    vo_encrypt_rlev(r, sk, m, b2, L2, d, g, rgsw + 2 * (size_t)L1 * w);
    free(nsm);
}

/* ---- gates (gates/mux.rs, gates/rotate.rs, gates/convert.rs) -------------- */

/* rotate (rotate.rs:52-57). */
void vo_rotate(const vo_ring *r, const uint64_t *ct, size_t k, uint64_t *out) {
    size_t w = vo_pw(r);
    vo_mul_x_pow(r, ct, k, out);
    vo_mul_x_pow(r, ct + w, k, out + w);
}

/* cmux (mux.rs:75-84): ct0 + RGSW (x) (ct1 - ct0). */
void vo_cmux(const vo_ring *r, const uint64_t *rgsw, uint32_t L, uint64_t base, const uint64_t *ct0,
             const uint64_t *ct1, uint64_t *out) {
    size_t w = vo_pw(r);
    uint64_t *diff = vo_alloc(2 * w), *ep = vo_alloc(2 * w);
    vo_poly_sub(r, diff, ct1, ct0);
    vo_poly_sub(r, diff + w, ct1 + w, ct0 + w);
    vo_external_product(r, rgsw, L, L, base, base, diff, ep);
    vo_poly_add(r, out, ct0, ep);
    vo_poly_add(r, out + w, ct0 + w, ep + w);
    free(diff);
    free(ep);
}

/* dmux (mux.rs:129-137): (ct - RGSW (x) ct, RGSW (x) ct). */
void vo_dmux(const vo_ring *r, const uint64_t *rgsw, uint32_t L, uint64_t base, const uint64_t *ct, uint64_t *out0,
             uint64_t *out1) {
    size_t w = vo_pw(r);
    uint64_t *p = vo_alloc(2 * w);
    vo_external_product(r, rgsw, L, L, base, base, ct, p);
    vo_poly_sub(r, out0, ct, p);
    vo_poly_sub(r, out0 + w, ct + w, p + w);
    memcpy(out1, p, sizeof(uint64_t) * 2 * w);
    free(p);
}

/* cmux_tree (mux.rs:167-198): LSB first, pairs (2j, 2j+1) -> j in place. */
void vo_cmux_tree(const vo_ring *r, const uint64_t *rgsws, uint32_t nbits, uint32_t L, uint64_t base,
                  uint64_t *inputs, uint64_t *out) {
    size_t w = vo_pw(r), cw = 2 * w, gw = 4 * (size_t)L * w;
    size_t active = (size_t)1 << nbits;
    uint64_t *t = vo_alloc(cw);
    for (uint32_t b = 0; b < nbits; b++) {
        size_t half = active >> 1;
        for (size_t j = 0; j < half; j++) {
            vo_cmux(r, rgsws + b * gw, L, base, inputs + 2 * j * cw, inputs + (2 * j + 1) * cw, t);
            memcpy(inputs + j * cw, t, cw * 8);
        }
        active = half;
    }
    memcpy(out, inputs, cw * 8);
    free(t);
}

/* dmux_tree (mux.rs:222-251): MSB first, j descending, out[2j] = ct - p,
 * out[2j+1] = p. */
void vo_dmux_tree(const vo_ring *r, const uint64_t *rgsws, uint32_t nbits, uint32_t L, uint64_t base,
                  const uint64_t *input, uint64_t *out) {
    size_t w = vo_pw(r), cw = 2 * w, gw = 4 * (size_t)L * w;
    uint64_t *src = vo_alloc(cw), *o0 = vo_alloc(cw), *o1 = vo_alloc(cw);
    memcpy(out, input, cw * 8);
    for (uint32_t i = 0; i < nbits; i++) {
        size_t active = (size_t)1 << i;
        for (size_t jj = active; jj-- > 0;) {
            memcpy(src, out + jj * cw, cw * 8);
            vo_dmux(r, rgsws + i * gw, L, base, src, o0, o1);
            memcpy(out + (2 * jj + 1) * cw, o1, cw * 8);
            memcpy(out + 2 * jj * cw, o0, cw * 8);
        }
    }
    free(src);
    free(o0);
    free(o1);
}

/* crot (rotate.rs:128-149). Forward: cmux(result, X^(2^i) result);
 * SlotExtract: cmux(result, -(X^(N-2^i) result)). */
void vo_crot(const vo_ring *r, int slot_extract, const uint64_t *rgsws, uint32_t nbits, uint32_t L, uint64_t base,
             const uint64_t *ct, uint64_t *out) {
    size_t w = vo_pw(r), cw = 2 * w, gw = 4 * (size_t)L * w;
    uint64_t *res = vo_alloc(cw), *rot = vo_alloc(cw), *t = vo_alloc(cw);
    memcpy(res, ct, cw * 8);
    for (uint32_t i = 0; i < nbits; i++) {
        if (slot_extract) {
            vo_rotate(r, res, r->n - ((size_t)1 << i), rot);
            vo_poly_neg(r, rot, rot);
            vo_poly_neg(r, rot + w, rot + w);
        } else {
            vo_rotate(r, res, (size_t)1 << i, rot);
        }
        vo_cmux(r, rgsws + i * gw, L, base, res, rot, t);
        memcpy(res, t, cw * 8);
    }
    memcpy(out, res, cw * 8);
    free(res);
    free(rot);
    free(t);
}

/* rlwe_to_rgsw (convert.rs:86-108): neg_s_m[i] = (B_i + (ck (x) A_i).mask,
 * (ck (x) A_i).body); m = the levels as an RLev. */
void vo_rlwe_to_rgsw(const vo_ring *r, const uint64_t *levels, const uint64_t *conv_key, const uint64_t *m_rlev,
                     uint64_t base_ck, uint32_t L_OUT, uint32_t L_CK, uint64_t *rgsw) {
    size_t w = vo_pw(r), cw = 2 * w;
    uint64_t *ev = vo_alloc(2 * L_CK * w), *gp = vo_alloc(cw);
    vo_rlev_to_eval(r, conv_key, L_CK, ev);
    for (uint32_t i = 0; i < L_OUT; i++) {
        vo_gadget_product_eval(r, ev, L_CK, base_ck, levels + i * cw, gp, gp + w, 0);
        vo_from_eval(r, gp);
        vo_from_eval(r, gp + w);
        vo_poly_add(r, rgsw + i * cw, levels + i * cw + w, gp);
        memcpy(rgsw + i * cw + w, gp + w, w * 8);
    }
    memcpy(rgsw + L_OUT * cw, m_rlev, sizeof(uint64_t) * L_OUT * cw);
    free(ev);
    free(gp);
}

/* ---- switching (switching/) ------------------------------------------ */

/* mod_switch_sym (mod_switch.rs:57-86) with RescaleConsts::scale
 * (kernels/mod_switch.rs:59-61): floor((c*q' + floor(q/2)) / q) on the
 * canonical value, then reduced into the destination ring. */
void vo_mod_switch_sym(const vo_ring *src, const uint64_t *ct, const vo_ring *dst, uint64_t *out) {
    size_t ws = vo_pw(src), wd = vo_pw(dst);
    u128 qs = vo_big_q(src), qd = vo_big_q(dst);
    for (int part = 0; part < 2; part++)
        for (size_t i = 0; i < src->n; i++) {
            const uint64_t *p = ct + part * ws;
            u128 c = vo_reconstruct(src, p[i], src->nl == 2 ? p[src->n + i] : 0);
            vo_set_u128(dst, out + part * wd, i, round_mul_div(c, qd, qs));
        }
}

/* mod_switch_asym (mod_switch.rs:124-154): mask to q_mask, body to q_body. */
void vo_mod_switch_asym(const vo_ring *src, const uint64_t *ct, uint64_t q_mask, uint64_t q_body, uint64_t *om,
                        uint64_t *ob) {
    size_t ws = vo_pw(src);
    u128 qs = vo_big_q(src);
    for (size_t i = 0; i < src->n; i++) {
        u128 c = vo_reconstruct(src, ct[i], src->nl == 2 ? ct[src->n + i] : 0);
        om[i] = vo_mod_u128(round_mul_div(c, q_mask, qs), q_mask);
        u128 b = vo_reconstruct(src, ct[ws + i], src->nl == 2 ? ct[ws + src->n + i] : 0);
        ob[i] = vo_mod_u128(round_mul_div(b, q_body, qs), q_body);
    }
}

/* mod_switch_rgsw (convert.rs:223-241). */
void vo_mod_switch_rgsw(const vo_ring *src, const uint64_t *rgsw, uint32_t L1, uint32_t L2, const vo_ring *dst,
                        uint64_t *out) {
    size_t ws = vo_pw(src), wd = vo_pw(dst);
    for (uint32_t k = 0; k < L1 + L2; k++) vo_mod_switch_sym(src, rgsw + k * 2 * ws, dst, out + k * 2 * wd);
}

/* rekey_secret_key (switching/rekey.rs): centred coefficients re-lifted. */
void vo_rekey(const vo_ring *src, const uint64_t *sk, const vo_ring *dst, uint64_t *out) {
    for (size_t i = 0; i < src->n; i++) vo_set_i128(dst, out, i, vo_centered(src, sk, i));
}

/* gen_rsk (ring_switch.rs:224-246): sample j encrypts under S2 the slot-0
 * projection of S1 (j = 0) or of -(S1 * X^(N1-j)). */
void vo_gen_rsk(const vo_ring *r1, const uint64_t *s1, const vo_ring *r2, const uint64_t *s2, uint64_t base,
                uint32_t L, vo_dist d, vo_prg *g, uint64_t *rsk) {
    uint32_t D = r1->n / r2->n;
    size_t w1 = vo_pw(r1), w2 = vo_pw(r2);
    uint64_t *rot = vo_alloc(w1), *proj = vo_alloc(w2);
    for (uint32_t j = 0; j < D; j++) {
        if (j == 0) {
            vo_project_at(r1->nl, r1->n, r2->n, s1, 0, proj);
        } else {
            vo_mul_x_pow(r1, s1, r1->n - j, rot);
            vo_poly_neg(r1, rot, rot);
            vo_project_at(r1->nl, r1->n, r2->n, rot, 0, proj);
        }
        vo_encrypt_rlev(r2, s2, proj, base, L, d, g, rsk + (size_t)j * L * 2 * w2);
    }
    free(rot);
    free(proj);
}

/* ring_switch (ring_switch.rs:283-312). */
void vo_ring_switch(const vo_ring *r1, const uint64_t *ct, const vo_ring *r2, const uint64_t *rsk, uint64_t base,
                    uint32_t L, uint64_t *out) {
    uint32_t D = r1->n / r2->n;
    size_t w1 = vo_pw(r1), w2 = vo_pw(r2);
    uint64_t *rot = vo_alloc(w1), *proj = vo_alloc(w2), *gp = vo_alloc(2 * w2);
    memset(out, 0, 8 * w2);
    vo_project_at(r1->nl, r1->n, r2->n, ct + w1, 0, out + w2);
    for (uint32_t j = 0; j < D; j++) {
        vo_mul_x_pow(r1, ct, j, rot);
        vo_project_at(r1->nl, r1->n, r2->n, rot, 0, proj);
        vo_gadget_product(r2, rsk + (size_t)j * L * 2 * w2, L, base, proj, gp);
        vo_poly_sub(r2, out, out, gp);
        vo_poly_sub(r2, out + w2, out + w2, gp + w2);
    }
    free(rot);
    free(proj);
    free(gp);
}

/* SecretKey::decrypt_asymmetric (rlwe.rs:187-224). */
void vo_decrypt_asym(const vo_ring *r3, const uint64_t *sk2, const uint64_t *mask, const uint64_t *body, uint64_t q4,
                     uint64_t p, uint64_t *pt) {
    size_t n = r3->n;
    uint64_t q3 = r3->q[0];
    uint64_t *prod = vo_alloc(n), *noisy = vo_alloc(n);
    vo_poly_mul(r3, prod, mask, sk2);
    vo_ring r4 = vo_ring_make((uint32_t)n, 1, q4, 0);
    for (size_t i = 0; i < n; i++) {
        u128 s = ((u128)prod[i] * q4 + q3 / 2) / q3;
        uint64_t sw = vo_mod_u128(s, q4);
        noisy[i] = body[i] >= sw ? body[i] - sw : body[i] + q4 - sw;
    }
    vo_decode(&r4, p, noisy, pt);
    free(prod);
    free(noisy);
}

/* ---- LWE and the cascade (conversion/) -------------------------------- */

static void lwe_encrypt_common(const vo_ring *r, const uint64_t *sk, u128 m_val, vo_dist d, vo_prg *g,
                               uint64_t *lwe) {
    size_t n = r->n;
    uint64_t *masks = vo_alloc(vo_pw(r));
    int64_t e;
    vo_poly_random(r, g, masks);
    vo_sample(g, d, &e, 1);
    for (uint32_t l = 0; l < r->nl; l++) {
        uint64_t q = r->q[l];
        /* dot_residues (kernels/lwe.rs:44-61) */
        u128 acc = 0;
        for (size_t i = 0; i < n; i++) acc += (u128)masks[l * n + i] * sk[l * n + i];
        uint64_t dot = vo_mod_u128(acc, q);
        uint64_t ev = vo_mod_i64(e, q), mv = vo_mod_u128(m_val, q);
        uint64_t b = vo_mod_u128((u128)dot + ev + mv, q);
        memcpy(lwe + l * (n + 1), masks + l * n, 8 * n);
        lwe[l * (n + 1) + n] = b;
    }
    free(masks);
}

/* encrypt_lwe_raw (mlwe_ops.rs:223-248). */
void vo_encrypt_lwe_raw(const vo_ring *r, const uint64_t *sk, uint64_t lo, uint64_t hi, vo_dist d, vo_prg *g,
                        uint64_t *lwe) {
    u128 m = ((u128)hi << 64) | lo;
    lwe_encrypt_common(r, sk, m % vo_big_q(r), d, g, lwe);
}

/* encrypt_lwe (mlwe_ops.rs:180-206): message scaled by ceil(Q/p). */
void vo_encrypt_lwe(const vo_ring *r, const uint64_t *sk, uint64_t msg, uint64_t p, vo_dist d, vo_prg *g,
                    uint64_t *lwe) {
    u128 Q = vo_big_q(r), delta = (Q + p - 1) / p;
    lwe_encrypt_common(r, sk, (delta * msg) % Q, d, g, lwe);
}

size_t vo_cascade_key_words(const vo_ring *r, uint32_t L) {
    unsigned logn = (unsigned)__builtin_ctz(r->n);
    return (size_t)logn * 2 * r->n * L * 2 * r->nl;
}

/* gen_lwe_to_rlwe_key_* (cascade.rs:90-110) via gen_conv_step_key_element
 * (conv.rs:161-190): steps in ascending output degree; key idx = m*group + j
 * encrypts, under S projected at slot j (degree N_OUT), the slot-(m*group+j)
 * projection of S (degree N_IN) embedded at slot 0. */
void vo_gen_cascade_key(const vo_ring *r, const uint64_t *sk, uint64_t base, uint32_t L, vo_dist d, vo_prg *g,
                        uint64_t *key) {
    uint32_t n = r->n, nl = r->nl;
    size_t off = 0;
    for (uint32_t nout = 2; nout <= n; nout <<= 1) {
        uint32_t nin = nout / 2, rank_in = n / nin, m = rank_in / 2;
        vo_ring ro = vo_ring_make(nout, nl, r->q[0], r->q[1]);
        uint64_t *src = vo_alloc(nl * nin), *srce = vo_alloc(nl * nout), *tgt = vo_alloc(nl * nout);
        for (uint32_t idx = 0; idx < rank_in; idx++) {
            uint32_t group = idx / m, j = idx % m;
            vo_project_at(nl, n, nin, sk, m * group + j, src);
            vo_embed_at(nl, nin, nout, src, 0, srce);
            vo_project_at(nl, n, nout, sk, j, tgt);
            vo_encrypt_rlev(&ro, tgt, srce, base, L, d, g, key + off);
            off += (size_t)L * 2 * nl * nout;
        }
        free(src);
        free(srce);
        free(tgt);
    }
}

