/*
 * repack.c — the VIA-B repacking step: extr, the cascade-key suffix
 * mod-switched to q2, and the binary tree of pair conversions.
 * Test infrastructure (see via_oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "vo_internal.h"

/* extr (conversion/extract.rs:61-75): RLWE (A, B) of degree n to the
 * (n/G, G)-MLWE (pi_0(A), pi_{D-1}(A)*X, ..., pi_1(A)*X; pi_0(B)) with
 * D = n/G, all projections n -> G. */
void vo_extr(const vo_ring *r, const uint64_t *ct, uint32_t G, uint64_t *masks, uint64_t *body) {
    uint32_t n = r->n, nl = r->nl, D = n / G;
    size_t w = vo_pw(r), wg = (size_t)nl * G;
    vo_ring rg = vo_ring_make(G, nl, r->q[0], r->q[1]);
    uint64_t *proj = vo_alloc(wg);
    vo_project_at(nl, n, G, ct, 0, masks);
    for (uint32_t k = 1; k < D; k++) {
        vo_project_at(nl, n, G, ct, D - k, proj);
        vo_mul_x_pow(&rg, proj, 1, masks + (size_t)k * wg);
    }
    vo_project_at(nl, n, G, ct + w, 0, body);
    free(proj);
}

/* bit_reverse_index (conversion/repack.rs:222-230). */
static size_t bit_reverse(size_t i, unsigned bits) {
    size_t r = 0;
    for (unsigned b = 0; b < bits; b++) {
        r = (r << 1) | (i & 1);
        i >>= 1;
    }
    return r;
}

size_t vo_repack_key_words(uint32_t n, uint32_t G, uint32_t L) {
    size_t words = 0;
    for (uint32_t nout = 2 * G; nout <= n; nout <<= 1) words += (size_t)(2 * n / nout) * L * 2 * nout;
    return words;
}

/* repack_keys_poly_2048_t{8,256}_from_rns_cascade_boxed (repack.rs:684-728,
 * 768-805) and repack_keys_*_from_cascade_modswitched (repack.rs:455-474):
 * the cascade steps with output degree 2G..n, every RLev sample
 * mod_switch_sym'ed from q1 to q2 (mod_switch_rlev, repack.rs:43-56). The
 * output holds, per step, [idx][k][part][nout] at q2. */
void vo_repack_key_from_cascade(const vo_ring *r1, const uint64_t *cascade_key, uint64_t q2, uint32_t G, uint32_t L,
                                uint64_t *out) {
    uint32_t n = r1->n, nl = r1->nl;
    size_t src = 0, dst = 0;
    for (uint32_t nout = 2; nout <= n; nout <<= 1) {
        size_t samples = (size_t)(2 * n / nout) * L, sw = 2 * (size_t)nl * nout;
        if (nout >= 2 * G) {
            vo_ring rs = vo_ring_make(nout, nl, r1->q[0], r1->q[1]);
            vo_ring rd = vo_ring_make(nout, 1, q2, 0);
            for (size_t k = 0; k < samples; k++)
                vo_mod_switch_sym(&rs, cascade_key + src + k * sw, &rd, out + dst + k * 2 * nout);
            dst += samples * 2 * nout;
        }
        src += samples * sw;
    }
}

/* mlwes_to_mlwe (repack.rs:146-188): the pair (a, b) of (rank_in, nin)-MLWEs
 * is interleaved at slots 0, 1 (embed_d, repack.rs:72-104); mask idx =
 * rank_out*group + j is key-switched with step key idx and its switched mask
 * (-A') accumulates into output mask j, its switched body (-B') into the
 * interleaved body. The sums run in eval form with one inverse each. */
static void pair_step(const vo_ring *ro, uint32_t rank_in, const uint64_t *a, const uint64_t *b,
                      const uint64_t *key_eval, uint64_t base, uint32_t L, uint64_t *out) {
    uint32_t nout = ro->n, nin = nout / 2, rank_out = rank_in / 2;
    size_t wi = nin, wo = nout, kw = (size_t)L * 2 * wo;
    uint64_t *emb = vo_alloc(wo), *tmp = vo_alloc(wo), *gm = vo_alloc(wo), *gb = vo_alloc(wo);
    uint64_t *acc_b = vo_alloc(wo);
    uint64_t *om = out, *ob = out + (size_t)rank_out * wo;
    memset(om, 0, 8 * rank_out * wo);
    for (uint32_t group = 0; group < 2; group++)
        for (uint32_t j = 0; j < rank_out; j++) {
            uint32_t idx = rank_out * group + j;
            vo_embed_at(1, nin, nout, a + (size_t)idx * wi, 0, emb);
            vo_embed_at(1, nin, nout, b + (size_t)idx * wi, 1, tmp);
            vo_poly_add(ro, emb, emb, tmp);
            vo_gadget_product_eval(ro, key_eval + idx * kw, L, base, emb, gm, gb, 0);
            vo_poly_add(ro, om + j * wo, om + j * wo, gm);
            vo_poly_add(ro, acc_b, acc_b, gb);
        }
    for (uint32_t j = 0; j < rank_out; j++) {
        vo_from_eval(ro, om + j * wo);
        vo_poly_neg(ro, om + j * wo, om + j * wo);
    }
    vo_from_eval(ro, acc_b);
    vo_embed_at(1, nin, nout, a + (size_t)rank_in * wi, 0, emb);
    vo_embed_at(1, nin, nout, b + (size_t)rank_in * wi, 1, tmp);
    vo_poly_add(ro, emb, emb, tmp);
    vo_poly_sub(ro, ob, emb, acc_b);
    free(emb);
    free(tmp);
    free(gm);
    free(gb);
    free(acc_b);
}

/* Repack of repack_engine! (repack.rs:331-369): Extr_G each of the T inputs,
 * reorder them by bit reversal, pad to D = n/G with zero MLWEs, then one
 * pair_convert per level (degree doubles, count halves) with the matching
 * step keys, and unwrap the rank-1 result. key is the output of
 * vo_repack_key_from_cascade (coefficient form at q2). */
int vo_repack(const vo_ring *r2, uint32_t T, uint32_t G, const uint64_t *rotateds, const uint64_t *key,
              uint64_t base, uint32_t L, uint64_t *out) {
    uint32_t n = r2->n;
    if (r2->nl != 1 || !G || n % G || !T || (T & (T - 1)) || T > n / G) return 1;
    uint32_t D = n / G;
    unsigned lt = (unsigned)__builtin_ctz(T);
    /* D MLWEs of rank D and degree G: D*G mask words + G body words each */
    size_t ew = (size_t)D * G + G;
    uint64_t *v = vo_alloc((size_t)D * ew);
    for (uint32_t i = 0; i < T; i++) vo_extr(r2, rotateds + (size_t)i * 2 * n, G, v + i * ew, v + i * ew + (size_t)D * G);
    for (uint32_t i = 0; i < T; i++) {
        size_t j = bit_reverse(i, lt);
        if (i < j) {
            uint64_t *t = vo_alloc(ew);
            memcpy(t, v + i * ew, 8 * ew);
            memcpy(v + i * ew, v + j * ew, 8 * ew);
            memcpy(v + j * ew, t, 8 * ew);
            free(t);
        }
    }
    /* entries T..D-1 stay zero (vo_alloc zeroes) */
    uint32_t rank = D, deg = G;
    size_t koff = 0;
    while (rank > 1) {
        uint32_t nout = 2 * deg, rank_out = rank / 2;
        vo_ring ro = vo_ring_make(nout, 1, r2->q[0], 0);
        size_t samples = (size_t)rank * L;
        uint64_t *kev = vo_alloc(samples * 2 * nout);
        vo_rlev_to_eval(&ro, key + koff, (uint32_t)samples, kev);
        size_t ow = (size_t)rank_out * nout + nout;
        uint64_t *nv = vo_alloc((size_t)(rank / 2) * ow);
        for (uint32_t p = 0; p < rank / 2; p++)
            pair_step(&ro, rank, v + (size_t)(2 * p) * ew, v + (size_t)(2 * p + 1) * ew, kev, base, L, nv + p * ow);
        free(kev);
        free(v);
        v = nv;
        ew = ow;
        koff += samples * 2 * nout;
        rank = rank_out;
        deg = nout;
    }
    /* mlwe_to_rlwe (mlwe_ops.rs:105-109) */
    memcpy(out, v, 8 * 2 * (size_t)n);
    free(v);
    return 0;
}
