/*
 * ring.c — Z_q / RNS arithmetic, the negacyclic NTT and polynomial kernels.
 * Test infrastructure (see via_oracle.h).
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "vo_internal.h"

/* ---- scalar arithmetic (algebra/zq/modulus.rs:45-211) ------------------- */

/* Barrett reduction of any u128 (zq/reduce.rs:59-141); the result is the
 * canonical residue, so a plain % gives the same value. The reduce is written
 * out for speed only. */
uint64_t vo_mod_u128(u128 x, uint64_t q) {
    if ((q & (q - 1)) == 0) return (uint64_t)x & (q - 1);
    return (uint64_t)(x % q);
}

uint64_t vo_mod_i64(int64_t x, uint64_t q) {
    /* Modulus::reduce_i64: reduce |x| then negate (modulus.rs:132-137). */
    uint64_t mag = x < 0 ? (uint64_t)(-(x + 1)) + 1 : (uint64_t)x;
    uint64_t r = vo_mod_u128(mag, q);
    if (x < 0 && r != 0) r = q - r;
    return r;
}

uint64_t vo_pow_mod(uint64_t b, uint64_t e, uint64_t q) {
    u128 r = 1 % q, bb = b % q;
    while (e) {
        if (e & 1) r = (r * bb) % q;
        bb = (bb * bb) % q;
        e >>= 1;
    }
    return (uint64_t)r;
}

/* mod_inverse_u64 (rns/reduce.rs) via extended Euclid. */
uint64_t vo_inv_mod(uint64_t a, uint64_t q) {
    int64_t t = 0, nt = 1;
    int64_t r = (int64_t)q, nr = (int64_t)(a % q);
    while (nr) {
        int64_t qq = r / nr, tmp;
        tmp = t - qq * nt; t = nt; nt = tmp;
        tmp = r - qq * nr; r = nr; nr = tmp;
    }
    if (r != 1) return 0;
    if (t < 0) t += (int64_t)q;
    return (uint64_t)t;
}

u128 vo_big_q(const vo_ring *r) {
    return r->nl == 1 ? (u128)r->q[0] : (u128)r->q[0] * r->q[1];
}

/* RnsBasis::reconstruct (rns/basis.rs:101-109), Garner:
 * x = a0 + q0 * ((a1 - a0) * q0^-1 mod q1). */
u128 vo_reconstruct(const vo_ring *r, uint64_t a0, uint64_t a1) {
    if (r->nl == 1) return a0;
    uint64_t q0 = r->q[0], q1 = r->q[1];
    uint64_t inv = vo_inv_mod(q0 % q1, q1);
    uint64_t a0m = a0 % q1;
    uint64_t diff = a1 >= a0m ? a1 - a0m : a1 + q1 - a0m;
    uint64_t t = (uint64_t)(((u128)diff * inv) % q1);
    return (u128)a0 + (u128)q0 * t;
}

/* Centred lift into (-Q/2, Q/2] (rns/element.rs:183-191, modulus.rs:167-175). */
i128 vo_centered(const vo_ring *r, const uint64_t *p, size_t i) {
    u128 Q = vo_big_q(r);
    u128 v = vo_reconstruct(r, p[i], r->nl == 2 ? p[r->n + i] : 0);
    return v <= Q / 2 ? (i128)v : (i128)v - (i128)Q;
}

/* from_u128 per limb (rns_element.rs:165-176). */
void vo_set_u128(const vo_ring *r, uint64_t *p, size_t i, u128 v) {
    for (uint32_t l = 0; l < r->nl; l++) p[(size_t)l * r->n + i] = vo_mod_u128(v, r->q[l]);
}

/* from_centered_i128s (ring/abstraction.rs:446-468): reduce |c|, negate. */
void vo_set_i128(const vo_ring *r, uint64_t *p, size_t i, i128 c) {
    u128 mag = c < 0 ? (u128)(-(c + 1)) + 1 : (u128)c;
    for (uint32_t l = 0; l < r->nl; l++) {
        uint64_t q = r->q[l];
        uint64_t v = vo_mod_u128(mag, q);
        if (c < 0 && v) v = q - v;
        p[(size_t)l * r->n + i] = v;
    }
}

/* ---- NTT tables (ring/ntt.rs:155-272) ----------------------------------- */

static int is_pow2(uint64_t x) { return x && !(x & (x - 1)); }

static uint32_t bitrev(uint32_t x, unsigned logn) {
    uint32_t r = 0;
    for (unsigned i = 0; i < logn; i++)
        if ((x >> i) & 1) r |= 1u << (logn - 1 - i);
    return r;
}

/* find_primitive_2n_th_root (ntt.rs:155-198): the smallest base g = 2, 3, ...
 * whose g^((q-1)/2N) has order exactly 2N; at most 1024 bases are tried. */
uint64_t vo_ntt_psi(uint64_t q, uint32_t n) {
    uint64_t two_n = 2ULL * n;
    if (q < 2 || (q - 1) % two_n) return 0;
    uint64_t quot = (q - 1) / two_n;
    unsigned log2n = (unsigned)__builtin_ctzll(two_n);
    uint64_t cap = q < 1026 ? q : 1026;
    for (uint64_t base = 2; base < cap; base++) {
        uint64_t cand = vo_pow_mod(base, quot, q), cur = cand;
        int valid = 1;
        for (unsigned i = 0; i < log2n; i++) {
            if (cur == 1) { valid = 0; break; }
            cur = (uint64_t)(((u128)cur * cur) % q);
        }
        if (valid && cur == 1) return cand;
    }
    return 0;
}

int vo_ntt_friendly(const vo_ring *r) {
    if (r->n < 2) return 0;
    for (uint32_t l = 0; l < r->nl; l++) {
        uint64_t q = r->q[l];
        if (is_pow2(q) || (q - 1) % (2ULL * r->n)) return 0;
        if (!vo_ntt_psi(q, r->n)) return 0;
    }
    return 1;
}

#define MAX_TABLES 128
static vo_ntt_table g_tables[MAX_TABLES];
static int g_ntables = 0;
static pthread_mutex_t g_tab_mu = PTHREAD_MUTEX_INITIALIZER;

static uint64_t shoup(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }

const vo_ntt_table *vo_table(uint64_t q, uint32_t n) {
    pthread_mutex_lock(&g_tab_mu);
    for (int i = 0; i < g_ntables; i++)
        if (g_tables[i].q == q && g_tables[i].n == n) {
            pthread_mutex_unlock(&g_tab_mu);
            return &g_tables[i];
        }
    if (g_ntables == MAX_TABLES) abort();
    vo_ntt_table *t = &g_tables[g_ntables];
    unsigned logn = (unsigned)__builtin_ctz(n);
    uint64_t psi = vo_ntt_psi(q, n);
    uint64_t psi_inv = vo_inv_mod(psi, q);
    t->q = q;
    t->n = n;
    t->psi = psi;
    t->n_inv = vo_inv_mod(n % q, q);
    t->n_inv_sh = shoup(t->n_inv, q);
    t->fwd = (uint64_t *)malloc(sizeof(uint64_t) * n * 4);
    t->fwd_sh = t->fwd + n;
    t->inv = t->fwd + 2 * n;
    t->inv_sh = t->fwd + 3 * n;
    /* build_twiddle_table_bit_reversed (ntt.rs:206-221). */
    uint64_t cf = 1, ci = 1;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t br = bitrev(k, logn);
        t->fwd[br] = cf;
        t->inv[br] = ci;
        cf = (uint64_t)(((u128)cf * psi) % q);
        ci = (uint64_t)(((u128)ci * psi_inv) % q);
    }
    for (uint32_t k = 0; k < n; k++) {
        t->fwd_sh[k] = shoup(t->fwd[k], q);
        t->inv_sh[k] = shoup(t->inv[k], q);
    }
    g_ntables++;
    pthread_mutex_unlock(&g_tab_mu);
    return t;
}

/* Shoup product w*a mod q with w' = floor(w 2^64 / q); a < 2^64, q < 2^63. */
static inline uint64_t mul_shoup(uint64_t a, uint64_t w, uint64_t wsh, uint64_t q) {
    uint64_t t = (uint64_t)(((u128)a * wsh) >> 64);
    uint64_t r = a * w - t * q;
    return r >= q ? r - q : r;
}

/* cooley_tukey_dit_forward (ntt.rs:305-343): natural in, bit-reversed out. */
void vo_ntt_fwd_limb(const vo_ntt_table *t, uint64_t *buf) {
    uint32_t n = t->n;
    uint64_t q = t->q;
    for (uint32_t m = 1, half = n >> 1; m < n; m <<= 1, half >>= 1) {
        for (uint32_t blk = 0; blk < m; blk++) {
            uint64_t w = t->fwd[m + blk], wsh = t->fwd_sh[m + blk];
            uint64_t *x = buf + (size_t)blk * 2 * half, *y = x + half;
            for (uint32_t j = 0; j < half; j++) {
                uint64_t a = x[j];
                uint64_t b = mul_shoup(y[j], w, wsh, q);
                uint64_t s = a + b;
                x[j] = s >= q ? s - q : s;
                y[j] = a >= b ? a - b : a + q - b;
            }
        }
    }
}

/* cooley_tukey_dif_inverse (ntt.rs:356-399): bit-reversed in, natural out,
 * then every lane times N^-1. */
void vo_ntt_inv_limb(const vo_ntt_table *t, uint64_t *buf) {
    uint32_t n = t->n;
    uint64_t q = t->q;
    for (uint32_t half = 1, m = n >> 1; half < n; half <<= 1, m >>= 1) {
        for (uint32_t blk = 0; blk < m; blk++) {
            uint64_t w = t->inv[m + blk], wsh = t->inv_sh[m + blk];
            uint64_t *x = buf + (size_t)blk * 2 * half, *y = x + half;
            for (uint32_t j = 0; j < half; j++) {
                uint64_t a = x[j], b = y[j];
                uint64_t s = a + b;
                x[j] = s >= q ? s - q : s;
                uint64_t d = a >= b ? a - b : a + q - b;
                y[j] = mul_shoup(d, w, wsh, q);
            }
        }
    }
    for (uint32_t i = 0; i < n; i++) buf[i] = mul_shoup(buf[i], t->n_inv, t->n_inv_sh, q);
}

void vo_ntt_fwd(const vo_ring *r, uint64_t *p) {
    for (uint32_t l = 0; l < r->nl; l++) vo_ntt_fwd_limb(vo_table(r->q[l], r->n), p + (size_t)l * r->n);
}

void vo_ntt_inv(const vo_ring *r, uint64_t *p) {
    for (uint32_t l = 0; l < r->nl; l++) vo_ntt_inv_limb(vo_table(r->q[l], r->n), p + (size_t)l * r->n);
}

/* ---- "eval" form: NTT for NTT-friendly rings, identity otherwise -------- */
/* (ring/ntt_abstraction.rs:61-158: the schoolbook backend is the identity
 * transform with negacyclic multiplication in place of the Hadamard one.) */

void vo_to_eval(const vo_ring *r, uint64_t *p) {
    if (vo_ntt_friendly(r)) vo_ntt_fwd(r, p);
}

void vo_from_eval(const vo_ring *r, uint64_t *p) {
    if (vo_ntt_friendly(r)) vo_ntt_inv(r, p);
}

/* Schoolbook negacyclic product (ring/ops.rs:65-111). */
static void schoolbook(uint32_t n, uint64_t q, uint64_t *out, const uint64_t *a, const uint64_t *b) {
    u128 *acc = (u128 *)calloc(n, sizeof(u128));
    int pow2 = is_pow2(q);
    for (uint32_t i = 0; i < n; i++) {
        if (!a[i]) continue;
        for (uint32_t j = 0; j < n; j++) {
            u128 prod = (u128)a[i] * b[j];
            uint64_t pr = pow2 ? (uint64_t)prod & (q - 1) : (uint64_t)(prod % q);
            uint32_t k = i + j;
            if (k < n) acc[k] += pr;
            else acc[k - n] += (u128)(q - pr) % q;
        }
    }
    for (uint32_t k = 0; k < n; k++) out[k] = vo_mod_u128(acc[k], q);
    free(acc);
}

/* out (+)= a (.) b in eval form. For NTT rings a Hadamard product per limb. */
void vo_eval_mul_acc(const vo_ring *r, uint64_t *out, const uint64_t *a, const uint64_t *b, int acc) {
    size_t n = r->n;
    if (vo_ntt_friendly(r)) {
        for (uint32_t l = 0; l < r->nl; l++) {
            uint64_t q = r->q[l];
            const uint64_t *al = a + l * n, *bl = b + l * n;
            uint64_t *ol = out + l * n;
            for (size_t i = 0; i < n; i++) {
                uint64_t v = (uint64_t)(((u128)al[i] * bl[i]) % q);
                if (acc) {
                    v += ol[i];
                    if (v >= q) v -= q;
                }
                ol[i] = v;
            }
        }
    } else {
        uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * n);
        for (uint32_t l = 0; l < r->nl; l++) {
            uint64_t q = r->q[l];
            schoolbook(r->n, q, tmp, a + l * n, b + l * n);
            uint64_t *ol = out + l * n;
            for (size_t i = 0; i < n; i++) {
                uint64_t v = tmp[i];
                if (acc) v = vo_mod_u128((u128)v + ol[i], q);
                ol[i] = v;
            }
        }
        free(tmp);
    }
}

void vo_poly_mul(const vo_ring *r, uint64_t *out, const uint64_t *a, const uint64_t *b) {
    size_t w = (size_t)r->nl * r->n;
    uint64_t *ta = vo_alloc(w), *tb = vo_alloc(w);
    memcpy(ta, a, w * 8);
    memcpy(tb, b, w * 8);
    vo_to_eval(r, ta);
    vo_to_eval(r, tb);
    vo_eval_mul_acc(r, out, ta, tb, 0);
    vo_from_eval(r, out);
    free(ta);
    free(tb);
}

/* ---- coefficient-wise ops ------------------------------------------------ */

void vo_poly_add(const vo_ring *r, uint64_t *out, const uint64_t *a, const uint64_t *b) {
    for (uint32_t l = 0; l < r->nl; l++) {
        uint64_t q = r->q[l];
        for (size_t i = 0; i < r->n; i++) {
            size_t k = (size_t)l * r->n + i;
            uint64_t s = a[k] + b[k];
            out[k] = s >= q ? s - q : s;
        }
    }
}

void vo_poly_sub(const vo_ring *r, uint64_t *out, const uint64_t *a, const uint64_t *b) {
    for (uint32_t l = 0; l < r->nl; l++) {
        uint64_t q = r->q[l];
        for (size_t i = 0; i < r->n; i++) {
            size_t k = (size_t)l * r->n + i;
            out[k] = a[k] >= b[k] ? a[k] - b[k] : a[k] + q - b[k];
        }
    }
}

void vo_poly_neg(const vo_ring *r, uint64_t *out, const uint64_t *a) {
    for (uint32_t l = 0; l < r->nl; l++) {
        uint64_t q = r->q[l];
        for (size_t i = 0; i < r->n; i++) {
            size_t k = (size_t)l * r->n + i;
            out[k] = a[k] ? q - a[k] : 0;
        }
    }
}

/* Multiply by a scalar given as u128 (reduced per limb first). */
void vo_poly_scale(const vo_ring *r, uint64_t *out, const uint64_t *a, u128 s) {
    for (uint32_t l = 0; l < r->nl; l++) {
        uint64_t q = r->q[l], sl = vo_mod_u128(s, q);
        for (size_t i = 0; i < r->n; i++) {
            size_t k = (size_t)l * r->n + i;
            out[k] = vo_mod_u128((u128)a[k] * sl, q);
        }
    }
}

/* rotate_slice (ring/ops.rs:147-180): multiply by X^k in Z[X]/(X^n+1). */
void vo_mul_x_pow(const vo_ring *r, const uint64_t *src, size_t k, uint64_t *dst) {
    size_t n = r->n;
    size_t keff = k % (2 * n), kred = keff % n;
    int negg = keff >= n;
    for (uint32_t l = 0; l < r->nl; l++) {
        uint64_t q = r->q[l];
        const uint64_t *s = src + l * n;
        uint64_t *d = dst + l * n;
        for (size_t i = 0; i < n; i++) {
            int wrapped = i + kred >= n;
            size_t pos = wrapped ? i + kred - n : i + kred;
            uint64_t v = s[i];
            d[pos] = (wrapped ^ negg) ? (v ? q - v : 0) : v;
        }
    }
}

/* embed_at_slice / project_at_slice (ring/reshape.rs:58-116) per limb.
 * src has degree ns, dst degree nd. */
void vo_embed_at(uint32_t nl, uint32_t ns, uint32_t nd, const uint64_t *src, size_t slot, uint64_t *dst) {
    uint32_t d = nd / ns;
    memset(dst, 0, sizeof(uint64_t) * nl * nd);
    for (uint32_t l = 0; l < nl; l++)
        for (uint32_t i = 0; i < ns; i++) dst[(size_t)l * nd + (size_t)d * i + slot] = src[(size_t)l * ns + i];
}

void vo_project_at(uint32_t nl, uint32_t ns, uint32_t nd, const uint64_t *src, size_t slot, uint64_t *dst) {
    uint32_t d = ns / nd;
    for (uint32_t l = 0; l < nl; l++)
        for (uint32_t i = 0; i < nd; i++) dst[(size_t)l * nd + i] = src[(size_t)l * ns + (size_t)d * i + slot];
}

/* RingPoly::random_uniform: Poly::random (element.rs:235) draws lanes in
 * order; PolyRns::random (rns_element.rs:201-212) interleaves limb draws per
 * lane: m0[0], m1[0], m0[1], m1[1], ... */
void vo_poly_random(const vo_ring *r, vo_prg *g, uint64_t *out) {
    for (size_t i = 0; i < r->n; i++)
        for (uint32_t l = 0; l < r->nl; l++) out[(size_t)l * r->n + i] = vo_prg_uniform_below(g, r->q[l]);
}

/* from_centered_i64s (ring/abstraction.rs:315-323, 432-444). */
void vo_poly_from_centered(const vo_ring *r, const int64_t *c, uint64_t *out) {
    for (uint32_t l = 0; l < r->nl; l++)
        for (size_t i = 0; i < r->n; i++) out[(size_t)l * r->n + i] = vo_mod_i64(c[i], r->q[l]);
}

uint64_t *vo_alloc(size_t words) {
    uint64_t *p = (uint64_t *)calloc(words ? words : 1, sizeof(uint64_t));
    if (!p) abort();
    return p;
}
