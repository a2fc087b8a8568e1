/* vo_internal.h — shared internals of the CPU restatement (test infrastructure). */
#ifndef VO_INTERNAL_H
#define VO_INTERNAL_H

#include "via_oracle.h"

typedef unsigned __int128 u128;
typedef __int128 i128;

typedef struct {
    uint64_t q;
    uint32_t n;
    uint64_t psi, n_inv, n_inv_sh;
    uint64_t *fwd, *fwd_sh, *inv, *inv_sh;
} vo_ntt_table;

uint64_t vo_mod_u128(u128 x, uint64_t q);
uint64_t vo_mod_i64(int64_t x, uint64_t q);
uint64_t vo_pow_mod(uint64_t b, uint64_t e, uint64_t q);
uint64_t vo_inv_mod(uint64_t a, uint64_t q);
u128 vo_big_q(const vo_ring *r);
u128 vo_reconstruct(const vo_ring *r, uint64_t a0, uint64_t a1);
i128 vo_centered(const vo_ring *r, const uint64_t *p, size_t i);
void vo_set_u128(const vo_ring *r, uint64_t *p, size_t i, u128 v);
void vo_set_i128(const vo_ring *r, uint64_t *p, size_t i, i128 c);

const vo_ntt_table *vo_table(uint64_t q, uint32_t n);
void vo_ntt_fwd_limb(const vo_ntt_table *t, uint64_t *buf);
void vo_ntt_inv_limb(const vo_ntt_table *t, uint64_t *buf);
void vo_to_eval(const vo_ring *r, uint64_t *p);
void vo_from_eval(const vo_ring *r, uint64_t *p);
void vo_eval_mul_acc(const vo_ring *r, uint64_t *out, const uint64_t *a, const uint64_t *b, int acc);

void vo_poly_add(const vo_ring *r, uint64_t *out, const uint64_t *a, const uint64_t *b);
void vo_poly_sub(const vo_ring *r, uint64_t *out, const uint64_t *a, const uint64_t *b);
void vo_poly_neg(const vo_ring *r, uint64_t *out, const uint64_t *a);
void vo_poly_scale(const vo_ring *r, uint64_t *out, const uint64_t *a, u128 s);
void vo_embed_at(uint32_t nl, uint32_t ns, uint32_t nd, const uint64_t *src, size_t slot, uint64_t *dst);
void vo_project_at(uint32_t nl, uint32_t ns, uint32_t nd, const uint64_t *src, size_t slot, uint64_t *dst);
uint64_t *vo_alloc(size_t words);

static inline size_t vo_pw(const vo_ring *r) { return (size_t)r->nl * r->n; }
static inline vo_ring vo_ring_make(uint32_t n, uint32_t nl, uint64_t q0, uint64_t q1) {
    vo_ring r;
    r.n = n;
    r.nl = nl;
    r.q[0] = q0;
    r.q[1] = nl == 2 ? q1 : 0;
    return r;
}

/* Gadget digits of one polynomial, LSB first: digits[k*n + i] is the k-th
 * extracted digit of coefficient i. */
void vo_gadget_digits_lsb(const vo_ring *r, const uint64_t *in, uint64_t base, uint32_t L, int64_t *digits);
/* Gadget product against an RLev already in eval form (RLevEval). */
void vo_gadget_product_eval(const vo_ring *r, const uint64_t *rlev_eval, uint32_t L, uint64_t base,
                            const uint64_t *in, uint64_t *acc_mask_eval, uint64_t *acc_body_eval, int acc);
void vo_rlev_to_eval(const vo_ring *r, const uint64_t *rlev, uint32_t L, uint64_t *out);

#endif
