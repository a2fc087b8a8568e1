/*
 * via_oracle.h — CPU restatement of the via-rs VIA-C and VIA-B server answer path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the bit-exactness checker for the
 * HIP implementation in via-rs_amd/ and the timed CPU baseline in bench.py.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker; the product path never links or calls it.
 *
 * Every function follows a reference file (paths relative to the via-rs
 * checkout); the citation sits next to each definition in via_oracle.c.
 *
 * Conventions
 *   - A ring is (n, nl, q[nl]): Z_Q[X]/(X^n+1) with Q = q[0] (nl == 1) or the
 *     two-prime RNS Q = q[0]*q[1] (nl == 2).
 *   - A polynomial is nl*n uint64 canonical residues, limb-major:
 *     p[l*n + i] is coefficient i mod q[l].
 *   - An RLWE ciphertext is mask then body: 2*nl*n words.
 *   - An RLev of depth L is L RLWE samples; sample k pairs with gadget entry
 *     g_k = round(Q / B^(k+1)) (MSB first).
 *   - An RGSW (L1, L2) is RLev(-s*m, L1) followed by RLev(m, L2).
 *   - An LWE of dimension n is, per limb, n masks followed by the body:
 *     lwe[l*(n+1) + i] (i < n) and lwe[l*(n+1) + n].
 *   - Eval (NTT) form of a polynomial is bit-reversed, as in ring/ntt.rs.
 *     For a modulus that is not NTT-friendly at n the "eval" form is the
 *     coefficient form and products are schoolbook (ntt_abstraction.rs).
 */
#ifndef VIA_ORACLE_H
#define VIA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t n;
    uint32_t nl;
    uint64_t q[2];
} vo_ring;

typedef struct vo_prg vo_prg;

/* Error-distribution selector: 0 = ternary, 1 = bounded uniform [-b, b]. */
typedef struct {
    int kind;
    uint32_t bound;
} vo_dist;

/* ---- PRG: SHAKE-256 counter mode (sampling/prg.rs) ---------------------- */
vo_prg *vo_prg_new(const uint8_t *seed, size_t len);
void vo_prg_free(vo_prg *g);
void vo_prg_fill(vo_prg *g, uint8_t *out, size_t len);
uint64_t vo_prg_uniform_below(vo_prg *g, uint64_t bound);
void vo_sample(vo_prg *g, vo_dist d, int64_t *out, size_t len);

/* ---- rings and NTT ------------------------------------------------------ */
int vo_ntt_friendly(const vo_ring *r);
uint64_t vo_ntt_psi(uint64_t q, uint32_t n);
void vo_ntt_fwd(const vo_ring *r, uint64_t *p);  /* all limbs, in place */
void vo_ntt_inv(const vo_ring *r, uint64_t *p);
void vo_poly_mul(const vo_ring *r, uint64_t *out, const uint64_t *a, const uint64_t *b);
void vo_poly_random(const vo_ring *r, vo_prg *g, uint64_t *out);
void vo_poly_from_centered(const vo_ring *r, const int64_t *c, uint64_t *out);
void vo_mul_x_pow(const vo_ring *r, const uint64_t *src, size_t k, uint64_t *dst);

/* ---- encryption layer --------------------------------------------------- */
void vo_keygen(const vo_ring *r, vo_dist d, vo_prg *g, uint64_t *sk);
void vo_encrypt(const vo_ring *r, const uint64_t *sk, const uint64_t *m, vo_dist d,
                vo_prg *g, uint64_t *ct);
void vo_encode(const vo_ring *r, uint64_t p, const uint64_t *pt, uint64_t *out);
void vo_decode(const vo_ring *r, uint64_t p, const uint64_t *in, uint64_t *pt);
void vo_gadget_vector(const vo_ring *r, uint64_t base, uint32_t L, uint64_t *out_lo_hi);
void vo_encrypt_rlev(const vo_ring *r, const uint64_t *sk, const uint64_t *m, uint64_t base,
                     uint32_t L, vo_dist d, vo_prg *g, uint64_t *rlev);
void vo_encrypt_rgsw(const vo_ring *r, const uint64_t *sk, const uint64_t *m, uint64_t b1,
                     uint64_t b2, uint32_t L1, uint32_t L2, vo_dist d, vo_prg *g,
                     uint64_t *rgsw);
void vo_gadget_decompose(const vo_ring *r, const uint64_t *in, uint64_t base, uint32_t L,
                         int64_t *digits /* [L][n], MSB first */);
void vo_gadget_product(const vo_ring *r, const uint64_t *rlev, uint32_t L, uint64_t base,
                       const uint64_t *in, uint64_t *ct);
void vo_external_product(const vo_ring *r, const uint64_t *rgsw, uint32_t L1, uint32_t L2,
                         uint64_t b1, uint64_t b2, const uint64_t *ct, uint64_t *out);
void vo_key_switch(const vo_ring *r, const uint64_t *rlev, uint32_t L, uint64_t base,
                   const uint64_t *ct, uint64_t *out);

/* ---- gates -------------------------------------------------------------- */
void vo_rotate(const vo_ring *r, const uint64_t *ct, size_t k, uint64_t *out);
void vo_cmux(const vo_ring *r, const uint64_t *rgsw, uint32_t L, uint64_t base,
             const uint64_t *ct0, const uint64_t *ct1, uint64_t *out);
void vo_dmux(const vo_ring *r, const uint64_t *rgsw, uint32_t L, uint64_t base,
             const uint64_t *ct, uint64_t *out0, uint64_t *out1);
void vo_cmux_tree(const vo_ring *r, const uint64_t *rgsws, uint32_t nbits, uint32_t L,
                  uint64_t base, uint64_t *inputs /* 2^nbits cts, clobbered */, uint64_t *out);
void vo_dmux_tree(const vo_ring *r, const uint64_t *rgsws, uint32_t nbits, uint32_t L,
                  uint64_t base, const uint64_t *input, uint64_t *out /* 2^nbits cts */);
void vo_crot(const vo_ring *r, int slot_extract, const uint64_t *rgsws, uint32_t nbits,
             uint32_t L, uint64_t base, const uint64_t *ct, uint64_t *out);
void vo_rlwe_to_rgsw(const vo_ring *r, const uint64_t *levels /* L_OUT cts */,
                     const uint64_t *conv_key /* RLev L_CK */, const uint64_t *m_rlev,
                     uint64_t base_ck, uint32_t L_OUT, uint32_t L_CK, uint64_t *rgsw);

/* ---- switching ---------------------------------------------------------- */
void vo_mod_switch_sym(const vo_ring *src, const uint64_t *ct, const vo_ring *dst,
                       uint64_t *out);
void vo_mod_switch_asym(const vo_ring *src, const uint64_t *ct, uint64_t q_mask,
                        uint64_t q_body, uint64_t *out_mask, uint64_t *out_body);
void vo_mod_switch_rgsw(const vo_ring *src, const uint64_t *rgsw, uint32_t L1, uint32_t L2,
                        const vo_ring *dst, uint64_t *out);
void vo_rekey(const vo_ring *src, const uint64_t *sk, const vo_ring *dst, uint64_t *out);
void vo_gen_rsk(const vo_ring *r1, const uint64_t *s1, const vo_ring *r2, const uint64_t *s2,
                uint64_t base, uint32_t L, vo_dist d, vo_prg *g, uint64_t *rsk);
void vo_ring_switch(const vo_ring *r1, const uint64_t *ct, const vo_ring *r2,
                    const uint64_t *rsk, uint64_t base, uint32_t L, uint64_t *out);
void vo_decrypt_asym(const vo_ring *r3, const uint64_t *sk2, const uint64_t *mask,
                     const uint64_t *body, uint64_t q4, uint64_t p, uint64_t *pt);

/* ---- LWE -> RLWE cascade ------------------------------------------------ */
void vo_encrypt_lwe(const vo_ring *r, const uint64_t *sk, uint64_t msg, uint64_t p, vo_dist d,
                    vo_prg *g, uint64_t *lwe);
void vo_encrypt_lwe_raw(const vo_ring *r, const uint64_t *sk, uint64_t msg_lo,
                        uint64_t msg_hi, vo_dist d, vo_prg *g, uint64_t *lwe);
size_t vo_cascade_key_words(const vo_ring *r, uint32_t L);
void vo_gen_cascade_key(const vo_ring *r, const uint64_t *sk, uint64_t base, uint32_t L,
                        vo_dist d, vo_prg *g, uint64_t *key);
void vo_conv_step(const vo_ring *r, uint32_t n_out, const uint64_t *in_masks,
                  const uint64_t *in_body, const uint64_t *step_key, uint64_t base,
                  uint32_t L, uint64_t *out_masks, uint64_t *out_body);
void vo_cascade(const vo_ring *r, const uint64_t *lwe, const uint64_t *key, uint64_t base,
                uint32_t L, uint64_t *ct);

/* ---- protocol: parameters, client, server ------------------------------- */
typedef struct {
    uint32_t n1, n2, n_rec;
    uint32_t q1_nl;
    uint64_t q1[2];
    uint64_t q2, q3, q4, p;
    uint64_t b1, b2, b_rsk, ck_base;
    uint32_t l_query, l_ck, l_rsk;
    uint32_t num_rows, num_cols;
} vo_params;

size_t vo_words_cascade_key(const vo_params *pp);
size_t vo_words_conv_key(const vo_params *pp);
size_t vo_words_rsk(const vo_params *pp);
size_t vo_words_lwe(const vo_params *pp);
uint32_t vo_query_lwes(const vo_params *pp);

/* Client::setup (via-client/src/client.rs:89-173). Outputs sk1 (n1 at q1),
 * sk2 (n2 at q3) and the public keys in coefficient form. */
int vo_client_setup(const vo_params *pp, vo_dist key1, vo_dist key2, vo_dist err, vo_prg *g,
                    uint64_t *sk1, uint64_t *sk2, uint64_t *cascade_key, uint64_t *conv_key,
                    uint64_t *rsk);
int vo_client_query(const vo_params *pp, const uint64_t *sk1, size_t index, vo_dist err,
                    vo_prg *g, uint64_t *lwes);
void vo_client_recover(const vo_params *pp, const uint64_t *sk2, const uint64_t *ans_mask,
                       const uint64_t *ans_body, uint64_t *record);

/* Server: setup_db + PreparedDb + PreparedKeys, then answer. */
typedef struct vo_server vo_server;
int vo_server_new(const vo_params *pp, const uint64_t *records, size_t n_records,
                  const uint64_t *cascade_key, const uint64_t *conv_key, const uint64_t *rsk,
                  vo_server **out);
void vo_server_free(vo_server *s);
/* Returns 0, or 1 = DimMismatch, 2 = QueryLengthMismatch (expected/got set). */
int vo_server_answer(const vo_server *s, const uint64_t *lwes, size_t n_lwe,
                     uint64_t *ans_mask, uint64_t *ans_body, size_t *expected, size_t *got);
/* Per-step wall time of the last answer, seconds, steps 1..7. */
void vo_server_step_times(const vo_server *s, double *t7);
/* Intermediate access for parity tests. */
int vo_server_answer_stages(const vo_server *s, const uint64_t *lwes, size_t n_lwe,
                            uint64_t *rgsw_bits, uint64_t *dmux_out, uint64_t *switched,
                            uint64_t *fd_out, uint64_t *cmux_out, uint64_t *crot_out,
                            uint64_t *ans_mask, uint64_t *ans_body);
/* Prepared (eval form at q2) database cells, [I][J][n1]. */
const uint64_t *vo_server_db(const vo_server *s);
void vo_setup_db(const vo_params *pp, const uint64_t *records, size_t n_records,
                 uint64_t *cells /* [I][J][n1] mod p */);
void vo_first_dim(const vo_params *pp, const uint64_t *switched /* I cts at q2 */,
                  const uint64_t *db_eval, int threads, uint64_t *out /* J cts */);

/* VIA-B: Server::answer_batch::<T> over T queries stored back to back (n_lwe
 * LWEs each); extraction degree n2/T. Returns 0, 1 = DimMismatch,
 * 2 = QueryLengthMismatch. */
int vo_server_answer_batch(const vo_server *s, uint32_t T, const uint64_t *lwes, size_t n_lwe, uint64_t *ans_mask,
                           uint64_t *ans_body, size_t *expected, size_t *got);
/* Repacking pieces (conversion/extract.rs, conversion/repack.rs). r is the
 * single-limb ring at the repack modulus. */
void vo_extr(const vo_ring *r, const uint64_t *ct, uint32_t G, uint64_t *masks, uint64_t *body);
size_t vo_repack_key_words(uint32_t n, uint32_t G, uint32_t L);
void vo_repack_key_from_cascade(const vo_ring *r1, const uint64_t *cascade_key, uint64_t q2, uint32_t G, uint32_t L,
                                uint64_t *out);
int vo_repack(const vo_ring *r2, uint32_t T, uint32_t G, const uint64_t *rotateds, const uint64_t *key,
              uint64_t base, uint32_t L, uint64_t *out);
void vo_embed_at(uint32_t nl, uint32_t ns, uint32_t nd, const uint64_t *src, size_t slot, uint64_t *dst);
void vo_project_at(uint32_t nl, uint32_t ns, uint32_t nd, const uint64_t *src, size_t slot, uint64_t *dst);

/* Threads the parallel parts use (the CPUs of this process's affinity mask). */
int vo_host_threads(void);
/* Threads for the query decomposition's per-LWE cascades and conversions
 * (default 1: the reference's serial loop, as the CPU baseline times it;
 * tests raise it, the values are the same). */
void vo_set_decomp_threads(int n);

/* Wide helpers exposed for tests of the 256-bit path. */
void vo_round_mul_div(uint64_t a_lo, uint64_t a_hi, uint64_t s_lo, uint64_t s_hi,
                      uint64_t d_lo, uint64_t d_hi, uint64_t *q_lo, uint64_t *q_hi);

#ifdef __cplusplus
}
#endif
#endif
