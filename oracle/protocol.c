/*
 * protocol.c — LWE->RLWE cascade, VIA-C client and server.
 * Test infrastructure (see via_oracle.h).
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "vo_internal.h"

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* Worker threads for the parallel parts: the CPUs this process may run on. */
int vo_host_threads(void) {
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof set, &set) == 0) {
        int c = CPU_COUNT(&set);
        if (c > 0) return c;
    }
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

/* ---- cascade -------------------------------------------------------------- */

/* Offset (words) of the step with output degree nout inside a cascade key. */
static size_t cascade_step_off(uint32_t n, uint32_t nl, uint32_t L, uint32_t nout) {
    size_t off = 0;
    for (uint32_t o = 2; o < nout; o <<= 1) off += (size_t)2 * n * L * 2 * nl;
    return off;
}

/* conv_step (conv.rs:62-95) against a step key already in eval form: the body
 * is embedded at slot 0; input mask idx = RANK_OUT*group + j is embedded at
 * slot 0 and key-switched; output mask j subtracts the switched masks of its
 * two groups, the body subtracts every switched body. The sums run in eval
 * form (the transform is linear), with one inverse per output polynomial. */
static void conv_step_eval(const vo_ring *r, uint32_t nout, const uint64_t *in_masks, const uint64_t *in_body,
                           const uint64_t *key_eval, uint64_t base, uint32_t L, uint64_t *out_masks,
                           uint64_t *out_body) {
    uint32_t nl = r->nl, nin = nout / 2, rank_in = r->n / nin, rank_out = rank_in / 2;
    vo_ring ro = vo_ring_make(nout, nl, r->q[0], r->q[1]);
    size_t wi = (size_t)nl * nin, wo = (size_t)nl * nout, kw = (size_t)L * 2 * wo;
    uint64_t *emb = vo_alloc(wo), *gm = vo_alloc(wo), *gb = vo_alloc(wo), *acc_b = vo_alloc(wo);
    for (uint32_t j = 0; j < rank_out; j++) memset(out_masks + j * wo, 0, 8 * wo);
    for (uint32_t group = 0; group < 2; group++)
        for (uint32_t j = 0; j < rank_out; j++) {
            uint32_t idx = rank_out * group + j;
            vo_embed_at(nl, nin, nout, in_masks + idx * wi, 0, emb);
            vo_gadget_product_eval(&ro, key_eval + idx * kw, L, base, emb, gm, gb, 0);
            vo_poly_add(&ro, out_masks + j * wo, out_masks + j * wo, gm);
            vo_poly_add(&ro, acc_b, acc_b, gb);
        }
    for (uint32_t j = 0; j < rank_out; j++) {
        vo_from_eval(&ro, out_masks + j * wo);
        vo_poly_neg(&ro, out_masks + j * wo, out_masks + j * wo);
    }
    vo_from_eval(&ro, acc_b);
    vo_embed_at(nl, nin, nout, in_body, 0, emb);
    vo_poly_sub(&ro, out_body, emb, acc_b);
    free(emb);
    free(gm);
    free(gb);
    free(acc_b);
}

/* Cascade key to eval form, step by step (cascade.rs:338-380). */
static void cascade_key_to_eval(const vo_ring *r, const uint64_t *key, uint32_t L, uint64_t *out) {
    uint32_t n = r->n, nl = r->nl;
    size_t off = 0;
    for (uint32_t nout = 2; nout <= n; nout <<= 1) {
        vo_ring ro = vo_ring_make(nout, nl, r->q[0], r->q[1]);
        size_t wo = (size_t)nl * nout, polys = (size_t)(2 * n / nout) * L * 2;
        memcpy(out + off, key + off, 8 * polys * wo);
        for (size_t k = 0; k < polys; k++) vo_to_eval(&ro, out + off + k * wo);
        off += polys * wo;
    }
}

void vo_conv_step(const vo_ring *r, uint32_t nout, const uint64_t *in_masks, const uint64_t *in_body,
                  const uint64_t *step_key, uint64_t base, uint32_t L, uint64_t *out_masks, uint64_t *out_body) {
    uint32_t nl = r->nl;
    vo_ring ro = vo_ring_make(nout, nl, r->q[0], r->q[1]);
    size_t wo = (size_t)nl * nout, polys = (size_t)(2 * r->n / nout) * L * 2;
    uint64_t *ev = vo_alloc(polys * wo);
    memcpy(ev, step_key, 8 * polys * wo);
    for (size_t k = 0; k < polys; k++) vo_to_eval(&ro, ev + k * wo);
    conv_step_eval(r, nout, in_masks, in_body, ev, base, L, out_masks, out_body);
    free(ev);
}

/* lwe_to_rlwe_* (cascade.rs:65-78, 388-405): an (n,1)-MLWE becomes an
 * (1,n)-MLWE through log2(n) conv steps. */
static void cascade_eval(const vo_ring *r, const uint64_t *lwe, const uint64_t *key_eval, uint64_t base, uint32_t L,
                         uint64_t *ct) {
    uint32_t n = r->n, nl = r->nl;
    /* rank-n degree-1 MLWE, masks [n][nl][1] */
    uint64_t *masks = vo_alloc((size_t)n * nl), *body = vo_alloc(nl);
    for (uint32_t l = 0; l < nl; l++) {
        for (uint32_t i = 0; i < n; i++) masks[(size_t)i * nl + l] = lwe[(size_t)l * (n + 1) + i];
        body[l] = lwe[(size_t)l * (n + 1) + n];
    }
    for (uint32_t nout = 2; nout <= n; nout <<= 1) {
        uint32_t rank_out = n / nout;
        uint64_t *om = vo_alloc((size_t)rank_out * nl * nout), *ob = vo_alloc((size_t)nl * nout);
        conv_step_eval(r, nout, masks, body, key_eval + cascade_step_off(n, nl, L, nout), base, L, om, ob);
        free(masks);
        free(body);
        masks = om;
        body = ob;
    }
    size_t w = (size_t)nl * n;
    memcpy(ct, masks, 8 * w);
    memcpy(ct + w, body, 8 * w);
    free(masks);
    free(body);
}

void vo_cascade(const vo_ring *r, const uint64_t *lwe, const uint64_t *key, uint64_t base, uint32_t L,
                uint64_t *ct) {
    uint64_t *ev = vo_alloc(vo_cascade_key_words(r, L));
    cascade_key_to_eval(r, key, L, ev);
    cascade_eval(r, lwe, ev, base, L, ct);
    free(ev);
}

/* ---- parameters ------------------------------------------------------------ */

static vo_ring ring_q1(const vo_params *pp) { return vo_ring_make(pp->n1, pp->q1_nl, pp->q1[0], pp->q1[1]); }
static vo_ring ring_q2(const vo_params *pp) { return vo_ring_make(pp->n1, 1, pp->q2, 0); }
static vo_ring ring_q3n1(const vo_params *pp) { return vo_ring_make(pp->n1, 1, pp->q3, 0); }
static vo_ring ring_q3n2(const vo_params *pp) { return vo_ring_make(pp->n2, 1, pp->q3, 0); }

static uint32_t ilog2(uint32_t x) { return x <= 1 ? 0 : 32 - (uint32_t)__builtin_clz(x - 1); }

size_t vo_words_cascade_key(const vo_params *pp) {
    vo_ring r = ring_q1(pp);
    return vo_cascade_key_words(&r, pp->l_ck);
}
size_t vo_words_conv_key(const vo_params *pp) { return (size_t)pp->l_ck * 2 * pp->q1_nl * pp->n1; }
size_t vo_words_rsk(const vo_params *pp) { return (size_t)(pp->n1 / pp->n2) * pp->l_rsk * 2 * pp->n2; }
size_t vo_words_lwe(const vo_params *pp) { return (size_t)pp->q1_nl * (pp->n1 + 1); }
uint32_t vo_query_lwes(const vo_params *pp) {
    return (ilog2(pp->num_rows) + ilog2(pp->num_cols) + ilog2(pp->n1 / pp->n_rec)) * pp->l_query;
}

/* ---- client ----------------------------------------------------------------- */

/* Client::setup (client.rs:89-173). PRG order: keygen S1, keygen S2, cascade
 * key, RLev_S1(S1^2), then the ring-switch key of S1 rekeyed to q3. */
int vo_client_setup(const vo_params *pp, vo_dist key1, vo_dist key2, vo_dist err, vo_prg *g, uint64_t *sk1,
                    uint64_t *sk2, uint64_t *cascade_key, uint64_t *conv_key, uint64_t *rsk) {
    if (!pp->num_rows || !pp->num_cols || (pp->num_rows & (pp->num_rows - 1)) ||
        (pp->num_cols & (pp->num_cols - 1)))
        return 1;
    vo_ring r1 = ring_q1(pp), r3n2 = ring_q3n2(pp), r3n1 = ring_q3n1(pp);
    vo_keygen(&r1, key1, g, sk1);
    vo_keygen(&r3n2, key2, g, sk2);
    vo_gen_cascade_key(&r1, sk1, pp->ck_base, pp->l_ck, err, g, cascade_key);
    /* gen_rlwe_to_rgsw_key (convert.rs:46-58): RLev_S(S^2). */
    size_t w = vo_pw(&r1);
    uint64_t *s2 = vo_alloc(w);
    vo_poly_mul(&r1, s2, sk1, sk1);
    vo_encrypt_rlev(&r1, sk1, s2, pp->ck_base, pp->l_ck, err, g, conv_key);
    free(s2);
    uint64_t *s1q3 = vo_alloc(pp->n1);
    vo_rekey(&r1, sk1, &r3n1, s1q3);
    vo_gen_rsk(&r3n1, s1q3, &r3n2, sk2, pp->b_rsk, pp->l_rsk, err, g, rsk);
    free(s1q3);
    return 0;
}

/* Client::query -> build_compressed_query (query.rs:75-130): (alpha, beta,
 * gamma) = decompose_index; DMux bits MSB-first with g(b1), CMux and CRot
 * bits LSB-first with g(b2), both gadgets at q1; each bit is L_QUERY raw
 * LWEs of bit * g_i. */
int vo_client_query(const vo_params *pp, const uint64_t *sk1, size_t index, vo_dist err, vo_prg *g,
                    uint64_t *lwes) {
    size_t I = pp->num_rows, J = pp->num_cols, d = pp->n1 / pp->n_rec;
    if (index >= d * I * J) return 3;
    vo_ring r1 = ring_q1(pp);
    u128 Q = vo_big_q(&r1);
    size_t group = I * J, gamma = index / group, alpha = (index % group) / J, beta = index % J;
    uint32_t nd = ilog2((uint32_t)I), nc = ilog2((uint32_t)J), nr = ilog2((uint32_t)d);
    size_t lw = vo_words_lwe(pp), c = 0;
    for (uint32_t grp = 0; grp < 3; grp++) {
        uint32_t nb = grp == 0 ? nd : (grp == 1 ? nc : nr);
        uint64_t base = grp == 0 ? pp->b1 : pp->b2;
        for (uint32_t i = 0; i < nb; i++) {
            int bit;
            if (grp == 0) bit = (int)((alpha >> (nd - 1 - i)) & 1);
            else if (grp == 1) bit = (int)((beta >> i) & 1);
            else bit = (int)((gamma >> i) & 1);
            u128 div = 1;
            for (uint32_t k = 0; k < pp->l_query; k++) {
                div *= base;
                u128 gi = (Q + div / 2) / div;
                u128 msg = bit ? gi % Q : 0;
                vo_encrypt_lwe_raw(&r1, sk1, (uint64_t)msg, (uint64_t)(msg >> 64), err, g, lwes + (c++) * lw);
            }
        }
    }
    return 0;
}

/* Client::recover (client.rs:222-239). */
void vo_client_recover(const vo_params *pp, const uint64_t *sk2, const uint64_t *ans_mask, const uint64_t *ans_body,
                       uint64_t *record) {
    vo_ring r3 = ring_q3n2(pp);
    vo_decrypt_asym(&r3, sk2, ans_mask, ans_body, pp->q4, pp->p, record);
}

/* ---- server ----------------------------------------------------------------- */

struct vo_server {
    vo_params pp;
    uint64_t *db_eval;      /* [I][J][n1] at q2, eval form */
    uint64_t *cascade_eval; /* per-step eval cascade key */
    uint64_t *conv_eval;    /* RLev(L_CK) eval at q1 */
    uint64_t *rsk_eval;     /* D x RLev(L_RSK) eval at (n2, q3) */
    uint64_t *cascade_coeff; /* the client's cascade key as shipped (VIA-B repack keys) */
    double t[7];
};

/* setup_db (setup_db.rs:39-75): record k*I*J + i*J + j lands in slot k of
 * cell (i, j); missing records are zero. */
void vo_setup_db(const vo_params *pp, const uint64_t *records, size_t n_records, uint64_t *cells) {
    size_t I = pp->num_rows, J = pp->num_cols, n1 = pp->n1, nr = pp->n_rec, d = n1 / nr;
    memset(cells, 0, 8 * I * J * n1);
    for (size_t i = 0; i < I; i++)
        for (size_t j = 0; j < J; j++)
            for (size_t k = 0; k < d; k++) {
                size_t flat = k * I * J + i * J + j;
                if (flat >= n_records) continue;
                uint64_t *cell = cells + (i * J + j) * n1;
                for (size_t c = 0; c < nr; c++) cell[d * c + k] = (cell[d * c + k] + records[flat * nr + c]) % pp->p;
            }
}

int vo_server_new(const vo_params *pp, const uint64_t *records, size_t n_records, const uint64_t *cascade_key,
                  const uint64_t *conv_key, const uint64_t *rsk, vo_server **out) {
    size_t I = pp->num_rows, J = pp->num_cols, n1 = pp->n1;
    if (n_records > I * J * (n1 / pp->n_rec)) return 1; /* answer.rs:503-508 */
    vo_server *s = (vo_server *)calloc(1, sizeof *s);
    s->pp = *pp;
    /* PreparedDb::from_encoded (prepared_db.rs:57-86): p -> q2 lift, NTT. */
    s->db_eval = vo_alloc(I * J * n1);
    vo_setup_db(pp, records, n_records, s->db_eval);
    vo_ring r2 = ring_q2(pp), r1 = ring_q1(pp), r3 = ring_q3n2(pp);
    for (size_t c = 0; c < I * J; c++) vo_to_eval(&r2, s->db_eval + c * n1);
    /* PreparedKeys::from_public_params (prepared_keys.rs:69-77). */
    s->cascade_eval = vo_alloc(vo_words_cascade_key(pp));
    cascade_key_to_eval(&r1, cascade_key, pp->l_ck, s->cascade_eval);
    s->cascade_coeff = vo_alloc(vo_words_cascade_key(pp));
    memcpy(s->cascade_coeff, cascade_key, 8 * vo_words_cascade_key(pp));
    s->conv_eval = vo_alloc(vo_words_conv_key(pp));
    vo_rlev_to_eval(&r1, conv_key, pp->l_ck, s->conv_eval);
    s->rsk_eval = vo_alloc(vo_words_rsk(pp));
    vo_rlev_to_eval(&r3, rsk, (pp->n1 / pp->n2) * pp->l_rsk, s->rsk_eval);
    *out = s;
    return 0;
}

void vo_server_free(vo_server *s) {
    if (!s) return;
    free(s->db_eval);
    free(s->cascade_eval);
    free(s->cascade_coeff);
    free(s->conv_eval);
    free(s->rsk_eval);
    free(s);
}

const uint64_t *vo_server_db(const vo_server *s) { return s->db_eval; }

void vo_server_step_times(const vo_server *s, double *t7) { memcpy(t7, s->t, sizeof s->t); }

/* Eval-form RGSW for a gate: every sample transformed once. */
static void rgsw_to_eval(const vo_ring *r, const uint64_t *rgsw, uint32_t L, uint64_t *out) {
    vo_rlev_to_eval(r, rgsw, 2 * L, out);
}

/* External product against an eval-form RGSW (rgsw.rs:108-120). */
static void ext_eval(const vo_ring *r, const uint64_t *rgsw_ev, uint32_t L, uint64_t base, const uint64_t *ct,
                     uint64_t *out) {
    size_t w = vo_pw(r);
    vo_gadget_product_eval(r, rgsw_ev, L, base, ct, out, out + w, 0);
    vo_gadget_product_eval(r, rgsw_ev + 2 * (size_t)L * w, L, base, ct + w, out, out + w, 1);
    vo_from_eval(r, out);
    vo_from_eval(r, out + w);
}

typedef struct {
    const vo_params *pp;
    const uint64_t *sel; /* I x (mask, body) eval at q2 */
    const uint64_t *db;
    uint64_t *out;
    size_t j0, j1;
} fd_job;

static void *fd_worker(void *arg) {
    fd_job *jb = (fd_job *)arg;
    const vo_params *pp = jb->pp;
    size_t I = pp->num_rows, J = pp->num_cols, n = pp->n1;
    vo_ring r2 = ring_q2(pp);
    uint64_t q = pp->q2;
    if (!vo_ntt_friendly(&r2)) {
        /* schoolbook backend: the "eval" form is the coefficient form */
        for (size_t j = jb->j0; j < jb->j1; j++) {
            uint64_t *om = jb->out + j * 2 * n, *ob = om + n;
            for (size_t i = 0; i < I; i++) {
                const uint64_t *c = jb->db + (i * J + j) * n, *sm = jb->sel + i * 2 * n;
                vo_eval_mul_acc(&r2, om, c, sm, i > 0);
                vo_eval_mul_acc(&r2, ob, c, sm + n, i > 0);
            }
        }
        return NULL;
    }
    for (size_t j = jb->j0; j < jb->j1; j++) {
        uint64_t *om = jb->out + j * 2 * n, *ob = om + n;
        u128 *am = (u128 *)calloc(n, sizeof(u128)), *ab = (u128 *)calloc(n, sizeof(u128));
        for (size_t i = 0; i < I; i++) {
            const uint64_t *c = jb->db + (i * J + j) * n, *sm = jb->sel + i * 2 * n, *sb = sm + n;
            for (size_t k = 0; k < n; k++) {
                am[k] += (u128)((u128)c[k] * sm[k] % q);
                ab[k] += (u128)((u128)c[k] * sb[k] % q);
            }
        }
        for (size_t k = 0; k < n; k++) {
            om[k] = (uint64_t)(am[k] % q);
            ob[k] = (uint64_t)(ab[k] % q);
        }
        vo_from_eval(&r2, om);
        vo_from_eval(&r2, ob);
        free(am);
        free(ab);
    }
    return NULL;
}

/* first_dim (first_dim.rs:94-184): sel_i = NTT(ct_i); out_j =
 * iNTT(sum_i db_ij (.) sel_i) for mask and body. Columns fan out over the
 * host's threads once I*J*N1 >= 2^16 (first_dim.rs:141-181). */
void vo_first_dim(const vo_params *pp, const uint64_t *switched, const uint64_t *db_eval, int threads,
                  uint64_t *out) {
    size_t I = pp->num_rows, J = pp->num_cols, n = pp->n1;
    vo_ring r2 = ring_q2(pp);
    uint64_t *sel = vo_alloc(I * 2 * n);
    memcpy(sel, switched, 8 * I * 2 * n);
    for (size_t i = 0; i < 2 * I; i++) vo_to_eval(&r2, sel + i * n);
    size_t workers = 1;
    if (I * J * n >= (1u << 16)) {
        long hc = threads > 0 ? threads : vo_host_threads();
        workers = hc < 1 ? 1 : (size_t)hc;
        if (workers > J) workers = J;
    }
    size_t chunk = (J + workers - 1) / workers;
    size_t nj = (J + chunk - 1) / chunk;
    fd_job *jobs = (fd_job *)calloc(nj, sizeof(fd_job));
    pthread_t *th = (pthread_t *)calloc(nj, sizeof(pthread_t));
    for (size_t t = 0; t < nj; t++) {
        jobs[t].pp = pp;
        jobs[t].sel = sel;
        jobs[t].db = db_eval;
        jobs[t].out = out;
        jobs[t].j0 = t * chunk;
        jobs[t].j1 = (t + 1) * chunk < J ? (t + 1) * chunk : J;
        if (nj > 1) pthread_create(&th[t], NULL, fd_worker, &jobs[t]);
        else fd_worker(&jobs[t]);
    }
    if (nj > 1)
        for (size_t t = 0; t < nj; t++) pthread_join(th[t], NULL);
    free(jobs);
    free(th);
    free(sel);
}

/* resp_comp (resp_comp.rs:46-76): q2 -> q3 at n1, ring switch to n2 with the
 * eval RSK, then the mask stays at q3 and the body goes to q4. */
static void resp_comp(const vo_server *s, const uint64_t *cr, uint64_t *ans_mask, uint64_t *ans_body) {
    const vo_params *pp = &s->pp;
    size_t n1 = pp->n1;
    vo_ring r2 = ring_q2(pp), r3n1 = ring_q3n1(pp), r3n2 = ring_q3n2(pp);
    uint64_t *c3 = vo_alloc(2 * n1);
    vo_mod_switch_sym(&r2, cr, &r3n1, c3);
    size_t n2 = pp->n2, D = n1 / n2;
    size_t kw = (size_t)pp->l_rsk * 2 * n2;
    uint64_t *out = vo_alloc(2 * n2), *rot = vo_alloc(n1), *proj = vo_alloc(n2), *acc = vo_alloc(2 * n2);
    vo_project_at(1, (uint32_t)n1, (uint32_t)n2, c3 + n1, 0, out + n2);
    for (size_t j = 0; j < D; j++) {
        vo_mul_x_pow(&r3n1, c3, j, rot);
        vo_project_at(1, (uint32_t)n1, (uint32_t)n2, rot, 0, proj);
        vo_gadget_product_eval(&r3n2, s->rsk_eval + j * kw, pp->l_rsk, pp->b_rsk, proj, acc, acc + n2, 0);
        vo_from_eval(&r3n2, acc);
        vo_from_eval(&r3n2, acc + n2);
        vo_poly_sub(&r3n2, out, out, acc);
        vo_poly_sub(&r3n2, out + n2, out + n2, acc + n2);
    }
    vo_mod_switch_asym(&r3n2, out, pp->q3, pp->q4, ans_mask, ans_body);
    free(c3);
    free(out);
    free(rot);
    free(proj);
    free(acc);
}

/* answer_through_crot (answer.rs:83-228) + answer_one_query (answer.rs:289-354). */
/* Threads for the query decomposition's per-LWE work. 1 (the default) is the
 * reference's serial loop, which the CPU baseline times; tests raise it to
 * check GPU answers sooner (the values do not depend on it). */
static int g_decomp_threads = 1;
void vo_set_decomp_threads(int n) { g_decomp_threads = n > 0 ? n : 1; }

/* Level i of bit b: the cascade of its LWE (cascade.rs:388-405) into the m
 * half of the RGSW, then rlwe_to_rgsw_eval (convert.rs:162-180) of that
 * level into the neg_s_m half; gp holds one ciphertext of scratch. */
static void decomp_level(const vo_server *s, const vo_ring *r1, const uint64_t *lwes, size_t lw, uint64_t *rgsw,
                         size_t gw1, uint32_t LQ, uint32_t b, uint32_t i, uint64_t *gp) {
    const vo_params *pp = &s->pp;
    size_t w1 = vo_pw(r1), cw1 = 2 * w1;
    uint64_t *g = rgsw + b * gw1;
    uint64_t *levels = g + 2 * (size_t)LQ * w1; /* the m half holds the RLWE levels */
    cascade_eval(r1, lwes + ((size_t)b * LQ + i) * lw, s->cascade_eval, pp->ck_base, pp->l_ck, levels + i * cw1);
    vo_gadget_product_eval(r1, s->conv_eval, pp->l_ck, pp->ck_base, levels + i * cw1, gp, gp + w1, 0);
    vo_from_eval(r1, gp);
    vo_from_eval(r1, gp + w1);
    vo_poly_add(r1, g + i * cw1, levels + i * cw1 + w1, gp);
    memcpy(g + i * cw1 + w1, gp + w1, 8 * w1);
}

typedef struct {
    const vo_server *s;
    const vo_ring *r1;
    const uint64_t *lwes;
    size_t lw;
    uint64_t *rgsw;
    size_t gw1;
    uint32_t LQ;
    size_t total, next;
} decomp_job;

static void *decomp_worker(void *arg) {
    decomp_job *j = (decomp_job *)arg;
    uint64_t *gp = vo_alloc(2 * vo_pw(j->r1));
    for (;;) {
        size_t k = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
        if (k >= j->total) break;
        decomp_level(j->s, j->r1, j->lwes, j->lw, j->rgsw, j->gw1, j->LQ, (uint32_t)(k / j->LQ),
                     (uint32_t)(k % j->LQ), gp);
    }
    free(gp);
    return NULL;
}

static int answer_stages(const vo_server *s, const uint64_t *lwes, size_t n_lwe, uint64_t *rgsw_bits,
                         uint64_t *dmux_out, uint64_t *switched, uint64_t *fd_out, uint64_t *cmux_out,
                         uint64_t *crot_out, uint64_t *ans_mask, uint64_t *ans_body, double *times) {
    const vo_params *pp = &s->pp;
    size_t I = pp->num_rows, J = pp->num_cols, n1 = pp->n1;
    uint32_t LQ = pp->l_query;
    vo_ring r1 = ring_q1(pp), r2 = ring_q2(pp);
    size_t w1 = vo_pw(&r1), w2 = n1, cw1 = 2 * w1, cw2 = 2 * w2, gw1 = 4 * (size_t)LQ * w1, gw2 = 4 * (size_t)LQ * w2;
    uint32_t nd = ilog2((uint32_t)I), nc = ilog2((uint32_t)J), nr = ilog2(pp->n1 / pp->n_rec);
    uint32_t bits = nd + nc + nr;
    double t0 = now_s();

    /* step 1: query_decomp (query_decomp.rs:39-95). */
    uint64_t *rgsw = rgsw_bits ? rgsw_bits : vo_alloc(bits * gw1);
    size_t lw = vo_words_lwe(pp);
    if (g_decomp_threads > 1) {
        /* test-only fan-out of the independent per-LWE work (same values) */
        decomp_job job = {s, &r1, lwes, lw, rgsw, gw1, LQ, (size_t)bits * LQ, 0};
        size_t nth = (size_t)g_decomp_threads < job.total ? (size_t)g_decomp_threads : job.total;
        pthread_t *th = (pthread_t *)calloc(nth, sizeof *th);
        for (size_t k = 0; k < nth; k++) pthread_create(&th[k], NULL, decomp_worker, &job);
        for (size_t k = 0; k < nth; k++) pthread_join(th[k], NULL);
        free(th);
    } else {
        uint64_t *gp = vo_alloc(cw1);
        for (uint32_t b = 0; b < bits; b++)
            for (uint32_t i = 0; i < LQ; i++) decomp_level(s, &r1, lwes, lw, rgsw, gw1, LQ, b, i, gp);
        free(gp);
    }
    double t1 = now_s();

    /* step 2: dmux_tree from trivial(Delta) (answer.rs:181-191). */
    uint64_t *dm = dmux_out ? dmux_out : vo_alloc(I * cw1);
    {
        uint64_t *triv = vo_alloc(cw1);
        u128 Q1 = vo_big_q(&r1), delta = (Q1 + pp->p - 1) / pp->p;
        vo_set_u128(&r1, triv + w1, 0, delta);
        uint64_t *ev = vo_alloc(gw1), *src = vo_alloc(cw1), *p = vo_alloc(cw1);
        memcpy(dm, triv, 8 * cw1);
        for (uint32_t i = 0; i < nd; i++) {
            rgsw_to_eval(&r1, rgsw + i * gw1, LQ, ev);
            size_t active = (size_t)1 << i;
            for (size_t jj = active; jj-- > 0;) {
                memcpy(src, dm + jj * cw1, 8 * cw1);
                ext_eval(&r1, ev, LQ, pp->b1, src, p);
                memcpy(dm + (2 * jj + 1) * cw1, p, 8 * cw1);
                vo_poly_sub(&r1, dm + 2 * jj * cw1, src, p);
                vo_poly_sub(&r1, dm + 2 * jj * cw1 + w1, src + w1, p + w1);
            }
        }
        free(triv);
        free(ev);
        free(src);
        free(p);
    }
    double t2 = now_s();

    /* step 3: mod_switch_sym q1 -> q2 per row. */
    uint64_t *sw = switched ? switched : vo_alloc(I * cw2);
    for (size_t i = 0; i < I; i++) vo_mod_switch_sym(&r1, dm + i * cw1, &r2, sw + i * cw2);
    double t3 = now_s();

    /* step 4: first_dim. */
    uint64_t *fd = fd_out ? fd_out : vo_alloc(J * cw2);
    vo_first_dim(pp, sw, s->db_eval, 0, fd);
    double t4 = now_s();

    /* step 5: CMux tree at q2 over mod-switched RGSWs. */
    uint64_t *cm = cmux_out ? cmux_out : vo_alloc(cw2);
    uint64_t *g2 = vo_alloc(gw2), *ev2 = vo_alloc(gw2), *d2 = vo_alloc(cw2), *e2 = vo_alloc(cw2);
    {
        uint64_t *inp = vo_alloc(J * cw2);
        memcpy(inp, fd, 8 * J * cw2);
        size_t active = J;
        for (uint32_t b = 0; b < nc; b++) {
            vo_mod_switch_rgsw(&r1, rgsw + (nd + b) * gw1, LQ, LQ, &r2, g2);
            rgsw_to_eval(&r2, g2, LQ, ev2);
            size_t half = active >> 1;
            for (size_t j = 0; j < half; j++) {
                uint64_t *a = inp + 2 * j * cw2, *bb = inp + (2 * j + 1) * cw2;
                vo_poly_sub(&r2, d2, bb, a);
                vo_poly_sub(&r2, d2 + w2, bb + w2, a + w2);
                ext_eval(&r2, ev2, LQ, pp->b2, d2, e2);
                vo_poly_add(&r2, inp + j * cw2, a, e2);
                vo_poly_add(&r2, inp + j * cw2 + w2, a + w2, e2 + w2);
            }
            active = half;
        }
        memcpy(cm, inp, 8 * cw2);
        free(inp);
    }
    double t5 = now_s();

    /* step 6: CRot SlotExtract (rotate.rs:128-149). */
    uint64_t *cr = crot_out ? crot_out : vo_alloc(cw2);
    {
        uint64_t *res = vo_alloc(cw2), *rot = vo_alloc(cw2);
        memcpy(res, cm, 8 * cw2);
        for (uint32_t b = 0; b < nr; b++) {
            vo_mod_switch_rgsw(&r1, rgsw + (nd + nc + b) * gw1, LQ, LQ, &r2, g2);
            rgsw_to_eval(&r2, g2, LQ, ev2);
            vo_rotate(&r2, res, n1 - ((size_t)1 << b), rot);
            vo_poly_neg(&r2, rot, rot);
            vo_poly_neg(&r2, rot + w2, rot + w2);
            vo_poly_sub(&r2, d2, rot, res);
            vo_poly_sub(&r2, d2 + w2, rot + w2, res + w2);
            ext_eval(&r2, ev2, LQ, pp->b2, d2, e2);
            vo_poly_add(&r2, res, res, e2);
            vo_poly_add(&r2, res + w2, res + w2, e2 + w2);
        }
        memcpy(cr, res, 8 * cw2);
        free(res);
        free(rot);
    }
    free(g2);
    free(ev2);
    free(d2);
    free(e2);
    double t6 = now_s();

    /* step 7: resp_comp (resp_comp.rs:46-76), skipped when only the CRot
     * output is wanted (the VIA-B prefix). */
    if (ans_mask) resp_comp(s, cr, ans_mask, ans_body);
    double t7 = now_s();
    times[0] = t1 - t0;
    times[1] = t2 - t1;
    times[2] = t3 - t2;
    times[3] = t4 - t3;
    times[4] = t5 - t4;
    times[5] = t6 - t5;
    times[6] = t7 - t6;
    if (!rgsw_bits) free(rgsw);
    if (!dmux_out) free(dm);
    if (!switched) free(sw);
    if (!fd_out) free(fd);
    if (!cmux_out) free(cm);
    if (!crot_out) free(cr);
    (void)n_lwe;
    return 0;
}

int vo_server_answer_stages(const vo_server *cs, const uint64_t *lwes, size_t n_lwe, uint64_t *rgsw_bits,
                            uint64_t *dmux_out, uint64_t *switched, uint64_t *fd_out, uint64_t *cmux_out,
                            uint64_t *crot_out, uint64_t *ans_mask, uint64_t *ans_body) {
    vo_server *s = (vo_server *)cs;
    return answer_stages(s, lwes, n_lwe, rgsw_bits, dmux_out, switched, fd_out, cmux_out, crot_out, ans_mask,
                         ans_body, s->t);
}

int vo_server_answer(const vo_server *s, const uint64_t *lwes, size_t n_lwe, uint64_t *ans_mask, uint64_t *ans_body,
                     size_t *expected, size_t *got) {
    const vo_params *pp = &s->pp;
    /* guards of answer_through_crot (answer.rs:131-157) */
    if (!pp->num_rows || !pp->num_cols || (pp->num_rows & (pp->num_rows - 1)) ||
        (pp->num_cols & (pp->num_cols - 1)))
        return 1;
    size_t exp = vo_query_lwes(pp);
    if (n_lwe != exp) {
        if (expected) *expected = exp;
        if (got) *got = n_lwe;
        return 2;
    }
    return vo_server_answer_stages(s, lwes, n_lwe, NULL, NULL, NULL, NULL, NULL, NULL, ans_mask, ans_body);
}

typedef struct {
    const vo_server *s;
    const uint64_t *lwes;
    size_t n_lwe, lw;
    uint64_t *rot;
    uint32_t T;
    size_t k, stride;
} prefix_job;

static void *prefix_worker(void *arg) {
    prefix_job *j = (prefix_job *)arg;
    size_t n1 = j->s->pp.n1;
    double times[7];
    for (size_t t = j->k; t < j->T; t += j->stride)
        answer_stages(j->s, j->lwes + t * j->n_lwe * j->lw, j->n_lwe, NULL, NULL, NULL, NULL, NULL,
                      j->rot + t * 2 * n1, NULL, NULL, times);
    return NULL;
}

/* Server::answer_batch::<T> (answer.rs:561-602) -> answer_batch (batch.rs:72-163):
 * answer_through_crot for each of the T queries (record degree n_rec), the
 * repack with the cascade-key suffix mod-switched to q2 (scheme.rs:244-266;
 * the extraction degree is G = n2/T, the presets' K/T), then one resp_comp.
 * lwes holds the T queries back to back, n_lwe LWEs each. Returns 0,
 * 1 = DimMismatch (T unsupported or mismatched) or 2 = QueryLengthMismatch. */
int vo_server_answer_batch(const vo_server *s, uint32_t T, const uint64_t *lwes, size_t n_lwe, uint64_t *ans_mask,
                           uint64_t *ans_body, size_t *expected, size_t *got) {
    const vo_params *pp = &s->pp;
    if (!pp->num_rows || !pp->num_cols || (pp->num_rows & (pp->num_rows - 1)) ||
        (pp->num_cols & (pp->num_cols - 1)))
        return 1;
    if (!T || (T & (T - 1)) || pp->n2 % T || (size_t)T * pp->n_rec > pp->n2) return 1;
    size_t exp = vo_query_lwes(pp);
    if (n_lwe != exp) {
        if (expected) *expected = exp;
        if (got) *got = n_lwe;
        return 2;
    }
    size_t n1 = pp->n1, lw = vo_words_lwe(pp);
    uint32_t G = pp->n2 / T;
    uint64_t *rot = vo_alloc((size_t)T * 2 * n1);
    /* the T prefixes are independent: one thread each, at most one per core */
    long ncpu = vo_host_threads();
    size_t nth = ncpu > 0 ? (size_t)ncpu : 1;
    if (nth > T) nth = T;
    prefix_job *jobs = (prefix_job *)calloc(nth, sizeof *jobs);
    pthread_t *th = (pthread_t *)calloc(nth, sizeof *th);
    for (size_t k = 0; k < nth; k++) {
        jobs[k] = (prefix_job){s, lwes, n_lwe, lw, rot, T, k, nth};
        pthread_create(&th[k], NULL, prefix_worker, &jobs[k]);
    }
    for (size_t k = 0; k < nth; k++) pthread_join(th[k], NULL);
    free(jobs);
    free(th);
    vo_ring r1 = ring_q1(pp), r2 = ring_q2(pp);
    uint64_t *key = vo_alloc(vo_repack_key_words((uint32_t)n1, G, pp->l_ck));
    vo_repack_key_from_cascade(&r1, s->cascade_coeff, pp->q2, G, pp->l_ck, key);
    uint64_t *packed = vo_alloc(2 * n1);
    int rc = vo_repack(&r2, T, G, rot, key, pp->ck_base, pp->l_ck, packed);
    if (!rc) resp_comp(s, packed, ans_mask, ans_body);
    free(rot);
    free(key);
    free(packed);
    return rc ? 1 : 0;
}
