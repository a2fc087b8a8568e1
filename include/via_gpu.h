/*
 * via_gpu.h — C ABI of the MI355X (gfx950) VIA-C server answer path.
 *
 * This is the drop-in boundary for the via-rs server crate: each entry point
 * replaces one reference interface (cited per declaration, paths relative to
 * the via-rs checkout). All buffers are dense little-endian uint64 canonical
 * residues with the layouts below; no Rust or torch types cross the boundary.
 *
 * Layouts
 *   poly   : nl x n residues, limb-major (coefficient i of limb l at l*n + i);
 *            q1 is the two-prime RNS (nl = 2), every other ring has nl = 1.
 *   RLWE   : mask poly then body poly.
 *   RLev L : L RLWE samples; sample k encrypts m * round(Q / B^(k+1)).
 *   LWE    : per limb, n masks then the body: lwe[l*(n+1) + i].
 *   record : n_rec coefficients in [0, p) (setup_db.rs:39-75).
 *   cascade key : steps in ascending output degree n_out = 2, 4, ..., n1; step
 *            s holds rank_in = 2*n1/n_out RLev(L_CK) at (n_out, q1)
 *            (cascade.rs:90-110, conv.rs:161-190).
 *   conv key : RLev(L_CK) of S1^2 at (n1, q1) (convert.rs:46-58).
 *   rsk    : d = n1/n2 RLev(L_RSK) at (n2, q3) (ring_switch.rs:224-246).
 *   answer : mask[n2] mod q3, body[n2] mod q4 (ModSwitchedCiphertext).
 */
#ifndef VIA_GPU_H
#define VIA_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ViaError (via-protocol/src/error.rs:20-64) plus the device-side failures. */
typedef enum {
    VIA_OK = 0,
    VIA_E_DIM = 1,         /* ViaError::DimMismatch */
    VIA_E_QUERY_LEN = 2,   /* ViaError::QueryLengthMismatch { expected, got } */
    VIA_E_INDEX = 3,       /* ViaError::IndexOutOfRange */
    VIA_E_NOMEM = 4,       /* HBM admission refused / allocation failed */
    VIA_E_INVALID = 5,     /* parameters outside what the device path supports */
    VIA_E_DEVICE = 6,      /* HIP runtime error */
    VIA_E_BATCH_LEN = 7,   /* answer_batch: wrong number of queries (batch.rs:124-128) */
    VIA_E_BUFFER_SHORT = 8, /* ViaError::BufferTooShort { needed, got } (err: expected, got) */
    VIA_E_FORMAT_TAG = 9,  /* ViaError::UnknownFormatTag(tag) (err: got = tag) */
    VIA_E_PRG_SEED = 10    /* ViaError::InvalidPrgSeed: a seed longer than 64 bytes */
} via_status;

/* PIRParams (via-protocol/src/params.rs) restricted to the server path, plus
 * the conversion-key base that PublicParams carries (wire/keys.rs:136-141). */
typedef struct {
    uint32_t n1, n2, n_rec;
    uint32_t q1_nl;
    uint64_t q1[2];
    uint64_t q2, q3, q4, p;
    uint64_t b1, b2, b_rsk, ck_base;
    uint32_t l_query, l_ck, l_rsk;
    uint32_t num_rows, num_cols;
} via_params;

typedef struct {
    uint64_t expected, got;
    char msg[128];
} via_err_detail;

typedef struct via_ctx via_ctx;
typedef struct via_db via_db;
typedef struct via_keys via_keys;

/* Device context: one HIP stream and the NTT tables of every ring in use.
 *
 * Threading (Server::answer(&self) is re-entrant, answer.rs:534-554): every
 * entry point may be called from several threads at once on one context and
 * shared databases and key sets. Host-buffer entry points (via_answer,
 * via_answer_many, via_answer_batch, via_first_dim, via_ntt) run on the
 * context when it is free and otherwise on one of up to seven lanes the
 * context creates on demand (own stream and workspace), so concurrent calls
 * overlap on the device. Device-buffer entry points run in call order on the
 * context's own stream, whose work the caller synchronises with
 * via_ctx_synchronize. A database or key set must outlive the calls that use
 * it and must not be destroyed while a call uses it.
 *
 * Inputs are canonical residues (the reference's Zq types hold nothing
 * else): a query, key or record word at or above its modulus is refused with
 * VIA_E_INVALID, by the host-buffer calls before they return and, for the
 * device-buffer answer calls, by the next via_ctx_synchronize (the word is
 * taken as 0 meanwhile, so no out-of-range value enters the arithmetic). On
 * any error the output buffers hold unspecified values. */
via_status via_ctx_create(int device, via_ctx **out);
void via_ctx_destroy(via_ctx *ctx);
/* Seconds spent in each of the seven answer steps of the last via_answer on
 * this context (answer.rs:166-343 span names step1..step7), by HIP events. */
via_status via_ctx_step_times(via_ctx *ctx, double *t7);
/* Seconds of the last FirstDim database-stream kernel alone (HIP events on
 * the context stream, around the one launch that reads the I*J cells for a
 * pass of up to 32 queries; for a VIA-B batch of more than 32 queries, its
 * last pass). */
via_status via_ctx_first_dim_time(via_ctx *ctx, double *seconds);
/* hipGraph mode (no counterpart in the reference), off by default. When on,
 * via_answer_device / via_answer_many_device run a call eagerly the first
 * time its database, batch size, key sets, query and answer buffers and query
 * length occur (and after the workspace grows or the autotuner stores new
 * shapes), capture the same launch sequence, and replay the capture for every
 * later identical call (up to eight captures per context). Step times are not
 * recorded by replays. Turning it off drops the captures. */
via_status via_ctx_set_graphs(via_ctx *ctx, int on);

/* Step timing (no counterpart in the reference): on (the default), every
 * answer records the HIP events behind via_ctx_step_times; off, only the
 * pair around the FirstDim pass (via_ctx_first_dim_time) is recorded, which
 * saves the few microseconds of idle device time each event record costs
 * between two kernels (about 40 us of a one-query answer). The autotuner
 * records them while it measures either way. */
via_status via_ctx_set_step_timing(via_ctx *ctx, int on);
/* Stream timer: which = 0 records the start event, 1 the stop event, on the
 * context stream; via_ctx_marked_time waits for the stop event and returns
 * the seconds between them. No counterpart in the reference (its benches
 * time with criterion on the host). */
via_status via_ctx_mark(via_ctx *ctx, int which);
via_status via_ctx_marked_time(via_ctx *ctx, double *seconds);

/* Server::setup, database half: setup_db + PreparedDb::from_encoded
 * (answer.rs:486-528, setup_db.rs:39-75, prepared_db.rs:57-86). Records are
 * [n_records][n_rec] coefficients; more records than I*J*(n1/n_rec) is
 * VIA_E_DIM. The prepared database stays resident in HBM. */
via_status via_db_create(via_ctx *ctx, const via_params *pp, const uint64_t *records, size_t n_records,
                         via_db **out);
/* Same, with records as packed nibbles: coefficient 2k = byte k & 0xF,
 * 2k+1 = byte k >> 4 (p = 16, n_rec/2 bytes per record). */
via_status via_db_create_bytes(via_ctx *ctx, const via_params *pp, const uint8_t *bytes, size_t n_records,
                               via_db **out);
void via_db_destroy(via_db *db);
/* Incremental record updates (no counterpart in the reference, whose
 * PreparedDb is rebuilt whole): replaces records indices[u] (record m is slot
 * m / (I*J) of cell m mod I*J, setup_db.rs:58-72) with records[u]
 * ([n][n_rec] coefficients, reduced mod p) in the HBM-resident database;
 * every touched cell is rebuilt on the device (inverse transform, slot
 * rewrite, transform, repack). A repeated index takes its last record;
 * VIA_E_INDEX for an index beyond I*J*(n1/n_rec). Answers afterwards are
 * those of a database set up with the updated records. */
via_status via_db_update(via_db *db, const uint64_t *indices, size_t n, const uint64_t *records);
/* The prepared cells, eval form at q2, [I][J][n1] (PreparedDb::cell). */
via_status via_db_read_cells(const via_db *db, uint64_t *out, size_t words);

/* PreparedKeys::from_public_params (prepared_keys.rs:69-77): cascade key,
 * conversion key and ring-switch key to eval form, resident in HBM. With
 * pp->n_rec < n2 the cascade steps of output degree >= 2 n_rec are also kept
 * mod-switched to q2 for VIA-B batches (the keys of repack.rs:684-805). The
 * keys serve any database whose record degree is >= pp->n_rec. */
via_status via_keys_create(via_ctx *ctx, const via_params *pp, const uint64_t *cascade_key,
                           const uint64_t *conv_key, const uint64_t *rsk, via_keys **out);
/* Zeroizes the device copies before release (prepared_keys.rs:89-125). */
void via_keys_destroy(via_keys *keys);

/* Server::answer (answer.rs:534-554) -> answer_one_query (answer.rs:289-354):
 * query_decomp, dmux_tree, mod switch, first_dim, cmux_tree, crot, resp_comp.
 * lwes = n_lwe LWEs at (n1, q1). Host buffers. */
via_status via_answer(via_ctx *ctx, const via_db *db, const via_keys *keys, const uint64_t *lwes, size_t n_lwe,
                      uint64_t *ans_mask, uint64_t *ans_body, via_err_detail *err);
/* Same with device-resident query and answer buffers; enqueued on the
 * context stream, returns without waiting (a non-canonical query word is
 * reported by the next via_ctx_synchronize as VIA_E_INVALID). */
via_status via_answer_device(via_ctx *ctx, const via_db *db, const via_keys *keys, const uint64_t *d_lwes,
                             size_t n_lwe, uint64_t *d_ans_mask, uint64_t *d_ans_body, via_err_detail *err);
/* Waits for the context stream; VIA_E_INVALID (once) when a device query
 * since the last call held a non-canonical word. */
via_status via_ctx_synchronize(via_ctx *ctx);

/* No counterpart in the reference: via_answer_device for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The whole workspace lives in a 96 MiB
 * array the library holds on each GPU it runs on (from the first context on
 * that device until the process ends), enough for one query against a grid
 * of up to I = J = 256 at the paper parameters; a larger workspace is
 * VIA_E_NOMEM (err: expected = bytes needed, got = bytes available), a
 * one-row grid VIA_E_INVALID, and parameters whose twiddle tables this
 * context has not built (via_db_create / via_keys_create on it build them)
 * VIA_E_INVALID. Nothing is allocated. Every call on one GPU shares the
 * array: calls whose device work overlaps (other contexts or streams on the
 * same GPU) race, so serialise them. Same query, answer and error contract
 * as via_answer_device. */
via_status via_answer_device_noalloc(via_ctx *ctx, const via_db *db, const via_keys *keys, const uint64_t *d_lwes,
                                     size_t n_lwe, uint64_t *d_mask, uint64_t *d_body, via_err_detail *err);

/* Server::answer for nq queries against one database, each with the keys of
 * its own client (keys[q]); every step runs batched across the queries, the
 * database is streamed once per group of up to 32 queries. Each answer is
 * bit-identical to answering the query alone. Host buffers: lwes[q] holds
 * n_lwe LWEs, masks[q] / bodies[q] receive n2 words. No counterpart in the
 * reference, which answers one query per call (answer.rs:534-554). */
via_status via_answer_many(via_ctx *ctx, const via_db *db, uint32_t nq, const via_keys *const *keys,
                            const uint64_t *const *lwes, size_t n_lwe, uint64_t *const *masks, uint64_t *const *bodies,
                            via_err_detail *err);
/* Same with device-resident buffers, enqueued on the context stream. */
via_status via_answer_many_device(via_ctx *ctx, const via_db *db, uint32_t nq, const via_keys *const *keys,
                                   const uint64_t *const *d_lwes, size_t n_lwe, uint64_t *const *d_masks,
                                   uint64_t *const *d_bodies, via_err_detail *err);

/* Autotuner (no counterpart in the reference): answers the nq device-resident
 * queries with each candidate launch shape of every query-decompression step
 * (the number of digit groups, i.e. workgroups splitting the 18 digits of one
 * gadget product, and the kernel variant: 256 threads with one digit
 * transform at a time in radix-8 or radix-16 passes, 512 threads in radix-8
 * passes, 1024 threads with several side by side, or the 256- and 512-thread
 * radix-8 kernels with one workgroup per RNS limb), reps timed
 * runs per candidate, and keeps the fastest per step for batches of nq
 * queries. groups_out and variant_out (12 entries each: the eleven
 * cascade steps and the RLWE -> RGSW conversion; variant 0 = 256 threads
 * radix 8, 2 = 256 threads radix 16, 3 = 512 threads radix 8, 1 = 1024
 * threads, 4 = 256 threads per limb, 5 = 512 threads per limb; for the
 * cascade steps, which take the fixed-shape kernel when its exactness bound
 * holds for the parameters, the variant mod 3 selects 1, 2 or 4 LWEs per
 * workgroup instead) receive the choice when not null. With those shapes it then measures, in turn, the
 * DMux and CMux trees as one persistent launch each against level by level
 * (via_ctx_tree_tuning), FirstDim's column-block groups per slot (1, 2, 4,
 * 8) and the gates' limb split (via_ctx_pass_tuning), keeping the faster of
 * each. The choices serve every context of the process on the same device
 * (the serving loop's included, whose trees always launch level by level);
 * an untuned batch size takes the choices of the largest tuned size below
 * it. The answers written meanwhile are those of via_answer_many_device. */
via_status via_ctx_autotune(via_ctx *ctx, const via_db *db, uint32_t nq, const via_keys *const *keys,
                            const uint64_t *const *d_lwes, size_t n_lwe, uint64_t *const *d_masks,
                            uint64_t *const *d_bodies, int reps, int *groups_out, int *variant_out);
/* Tree launch mode (no counterpart in the reference, whose trees are loops,
 * mux.rs:167-251): mode 1 runs every level of the DMux tree, and of the CMux
 * tree, in one cooperative launch whose workgroups wait for each level at a
 * bounded grid barrier (k_tree_persistent); mode 0 launches level by level;
 * mode -1 (the default) takes the choice via_ctx_autotune stored for the
 * batch size, or level by level when none is stored. A barrier wait that
 * gives up is reported by the call (host buffers) or the next
 * via_ctx_synchronize as VIA_E_DEVICE. Answers are the same in every mode. */
via_status via_ctx_set_tree_mode(via_ctx *ctx, int mode);
/* The tree modes via_ctx_autotune chose for batches of nq queries (1 =
 * persistent) for the DMux and the CMux tree; VIA_E_INVALID when none. */
via_status via_ctx_tree_tuning(via_ctx *ctx, uint32_t nq, int *dmux, int *cmux);
/* The FirstDim column-block groups per slot (workgroups sharing one slot's
 * database block) and the gate limb split (-1: split two-limb tree levels
 * narrower than the CU count, 0: never, 1: always) that via_ctx_autotune
 * measured and chose for batches of nq queries; VIA_E_INVALID when none. */
via_status via_ctx_pass_tuning(via_ctx *ctx, uint32_t nq, int *fd_groups, int *gate_split);
/* Test switch: polls a persistent tree's barrier wait makes before it gives
 * up (default 2^24; 0 gives up at the first wait that finds the level
 * unfinished). */
via_status via_debug_set_tree_spin(via_ctx *ctx, long spin);
/* No counterpart in the reference: stores launch shapes for batches of nq
 * queries as via_ctx_autotune would (12 entries each, the same encoding as
 * its outputs), e.g. a choice saved from an earlier tuning run; groups null
 * erases the entry. VIA_E_INVALID for a shape the tuner never offers. */
via_status via_ctx_set_tuning(via_ctx *ctx, uint32_t nq, const int *groups, const int *variants);

/* Server::answer_batch::<T> (answer.rs:561-602) -> answer_batch
 * (batch.rs:72-163), VIA-B: answer_through_crot for each of the T queries of
 * one client at the database's record degree n_rec, Repack_{n2} of the T
 * outputs (repack.rs:331-369, extraction degree n2/T) with the client's
 * cascade keys mod-switched to q2, then one resp_comp. lwes[t] holds
 * n_lwe[t] LWEs. Errors: VIA_E_DIM if n_queries != T (batch.rs:124-128),
 * VIA_E_QUERY_LEN for the first query of the wrong length, VIA_E_INVALID
 * for a T the record degree does not fit (T a power of two, T * n_rec <= n2,
 * n2/T >= 2, T <= 256; the reference fixes T at compile time) or keys
 * created with a record degree above the database's. Host buffers. */
via_status via_answer_batch(via_ctx *ctx, const via_db *db, const via_keys *keys, uint32_t T,
                            const uint64_t *const *lwes, const size_t *n_lwe, size_t n_queries, uint64_t *ans_mask,
                            uint64_t *ans_body, via_err_detail *err);
/* Same with the T queries back to back in one device buffer (n_lwe LWEs
 * each) and device answer buffers, enqueued on the context stream. */
via_status via_answer_batch_device(via_ctx *ctx, const via_db *db, const via_keys *keys, uint32_t T,
                                   const uint64_t *d_lwes, size_t n_lwe, uint64_t *d_ans_mask, uint64_t *d_ans_body,
                                   via_err_detail *err);

/* Intermediate values of the last answer for parity tests, coefficient form:
 * stage 1 RGSW bits [bit][half][i][part][limb][n1] at q1, 2 DMux outputs
 * [I][part][limb][n1] at q1, 3 mod-switched rows [I][part][n1] at q2,
 * 4 FirstDim outputs [J][part][n1] at q2, 5 CMux output, 6 CRot output,
 * 7 the repacked RLWE [part][n1] at q2 of the last VIA-B answer. Stages 1-6
 * of a VIA-B answer are those of its last group of queries. */
via_status via_debug_stage(via_ctx *ctx, int stage, uint64_t *out, size_t words);

/* Test switch with no counterpart in the reference: on != 0 makes every
 * later launch of the process take the exact 128/192-bit paths of the gadget
 * scale and of the q1 -> q2 mod switch instead of their fp64 fast paths (the
 * paths the fast paths hand the values they decline to). Answers are the same
 * either way, only slower; graphs captured before the switch keep the paths
 * they were captured with. */
via_status via_debug_set_exact_paths(int on);

/* RingPolyEval batch (ntt_abstraction.rs:61-78): c = a * b in
 * Z_q[X]/(X^n+1) for count polys; poly p uses modulus moduli[(p / group) % nmod].
 * Device pointers, enqueued on the context stream. */
via_status via_poly_mul_device(via_ctx *ctx, uint32_t n, const uint64_t *moduli, uint32_t nmod, size_t group,
                               const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_c, size_t count);
/* Forward / inverse negacyclic NTT of count polys in place (bit-reversed
 * eval order, ring/ntt.rs:305-399), host buffers. */
via_status via_ntt(via_ctx *ctx, uint32_t n, const uint64_t *moduli, uint32_t nmod, size_t group, uint64_t *polys,
                   size_t count, int inverse);

/* FirstDim (via-server/src/first_dim.rs:94-184, `first_dim`) for nq <= 32
 * queries in one database pass, the batched GPU form the reference's
 * first_dim.rs:81-88 describes: switched [q][I][part][n1] holds each query's
 * I mod-switched RLWE selection ciphertexts, out [q][J][part][n1] receives
 * the J output ciphertexts, both coefficient form at q2, host buffers. The
 * database must belong to ctx. */
via_status via_first_dim(via_ctx *ctx, const via_db *db, uint32_t nq, const uint64_t *switched, uint64_t *out);

/* ---- wire formats (via-protocol wire/format.rs) ----------------------------- */

/* Uncompressed (format.rs:93-132): [0x00][count u64 LE][count coefficients
 * u64 LE]. via_wire_uncompressed_size gives the packet length; encode
 * writes it into out (VIA_E_INVALID when out_len is smaller). decode
 * (WireFormat::deserialize) returns VIA_E_BUFFER_SHORT {needed, got} and
 * VIA_E_FORMAT_TAG {tag} as the reference's ViaError does; with out null it
 * only reports *count. An answer travels as the Uncompressed mask followed by
 * the Uncompressed body. */
size_t via_wire_uncompressed_size(size_t count);
via_status via_wire_encode_uncompressed(const uint64_t *coeffs, size_t count, uint8_t *out, size_t out_len);
via_status via_wire_decode_uncompressed(const uint8_t *buf, size_t len, uint64_t *out, size_t out_cap,
                                        size_t *count, via_err_detail *err);
/* PrgCompressed<32>::deserialize (format.rs:163-190): [0x01][seed, 32
 * bytes][count u64 LE][count body coefficients]; the seed goes to seed_out
 * (32 bytes) when not null, the body coefficients as in the Uncompressed
 * decode. */
via_status via_wire_decode_prg(const uint8_t *buf, size_t len, uint8_t *seed_out, uint64_t *out, size_t out_cap,
                               size_t *count, via_err_detail *err);
/* PrgCompressed::regenerate_masks (format.rs:143-160) on the GPU: count draws
 * Shake256Prg(seed).uniform_below(q) (prg.rs) into out (host buffer).
 * VIA_E_PRG_SEED for a seed longer than 64 bytes (the reference asserts). */
via_status via_prg_regenerate(via_ctx *ctx, const uint8_t *seed, size_t seed_len, uint64_t q, size_t count,
                              uint64_t *out);
/* A query whose LWE masks travel as a PRG seed (the regenerate_masks
 * contract: masks row-major, one draw per q1 limb, limb 0 first): writes the
 * device LWE array d_lwes [n_lwe][limb][n1 + 1] with the masks regenerated on
 * the GPU and the bodies [n_lwe][limb] (host) verbatim, ready for
 * via_answer_device. No counterpart beyond regenerate_masks: it saves the
 * 2 x n1 x 8 bytes per LWE of mask transfer. */
via_status via_query_from_prg(via_ctx *ctx, const via_params *pp, const uint8_t *seed, size_t seed_len,
                              const uint64_t *bodies, size_t n_lwe, uint64_t *d_lwes);

/* ---- client key generation and queries (via-client) -------------------------- */

/* Distribution (sampling/distribution.rs:38-60): kind 0 Ternary, kind 1
 * BoundedUniform { bound } ([-bound, bound]); the Gaussian arm is not
 * implemented (VIA_E_INVALID). */
typedef struct {
    int32_t kind;
    uint32_t bound;
} via_dist;

/* Shake256Prg (sampling/prg.rs): the stream of a seed of at most 64 bytes
 * (VIA_E_PRG_SEED beyond), consumed by via_client_setup and
 * via_client_query in the reference's draw order; position = bytes used.
 * One call at a time per via_prg (it is a sequential stream); a failed call
 * leaves its position unchanged. */
typedef struct via_prg via_prg;
via_status via_prg_create(const uint8_t *seed, size_t seed_len, via_prg **out);
void via_prg_destroy(via_prg *prg);
uint64_t via_prg_position(const via_prg *prg);
/* Shake256Prg::fill_bytes (prg.rs:143-158): the next len bytes of the
 * stream (computed on the GPU). */
via_status via_prg_fill(via_ctx *ctx, via_prg *prg, uint8_t *out, size_t len);
/* Shake256Prg::uniform_below (prg.rs:184-207): one draw below bound (>= 1). */
via_status via_prg_uniform_below(via_ctx *ctx, via_prg *prg, uint64_t bound, uint64_t *out);

/* Words of the client's keys: cascade key (log2(n1) steps of 2*n1 RLev
 * samples' worth, [k][part][limb][n_out] per key), conversion key
 * RLev_S1(S1^2) [k][part][limb][n1], ring-switch key [j][k][part][n2]. */
via_status via_client_key_words(const via_params *pp, size_t *cascade, size_t *conv, size_t *rsk);

/* Client::setup (client.rs:89-173): S1 (key1, at q1, [limb][n1]), S2 (key2,
 * at q3, [n2]), the cascade key, RLev_S1(S1^2) and the ring-switch key, all
 * drawn from prg exactly as the reference draws them. The stream's blocks
 * are computed on the GPU and parsed on the host; the ring products run on
 * the GPU. VIA_E_DIM for a non-power-of-two num_rows or num_cols
 * (client.rs:125-131); VIA_E_INVALID for parameters outside the device path
 * (n1 = 2048 over two NTT primes, q3 NTT-friendly at n2). The PRG advances
 * only on success; on an error the output buffers hold partial values.
 * Host buffers. */
via_status via_client_setup(via_ctx *ctx, const via_params *pp, via_dist key1, via_dist key2, via_dist err,
                            via_prg *prg, uint64_t *sk1, uint64_t *sk2, uint64_t *cascade_key, uint64_t *conv_key,
                            uint64_t *rsk);

/* Client::setup for n clients at once (no counterpart in the reference,
 * whose Client::setup is one client per call): client i draws from prgs[i]
 * (distinct streams) into sk1[i], sk2[i], cascade_key[i], conv_key[i] and
 * rsk[i], each word for word what via_client_setup gives for that stream.
 * Up to eight host threads take whole clients in turn, each on a lane of the
 * context (its own stream), so one client's stream parse and body assembly
 * overlap other clients' device work. status[i] receives client i's status;
 * the return value is the first failure (VIA_OK when all succeed). */
via_status via_client_setup_many(via_ctx *ctx, const via_params *pp, uint32_t n, via_dist key1, via_dist key2,
                                 via_dist err, via_prg *const *prgs, uint64_t *const *sk1, uint64_t *const *sk2,
                                 uint64_t *const *cascade_key, uint64_t *const *conv_key, uint64_t *const *rsk,
                                 via_status *status);

/* Client::query (client.rs:183-207 -> build_compressed_query, query.rs:
 * 75-130): the query's raw LWEs ([lwe][limb][n1 + 1], bits in DMux, CMux,
 * CRot order, L_QUERY per bit) under S1, drawn from prg. VIA-B batch queries
 * (client.rs:256-300) are T calls in batch order with n_rec the record
 * degree. VIA_E_INDEX for index >= (n1/n_rec) * num_rows * num_cols,
 * VIA_E_DIM when num_rows, num_cols or n1/n_rec is not a power of two. */
via_status via_client_query(via_ctx *ctx, const via_params *pp, const uint64_t *sk1, size_t index, via_dist err,
                            via_prg *prg, uint64_t *lwes);

/* ---- client recovery (via-client), host code, no device ---------------------- */

/* Client::recover (client.rs:222-239 -> decrypt_asymmetric, rlwe.rs:187-224,
 * and decode, rlwe.rs:296-330): the n2 plaintext coefficients (mod p) of an
 * answer (mask [n2] < q3, body [n2] < q4) under the client's S2 [n2] < q3.
 * VIA_E_INVALID for non-canonical inputs. */
via_status via_client_recover(const via_params *pp, const uint64_t *sk2, const uint64_t *mask, const uint64_t *body,
                              uint64_t *plain);
/* deinterleave_batch (via-client batch.rs:31-52): record t < T of a VIA-B
 * plaintext is coefficient d*c + t, c < n_rec, d = n2 / n_rec; records
 * [T][n_rec]. VIA_E_INVALID when T * n_rec > n2. */
via_status via_client_deinterleave(const via_params *pp, uint32_t T, const uint64_t *plain, uint64_t *records);

/* ---- serving loop (no counterpart in the reference, which has no server
 * process; the answers it produces are those of Server::answer) ----------- */
typedef struct via_server via_server;
typedef struct {
    uint64_t answered;       /* requests answered (a VIA-B batch counts once) */
    uint64_t batches;        /* launches */
    uint64_t max_batch_seen; /* largest fused VIA-C batch */
    uint64_t clients, key_bytes, pending;
    uint64_t records;        /* records delivered: 1 per VIA-C answer, T per VIA-B answer */
    uint64_t answered_b;     /* VIA-B requests answered */
    double p50_s, p99_s;     /* submit-to-result latency percentiles, all requests */
    double p50_b_s, p99_b_s; /* the same over VIA-B requests */
} via_server_stats;

/* Starts a request loop over `db` (VIA-C, record degree pp->n_rec) and, when
 * db_b is not NULL, `db_b` (VIA-B: the same grid and client parameters at a
 * finer record degree) on `device` with `slots` device contexts (launches in
 * flight at once) and VIA-C batches of at most `max_batch` queries. Client
 * keys are admitted while their HBM footprint fits key_budget_bytes. */
via_status via_server_create(int device, const via_params *pp, const via_db *db, const via_db *db_b,
                             size_t key_budget_bytes, uint32_t max_batch, uint32_t slots, via_server **out);
void via_server_destroy(via_server *srv);
/* PreparedKeys of one client, resident until evicted; VIA_E_NOMEM when the
 * budget is exhausted. */
via_status via_server_admit(via_server *srv, uint64_t client_id, const uint64_t *cascade_key,
                            const uint64_t *conv_key, const uint64_t *rsk);
/* Waits for the client's outstanding queries, then frees (and zeroizes) its keys. */
via_status via_server_evict(via_server *srv, uint64_t client_id);
/* Queues one query of an admitted client; returns at once with a ticket. */
via_status via_server_submit(via_server *srv, uint64_t client_id, const uint64_t *lwes, size_t n_lwe,
                             uint64_t *ticket);
/* Queues one VIA-B batch (answer_batch::<T> against db_b): T queries of the
 * client back to back, n_lwe LWEs each; the ticket's answer is the one
 * repacked ciphertext. T at most 64 and accepted by via_answer_batch. */
via_status via_server_submit_batch(via_server *srv, uint64_t client_id, uint32_t T, const uint64_t *lwes,
                                   size_t n_lwe, uint64_t *ticket);
via_status via_server_poll(via_server *srv, uint64_t ticket, int *done);
/* Blocks until the ticket's answer is ready and copies it out. */
via_status via_server_wait(via_server *srv, uint64_t ticket, uint64_t *mask, uint64_t *body, double *latency_s);
via_status via_server_get_stats(via_server *srv, via_server_stats *out);
via_status via_server_reset_stats(via_server *srv);

/* Device memory helpers for callers without a runtime of their own. */
via_status via_device_alloc(via_ctx *ctx, size_t bytes, void **d_ptr);
via_status via_device_free(via_ctx *ctx, void *d_ptr);
via_status via_copy_to_device(via_ctx *ctx, void *d_dst, const void *src, size_t bytes);
via_status via_copy_to_host(via_ctx *ctx, void *dst, const void *d_src, size_t bytes);
/* Device-to-device copy enqueued on the context stream (asynchronous). */
via_status via_copy_device(via_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);

const char *via_status_str(via_status s);

#ifdef __cplusplus
}
#endif
#endif
