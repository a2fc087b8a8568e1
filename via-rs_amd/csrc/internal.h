// internal.h — device data structures and launchers shared by the HIP kernels
// and the host library. Not part of the C ABI (see include/via_gpu.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "modarith.h"
#include "xcd_map.h"

namespace via {

constexpr int kSlots = 2048;    // coefficients handled by one workgroup
constexpr int kSlotsLog = 11;
constexpr int kThreads = 256;   // workgroup size of the polynomial kernels
constexpr int kPer = kSlots / kThreads;
constexpr int kMaxQ = 32;       // queries answered together in one batched launch
// NTT moduli stay below 2^38: the lazy transforms keep residues below 64q,
// exact in fp64 lanes (ntt_lazy.h).
constexpr uint64_t kMaxModulus = 1ULL << 38;

constexpr int kDirectWords = 24;  // direct first-pass constants stored after each forward table

// Twiddles of one (q, n): bit-reversed psi^k and psi^-k as exact doubles,
// the layout of via-rs ring/ntt.rs:206-221.
struct NttTab {
    const double *fwd;
    const double *dfw;  // kDirectWords constants of the direct first forward pass
    const double *inv;
    double n_inv;
    FMod m;
};

// Modulus context of a ring with NL limbs (1 or 2) at degree n.
struct RingDev {
    int nl;
    int n;
    int logn;
    NttTab t[2];
    Rns2 rns;          // valid when nl == 2
    WideDiv qdiv;      // Q as a wide divisor
    uint64_t q_single; // Q when nl == 1
};

// Gadget parameters: base, depth, B^L. wide: single-modulus scaling needs
// the 192-bit path ((q/2) * B^L beyond 63 bits).
struct Gadget {
    uint32_t base;
    int L;
    via_u128 bl;
    int wide;
    // fp64 fast path of the scale (set by the gate launcher): 0 none,
    // 1 single modulus with an exact fp64 numerator, 2 two-limb estimate
    // with the exact path as fallback (modarith.h)
    int f64;
    ScaleF64 sf;
    ScaleW sw;  // f64 == 3: the wide split of gadget_split_w
};

// The wide fp64 split of a gadget over the ring r (two limbs, or one): true
// and *sw filled when its bounds hold (K, q1 < 2^39 for two limbs; q < 2^39
// and K < 2^53 for one; B^(L-h) < 2^40).
bool scale_w_make(const RingDev &r, const Gadget &g, const DigitPlan &dp, ScaleW *sw);

// Test switch (via_debug_set_exact_paths): when set, every launcher takes
// the exact 128/192-bit paths of the gadget scale and the q1 -> q2 mod switch
// in place of their fp64 fast paths, so the paths those fast paths fall back
// to run on every coefficient. The values are the same either way.
bool exact_paths_only();
void set_exact_paths_only(bool on);

// The q1 -> q2 switch of two-limb coefficients as kernels take it: the fp64
// estimate (fast, for q2 < q1 and both limbs within the fp64 path) with the
// exact wide division as the fallback of the coefficients it leaves open.
struct SwitchIn {
    Rns2 rns;
    WideDiv qd;
    SwitchF sf;
    uint64_t q2;
    int fast;
};
SwitchIn switch_in(const RingDev &r1, uint64_t q2);

// ---- launchers (ntt.hip / kernels.hip) ----------------------------------------

// In-place batched NTT over `npolys` polys of degree r.n stored contiguously;
// poly p uses limb (p / limb_group) % r.nl.
void launch_ntt(const RingDev &r, double *data, size_t npolys, size_t limb_group, bool inverse,
                hipStream_t s);
// The same forward transform from src into dst (src may equal dst).
void launch_ntt_to(const RingDev &r, const double *src, double *dst, size_t npolys, size_t limb_group, hipStream_t s);
// Forward transforms of `groups` groups of gpolys polys: group g's poly i at
// src + g*src_gstride + i*n into dst + g*dst_gstride + i*n; the limb of poly
// i is (i / limb_group) % r.nl.
void launch_ntt_groups(const RingDev &r, const double *src, double *dst, size_t gpolys, size_t groups,
                       size_t src_gstride, size_t dst_gstride, size_t limb_group, hipStream_t s);
// u64 -> f64 of `count` arrays of `words` words, array q to dst + q*words.
struct U64Srcs {
    const uint64_t *p[kMaxQ];
};
// The LWE arrays of a batch (two limbs of n1 + 1 words per LWE, limb l below
// q[l]); a non-canonical word is taken as 0 and sets *bad (device pointer).
void launch_u64_to_f64_multi(const U64Srcs &srcs, int count, double *dst, size_t words, const uint64_t q[2], int n1,
                             int *bad, hipStream_t s);
void launch_u64_to_f64(const uint64_t *src, double *dst, size_t n, hipStream_t s);
void launch_f64_to_u64(const double *src, uint64_t *dst, size_t n, hipStream_t s);
// Negacyclic products c = iNTT(NTT(a) * NTT(b)) of `count` u64 polys per limb
// group; poly p uses limb (p / limb_group) % r.nl.
// (bidx: product i's b operand is entry bidx[i] of b, nl polys each, so a
// shared operand is stored once; null: b holds one entry per product)
void launch_polymul_u64(const RingDev &r, const uint64_t *a, const uint64_t *b, uint64_t *c, size_t count,
                        size_t limb_group, hipStream_t s, const uint32_t *bidx = nullptr);

// Digits of the LWE masks of `items` LWEs (int16 planes [item][k][n]).
void launch_lwe_digits(const RingDev &r1, const Gadget &g, const DigitPlan &dp, const double *lwe, int items,
                       int16_t *dig, hipStream_t s);
// Digits of `items` vectors of n coefficients (limb l of item i at
// src[i*item_stride + l*limb_stride]) into int16 planes [item][k][n].
void launch_poly_digits(const RingDev &r1, const Gadget &g, const DigitPlan &dp, const double *src, size_t item_stride,
                        size_t limb_stride, int n, int items, int16_t *dig, hipStream_t s);

// Gadget-product multiply-accumulate over digit planes.
struct MacArgs {
    const int16_t *dig;      // [item][k][dig_coeffs] digits of the inputs
    size_t dig_coeffs;       // coefficients per digit plane (tiles * S or tiles * S/2)
    int embed;               // 1: inputs of degree n/2 embedded at slot 0 (x -> x^2)
    const double *key[kMaxQ];  // eval-form RLev per query (client), [k][part][limb][key_slots]
    int items_per_key;       // consecutive items that share one key
    size_t key_slots;        // slots per (k, part, limb) plane
    int tiles;               // workgroups per item along the slots (tile t covers [t*S, t*S+S))
    int groups;              // workgroups per item along the digits
    double *out_mask;        // [group][item][tile][limb][S] eval
    double *out_body;        // same, or [group][item][tile][limb][n] when reduce_body
    int reduce_body;
    int items;
    int wide;                // 1: digit transforms of a group side by side (k_gadget_mac_wide)
    int radix16;             // 256-thread kernel: radix-16 transform passes (else radix 8)
    int t512;                // 512-thread kernel with radix-8 passes
    int lsplit;              // two limbs: one workgroup per limb (not with `wide`)
    int ipw;                 // fixed-shape cascade MAC: items per workgroup (1, 2 or 4)
};
void launch_gadget_mac(const RingDev &r, const Gadget &g, const MacArgs &a, hipStream_t s);
// The cascade-step MAC in its fixed shape (cascade_mac.hip): two q1 limbs,
// embedded inputs, two tiles of 2048 slots, or the RLWE -> RGSW conversion
// (2048 inputs, one tile); keys in the cascade_key_pos order.
// cascade_mac_fits says whether a launch meets those conditions and the
// exactness bound of its direct first pass.
bool cascade_mac_fits(const RingDev &r, const Gadget &g, const MacArgs &a);
void launch_cascade_mac(const RingDev &r, const Gadget &g, const MacArgs &a, hipStream_t s);
__host__ __device__ int cascade_key_pos(int sg);
// rows of tiles x 2048 slots from slot order into the cascade_key_pos order
void launch_cascade_key_permute(const double *src, double *dst, size_t rows, int tiles, hipStream_t s);

// One cascade combine: masks_next = -iNTT(sum pm), body_next =
// embed(body_prev) - iNTT(sum pb), then the digits of masks_next into
// dig_next (rq: the q1 ring whose RNS and gadget constants the scale uses).
struct CombineArgs {
    const double *pm, *pb;    // MAC outputs with tiles == 2
    int groups, items;
    const double *body_prev;  // [item] stride body_item_stride, limb stride body_limb_stride
    size_t body_item_stride, body_limb_stride;
    double *masks_next;       // [item][limb][S]
    double *body_next;        // [item][limb][n]
    int16_t *dig_next;        // [item][k][S] or null
};
void launch_cascade_combine(const RingDev &r, const Gadget &g, const DigitPlan &dp, const RingDev &rq,
                            const CombineArgs &a, hipStream_t s);

// rlwe_to_rgsw combine: neg_s_m[i] = (B_i + iNTT(pm), iNTT(pb)); m[i] = level i.
void launch_r2g_combine(const RingDev &r, const double *pm, const double *pb, int groups, const double *rlwe_masks,
                        const double *rlwe_bodies, double *rgsw, int bits, int lq, hipStream_t s);

// Gate kernels over an eval-form RGSW (layout [half][i][part][limb][n]). Item
// b of the launch is output j = b % items_per_query of query b / items_per_query.
// GATE_CMUX_EV: the first CMux level over FirstDim outputs in pair form
// (launch_fd_intt_pairs), fixed-shape gates only (gate_takes_fx).
enum GateMode { GATE_DMUX = 0, GATE_CMUX = 1, GATE_CROT = 2, GATE_CMUX_EV = 3 };
struct GateArgs {
    int mode;
    const double *rgsw;  // query q's RGSW at rgsw + q * rgsw_qstride
    size_t rgsw_qstride;
    const double *in;
    size_t in_qstride;
    double *out;
    size_t out_qstride;
    int items_per_query;
    int queries;
    int shift;           // CRot rotation N - 2^i
    DigitPlan dp;        // digit plan of the gadget (set by launch_gate)
    int split = -1;      // two limbs, one workgroup per limb: -1 below 256 gates, 0 never, 1 always
    // narrow levels of at most kGateNxSplitMax gate limbs: device words for
    // the split gate's partial products (gates * limbs * 4 * n), or null
    double *partials = nullptr;
    size_t partials_words = 0;
};
void launch_gate(const RingDev &r, const Gadget &g, const GateArgs &a, hipStream_t s);
// whether launch_gate runs a one-limb level of `blocks` gates on the
// fixed-shape gate (the condition of the pair-form FirstDim outputs)
bool gate_takes_fx(const RingDev &r, const Gadget &g, int blocks);
// The fixed-shape gate (gate_fx.hip) for levels of at least kGateFxMin
// gates (one workgroup per limb with two limbs): degree 2048, gadget depth
// 2; *direct when its first forward pass may run in direct form.
constexpr int kGateFxMin = 256;
bool gate_fx_fits(const RingDev &r, const Gadget &g, bool *direct);
void launch_gate_fx(const RingDev &r, const Gadget &g, const GateArgs &a, bool direct, hipStream_t s);
// The same shape for narrow levels (fewer than kGateFxMin gates): one gate,
// or one limb of a two-limb gate, per 1024-thread workgroup (not for
// GATE_CMUX_EV); with a.partials, levels of at most kGateNxSplitMax gate
// limbs (gates times limbs) take two launches of twice the workgroups
// (halves, then product parts); above that the one launch is faster.
constexpr int kGateNxSplitMax = 64;
void launch_gate_nx(const RingDev &r, const Gadget &g, const GateArgs &a, bool direct, hipStream_t s);

// Every level of one DMux or CMux tree in one cooperative launch
// (k_tree_persistent): level v's gates read lev[v].in and write lev[v].out
// with the RGSW lev[v].rgsw; the grid waits between levels at a barrier
// (bar: two device words, zero before the first use and left consistent by
// every launch that does not fail) whose waits give up after `spin` polls and
// set *fail (host-visible).
constexpr int kMaxTreeLevels = 16;
struct TreeLevel {
    const double *rgsw;
    const double *in;
    double *out;
    int items_per_query;
};
struct TreeArgs {
    GateArgs gate;  // mode, strides, queries, shift (the per-level fields are in lev)
    int nlev;
    TreeLevel lev[kMaxTreeLevels];
    unsigned *bar;
    int *fail;
    long spin;
};
// false when the device cannot run the grid co-resident (the caller then
// launches level by level)
bool launch_tree_persistent(const RingDev &r, const Gadget &g, const TreeArgs &a, hipStream_t s);

// Gadget digits of Delta (the trivial DMux root) and Delta per limb.
struct DeltaDigits {
    int32_t d[8];
    double delta[2];
};
void launch_dmux_trivial(const RingDev &r, const double *rgsw_coeff, size_t rgsw_qstride, int L, const DeltaDigits &dd,
                         double *out, size_t out_qstride, int queries, hipStream_t s);

// Mod switch of a coefficient-form array at (n, Q1 two limbs) to q2: rows of
// 2 polys [row][part][limb][n] -> [row][part][n].
void launch_modswitch_q1_q2(const RingDev &r1, uint64_t q2, const double *in, double *out, size_t polys,
                            hipStream_t s);
// The same for `groups` groups of polys: group g reads in + g*in_gstride and
// writes out + g*out_gstride.
void launch_modswitch_q1_q2_groups(const RingDev &r1, uint64_t q2, const double *in, size_t in_gstride, double *out,
                                   size_t out_gstride, size_t polys, int groups, hipStream_t s);
// Mod switch q1 -> q2 and forward transform in one pass: the two-limb polys
// [poly][limb][n] of src (at the degree kSlots) into the evaluation form mod q2
// [poly][n] of dst; the array launch_modswitch_q1_q2 writes is not formed.
bool ntt_switch_fits(const RingDev &r1, const RingDev &r2);
void launch_ntt_switch(const RingDev &r1, const RingDev &r2, uint64_t q2, const double *src, double *dst,
                       size_t npolys, hipStream_t s);
// The same for `groups` groups of gpolys polys: group g's two-limb polys at
// src + g*src_gstride into dst + g*dst_gstride.
void launch_ntt_switch_groups(const RingDev &r1, const RingDev &r2, uint64_t q2, const double *src, size_t src_gstride,
                              double *dst, size_t dst_gstride, size_t gpolys, size_t groups, hipStream_t s);

// FirstDim on the int8 matrix cores (first_dim.hip). The prepared database
// is held as 7-bit limb fragments (5 bytes per residue); a pass packs the
// selection vectors the same way, runs one workgroup per (slot, segment of
// kFdSegRows rows) and gathers the segment outputs to [q][j][part][k].
constexpr int kFdSegRows = 256;
struct FdPlan {
    int n, I, J, Q;
    int Ipad, Jpad;  // rows padded to KC (at least 32), columns to 16
    int KC;          // rows per MFMA fragment: 64, or 32 when Ipad is 32
    int MB, KCN;     // column blocks of 16, row chunks of KC
    int QP, NB;      // selection columns 2Q, blocks of 16
    int nseg;        // row segments
    int G = 0;       // column-block groups per slot (0: the launcher's heuristic)
};
FdPlan fd_plan(int I, int J, int n, int Q);
size_t fd_db_bytes(const FdPlan &p);   // prepared database
size_t fd_sel_bytes(const FdPlan &p);  // packed selection of a pass
size_t fd_out_words(const FdPlan &p);  // segment outputs (doubles)
void launch_pack_db(const double *cells, uint8_t *dbl, const FdPlan &p, hipStream_t s);
void launch_unpack_db(const uint8_t *dbl, uint64_t *out, const FdPlan &p, hipStream_t s);
void launch_sel_pack(const double *sel_eval, uint8_t *sell, const FdPlan &p, hipStream_t s);
void launch_fd_mfma(const uint8_t *dbl, const uint8_t *sell, double *out_k, const FdPlan &p, FMod m, hipStream_t s);
// The segment sum, transpose and inverse transform (r2: q2 at n) of the
// FirstDim outputs in one launch: out [q][j][part][n] in coefficient form.
void launch_fd_intt(const double *out_k, double *out, const FdPlan &p, const RingDev &r2, hipStream_t s);
// The same outputs in pair form for a first CMux level of GATE_CMUX_EV
// gates: column 2i holds c1 - c0 in coefficient form, column 2i + 1 c0 in
// eval form (fd_pairs_fits: one row segment, n = 2048, J a multiple of 16).
bool fd_pairs_fits(const FdPlan &p);
void launch_fd_intt_pairs(const double *out_k, double *out, const FdPlan &p, const RingDev &r2, hipStream_t s);
// query 0's pair-form outputs (J columns) back to coefficient form
void launch_fd_unpair(const double *in, double *out, int J, const RingDev &r2, hipStream_t s);
// Rewrites record slots of whole cells in place (r2: q2 at n1): cell c =
// cells[x] takes, for slot s < d, record slot_src[x*d + s] of recs (n_rec
// coefficients each, reduced mod pmod) or keeps its own (-1).
void launch_db_update(uint8_t *dbl, const FdPlan &p, const RingDev &r2, const int64_t *cells, int ncells,
                      const int32_t *slot_src, int d, const uint64_t *recs, int n_rec, uint64_t pmod, hipStream_t s);

// RespComp for Q ciphertexts at (n1, q2): mod switch to q3, ring switch to n2
// with each query's eval RSK ([j][k][part][n2]), then mask at q3 / body at q4.
struct RespArgs {
    const double *ct;  // query q at ct + q * 2 * n1, [part][n1] at q2
    const double *rsk_eval[kMaxQ];
    uint64_t *out_mask[kMaxQ];
    uint64_t *out_body[kMaxQ];
    int n1, n2, d, l_rsk;
    uint32_t b_rsk;
    uint64_t q2, q3, q4;
    uint64_t minv2 = 0, minv3 = 0;  // floor((2^64 - 1) / q2), / q3 (set by launch_resp_comp)
    NttTab t3;        // q3 at n2
    int queries;
    DigitPlan dp;     // digit plan of (b_rsk, l_rsk) (set by launch_resp_comp)
    // device words for the ring switch's per-j partial sums
    // (queries * d * 2 * n2), or null: one workgroup per query then
    double *partials = nullptr;
    size_t partials_words = 0;
};
void launch_resp_comp(const RespArgs &a, hipStream_t s);

// ---- Shake256Prg draws (prg.hip) ----
// Seeds of up to kMaxQ streams (at most 64 bytes each, prg.rs SEED_CAP).
struct PrgSeeds {
    uint8_t bytes[kMaxQ][64];
    int len[kMaxQ];
};
// count uniform_below draws per stream, draw d at modulus q[d % nmod]
// (nbytes bytes per attempt for every modulus); out[s*out_sstride + d], or
// with lwe set, draw d of row d / (n*nmod), limb d % nmod, coefficient
// (d / nmod) % n at row*row_stride + limb*limb_stride + coefficient.
struct PrgPlan {
    uint64_t q[2], mask[2];
    int nmod, nbytes;
    size_t count, attempts, blocks;
    int lwe, n;
    size_t row_stride, limb_stride, out_sstride;
};
// stream: streams * blocks * 17 words; short_flag set when a stream had
// fewer accepted attempts than draws (the caller sizes attempts with slack).
void launch_prg_draws(const PrgSeeds &seeds, int streams, const PrgPlan &pl, uint64_t *stream, uint64_t *out,
                      int *short_flag, hipStream_t s);
// Blocks first .. first + count - 1 of stream 0 of `seeds`, 17 little-endian
// lanes each, into out (count * 17 words).
void launch_prg_window(const PrgSeeds &seeds, uint64_t first, size_t count, uint64_t *out, hipStream_t s);

// ---- VIA-B repack (repack.hip) ----
// Level-0 digits [pair][k][2*n1] and interleaved bodies [pair][body_stride]
// from the T post-CRot ciphertexts rot [T][part][n1] (ring r: q2 at n1).
void launch_repack_extract(const RingDev &r, const Gadget &g, const DigitPlan &dp, const double *rot, int T, int G,
                           int pairs, double *body0, size_t body_stride, int16_t *dig, hipStream_t s);
struct RepackCombineArgs {
    const double *pm, *pb;  // MAC outputs (tiles == 2, reduced bodies)
    int groups, items;
    const double *body_cur;  // [item][body_stride], degree n
    size_t body_stride;
    int final_level;
    double *out;            // final level: RLWE [part][n1]
    int16_t *dig_next;      // next level: [item/2][k][2*n1]
    double *body_next;      // next level: [item/2][body_stride]
    int zero_partner;       // one live pair: the next pair's odd entries are zero
};
void launch_repack_combine(const RingDev &r, const Gadget &g, const DigitPlan &dp, const RepackCombineArgs &a,
                           hipStream_t s);
void launch_key_relayout_q2(const double *src, double *dst, int rank_in, int L, int n, hipStream_t s);

// Relayout of a cascade step key [idx][k][part][limb][n] (u64) to the
// eval-ready layout [k][part][limb][idx*n + c] (doubles).
void launch_key_relayout(const uint64_t *src, double *dst, int rank_in, int L, int nl, int n, hipStream_t s);
// DB packing: records [rec][n_rec] (u64 < p) -> cells [i][j][n1] doubles (setup_db.rs:39-75).
void launch_setup_db(const uint64_t *records, size_t n_records, double *cells, int I, int J, int n1, int n_rec,
                     uint64_t p, int *bad, hipStream_t s);
void launch_setup_db_bytes(const uint8_t *bytes, size_t n_records, double *cells, int I, int J, int n1,
                           int n_rec, hipStream_t s);

}  // namespace via
