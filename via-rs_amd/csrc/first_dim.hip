// first_dim.hip — FirstDim (via-rs via-server/src/first_dim.rs:94-184) on the
// int8 matrix cores, and the prepared-database layout it streams.
//
// For every slot k the pass is a matrix product over the database rows i,
//   out[q][j][part][k] = sum_i db[i][j][k] * sel[q][i][part][k]  (mod q2),
// of slot k's J x I database block with the I x 2Q selection block of the
// pass's Q queries. Residues are split into five 7-bit limbs (fd_limbs.h);
// the 25 limb products of every (database, selection) pair accumulate exactly
// in int32 on v_mfma_i32_16x16x64_i8 (16x16x32 for 32-row databases), into
// nine sums by limb shift, which limb_fold brings back to a residue.
//
// Prepared database, 5 bytes per residue: for slot k, column block mb (16
// columns j), row chunk kc (KC rows i) and limb l, one MFMA operand fragment
// of 16 x KC bytes in which lane 16 g + r holds column j = 16 mb + r, rows
// i = KC kc + (KC/4) g + [0, KC/4). Fragments run [k][mb][kc][l], so slot k's
// block is contiguous and its workgroup streams it once.
// Selection, packed per pass: [k][kc][l][g][qp][KC/4 bytes] with qp = 2q +
// part and the same row order, staged through LDS by each workgroup.
// The MFMA B operand's lane 16 g + c supplies column c for the same rows, so
// a fragment pair multiplies exactly the rows it should whatever order the
// hardware gives the bytes inside a lane group.
// Output of a workgroup (slot k, segment of kFdSegRows rows): rows of J
// contiguous words, out_k[seg][k / 16][q][part][k % 16][j] (fd_out_at): the
// 16 q-part rows a workgroup writes lie within 16 * 2Q * J words, and the
// inverse transform reads each output poly in 128 runs of 16 rows at a
// stride of J words. k_fd_intt adds the segments, transposes to
// [q][j][part][k] and transforms back.
#include "internal.h"
#include "fd_limbs.h"
#include "ntt_fixed.h"

namespace via {

typedef int fd_v4i __attribute__((ext_vector_type(4)));

// Word of FirstDim output (segment seg, slot k, column qp = 2q + part, row
// j) in the layout the MFMA pass writes: blocks of kFdKlo slots, then qp,
// then the slot within the block, then j (n a multiple of kFdKlo). The
// contributions of the arguments add: fd_out_at(p, s, k, qp, j) =
// fd_out_at(p, s, k, 0, 0) + fd_out_at(p, 0, 0, qp, j).
constexpr int kFdKlo = 16;
__host__ __device__ __forceinline__ size_t fd_out_at(const FdPlan &p, int seg, int k, int qp, int j) {
    return (((size_t)seg * (p.n / kFdKlo) + (size_t)(k / kFdKlo)) * p.QP + (size_t)qp) * kFdKlo * p.J +
           (size_t)(k % kFdKlo) * p.J + (size_t)j;
}


template <int KC>
struct FdOp;
template <>
struct FdOp<64> {
    typedef fd_v4i T;  // 16 bytes per lane
    static __device__ __forceinline__ T zero() { return fd_v4i{0, 0, 0, 0}; }
    static __device__ __forceinline__ fd_v4i mac(T a, T b, fd_v4i c) {
        return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
    }
};
template <>
struct FdOp<32> {
    typedef long T;  // 8 bytes per lane
    static __device__ __forceinline__ T zero() { return 0; }
    static __device__ __forceinline__ fd_v4i mac(T a, T b, fd_v4i c) {
        return __builtin_amdgcn_mfma_i32_16x16x32_i8(a, b, c, 0, 0, 0);
    }
};

FdPlan fd_plan(int I, int J, int n, int Q) {
    FdPlan p;
    p.n = n;
    p.I = I;
    p.J = J;
    p.Q = Q;
    p.Ipad = I <= 32 ? 32 : (I + 63) / 64 * 64;
    p.KC = p.Ipad % 64 == 0 ? 64 : 32;
    p.Jpad = (J + 15) / 16 * 16;
    p.MB = p.Jpad / 16;
    p.KCN = p.Ipad / p.KC;
    p.QP = 2 * Q;
    p.NB = (p.QP + 15) / 16;
    const int ks = kFdSegRows / p.KC;
    p.nseg = (p.KCN + ks - 1) / ks;
    return p;
}

size_t fd_db_bytes(const FdPlan &p) { return (size_t)p.n * p.Ipad * p.Jpad * kLimbs; }
size_t fd_sel_bytes(const FdPlan &p) { return (size_t)p.n * p.Ipad * kLimbs * p.QP; }
size_t fd_out_words(const FdPlan &p) { return (size_t)p.nseg * p.n * p.Q * 2 * p.J; }

// KB limb bytes of each of the five limbs of values v[0..KB), packed little
// end first into words.
template <int KB>
__device__ __forceinline__ void pack_limbs(const double *v, uint32_t (&wd)[kLimbs][KB / 4]) {
#pragma unroll
    for (int a = 0; a < kLimbs; a++)
#pragma unroll
        for (int w = 0; w < KB / 4; w++) wd[a][w] = 0;
#pragma unroll
    for (int b = 0; b < KB; b++) {
        uint32_t l[kLimbs];
        limbs7(v[b], l);
#pragma unroll
        for (int a = 0; a < kLimbs; a++) wd[a][b >> 2] |= l[a] << (8 * (b & 3));
    }
}

template <int KB>
__device__ __forceinline__ void store_words(uint8_t *dst, const uint32_t (&w)[KB / 4]) {
    if constexpr (KB == 16) *(uint4 *)dst = make_uint4(w[0], w[1], w[2], w[3]);
    else *(uint2 *)dst = make_uint2(w[0], w[1]);
}

// ---- prepared database ------------------------------------------------------------

// cells [I][J][n] (canonical residues) -> fragments; rows i >= I and columns
// j >= J are zero. Eight consecutive slots per lane group read 64-byte runs.
template <int KC>
__global__ __launch_bounds__(256) void k_pack_db(const double *cells, uint8_t *dbl, FdPlan p) {
    constexpr int KB = KC / 4, FRAG = 64 * KB;
    const size_t total = (size_t)p.n * p.MB * p.KCN * 64;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int kl = (int)(t & 7), lane = (int)((t >> 3) & 63);
        size_t rest = t >> 9;
        const int kc = (int)(rest % p.KCN);
        rest /= p.KCN;
        const int mb = (int)(rest % p.MB);
        const int k = (int)(rest / p.MB) * 8 + kl;
        const int r = lane & 15, g = lane >> 4, j = mb * 16 + r;
        double v[KB];
#pragma unroll
        for (int b = 0; b < KB; b++) {
            const int i = kc * KC + g * KB + b;
            v[b] = (i < p.I && j < p.J) ? cells[((size_t)i * p.J + j) * p.n + k] : 0.0;
        }
        uint32_t wd[kLimbs][KB / 4];
        pack_limbs<KB>(v, wd);
        uint8_t *dst = dbl + ((((size_t)k * p.MB + mb) * p.KCN + kc) * kLimbs) * FRAG + (size_t)lane * KB;
#pragma unroll
        for (int a = 0; a < kLimbs; a++) store_words<KB>(dst + (size_t)a * FRAG, wd[a]);
    }
}

// fragments -> u64 cells [I][J][n] (the prepared database as the reference
// holds it, for inspection).
template <int KC>
__global__ __launch_bounds__(256) void k_unpack_db(const uint8_t *dbl, uint64_t *out, FdPlan p) {
    constexpr int KB = KC / 4, FRAG = 64 * KB;
    const size_t total = (size_t)p.I * p.J * p.n;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(t % p.n);
        const size_t rest = t / p.n;
        const int j = (int)(rest % p.J), i = (int)(rest / p.J);
        const int mb = j >> 4, r = j & 15, kc = i / KC, g = (i % KC) / KB, b = i % KB;
        const uint8_t *src = dbl + ((((size_t)k * p.MB + mb) * p.KCN + kc) * kLimbs) * FRAG + (size_t)(g * 16 + r) * KB + b;
        uint64_t v = 0;
        for (int a = kLimbs - 1; a >= 0; a--) v = v * 128 + src[(size_t)a * FRAG];
        out[t] = v;
    }
}

void launch_pack_db(const double *cells, uint8_t *dbl, const FdPlan &p, hipStream_t st) {
    const size_t total = (size_t)p.n * p.MB * p.KCN * 64;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
    if (p.KC == 64) hipLaunchKernelGGL(k_pack_db<64>, dim3(blocks), dim3(256), 0, st, cells, dbl, p);
    else hipLaunchKernelGGL(k_pack_db<32>, dim3(blocks), dim3(256), 0, st, cells, dbl, p);
}

void launch_unpack_db(const uint8_t *dbl, uint64_t *out, const FdPlan &p, hipStream_t st) {
    const size_t total = (size_t)p.I * p.J * p.n;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
    if (p.KC == 64) hipLaunchKernelGGL(k_unpack_db<64>, dim3(blocks), dim3(256), 0, st, dbl, out, p);
    else hipLaunchKernelGGL(k_unpack_db<32>, dim3(blocks), dim3(256), 0, st, dbl, out, p);
}

// Incremental update of whole cells (one workgroup of 256 threads per
// touched cell): gather the cell's 2048 eval residues from the fragments,
// inverse-transform at q2 to the coefficient form setup_db builds
// (setup_db.rs:58-72: record slot s holds coefficient c at d*c + s), rewrite
// the slots listed in slot_src (index of the new record in recs, or -1 to
// keep), transform back and scatter the limb bytes. Distinct cells own
// distinct bytes of the fragments.
template <int KC>
__global__ __launch_bounds__(256) void k_db_update(uint8_t *dbl, FdPlan p, RingDev r2, const int64_t *cells,
                                                   const int32_t *slot_src, int d, const uint64_t *recs, int n_rec,
                                                   uint64_t pmod) {
    constexpr int KB = KC / 4, FRAG = 64 * KB;
    __shared__ double s[lds_words(kSlots)];
    const int64_t cell = cells[blockIdx.x];
    const int i = (int)(cell / p.J), j = (int)(cell - (int64_t)i * p.J);
    const int mb = j >> 4, rr = j & 15, kc = i / KC, g = (i % KC) / KB, b = i % KB;
    const size_t lane_off = (size_t)(g * 16 + rr) * KB + b;
    const size_t kstride = (size_t)p.MB * p.KCN * kLimbs * FRAG;
    const size_t frag0 = (((size_t)mb * p.KCN + kc) * kLimbs) * FRAG + lane_off;
    const int n = p.n;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const uint8_t *src = dbl + (size_t)k * kstride + frag0;
        double v = 0.0;
        for (int a = kLimbs - 1; a >= 0; a--) v = v * 128.0 + (double)src[(size_t)a * FRAG];
        s[lds_pad(k)] = v;
    }
    __syncthreads();
    lds_ntt_inv_noscale<1, 4>(s, n, r2);
    __syncthreads();
    const FMod m = r2.t[0].m;
    for (int c = threadIdx.x; c < n; c += blockDim.x) s[lds_pad(c)] = mod_mul(s[lds_pad(c)], r2.t[0].n_inv, m);
    __syncthreads();
    for (int t = threadIdx.x; t < d * n_rec; t += blockDim.x) {
        const int slot = t / n_rec, c = t - slot * n_rec;
        const int src = slot_src[(size_t)blockIdx.x * d + slot];
        if (src >= 0) s[lds_pad(d * c + slot)] = (double)(recs[(size_t)src * n_rec + c] % pmod);
    }
    __syncthreads();
    lds_ntt_fwd<1, 4>(s, n, r2);
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        uint32_t l[kLimbs];
        limbs7(mod_reduce_small(s[lds_pad(k)], m), l);
        uint8_t *dst = dbl + (size_t)k * kstride + frag0;
        for (int a = 0; a < kLimbs; a++) dst[(size_t)a * FRAG] = (uint8_t)l[a];
    }
}

void launch_db_update(uint8_t *dbl, const FdPlan &p, const RingDev &r2, const int64_t *cells, int ncells,
                      const int32_t *slot_src, int d, const uint64_t *recs, int n_rec, uint64_t pmod, hipStream_t st) {
    if (ncells < 1) return;
    if (p.KC == 64)
        hipLaunchKernelGGL(k_db_update<64>, dim3(ncells), dim3(256), 0, st, dbl, p, r2, cells, slot_src, d, recs,
                           n_rec, pmod);
    else
        hipLaunchKernelGGL(k_db_update<32>, dim3(ncells), dim3(256), 0, st, dbl, p, r2, cells, slot_src, d, recs,
                           n_rec, pmod);
}

// ---- selection pack -------------------------------------------------------------------

// sel [q][i][part][n] (canonical residues at q2) -> [k][kc][l][g][qp][KB].
// A wavefront covers eight slots x eight columns qp: 64-byte read runs and
// full 128-byte write lines.
template <int KC>
__global__ __launch_bounds__(256) void k_sel_pack(const double *sel, uint8_t *sell, FdPlan p) {
    constexpr int KB = KC / 4;
    const unsigned QP4 = (p.QP + 3) / 4, KCN = p.KCN;
    // 64 threads per (slot group of sixteen, chunk, lane group, column group
    // of four): each load of a wavefront reads 16 consecutive slots (a
    // 128-byte line) of four columns; below 2^32 threads in all, so the
    // index splits are 32-bit divisions
    const unsigned total = (unsigned)((p.n + 15) / 16) * KCN * 4 * QP4 * 64;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int kl = (int)(t & 15), qpl = (int)((t >> 4) & 3);
        unsigned rest = t >> 6;
        const int qph = (int)(rest % QP4);
        rest /= QP4;
        const int g = (int)(rest & 3);
        rest >>= 2;
        const int kc = (int)(rest % KCN);
        const int k = (int)(rest / KCN) * 16 + kl;
        const int qp = qph * 4 + qpl;
        if (qp >= p.QP || k >= p.n) continue;
        const int q = qp >> 1, part = qp & 1;
        double v[KB];
#pragma unroll
        for (int b = 0; b < KB; b++) {
            const int i = kc * KC + g * KB + b;
            v[b] = i < p.I ? sel[(((size_t)q * p.I + i) * 2 + part) * p.n + k] : 0.0;
        }
        uint32_t wd[kLimbs][KB / 4];
        pack_limbs<KB>(v, wd);
        uint8_t *dst = sell + (size_t)k * p.KCN * kLimbs * p.QP * KC +
                       ((((size_t)kc * kLimbs) * 4 + g) * p.QP + qp) * KB;
#pragma unroll
        for (int a = 0; a < kLimbs; a++) store_words<KB>(dst + (size_t)a * 4 * p.QP * KB, wd[a]);
    }
}

void launch_sel_pack(const double *sel, uint8_t *sell, const FdPlan &p, hipStream_t st) {
    const size_t total = (size_t)((p.n + 15) / 16) * p.KCN * 4 * ((p.QP + 3) / 4) * 64;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
    if (p.KC == 64) hipLaunchKernelGGL(k_sel_pack<64>, dim3(blocks), dim3(256), 0, st, sel, sell, p);
    else hipLaunchKernelGGL(k_sel_pack<32>, dim3(blocks), dim3(256), 0, st, sel, sell, p);
}

// ---- the matrix-core pass ------------------------------------------------------------

// Workgroup (slot k, column-block group, row segment): the segment's packed
// selection goes to LDS once; wavefront w takes the group's column blocks
// w, w + 4, ... (G groups of ceil(MB / G) blocks per slot: with one
// selection block, G = 4 makes four times as many, shorter workgroups, so
// the last round of the grid leaves fewer CUs idle); for each row
// chunk it holds the five database limb fragments (two more chunks' loads in
// flight) and runs 25 MFMAs per selection block of 16 columns into the nine
// shift sums, then folds and stores its 16 x 2Q outputs.
template <int KC, int NBW, int SPLIT>
__global__ __launch_bounds__(256, 2) void k_fd_mfma(const uint8_t *__restrict__ dbl, const uint8_t *__restrict__ sell,
                                                 double *__restrict__ out, FdPlan p, FMod m, int G) {
    typedef FdOp<KC> Op;
    typedef typename Op::T T;
    constexpr int KB = KC / 4, FRAG = 64 * KB;
    constexpr int KS = kFdSegRows / KC;
    constexpr int NB = NBW * SPLIT;  // selection blocks the LDS holds
    __shared__ __attribute__((aligned(16))) uint8_t sb[KS * kLimbs * 16 * NB * KC];
    const int k = blockIdx.x / G, gi = blockIdx.x - k * G, seg = blockIdx.y;
    const int mbper = (p.MB + G - 1) / G;
    const int mb0 = min(p.MB, gi * mbper), mb1 = min(p.MB, mb0 + mbper);
    const int kc0 = seg * KS, nkc = min(KS, p.KCN - kc0);
    const size_t chunk = (size_t)kLimbs * p.QP * KC;  // packed selection bytes per row chunk
    {
        const uint8_t *src = sell + ((size_t)k * p.KCN + kc0) * chunk;
        const int bytes = nkc * (int)chunk;  // a multiple of 16
        for (int o = threadIdx.x * 16; o < bytes; o += 256 * 16) *(uint4 *)(sb + o) = *(const uint4 *)(src + o);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // SPLIT wavefronts share a column block, each with NBW selection blocks
    const int w = wv / SPLIT, nb0 = (wv % SPLIT) * NBW;
    constexpr int MBW = 4 / SPLIT;  // column blocks in flight per workgroup
    const int c = lane & 15, g = lane >> 4;
    const size_t mbstride = (size_t)p.KCN * kLimbs * FRAG;
    const uint8_t *dk = dbl + (size_t)k * p.MB * mbstride + (size_t)kc0 * kLimbs * FRAG + (size_t)lane * KB;
    double *ok = out + fd_out_at(p, seg, k, 0, 0);
    // the wave's steps: (column block mb = w + 4 u, row chunk t) for u, t in order
    const int nmb = mb1 - mb0 - w > 0 ? (mb1 - mb0 - w + MBW - 1) / MBW : 0;
    const int steps = nmb > 0 ? nmb * nkc : 0;
    auto load = [&](T (&dst)[kLimbs], int s) {
        const int u = s / nkc, t = s - u * nkc;
        const uint8_t *da = dk + (size_t)(mb0 + w + MBW * u) * mbstride + (size_t)t * kLimbs * FRAG;
#pragma unroll
        for (int l = 0; l < kLimbs; l++) dst[l] = __builtin_nontemporal_load((const T *)(da + (size_t)l * FRAG));
    };
    fd_v4i acc[NBW][kLimbShifts];
#pragma unroll
    for (int nb = 0; nb < NBW; nb++)
#pragma unroll
        for (int sh = 0; sh < kLimbShifts; sh++) acc[nb][sh] = fd_v4i{0, 0, 0, 0};
    auto step = [&](const T (&a)[kLimbs], int s) {
        const int u = s / nkc, t = s - u * nkc;
        const uint8_t *bt = sb + (size_t)t * chunk;
#pragma unroll
        for (int nb = 0; nb < NBW; nb++) {
            const int qp = (nb0 + nb) * 16 + c;
#pragma unroll
            for (int bl = 0; bl < kLimbs; bl++) {
                const T b = qp < p.QP ? *(const T *)(bt + ((size_t)(bl * 4 + g) * p.QP + qp) * KB) : Op::zero();
#pragma unroll
                for (int al = 0; al < kLimbs; al++) acc[nb][al + bl] = Op::mac(a[al], b, acc[nb][al + bl]);
            }
        }
        if (t != nkc - 1) return;
        // lane (g, c) of block nb holds rows j = 16 mb + 4 g + [0, 4) of column qp
        const int j0 = (mb0 + w + MBW * u) * 16 + 4 * g;
#pragma unroll
        for (int nb = 0; nb < NBW; nb++) {
            const int qp = (nb0 + nb) * 16 + c;
            if (qp >= p.QP) continue;
            double v[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int32_t C[kLimbShifts];
#pragma unroll
                for (int sh = 0; sh < kLimbShifts; sh++) C[sh] = acc[nb][sh][r];
                v[r] = limb_fold(C, m);
            }
            double *o = ok + (size_t)qp * kFdKlo * p.J + j0;  // qp = 2q + part
            if (j0 + 4 <= p.J) {
                typedef double dbl2 __attribute__((ext_vector_type(2)));
                *(dbl2 *)o = dbl2{v[0], v[1]};
                *(dbl2 *)(o + 2) = dbl2{v[2], v[3]};
            } else {
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (j0 + r < p.J) o[r] = v[r];
            }
        }
#pragma unroll
        for (int nb = 0; nb < NBW; nb++)
#pragma unroll
            for (int sh = 0; sh < kLimbShifts; sh++) acc[nb][sh] = fd_v4i{0, 0, 0, 0};
    };
    // three fragment sets in rotation: the loads of steps s + 1 and s + 2 are
    // in flight while step s multiplies (across column blocks too)
    T f0[kLimbs], f1[kLimbs], f2[kLimbs];
    if (steps > 0) load(f0, 0);
    if (steps > 1) load(f1, 1);
    for (int s = 0; s < steps; s += 3) {
        if (s + 2 < steps) load(f2, s + 2);
        step(f0, s);
        if (s + 1 >= steps) break;
        if (s + 3 < steps) load(f0, s + 3);
        step(f1, s + 1);
        if (s + 2 >= steps) break;
        if (s + 4 < steps) load(f1, s + 4);
        step(f2, s + 2);
    }
}

template <int KC>
static void fd_mfma_kc(const uint8_t *dbl, const uint8_t *sell, double *out_k, const FdPlan &p, FMod m,
                       hipStream_t st) {
    // column-block groups per slot: with one selection block (its staging is
    // small) four, or on wide grids one group per 8 column blocks (two per
    // wavefront: measured best at J = 4096, 414 against 442 us with four
    // groups), else one (the staging is shared by every block)
    const int G = p.G > 0 ? std::min(p.G, p.MB) : (p.NB == 1 ? (p.MB >= 64 ? p.MB / 8 : p.MB >= 16 ? 4 : 1) : 1);
    const dim3 grid(p.n * G, p.nseg);
    switch (p.NB) {
        case 1: hipLaunchKernelGGL((k_fd_mfma<KC, 1, 1>), grid, dim3(256), 0, st, dbl, sell, out_k, p, m, G); break;
        case 2: hipLaunchKernelGGL((k_fd_mfma<KC, 2, 1>), grid, dim3(256), 0, st, dbl, sell, out_k, p, m, G); break;
        // three or four blocks: two wavefronts per column block (the block past
        // 2Q columns, if any, multiplies zeros)
        default: hipLaunchKernelGGL((k_fd_mfma<KC, 2, 2>), grid, dim3(256), 0, st, dbl, sell, out_k, p, m, G); break;
    }
}

void launch_fd_mfma(const uint8_t *dbl, const uint8_t *sell, double *out_k, const FdPlan &p, FMod m, hipStream_t st) {
    if (p.KC == 64) fd_mfma_kc<64>(dbl, sell, out_k, p, m, st);
    else fd_mfma_kc<32>(dbl, sell, out_k, p, m, st);
}

// The segment sum, the transpose and the inverse transform of the FirstDim
// outputs in one pass (first_dim.rs:174-181 inverts each output ciphertext):
// workgroup b takes IPP consecutive output polys idx = qp * J + j (qp = 2q +
// part), reads their n slots straight from the slot-major segment outputs
// (IPP * 8 bytes per slot row and segment), adds the segments, transforms
// the polys back in LDS and writes them coefficient form to [q][j][part][k].
// The ROWP / IPP workgroups that read one row of ROWP polys (8 ROWP bytes of
// each slot) get block ids congruent mod 8, i.e. one XCD and its L2, and run
// side by side (groups of 8 rows), so each row is fetched from HBM once, in
// runs of 8 ROWP bytes.
template <int IPP, int TPB, int ROWP>
__global__ __launch_bounds__(TPB) void k_fd_intt(const double *__restrict__ ok, double *__restrict__ out, FdPlan p,
                                                 RingDev r2) {
    __shared__ double s[lds_words(IPP * kSlots)];
    constexpr int LPW = ROWP / IPP;  // workgroups per row group of ROWP polys
    const size_t P = (size_t)p.Q * 2 * p.J;
    const int n = p.n;
    const size_t b = blockIdx.x, grp = b / (8 * LPW), wi = b % (8 * LPW);
    const size_t tile = grp * 8 * LPW + (wi & 7) * LPW + (wi >> 3);
    const size_t idx0 = tile * IPP;
    if (idx0 >= P) return;  // the whole workgroup, before any barrier
    const int np = (int)(P - idx0 < (size_t)IPP ? P - idx0 : (size_t)IPP);
    const FMod m = r2.t[0].m;
    if (p.nseg == 1 && np == IPP && n == kSlots) {
        // one segment, a whole tile: every load of the thread issued before
        // the first store (independent loads in flight)
        constexpr int PT = IPP * kSlots / TPB;
        // element e: slot e / IPP of poly idx0 + e % IPP (the same poly for every e of a thread)
        const size_t idx = idx0 + threadIdx.x % IPP;
        const double *src = ok + fd_out_at(p, 0, 0, (int)(idx / p.J), (int)(idx % p.J));
        double v[PT];
#pragma unroll
        for (int x = 0; x < PT; x++) {
            const int e = threadIdx.x + x * TPB;
            v[x] = src[fd_out_at(p, 0, e / IPP, 0, 0)];
        }
#pragma unroll
        for (int x = 0; x < PT; x++) {
            const int e = threadIdx.x + x * TPB;
            s[lds_pad((e % IPP) * n + e / IPP)] = v[x];
        }
    } else {
        for (int e = threadIdx.x; e < IPP * n; e += TPB) {
            const int pp = e % IPP, k = e / IPP;
            double v = 0.0;
            if (pp < np) {
                const size_t idx = idx0 + pp;
                for (int sg = 0; sg < p.nseg; sg++) v += ok[fd_out_at(p, sg, k, (int)(idx / p.J), (int)(idx % p.J))];
            }
            s[lds_pad(pp * n + k)] = mod_reduce_small(v, m);
        }
    }
    __syncthreads();
    lds_ntt_inv_noscale<1, 4>(s, IPP * n, r2);
    const double ni = r2.t[0].n_inv;
    for (int pp = 0; pp < np; pp++) {
        const size_t idx = idx0 + pp, qp = idx / p.J, j = idx % p.J;
        double *o = out + (((qp >> 1) * p.J + j) * 2 + (qp & 1)) * n;
        for (int k = threadIdx.x; k < n; k += TPB) o[k] = mod_mul(s[lds_pad(pp * n + k)], ni, m);
    }
}

// The same for one row segment and n = 2048 with fixed-shape passes: four
// polys per workgroup, 128 threads each. The slot-major outputs are staged
// into LDS transposed as in k_fd_intt (consecutive lanes on consecutive words
// of a slot row, every load in flight before the first store); inverse
// stages 0..3 and 4..7 run over LDS with compile-time geometry; thread u of
// poly pp then holds coefficients u + 256 t and u + 128 + 256 t (two units of
// stages 8..10) in registers, scales them by N^-1 and stores them. Row groups
// of 128 polys per XCD. The polys' LDS buffers are two words more than a
// buffer apart, so the four lanes that stage one slot of the four polys hit
// different banks.
__global__ __launch_bounds__(512) void k_fd_intt_fx(const double *__restrict__ ok, double *__restrict__ out, FdPlan p,
                                                    RingDev r2) {
    constexpr int IPP = 4, LPW = 128 / IPP, BUF = lds_words(kSlots) + 2, TPB = 512;
    __shared__ double s[IPP * BUF];
    const size_t P = (size_t)p.Q * 2 * p.J;
    const size_t b = blockIdx.x, grp = b / (8 * LPW), wi = b % (8 * LPW);
    const size_t tile = grp * 8 * LPW + (wi & 7) * LPW + (wi >> 3);
    const size_t idx0 = tile * IPP;
    if (idx0 >= P) return;  // the whole workgroup, before any barrier
    const FMod m = r2.t[0].m;
    const double *twi = r2.t[0].inv;
    {
        constexpr int PT = IPP * kSlots / TPB;
        const size_t idx = idx0 + threadIdx.x % IPP;
        const double *src = ok + fd_out_at(p, 0, 0, (int)(idx / p.J), (int)(idx % p.J));
        double v[PT];
#pragma unroll
        for (int x = 0; x < PT; x++) {
            const int e = threadIdx.x + x * TPB;
            v[x] = src[fd_out_at(p, 0, e / IPP, 0, 0)];
        }
#pragma unroll
        for (int x = 0; x < PT; x++) {
            const int e = threadIdx.x + x * TPB;
            s[(e % IPP) * BUF + lds_pad(e / IPP)] = v[x];  // canonical already
        }
    }
    __syncthreads();
    const int pp = threadIdx.x >> 7, u = threadIdx.x & 127;
    double *sp = s + pp * BUF;
    {
        // inverse stages 0..3 on slots 16u .. 16u+15 (block SB = u)
        double w[15];
        load_inv_unit_tw<kSlotsLog, 0, 4>(w, twi, u);
        double *q = sp + lds_pad(16 * u);
        double x[16];
#pragma unroll
        for (int t = 0; t < 16; t++) x[t] = q[t];
        inv_unit_stages<4>(x, w, m);
#pragma unroll
        for (int t = 0; t < 16; t++) q[t] = x[t];
    }
    __syncthreads();
    {
        // inverse stages 4..7
        int base, SB;
        InvGeo<kSlotsLog, 4, 4>::at(u, base, SB);
        double w[15];
        load_inv_unit_tw<kSlotsLog, 4, 4>(w, twi, SB);
        double *q = sp + lds_pad(base);
        double x[16];
#pragma unroll
        for (int t = 0; t < 16; t++) x[t] = q[unit_off<16>(t)];
        inv_unit_stages<4>(x, w, m);
#pragma unroll
        for (int t = 0; t < 16; t++) q[unit_off<16>(t)] = x[t];
    }
    __syncthreads();
    {
        // inverse stages 8..10 on coefficients c + 256 t for c = u and u + 128,
        // the N^-1 scaling, the store to [q][j][part][k]
        double w[7];
        load_inv_unit_tw<kSlotsLog, 8, 3>(w, twi, 0);
        const double ni = r2.t[0].n_inv;
        const size_t idx = idx0 + pp, qp = idx / p.J, j = idx % p.J;
        double *o = out + (((qp >> 1) * p.J + j) * 2 + (qp & 1)) * kSlots;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int c = u + 128 * h;
            const double *q = sp + lds_pad(c);
            double x[8];
#pragma unroll
            for (int t = 0; t < 8; t++) x[t] = q[unit_off<256>(t)];
            inv_unit_stages<3>(x, w, m);
#pragma unroll
            for (int t = 0; t < 8; t++) o[c + 256 * t] = mod_mul(x[t], ni, m);
        }
    }
}

// The FirstDim outputs in the form the first CMux level takes from them
// (cmux(b, c0, c1) = c0 + b (x) (c1 - c0), mux.rs:75-84, on columns
// 2i, 2i + 1): one inverse transform per pair instead of two. Workgroup b
// takes the polys of 2 NPR consecutive columns j (J a multiple of 2 NPR, so
// the same q and part), i.e. NPR pairs: thread t reads slot t / NPR + 128 x
// of pair t % NPR as one 16-byte (c0, c1) load (the workgroup reads
// 16 NPR-byte runs of each slot row; row groups of 128 polys per XCD as in
// k_fd_intt_fx), stores c0 in eval form to column 2i + 1 of out (runs of
// consecutive slots of each pair per wavefront store) and stages the
// difference c1 - c0 in LDS. The differences are transformed back as in
// k_fd_intt_fx (128 threads each: stages 0..3 and 4..7 over LDS, 8..10 in
// registers, N^-1) into column 2i. The gate adds c0 to its product before
// its own inverse transform. NPR = 8 (128-byte runs, 1024 threads, one
// workgroup per CU) reads faster than NPR = 4 (64-byte runs, two per CU).
template <int NPR>
__global__ __launch_bounds__(128 * NPR) void k_fd_intt_pairs(const double *__restrict__ ok, double *__restrict__ out,
                                                             FdPlan p, RingDev r2) {
    constexpr int IPP = 2 * NPR, LPW = 128 / IPP, BUF = lds_words(kSlots) + 2, TPB = 128 * NPR;
    constexpr int PT = NPR * kSlots / TPB;  // slots per thread (one pair)
    __shared__ double s[NPR * BUF];
    const size_t P = (size_t)p.Q * 2 * p.J;
    const size_t b = blockIdx.x, grp = b / (8 * LPW), wi = b % (8 * LPW);
    const size_t tile = grp * 8 * LPW + (wi & 7) * LPW + (wi >> 3);
    const size_t idx0 = tile * IPP;
    if (idx0 >= P) return;  // the whole workgroup, before any barrier
    const FMod m = r2.t[0].m;
    const double *twi = r2.t[0].inv;
    const size_t qp = idx0 / p.J, j0 = idx0 % p.J;
    const size_t q = qp >> 1, part = qp & 1;
    {
        const int pr = threadIdx.x % NPR, k0 = threadIdx.x / NPR;
        const double *src = ok + fd_out_at(p, 0, 0, (int)qp, (int)(j0 + 2 * pr));
        double *oe = out + ((q * p.J + j0 + 2 * pr + 1) * 2 + part) * kSlots;
        double2 v[PT];
#pragma unroll
        for (int x = 0; x < PT; x++)
            v[x] = *reinterpret_cast<const double2 *>(src + fd_out_at(p, 0, k0 + 128 * x, 0, 0));
#pragma unroll
        for (int x = 0; x < PT; x++) {
            oe[k0 + 128 * x] = v[x].x;
            s[pr * BUF + lds_pad(k0 + 128 * x)] = mod_sub(v[x].y, v[x].x, m);
        }
    }
    __syncthreads();
    const int pp = threadIdx.x >> 7, u = threadIdx.x & 127;
    double *sp = s + pp * BUF;
    {
        // inverse stages 0..3 on slots 16u .. 16u+15 (block SB = u)
        double w[15];
        load_inv_unit_tw<kSlotsLog, 0, 4>(w, twi, u);
        double *qq = sp + lds_pad(16 * u);
        double x[16];
#pragma unroll
        for (int t = 0; t < 16; t++) x[t] = qq[t];
        inv_unit_stages<4>(x, w, m);
#pragma unroll
        for (int t = 0; t < 16; t++) qq[t] = x[t];
    }
    __syncthreads();
    {
        // inverse stages 4..7
        int base, SB;
        InvGeo<kSlotsLog, 4, 4>::at(u, base, SB);
        double w[15];
        load_inv_unit_tw<kSlotsLog, 4, 4>(w, twi, SB);
        double *qq = sp + lds_pad(base);
        double x[16];
#pragma unroll
        for (int t = 0; t < 16; t++) x[t] = qq[unit_off<16>(t)];
        inv_unit_stages<4>(x, w, m);
#pragma unroll
        for (int t = 0; t < 16; t++) qq[unit_off<16>(t)] = x[t];
    }
    __syncthreads();
    const size_t jp = j0 + 2 * pp;  // this thread's pair: columns jp, jp + 1
    {
        // inverse stages 8..10 on coefficients c + 256 t for c = u and u + 128,
        // N^-1, the difference to column jp
        double w[7];
        load_inv_unit_tw<kSlotsLog, 8, 3>(w, twi, 0);
        const double ni = r2.t[0].n_inv;
        double *od = out + ((q * p.J + jp) * 2 + part) * kSlots;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int c = u + 128 * h;
            const double *qq = sp + lds_pad(c);
            double x[8];
#pragma unroll
            for (int t = 0; t < 8; t++) x[t] = qq[unit_off<256>(t)];
            inv_unit_stages<3>(x, w, m);
#pragma unroll
            for (int t = 0; t < 8; t++) od[c + 256 * t] = mod_mul(x[t], ni, m);
        }
    }
}

constexpr int kFdPairsNpr = 8;  // column pairs per workgroup of k_fd_intt_pairs

bool fd_pairs_fits(const FdPlan &p) { return p.nseg == 1 && p.n == kSlots && p.J % (2 * kFdPairsNpr) == 0; }

void launch_fd_intt_pairs(const double *out_k, double *out, const FdPlan &p, const RingDev &r2, hipStream_t st) {
    constexpr int IPP = 2 * kFdPairsNpr;
    const size_t P = (size_t)p.Q * 2 * p.J;
    const size_t g = (size_t)8 * 128 / IPP;  // whole XCD row groups of 128 polys
    const size_t blocks = ((P + IPP - 1) / IPP + g - 1) / g * g;
    hipLaunchKernelGGL(k_fd_intt_pairs<kFdPairsNpr>, dim3((unsigned)blocks), dim3(128 * kFdPairsNpr), 0, st, out_k,
                       out, p, r2);
}

// Coefficient form [q][j][part][n] of query 0's FirstDim outputs from the
// pair form k_fd_intt_pairs writes (c0 = iNTT(column 2i + 1), c1 = c0 +
// column 2i), for inspection: one workgroup per (pair, part).
__global__ __launch_bounds__(256) void k_fd_unpair(const double *in, double *out, int J, RingDev r2) {
    __shared__ double s[lds_words(kSlots)];
    const int i = blockIdx.x >> 1, part = blockIdx.x & 1, n = kSlots;
    const double *e0 = in + ((size_t)(2 * i + 1) * 2 + part) * n;
    const double *d = in + ((size_t)(2 * i) * 2 + part) * n;
    for (int k = threadIdx.x; k < n; k += blockDim.x) s[lds_pad(k)] = e0[k];
    __syncthreads();
    lds_ntt_inv_noscale<1, 4>(s, n, r2);
    __syncthreads();
    const FMod m = r2.t[0].m;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const double c0 = mod_mul(s[lds_pad(k)], r2.t[0].n_inv, m);
        out[((size_t)(2 * i) * 2 + part) * n + k] = c0;
        out[((size_t)(2 * i + 1) * 2 + part) * n + k] = mod_add(c0, d[k], m);
    }
    (void)J;
}

void launch_fd_unpair(const double *in, double *out, int J, const RingDev &r2, hipStream_t st) {
    hipLaunchKernelGGL(k_fd_unpair, dim3((unsigned)J), dim3(256), 0, st, in, out, J, r2);
}

void launch_fd_intt(const double *out_k, double *out, const FdPlan &p, const RingDev &r2, hipStream_t st) {
    const size_t P = (size_t)p.Q * 2 * p.J;
    // four polys per workgroup when that still gives a thousand workgroups,
    // else fewer, so small passes fill the 256 CUs
    // (grids in whole XCD groups of 8 * ROWP / ipp workgroups)
    auto blocks = [&](int ipp, int rowp) {
        const size_t g = (size_t)8 * rowp / ipp;
        return ((P + ipp - 1) / ipp + g - 1) / g * g;
    };
    if (P >= 4 * 1024 && p.nseg == 1 && p.n == kSlots && P % 4 == 0)
        hipLaunchKernelGGL(k_fd_intt_fx, dim3((unsigned)blocks(4, 128)), dim3(512), 0, st, out_k, out, p, r2);
    else if (P >= 4 * 1024)
        hipLaunchKernelGGL((k_fd_intt<4, 512, 64>), dim3((unsigned)blocks(4, 64)), dim3(512), 0, st, out_k, out, p,
                           r2);
    else if (P >= 2 * 512)
        hipLaunchKernelGGL((k_fd_intt<2, 256, 16>), dim3((unsigned)blocks(2, 16)), dim3(256), 0, st, out_k, out, p,
                           r2);
    else
        hipLaunchKernelGGL((k_fd_intt<1, 256, 16>), dim3((unsigned)blocks(1, 16)), dim3(256), 0, st, out_k, out, p,
                           r2);
}

}  // namespace via
