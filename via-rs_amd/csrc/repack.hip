// repack.hip — the VIA-B repack (conversion/repack.rs:331-369) on gfx950.
//
// The T post-CRot ciphertexts at (n1, q2) become one RLWE through log2(D)
// pair conversions, D = n1/G. Level l holds (rank n1/nin, degree nin) MLWEs
// with nin = G * 2^l; converting a pair interleaves the two inputs at slots
// 0 and 1 of degree nout = 2 nin and key-switches its rank_in masks with the
// cascade step keys of output degree nout. Every level therefore has
// rank_in * nout = 2 n1 mask coefficients per pair: two 2048-slot tiles,
// the shape the cascade's gadget-product kernel already takes.
//
// Only pairs that hold a real input are computed. The T inputs sit at
// entries 0..T-1 after the bit-reversal reorder, the padding is zero, and the
// conversion of a pair of zero MLWEs is exactly zero (zero digits, zero
// products), so level l has max(1, T >> (l+1)) live pairs.
#include "digits.h"
#include "ntt_device.h"

namespace via {

__device__ __forceinline__ uint32_t bit_reverse_bits(uint32_t i, int bits) {
    return bits ? (__brev(i) >> (32 - bits)) : 0u;
}

// Extr_G of each input (extract.rs:61-75), bit-reversal reorder
// (repack.rs:346-353), zero padding, and the level-0 pair interleave
// (embed_d, repack.rs:72-104): digits of every interleaved mask coefficient
// and the interleaved body. Workgroup (p, y) covers 256 of pair p's 2 n1
// mask slots; slot u = idx * nout + c holds coefficient c of mask idx.
__global__ __launch_bounds__(256) void k_repack_extract(RingDev r, Gadget g, DigitPlan dp, const double *rot, int T,
                                                        int lt, int G, double *body0, size_t body_stride,
                                                        int16_t *dig) {
    const int p = blockIdx.x;
    const int n1 = r.n;
    const int D = n1 / G;
    const int nout = 2 * G;
    const int lognout = __builtin_ctz(nout);
    const FMod m = r.t[0].m;
    const int slots = 2 * n1;
    const int u = blockIdx.y * blockDim.x + threadIdx.x;
    if (u < slots) {
        const int idx = u >> lognout, c = u & (nout - 1);
        const int s = c & 1, i = c >> 1;
        const int mem = 2 * p + s;
        double val = 0.0;
        if (mem < T) {
            const double *A = rot + (size_t)bit_reverse_bits((uint32_t)mem, lt) * 2 * n1;
            if (idx == 0)
                val = A[D * i];
            else if (i == 0)
                val = mod_neg(A[D * (G - 1) + D - idx], m);  // pi_{D-idx}(A) * X, wrapped term
            else
                val = A[D * (i - 1) + D - idx];
        }
        write_digits<1>(val, 0.0, r, g, dp, dig + (size_t)p * g.L * slots + u, (size_t)slots);
    }
    if (blockIdx.y == 0) {
        for (int c = threadIdx.x; c < nout; c += blockDim.x) {
            const int s = c & 1, i = c >> 1;
            const int mem = 2 * p + s;
            double val = 0.0;
            if (mem < T) val = rot[(size_t)bit_reverse_bits((uint32_t)mem, lt) * 2 * n1 + n1 + D * i];
            body0[(size_t)p * body_stride + c] = val;
        }
    }
}

// Wide gadgets take the fp64 split of gadget_split_w where its bounds hold.
static Gadget with_wide_split(const RingDev &r, const Gadget &g_in, const DigitPlan &dp) {
    Gadget g = g_in;
    if (g.f64 == 0 && g.bl >= ((via_u128)1 << 40) && scale_w_make(r, g, dp, &g.sw)) g.f64 = 3;
    return g;
}

void launch_repack_extract(const RingDev &r, const Gadget &g_in, const DigitPlan &dp, const double *rot, int T, int G,
                           int pairs, double *body0, size_t body_stride, int16_t *dig, hipStream_t st) {
    const Gadget g = with_wide_split(r, g_in, dp);
    const int lt = __builtin_ctz((unsigned)T);
    dim3 grid(pairs, (2 * r.n + 255) / 256);
    hipLaunchKernelGGL(k_repack_extract, grid, dim3(256), 0, st, r, g, dp, rot, T, lt, G, body0, body_stride, dig);
}

// One level's combine (the reduction of mlwes_to_mlwe, repack.rs:169-187):
// output mask j = -iNTT(sum of the switched masks of idx = j and
// idx = rank_out + j), output body = interleaved body - iNTT(sum of every
// switched body). Non-final levels write the results straight into the next
// level's pair interleave: digits of output mask j of pair p land at slot
// j * 2n + 2c + (p & 1) of pair p / 2, and the body at 2c + (p & 1). With a
// single live pair the partner entries are written as zero. The final level
// writes the RLWE (mask, body). Workgroup (p, chunk) handles a chunk of the
// n1 mask slots; chunk 0 also the body.
__global__ __launch_bounds__(kThreads) void k_repack_combine(RingDev r, Gadget g, DigitPlan dp, RepackCombineArgs a) {
    __shared__ double s[lds_words(kSlots)];
    const int item = blockIdx.x, chunk = blockIdx.y, chunks = gridDim.y;
    const int S = kSlots, n = r.n;
    const int CS = S / chunks;  // a multiple of n
    const int c0 = chunk * CS;
    const FMod m = r.t[0].m;
    const NttLimbs T = limbs_inv(r);
    const double ninv = r.t[0].n_inv;
    const size_t gstride = (size_t)a.items * 2 * S;
    for (int t = threadIdx.x; t < CS; t += blockDim.x) {
        const double *p = a.pm + (size_t)item * 2 * S + c0 + t;
        double acc = mod_add(p[0], p[S], m);
        for (int gi = 1; gi < a.groups; gi++) acc = mod_add(acc, mod_add(p[gi * gstride], p[gi * gstride + S], m), m);
        s[lds_pad(t)] = acc;
    }
    __syncthreads();
    ntt_inv_noscale<1>(s, CS, n, r.logn, T);
    const int logn = r.logn;
    const size_t next_slots = 2 * (size_t)S;
    int16_t *dn = a.dig_next ? a.dig_next + (size_t)(item >> 1) * g.L * next_slots : nullptr;
    for (int t = threadIdx.x; t < CS; t += blockDim.x) {
        const double x = mod_neg(mod_mul(s[lds_pad(t)], ninv, m), m);
        const int u = c0 + t;
        if (a.final_level) {
            a.out[u] = x;
        } else {
            const int j = u >> logn, c = u & (n - 1);
            const size_t un = (size_t)j * 2 * n + 2 * c + (item & 1);
            write_digits<1>(x, 0.0, r, g, dp, dn + un, next_slots);
            if (a.zero_partner)
                for (int k = 0; k < g.L; k++) dn[(size_t)k * next_slots + un + 1] = 0;
        }
    }
    if (chunk != 0) return;
    __syncthreads();
    const size_t gstride_b = (size_t)a.items * 2 * n;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const double *p = a.pb + (size_t)item * 2 * n + t;
        double acc = mod_add(p[0], p[n], m);
        for (int gi = 1; gi < a.groups; gi++)
            acc = mod_add(acc, mod_add(p[gi * gstride_b], p[gi * gstride_b + n], m), m);
        s[lds_pad(t)] = acc;
    }
    __syncthreads();
    ntt_inv_noscale<1>(s, n, n, r.logn, T);
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const double x = mod_mul(s[lds_pad(t)], ninv, m);
        const double v = mod_sub(a.body_cur[(size_t)item * a.body_stride + t], x, m);
        if (a.final_level) {
            a.out[S + t] = v;
        } else {
            double *bn = a.body_next + (size_t)(item >> 1) * a.body_stride;
            bn[2 * t + (item & 1)] = v;
            if (a.zero_partner) bn[2 * t + 1] = 0.0;
        }
    }
}

void launch_repack_combine(const RingDev &r, const Gadget &g_in, const DigitPlan &dp, const RepackCombineArgs &a,
                           hipStream_t st) {
    const Gadget g = with_wide_split(r, g_in, dp);
    int chunks = kSlots / r.n;
    if (chunks > 8) chunks = 8;
    dim3 grid(a.items, chunks);
    hipLaunchKernelGGL(k_repack_combine, grid, dim3(kThreads), 0, st, r, g, dp, a);
}

// Relayout of a step key already at a single modulus, [idx][k][part][n] ->
// [k][part][idx*n + c], the layout the gadget-product kernel reads.
__global__ void k_key_relayout_q2(const double *src, double *dst, int rank_in, int L, int n) {
    const size_t total = (size_t)rank_in * L * 2 * n;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        size_t x = t;
        const int c = (int)(x % n);
        x /= n;
        const int part = (int)(x % 2);
        x /= 2;
        const int k = (int)(x % L);
        const size_t idx = x / L;
        dst[((size_t)k * 2 + part) * ((size_t)rank_in * n) + idx * n + c] = src[t];
    }
}

void launch_key_relayout_q2(const double *src, double *dst, int rank_in, int L, int n, hipStream_t st) {
    const size_t total = (size_t)rank_in * L * 2 * n;
    unsigned blocks = (unsigned)((total + 255) / 256);
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_key_relayout_q2, dim3(blocks), dim3(256), 0, st, src, dst, rank_in, L, n);
}

}  // namespace via
