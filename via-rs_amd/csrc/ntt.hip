// ntt.hip — batched negacyclic NTT, u64 <-> f64 conversion and the batched
// negacyclic product (RingPolyEval of via-rs ring/ntt_abstraction.rs:61-78).
#include "ntt_fixed.h"

namespace via {

// One workgroup transforms kSlots/n polys of one limb (or one poly when
// n == kSlots). The polys of a workgroup share a limb (checked by the host).
// Radix-16 passes: these kernels hold little else in registers. Polys come in
// groups of gpolys (group g's poly r at src + g*src_gstride + r*n, written to
// data + g*dst_gstride + r*n): several queries' arrays in one launch.
template <bool INV>
__global__ __launch_bounds__(kThreads) void k_ntt_batch(RingDev r, const double *src, double *data, size_t npolys,
                                                        size_t limb_group, size_t gpolys, size_t src_gstride,
                                                        size_t dst_gstride) {
    __shared__ double s[lds_words(kSlots)];
    const int n = r.n;
    const int P = kSlots / n;
    const size_t p0 = (size_t)blockIdx.x * P;
    const int cnt = (int)((npolys - p0) < (size_t)P ? (npolys - p0) : (size_t)P);
    const size_t gq = p0 / gpolys, pr = p0 - gq * gpolys;
    const int l = (int)((pr / limb_group) % (size_t)r.nl);
    RingDev r1 = r;
    if (l == 1) r1.t[0] = r.t[1];
    double *g = data + gq * dst_gstride + pr * n;
    const double *gs = src + gq * src_gstride + pr * n;
    const int S = cnt * n;
    for (int i = threadIdx.x; i < S; i += blockDim.x) s[lds_pad(i)] = gs[i];
    __syncthreads();
    if (!INV) {
        lds_ntt_fwd<1, 4>(s, S, r1);
        for (int i = threadIdx.x; i < S; i += blockDim.x) g[i] = mod_reduce_small(s[lds_pad(i)], r1.t[0].m);
    } else {
        lds_ntt_inv_noscale<1, 4>(s, S, r1);
        const double ni = r1.t[0].n_inv;
        for (int i = threadIdx.x; i < S; i += blockDim.x) g[i] = mod_mul(s[lds_pad(i)], ni, r1.t[0].m);
    }
}

// The exact q1 -> q2 switch of one coefficient: round(x q2 / q1) mod q2
// for x the CRT value of its limbs (mod_switch.rs:57-86).
__device__ double switch_exact(double a0, double a1, const SwitchIn &si) {
    const via_u128 x = rns_reconstruct(a0, a1, si.rns);
    uint64_t y = (uint64_t)round_mul_div(x, (via_u128)si.q2, si.qd);
    if (y >= si.q2) y -= si.q2;
    return (double)y;
}

// Forward transforms at n = 2048 with compile-time passes: two polys per
// workgroup, 128 threads each (the halves may take different limbs). Thread
// u reads elements u + 128 t of its poly straight into stages 0..3, runs
// stages 4..7 over LDS, and holds slots 8 v .. 8 v + 7 for v = u, u + 128 in
// the last pass, which it stores canonical. Polys come in groups as in
// k_ntt_batch; src may equal dst (every poly is read before it is written).
// SW: the polys are two-limb q1 polys (poly r of group g with its limbs at
// src + g*src_gstride + 2 n r and n words further) switched to q2 as they
// are loaded (mod_switch.rs:57-86), the switch of k_modswitch_q1_q2 without
// its intermediate array.
template <bool SW>
__global__ __launch_bounds__(kThreads) void k_ntt_fx(RingDev r, const double *src, double *dst, size_t npolys,
                                                     size_t limb_group, size_t gpolys, size_t src_gstride,
                                                     size_t dst_gstride, SwitchIn si) {
    constexpr int BUF = lds_words(kSlots);
    __shared__ double s[2 * BUF];
    const int h = threadIdx.x >> 7, u = threadIdx.x & 127;
    const size_t p = (size_t)blockIdx.x * 2 + h;
    const bool live = p < npolys;  // the same for a whole wavefront
    const size_t gq = live ? p / gpolys : 0, pr = live ? p - gq * gpolys : 0;
    const int l = __builtin_amdgcn_readfirstlane((int)((pr / limb_group) % (size_t)r.nl));
    const FMod m = r.t[l].m;
    const double *twf = r.t[l].fwd;
    double *sp = s + h * BUF;
    if (live) {
        double x[16];
        if constexpr (SW) {
            const double *g = src + gq * src_gstride + pr * (2 * kSlots) + u;
            unsigned bad = si.fast ? 0u : 0xffffu;
            if (si.fast) {
#pragma unroll
                for (int t = 0; t < 16; t++)
                    if (!modswitch_f64(g[128 * t], g[kSlots + 128 * t], si.rns, si.sf, x[t])) bad |= 1u << t;
            }
            if (bad) {  // rare: the exact switch, one coefficient at a time, through this thread's LDS words
                double *q = sp + lds_pad(u);
#pragma unroll
                for (int t = 0; t < 16; t++) q[unit_off<128>(t)] = x[t];
                while (bad) {
                    const int t = __builtin_ctz(bad);
                    bad &= bad - 1;
                    q[unit_off<128>(t)] = switch_exact(g[128 * t], g[kSlots + 128 * t], si);
                }
#pragma unroll
                for (int t = 0; t < 16; t++) x[t] = q[unit_off<128>(t)];
            }
        } else {
            const double *g = src + gq * src_gstride + pr * kSlots + u;
#pragma unroll
            for (int t = 0; t < 16; t++) x[t] = g[128 * t];
        }
        double w[15];
        load_unit_tw<4, 0>(w, twf, 0);
        unit_stages<4>(x, w, m);
        double *q = sp + lds_pad(u);
#pragma unroll
        for (int t = 0; t < 16; t++) q[unit_off<128>(t)] = x[t];
    }
    __syncthreads();
    if (live) {
        int base, SB;
        UnitGeo<kSlotsLog, 4, 4>::at(u, base, SB);
        double w[15];
        load_unit_tw<4, 4>(w, twf, SB);
        double *q = sp + lds_pad(base);
        double x[16];
#pragma unroll
        for (int t = 0; t < 16; t++) x[t] = q[unit_off<8>(t)];
        unit_stages<4>(x, w, m);
#pragma unroll
        for (int t = 0; t < 16; t++) q[unit_off<8>(t)] = x[t];
    }
    __syncthreads();
    if (live) {
        double *g = dst + gq * dst_gstride + pr * kSlots;
#pragma unroll
        for (int v = 0; v < 2; v++) {
            const int unit = u + 128 * v;
            double w[7];
            load_unit_tw<3, 8>(w, twf, unit);
            const double *q = sp + lds_pad(8 * unit);
            double x[8];
#pragma unroll
            for (int t = 0; t < 8; t++) x[t] = q[t];
            unit_stages<3>(x, w, m);
            double2 *o = reinterpret_cast<double2 *>(g + 8 * unit);
#pragma unroll
            for (int t = 0; t < 4; t++) o[t] = make_double2(mod_reduce_small(x[2 * t], m), mod_reduce_small(x[2 * t + 1], m));
        }
    }
}

static void launch_ntt_fwd(const RingDev &r, const double *src, double *dst, size_t npolys, size_t limb_group,
                           size_t gpolys, size_t src_gstride, size_t dst_gstride, hipStream_t st) {
    if (r.n == kSlots)
        hipLaunchKernelGGL(k_ntt_fx<false>, dim3((unsigned)((npolys + 1) / 2)), dim3(kThreads), 0, st, r, src, dst,
                           npolys, limb_group, gpolys, src_gstride, dst_gstride, SwitchIn{});
    else
        hipLaunchKernelGGL(k_ntt_batch<false>, dim3((unsigned)((npolys * r.n + kSlots - 1) / kSlots)), dim3(kThreads), 0,
                           st, r, src, dst, npolys, limb_group, gpolys, src_gstride, dst_gstride);
}

void launch_ntt(const RingDev &r, double *data, size_t npolys, size_t limb_group, bool inverse, hipStream_t st) {
    if (!npolys) return;
    const size_t P = kSlots / r.n;
    const size_t blocks = (npolys + P - 1) / P;
    if (inverse)
        hipLaunchKernelGGL(k_ntt_batch<true>, dim3(blocks), dim3(kThreads), 0, st, r, data, data, npolys, limb_group,
                           npolys, (size_t)0, (size_t)0);
    else
        launch_ntt_fwd(r, data, data, npolys, limb_group, npolys, 0, 0, st);
}

void launch_ntt_to(const RingDev &r, const double *src, double *dst, size_t npolys, size_t limb_group,
                   hipStream_t st) {
    if (!npolys) return;
    launch_ntt_fwd(r, src, dst, npolys, limb_group, npolys, 0, 0, st);
}

bool ntt_switch_fits(const RingDev &r1, const RingDev &r2) { return r1.nl == 2 && r2.nl == 1 && r2.n == kSlots && r1.n == kSlots; }

void launch_ntt_switch_groups(const RingDev &r1, const RingDev &r2, uint64_t q2, const double *src, size_t src_gstride,
                              double *dst, size_t dst_gstride, size_t gpolys, size_t groups, hipStream_t st) {
    const size_t npolys = gpolys * groups;
    if (!npolys) return;
    hipLaunchKernelGGL(k_ntt_fx<true>, dim3((unsigned)((npolys + 1) / 2)), dim3(kThreads), 0, st, r2, src, dst,
                       npolys, (size_t)1, gpolys, src_gstride, dst_gstride, switch_in(r1, q2));
}

void launch_ntt_switch(const RingDev &r1, const RingDev &r2, uint64_t q2, const double *src, double *dst,
                       size_t npolys, hipStream_t st) {
    launch_ntt_switch_groups(r1, r2, q2, src, 0, dst, 0, npolys, 1, st);
}

void launch_ntt_groups(const RingDev &r, const double *src, double *dst, size_t gpolys, size_t groups,
                       size_t src_gstride, size_t dst_gstride, size_t limb_group, hipStream_t st) {
    const size_t npolys = gpolys * groups;
    if (!npolys) return;
    if (r.n == kSlots) {  // every poly addressed by its own group
        launch_ntt_fwd(r, src, dst, npolys, limb_group, gpolys, src_gstride, dst_gstride, st);
        return;
    }
    const size_t P = kSlots / r.n;
    if (gpolys % P) {  // a workgroup's polys must not straddle two groups: one launch per group
        for (size_t g = 0; g < groups; g++)
            hipLaunchKernelGGL(k_ntt_batch<false>, dim3((gpolys + P - 1) / P), dim3(kThreads), 0, st, r,
                               src + g * src_gstride, dst + g * dst_gstride, gpolys, limb_group, gpolys, (size_t)0,
                               (size_t)0);
        return;
    }
    hipLaunchKernelGGL(k_ntt_batch<false>, dim3(npolys / P), dim3(kThreads), 0, st, r, src, dst, npolys, limb_group,
                       gpolys, src_gstride, dst_gstride);
}

__global__ void k_u64_to_f64(const uint64_t *src, double *dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (double)src[i];
}

__global__ void k_f64_to_u64(const double *src, uint64_t *dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (uint64_t)src[i];
}

static unsigned grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 8192 ? 8192 : (b ? b : 1));
}

// LWE arrays of `count` queries to fp64. Word i of an LWE (n + 1 words per
// limb) is a residue of limb (i mod lw) / (n + 1); a word at or above its
// modulus (the reference's Zq types hold canonical residues only) is taken
// as 0 and raises *bad, so no value outside the exactness bounds of the fp64
// arithmetic enters the pipeline.
__global__ void k_u64_to_f64_multi(U64Srcs srcs, double *dst, size_t words, uint64_t q0, uint64_t q1, int n1p,
                                   int *bad) {
    const uint64_t *src = srcs.p[blockIdx.y];
    double *d = dst + (size_t)blockIdx.y * words;
    const unsigned lw = 2 * n1p;
    // one query's words (below 2^32): 32-bit index arithmetic
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < (unsigned)words; i += gridDim.x * blockDim.x) {
        const uint64_t x = src[i];
        const uint64_t q = i % lw < (unsigned)n1p ? q0 : q1;
        if (x >= q) {
            *bad = 1;
            d[i] = 0.0;
        } else {
            d[i] = (double)x;
        }
    }
}

void launch_u64_to_f64_multi(const U64Srcs &srcs, int count, double *dst, size_t words, const uint64_t q[2], int n1,
                             int *bad, hipStream_t st) {
    if (!words || count < 1) return;
    const unsigned bx = std::max(1u, grid_for(words) / (unsigned)count);
    hipLaunchKernelGGL(k_u64_to_f64_multi, dim3(bx, count), dim3(256), 0, st, srcs, dst, words, q[0], q[1], n1 + 1,
                       bad);
}

void launch_u64_to_f64(const uint64_t *src, double *dst, size_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_u64_to_f64, dim3(grid_for(n)), dim3(256), 0, st, src, dst, n);
}

void launch_f64_to_u64(const double *src, uint64_t *dst, size_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_f64_to_u64, dim3(grid_for(n)), dim3(256), 0, st, src, dst, n);
}

// c = iNTT(NTT(a) (.) NTT(b)) for one poly per workgroup (n <= kSlots/2 per
// operand in LDS: both operands share the buffer as two "limbs" of one
// modulus).
__global__ __launch_bounds__(kThreads) void k_polymul(RingDev r, const uint64_t *a, const uint64_t *b, uint64_t *c,
                                                      size_t limb_group, const uint32_t *bidx) {
    __shared__ double s[lds_words(2 * kSlots)];
    const int n = r.n;
    const size_t p = blockIdx.x;
    const int l = (int)((p / limb_group) % (size_t)r.nl);
    RingDev r2 = r;
    r2.t[0] = l ? r.t[1] : r.t[0];
    r2.t[1] = r2.t[0];
    // bidx: operand b of product p / nl is b's entry bidx[p / nl] (limb p % nl)
    const size_t pb = bidx ? (size_t)bidx[p / r.nl] * r.nl + p % r.nl : p;
    const uint64_t *ga = a + p * n, *gb = b + pb * n;
    const FMod md = r2.t[0].m;
    const double ni = r2.t[0].n_inv;
    uint64_t *gc = c + p * n;
    if (n == kSlots) {
        // the common degree: every element phase unrolled, all loads of a
        // phase in flight together, canonical residues converted exactly
        // through the 2^52 bias (two instructions each way)
        constexpr int PT = kSlots / kThreads;
        uint64_t va[PT], vb[PT];
#pragma unroll
        for (int x = 0; x < PT; x++) {
            va[x] = ga[threadIdx.x + x * kThreads];
            vb[x] = gb[threadIdx.x + x * kThreads];
        }
#pragma unroll
        for (int x = 0; x < PT; x++) {
            const int i = threadIdx.x + x * kThreads;
            s[lds_pad(i)] = small_u64_to_f64(va[x]);
            s[lds_pad(kSlots + i)] = small_u64_to_f64(vb[x]);
        }
        __syncthreads();
        ntt_fwd<2, 4>(s, kSlots, kSlots, kSlotsLog, limbs_fwd(r2));
#pragma unroll
        for (int x = 0; x < PT; x++) {
            const int i = threadIdx.x + x * kThreads;
            s[lds_pad(i)] = mod_mul(mod_reduce_small(s[lds_pad(i)], md), s[lds_pad(kSlots + i)], md);
        }
        __syncthreads();
        ntt_inv_noscale<1, 4>(s, kSlots, kSlots, kSlotsLog, limbs_inv(r2));
#pragma unroll
        for (int x = 0; x < PT; x++) {
            const int i = threadIdx.x + x * kThreads;
            gc[i] = small_f64_to_u64(mod_mul(s[lds_pad(i)], ni, md));
        }
        return;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        s[lds_pad(i)] = (double)ga[i];
        s[lds_pad(n + i)] = (double)gb[i];
    }
    __syncthreads();
    lds_ntt_fwd<2, 4>(s, n, r2);
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        s[lds_pad(i)] = mod_mul(mod_reduce_small(s[lds_pad(i)], md), s[lds_pad(n + i)], md);
    __syncthreads();
    lds_ntt_inv_noscale<1, 4>(s, n, r2);
    for (int i = threadIdx.x; i < n; i += blockDim.x) gc[i] = (uint64_t)mod_mul(s[lds_pad(i)], ni, md);
}

// The same product at n = 2048 with the first and last passes of every
// transform in registers: thread j takes, in the forward transforms, unit
// j & 127 of operand j >> 7 (elements u + 128 t, read straight from memory)
// and then slots 8j .. 8j+7 of both operands (the last forward pass), where
// it forms the pointwise product and runs the first three inverse stages;
// the inverse ends on threads j < 128 holding coefficients j + 128 t, which
// they scale and store. Four barriers per product (radix-16 middle passes).
__global__ __launch_bounds__(kThreads) void k_polymul_fx(RingDev r, const uint64_t *a, const uint64_t *b,
                                                         uint64_t *c, size_t limb_group, const uint32_t *bidx) {
    constexpr int BUF = lds_words(kSlots);
    __shared__ double s[2 * BUF];
    const size_t p = blockIdx.x;
    const int l = (int)((p / limb_group) % (size_t)r.nl);
    const size_t pb = bidx ? (size_t)bidx[p / r.nl] * r.nl + p % r.nl : p;
    const FMod m = r.t[l].m;
    const double *twf = r.t[l].fwd, *twi = r.t[l].inv;
    const int j = threadIdx.x;
    const int op = j >> 7, u = j & 127;  // the forward passes' operand and unit
    {
        // forward stages 0..3: elements u + 128 t of the operand
        const uint64_t *src = (op ? b + pb * kSlots : a + p * kSlots) + u;
        double x[16];
#pragma unroll
        for (int t = 0; t < 16; t++) x[t] = small_u64_to_f64(src[128 * t]);
        double w[15];
        load_unit_tw<4, 0>(w, twf, 0);
        unit_stages<4>(x, w, m);
        double *q = s + op * BUF + lds_pad(u);
#pragma unroll
        for (int t = 0; t < 16; t++) q[unit_off<128>(t)] = x[t];
    }
    __syncthreads();
    {
        // forward stages 4..7
        int base, SB;
        UnitGeo<kSlotsLog, 4, 4>::at(u, base, SB);
        double w[15];
        load_unit_tw<4, 4>(w, twf, SB);
        double *q = s + op * BUF + lds_pad(base);
        double x[16];
#pragma unroll
        for (int t = 0; t < 16; t++) x[t] = q[unit_off<8>(t)];
        unit_stages<4>(x, w, m);
#pragma unroll
        for (int t = 0; t < 16; t++) q[unit_off<8>(t)] = x[t];
    }
    __syncthreads();
    double y[8];
    {
        // forward stages 8..10 of both operands on slots 8j .. 8j+7, the
        // pointwise product, inverse stages 0..2 on the same slots
        double w[7];
        load_unit_tw<3, 8>(w, twf, j);
        double xa[8], xb[8];
        const double *qa = s + lds_pad(8 * j), *qb = s + BUF + lds_pad(8 * j);
#pragma unroll
        for (int t = 0; t < 8; t++) {
            xa[t] = qa[t];
            xb[t] = qb[t];
        }
        unit_stages<3>(xa, w, m);
        unit_stages<3>(xb, w, m);
#pragma unroll
        // lazy product of the two lazy transforms (|x| < 24q): |xa xb| / q stays
        // below 2^47, so the fp64 quotient estimate is within one and y lies in
        // [-q, 2q), the inverse transform's input range
        for (int t = 0; t < 8; t++) y[t] = mod_mul_lazy(xa[t], xb[t], m);
        double wi[7];
        load_inv_unit_tw<kSlotsLog, 0, 3>(wi, twi, j);
        inv_unit_stages<3>(y, wi, m);
    }
    __syncthreads();  // every thread has read both operands before the product overwrites operand a
    {
        double *q = s + lds_pad(8 * j);
#pragma unroll
        for (int t = 0; t < 8; t++) q[t] = y[t];
    }
    __syncthreads();
    if (j < 128) {
        // inverse stages 3..6 (one product, 128 units of 16)
        int base, SB;
        InvGeo<kSlotsLog, 3, 4>::at(j, base, SB);
        double w[15];
        load_inv_unit_tw<kSlotsLog, 3, 4>(w, twi, SB);
        double *q = s + lds_pad(base);
        double x[16];
#pragma unroll
        for (int t = 0; t < 16; t++) x[t] = q[unit_off<8>(t)];
        inv_unit_stages<4>(x, w, m);
#pragma unroll
        for (int t = 0; t < 16; t++) q[unit_off<8>(t)] = x[t];
    }
    __syncthreads();
    if (j < 128) {
        // inverse stages 7..10 on coefficients j + 128 t, the N^-1 scaling, the store
        double w[15];
        load_inv_unit_tw<kSlotsLog, 7, 4>(w, twi, 0);
        const double *q = s + lds_pad(j);
        double x[16];
#pragma unroll
        for (int t = 0; t < 16; t++) x[t] = q[unit_off<128>(t)];
        inv_unit_stages<4>(x, w, m);
        const double ni = r.t[l].n_inv;
        uint64_t *dst = c + p * kSlots + j;
#pragma unroll
        for (int t = 0; t < 16; t++) dst[128 * t] = small_f64_to_u64(mod_mul(x[t], ni, m));
    }
}

void launch_polymul_u64(const RingDev &r, const uint64_t *a, const uint64_t *b, uint64_t *c, size_t count,
                        size_t limb_group, hipStream_t st, const uint32_t *bidx) {
    if (!count) return;
    if (r.n == kSlots)
        hipLaunchKernelGGL(k_polymul_fx, dim3(count), dim3(kThreads), 0, st, r, a, b, c, limb_group, bidx);
    else
        hipLaunchKernelGGL(k_polymul, dim3(count), dim3(kThreads), 0, st, r, a, b, c, limb_group, bidx);
}

}  // namespace via
