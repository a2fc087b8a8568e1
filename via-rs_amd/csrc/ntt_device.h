// ntt_device.h — LDS-resident negacyclic NTT for workgroup-cooperative kernels.
//
// Forward: iterative Cooley-Tukey decimation in time, natural order in and
// bit-reversed order out, twiddle of block b in stage m = psi^br(m+b)
// (via-rs ring/ntt.rs:305-343). Inverse: decimation in frequency from
// bit-reversed order, then the N^-1 scaling (ring/ntt.rs:356-399). Both are
// exact, so the result is bit-identical to the reference; the stages run in
// the reference order, two at a time in registers (radix-4 passes, see
// ntt_index.h) with one barrier per pass.
//
// The LDS buffer holds NL limbs of S slots; each limb is S/n independent
// polynomials of degree n (n a power of two, 2 <= n <= S). The twiddle
// table of limb l is tw[l] (bit-reversed powers for degree n).
//
// Embedded inputs m(X^2) of degree n: the first n/2 entries of the degree-n
// forward table are the degree-n/2 table for psi^2, and the evaluations of
// m(X^2) at bit-reversed slot p equal the degree-n/2 transform of m at slot
// p >> 1. Callers transform n/2 coefficients and read slot p >> 1.
#pragma once

#include "internal.h"
#include "ntt_index.h"

namespace via {

struct NttLimbs {
    const double *tw[2];
    FMod m[2];
};

__device__ __forceinline__ NttLimbs limbs_fwd(const RingDev &r) {
    NttLimbs t;
    t.tw[0] = r.t[0].fwd;
    t.tw[1] = r.t[1].fwd;
    t.m[0] = r.t[0].m;
    t.m[1] = r.t[1].m;
    return t;
}

__device__ __forceinline__ NttLimbs limbs_inv(const RingDev &r) {
    NttLimbs t;
    t.tw[0] = r.t[0].inv;
    t.tw[1] = r.t[1].inv;
    t.m[0] = r.t[0].m;
    t.m[1] = r.t[1].m;
    return t;
}

// Forward transform of every degree-n segment of the NL x S buffer.
template <int NL>
__device__ void ntt_fwd(double *s, int S, int n, int logn, const NttLimbs &T) {
    const int quarterS = S >> 2;
    int lm = 0;
    if (logn >= 2) {
        const int un = n >> 2;  // units per segment
        const int lun = logn - 2;
        for (; lm + 1 < logn; lm += 2) {
            for (int t = threadIdx.x; t < NL * quarterS; t += blockDim.x) {
                const int l = (NL == 1) ? 0 : (t >= quarterS ? 1 : 0);
                const int ua = t - l * quarterS;
                const int seg = ua >> lun;
                const int u = ua & (un - 1);
                const R4Idx r = r4_fwd(u, lm, logn);
                double *p = s + l * S + seg * n;
                const double *tw = l ? T.tw[1] : T.tw[0];
                const FMod md = l ? T.m[1] : T.m[0];
                const double w1 = tw[r.w1];
                const double a = p[r.a], b = p[r.b];
                const double t1 = mod_mul(p[r.c], w1, md), t2 = mod_mul(p[r.d], w1, md);
                const double a1 = mod_add(a, t1, md), c1 = mod_sub(a, t1, md);
                const double b1 = mod_add(b, t2, md), d1 = mod_sub(b, t2, md);
                const double t3 = mod_mul(b1, tw[r.w2], md), t4 = mod_mul(d1, tw[r.w3], md);
                p[r.a] = mod_add(a1, t3, md);
                p[r.b] = mod_sub(a1, t3, md);
                p[r.c] = mod_add(c1, t4, md);
                p[r.d] = mod_sub(c1, t4, md);
            }
            __syncthreads();
        }
    }
    if (lm < logn) {  // last radix-2 stage (h = 1): pairs (2i, 2i+1)
        const int m = 1 << lm;
        const int halfS = S >> 1;
        for (int t = threadIdx.x; t < NL * halfS; t += blockDim.x) {
            const int l = (NL == 1) ? 0 : (t >= halfS ? 1 : 0);
            const int u = t - l * halfS;
            const int seg = u >> (logn - 1);
            const int blk = u & ((n >> 1) - 1);
            double *p = s + l * S + seg * n + 2 * blk;
            const double *tw = l ? T.tw[1] : T.tw[0];
            const FMod md = l ? T.m[1] : T.m[0];
            const double x = p[0];
            const double y = mod_mul(p[1], tw[m + blk], md);
            p[0] = mod_add(x, y, md);
            p[1] = mod_sub(x, y, md);
        }
        __syncthreads();
    }
}

// Inverse transform without the N^-1 scaling (callers fold it into their store).
template <int NL>
__device__ void ntt_inv_noscale(double *s, int S, int n, int logn, const NttLimbs &T) {
    const int quarterS = S >> 2;
    int lh = 0;
    if (logn >= 2) {
        const int un = n >> 2;
        const int lun = logn - 2;
        for (; lh + 1 < logn; lh += 2) {
            for (int t = threadIdx.x; t < NL * quarterS; t += blockDim.x) {
                const int l = (NL == 1) ? 0 : (t >= quarterS ? 1 : 0);
                const int ua = t - l * quarterS;
                const int seg = ua >> lun;
                const int u = ua & (un - 1);
                const R4Idx r = r4_inv(u, lh, logn);
                double *p = s + l * S + seg * n;
                const double *tw = l ? T.tw[1] : T.tw[0];
                const FMod md = l ? T.m[1] : T.m[0];
                const double a = p[r.a], b = p[r.b], c = p[r.c], d = p[r.d];
                const double a1 = mod_add(a, b, md), b1 = mod_mul(mod_sub(a, b, md), tw[r.w1], md);
                const double c1 = mod_add(c, d, md), d1 = mod_mul(mod_sub(c, d, md), tw[r.w2], md);
                const double w3 = tw[r.w3];
                p[r.a] = mod_add(a1, c1, md);
                p[r.c] = mod_mul(mod_sub(a1, c1, md), w3, md);
                p[r.b] = mod_add(b1, d1, md);
                p[r.d] = mod_mul(mod_sub(b1, d1, md), w3, md);
            }
            __syncthreads();
        }
    }
    if (lh < logn) {  // last radix-2 stage: half n/2, one block per segment
        const int h = 1 << lh;
        const int halfS = S >> 1;
        for (int t = threadIdx.x; t < NL * halfS; t += blockDim.x) {
            const int l = (NL == 1) ? 0 : (t >= halfS ? 1 : 0);
            const int u = t - l * halfS;
            const int seg = u >> (logn - 1);
            const int j = u & (h - 1);
            double *p = s + l * S + seg * n + j;
            const double *tw = l ? T.tw[1] : T.tw[0];
            const FMod md = l ? T.m[1] : T.m[0];
            const double x = p[0], y = p[h];
            p[0] = mod_add(x, y, md);
            p[h] = mod_mul(mod_sub(x, y, md), tw[1], md);
        }
        __syncthreads();
    }
}

// Wrappers with the ring's own degree.
template <int NL>
__device__ __forceinline__ void lds_ntt_fwd(double *s, int S, const RingDev &r) {
    ntt_fwd<NL>(s, S, r.n, r.logn, limbs_fwd(r));
}

template <int NL>
__device__ __forceinline__ void lds_ntt_inv_noscale(double *s, int S, const RingDev &r) {
    ntt_inv_noscale<NL>(s, S, r.n, r.logn, limbs_inv(r));
}

}  // namespace via
