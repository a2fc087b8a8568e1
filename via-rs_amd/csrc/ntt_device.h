// ntt_device.h — LDS-resident negacyclic NTT for workgroup-cooperative kernels.
//
// Forward: iterative Cooley-Tukey decimation in time, natural order in and
// bit-reversed order out, twiddle of block b in stage m = psi^br(m+b)
// (via-rs ring/ntt.rs:305-343). Inverse: decimation in frequency from
// bit-reversed order, then the N^-1 scaling (ring/ntt.rs:356-399). Both are
// exact, so the result is bit-identical to the reference; the stages run in
// the reference order, up to four at a time in registers (radix-16 passes
// over lazy residues, ntt_lazy.h) with one barrier per pass.
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
#include "ntt_lazy.h"

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

// LDS layout of every transform buffer: one pad word after every 16, so
// that the strided element groups of the radix-8 passes (strides of 4 and
// 32 words, 16 lanes per LDS cycle) fall into distinct banks. Logical word i
// of a buffer lives at lds_pad(i); a buffer of w logical words takes
// lds_words(w).
__host__ __device__ constexpr int lds_words(int w) { return w + w / 16; }
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }

// LDS view for the pass kernels of ntt_lazy.h: a segment starting at logical
// word `base` of a padded buffer.
struct LdsMem {
    double *s;
    int base;
    __device__ __forceinline__ double ld(int i) const { return s[lds_pad(base + i)]; }
    __device__ __forceinline__ void st(int i, double v) const { s[lds_pad(base + i)] = v; }
};

constexpr int kNttMaxR = 3;  // default stages per register pass (radix 8)

// Forward transform of every degree-n segment of the NL x S buffer (padded
// layout, lds_words(NL * S) words) by threads
// tid = 0 .. nth-1 (a subset of the workgroup may transform its own buffer;
// every thread of the workgroup must reach the pass barriers). Inputs
// are exact integers in (-2q, 2q); outputs are lazy, in (-24q, 24q) for
// n <= 2048 (ntt_lazy.h): reduce before storing canonical residues.
template <int NL, int MAXR = kNttMaxR>
__device__ void ntt_fwd(double *s, int S, int n, int logn, const NttLimbs &T, int tid, int nth) {
    int rs[12];
    const int np = ntt_schedule(logn, MAXR, rs);
    int L0 = 0;
    for (int pi = 0; pi < np; pi++) {
        const int R = rs[pi];
        const int upl = S >> R;  // units per limb
        const int useg = logn - R;
        for (int t = tid; t < NL * upl; t += nth) {
            const int l = (NL == 1) ? 0 : (t >= upl ? 1 : 0);
            const int ua = t - l * upl;
            const int seg = ua >> useg;
            const int u = ua & ((1 << useg) - 1);
            const LdsMem p{s, l * S + seg * n};
            const double *tw = l ? T.tw[1] : T.tw[0];
            const FMod md = l ? T.m[1] : T.m[0];
            if (MAXR >= 4 && R == 4)
                ntt_fwd_unit<(MAXR >= 4 ? 4 : 1)>(p, logn, L0, u, tw, md);
            else if (MAXR >= 3 && R == 3)
                ntt_fwd_unit<(MAXR >= 3 ? 3 : 1)>(p, logn, L0, u, tw, md);
            else if (R == 2)
                ntt_fwd_unit<2>(p, logn, L0, u, tw, md);
            else
                ntt_fwd_unit<1>(p, logn, L0, u, tw, md);
        }
        __syncthreads();
        L0 += R;
    }
}

// Inverse transform without the N^-1 scaling (callers fold it into their
// store with mod_mul, which also makes the lazy outputs canonical), by
// threads tid = 0 .. nth-1 as for ntt_fwd. Inputs are exact integers in
// (-4q, 4q).
template <int NL, int MAXR = kNttMaxR>
__device__ void ntt_inv_noscale(double *s, int S, int n, int logn, const NttLimbs &T, int tid, int nth) {
    int rs[12];
    const int np = ntt_schedule(logn, MAXR, rs);
    int lh0 = 0;
    for (int pi = 0; pi < np; pi++) {
        const int R = rs[pi];
        const int upl = S >> R;
        const int useg = logn - R;
        for (int t = tid; t < NL * upl; t += nth) {
            const int l = (NL == 1) ? 0 : (t >= upl ? 1 : 0);
            const int ua = t - l * upl;
            const int seg = ua >> useg;
            const int u = ua & ((1 << useg) - 1);
            const LdsMem p{s, l * S + seg * n};
            const double *tw = l ? T.tw[1] : T.tw[0];
            const FMod md = l ? T.m[1] : T.m[0];
            if (MAXR >= 4 && R == 4)
                ntt_inv_unit<(MAXR >= 4 ? 4 : 1)>(p, logn, lh0, u, tw, md);
            else if (MAXR >= 3 && R == 3)
                ntt_inv_unit<(MAXR >= 3 ? 3 : 1)>(p, logn, lh0, u, tw, md);
            else if (R == 2)
                ntt_inv_unit<2>(p, logn, lh0, u, tw, md);
            else
                ntt_inv_unit<1>(p, logn, lh0, u, tw, md);
        }
        __syncthreads();
        lh0 += R;
    }
}

// Every thread of the workgroup takes part.
template <int NL, int MAXR = kNttMaxR>
__device__ __forceinline__ void ntt_fwd(double *s, int S, int n, int logn, const NttLimbs &T) {
    ntt_fwd<NL, MAXR>(s, S, n, logn, T, threadIdx.x, blockDim.x);
}

template <int NL, int MAXR = kNttMaxR>
__device__ __forceinline__ void ntt_inv_noscale(double *s, int S, int n, int logn, const NttLimbs &T) {
    ntt_inv_noscale<NL, MAXR>(s, S, n, logn, T, threadIdx.x, blockDim.x);
}

// Wrappers with the ring's own degree.
template <int NL, int MAXR = kNttMaxR>
__device__ __forceinline__ void lds_ntt_fwd(double *s, int S, const RingDev &r) {
    ntt_fwd<NL, MAXR>(s, S, r.n, r.logn, limbs_fwd(r));
}

template <int NL, int MAXR = kNttMaxR>
__device__ __forceinline__ void lds_ntt_inv_noscale(double *s, int S, const RingDev &r) {
    ntt_inv_noscale<NL, MAXR>(s, S, r.n, r.logn, limbs_inv(r));
}

}  // namespace via
