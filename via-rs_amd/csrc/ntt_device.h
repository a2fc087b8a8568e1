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
// (lds_words and lds_pad are defined in ntt_lazy.h.)

constexpr int kNttMaxR = 3;  // default stages per register pass (radix 8)

// Passes P, P+1, ... of a degree-2^LOGN transform over every segment of the
// NL x S buffer, with the degree and schedule fixed at compile time.
// With two limbs and a thread count that splits into wave-aligned halves
// (nth a multiple of 128), the first half of the threads takes the units of
// limb 0 and the second half those of limb 1: the limb, its twiddle table
// and its modulus are then the same across a wavefront and live in scalar
// registers instead of being selected per lane for every unit. Otherwise
// the units of both limbs are spread over all threads.
template <int NL, int LOGN, int MAXR, int P>
__device__ __forceinline__ void fwd_passes(double *s, int S, const NttLimbs &T, int tid, int nth) {
    if constexpr (P < sched_np(LOGN, MAXR)) {
        constexpr int R = sched_r(LOGN, MAXR, P), L0 = sched_l0(LOGN, MAXR, P), USEG = LOGN - R;
        const int upl = S >> R;  // units per limb
        if (NL == 2 && (nth & 127) == 0) {
            const int half = nth >> 1;
            const int l = __builtin_amdgcn_readfirstlane(tid >= half ? 1 : 0);
            const double *tw = l ? T.tw[1] : T.tw[0];
            const FMod m = l ? T.m[1] : T.m[0];
            for (int ua = tid - l * half; ua < upl; ua += half)
                ntt_fwd_unit_p<R, LOGN, L0>(s, l * S + ((ua >> USEG) << LOGN), ua & ((1 << USEG) - 1), tw, m);
        } else
        for (int t = tid; t < NL * upl; t += nth) {
            const int l = (NL == 1) ? 0 : (t >= upl ? 1 : 0);
            const int ua = t - l * upl;
            const int seg = ua >> USEG;
            const int u = ua & ((1 << USEG) - 1);
            ntt_fwd_unit_p<R, LOGN, L0>(s, l * S + (seg << LOGN), u, l ? T.tw[1] : T.tw[0], l ? T.m[1] : T.m[0]);
        }
        __syncthreads();
        fwd_passes<NL, LOGN, MAXR, P + 1>(s, S, T, tid, nth);
    }
}

template <int NL, int LOGN, int MAXR, int P>
__device__ __forceinline__ void inv_passes(double *s, int S, const NttLimbs &T, int tid, int nth) {
    if constexpr (P < sched_np(LOGN, MAXR)) {
        constexpr int R = sched_r(LOGN, MAXR, P), LH0 = sched_l0(LOGN, MAXR, P), USEG = LOGN - R;
        const int upl = S >> R;
        if (NL == 2 && (nth & 127) == 0) {
            const int half = nth >> 1;
            const int l = __builtin_amdgcn_readfirstlane(tid >= half ? 1 : 0);
            const double *tw = l ? T.tw[1] : T.tw[0];
            const FMod m = l ? T.m[1] : T.m[0];
            for (int ua = tid - l * half; ua < upl; ua += half)
                ntt_inv_unit_p<R, LOGN, LH0>(s, l * S + ((ua >> USEG) << LOGN), ua & ((1 << USEG) - 1), tw, m);
        } else
        for (int t = tid; t < NL * upl; t += nth) {
            const int l = (NL == 1) ? 0 : (t >= upl ? 1 : 0);
            const int ua = t - l * upl;
            const int seg = ua >> USEG;
            const int u = ua & ((1 << USEG) - 1);
            ntt_inv_unit_p<R, LOGN, LH0>(s, l * S + (seg << LOGN), u, l ? T.tw[1] : T.tw[0], l ? T.m[1] : T.m[0]);
        }
        __syncthreads();
        inv_passes<NL, LOGN, MAXR, P + 1>(s, S, T, tid, nth);
    }
}

// Forward transform of every degree-n segment (n = 2^logn <= 2048) of the
// NL x S buffer (padded layout, lds_words(NL * S) words; S a multiple of n)
// by threads tid = 0 .. nth-1 (a subset of the workgroup may transform its
// own buffer; every thread of the workgroup must reach the pass barriers).
// Inputs are exact integers in (-2q, 2q); outputs are lazy, in (-24q, 24q)
// (ntt_lazy.h): reduce before storing canonical residues.
template <int NL, int MAXR = kNttMaxR>
__device__ void ntt_fwd(double *s, int S, int n, int logn, const NttLimbs &T, int tid, int nth) {
    (void)n;
    switch (logn) {
        case 1: fwd_passes<NL, 1, MAXR, 0>(s, S, T, tid, nth); break;
        case 2: fwd_passes<NL, 2, MAXR, 0>(s, S, T, tid, nth); break;
        case 3: fwd_passes<NL, 3, MAXR, 0>(s, S, T, tid, nth); break;
        case 4: fwd_passes<NL, 4, MAXR, 0>(s, S, T, tid, nth); break;
        case 5: fwd_passes<NL, 5, MAXR, 0>(s, S, T, tid, nth); break;
        case 6: fwd_passes<NL, 6, MAXR, 0>(s, S, T, tid, nth); break;
        case 7: fwd_passes<NL, 7, MAXR, 0>(s, S, T, tid, nth); break;
        case 8: fwd_passes<NL, 8, MAXR, 0>(s, S, T, tid, nth); break;
        case 9: fwd_passes<NL, 9, MAXR, 0>(s, S, T, tid, nth); break;
        case 10: fwd_passes<NL, 10, MAXR, 0>(s, S, T, tid, nth); break;
        case 11: fwd_passes<NL, 11, MAXR, 0>(s, S, T, tid, nth); break;
        default: break;  // degree 1: the identity
    }
}

// Inverse transform without the N^-1 scaling (callers fold it into their
// store with mod_mul, which also makes the lazy outputs canonical), by
// threads tid = 0 .. nth-1 as for ntt_fwd. Inputs are exact integers in
// (-4q, 4q).
template <int NL, int MAXR = kNttMaxR>
__device__ void ntt_inv_noscale(double *s, int S, int n, int logn, const NttLimbs &T, int tid, int nth) {
    (void)n;
    switch (logn) {
        case 1: inv_passes<NL, 1, MAXR, 0>(s, S, T, tid, nth); break;
        case 2: inv_passes<NL, 2, MAXR, 0>(s, S, T, tid, nth); break;
        case 3: inv_passes<NL, 3, MAXR, 0>(s, S, T, tid, nth); break;
        case 4: inv_passes<NL, 4, MAXR, 0>(s, S, T, tid, nth); break;
        case 5: inv_passes<NL, 5, MAXR, 0>(s, S, T, tid, nth); break;
        case 6: inv_passes<NL, 6, MAXR, 0>(s, S, T, tid, nth); break;
        case 7: inv_passes<NL, 7, MAXR, 0>(s, S, T, tid, nth); break;
        case 8: inv_passes<NL, 8, MAXR, 0>(s, S, T, tid, nth); break;
        case 9: inv_passes<NL, 9, MAXR, 0>(s, S, T, tid, nth); break;
        case 10: inv_passes<NL, 10, MAXR, 0>(s, S, T, tid, nth); break;
        case 11: inv_passes<NL, 11, MAXR, 0>(s, S, T, tid, nth); break;
        default: break;
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
