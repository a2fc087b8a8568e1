// ntt_lazy.h — radix-2^R passes of the negacyclic NTT over lazy residues.
//
// The transform is the reference's (via-rs ring/ntt.rs:305-399): forward
// Cooley-Tukey decimation in time from natural order to bit-reversed order
// with twiddle psi^br(m+b) for block b of stage m, inverse Gentleman-Sande
// decimation in frequency with the inverse table. A pass applies R
// consecutive radix-2 stages to 2^R elements held in registers, so one LDS
// round trip (and one barrier) covers R stages. Every butterfly and twiddle is
// the reference's, so after canonicalisation the result is bit-identical.
//
// Residues are exact integers in fp64 lanes, kept lazily (not reduced to
// [0, q) between stages). With q < 2^38:
//   lazy_mul(a, w)  = a*w - t*q, t = floor(a*w/q) or one off, in [-q, 2q)
//                     for |a| < 2^45 and 0 <= w < q (the fp64 quotient
//                     estimate is within 2^-7 of a*w/q below 2^45 * 2^38 / q);
//   lazy_reduce(v)  = v - t*q, same estimate, in [-q, 2q) for |v| < 2^50.
// Forward: a butterfly (x + y*w, x - y*w) grows the bound by 2q, so inputs in
// (-2q, 2q) leave n = 2^11 stages below 24q (< 2^43): no reduction at all.
// Inverse: (x + y, (x - y)*w) doubles the sum path; each pass reduces the
// outputs whose bound exceeds 4q, so pass inputs stay below 4q and values
// inside a pass below 64q. Outputs are lazy; mod_reduce_small or mod_mul make
// them canonical (both accept |v| < 2^45).
#pragma once

#include "modarith.h"

#if defined(__HIPCC__)
#define NTT_HD __host__ __device__ __forceinline__
#else
#define NTT_HD static inline
#endif

VIA_HD double lazy_reduce(double v, const FMod &m) {
    const double t = floor(v * m.qinv);
    return fma(-t, m.q, v);
}

// Forward pass from stage L0 over R stages for unit u of a degree-2^logn
// segment p (units per segment: 2^(logn-R)). The unit's elements are
// base + t*stride, t < 2^R; at sub-stage s the pairs are (t, t + 2^(R-1-s))
// and the block of stage L0+s is (SB << s) + (t >> (R-s)).
template <int R, typename Mem>
NTT_HD void ntt_fwd_unit(Mem p, int logn, int L0, int u, const double *tw, const FMod &m) {
    constexpr int E = 1 << R;
    const int jbits = logn - L0 - R;
    const int SB = u >> jbits;
    const int j = u & ((1 << jbits) - 1);
    const int base = (SB << (logn - L0)) + j;
    const int stride = 1 << jbits;
    double x[E];
#pragma unroll
    for (int t = 0; t < E; t++) x[t] = p.ld(base + t * stride);
#pragma unroll
    for (int s = 0; s < R; s++) {
        const int half = E >> (s + 1);
        const int wbase = (1 << (L0 + s)) + (SB << s);
#pragma unroll
        for (int t = 0; t < E; t++) {
            if (t & half) continue;
            const double w = tw[wbase + (t >> (R - s))];
            const double y = mod_mul_lazy(x[t + half], w, m);
            const double a = x[t];
            x[t] = a + y;
            x[t + half] = a - y;
        }
    }
#pragma unroll
    for (int t = 0; t < E; t++) p.st(base + t * stride, x[t]);
}

// Inverse pass over stages lh0 .. lh0+R-1 (half 2^lh at stage lh) for unit u:
// elements base + t*2^lh0 with base = (SB << (lh0+R)) + j; at sub-stage s the
// pairs are (t, t + 2^s), block (SB << (R-1-s)) + (t >> (s+1)) of the stage's
// n >> (lh+1) blocks. Outputs t < max(1, 2^R/4) are reduced (their bound
// would exceed 4q).
template <int R, typename Mem>
NTT_HD void ntt_inv_unit(Mem p, int logn, int lh0, int u, const double *tw, const FMod &m) {
    constexpr int E = 1 << R;
    constexpr int NRED = E >= 4 ? E / 4 : 1;
    const int SB = u >> lh0;
    const int j = u & ((1 << lh0) - 1);
    const int base = (SB << (lh0 + R)) + j;
    const int stride = 1 << lh0;
    double x[E];
#pragma unroll
    for (int t = 0; t < E; t++) x[t] = p.ld(base + t * stride);
#pragma unroll
    for (int s = 0; s < R; s++) {
        const int half = 1 << s;
        const int wbase = (1 << (logn - lh0 - s - 1)) + (SB << (R - 1 - s));
#pragma unroll
        for (int t = 0; t < E; t++) {
            if (t & half) continue;
            const double w = tw[wbase + (t >> (s + 1))];
            const double a = x[t], b = x[t + half];
            x[t] = a + b;
            x[t + half] = mod_mul_lazy(a - b, w, m);
        }
    }
#pragma unroll
    for (int t = 0; t < NRED; t++) x[t] = lazy_reduce(x[t], m);
#pragma unroll
    for (int t = 0; t < E; t++) p.st(base + t * stride, x[t]);
}

// ---- fixed-shape units over the padded LDS layout -------------------------------
//
// Transform buffers in LDS carry one pad word per 16 (logical word i at
// lds_pad(i)). For a unit of 2^R elements at logical base + t*stride, with
// the segment start a multiple of the degree, the elements never straddle a
// pad differently from the unit's first one, so element t sits at
// lds_pad(base) + t*stride + ((t*stride) >> 4): with the degree and the pass
// known at compile time these are immediate offsets.

NTT_HD constexpr int lds_words(int w) { return w + w / 16; }
NTT_HD int lds_pad(int i) { return i + (i >> 4); }

template <int STRIDE>
NTT_HD constexpr int unit_off(int t) {
    return t * STRIDE + ((t * STRIDE) >> 4);
}

// Forward pass from stage L0 over R stages of a degree-2^LOGN segment whose
// logical start is seg_base (a multiple of 2^LOGN), unit u: the same
// butterflies and twiddles as ntt_fwd_unit.
template <int R, int LOGN, int L0>
NTT_HD void ntt_fwd_unit_p(double *s, int seg_base, int u, const double *tw, const FMod &m) {
    constexpr int E = 1 << R;
    constexpr int JB = LOGN - L0 - R;
    constexpr int STRIDE = 1 << JB;
    const int SB = u >> JB;
    const int j = u & (STRIDE - 1);
    double *p = s + lds_pad(seg_base + (SB << (LOGN - L0)) + j);
    double x[E];
#pragma unroll
    for (int t = 0; t < E; t++) x[t] = p[unit_off<STRIDE>(t)];
#pragma unroll
    for (int st = 0; st < R; st++) {
        const int half = E >> (st + 1);
        const double *w = tw + (1 << (L0 + st)) + (SB << st);
#pragma unroll
        for (int t = 0; t < E; t++) {
            if (t & half) continue;
            const double y = mod_mul_lazy(x[t + half], w[t >> (R - st)], m);
            const double a = x[t];
            x[t] = a + y;
            x[t + half] = a - y;
        }
    }
#pragma unroll
    for (int t = 0; t < E; t++) p[unit_off<STRIDE>(t)] = x[t];
}

// Inverse pass over stages LH0 .. LH0+R-1, unit u: the same butterflies,
// twiddles and reductions as ntt_inv_unit.
template <int R, int LOGN, int LH0>
NTT_HD void ntt_inv_unit_p(double *s, int seg_base, int u, const double *tw, const FMod &m) {
    constexpr int E = 1 << R;
    constexpr int NRED = E >= 4 ? E / 4 : 1;
    constexpr int STRIDE = 1 << LH0;
    const int SB = u >> LH0;
    const int j = u & (STRIDE - 1);
    double *p = s + lds_pad(seg_base + (SB << (LH0 + R)) + j);
    double x[E];
#pragma unroll
    for (int t = 0; t < E; t++) x[t] = p[unit_off<STRIDE>(t)];
#pragma unroll
    for (int st = 0; st < R; st++) {
        const int half = 1 << st;
        const double *w = tw + (1 << (LOGN - LH0 - st - 1)) + (SB << (R - 1 - st));
#pragma unroll
        for (int t = 0; t < E; t++) {
            if (t & half) continue;
            const double a = x[t], b = x[t + half];
            x[t] = a + b;
            x[t + half] = mod_mul_lazy(a - b, w[t >> (st + 1)], m);
        }
    }
#pragma unroll
    for (int t = 0; t < NRED; t++) x[t] = lazy_reduce(x[t], m);
#pragma unroll
    for (int t = 0; t < E; t++) p[unit_off<STRIDE>(t)] = x[t];
}

// Compile-time form of ntt_schedule: pass i of a degree-2^logn transform in
// passes of at most maxr stages has sched_r stages starting at sched_l0.
NTT_HD constexpr int sched_np(int logn, int maxr) { return logn <= 0 ? 0 : (logn + maxr - 1) / maxr; }
NTT_HD constexpr int sched_r(int logn, int maxr, int i) {
    return logn / sched_np(logn, maxr) + (i < logn - (logn / sched_np(logn, maxr)) * sched_np(logn, maxr) ? 1 : 0);
}
NTT_HD constexpr int sched_l0(int logn, int maxr, int i) { return i <= 0 ? 0 : sched_l0(logn, maxr, i - 1) + sched_r(logn, maxr, i - 1); }

// Pass schedule: ceil(logn / maxr) passes with stage counts as even as
// possible, larger first (e.g. 11 -> 4,4,3; 10 -> 4,3,3). Returns the count.
NTT_HD int ntt_schedule(int logn, int maxr, int *rs) {
    if (logn <= 0) return 0;
    const int np = (logn + maxr - 1) / maxr;
    const int lo = logn / np, extra = logn - lo * np;
    for (int i = 0; i < np; i++) rs[i] = lo + (i < extra ? 1 : 0);
    return np;
}
