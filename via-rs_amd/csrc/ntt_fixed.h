// ntt_fixed.h — register-level pieces of fixed-shape transforms: a pass
// whose unit (2^R elements) a thread holds in registers, with compile-time
// geometry, used by kernels that keep the first or last pass of a transform
// out of LDS (the cascade MAC, the fixed-shape gates).
//
// The butterflies and twiddles are those of ntt_lazy.h (the reference's DIT
// forward transform, ring/ntt.rs:305-343, and DIF inverse, ring/ntt.rs:356-399);
// only the grouping of exact integer operations differs in direct_pass.
#pragma once

#include "ntt_device.h"

namespace via {

// Passes of a 2^LOGNT-point transform whose last pass has RL stages
// (outputs in registers): none for LOGNT = 0, one for LOGNT <= RL, else
// passes of at most three stages ending with the pass of RL, the stages
// before it split as evenly as possible, larger first.
__host__ __device__ constexpr int cm_np(int logn, int rl) {
    return logn == 0 ? 0 : logn <= rl ? 1 : 1 + (logn - rl + 2) / 3;
}
__host__ __device__ constexpr int cm_r(int logn, int rl, int p) {
    return logn <= rl ? logn
           : p == cm_np(logn, rl) - 1
               ? rl
               : (logn - rl) / (cm_np(logn, rl) - 1) + (p < (logn - rl) % (cm_np(logn, rl) - 1) ? 1 : 0);
}
__host__ __device__ constexpr int cm_l0(int logn, int rl, int p) {
    return p == 0 ? 0 : cm_l0(logn, rl, p - 1) + cm_r(logn, rl, p - 1);
}

// Direct form of stages 0..R-1 of block 0 on x[0 .. 2^R) (element t at
// offset t * stride of a unit that starts its segment's block 0), with the
// ring's direct constants c (direct_consts on the host). Exact: every term is
// a constant below q times an input of magnitude at most B/2.
template <int R>
__device__ __forceinline__ void direct_pass(double *x, const double *c) {
    if constexpr (R == 1) {
        const double y = c[0] * x[1];
        const double a = x[0];
        x[0] = a + y;
        x[1] = a - y;
    } else if constexpr (R == 2) {
        const double p = fma(c[0], x[2], x[0]), m = fma(-c[0], x[2], x[0]);
        const double t1 = fma(c[2], x[3], c[1] * x[1]);
        const double t2 = fma(-c[4], x[3], c[3] * x[1]);
        x[0] = p + t1;
        x[1] = p - t1;
        x[2] = m + t2;
        x[3] = m - t2;
    } else if constexpr (R == 3) {
        const double p = fma(c[0], x[4], x[0]), m = fma(-c[0], x[4], x[0]);
        const double t1 = fma(c[2], x[6], c[1] * x[2]);
        const double t2 = fma(-c[4], x[6], c[3] * x[2]);
        double b[4] = {p + t1, p - t1, m + t2, m - t2};
        double o[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double *d = c + 5 + 4 * k;
            o[k] = fma(d[3], x[7], fma(d[2], x[3], fma(d[1], x[5], d[0] * x[1])));
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            x[2 * k] = b[k] + o[k];
            x[2 * k + 1] = b[k] - o[k];
        }
    }
}

// R lazy stages of a unit held in registers: stage s of the unit's block SB
// takes twiddle w[(1 << s) - 1 + (t >> (R - s))] (the 2^R - 1 twiddles the
// unit needs, prefetched by load_unit_tw); the butterflies of ntt_fwd_unit_p.
template <int R>
__device__ __forceinline__ void unit_stages(double *x, const double *w, const FMod &m) {
    constexpr int E = 1 << R;
#pragma unroll
    for (int st = 0; st < R; st++) {
        const int half = E >> (st + 1);
#pragma unroll
        for (int t = 0; t < E; t++) {
            if (t & half) continue;
            const double y = mod_mul_lazy(x[t + half], w[(1 << st) - 1 + (t >> (R - st))], m);
            const double a = x[t];
            x[t] = a + y;
            x[t + half] = a - y;
        }
    }
}

// The 2^R - 1 twiddles of block SB over stages L0 .. L0+R-1.
template <int R, int L0>
__device__ __forceinline__ void load_unit_tw(double *w, const double *tw, int SB) {
#pragma unroll
    for (int st = 0; st < R; st++)
#pragma unroll
        for (int e = 0; e < (1 << st); e++) w[(1 << st) - 1 + e] = tw[(1 << (L0 + st)) + (SB << st) + e];
}

// Unit geometry of pass (L0, R) of a 2^LOGNT-point transform over the 1024
// inputs of a limb: unit u's elements are base + t * 2^JB, its block SB.
template <int LOGNT, int L0, int R>
struct UnitGeo {
    static constexpr int JB = LOGNT - L0 - R;
    static constexpr int USEG = LOGNT - R;  // log2 units per segment
    __device__ static __forceinline__ void at(int u, int &base, int &SB) {
        const int seg = u >> USEG;
        const int uu = u & ((1 << USEG) - 1);
        SB = uu >> JB;
        base = (seg << LOGNT) + (SB << (LOGNT - L0)) + (uu & ((1 << JB) - 1));
    }
};

// The inverse (DIF) counterparts: unit u of pass (LH0, R) over a degree-2^LOGN
// segment holds elements base + t * 2^LH0, base = (SB << (LH0 + R)) + j.
template <int LOGN, int LH0, int R>
struct InvGeo {
    static constexpr int USEG = LOGN - R;  // log2 units per segment
    __device__ static __forceinline__ void at(int u, int &base, int &SB) {
        const int seg = u >> USEG;
        const int uu = u & ((1 << USEG) - 1);
        SB = uu >> LH0;
        base = (seg << LOGN) + (SB << (LH0 + R)) + (uu & ((1 << LH0) - 1));
    }
};

// The 2^R - 1 inverse twiddles of block SB over stages LH0 .. LH0+R-1 of a
// degree-2^LOGN transform: stage st's 2^(R-1-st) at w[2^R - 2^(R-st) + e].
template <int LOGN, int LH0, int R>
__device__ __forceinline__ void load_inv_unit_tw(double *w, const double *tw, int SB) {
#pragma unroll
    for (int st = 0; st < R; st++)
#pragma unroll
        for (int e = 0; e < (1 << (R - 1 - st)); e++)
            w[(1 << R) - (1 << (R - st)) + e] = tw[(1 << (LOGN - LH0 - st - 1)) + (SB << (R - 1 - st)) + e];
}

// R inverse stages of a unit in registers (ntt_inv_unit_p's butterflies),
// then the reduction of the outputs whose bound would exceed 4q.
template <int R>
__device__ __forceinline__ void inv_unit_stages(double *x, const double *w, const FMod &m) {
    constexpr int E = 1 << R;
    constexpr int NRED = E >= 4 ? E / 4 : 1;
#pragma unroll
    for (int st = 0; st < R; st++) {
        const int half = 1 << st;
#pragma unroll
        for (int t = 0; t < E; t++) {
            if (t & half) continue;
            const double a = x[t], b = x[t + half];
            x[t] = a + b;
            x[t + half] = mod_mul_lazy(a - b, w[(1 << R) - (1 << (R - st)) + (t >> (st + 1))], m);
        }
    }
#pragma unroll
    for (int t = 0; t < NRED; t++) x[t] = lazy_reduce(x[t], m);
}

}  // namespace via
