// ntt_device.h — LDS-resident negacyclic NTT for workgroup-cooperative kernels.
//
// Forward: iterative Cooley-Tukey decimation in time, natural order in and
// bit-reversed order out, twiddle of block b in stage m = psi^br(m+b)
// (via-rs ring/ntt.rs:305-343). Inverse: decimation in frequency from
// bit-reversed order, then the N^-1 scaling (ring/ntt.rs:356-399). Both are
// exact, so any schedule producing these values is bit-identical to the
// reference; this one keeps the reference's butterfly order per stage.
//
// The LDS buffer holds NL limbs of S slots; each limb is S/n independent
// polynomials of degree n (n a power of two, 2 <= n <= S).
#pragma once

#include "internal.h"

namespace via {

__device__ __forceinline__ const double *tw_fwd(const RingDev &r, int l) { return l ? r.t[1].fwd : r.t[0].fwd; }
__device__ __forceinline__ const double *tw_inv(const RingDev &r, int l) { return l ? r.t[1].inv : r.t[0].inv; }
__device__ __forceinline__ FMod fm(const RingDev &r, int l) { return l ? r.t[1].m : r.t[0].m; }
__device__ __forceinline__ double ninv(const RingDev &r, int l) { return l ? r.t[1].n_inv : r.t[0].n_inv; }

template <int NL>
__device__ void lds_ntt_fwd(double *s, int S, const RingDev &r) {
    const int logn = r.logn, n = r.n;
    const int halfS = S >> 1;
    const int nh_mask = (n >> 1) - 1;
    for (int lm = 0; lm < logn; lm++) {
        const int lh = logn - lm - 1;
        const int half = 1 << lh;
        const int m = 1 << lm;
        for (int t = threadIdx.x; t < NL * halfS; t += blockDim.x) {
            const int l = (NL == 1) ? 0 : (t >= halfS ? 1 : 0);
            const int u = t - l * halfS;
            const int p = u >> (logn - 1);
            const int v = u & nh_mask;
            const int blk = v >> lh;
            const int j = v & (half - 1);
            const int base = l * S + p * n + (blk << (lh + 1)) + j;
            const FMod md = fm(r, l);
            const double w = tw_fwd(r, l)[m + blk];
            const double x = s[base];
            const double y = mod_mul(s[base + half], w, md);
            s[base] = mod_add(x, y, md);
            s[base + half] = mod_sub(x, y, md);
        }
        __syncthreads();
    }
}

// Inverse without the final N^-1 scaling (callers fold it into their store).
template <int NL>
__device__ void lds_ntt_inv_noscale(double *s, int S, const RingDev &r) {
    const int logn = r.logn, n = r.n;
    const int halfS = S >> 1;
    const int nh_mask = (n >> 1) - 1;
    for (int lh = 0; lh < logn; lh++) {
        const int half = 1 << lh;
        const int m = n >> (lh + 1);
        for (int t = threadIdx.x; t < NL * halfS; t += blockDim.x) {
            const int l = (NL == 1) ? 0 : (t >= halfS ? 1 : 0);
            const int u = t - l * halfS;
            const int p = u >> (logn - 1);
            const int v = u & nh_mask;
            const int blk = v >> lh;
            const int j = v & (half - 1);
            const int base = l * S + p * n + (blk << (lh + 1)) + j;
            const FMod md = fm(r, l);
            const double w = tw_inv(r, l)[m + blk];
            const double x = s[base];
            const double y = s[base + half];
            s[base] = mod_add(x, y, md);
            s[base + half] = mod_mul(mod_sub(x, y, md), w, md);
        }
        __syncthreads();
    }
}

}  // namespace via
