// ntt.hip — batched negacyclic NTT, u64 <-> f64 conversion and the batched
// negacyclic product (RingPolyEval of via-rs ring/ntt_abstraction.rs:61-78).
#include "ntt_device.h"

namespace via {

// One workgroup transforms kSlots/n polys of one limb (or one poly when
// n == kSlots). The polys of a workgroup share a limb (checked by the host).
// Radix-16 passes: these kernels hold little else in registers.
template <bool INV>
__global__ __launch_bounds__(kThreads) void k_ntt_batch(RingDev r, const double *src, double *data, size_t npolys,
                                                        size_t limb_group) {
    __shared__ double s[lds_words(kSlots)];
    const int n = r.n;
    const int P = kSlots / n;
    const size_t p0 = (size_t)blockIdx.x * P;
    const int cnt = (int)((npolys - p0) < (size_t)P ? (npolys - p0) : (size_t)P);
    const int l = (int)((p0 / limb_group) % (size_t)r.nl);
    RingDev r1 = r;
    if (l == 1) r1.t[0] = r.t[1];
    double *g = data + p0 * n;
    const double *gs = src + p0 * n;
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

void launch_ntt(const RingDev &r, double *data, size_t npolys, size_t limb_group, bool inverse, hipStream_t st) {
    if (!npolys) return;
    const size_t P = kSlots / r.n;
    const size_t blocks = (npolys + P - 1) / P;
    if (inverse)
        hipLaunchKernelGGL(k_ntt_batch<true>, dim3(blocks), dim3(kThreads), 0, st, r, data, data, npolys, limb_group);
    else
        hipLaunchKernelGGL(k_ntt_batch<false>, dim3(blocks), dim3(kThreads), 0, st, r, data, data, npolys, limb_group);
}

void launch_ntt_to(const RingDev &r, const double *src, double *dst, size_t npolys, size_t limb_group,
                   hipStream_t st) {
    if (!npolys) return;
    const size_t P = kSlots / r.n;
    const size_t blocks = (npolys + P - 1) / P;
    hipLaunchKernelGGL(k_ntt_batch<false>, dim3(blocks), dim3(kThreads), 0, st, r, src, dst, npolys, limb_group);
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
                                                      size_t limb_group) {
    __shared__ double s[lds_words(2 * kSlots)];
    const int n = r.n;
    const size_t p = blockIdx.x;
    const int l = (int)((p / limb_group) % (size_t)r.nl);
    RingDev r2 = r;
    r2.t[0] = l ? r.t[1] : r.t[0];
    r2.t[1] = r2.t[0];
    const uint64_t *ga = a + p * n, *gb = b + p * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        s[lds_pad(i)] = (double)ga[i];
        s[lds_pad(n + i)] = (double)gb[i];
    }
    __syncthreads();
    lds_ntt_fwd<2, 4>(s, n, r2);
    const FMod md = r2.t[0].m;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        s[lds_pad(i)] = mod_mul(mod_reduce_small(s[lds_pad(i)], md), s[lds_pad(n + i)], md);
    __syncthreads();
    lds_ntt_inv_noscale<1, 4>(s, n, r2);
    const double ni = r2.t[0].n_inv;
    uint64_t *gc = c + p * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) gc[i] = (uint64_t)mod_mul(s[lds_pad(i)], ni, md);
}

void launch_polymul_u64(const RingDev &r, const uint64_t *a, const uint64_t *b, uint64_t *c, size_t count,
                        size_t limb_group, hipStream_t st) {
    if (count) hipLaunchKernelGGL(k_polymul, dim3(count), dim3(kThreads), 0, st, r, a, b, c, limb_group);
}

}  // namespace via
