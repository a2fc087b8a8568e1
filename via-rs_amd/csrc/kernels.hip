// kernels.hip — the VIA-C server answer path on gfx950.
//
// Every kernel computes exactly the integers the reference computes; the
// comments name the via-rs function each one restates.
#include <atomic>

#include "digits.h"
#include "ntt_device.h"

namespace via {

// ---- LWE -> RLWE cascade (conv.rs:62-95, cascade.rs:388-405) -----------------------

// Gadget digits of `items` coefficient vectors of n coefficients: item i,
// limb l, coefficient c at src[i*item_stride + l*limb_stride + c], digits to
// int16 planes dig[i][k][n]. One thread per coefficient over the whole grid
// (the wide gadget scale dominates). Used for the LWE masks (the rank-n
// degree-1 input of the first conv step); the digits of every cascade
// output come from k_cascade_combine, which computes the same values with
// the same scale in the workgroup that produces the coefficient.
template <int NL>
__global__ __launch_bounds__(256) void k_poly_digits(RingDev r1, Gadget g, DigitPlan dp, const double *src,
                                                     size_t item_stride, size_t limb_stride, int n, int16_t *dig) {
    const int item = blockIdx.y;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double *x = src + (size_t)item * item_stride;
    const double a0 = x[c], a1 = NL == 2 ? x[limb_stride + c] : 0.0;
    write_digits<NL>(a0, a1, r1, g, dp, dig + (size_t)item * g.L * n + c, (size_t)n);
}

static std::atomic<int> g_exact_only{0};
bool exact_paths_only() { return g_exact_only.load(std::memory_order_relaxed) != 0; }
void set_exact_paths_only(bool on) { g_exact_only.store(on ? 1 : 0, std::memory_order_relaxed); }

bool scale_w_make(const RingDev &r, const Gadget &g, const DigitPlan &dp, ScaleW *sw) {
    if (exact_paths_only()) return false;
    const via_u128 q1 = r.nl == 2 ? (via_u128)r.rns.q1 : (via_u128)r.q_single;
    const via_u128 Q = r.nl == 2 ? r.rns.big_q : q1;
    via_u128 bhi = 1;  // B^(L-h) bounds the high half
    for (int k = dp.h; k < g.L; k++) {
        bhi *= g.base;
        if (bhi >= ((via_u128)1 << 40)) return false;
    }
    if (!scale_w_consts(g.bl, q1, Q, r.nl, *sw)) return false;
    sw->bh = (double)dp.bh;
    sw->inv_bh = dp.inv_bh;
    return true;
}

void launch_poly_digits(const RingDev &r1, const Gadget &g_in, const DigitPlan &dp, const double *src,
                        size_t item_stride, size_t limb_stride, int n, int items, int16_t *dig, hipStream_t st) {
    // wide gadgets (the base-18 depth-18 cascade gadget) take the fp64 split
    Gadget g = g_in;
    if (g.f64 == 0 && g.bl >= ((via_u128)1 << 40) && scale_w_make(r1, g, dp, &g.sw)) g.f64 = 3;
    dim3 grid((n + 255) / 256, items);
    if (r1.nl == 2)
        hipLaunchKernelGGL(k_poly_digits<2>, grid, dim3(256), 0, st, r1, g, dp, src, item_stride, limb_stride, n, dig);
    else
        hipLaunchKernelGGL(k_poly_digits<1>, grid, dim3(256), 0, st, r1, g, dp, src, item_stride, limb_stride, n, dig);
}

void launch_lwe_digits(const RingDev &r1, const Gadget &g, const DigitPlan &dp, const double *lwe, int items,
                       int16_t *dig, hipStream_t st) {
    launch_poly_digits(r1, g, dp, lwe, (size_t)r1.nl * (r1.n + 1), (size_t)r1.n + 1, r1.n, items, dig, st);
}

// Gadget-product multiply-accumulate over a digit range (RLevEval::
// gadget_product, rlev.rs:259-283, without the inverse): workgroup (item,
// group, tile) adds NTT(digit_k) (.) key[L-1-k] for k in its group. Inputs
// that are embedded (x -> x^2) transform n/2 coefficients. LSPLIT (NL = 2):
// workgroup row blockIdx.y handles limb blockIdx.y alone (the two limbs'
// transforms and products are independent), holding half the accumulators
// and LDS of the two-limb form for the same total work; it also issues each
// digit's key-row loads before that digit's transform (the transform's
// barriers wait on LDS only), so their latency hides behind it.
template <int NL, int MAXR, int TPB, int LSPLIT = 0>
__global__ __launch_bounds__(TPB) void k_gadget_mac(RingDev r, Gadget g, MacArgs a) {
    constexpr bool PF = LSPLIT != 0;
    constexpr int PER = kSlots / TPB;  // slots per thread
    constexpr int NC = LSPLIT ? 1 : NL;  // limbs this workgroup computes
    __shared__ double s[lds_words(NC * kSlots)];
    const MacTask task = mac_task(blockIdx.x, a.items, a.items_per_key, a.groups, a.tiles);
    const int item = task.item, grp = task.grp, tile = task.tile;
    const int lb = LSPLIT ? (int)blockIdx.y : 0;  // first limb of the workgroup
    const int S = kSlots;
    const int cin = a.embed ? S / 2 : S;
    const int nt = a.embed ? r.n / 2 : r.n;
    const int lognt = a.embed ? r.logn - 1 : r.logn;
    NttLimbs T = limbs_fwd(r);
    if (LSPLIT) {
        T.tw[0] = T.tw[1] = r.t[lb].fwd;
        T.m[0] = T.m[1] = r.t[lb].m;
    }
    const FMod m0 = r.t[lb].m, m1 = r.t[NC == 2 ? 1 : lb].m;
    const int k0 = grp * g.L / a.groups, k1 = (grp + 1) * g.L / a.groups;
    double am[NC][PER], ab[NC][PER];
#pragma unroll
    for (int l = 0; l < NC; l++)
#pragma unroll
        for (int e = 0; e < PER; e++) am[l][e] = ab[l][e] = 0.0;
    const int16_t *dig = a.dig + (size_t)item * g.L * a.dig_coeffs + (size_t)tile * cin;
    const double *key = a.key[item / a.items_per_key] + (size_t)tile * S;
    for (int k = k0; k < k1; k++) {
        const int ks = g.L - 1 - k;
        double kmr[PF ? NC : 1][PF ? PER : 1], kbr[PF ? NC : 1][PF ? PER : 1];
        if constexpr (PF) {
#pragma unroll
            for (int l = 0; l < NC; l++) {
                const double *km = key + ((size_t)(ks * 2 + 0) * NL + lb + l) * a.key_slots;
                const double *kb = key + ((size_t)(ks * 2 + 1) * NL + lb + l) * a.key_slots;
#pragma unroll
                for (int e = 0; e < PER; e++) {
                    kmr[l][e] = km[threadIdx.x + e * TPB];
                    kbr[l][e] = kb[threadIdx.x + e * TPB];
                }
            }
        }
        const int16_t *dk = dig + (size_t)k * a.dig_coeffs;
        for (int c = threadIdx.x; c < cin; c += blockDim.x) {
            // signed digits are valid NTT inputs (the lazy transform takes (-2q, 2q))
            const double d = (double)dk[c];
            s[lds_pad(c)] = d;
            if (NC == 2) s[lds_pad(cin + c)] = d;
        }
        __syncthreads();
        ntt_fwd<NC, MAXR>(s, cin, nt, lognt, T);
#pragma unroll
        for (int l = 0; l < NC; l++) {
            const FMod md = l ? m1 : m0;
            const double *km = key + ((size_t)(ks * 2 + 0) * NL + lb + l) * a.key_slots;
            const double *kb = key + ((size_t)(ks * 2 + 1) * NL + lb + l) * a.key_slots;
#pragma unroll
            for (int e = 0; e < PER; e++) {
                const int sg = threadIdx.x + e * TPB;
                const double x = s[lds_pad(l * cin + (a.embed ? (sg >> 1) : sg))];
                if constexpr (PF) {
                    am[l][e] += mod_mul_lazy(x, kmr[l][e], md);
                    ab[l][e] += mod_mul_lazy(x, kbr[l][e], md);
                } else {
                    am[l][e] += mod_mul_lazy(x, km[sg], md);
                    ab[l][e] += mod_mul_lazy(x, kb[sg], md);
                }
            }
        }
        __syncthreads();
    }
    const size_t ot = ((size_t)grp * a.items + item) * a.tiles + tile;
#pragma unroll
    for (int l = 0; l < NC; l++)
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int sg = threadIdx.x + e * TPB;
            const FMod md = l ? m1 : m0;
            am[l][e] = mod_reduce_small(am[l][e], md);
            ab[l][e] = mod_reduce_small(ab[l][e], md);
            a.out_mask[(ot * NL + lb + l) * S + sg] = am[l][e];
            if (!a.reduce_body) a.out_body[(ot * NL + lb + l) * S + sg] = ab[l][e];
            else s[lds_pad(l * S + sg)] = ab[l][e];
        }
    if (!a.reduce_body) return;
    __syncthreads();
    // Sum the S/n segment bodies of the tile (exact modular additions).
    const int n = r.n;
    for (int half = (S / n) >> 1; half >= 1; half >>= 1) {
        for (int t = threadIdx.x; t < NC * half * n; t += blockDim.x) {
            const int l = t / (half * n);
            const int rem = t - l * half * n;
            const FMod md = l ? m1 : m0;
            s[lds_pad(l * S + rem)] = mod_add(s[lds_pad(l * S + rem)], s[lds_pad(l * S + half * n + rem)], md);
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < NC * n; t += blockDim.x) {
        const int l = t / n, c = t - l * n;
        a.out_body[(ot * NL + lb + l) * n + c] = s[lds_pad(l * S + c)];
    }
}

// The same multiply-accumulate with G digit transforms side by side: 1024
// threads, one LDS buffer per digit of a round (G digits per round, rounds
// until the group's digits are done), each transform by 1024/G threads,
// then one multiply-accumulate pass over the G transforms. Fewer
// workgroups carry more digits each, with one transform's latency per round
// instead of one per digit.
template <int NL, int G>
__global__ __launch_bounds__(1024) void k_gadget_mac_wide(RingDev r, Gadget g, MacArgs a) {
    constexpr int S = kSlots, TPB = 1024, BW = lds_words(NL * kSlots), PER = NL * kSlots / TPB;
    __shared__ double s[G * BW];
    const MacTask task = mac_task(blockIdx.x, a.items, a.items_per_key, a.groups, a.tiles);
    const int item = task.item, grp = task.grp, tile = task.tile;
    const int cin = a.embed ? S / 2 : S;
    const int nt = a.embed ? r.n / 2 : r.n;
    const int lognt = a.embed ? r.logn - 1 : r.logn;
    const NttLimbs T = limbs_fwd(r);
    const FMod m0 = r.t[0].m, m1 = r.t[1].m;
    const int k0 = grp * g.L / a.groups, k1 = (grp + 1) * g.L / a.groups;
    const int tid = threadIdx.x;
    double am[PER], ab[PER];
#pragma unroll
    for (int e = 0; e < PER; e++) am[e] = ab[e] = 0.0;
    const int16_t *dig = a.dig + (size_t)item * g.L * a.dig_coeffs + (size_t)tile * cin;
    const double *key = a.key[item / a.items_per_key] + (size_t)tile * S;
    for (int kb = k0; kb < k1; kb += G) {
        const int nk = min(G, k1 - kb);
        for (int t = tid; t < nk * cin; t += TPB) {
            const int kk = t / cin, c = t - kk * cin;
            const double d = (double)dig[(size_t)(kb + kk) * a.dig_coeffs + c];
            s[kk * BW + lds_pad(c)] = d;
            if (NL == 2) s[kk * BW + lds_pad(cin + c)] = d;
        }
        __syncthreads();
        {
            constexpr int per = TPB / G;
            const int gi = tid / per;
            // buffers past the round's digits only take part in the barriers
            ntt_fwd<NL>(s + gi * BW, cin, nt, lognt, T, gi < nk ? tid - gi * per : (1 << 30), per);
        }
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int t = tid + e * TPB;  // limb-major slot
            const int l = NL == 2 && t >= S ? 1 : 0;
            const int sg = t - l * S;
            const FMod md = l ? m1 : m0;
            const int xi = l * cin + (a.embed ? (sg >> 1) : sg);
            for (int kk = 0; kk < nk; kk++) {
                const int ks = g.L - 1 - (kb + kk);
                const double x = s[kk * BW + lds_pad(xi)];
                const double *km = key + ((size_t)(ks * 2 + 0) * NL + l) * a.key_slots;
                const double *kbp = key + ((size_t)(ks * 2 + 1) * NL + l) * a.key_slots;
                am[e] += mod_mul_lazy(x, km[sg], md);
                ab[e] += mod_mul_lazy(x, kbp[sg], md);
            }
        }
        __syncthreads();
    }
    const size_t ot = ((size_t)grp * a.items + item) * a.tiles + tile;
#pragma unroll
    for (int e = 0; e < PER; e++) {
        const int t = tid + e * TPB;
        const int l = NL == 2 && t >= S ? 1 : 0;
        const int sg = t - l * S;
        const FMod md = l ? m1 : m0;
        a.out_mask[(ot * NL + l) * S + sg] = mod_reduce_small(am[e], md);
        const double b = mod_reduce_small(ab[e], md);
        if (!a.reduce_body) a.out_body[(ot * NL + l) * S + sg] = b;
        else s[lds_pad(l * S + sg)] = b;
    }
    if (!a.reduce_body) return;
    __syncthreads();
    // Sum the S/n segment bodies of the tile (exact modular additions).
    const int n = r.n;
    for (int half = (S / n) >> 1; half >= 1; half >>= 1) {
        for (int t = tid; t < NL * half * n; t += TPB) {
            const int l = t / (half * n);
            const int rem = t - l * half * n;
            const FMod md = l ? m1 : m0;
            s[lds_pad(l * S + rem)] = mod_add(s[lds_pad(l * S + rem)], s[lds_pad(l * S + half * n + rem)], md);
        }
        __syncthreads();
    }
    for (int t = tid; t < NL * n; t += TPB) {
        const int l = t / n, c = t - l * n;
        a.out_body[(ot * NL + l) * n + c] = s[lds_pad(l * S + c)];
    }
}

void launch_gadget_mac(const RingDev &r, const Gadget &g, const MacArgs &a, hipStream_t st) {
    dim3 grid(a.items * a.groups * a.tiles);
    if (a.wide) {
        if (r.nl == 2)
            hipLaunchKernelGGL((k_gadget_mac_wide<2, 4>), grid, dim3(1024), 0, st, r, g, a);
        else
            hipLaunchKernelGGL((k_gadget_mac_wide<1, 8>), grid, dim3(1024), 0, st, r, g, a);
        return;
    }
    // radix-16 passes (a.radix16) trade registers for fewer LDS round trips;
    // 512 threads (a.t512) hold half the accumulators per thread, so more
    // wavefronts fit each SIMD
    if (r.nl == 2 && a.lsplit) {
        // one workgroup per limb (grid row = limb)
        const dim3 g2(grid.x, 2);
        if (a.t512)
            hipLaunchKernelGGL((k_gadget_mac<2, 3, 512, 1>), g2, dim3(512), 0, st, r, g, a);
        else if (a.radix16)
            hipLaunchKernelGGL((k_gadget_mac<2, 4, kThreads, 1>), g2, dim3(kThreads), 0, st, r, g, a);
        else
            hipLaunchKernelGGL((k_gadget_mac<2, 3, kThreads, 1>), g2, dim3(kThreads), 0, st, r, g, a);
    } else if (r.nl == 2) {
        if (a.t512)
            hipLaunchKernelGGL((k_gadget_mac<2, 3, 512>), grid, dim3(512), 0, st, r, g, a);
        else if (a.radix16)
            hipLaunchKernelGGL((k_gadget_mac<2, 4, kThreads>), grid, dim3(kThreads), 0, st, r, g, a);
        else
            hipLaunchKernelGGL((k_gadget_mac<2, 3, kThreads>), grid, dim3(kThreads), 0, st, r, g, a);
    } else {
        if (a.t512)
            hipLaunchKernelGGL((k_gadget_mac<1, 3, 512>), grid, dim3(512), 0, st, r, g, a);
        else if (a.radix16)
            hipLaunchKernelGGL((k_gadget_mac<1, 4, kThreads>), grid, dim3(kThreads), 0, st, r, g, a);
        else
            hipLaunchKernelGGL((k_gadget_mac<1, 3, kThreads>), grid, dim3(kThreads), 0, st, r, g, a);
    }
}

// Cascade combine: masks_next = -iNTT(sum of mask partials), body_next =
// embed(body_prev) - iNTT(sum of body partials); each thread that writes a
// mask coefficient (both limbs) also writes its digits for the next gadget
// product into dig_next (when set). Workgroup (item, chunk)
// handles the chunk's mask polynomials (chunks of at least one degree-n
// polynomial). The partial sums of all digit groups and both tiles are added lazily
// (canonical terms, at most 2 * 18 of them: below 2^44) and brought into
// the inverse transform's input range with one lazy reduction. Workgroups
// y < chunks handle the masks, workgroup y == chunks the body. TPB = 1024
// at the large output degrees (one or two polynomials per item, where the
// workgroups rather than the work per workgroup are scarce), 256 otherwise.
template <int NL, int TPB>
__global__ __launch_bounds__(TPB) void k_cascade_combine(RingDev r, Gadget g, DigitPlan dp, RingDev rq,
                                                         CombineArgs a) {
    // lds_words(NL * max(chunk, n)) words, sized by the launcher: small at
    // the low degrees, so more workgroups share a CU
    extern __shared__ double s[];
    const int item = blockIdx.x, chunk = blockIdx.y, chunks = gridDim.y - 1;
    const int S = kSlots, n = r.n;
    const FMod m0 = r.t[0].m, m1 = r.t[1].m;
    const NttLimbs T = limbs_inv(r);
    const double ni0 = r.t[0].n_inv, ni1 = r.t[1].n_inv;
    if (chunk < chunks) {
        const int CS = S / chunks;  // slots of this chunk (a multiple of n)
        const int c0 = chunk * CS;
        const size_t gstride = (size_t)a.items * 2 * NL * S;
        for (int t = threadIdx.x; t < NL * CS; t += blockDim.x) {
            const int l = (NL == 2 && t >= CS) ? 1 : 0;
            const int c = t - l * CS;
            const double *p = a.pm + (size_t)item * 2 * NL * S + (size_t)l * S + c0 + c;
            double acc = p[0] + p[(size_t)NL * S];
            for (int gi = 1; gi < a.groups; gi++) acc += p[gi * gstride] + p[gi * gstride + (size_t)NL * S];
            s[lds_pad(t)] = lazy_reduce(acc, l ? m1 : m0);
        }
        __syncthreads();
        ntt_inv_noscale<NL>(s, CS, n, r.logn, T);
        for (int c = threadIdx.x; c < CS; c += blockDim.x) {
            const double x0 = mod_neg(mod_mul(s[lds_pad(c)], ni0, m0), m0);
            const double x1 = NL == 2 ? mod_neg(mod_mul(s[lds_pad(CS + c)], ni1, m1), m1) : 0.0;
            double *mo = a.masks_next + (size_t)item * NL * S + c0 + c;
            mo[0] = x0;
            if (NL == 2) mo[S] = x1;
            // the next step's gadget digits of this coefficient (both limbs
            // are in hand; rq carries the RNS constants of the q1 ring)
            if (a.dig_next) write_digits<NL>(x0, x1, rq, g, dp, a.dig_next + (size_t)item * g.L * S + c0 + c, (size_t)S);
        }
        return;
    }
    const size_t gstride_b = (size_t)a.items * 2 * NL * n;
    for (int t = threadIdx.x; t < NL * n; t += blockDim.x) {
        const double *p = a.pb + (size_t)item * 2 * NL * n + t;
        double acc = p[0] + p[(size_t)NL * n];
        for (int gi = 1; gi < a.groups; gi++) acc += p[gi * gstride_b] + p[gi * gstride_b + (size_t)NL * n];
        s[lds_pad(t)] = lazy_reduce(acc, t >= n ? m1 : m0);
    }
    __syncthreads();
    ntt_inv_noscale<NL>(s, n, n, r.logn, T);
    for (int t = threadIdx.x; t < NL * n; t += blockDim.x) {
        const int l = t >= n ? 1 : 0, c = t - l * n;
        const FMod md = l ? m1 : m0;
        const double x = mod_mul(s[lds_pad(t)], l ? ni1 : ni0, md);
        double e = 0.0;
        if ((c & 1) == 0) e = a.body_prev[(size_t)item * a.body_item_stride + (size_t)l * a.body_limb_stride + (c >> 1)];
        a.body_next[((size_t)item * NL + l) * n + c] = mod_sub(e, x, md);
    }
}

void launch_cascade_combine(const RingDev &r, const Gadget &g_in, const DigitPlan &dp, const RingDev &rq,
                            const CombineArgs &a, hipStream_t st) {
    // chunks of whole degree-n polys per item: up to 16 when few items
    // leave CUs idle (one query: 36 LWEs), else up to 8
    const int cap = a.items < 128 ? 16 : 8;
    int chunks = kSlots / r.n;
    if (chunks > cap) chunks = cap;
    dim3 grid(a.items, chunks + 1);
    const bool wide = chunks <= 2;
    const int cs = kSlots / chunks;
    const size_t lds = sizeof(double) * lds_words(r.nl * (cs > (int)r.n ? cs : (int)r.n));
    // digits of every output coefficient with the scale of launch_poly_digits
    // (rq: the q1 ring at n1, whose RNS and gadget constants the scale uses)
    Gadget g = g_in;
    if (a.dig_next && g.f64 == 0 && g.bl >= ((via_u128)1 << 40) && scale_w_make(rq, g, dp, &g.sw)) g.f64 = 3;
    if (r.nl == 2) {
        if (wide)
            hipLaunchKernelGGL((k_cascade_combine<2, 1024>), grid, dim3(1024), lds, st, r, g, dp, rq, a);
        else
            hipLaunchKernelGGL((k_cascade_combine<2, 256>), grid, dim3(256), lds, st, r, g, dp, rq, a);
    } else {
        if (wide)
            hipLaunchKernelGGL((k_cascade_combine<1, 1024>), grid, dim3(1024), lds, st, r, g, dp, rq, a);
        else
            hipLaunchKernelGGL((k_cascade_combine<1, 256>), grid, dim3(256), lds, st, r, g, dp, rq, a);
    }
}

// ---- rlwe_to_rgsw combine (convert.rs:162-180) -----------------------------------

template <int NL>
__global__ __launch_bounds__(1024) void k_r2g_combine(RingDev r, const double *pm, const double *pb, int groups,
                                                          int items, const double *rlwe_masks,
                                                          const double *rlwe_bodies, double *rgsw, int lq) {
    __shared__ double s[lds_words(kSlots)];
    const int item = blockIdx.x;  // bit * lq + i
    const int part = blockIdx.y;  // 0: mask of neg_s_m[i] (and the copy of m[i]), 1: its body
    const int l = blockIdx.z;     // the limb: one workgroup per limb
    const int bit = item / lq, i = item - bit * lq;
    const int n = r.n;  // == kSlots
    const FMod md = r.t[l].m;
    const size_t pw = (size_t)NL * n;
    double *gp = rgsw + (size_t)bit * 4 * lq * pw;
    double *nsm = gp + (size_t)i * 2 * pw;        // neg_s_m[i]
    double *mm = gp + (size_t)(lq + i) * 2 * pw;  // m[i]
    const double *rm = rlwe_masks + (size_t)item * pw, *rb = rlwe_bodies + (size_t)item * pw;
    const double *src = (part ? pb : pm) + (size_t)item * pw;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const int t = l * n + c;
        double acc = src[t];  // canonical partials, at most 18: a lazy sum below 2^43
        for (int gi = 1; gi < groups; gi++) acc += src[(size_t)gi * items * pw + t];
        s[lds_pad(c)] = lazy_reduce(acc, md);
    }
    __syncthreads();
    NttLimbs T;
    T.tw[0] = T.tw[1] = r.t[l].inv;
    T.m[0] = T.m[1] = md;
    ntt_inv_noscale<1>(s, n, n, r.logn, T);
    const double ni = r.t[l].n_inv;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const int t = l * n + c;
        const double x = mod_mul(s[lds_pad(c)], ni, md);
        if (part == 0) {
            nsm[t] = mod_add(rb[t], x, md);
            mm[t] = rm[t];
            mm[pw + t] = rb[t];
        } else {
            nsm[pw + t] = x;
        }
    }
}

void launch_r2g_combine(const RingDev &r, const double *pm, const double *pb, int groups, const double *rlwe_masks,
                        const double *rlwe_bodies, double *rgsw, int bits, int lq, hipStream_t st) {
    const int items = bits * lq;
    if (r.nl == 2)
        hipLaunchKernelGGL(k_r2g_combine<2>, dim3(items, 2, 2), dim3(1024), 0, st, r, pm, pb, groups, items,
                           rlwe_masks, rlwe_bodies, rgsw, lq);
    else
        hipLaunchKernelGGL(k_r2g_combine<1>, dim3(items, 2, 1), dim3(1024), 0, st, r, pm, pb, groups, items,
                           rlwe_masks, rlwe_bodies, rgsw, lq);
}

// ---- gates: external product with DMux / CMux / CRot combine -------------------

// RGSWCiphertext::external_product (rgsw.rs:108-120) over an eval-form RGSW
// [half][i][part][limb][n], then dmux (mux.rs:129-137), cmux (mux.rs:75-84)
// or the SlotExtract step of crot. One workgroup per output item, n == kSlots;
// TPB threads with kSlots/TPB slots each (more threads for narrow tree levels).
template <int NL, int TPB>
__global__ __launch_bounds__(TPB) void k_gate(RingDev r, Gadget g, GateArgs ga) {
    constexpr int PER = kSlots / TPB;
    __shared__ double s[lds_words(NL * kSlots)];
    const int q = blockIdx.x / ga.items_per_query;
    const int j = blockIdx.x - q * ga.items_per_query;
    const int mode = ga.mode, shift = ga.shift;
    const double *rgsw = ga.rgsw + (size_t)q * ga.rgsw_qstride;
    const double *in = ga.in + (size_t)q * ga.in_qstride;
    double *out = ga.out + (size_t)q * ga.out_qstride;
    const int n = r.n, S = kSlots;
    const size_t pw = (size_t)NL * n;
    const FMod m0 = r.t[0].m, m1 = r.t[1].m;
    double am[NL][PER], ab[NL][PER];
#pragma unroll
    for (int l = 0; l < NL; l++)
#pragma unroll
        for (int e = 0; e < PER; e++) am[l][e] = ab[l][e] = 0.0;
    for (int half = 0; half < 2; half++) {  // 0: ct.mask with -s*m, 1: ct.body with m
        double vlo[PER], vhi[PER];
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int sg = threadIdx.x + e * TPB;
            double b0, c0, b1 = 0.0, c1 = 0.0;
            gate_input<NL>(mode, in, j, half, 0, sg, n, shift, m0, b0, c0);
            if (NL == 2) gate_input<NL>(mode, in, j, half, 1, sg, n, shift, m1, b1, c1);
            digit_split_f64(gadget_scale<NL>(c0, c1, r, g), ga.dp, vlo[e], vhi[e]);
        }
        const double *key = rgsw + (size_t)half * g.L * 2 * pw;
        for (int k = 0; k < g.L; k++) {
#pragma unroll
            for (int e = 0; e < PER; e++) {
                const int sg = threadIdx.x + e * TPB;
                const double d = (double)next_digit(vlo[e], vhi[e], k, ga.dp);
                s[lds_pad(sg)] = d;
                if (NL == 2) s[lds_pad(S + sg)] = d;
            }
            __syncthreads();
            lds_ntt_fwd<NL>(s, S, r);
            const int ks = g.L - 1 - k;
#pragma unroll
            for (int l = 0; l < NL; l++) {
                const FMod md = l ? m1 : m0;
                const double *km = key + (size_t)ks * 2 * pw + (size_t)l * n;
                const double *kb = km + pw;
#pragma unroll
                for (int e = 0; e < PER; e++) {
                    const int sg = threadIdx.x + e * TPB;
                    const double x = s[lds_pad(l * S + sg)];
                    am[l][e] += mod_mul_lazy(x, km[sg], md);
                    ab[l][e] += mod_mul_lazy(x, kb[sg], md);
                }
            }
            __syncthreads();
        }
    }
    // inverse transforms of the product, then the gate combine
    for (int part = 0; part < 2; part++) {
#pragma unroll
        for (int l = 0; l < NL; l++) {
            const FMod md = l ? m1 : m0;
#pragma unroll
            for (int e = 0; e < PER; e++)
                s[lds_pad(l * S + threadIdx.x + e * TPB)] = mod_reduce_small(part ? ab[l][e] : am[l][e], md);
        }
        __syncthreads();
        lds_ntt_inv_noscale<NL>(s, S, r);
#pragma unroll
        for (int l = 0; l < NL; l++) {
            const FMod md = l ? m1 : m0;
            const double ni = l ? r.t[1].n_inv : r.t[0].n_inv;
#pragma unroll
            for (int e = 0; e < PER; e++) {
                const int sg = threadIdx.x + e * TPB;
                const double p = mod_mul(s[lds_pad(l * S + sg)], ni, md);
                double bv, cv;
                gate_input<NL>(mode, in, j, part, l, sg, n, shift, md, bv, cv);
                const size_t o = (size_t)part * pw + (size_t)l * n + sg;
                if (mode == GATE_DMUX) {
                    out[(size_t)(2 * j) * 2 * pw + o] = mod_sub(cv, p, md);
                    out[(size_t)(2 * j + 1) * 2 * pw + o] = p;
                } else {
                    out[(size_t)j * 2 * pw + o] = mod_add(bv, p, md);
                }
            }
        }
        __syncthreads();
    }
}

// The same gate with the 2L <= 4 digit transforms of one item running at
// once: 1024 threads, one LDS buffer per (half, digit) polynomial. Phase 1
// scales each input coefficient once and writes its L digits; phase 2
// transforms the 2L digit polynomials concurrently (1024 / 2L threads each);
// phase 3 multiplies and sums against the RGSW rows; phase 4 runs the two
// inverse transforms (mask and body parts) concurrently; phase 5 combines.
// The critical path is one forward and one inverse transform instead of
// 2L + 2 in sequence, which is what the narrow tree levels wait on.
// SPLIT (two limbs): the limbs of a gate are independent once the digits are
// known, so workgroup 2g + l computes gate g's digits from both limbs and
// transforms, multiplies and writes limb l only: half the work and latency
// per workgroup, twice the workgroups on the narrow levels.
// The body of one fused gate: `blk` is the gate's workgroup index (gate blk,
// or with SPLIT gate blk >> 1 and limb blk & 1), `s` the workgroup's LDS
// (lds_words(4 * NLC * kSlots) words). Shared by k_gate_fused (one launch per
// tree level) and k_tree_persistent (every level in one launch).
template <int NL, int SPLIT>
__device__ __forceinline__ void gate_fused_body(double *s, const RingDev &r, const Gadget &g, const GateArgs &ga,
                                                int blk) {
    constexpr int S = kSlots;
    constexpr int TPB = 1024;
    constexpr int NLC = SPLIT ? 1 : NL;  // limbs this workgroup computes
    const int gate = SPLIT ? (blk >> 1) : blk;
    const int lsel = SPLIT ? (blk & 1) : 0;
    const int q = gate / ga.items_per_query;
    const int j = gate - q * ga.items_per_query;
    const int mode = ga.mode, shift = ga.shift;
    const double *rgsw = ga.rgsw + (size_t)q * ga.rgsw_qstride;
    const double *in = ga.in + (size_t)q * ga.in_qstride;
    double *out = ga.out + (size_t)q * ga.out_qstride;
    const int n = r.n;
    const size_t pw = (size_t)NL * n;
    const FMod m0 = r.t[0].m, m1 = r.t[1].m;
    RingDev rl = r;  // the transforms of the computed limb(s)
    if (SPLIT && lsel) rl.t[0] = r.t[1];
    const int L = g.L;
    const int tid = threadIdx.x;
    // phase 1: digits of ct's two halves (0: mask, with -s*m; 1: body, with m)
    {
        const int half = tid >> 9;
        for (int e = 0; e < S / 512; e++) {
            const int c = (tid & 511) + 512 * e;
            double b0, c0, b1 = 0.0, c1 = 0.0;
            gate_input<NL>(mode, in, j, half, 0, c, n, shift, m0, b0, c0);
            if (NL == 2) gate_input<NL>(mode, in, j, half, 1, c, n, shift, m1, b1, c1);
            double lo, hi;
            digit_split_f64(gadget_scale<NL>(c0, c1, r, g), ga.dp, lo, hi);
            for (int k = 0; k < L; k++) {
                const double d = (double)next_digit(lo, hi, k, ga.dp);
                const int b0 = (half * L + k) * NLC * S;
                s[lds_pad(b0 + c)] = d;
                if (NLC == 2) s[lds_pad(b0 + S + c)] = d;
            }
        }
    }
    // One computed limb (2 slots per thread): the RGSW rows of phase 3 and
    // the inputs of phase 5 are loaded ahead, so that their latency hides
    // behind the transforms instead of following them.
    constexpr int PER = NLC * S / TPB;
    constexpr bool PF = NLC == 1;
    double kv[PF ? PER : 1][2][2][2];  // [slot][half][digit][part]
    if constexpr (PF) {
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int t = tid + e * TPB;
#pragma unroll
            for (int half = 0; half < 2; half++)
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    if (k >= L) continue;
                    const double *key =
                        rgsw + ((size_t)half * L + (L - 1 - k)) * 2 * pw + (size_t)(SPLIT ? lsel : 0) * n + t;
                    kv[e][half][k][0] = key[0];
                    kv[e][half][k][1] = key[pw];
                }
        }
    }
    __syncthreads();
    // phase 2: the 2L forward transforms side by side
    {
        const int np = 2 * L;
        const int per = TPB / np;
        const int gi = tid / per;
        ntt_fwd<NLC>(s + lds_words(gi * NLC * S), S, n, r.logn, limbs_fwd(rl), tid - gi * per, per);
    }
    // phase 3: external product sums (both parts) for NLC * S slots per part
    double am[PER], ab[PER];
#pragma unroll
    for (int e = 0; e < PER; e++) {
        const int t = tid + e * TPB;  // limb-major slot
        const int l = SPLIT ? lsel : (NL == 2 && t >= S ? 1 : 0);
        const int c = SPLIT ? t : t - l * S;
        const FMod md = l ? m1 : m0;
        double sm = 0.0, sb = 0.0;
        if constexpr (PF) {
#pragma unroll
            for (int half = 0; half < 2; half++)
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    if (k >= L) continue;
                    const double x = s[lds_pad((half * L + k) * NLC * S + t)];
                    sm += mod_mul_lazy(x, kv[e][half][k][0], md);
                    sb += mod_mul_lazy(x, kv[e][half][k][1], md);
                }
        } else {
            for (int half = 0; half < 2; half++)
                for (int k = 0; k < L; k++) {
                    const double x = s[lds_pad((half * L + k) * NLC * S + t)];
                    const double *key = rgsw + ((size_t)half * L + (L - 1 - k)) * 2 * pw + (size_t)l * n + c;
                    sm += mod_mul_lazy(x, key[0], md);
                    sb += mod_mul_lazy(x, key[pw], md);
                }
        }
        am[e] = mod_reduce_small(sm, md);
        ab[e] = mod_reduce_small(sb, md);
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; e++) {
        const int t = tid + e * TPB;
        s[lds_pad(t)] = am[e];
        s[lds_pad(NLC * S + t)] = ab[e];
    }
    double bvp[2][PF ? PER : 1], cvp[2][PF ? PER : 1];
    if constexpr (PF) {
#pragma unroll
        for (int part = 0; part < 2; part++)
#pragma unroll
            for (int e = 0; e < PER; e++) {
                const int t = tid + e * TPB;
                const int l = SPLIT ? lsel : 0;
                gate_input<NL>(mode, in, j, part, l, t, n, shift, l ? m1 : m0, bvp[part][e], cvp[part][e]);
            }
    }
    __syncthreads();
    // phase 4: inverse transforms of the two parts side by side
    {
        const int part = tid >> 9;
        ntt_inv_noscale<NLC>(s + lds_words(part * NLC * S), S, n, r.logn, limbs_inv(rl), tid & 511, 512);
    }
    // phase 5: gate combine
#pragma unroll
    for (int part = 0; part < 2; part++)
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int t = tid + e * TPB;
            const int l = SPLIT ? lsel : (NL == 2 && t >= S ? 1 : 0);
            const int c = SPLIT ? t : t - l * S;
            const FMod md = l ? m1 : m0;
            const double ni = l ? r.t[1].n_inv : r.t[0].n_inv;
            const double p = mod_mul(s[lds_pad(part * NLC * S + t)], ni, md);
            double bv, cv;
            if constexpr (PF) {
                bv = bvp[part][e];
                cv = cvp[part][e];
            } else {
                gate_input<NL>(mode, in, j, part, l, c, n, shift, md, bv, cv);
            }
            const size_t o = (size_t)part * pw + (size_t)l * n + c;
            if (mode == GATE_DMUX) {
                out[(size_t)(2 * j) * 2 * pw + o] = mod_sub(cv, p, md);
                out[(size_t)(2 * j + 1) * 2 * pw + o] = p;
            } else {
                out[(size_t)j * 2 * pw + o] = mod_add(bv, p, md);
            }
        }
}

template <int NL, int SPLIT>
__global__ __launch_bounds__(1024) void k_gate_fused(RingDev r, Gadget g, GateArgs ga) {
    __shared__ double s[lds_words(4 * (SPLIT ? 1 : NL) * kSlots)];
    gate_fused_body<NL, SPLIT>(s, r, g, ga, (int)blockIdx.x);
}

// ---- persistent tree: every level of a DMux or CMux tree in one launch --------------

// Grid-wide barrier over co-resident workgroups (cooperative launch):
// bar[0] counts arrivals, bar[1] is the generation. Each wait polls at most
// `spin` times; a longer wait sets *fail and the workgroup goes on (the host
// discards the results and reports the failure), so every wave always exits.
__device__ __forceinline__ void tree_sync(unsigned *bar, unsigned &gen, long spin, int *fail) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned g = gen;
        if (__hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long i = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (++i > spin) {
                    __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        gen = g + 1;
    }
    __syncthreads();
}

// Levels lev0 .. of one tree for every query: level v runs its gates over the
// grid (workgroup b takes gates b, b + gridDim.x, ...), then the grid waits
// for the level before the next one reads its outputs.
template <int NL, int SPLIT>
__global__ __launch_bounds__(1024) void k_tree_persistent(RingDev r, Gadget g, TreeArgs ta) {
    __shared__ double s[lds_words(4 * (SPLIT ? 1 : NL) * kSlots)];
    // the generation before this launch's first barrier (no workgroup can
    // complete that barrier before every one has read it)
    unsigned gen = __hip_atomic_load(&ta.bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int v = 0; v < ta.nlev; v++) {
        GateArgs ga = ta.gate;
        ga.rgsw = ta.lev[v].rgsw;
        ga.in = ta.lev[v].in;
        ga.out = ta.lev[v].out;
        ga.items_per_query = ta.lev[v].items_per_query;
        const int blocks = ga.items_per_query * ga.queries * (SPLIT ? 2 : 1);
        for (int b = blockIdx.x; b < blocks; b += gridDim.x) {
            gate_fused_body<NL, SPLIT>(s, r, g, ga, b);
            __syncthreads();  // the next gate's phase 1 rewrites the LDS
        }
        if (v + 1 < ta.nlev) tree_sync(ta.bar, gen, ta.spin, ta.fail);
    }
}

// The gate gadget with its fp64 scale where that is exact (one limb) or
// decided (two limbs).
static Gadget gate_gadget(const RingDev &r, const Gadget &g_in) {
    Gadget g = g_in;
    g.f64 = 0;
    if (exact_paths_only()) {
        // the 128-bit scale for every coefficient
    } else if (r.nl == 2 && g.bl < ((via_u128)1 << 40)) {
        g.f64 = 2;
        g.sf = scale_f64_make((double)(uint64_t)g.bl, r.rns.q1, u128_to_f64(r.rns.big_q));
    } else if (r.nl == 1 && !g.wide && ((via_u128)(r.q_single / 2 + 1) * g.bl + r.q_single) < ((via_u128)1 << 53)) {
        g.f64 = 1;
        g.sf.bl = (double)(uint64_t)g.bl;
    }
    return g;
}

bool launch_tree_persistent(const RingDev &r, const Gadget &g_in, const TreeArgs &a_in, hipStream_t st) {
    if (g_in.L > 2 || r.n != kSlots || a_in.nlev < 1 || a_in.nlev > kMaxTreeLevels) return false;
    TreeArgs a = a_in;
    a.gate.dp = digit_plan(g_in.base, g_in.L);
    const Gadget g = gate_gadget(r, g_in);
    // two limbs: one workgroup per limb of each gate (every level)
    const void *fn = r.nl == 2 ? (const void *)k_tree_persistent<2, 1> : (const void *)k_tree_persistent<1, 0>;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 1024, 0) != hipSuccess || per_cu < 1)
        return false;
    int most = 0;  // workgroups of the widest level
    for (int v = 0; v < a.nlev; v++)
        most = std::max(most, a.lev[v].items_per_query * a.gate.queries * (r.nl == 2 ? 2 : 1));
    const int grid = std::max(1, std::min(most, cus * per_cu));
    void *args[] = {(void *)&r, (void *)&g, (void *)&a};
    return hipLaunchCooperativeKernel(fn, dim3(grid), dim3(1024), args, 0, st) == hipSuccess;
}

bool gate_takes_fx(const RingDev &r, const Gadget &g_in, int blocks) {
    bool direct = false;
    return r.nl == 1 && blocks >= kGateFxMin && gate_fx_fits(r, gate_gadget(r, g_in), &direct);
}

void launch_gate(const RingDev &r, const Gadget &g_in, const GateArgs &a_in, hipStream_t st) {
    GateArgs a = a_in;
    a.dp = digit_plan(g_in.base, g_in.L);
    const Gadget g = gate_gadget(r, g_in);
    const int blocks = a.items_per_query * a.queries;
    bool direct = false;
    if (blocks >= kGateFxMin && gate_fx_fits(r, g, &direct)) {
        // wide levels: the fixed-shape gate (two limbs: one workgroup per
        // limb, each computing the digits from both limbs)
        launch_gate_fx(r, g, a, direct, st);
        return;
    }
    if (g.L <= 2 && r.n == kSlots) {
        // two limbs: one workgroup per limb of each gate on levels narrower
        // than the CU count (latency-bound); wider levels keep both limbs in
        // one workgroup (the digits are extracted once)
        const bool split = a.split < 0 ? blocks < 256 : a.split == 1;
        // narrow levels of the fixed shape: the compile-time passes of the
        // fixed gate with 1024 threads per gate (per limb with two limbs)
        if (blocks < kGateFxMin && a.mode != GATE_CMUX_EV && (r.nl == 1 || split) && gate_fx_fits(r, g, &direct)) {
            launch_gate_nx(r, g, a, direct, st);
            return;
        }
        if (r.nl == 2 && split)
            hipLaunchKernelGGL((k_gate_fused<2, 1>), dim3(2 * blocks), dim3(1024), 0, st, r, g, a);
        else if (r.nl == 2)
            hipLaunchKernelGGL((k_gate_fused<2, 0>), dim3(blocks), dim3(1024), 0, st, r, g, a);
        else
            hipLaunchKernelGGL((k_gate_fused<1, 0>), dim3(blocks), dim3(1024), 0, st, r, g, a);
        return;
    }
    const bool wide = blocks < 64;  // narrow levels: shorter per-workgroup latency
    if (r.nl == 2) {
        if (wide)
            hipLaunchKernelGGL((k_gate<2, 1024>), dim3(blocks), dim3(1024), 0, st, r, g, a);
        else
            hipLaunchKernelGGL((k_gate<2, 256>), dim3(blocks), dim3(256), 0, st, r, g, a);
    } else {
        if (wide)
            hipLaunchKernelGGL((k_gate<1, 1024>), dim3(blocks), dim3(1024), 0, st, r, g, a);
        else
            hipLaunchKernelGGL((k_gate<1, 256>), dim3(blocks), dim3(256), 0, st, r, g, a);
    }
}

// First DMux level on trivial(Delta) = (0, Delta): the mask's digits are all
// zero and the body's digits are the constants d_k of Delta at coefficient 0,
// so the external product is sum_k d_k * m[L-1-k] in coefficient form.
__global__ void k_dmux_trivial(RingDev r, const double *rgsw_all, size_t rgsw_qstride, int L, DeltaDigits dd,
                               double *out_all, size_t out_qstride) {
    const int n = r.n;
    const size_t pw = (size_t)2 * n;  // NL = 2
    const size_t total = 2 * pw;      // mask and body
    const int q = blockIdx.y;
    const double *rgsw_coeff = rgsw_all + (size_t)q * rgsw_qstride;
    double *out = out_all + (size_t)q * out_qstride;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int part = (int)(t / pw);
        const size_t rem = t - part * pw;
        const int l = (int)(rem / n), c = (int)(rem - (size_t)l * n);
        const FMod md = l ? r.t[1].m : r.t[0].m;
        double acc = 0.0;
        for (int k = 0; k < L; k++) {
            const double *m_k = rgsw_coeff + (size_t)(L + (L - 1 - k)) * 2 * pw;  // m half, sample L-1-k
            acc += mod_mul_lazy(digit_residue(dd.d[k], md), m_k[part * pw + rem], md);
        }
        const double p = mod_reduce_small(acc, md);
        double base = 0.0;
        if (part == 1 && c == 0) base = l ? dd.delta[1] : dd.delta[0];
        out[t] = mod_sub(base, p, md);  // out[0] = trivial - p
        out[total + t] = p;             // out[1] = p
    }
}

void launch_dmux_trivial(const RingDev &r, const double *rgsw_coeff, size_t rgsw_qstride, int L, const DeltaDigits &dd,
                         double *out, size_t out_qstride, int queries, hipStream_t st) {
    hipLaunchKernelGGL(k_dmux_trivial, dim3(32, queries), dim3(256), 0, st, r, rgsw_coeff, rgsw_qstride, L, dd, out,
                       out_qstride);
}

// ---- mod switch q1 -> q2 (mod_switch.rs:57-86) ----------------------------------

// Group blockIdx.y reads in + y*in_gstride, writes out + y*out_gstride.
// fast: modswitch_f64 first (the 128-bit path for the values it declines)
__global__ void k_modswitch_q1_q2(RingDev r1, SwitchIn si, const double *in, size_t in_gstride, double *out,
                                  size_t out_gstride, size_t polys) {
    const int n = r1.n;
    const size_t total = polys * n;
    in += (size_t)blockIdx.y * in_gstride;
    out += (size_t)blockIdx.y * out_gstride;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t p = t >> r1.logn, c = t & (size_t)(n - 1);  // n a power of two
        const double a0 = in[p * 2 * n + c], a1 = in[p * 2 * n + n + c];
        double yf;
        if (si.fast && modswitch_f64(a0, a1, r1.rns, si.sf, yf)) {
            out[t] = yf;
            continue;
        }
        const via_u128 x = rns_reconstruct(a0, a1, r1.rns);
        uint64_t y = (uint64_t)round_mul_div(x, (via_u128)si.q2, si.qd);
        if (y >= si.q2) y -= si.q2;
        out[t] = (double)y;
    }
}

SwitchIn switch_in(const RingDev &r1, uint64_t q2) {
    SwitchIn si;
    si.rns = r1.rns;
    si.qd = r1.qdiv;
    si.q2 = q2;
    si.sf.q2 = (double)q2;
    si.sf.inv_q1 = 1.0 / (double)r1.rns.q1;
    si.sf.q2_Q = (double)q2 / u128_to_f64(r1.rns.big_q);
    si.fast = !exact_paths_only() && r1.nl == 2 && q2 < r1.rns.q1 && r1.rns.q1 < kMaxModulus &&
              r1.rns.q0 < kMaxModulus;
    return si;
}

void launch_modswitch_q1_q2_groups(const RingDev &r1, uint64_t q2, const double *in, size_t in_gstride, double *out,
                                   size_t out_gstride, size_t polys, int groups, hipStream_t st) {
    const size_t total = polys * r1.n;
    if (!total || groups < 1) return;
    size_t blocks = (total + 255) / 256;
    const size_t cap = std::max<size_t>(1, 16384 / (size_t)groups);
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_modswitch_q1_q2, dim3((unsigned)blocks, groups), dim3(256), 0, st, r1, switch_in(r1, q2), in,
                       in_gstride, out, out_gstride, polys);
}

void launch_modswitch_q1_q2(const RingDev &r1, uint64_t q2, const double *in, double *out, size_t polys,
                            hipStream_t st) {
    launch_modswitch_q1_q2_groups(r1, q2, in, 0, out, 0, polys, 1, st);
}

// ---- RespComp (resp_comp.rs:46-76) -----------------------------------------------------

// One workgroup of 1024 threads per query. mod_switch_sym q2 -> q3 of the
// mask (n1 coefficients, kept in LDS) and of the projected body; then the
// ring switch (ring_switch.rs:358-385): for every j < d the projection
// pi_0(mask * X^j) is gadget-decomposed (base b_rsk, l_rsk digits) and its
// digit polynomials of degree n2 are multiplied with rsk_j. All d * l_rsk
// digit transforms of a round (as many as the LDS holds, 32 at the paper
// sizes) run side by side, then one multiply-accumulate pass and the two
// inverse transforms (mask and body parts) side by side; finally
// mod_switch_asym (mod_switch.rs:124-154) to mask mod q3, body mod q4.
constexpr int kRespBufWords = 16384;  // digit polynomials of one round

// floor(N / d) for any 64-bit N, with minv = floor((2^64 - 1) / d): the
// estimate umulhi(N, minv) is at most two below the quotient (minv is below
// 2^64 / d by less than one, so N minv / 2^64 is below N / d by less than
// N / 2^64 < 1). A few instructions instead of a 64-bit division.
__device__ __forceinline__ uint64_t udiv_pre(uint64_t N, uint64_t d, uint64_t minv) {
    uint64_t qh = __umul64hi(N, minv);
    uint64_t r = N - qh * d;
    if (r >= d) {
        qh++;
        r -= d;
    }
    if (r >= d) qh++;
    return qh;
}

// round(x * qt / qs) mod qt for x < qs, i.e. floor((x qt + qs / 2) / qs),
// which is at most qt (mod_switch.rs rounding)
__device__ __forceinline__ uint64_t switch_round(uint64_t x, uint64_t qt, uint64_t qs, uint64_t minv_s) {
    const uint64_t v = udiv_pre(x * qt + (qs >> 1), qs, minv_s);
    return v >= qt ? v - qt : v;
}

__global__ __launch_bounds__(1024) void k_resp_comp(RespArgs a) {
    __shared__ double c3m[kSlots];           // mask at q3, degree n1
    __shared__ double buf[lds_words(kRespBufWords)];  // digit polynomials, then the two parts
    const int qy = blockIdx.x;
    const int n1 = a.n1, n2 = a.n2, d = a.d, L = a.l_rsk;
    const int logn2 = 31 - __clz(n2);
    const uint64_t q2 = a.q2, q3 = a.q3, q4 = a.q4;
    const FMod m3 = a.t3.m;
    const double *ct = a.ct + (size_t)qy * 2 * n1;
    for (int c = threadIdx.x; c < n1; c += blockDim.x) {
        c3m[c] = (double)switch_round((uint64_t)ct[c], q3, q2, a.minv2);
    }
    __syncthreads();
    NttLimbs Tf, Ti;
    Tf.tw[0] = Tf.tw[1] = a.t3.fwd;
    Ti.tw[0] = Ti.tw[1] = a.t3.inv;
    Tf.m[0] = Tf.m[1] = Ti.m[0] = Ti.m[1] = m3;
    uint64_t bl = 1;
    for (int k = 0; k < L; k++) bl *= a.b_rsk;
    const int npoly = d * L;
    const int per_round = kRespBufWords / n2;
    const size_t kw = (size_t)L * 2 * n2;  // key words per j
    const double *rsk = a.rsk_eval[qy];
    // accumulators of the outputs (part, i) this thread owns: t = part * n2 + i
    double acc[2] = {0.0, 0.0};
    for (int p0 = 0; p0 < npoly; p0 += per_round) {
        const int p1 = min(npoly, p0 + per_round);
        // digits of pi_0(mask * X^j) for the (j, k) of this round
        for (int t = threadIdx.x; t < d * n2; t += blockDim.x) {
            const int j = t / n2, i = t - j * n2;
            if ((j + 1) * L <= p0 || j * L >= p1) continue;
            const int idx = d * i - j;
            const double pv = idx >= 0 ? c3m[idx] : mod_neg(c3m[idx + n1], m3);
            const uint64_t x = (uint64_t)pv;
            const int neg = x > (q3 >> 1);
            const uint64_t mag = neg ? q3 - x : x;
            SMag v;
            v.mag = (via_u128)udiv_pre(mag * bl + (q3 >> 1), q3, a.minv3);
            v.neg = neg && v.mag != 0;
            double lo, hi;
            digit_split_f64(v, a.dp, lo, hi);
            for (int k = 0; k < L; k++) {
                const int dg = next_digit(lo, hi, k, a.dp);
                const int pidx = j * L + k;
                if (pidx >= p0 && pidx < p1) buf[lds_pad((pidx - p0) * n2 + i)] = (double)dg;
            }
        }
        __syncthreads();
        ntt_fwd<1>(buf, (p1 - p0) * n2, n2, logn2, Tf);
        for (int e = 0; e < 2; e++) {
            const int t = threadIdx.x + e * blockDim.x;
            if (t >= 2 * n2) break;
            const int part = t >= n2 ? 1 : 0, i = t - part * n2;
            double sacc = 0.0;
            for (int pidx = p0; pidx < p1; pidx++) {
                const int j = pidx / L, k = pidx - j * L;
                const int ks = L - 1 - k;
                sacc += mod_mul_lazy(buf[lds_pad((pidx - p0) * n2 + i)],
                                     rsk[(size_t)j * kw + (size_t)ks * 2 * n2 + part * n2 + i], m3);
            }
            acc[e] = mod_add(acc[e], mod_reduce_small(sacc, m3), m3);
        }
        __syncthreads();
    }
    for (int e = 0; e < 2; e++) {
        const int t = threadIdx.x + e * blockDim.x;
        if (t < 2 * n2) buf[lds_pad(t)] = acc[e];
    }
    __syncthreads();
    ntt_inv_noscale<1>(buf, 2 * n2, n2, logn2, Ti);
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        const double rm = mod_mul(buf[lds_pad(i)], a.t3.n_inv, m3);
        const double rb = mod_mul(buf[lds_pad(n2 + i)], a.t3.n_inv, m3);
        const double body_proj = (double)switch_round((uint64_t)ct[n1 + d * i], q3, q2, a.minv2);
        const uint64_t mask = (uint64_t)mod_neg(rm, m3);
        const uint64_t body = (uint64_t)mod_sub(body_proj, rb, m3);
        a.out_mask[qy][i] = switch_round(mask, q3, q3, a.minv3);
        a.out_body[qy][i] = switch_round(body, q4, q3, a.minv3);
    }
}

// RespComp over d + 1 workgroups per query where a partial buffer is given:
// k_resp_part, one 1024-thread workgroup per (query, j), switches the n2
// coefficients of pi_0(mask * X^j) it needs straight from the q2 mask, forms
// their l_rsk digit polynomials, transforms them and writes the reduced sum
// of their products with rsk_j (both parts) to the partial buffer; then
// k_resp_fin, one per query, adds the d partials (the canonical residue of
// the same sum k_resp_comp forms), runs the two inverse transforms and the
// final switches as k_resp_comp does.
constexpr int kRespPartWords = 8192;  // digit polynomials of one j

__global__ __launch_bounds__(1024) void k_resp_part(RespArgs a) {
    __shared__ double buf[lds_words(kRespPartWords)];
    const int d = a.d, L = a.l_rsk, n1 = a.n1, n2 = a.n2;
    const int qy = blockIdx.x / d, j = blockIdx.x - qy * d;
    const int logn2 = 31 - __clz(n2);
    const uint64_t q2 = a.q2, q3 = a.q3;
    const FMod m3 = a.t3.m;
    const double *ct = a.ct + (size_t)qy * 2 * n1;
    uint64_t bl = 1;
    for (int k = 0; k < L; k++) bl *= a.b_rsk;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        const int idx = d * i - j;
        const double sv = (double)switch_round((uint64_t)ct[idx >= 0 ? idx : idx + n1], q3, q2, a.minv2);
        const double pv = idx >= 0 ? sv : mod_neg(sv, m3);
        const uint64_t x = (uint64_t)pv;
        const int neg = x > (q3 >> 1);
        const uint64_t mag = neg ? q3 - x : x;
        SMag v;
        v.mag = (via_u128)udiv_pre(mag * bl + (q3 >> 1), q3, a.minv3);
        v.neg = neg && v.mag != 0;
        double lo, hi;
        digit_split_f64(v, a.dp, lo, hi);
        for (int k = 0; k < L; k++) buf[lds_pad(k * n2 + i)] = (double)next_digit(lo, hi, k, a.dp);
    }
    __syncthreads();
    NttLimbs Tf;
    Tf.tw[0] = Tf.tw[1] = a.t3.fwd;
    Tf.m[0] = Tf.m[1] = m3;
    ntt_fwd<1>(buf, L * n2, n2, logn2, Tf);
    const size_t kw = (size_t)L * 2 * n2;  // key words per j
    const double *rsk = a.rsk_eval[qy] + (size_t)j * kw;
    double *out = a.partials + ((size_t)qy * d + j) * 2 * n2;
    for (int t = threadIdx.x; t < 2 * n2; t += blockDim.x) {
        const int part = t >= n2 ? 1 : 0, i = t - part * n2;
        double sacc = 0.0;
        for (int k = 0; k < L; k++)
            sacc += mod_mul_lazy(buf[lds_pad(k * n2 + i)], rsk[(size_t)(L - 1 - k) * 2 * n2 + part * n2 + i], m3);
        out[t] = mod_reduce_small(sacc, m3);
    }
}

__global__ __launch_bounds__(1024) void k_resp_fin(RespArgs a) {
    __shared__ double buf[lds_words(2 * 2048)];
    const int qy = blockIdx.x;
    const int d = a.d, n1 = a.n1, n2 = a.n2;
    const int logn2 = 31 - __clz(n2);
    const uint64_t q2 = a.q2, q3 = a.q3, q4 = a.q4;
    const FMod m3 = a.t3.m;
    const double *ct = a.ct + (size_t)qy * 2 * n1;
    const double *pj = a.partials + (size_t)qy * d * 2 * n2;
    for (int t = threadIdx.x; t < 2 * n2; t += blockDim.x) {
        double acc = pj[t];
        for (int j = 1; j < d; j++) acc = mod_add(acc, pj[(size_t)j * 2 * n2 + t], m3);
        buf[lds_pad(t)] = acc;
    }
    __syncthreads();
    NttLimbs Ti;
    Ti.tw[0] = Ti.tw[1] = a.t3.inv;
    Ti.m[0] = Ti.m[1] = m3;
    ntt_inv_noscale<1>(buf, 2 * n2, n2, logn2, Ti);
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        const double rm = mod_mul(buf[lds_pad(i)], a.t3.n_inv, m3);
        const double rb = mod_mul(buf[lds_pad(n2 + i)], a.t3.n_inv, m3);
        const double body_proj = (double)switch_round((uint64_t)ct[n1 + d * i], q3, q2, a.minv2);
        const uint64_t mask = (uint64_t)mod_neg(rm, m3);
        const uint64_t body = (uint64_t)mod_sub(body_proj, rb, m3);
        a.out_mask[qy][i] = switch_round(mask, q3, q3, a.minv3);
        a.out_body[qy][i] = switch_round(body, q4, q3, a.minv3);
    }
}

void launch_resp_comp(const RespArgs &a_in, hipStream_t st) {
    RespArgs a = a_in;
    a.dp = digit_plan(a.b_rsk, a.l_rsk);
    a.minv2 = ~0ULL / a.q2;
    a.minv3 = ~0ULL / a.q3;
    const bool split = a.partials && a.d >= 2 && a.n2 <= 2048 && (size_t)a.l_rsk * a.n2 <= (size_t)kRespPartWords &&
                       a.partials_words >= (size_t)a.queries * a.d * 2 * a.n2;
    if (split) {
        hipLaunchKernelGGL(k_resp_part, dim3(a.queries * a.d), dim3(1024), 0, st, a);
        hipLaunchKernelGGL(k_resp_fin, dim3(a.queries), dim3(1024), 0, st, a);
        return;
    }
    hipLaunchKernelGGL(k_resp_comp, dim3(a.queries), dim3(1024), 0, st, a);
}

// ---- setup: key relayout and DB packing ----------------------------------------------

__global__ void k_key_relayout(const uint64_t *src, double *dst, int rank_in, int L, int nl, int n) {
    const size_t total = (size_t)rank_in * L * 2 * nl * n;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        // src index t = (((idx*L + k)*2 + part)*nl + l)*n + c
        size_t x = t;
        const int c = (int)(x % n);
        x /= n;
        const int l = (int)(x % nl);
        x /= nl;
        const int part = (int)(x % 2);
        x /= 2;
        const int k = (int)(x % L);
        const size_t idx = x / L;
        const size_t o = (((size_t)k * 2 + part) * nl + l) * ((size_t)rank_in * n) + idx * n + c;
        dst[o] = (double)src[t];
    }
}

void launch_key_relayout(const uint64_t *src, double *dst, int rank_in, int L, int nl, int n, hipStream_t st) {
    const size_t total = (size_t)rank_in * L * 2 * nl * n;
    unsigned blocks = (unsigned)((total + 255) / 256);
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_key_relayout, dim3(blocks), dim3(256), 0, st, src, dst, rank_in, L, nl, n);
}

// records [rec][n_rec] must hold values below p (Poly<N_REC, p>); a word at
// or above p is taken as 0 and sets *bad.
__global__ void k_setup_db(const uint64_t *records, const uint8_t *bytes, size_t n_records, double *cells, int I,
                           int J, int n1, int n_rec, uint64_t p, int *bad) {
    const int d = n1 / n_rec;
    const size_t total = (size_t)I * J * n1;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t cell = t / n1;
        const int pos = (int)(t - cell * n1);
        const int c = pos / d, k = pos - c * d;
        const size_t i = cell / J, j = cell - i * J;
        const size_t flat = (size_t)k * I * J + i * J + j;
        double v = 0.0;
        if (flat < n_records) {
            if (records) {
                const uint64_t x = records[flat * n_rec + c];
                if (x >= p) *bad = 1;
                else v = (double)x;
            } else {
                const uint8_t b = bytes[flat * (n_rec / 2) + (c >> 1)];
                v = (double)((c & 1) ? (b >> 4) : (b & 0xF));
            }
        }
        cells[t] = v;
    }
}

void launch_setup_db(const uint64_t *records, size_t n_records, double *cells, int I, int J, int n1, int n_rec,
                     uint64_t p, int *bad, hipStream_t st) {
    const size_t total = (size_t)I * J * n1;
    unsigned blocks = (unsigned)((total + 255) / 256);
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_setup_db, dim3(blocks), dim3(256), 0, st, records, (const uint8_t *)nullptr, n_records,
                       cells, I, J, n1, n_rec, p, bad);
}

void launch_setup_db_bytes(const uint8_t *bytes, size_t n_records, double *cells, int I, int J, int n1, int n_rec,
                           hipStream_t st) {
    const size_t total = (size_t)I * J * n1;
    unsigned blocks = (unsigned)((total + 255) / 256);
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_setup_db, dim3(blocks), dim3(256), 0, st, (const uint64_t *)nullptr, bytes, n_records,
                       cells, I, J, n1, n_rec, (uint64_t)16, (int *)nullptr);
}

}  // namespace via
