// cascade_mac.hip — the gadget-product multiply-accumulate of the LWE -> RLWE
// cascade steps (RLevEval::gadget_product, rlev.rs:259-283, inside
// conv_step_eval, conv.rs:122-150) in a fixed shape for n1 = 2048 and two
// q1 limbs.
//
// A workgroup (item, digit group, tile) adds NTT(digit_k) (.) key[L-1-k] over
// its digits k, for the 2048 slots of its tile, both RNS limbs: threads
// 0..127 take limb 0 and 128..255 limb 1, so the limb's modulus, twiddle
// table and key rows are wave-uniform. The inputs are embedded (degree
// nt = n_out / 2, 1024 per limb and tile), so thread i ends every transform
// holding outputs 8i .. 8i+7 in registers: slots 16i .. 16i+15, whose key
// words the key layout of this kernel stores so that every load of a
// wavefront is contiguous (cascade_key_pos). Per digit:
//  * the first pass reads the thread's digits straight from the int16 plane
//    and evaluates its stages in direct form (the digits are at most B/2 in
//    magnitude, so every product with a twiddle product is exact and the sums
//    stay below 2^45 without a reduction);
//  * the middle passes run over LDS (two buffers, used by alternate digits,
//    so the last pass of one digit and the first of the next need no barrier);
//  * the last pass (three stages, contiguous outputs) stays in registers and
//    feeds the multiply-accumulate directly.
// The butterflies, twiddles and transform order are those of ntt_lazy.h (the
// reference's DIT forward transform, ring/ntt.rs:305-343); only the grouping of
// exact integer operations differs, so every canonical result is identical.
#include "ntt_fixed.h"

namespace via {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));


// Middle passes P, P+1, .. NP-2 over the LDS buffers of the IPW items (item
// `it` at buf + it * lds_words(2^LOGNIN)): the thread's units w = j + TL v of
// the items' 2^LOGNIN >> R units each, twiddles read per unit, a barrier
// after each.
template <int LOGNT, int RL, int IPW, int LOGNIN, int P>
__device__ __forceinline__ void cm_middle(double *buf, int j, const double *tw, const FMod &m) {
    constexpr int NP = cm_np(LOGNT, RL);
    if constexpr (P < NP - 1) {
        constexpr int R = cm_r(LOGNT, RL, P), L0 = cm_l0(LOGNT, RL, P), E = 1 << R, TL = 128 * IPW;
        constexpr int EPT = (1 << LOGNIN) / 128, IB = lds_words(1 << LOGNIN);
        using G = UnitGeo<LOGNT, L0, R>;
#pragma unroll
        for (int v = 0; v < (EPT >> R); v++) {
            const int w = j + TL * v;
            const int it = w >> (LOGNIN - R), u = w & ((1 << (LOGNIN - R)) - 1);
            int base, SB;
            G::at(u, base, SB);
            double wt[E - 1];
            load_unit_tw<R, L0>(wt, tw, SB);
            double *p = buf + it * IB + lds_pad(base);
            double x[E];
#pragma unroll
            for (int t = 0; t < E; t++) x[t] = p[unit_off<(1 << G::JB)>(t)];
            unit_stages<R>(x, wt, m);
#pragma unroll
            for (int t = 0; t < E; t++) p[unit_off<(1 << G::JB)>(t)] = x[t];
        }
        __syncthreads();
        cm_middle<LOGNT, RL, IPW, LOGNIN, P + 1>(buf, j, tw, m);
    }
}

}  // namespace

// Position of slot sg (0..2047) of a tile in the key layout of
// k_cascade_mac: slot 8j + m sits at (m >> 1) * 512 + 2j + (m & 1). Thread j
// of the two-item kernel owns slots 8j .. 8j+7: its slot pair t is the 16
// bytes at 4096 t + 16 j. Thread i of the one-item kernel owns slots
// 16i .. 16i+15: its slot pairs t and t + 4 are the 32 bytes at
// 4096 t + 32 i. Every load of a wavefront is contiguous either way.
__host__ __device__ int cascade_key_pos(int sg) {
    const int j = sg >> 3, m = sg & 7;
    return ((m >> 1) << 9) | (j << 1) | (m & 1);
}

// One RNS limb (blockIdx.y) of IPW items (LWEs sharing one key) per
// workgroup, TL = 128 IPW threads. EMB: the cascade steps (embedded inputs,
// 1024 per tile, slots 2p and 2p + 1 take transform output p); otherwise the
// RLWE -> RGSW conversion (2048 inputs, one slot per output, no segment
// sums). Thread j ends each transform holding outputs OPT j .. OPT j + OPT - 1
// of every item, i.e. slots SPT j .. SPT j + SPT - 1 (SPT = 16 / IPW either
// way), and each key word it loads serves all IPW items (the key stream per
// item falls with IPW; the last pass shrinks to log2(OPT) stages). Passes
// before the last give every thread EPT elements: units w = j + TL v of the
// items' units.
template <int LOGNT, int IPW, bool EMB>
__global__ __launch_bounds__(128 * IPW, !EMB ? 2 : IPW == 4 ? 4 : 3) void k_cascade_mac(RingDev r, Gadget g, MacArgs a) {
    constexpr int LOGNIN = EMB ? 10 : 11, NIN = 1 << LOGNIN;         // transform inputs per item and tile
    constexpr int TL = 128 * IPW, EPT = NIN / 128;                   // threads; elements per thread and pass
    constexpr int OPT = NIN / TL, SPT = EMB ? 2 * OPT : OPT;         // outputs and slots per item
    constexpr int RL = OPT == 16 ? 4 : OPT == 8 ? 3 : OPT == 4 ? 2 : 1;  // stages of the last pass
    constexpr int NP = cm_np(LOGNT, RL);
    constexpr int S = kSlots;
    constexpr int IB = lds_words(NIN);  // LDS words per item and buffer
    // key words loaded ahead of the transform where that fits the register
    // budget of the occupancy target without spills
    constexpr bool KPF = EMB && (NP <= 1 || (IPW == 2 && LOGNT < 10) || (IPW == 4 && LOGNT < 5));
    // [buffer][item] transform buffers of the limb; at the end the body sums
    // (2048 padded words per item)
    __shared__ double s[2 * IPW * IB];
    const int j = threadIdx.x;
    const int l = blockIdx.y;
    const MacTask task = mac_task(blockIdx.x, a.items / IPW, a.items_per_key / IPW, a.groups, a.tiles);
    const int item0 = task.item * IPW, grp = task.grp, tile = task.tile;
    const FMod m = r.t[l].m;
    const double *tw = r.t[l].fwd;
    const double *dc = r.t[l].dfw;
    const int k0 = grp * g.L / a.groups, k1 = (grp + 1) * g.L / a.groups;
    const size_t dplane = (size_t)g.L * a.dig_coeffs;  // digit words per item
    const int16_t *dig = a.dig + (size_t)item0 * dplane + (size_t)tile * NIN;
    const double2 *key = reinterpret_cast<const double2 *>(a.key[item0 / a.items_per_key] + (size_t)tile * S +
                                                           (size_t)l * a.key_slots);
    const size_t krow = a.key_slots;  // double2 stride between part rows (two limbs of key_slots words each)
    double am[IPW][SPT], ab[IPW][SPT];
#pragma unroll
    for (int it = 0; it < IPW; it++)
#pragma unroll
        for (int e = 0; e < SPT; e++) am[it][e] = ab[it][e] = 0.0;
    // the digits of the thread's first-pass elements, one digit ahead
    constexpr int R1 = cm_r(LOGNT, RL, 0);
    int dcur[EPT], dnxt[EPT];
    auto load_digits = [&](int k, int *d) {
        const int16_t *dk = dig + (size_t)k * a.dig_coeffs;
        if constexpr (NP <= 1) {
            // OPT contiguous digits of each item
#pragma unroll
            for (int it = 0; it < IPW; it++) {
                const int16_t *p = dk + it * dplane + OPT * j;
#pragma unroll
                for (int t = 0; t < OPT; t++) d[it * OPT + t] = p[t];
            }
        } else {
            using Gm = UnitGeo<LOGNT, 0, R1>;
#pragma unroll
            for (int v = 0; v < (EPT >> R1); v++) {
                const int w = j + TL * v;
                const int it = w >> (LOGNIN - R1), u = w & ((NIN >> R1) - 1);
                int base, SB;
                Gm::at(u, base, SB);
#pragma unroll
                for (int t = 0; t < (1 << R1); t++) d[(v << R1) + t] = dk[it * dplane + base + (t << Gm::JB)];
            }
        }
    };
    if (k0 < k1) load_digits(k0, dcur);
    for (int k = k0; k < k1; k++) {
        const int ks = g.L - 1 - k;
        // this digit's key words (the thread's slot pairs): loaded here, in
        // flight during the transform, where the registers allow (KPF), else
        // at the multiply-accumulate
        const double2 *km = key + (size_t)ks * 2 * krow;
        const double2 *kb = km + krow;
        constexpr int NPAIR = SPT / 2;
        double2 vm[NPAIR], vb[NPAIR];
        auto load_keys = [&]() {
#pragma unroll
            for (int t = 0; t < NPAIR; t++) {
                const int pos = cascade_key_pos(SPT * j + 2 * t) >> 1;
                vm[t] = km[pos];
                vb[t] = kb[pos];
            }
        };
        if constexpr (KPF) load_keys();
        if (k + 1 < k1) load_digits(k + 1, dnxt);
        double x[IPW][OPT];
        if constexpr (NP <= 1) {
#pragma unroll
            for (int it = 0; it < IPW; it++) {
#pragma unroll
                for (int t = 0; t < OPT; t++) x[it][t] = (double)dcur[it * OPT + t];
                if constexpr (NP == 1) {
#pragma unroll
                    for (int sgm = 0; sgm < (OPT >> LOGNT); sgm++) direct_pass<LOGNT>(x[it] + (sgm << LOGNT), dc);
                }
            }
        } else {
            double *buf = s + (k & 1) * IPW * IB;
            {
                // first pass: stages 0 .. R1-1 in direct form
                using Gm = UnitGeo<LOGNT, 0, R1>;
#pragma unroll
                for (int v = 0; v < (EPT >> R1); v++) {
                    const int w = j + TL * v;
                    const int it = w >> (LOGNIN - R1), u = w & ((NIN >> R1) - 1);
                    int base, SB;
                    Gm::at(u, base, SB);
                    double y[1 << R1];
#pragma unroll
                    for (int t = 0; t < (1 << R1); t++) y[t] = (double)dcur[(v << R1) + t];
                    direct_pass<R1>(y, dc);
                    double *p = buf + it * IB + lds_pad(base);
#pragma unroll
                    for (int t = 0; t < (1 << R1); t++) p[unit_off<(1 << Gm::JB)>(t)] = y[t];
                }
            }
            __syncthreads();
            cm_middle<LOGNT, RL, IPW, LOGNIN, 1>(buf, j, tw, m);
            // last pass: the thread's OPT contiguous outputs of every item, in registers
            {
                constexpr int L0 = LOGNT - RL;
                int base, SB;
                UnitGeo<LOGNT, L0, RL>::at(j, base, SB);
                double w[(1 << RL) - 1];
                load_unit_tw<RL, L0>(w, tw, SB);
#pragma unroll
                for (int it = 0; it < IPW; it++) {
                    const double *p = buf + it * IB + lds_pad(OPT * j);
#pragma unroll
                    for (int t = 0; t < OPT; t++) x[it][t] = p[t];
                    unit_stages<RL>(x[it], w, m);
                }
            }
        }
        // multiply-accumulate: slot pair t of the thread is transform output t
        if constexpr (!KPF) load_keys();
#pragma unroll
        for (int t = 0; t < SPT / 2; t++)
#pragma unroll
            for (int it = 0; it < IPW; it++) {
                // slots 2t, 2t + 1: one transform output (embedded) or two
                const double x0 = EMB ? x[it][t] : x[it][2 * t], x1 = EMB ? x[it][t] : x[it][2 * t + 1];
                am[it][2 * t] += mod_mul_lazy(x0, vm[t].x, m);
                am[it][2 * t + 1] += mod_mul_lazy(x1, vm[t].y, m);
                ab[it][2 * t] += mod_mul_lazy(x0, vb[t].x, m);
                ab[it][2 * t + 1] += mod_mul_lazy(x1, vb[t].y, m);
            }
#pragma unroll
        for (int e = 0; e < EPT; e++) dcur[e] = dnxt[e];
    }
    if constexpr (!EMB) {
        // the conversion's products, slot by slot
#pragma unroll
        for (int it = 0; it < IPW; it++) {
            const size_t ot = ((size_t)grp * a.items + item0 + it) * a.tiles + tile;
            double2 *om = reinterpret_cast<double2 *>(a.out_mask + (ot * 2 + l) * S + SPT * j);
            double2 *ob = reinterpret_cast<double2 *>(a.out_body + (ot * 2 + l) * S + SPT * j);
#pragma unroll
            for (int t = 0; t < SPT / 2; t++) {
                om[t] = make_double2(mod_reduce_small(am[it][2 * t], m), mod_reduce_small(am[it][2 * t + 1], m));
                ob[t] = make_double2(mod_reduce_small(ab[it][2 * t], m), mod_reduce_small(ab[it][2 * t + 1], m));
            }
        }
        return;
    }
    // Body: the sum over the tile's 2048 / n segments of degree n = 2 nt
    // (exact fp64 sums of canonical residues, below 2^49), reduced once.
    constexpr int N = 2 << LOGNT;
    constexpr int G = N >= SPT ? S / N : TL;  // terms per output
    constexpr int BW = lds_words(S);         // body words per item
    static_assert(!EMB || IPW * BW <= 2 * IPW * IB, "body sums fit the transform buffers");
    __syncthreads();  // the transform buffers become the body buffers
#pragma unroll
    for (int it = 0; it < IPW; it++) {
        const size_t ot = ((size_t)grp * a.items + item0 + it) * a.tiles + tile;
        double2 *om = reinterpret_cast<double2 *>(a.out_mask + (ot * 2 + l) * S + SPT * j);
#pragma unroll
        for (int t = 0; t < SPT / 2; t++)
            om[t] = make_double2(mod_reduce_small(am[it][2 * t], m), mod_reduce_small(am[it][2 * t + 1], m));
        double *sb = s + it * BW;
        if constexpr (N >= SPT) {
#pragma unroll
            for (int e = 0; e < SPT; e++) sb[lds_pad(SPT * j + e)] = mod_reduce_small(ab[it][e], m);
        } else {
            // the thread's slots are SPT / N whole segments: sum them first
#pragma unroll
            for (int c = 0; c < N; c++) {
                double acc = mod_reduce_small(ab[it][c], m);
#pragma unroll
                for (int sgm = 1; sgm < SPT / N; sgm++) acc += mod_reduce_small(ab[it][sgm * N + c], m);
                sb[lds_pad(j * N + c)] = acc;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int half = G >> 1; half >= 1; half >>= 1) {
#pragma unroll
        for (int it = 0; it < IPW; it++) {
            double *sb = s + it * BW;
            for (int t = j; t < half * N; t += TL) sb[lds_pad(t)] += sb[lds_pad(t + half * N)];
        }
        __syncthreads();
    }
#pragma unroll
    for (int it = 0; it < IPW; it++) {
        const size_t ot = ((size_t)grp * a.items + item0 + it) * a.tiles + tile;
        const double *sb = s + it * BW;
        for (int c = j; c < N; c += TL) a.out_body[(ot * 2 + l) * N + c] = mod_reduce_small(sb[lds_pad(c)], m);
    }
}

bool cascade_mac_fits(const RingDev &r, const Gadget &g, const MacArgs &a) {
    // n1 = 2048 with two limbs: the cascade steps (embedded inputs, two tiles
    // of 2048 slots, degrees up to 2048, body segment sums) or the RLWE ->
    // RGSW conversion (degree 2048, one tile, products slot by slot); digits
    // small enough for the exact direct pass: (B/2) (1 + 7 (q - 1)) + 14 q
    // below 2^45 for the lazy products after it
    if (r.nl != 2 || a.dig_coeffs != (size_t)kSlots) return false;
    if (a.embed) {
        if (a.tiles != 2 || !a.reduce_body || r.logn < 1 || r.logn > kSlotsLog || a.key_slots != (size_t)2 * kSlots)
            return false;
    } else if (a.tiles != 1 || a.reduce_body || r.logn != kSlotsLog || a.key_slots != (size_t)kSlots) {
        return false;
    }
    const double qmax = r.t[0].m.q > r.t[1].m.q ? r.t[0].m.q : r.t[1].m.q;
    const double half_b = (double)(g.base / 2 + 1);
    return half_b * 8.0 * qmax + 14.0 * qmax < 35184372088832.0;  // 2^45
}

template <int IPW>
static void launch_cm(const RingDev &r, const Gadget &g, const MacArgs &a, hipStream_t st) {
    const dim3 grid(a.items / IPW * a.groups * a.tiles, 2), blk(128 * IPW);
    if (!a.embed) {
        hipLaunchKernelGGL((k_cascade_mac<11, IPW, false>), grid, blk, 0, st, r, g, a);
        return;
    }
    switch (r.logn - 1) {
        case 0: hipLaunchKernelGGL((k_cascade_mac<0, IPW, true>), grid, blk, 0, st, r, g, a); break;
        case 1: hipLaunchKernelGGL((k_cascade_mac<1, IPW, true>), grid, blk, 0, st, r, g, a); break;
        case 2: hipLaunchKernelGGL((k_cascade_mac<2, IPW, true>), grid, blk, 0, st, r, g, a); break;
        case 3: hipLaunchKernelGGL((k_cascade_mac<3, IPW, true>), grid, blk, 0, st, r, g, a); break;
        case 4: hipLaunchKernelGGL((k_cascade_mac<4, IPW, true>), grid, blk, 0, st, r, g, a); break;
        case 5: hipLaunchKernelGGL((k_cascade_mac<5, IPW, true>), grid, blk, 0, st, r, g, a); break;
        case 6: hipLaunchKernelGGL((k_cascade_mac<6, IPW, true>), grid, blk, 0, st, r, g, a); break;
        case 7: hipLaunchKernelGGL((k_cascade_mac<7, IPW, true>), grid, blk, 0, st, r, g, a); break;
        case 8: hipLaunchKernelGGL((k_cascade_mac<8, IPW, true>), grid, blk, 0, st, r, g, a); break;
        case 9: hipLaunchKernelGGL((k_cascade_mac<9, IPW, true>), grid, blk, 0, st, r, g, a); break;
        case 10: hipLaunchKernelGGL((k_cascade_mac<10, IPW, true>), grid, blk, 0, st, r, g, a); break;
        default: break;
    }
}

void launch_cascade_mac(const RingDev &r, const Gadget &g, const MacArgs &a, hipStream_t st) {
    // a.ipw items per workgroup where that many consecutive items share a key
    if (a.ipw >= 4 && a.items_per_key % 4 == 0)
        launch_cm<4>(r, g, a, st);
    else if (a.ipw >= 2 && a.items_per_key % 2 == 0)
        launch_cm<2>(r, g, a, st);
    else
        launch_cm<1>(r, g, a, st);
}

// Key rows in the layout of k_cascade_mac: every (k, part, limb) row of
// tiles * 2048 slots, tile by tile, slot sg of a tile moved to
// cascade_key_pos(sg).
__global__ void k_cascade_key_permute(const double *src, double *dst, size_t rows, int tiles) {
    const size_t total = rows * tiles * kSlots;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t tile_base = t & ~(size_t)(kSlots - 1);
        dst[tile_base + cascade_key_pos((int)(t & (kSlots - 1)))] = src[t];
    }
}

void launch_cascade_key_permute(const double *src, double *dst, size_t rows, int tiles, hipStream_t st) {
    if (!rows) return;
    const size_t total = rows * tiles * kSlots;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_cascade_key_permute, dim3(blocks), dim3(256), 0, st, src, dst, rows, tiles);
}

}  // namespace via
