// gate_fx.hip — the tree gates (RGSWCiphertext::external_product,
// rgsw.rs:108-120, then dmux mux.rs:129-137, cmux mux.rs:75-84, or the
// SlotExtract step of crot, rotate.rs:128-149) in a fixed shape for the wide
// tree levels: degree 2048, gadget depth 2, one RNS limb of one gate per
// workgroup of 256 threads.
//
// Thread j computes the gadget digits of the input coefficients
// c = j + 256 t (t < 8) of one ciphertext half at a time, which are exactly
// the elements of its units in the first forward pass, so digits never pass
// through LDS; it ends the forward transforms holding eval slots
// 8j .. 8j+7 of both digit polys, multiplies them with the RGSW rows in
// registers, and after both halves starts the inverse transforms of the two
// product parts on those same slots (the first inverse stages pair adjacent
// slots); the last inverse pass again holds coefficients j + 256 t, where
// the gate combine reads its inputs. Per half: three barriers; inverse:
// three. With one-limb digits at most B/2 in magnitude (the CMux and CRot
// gadget at q2), the first forward pass runs in direct form (ntt_fixed.h).
#include "digits.h"
#include "ntt_fixed.h"

namespace via {

namespace {

constexpr int kGxBuf = lds_words(kSlots);

// The signed gadget scale round-half-away(c_centred * B^L / q) of a
// one-limb coefficient as an exact double (gadget_scale's fp64 path,
// g.f64 = 1, which the launcher requires). Two-limb coefficients take
// gadget_scale inline: an out-of-line exact fallback inside the unrolled
// digit loops measured slower than the whole inlined path.
__device__ __forceinline__ double gx_scaled1(double a0, const RingDev &r, const Gadget &g) {
    const uint64_t q = r.q_single, x = (uint64_t)a0;
    const bool neg = x > (q >> 1);
    const double v = (double)scale_div_f64(neg ? q - x : x, g.sf.bl, r.t[0].m.q, r.t[0].m.qinv, q >> 1);
    return neg ? -v : v;
}

// Forward middle pass (L0, R) over the two digit polys in LDS (poly p at
// s + p * kGxBuf): units w = j + 256 v of the polys' 2048 >> R units each.
template <int L0, int R>
__device__ __forceinline__ void gx_fwd_mid(double *s, int j, const double *tw, const FMod &m) {
    constexpr int E = 1 << R, UPP = kSlots >> R;
    using G = UnitGeo<kSlotsLog, L0, R>;
#pragma unroll
    for (int v = 0; v < (16 >> R); v++) {
        const int w = j + 256 * v;
        const int poly = w / UPP, u = w % UPP;
        int base, SB;
        G::at(u, base, SB);
        double wt[E - 1];
        load_unit_tw<R, L0>(wt, tw, SB);
        double *p = s + poly * kGxBuf + lds_pad(base);
        double x[E];
#pragma unroll
        for (int t = 0; t < E; t++) x[t] = p[unit_off<(1 << G::JB)>(t)];
        unit_stages<R>(x, wt, m);
#pragma unroll
        for (int t = 0; t < E; t++) p[unit_off<(1 << G::JB)>(t)] = x[t];
    }
}

// Inverse middle pass (LH0, R) over the two product parts in LDS.
template <int LH0, int R>
__device__ __forceinline__ void gx_inv_mid(double *s, int j, const double *tw, const FMod &m) {
    constexpr int E = 1 << R, UPP = kSlots >> R;
    using G = InvGeo<kSlotsLog, LH0, R>;
#pragma unroll
    for (int v = 0; v < (16 >> R); v++) {
        const int w = j + 256 * v;
        const int poly = w / UPP, u = w % UPP;
        int base, SB;
        G::at(u, base, SB);
        double wt[E - 1];
        load_inv_unit_tw<kSlotsLog, LH0, R>(wt, tw, SB);
        double *p = s + poly * kGxBuf + lds_pad(base);
        double x[E];
#pragma unroll
        for (int t = 0; t < E; t++) x[t] = p[unit_off<(1 << LH0)>(t)];
        inv_unit_stages<R>(x, wt, m);
#pragma unroll
        for (int t = 0; t < E; t++) p[unit_off<(1 << LH0)>(t)] = x[t];
    }
}

}  // namespace

template <int NL, bool DIRECT, int MODE>
__global__ __launch_bounds__(256, 3) void k_gate_fx(RingDev r, Gadget g, GateArgs ga) {
    constexpr int S = kSlots, L = 2;
    __shared__ double s[2 * kGxBuf];  // two polys: the digit polys of a half, then the product parts
    const int j = threadIdx.x;
    const int gate = blockIdx.x, l = NL == 2 ? (int)blockIdx.y : 0;
    const int q = gate / ga.items_per_query;
    const int jg = gate - q * ga.items_per_query;
    constexpr int mode = MODE;
    const int shift = ga.shift;
    const double *rgsw = ga.rgsw + (size_t)q * ga.rgsw_qstride;
    const double *in = ga.in + (size_t)q * ga.in_qstride;
    double *out = ga.out + (size_t)q * ga.out_qstride;
    const size_t pw = (size_t)NL * S;
    const FMod m = r.t[l].m, m0 = r.t[0].m, m1 = r.t[1].m;
    const double *twf = r.t[l].fwd, *twi = r.t[l].inv;
    double am[8], ab[8];
#pragma unroll
    for (int e = 0; e < 8; e++) am[e] = ab[e] = 0.0;
    for (int half = 0; half < 2; half++) {  // 0: ct.mask with the -s*m rows, 1: ct.body with the m rows
        // digits of coefficients j + 256 t (both limbs enter the scale), then
        // stages 0..2 of both digit polys on them
        {
            double y[L][8];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int c = j + 256 * t;
                double b0, c0, b1 = 0.0, c1 = 0.0;
                gate_input<NL>(mode, in, jg, half, 0, c, S, shift, m0, b0, c0);
                if (NL == 2) gate_input<NL>(mode, in, jg, half, 1, c, S, shift, m1, b1, c1);
                // both digits from the low part (B^2 < 2^50), LSB first (gadget.rs:168-190)
                if constexpr (NL == 2) {
                    double lo, hi;
                    digit_split_f64(gadget_scale<NL>(c0, c1, r, g), ga.dp, lo, hi);
#pragma unroll
                    for (int k = 0; k < L; k++) y[k][t] = (double)next_digit(lo, hi, k, ga.dp);
                } else {
                    double v = gx_scaled1(c0, r, g);
#pragma unroll
                    for (int k = 0; k < L; k++) y[k][t] = (double)digit_f64(v, ga.dp);
                }
            }
#pragma unroll
            for (int k = 0; k < L; k++) {
                if constexpr (DIRECT) {
                    direct_pass<3>(y[k], r.t[l].dfw);
                } else {
                    double w0[7];
                    load_unit_tw<3, 0>(w0, twf, 0);
                    unit_stages<3>(y[k], w0, m);
                }
                double *p = s + k * kGxBuf + lds_pad(j);
#pragma unroll
                for (int t = 0; t < 8; t++) p[unit_off<256>(t)] = y[k][t];
            }
        }
        __syncthreads();
        gx_fwd_mid<3, 3>(s, j, twf, m);
        __syncthreads();
        gx_fwd_mid<6, 2>(s, j, twf, m);
        __syncthreads();
        // last pass: slots 8j .. 8j+7 of both digit polys, then the RGSW
        // products (rows [half][i][part][limb][n], digit k with row L-1-k)
        {
            int base, SB;
            UnitGeo<kSlotsLog, 8, 3>::at(j, base, SB);
            double w[7];
            load_unit_tw<3, 8>(w, twf, SB);
#pragma unroll
            for (int k = 0; k < L; k++) {
                double x[8];
                const double *p = s + k * kGxBuf + lds_pad(8 * j);
#pragma unroll
                for (int t = 0; t < 8; t++) x[t] = p[t];
                unit_stages<3>(x, w, m);
                const double2 *km = reinterpret_cast<const double2 *>(
                    rgsw + ((size_t)half * L + (L - 1 - k)) * 2 * pw + (size_t)l * S + 8 * j);
                const double2 *kb = reinterpret_cast<const double2 *>(reinterpret_cast<const double *>(km) + pw);
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const double2 vm = km[t], vb = kb[t];
                    am[2 * t] += mod_mul_lazy(x[2 * t], vm.x, m);
                    am[2 * t + 1] += mod_mul_lazy(x[2 * t + 1], vm.y, m);
                    ab[2 * t] += mod_mul_lazy(x[2 * t], vb.x, m);
                    ab[2 * t + 1] += mod_mul_lazy(x[2 * t + 1], vb.y, m);
                }
            }
        }
        __syncthreads();  // the LDS is rewritten next (the other half, or the inverse)
    }
    // inverse transforms of the two product parts: stages 0..2 on slots
    // 8j .. 8j+7 in registers (inputs brought into (-q, 2q))
    {
        double w[7];
        load_inv_unit_tw<kSlotsLog, 0, 3>(w, twi, j);
        if constexpr (MODE == GATE_CMUX_EV) {
            // c0 of the pair (column 2jg + 1, eval form), slots 8j .. 8j+7
            const double2 *e0m = reinterpret_cast<const double2 *>(in + (size_t)(2 * jg + 1) * 2 * pw + 8 * j);
            const double2 *e0b = reinterpret_cast<const double2 *>(in + (size_t)(2 * jg + 1) * 2 * pw + pw + 8 * j);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const double2 vm = e0m[t], vb = e0b[t];
                am[2 * t] += vm.x;
                am[2 * t + 1] += vm.y;
                ab[2 * t] += vb.x;
                ab[2 * t + 1] += vb.y;
            }
        }
#pragma unroll
        for (int e = 0; e < 8; e++) {
            am[e] = lazy_reduce(am[e], m);
            ab[e] = lazy_reduce(ab[e], m);
        }
        inv_unit_stages<3>(am, w, m);
        inv_unit_stages<3>(ab, w, m);
        double *p0 = s + lds_pad(8 * j), *p1 = s + kGxBuf + lds_pad(8 * j);
#pragma unroll
        for (int t = 0; t < 8; t++) {
            p0[t] = am[t];
            p1[t] = ab[t];
        }
    }
    __syncthreads();
    gx_inv_mid<3, 3>(s, j, twi, m);
    __syncthreads();
    gx_inv_mid<6, 2>(s, j, twi, m);
    __syncthreads();
    // last inverse pass: coefficients j + 256 t of both parts, the N^-1
    // scaling, the gate combine
    {
        double w[7];
        load_inv_unit_tw<kSlotsLog, 8, 3>(w, twi, 0);
        const double ni = r.t[l].n_inv;
#pragma unroll
        for (int part = 0; part < 2; part++) {
            double x[8];
            const double *p = s + part * kGxBuf + lds_pad(j);
#pragma unroll
            for (int t = 0; t < 8; t++) x[t] = p[unit_off<256>(t)];
            inv_unit_stages<3>(x, w, m);
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int c = j + 256 * t;
                const double pv = mod_mul(x[t], ni, m);
                const size_t o = (size_t)part * pw + (size_t)l * S + c;
                if constexpr (MODE == GATE_CMUX_EV) {
                    out[(size_t)jg * 2 * pw + o] = pv;  // c0 entered before the transform
                    continue;
                }
                double bv, cv;
                gate_input<NL>(mode, in, jg, part, l, c, S, shift, m, bv, cv);
                if (mode == GATE_DMUX) {
                    out[(size_t)(2 * jg) * 2 * pw + o] = mod_sub(cv, pv, m);
                    out[(size_t)(2 * jg + 1) * 2 * pw + o] = pv;
                } else {
                    out[(size_t)jg * 2 * pw + o] = mod_add(bv, pv, m);
                }
            }
        }
    }
}

// Narrow levels (fewer than kGateFxMin gates, where one gate's latency is
// the level's): one gate (one limb of it with two limbs) per 1024-thread
// workgroup with compile-time passes. Threads 256 h + j (h = 2 half + k) compute digit k of
// coefficients j + 256 t of the ciphertext half and run the first forward
// pass on them in registers; stages 3..5 and 6..7 run over LDS; the
// last pass (8..10) leaves slots 8j .. 8j+7 of poly h in LDS. Threads
// 512 p + u then multiply slots 4u .. 4u+3 of the four digit polys with the
// RGSW rows of part p (loaded during the forward passes), run inverse stages 0..1 on the
// sums in registers, 2..4 and 5..7 over LDS (256 threads per part), and
// 8..10 with N^-1 and the gate combine on coefficients u + 256 t. Eight
// barriers per gate. Two limbs (the DMux at q1): one limb per workgroup
// (workgroup 2 gate + l), each coefficient's two-limb scale computed once
// per workgroup (the threads of digit k take half of the coefficients and
// hand the other digit over through LDS) with the scale inlined: an
// out-of-line exact fallback in that loop cost more than the whole gain.
template <int NL, bool DIRECT, int MODE>
__global__ __launch_bounds__(1024) void k_gate_nx(RingDev r, Gadget g, GateArgs ga) {
    constexpr int S = kSlots, L = 2;
    __shared__ double s[4 * kGxBuf];  // the four digit polys, then the two product parts
    const int tid = threadIdx.x;
    const int gate = NL == 2 ? (int)(blockIdx.x >> 1) : (int)blockIdx.x;
    const int l = NL == 2 ? (int)(blockIdx.x & 1) : 0;
    const int q = gate / ga.items_per_query;
    const int jg = gate - q * ga.items_per_query;
    const int shift = ga.shift;
    const double *rgsw = ga.rgsw + (size_t)q * ga.rgsw_qstride;
    const double *in = ga.in + (size_t)q * ga.in_qstride;
    double *out = ga.out + (size_t)q * ga.out_qstride;
    const size_t pw = (size_t)NL * S;
    const FMod m = r.t[l].m;
    const double *twf = r.t[l].fwd, *twi = r.t[l].inv;
    // product role: part p, slots 4u .. 4u+3; combine role (u < 256):
    // coefficients u + 256 t of part p
    const int p = tid >> 9, u = tid & 511;
    {
        // digit k of coefficients j + 256 t of the half, first forward pass
        const int h = tid >> 8, j = tid & 255, half = h >> 1, k = h & 1;
        double y[8];
        if constexpr (NL == 1) {
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int c = j + 256 * t;
                double b0, c0;
                gate_input<NL>(MODE, in, jg, half, 0, c, S, shift, m, b0, c0);
                double v = gx_scaled1(c0, r, g);
                double d = (double)digit_f64(v, ga.dp);  // LSB first (gadget.rs:168-190)
                if (k == 1) d = (double)digit_f64(v, ga.dp);
                y[t] = d;
            }
        } else {
            // the two-limb scale of a coefficient once per workgroup: the
            // threads of digit k scale t = 4k .. 4k+3, keep digit k and hand
            // the other digit to the partner thread through its buffer
            const FMod m0 = r.t[0].m, m1 = r.t[1].m;
            double *mine = s + h * kGxBuf + lds_pad(j), *other = s + (h ^ 1) * kGxBuf + lds_pad(j);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int t = 4 * k + e, c = j + 256 * t;
                double b0, c0, b1, c1;
                gate_input<NL>(MODE, in, jg, half, 0, c, S, shift, m0, b0, c0);
                gate_input<NL>(MODE, in, jg, half, 1, c, S, shift, m1, b1, c1);
                double lo, hi;
                digit_split_f64(gadget_scale<NL>(c0, c1, r, g), ga.dp, lo, hi);
                const double d0 = (double)next_digit(lo, hi, 0, ga.dp);
                const double d1 = (double)next_digit(lo, hi, 1, ga.dp);
                y[t] = k ? d1 : d0;
                other[unit_off<256>(t)] = k ? d0 : d1;
            }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int t = 4 * (1 - k) + e;
                y[t] = mine[unit_off<256>(t)];
            }
        }
        if constexpr (DIRECT) {
            direct_pass<3>(y, r.t[l].dfw);
        } else {
            double w0[7];
            load_unit_tw<3, 0>(w0, twf, 0);
            unit_stages<3>(y, w0, m);
        }
        double *pp = s + h * kGxBuf + lds_pad(j);
#pragma unroll
        for (int t = 0; t < 8; t++) pp[unit_off<256>(t)] = y[t];
    }
    // the RGSW row words of the products, in flight during the forward passes
    double kr[4][4];  // RGSW row words [poly h][slot]
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const int half = h >> 1, k = h & 1;
        const double2 *row = reinterpret_cast<const double2 *>(
            rgsw + ((size_t)half * L + (L - 1 - k)) * 2 * pw + (size_t)p * pw + (size_t)l * S + 4 * u);
        const double2 a = row[0], b = row[1];
        kr[h][0] = a.x;
        kr[h][1] = a.y;
        kr[h][2] = b.x;
        kr[h][3] = b.y;
    }
    __syncthreads();
    {
        const int h = tid >> 8, j = tid & 255;
        {
            using G = UnitGeo<kSlotsLog, 3, 3>;
            int base, SB;
            G::at(j, base, SB);
            double w[7];
            load_unit_tw<3, 3>(w, twf, SB);
            double *pp = s + h * kGxBuf + lds_pad(base);
            double x[8];
#pragma unroll
            for (int t = 0; t < 8; t++) x[t] = pp[unit_off<(1 << G::JB)>(t)];
            unit_stages<3>(x, w, m);
#pragma unroll
            for (int t = 0; t < 8; t++) pp[unit_off<(1 << G::JB)>(t)] = x[t];
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < 2; v++) {
            using G = UnitGeo<kSlotsLog, 6, 2>;
            int base, SB;
            G::at(j + 256 * v, base, SB);
            double w[3];
            load_unit_tw<2, 6>(w, twf, SB);
            double *pp = s + h * kGxBuf + lds_pad(base);
            double x[4];
#pragma unroll
            for (int t = 0; t < 4; t++) x[t] = pp[unit_off<(1 << G::JB)>(t)];
            unit_stages<2>(x, w, m);
#pragma unroll
            for (int t = 0; t < 4; t++) pp[unit_off<(1 << G::JB)>(t)] = x[t];
        }
        __syncthreads();
        {
            int base, SB;
            UnitGeo<kSlotsLog, 8, 3>::at(j, base, SB);
            double w[7];
            load_unit_tw<3, 8>(w, twf, SB);
            double *pp = s + h * kGxBuf + lds_pad(8 * j);
            double x[8];
#pragma unroll
            for (int t = 0; t < 8; t++) x[t] = pp[t];
            unit_stages<3>(x, w, m);
#pragma unroll
            for (int t = 0; t < 8; t++) pp[t] = x[t];
        }
    }
    // the combine's inputs, in flight during the inverse passes
    double bvp[8], cvp[8];
    if (u < 256) {
#pragma unroll
        for (int t = 0; t < 8; t++) gate_input<NL>(MODE, in, jg, p, l, u + 256 * t, S, shift, m, bvp[t], cvp[t]);
    }
    __syncthreads();
    // products of part p at slots 4u .. 4u+3 (digit k with row L-1-k), then
    // inverse stages 0..1 on them
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const double *src = s + h * kGxBuf + lds_pad(4 * u);
#pragma unroll
        for (int t = 0; t < 4; t++) acc[t] += mod_mul_lazy(src[t], kr[h][t], m);
    }
    {
        double w[3];
        load_inv_unit_tw<kSlotsLog, 0, 2>(w, twi, u);
#pragma unroll
        for (int t = 0; t < 4; t++) acc[t] = lazy_reduce(acc[t], m);
        inv_unit_stages<2>(acc, w, m);
    }
    __syncthreads();  // every product read the digit polys
    double *sp = s + p * kGxBuf;
#pragma unroll
    for (int t = 0; t < 4; t++) sp[lds_pad(4 * u) + t] = acc[t];
    __syncthreads();
    if (u < 256) {
        using G = InvGeo<kSlotsLog, 2, 3>;
        int base, SB;
        G::at(u, base, SB);
        double w[7];
        load_inv_unit_tw<kSlotsLog, 2, 3>(w, twi, SB);
        double *pp = sp + lds_pad(base);
        double x[8];
#pragma unroll
        for (int t = 0; t < 8; t++) x[t] = pp[unit_off<4>(t)];
        inv_unit_stages<3>(x, w, m);
#pragma unroll
        for (int t = 0; t < 8; t++) pp[unit_off<4>(t)] = x[t];
    }
    __syncthreads();
    if (u < 256) {
        using G = InvGeo<kSlotsLog, 5, 3>;
        int base, SB;
        G::at(u, base, SB);
        double w[7];
        load_inv_unit_tw<kSlotsLog, 5, 3>(w, twi, SB);
        double *pp = sp + lds_pad(base);
        double x[8];
#pragma unroll
        for (int t = 0; t < 8; t++) x[t] = pp[unit_off<32>(t)];
        inv_unit_stages<3>(x, w, m);
#pragma unroll
        for (int t = 0; t < 8; t++) pp[unit_off<32>(t)] = x[t];
    }
    __syncthreads();
    if (u < 256) {
        // inverse stages 8..10 on coefficients u + 256 t, N^-1, the combine
        double w[7];
        load_inv_unit_tw<kSlotsLog, 8, 3>(w, twi, 0);
        const double ni = r.t[l].n_inv;
        const double *pp = sp + lds_pad(u);
        double x[8];
#pragma unroll
        for (int t = 0; t < 8; t++) x[t] = pp[unit_off<256>(t)];
        inv_unit_stages<3>(x, w, m);
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int c = u + 256 * t;
            const double pv = mod_mul(x[t], ni, m);
            const size_t o = (size_t)p * pw + (size_t)l * S + c;
            if (MODE == GATE_DMUX) {
                out[(size_t)(2 * jg) * 2 * pw + o] = mod_sub(cvp[t], pv, m);
                out[(size_t)(2 * jg + 1) * 2 * pw + o] = pv;
            } else {
                out[(size_t)jg * 2 * pw + o] = mod_add(bvp[t], pv, m);
            }
        }
    }
}

// Narrow levels across twice the compute units: each gate (each limb of it
// with two limbs) is split into two launches, since one gate per compute
// unit leaves the level's latency to that unit's VALU throughput. First
// k_gate_nxa, one 512-thread workgroup per ciphertext half: the digits of
// the half and their forward transforms as in k_gate_nx (threads 256 k + j
// hold digit k), then thread u multiplies slots 4u .. 4u+3 of both digit
// polys with the RGSW rows of both product parts and writes the two lazy
// sums to the partial buffer P[gate][limb][half][part][slot]. Then
// k_gate_nxb, one 512-thread workgroup per product part: the sum of the two
// halves' partials is the gate's product sum (exact integers in fp64, so
// the same value as k_gate_nx's sum over all four digit polys), then the
// inverse transform and the combine as in k_gate_nx.
template <int NL, bool DIRECT, int MODE>
__global__ __launch_bounds__(512) void k_gate_nxa(RingDev r, Gadget g, GateArgs ga) {
    constexpr int S = kSlots, L = 2;
    __shared__ double s[2 * kGxBuf];  // the two digit polys of the half
    const int tid = threadIdx.x;
    const int blk = (int)blockIdx.x, half = blk & 1, gl = blk >> 1;
    const int gate = NL == 2 ? gl >> 1 : gl;
    const int l = NL == 2 ? (gl & 1) : 0;
    const int q = gate / ga.items_per_query;
    const int jg = gate - q * ga.items_per_query;
    const int shift = ga.shift;
    const double *rgsw = ga.rgsw + (size_t)q * ga.rgsw_qstride;
    const double *in = ga.in + (size_t)q * ga.in_qstride;
    const size_t pw = (size_t)NL * S;
    const FMod m = r.t[l].m;
    const double *twf = r.t[l].fwd;
    const int k = tid >> 8, j = tid & 255;
    {
        double y[8];
        if constexpr (NL == 1) {
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int c = j + 256 * t;
                double b0, c0;
                gate_input<NL>(MODE, in, jg, half, 0, c, S, shift, m, b0, c0);
                double v = gx_scaled1(c0, r, g);
                double d = (double)digit_f64(v, ga.dp);  // LSB first (gadget.rs:168-190)
                if (k == 1) d = (double)digit_f64(v, ga.dp);
                y[t] = d;
            }
        } else {
            // the threads of digit k scale t = 4k .. 4k+3, keep digit k and
            // hand the other digit to the partner thread through its buffer
            const FMod m0 = r.t[0].m, m1 = r.t[1].m;
            double *mine = s + k * kGxBuf + lds_pad(j), *other = s + (k ^ 1) * kGxBuf + lds_pad(j);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int t = 4 * k + e, c = j + 256 * t;
                double b0, c0, b1, c1;
                gate_input<NL>(MODE, in, jg, half, 0, c, S, shift, m0, b0, c0);
                gate_input<NL>(MODE, in, jg, half, 1, c, S, shift, m1, b1, c1);
                double lo, hi;
                digit_split_f64(gadget_scale<NL>(c0, c1, r, g), ga.dp, lo, hi);
                const double d0 = (double)next_digit(lo, hi, 0, ga.dp);
                const double d1 = (double)next_digit(lo, hi, 1, ga.dp);
                y[t] = k ? d1 : d0;
                other[unit_off<256>(t)] = k ? d0 : d1;
            }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int t = 4 * (1 - k) + e;
                y[t] = mine[unit_off<256>(t)];
            }
        }
        if constexpr (DIRECT) {
            direct_pass<3>(y, r.t[l].dfw);
        } else {
            double w0[7];
            load_unit_tw<3, 0>(w0, twf, 0);
            unit_stages<3>(y, w0, m);
        }
        double *pp = s + k * kGxBuf + lds_pad(j);
#pragma unroll
        for (int t = 0; t < 8; t++) pp[unit_off<256>(t)] = y[t];
    }
    // the RGSW row words of both parts, in flight during the forward passes
    const int u = tid;
    double kr[2][2][4];  // [digit k][part][slot]
#pragma unroll
    for (int kk = 0; kk < 2; kk++)
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const double2 *row = reinterpret_cast<const double2 *>(
                rgsw + ((size_t)half * L + (L - 1 - kk)) * 2 * pw + (size_t)p * pw + (size_t)l * S + 4 * u);
            const double2 a = row[0], b = row[1];
            kr[kk][p][0] = a.x;
            kr[kk][p][1] = a.y;
            kr[kk][p][2] = b.x;
            kr[kk][p][3] = b.y;
        }
    __syncthreads();
    {
        using G = UnitGeo<kSlotsLog, 3, 3>;
        int base, SB;
        G::at(j, base, SB);
        double w[7];
        load_unit_tw<3, 3>(w, twf, SB);
        double *pp = s + k * kGxBuf + lds_pad(base);
        double x[8];
#pragma unroll
        for (int t = 0; t < 8; t++) x[t] = pp[unit_off<(1 << G::JB)>(t)];
        unit_stages<3>(x, w, m);
#pragma unroll
        for (int t = 0; t < 8; t++) pp[unit_off<(1 << G::JB)>(t)] = x[t];
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 2; v++) {
        using G = UnitGeo<kSlotsLog, 6, 2>;
        int base, SB;
        G::at(j + 256 * v, base, SB);
        double w[3];
        load_unit_tw<2, 6>(w, twf, SB);
        double *pp = s + k * kGxBuf + lds_pad(base);
        double x[4];
#pragma unroll
        for (int t = 0; t < 4; t++) x[t] = pp[unit_off<(1 << G::JB)>(t)];
        unit_stages<2>(x, w, m);
#pragma unroll
        for (int t = 0; t < 4; t++) pp[unit_off<(1 << G::JB)>(t)] = x[t];
    }
    __syncthreads();
    {
        int base, SB;
        UnitGeo<kSlotsLog, 8, 3>::at(j, base, SB);
        double w[7];
        load_unit_tw<3, 8>(w, twf, SB);
        double *pp = s + k * kGxBuf + lds_pad(8 * j);
        double x[8];
#pragma unroll
        for (int t = 0; t < 8; t++) x[t] = pp[t];
        unit_stages<3>(x, w, m);
#pragma unroll
        for (int t = 0; t < 8; t++) pp[t] = x[t];
    }
    __syncthreads();
    // the half's products at slots 4u .. 4u+3 (digit k with row L-1-k)
    double *P = ga.partials + ((size_t)gl * 2 + half) * 2 * S;
#pragma unroll
    for (int p = 0; p < 2; p++) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 2; kk++) {
            const double *src = s + kk * kGxBuf + lds_pad(4 * u);
#pragma unroll
            for (int t = 0; t < 4; t++) acc[t] += mod_mul_lazy(src[t], kr[kk][p][t], m);
        }
        double2 *o = reinterpret_cast<double2 *>(P + (size_t)p * S + 4 * u);
        o[0] = make_double2(acc[0], acc[1]);
        o[1] = make_double2(acc[2], acc[3]);
    }
}

template <int NL, int MODE>
__global__ __launch_bounds__(512) void k_gate_nxb(RingDev r, GateArgs ga) {
    constexpr int S = kSlots;
    __shared__ double s[kGxBuf];  // the product part
    const int tid = threadIdx.x;
    const int blk = (int)blockIdx.x, p = blk & 1, gl = blk >> 1;
    const int gate = NL == 2 ? gl >> 1 : gl;
    const int l = NL == 2 ? (gl & 1) : 0;
    const int q = gate / ga.items_per_query;
    const int jg = gate - q * ga.items_per_query;
    const int shift = ga.shift;
    const double *in = ga.in + (size_t)q * ga.in_qstride;
    double *out = ga.out + (size_t)q * ga.out_qstride;
    const size_t pw = (size_t)NL * S;
    const FMod m = r.t[l].m;
    const double *twi = r.t[l].inv;
    const int u = tid;
    // the sum over both halves at slots 4u .. 4u+3, inverse stages 0..1
    const double *P = ga.partials + (size_t)gl * 4 * S + (size_t)p * S + 4 * u;
    double acc[4];
    {
        const double2 *h0 = reinterpret_cast<const double2 *>(P);
        const double2 *h1 = reinterpret_cast<const double2 *>(P + 2 * S);
        const double2 a0 = h0[0], a1 = h0[1], b0 = h1[0], b1 = h1[1];
        acc[0] = a0.x + b0.x;
        acc[1] = a0.y + b0.y;
        acc[2] = a1.x + b1.x;
        acc[3] = a1.y + b1.y;
    }
    // the combine's inputs, in flight during the inverse passes
    double bvp[8], cvp[8];
    if (u < 256) {
#pragma unroll
        for (int t = 0; t < 8; t++) gate_input<NL>(MODE, in, jg, p, l, u + 256 * t, S, shift, m, bvp[t], cvp[t]);
    }
    {
        double w[3];
        load_inv_unit_tw<kSlotsLog, 0, 2>(w, twi, u);
#pragma unroll
        for (int t = 0; t < 4; t++) acc[t] = lazy_reduce(acc[t], m);
        inv_unit_stages<2>(acc, w, m);
    }
#pragma unroll
    for (int t = 0; t < 4; t++) s[lds_pad(4 * u) + t] = acc[t];
    __syncthreads();
    if (u < 256) {
        using G = InvGeo<kSlotsLog, 2, 3>;
        int base, SB;
        G::at(u, base, SB);
        double w[7];
        load_inv_unit_tw<kSlotsLog, 2, 3>(w, twi, SB);
        double *pp = s + lds_pad(base);
        double x[8];
#pragma unroll
        for (int t = 0; t < 8; t++) x[t] = pp[unit_off<4>(t)];
        inv_unit_stages<3>(x, w, m);
#pragma unroll
        for (int t = 0; t < 8; t++) pp[unit_off<4>(t)] = x[t];
    }
    __syncthreads();
    if (u < 256) {
        using G = InvGeo<kSlotsLog, 5, 3>;
        int base, SB;
        G::at(u, base, SB);
        double w[7];
        load_inv_unit_tw<kSlotsLog, 5, 3>(w, twi, SB);
        double *pp = s + lds_pad(base);
        double x[8];
#pragma unroll
        for (int t = 0; t < 8; t++) x[t] = pp[unit_off<32>(t)];
        inv_unit_stages<3>(x, w, m);
#pragma unroll
        for (int t = 0; t < 8; t++) pp[unit_off<32>(t)] = x[t];
    }
    __syncthreads();
    if (u < 256) {
        // inverse stages 8..10 on coefficients u + 256 t, N^-1, the combine
        double w[7];
        load_inv_unit_tw<kSlotsLog, 8, 3>(w, twi, 0);
        const double ni = r.t[l].n_inv;
        const double *pp = s + lds_pad(u);
        double x[8];
#pragma unroll
        for (int t = 0; t < 8; t++) x[t] = pp[unit_off<256>(t)];
        inv_unit_stages<3>(x, w, m);
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int c = u + 256 * t;
            const double pv = mod_mul(x[t], ni, m);
            const size_t o = (size_t)p * pw + (size_t)l * S + c;
            if (MODE == GATE_DMUX) {
                out[(size_t)(2 * jg) * 2 * pw + o] = mod_sub(cvp[t], pv, m);
                out[(size_t)(2 * jg + 1) * 2 * pw + o] = pv;
            } else {
                out[(size_t)jg * 2 * pw + o] = mod_add(bvp[t], pv, m);
            }
        }
    }
}

// Whether the fixed-shape gate takes a level: degree 2048, gadget depth 2
// with B^2 below 2^50 (both digits from the low part), the fp64 scale (one
// limb: exact, g.f64 = 1; two: decided or exact fallback, g.f64 = 2);
// `direct` when one limb's digits (at most B/2) allow the direct first pass
// with the lazy stages after it below 2^45.
bool gate_fx_fits(const RingDev &r, const Gadget &g, bool *direct) {
    if (r.n != kSlots || g.L != 2 || (double)g.base * g.base >= 1125899906842624.0) return false;
    if (g.f64 != (r.nl == 2 ? 2 : 1)) return false;
    const double qmax = r.t[0].m.q > r.t[1].m.q ? r.t[0].m.q : r.t[1].m.q;
    *direct = r.nl == 1 && ((double)(g.base / 2 + 1) * 8.0 + 14.0) * qmax < 35184372088832.0;  // 2^45
    return true;
}

template <int NL, bool DIRECT>
static void launch_gx(const dim3 &grid, const RingDev &r, const Gadget &g, const GateArgs &a, hipStream_t st) {
    switch (a.mode) {
        case GATE_DMUX: hipLaunchKernelGGL((k_gate_fx<NL, DIRECT, GATE_DMUX>), grid, dim3(256), 0, st, r, g, a); break;
        case GATE_CMUX: hipLaunchKernelGGL((k_gate_fx<NL, DIRECT, GATE_CMUX>), grid, dim3(256), 0, st, r, g, a); break;
        case GATE_CMUX_EV:
            hipLaunchKernelGGL((k_gate_fx<NL, DIRECT, GATE_CMUX_EV>), grid, dim3(256), 0, st, r, g, a);
            break;
        default: hipLaunchKernelGGL((k_gate_fx<NL, DIRECT, GATE_CROT>), grid, dim3(256), 0, st, r, g, a); break;
    }
}

template <int NL, bool DIRECT>
static void launch_nx(const dim3 &grid, const RingDev &r, const Gadget &g, const GateArgs &a, hipStream_t st) {
    switch (a.mode) {
        case GATE_DMUX: hipLaunchKernelGGL((k_gate_nx<NL, DIRECT, GATE_DMUX>), grid, dim3(1024), 0, st, r, g, a); break;
        case GATE_CMUX: hipLaunchKernelGGL((k_gate_nx<NL, DIRECT, GATE_CMUX>), grid, dim3(1024), 0, st, r, g, a); break;
        default: hipLaunchKernelGGL((k_gate_nx<NL, DIRECT, GATE_CROT>), grid, dim3(1024), 0, st, r, g, a); break;
    }
}

template <int NL, bool DIRECT>
static void launch_nx_split(int blocks, const RingDev &r, const Gadget &g, const GateArgs &a, hipStream_t st) {
    const dim3 grid(2 * NL * blocks);
    switch (a.mode) {
        case GATE_DMUX:
            hipLaunchKernelGGL((k_gate_nxa<NL, DIRECT, GATE_DMUX>), grid, dim3(512), 0, st, r, g, a);
            hipLaunchKernelGGL((k_gate_nxb<NL, GATE_DMUX>), grid, dim3(512), 0, st, r, a);
            break;
        case GATE_CMUX:
            hipLaunchKernelGGL((k_gate_nxa<NL, DIRECT, GATE_CMUX>), grid, dim3(512), 0, st, r, g, a);
            hipLaunchKernelGGL((k_gate_nxb<NL, GATE_CMUX>), grid, dim3(512), 0, st, r, a);
            break;
        default:
            hipLaunchKernelGGL((k_gate_nxa<NL, DIRECT, GATE_CROT>), grid, dim3(512), 0, st, r, g, a);
            hipLaunchKernelGGL((k_gate_nxb<NL, GATE_CROT>), grid, dim3(512), 0, st, r, a);
            break;
    }
}

void launch_gate_nx(const RingDev &r, const Gadget &g, const GateArgs &a, bool direct, hipStream_t st) {
    const int blocks = a.items_per_query * a.queries;
    if (a.partials && blocks * r.nl <= kGateNxSplitMax && a.partials_words >= (size_t)blocks * r.nl * 4 * kSlots) {
        if (r.nl == 2)
            launch_nx_split<2, false>(blocks, r, g, a, st);
        else if (direct)
            launch_nx_split<1, true>(blocks, r, g, a, st);
        else
            launch_nx_split<1, false>(blocks, r, g, a, st);
        return;
    }
    if (r.nl == 2)
        launch_nx<2, false>(dim3(2 * blocks), r, g, a, st);
    else if (direct)
        launch_nx<1, true>(dim3(blocks), r, g, a, st);
    else
        launch_nx<1, false>(dim3(blocks), r, g, a, st);
}

void launch_gate_fx(const RingDev &r, const Gadget &g, const GateArgs &a, bool direct, hipStream_t st) {
    const int blocks = a.items_per_query * a.queries;
    if (r.nl == 2)
        launch_gx<2, false>(dim3(blocks, 2), r, g, a, st);
    else if (direct)
        launch_gx<1, true>(dim3(blocks), r, g, a, st);
    else
        launch_gx<1, false>(dim3(blocks), r, g, a, st);
}

}  // namespace via
