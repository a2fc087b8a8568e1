// digits.h — gadget scaling and signed-balanced digit extraction on the
// device, shared by the answer kernels and the VIA-B repack.
#pragma once

#include "internal.h"

namespace via {

// ---- gadget digits ---------------------------------------------------------------

// Centred lift and gadget scale of one coefficient (gadget.rs:115-151):
// val = round-half-away(c_centred * B^L / Q), returned as sign and magnitude.
template <int NL>
__device__ __forceinline__ SMag gadget_scale(double a0, double a1, const RingDev &r, const Gadget &g) {
    SMag v;
    if (NL == 2) {
        if (g.f64 == 2 && gadget_scale2_f64(a0, a1, r.rns, g.sf, v)) return v;
        via_u128 x = rns_reconstruct(a0, a1, r.rns);
        via_u128 Q = r.rns.big_q;
        int neg = x > (Q >> 1);
        via_u128 mag = neg ? Q - x : x;
        v.mag = round_mul_div(mag, g.bl, r.qdiv);
        v.neg = neg && v.mag != 0;
    } else {
        uint64_t q = r.q_single;
        uint64_t x = (uint64_t)a0;
        int neg = x > (q >> 1);
        uint64_t mag = neg ? q - x : x;
        if (g.wide) {
            // B^L beyond 64-bit products (the base-18 depth-18 gadget at q2)
            v.mag = round_mul_div((via_u128)mag, g.bl, r.qdiv);
        } else if (g.f64 == 1) {
            v.mag = (via_u128)scale_div_f64(mag, g.sf.bl, r.t[0].m.q, r.t[0].m.qinv, q >> 1);
        } else {
            // (mag * B^L + floor(q/2)) / q fits in 64 bits (checked by the host).
            uint64_t bl = (uint64_t)g.bl;
            v.mag = (via_u128)((mag * bl + (q >> 1)) / q);
        }
        v.neg = neg && v.mag != 0;
    }
    return v;
}

__device__ __forceinline__ double digit_residue(int32_t d, const FMod &m) { return d >= 0 ? (double)d : (double)d + m.q; }

// Digit k of a value split by digit_split_f64 into (lo, hi) (LSB first,
// gadget.rs:168-190): the first dp.h digits come from lo, whose final carry
// joins hi, the rest from hi. Call for k = 0, 1, ... in order. Exact fp64
// arithmetic on both halves replaces the 128-bit divisions of extract_digit.
__device__ __forceinline__ int32_t next_digit(double &lo, double &hi, int k, const DigitPlan &dp) {
    int32_t d;
    if (k < dp.h) {
        d = digit_f64(lo, dp);
        if (k == dp.h - 1) hi += lo;
    } else {
        d = digit_f64(hi, dp);
    }
    return d;
}

// Digits of one coefficient into int16 planes dig[k * stride] (LSB first,
// gadget.rs:168-190), from the gadget scale of gadget.rs:115-151.
template <int NL>
__device__ __forceinline__ void write_digits(double a0, double a1, const RingDev &r, const Gadget &g,
                                             const DigitPlan &dp, int16_t *dig, size_t stride) {
    double lo, hi;
    if (!(g.f64 == 3 && gadget_split_w<NL>(a0, a1, r.rns, (double)r.q_single, g.sw, lo, hi)))
        digit_split_f64(gadget_scale<NL>(a0, a1, r, g), dp, lo, hi);
    // next_digit's order with its branches hoisted: the low part's digits,
    // its carry into the high part, the high part's digits
    const int hl = dp.h < g.L ? dp.h : g.L;
    int k = 0;
    for (; k < hl; k++, dig += stride) *dig = (int16_t)digit_f64(lo, dp);
    hi += lo;
    for (; k < g.L; k++, dig += stride) *dig = (int16_t)digit_f64(hi, dp);
}

// ---- gate inputs -----------------------------------------------------------------

// Coefficient sg of the gate input ct for item j (see launch_gate).
template <int NL>
__device__ __forceinline__ void gate_input(int mode, const double *in, int j, int part, int l, int sg, int n,
                                           int shift, const FMod &md, double &base_v, double &ct_v) {
    const size_t pw = (size_t)NL * n;
    if (mode == GATE_DMUX) {
        const double x = in[(size_t)j * 2 * pw + part * pw + (size_t)l * n + sg];
        base_v = x;
        ct_v = x;
    } else if (mode == GATE_CMUX_EV) {
        // column 2j holds c1 - c0 (coefficient form); c0 is added in eval form
        const double x = in[(size_t)(2 * j) * 2 * pw + part * pw + (size_t)l * n + sg];
        base_v = 0.0;
        ct_v = x;
    } else if (mode == GATE_CMUX) {
        const double x0 = in[(size_t)(2 * j) * 2 * pw + part * pw + (size_t)l * n + sg];
        const double x1 = in[(size_t)(2 * j + 1) * 2 * pw + part * pw + (size_t)l * n + sg];
        base_v = x0;
        ct_v = mod_sub(x1, x0, md);
    } else {
        // CRot SlotExtract (rotate.rs:128-149): rot = -(res * X^shift),
        // ct = rot - res, out = res + RGSW (x) ct.
        const double *p = in + part * pw + (size_t)l * n;
        const double x = p[sg];
        double rv;
        if (sg >= shift) rv = p[sg - shift];
        else rv = mod_neg(p[sg - shift + n], md);
        const double rneg = mod_neg(rv, md);
        base_v = x;
        ct_v = mod_sub(rneg, x, md);
    }
}

}  // namespace via
