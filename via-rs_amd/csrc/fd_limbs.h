// fd_limbs.h — residues as unsigned 7-bit limbs for the int8 matrix cores,
// and the exact recombination of limb-product sums (host and device).
//
// FirstDim (via-rs via-server/src/first_dim.rs:94-184) multiplies database
// residues x by selection residues y, both below q2 < 2^35, and sums over the
// database rows. With x = sum_a x_a 128^a and y = sum_b y_b 128^b (five
// limbs each, 0 <= x_a, y_b < 128) every limb product fits int8 x int8 ->
// int32, and
//   sum_i x_i y_i = sum_{s=0..8} 128^s C_s,  C_s = sum_i sum_{a+b=s} x_{i,a} y_{i,b}.
// A sum over R rows has C_s <= 5 * R * 127^2, below 2^31 for R < 26630.
#pragma once

#include "modarith.h"

constexpr int kLimbs = 5;        // 7-bit limbs of a residue below 2^35
constexpr int kLimbShifts = 9;   // shifts a + b of a limb product
constexpr int kLimbMaxRows = 16384;  // rows of one sum: C_s < 2^31

// v (an exact integer in [0, 2^35)) -> limb[l], v = sum_l limb[l] * 128^l.
VIA_HD void limbs7(double v, uint32_t limb[kLimbs]) {
    const double h = floor(v * 0x1p-21);           // < 2^14
    const uint32_t lo = (uint32_t)(v - h * 0x1p21);  // exact, < 2^21
    const uint32_t hi = (uint32_t)h;
    limb[0] = lo & 127u;
    limb[1] = (lo >> 7) & 127u;
    limb[2] = lo >> 14;
    limb[3] = hi & 127u;
    limb[4] = hi >> 7;
}

// sum_s C[s] * 128^s mod q in [0, q), for 0 <= C[s] < 2^31 and q < 2^35.
// Horner in exact fp64: every intermediate stays an integer below 2^53, with
// a lazy reduction (to [-q, 2q), below 2^36) before a step could pass it:
// 2^31 -> 2^38 -> 2^45 -> 2^52, reduce; 2^43 -> 2^50, reduce; 2^43 -> 2^50,
// reduce; 2^43, canonical.
VIA_HD double limb_fold(const int32_t C[kLimbShifts], const FMod &m) {
    double v = (double)C[8];
    v = fma(v, 128.0, (double)C[7]);
    v = fma(v, 128.0, (double)C[6]);
    v = fma(v, 128.0, (double)C[5]);
    v = fma(-floor(v * m.qinv), m.q, v);
    v = fma(v, 128.0, (double)C[4]);
    v = fma(v, 128.0, (double)C[3]);
    v = fma(-floor(v * m.qinv), m.q, v);
    v = fma(v, 128.0, (double)C[2]);
    v = fma(v, 128.0, (double)C[1]);
    v = fma(-floor(v * m.qinv), m.q, v);
    v = fma(v, 128.0, (double)C[0]);
    return mod_reduce_small(v, m);
}
