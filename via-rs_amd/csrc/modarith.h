// modarith.h — exact modular arithmetic for moduli below 2^50, host and device.
//
// Residues live in fp64 registers as exact integers in [0, q). A product is
// reduced with one fp64 Barrett estimate and an FMA-exact remainder:
//   h = a*b (rounded), l = fma(a, b, -h) (the exact rounding error),
//   t = floor(h * (1/q)), r = fma(-t, q, h) + l,
// and r lands in [-q, 2q) because t is within one of floor(a*b/q) for
// a, b < q < 2^50; two conditional corrections make it canonical. Every step
// is exact: fma(-t, q, h) and r + l are integers below 2^53 in magnitude.
// Compile with -ffp-contract=off so that h stays a rounded product.
//
// The wide helpers implement the 128/256-bit integer steps of the gadget
// scale (via-rs encryption/gadget.rs:115-151, algebra/wide.rs:145-153) and
// the mod switch (switching/kernels/mod_switch.rs:59-61) without 128-bit
// division: the quotient is estimated in fp64 and corrected exactly.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define VIA_HD __host__ __device__ __forceinline__
#else
#define VIA_HD static inline
#endif

typedef unsigned __int128 via_u128;
typedef __int128 via_i128;

struct FMod {
    double q;     // modulus
    double qinv;  // 1/q rounded
};

VIA_HD FMod fmod_make(uint64_t q) {
    FMod m;
    m.q = (double)q;
    m.qinv = 1.0 / (double)q;
    return m;
}

VIA_HD double mod_add(double a, double b, const FMod &m) {
    double s = a + b;
    return s >= m.q ? s - m.q : s;
}

VIA_HD double mod_sub(double a, double b, const FMod &m) {
    double d = a - b;
    return d < 0.0 ? d + m.q : d;
}

VIA_HD double mod_neg(double a, const FMod &m) { return a == 0.0 ? 0.0 : m.q - a; }

VIA_HD double mod_mul(double a, double b, const FMod &m) {
    double h = a * b;
    double l = fma(a, b, -h);
    double t = floor(h * m.qinv);
    double r = fma(-t, m.q, h) + l;
    r = r < 0.0 ? r + m.q : r;
    r = r >= m.q ? r - m.q : r;
    return r;
}

// Product a*b less a multiple of q, left in (-q, 2q) without the final
// corrections: an exact integer. Sums of up to 2^16 such terms stay exact in
// fp64 (|sum| < 2^53 for q < 2^36, or 2^14 terms for q < 2^38) and are
// brought into [0, q) once with mod_reduce_small.
VIA_HD double mod_mul_lazy(double a, double b, const FMod &m) {
    double h = a * b;
    double l = fma(a, b, -h);
    double t = floor(h * m.qinv);
    return fma(-t, m.q, h) + l;
}

// Reduce an exact integer value v (|v| < 2^52) into [0, q).
VIA_HD double mod_reduce_small(double v, const FMod &m) {
    double t = floor(v * m.qinv);
    double r = fma(-t, m.q, v);
    r = r < 0.0 ? r + m.q : r;
    r = r >= m.q ? r - m.q : r;
    return r;
}

// Exact conversions of canonical residues (integers in [0, 2^52)) through the
// 2^52 bias: the double with exponent 2^52 and mantissa x is 2^52 + x.
VIA_HD double small_u64_to_f64(uint64_t x) {
    const uint64_t b = x | 0x4330000000000000ULL;
    double d;
    __builtin_memcpy(&d, &b, sizeof d);
    return d - 4503599627370496.0;
}

VIA_HD uint64_t small_f64_to_u64(double v) {
    const double d = v + 4503599627370496.0;
    uint64_t b;
    __builtin_memcpy(&b, &d, sizeof b);
    return b & 0x000FFFFFFFFFFFFFULL;
}

// ---- 128-bit helpers ---------------------------------------------------------

VIA_HD double u128_to_f64(via_u128 x) {
    return (double)(uint64_t)(x >> 64) * 18446744073709551616.0 + (double)(uint64_t)x;
}

// Exact conversion of an integral double in [0, 2^106) to u128.
VIA_HD via_u128 f64_to_u128(double v) {
    double hi = floor(v * (1.0 / 18446744073709551616.0));
    double lo = fma(-hi, 18446744073709551616.0, v);  // exact: v - hi*2^64 in [0, 2^64)
    if (lo < 0.0) {
        hi -= 1.0;
        lo += 18446744073709551616.0;
    }
    return ((via_u128)(uint64_t)hi << 64) + (via_u128)(uint64_t)lo;
}

// 3-word little-endian 192-bit value.
struct U192 {
    uint64_t w0, w1, w2;
};

VIA_HD U192 mul_u128_u128_192(via_u128 a, via_u128 b) {
    // a, b < 2^96 on every call site; the product fits in 192 bits.
    uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64);
    uint64_t b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
    via_u128 p00 = (via_u128)a0 * b0;
    via_u128 p01 = (via_u128)a0 * b1;
    via_u128 p10 = (via_u128)a1 * b0;
    via_u128 p11 = (via_u128)a1 * b1;
    via_u128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;
    U192 r;
    r.w0 = (uint64_t)p00;
    r.w1 = (uint64_t)mid;
    r.w2 = (uint64_t)((mid >> 64) + (p01 >> 64) + (p10 >> 64) + p11);
    return r;
}

VIA_HD U192 add_u192_u128(U192 x, via_u128 y) {
    via_u128 lo = ((via_u128)x.w1 << 64) | x.w0;
    via_u128 s = lo + y;
    U192 r;
    r.w0 = (uint64_t)s;
    r.w1 = (uint64_t)(s >> 64);
    r.w2 = x.w2 + (s < lo ? 1 : 0);
    return r;
}

// Divisor description for floor(x / D) with D < 2^96.
struct WideDiv {
    via_u128 d;
    double df;    // (double)d
    double dinv;  // 1/(double)d
};

VIA_HD WideDiv widediv_make(via_u128 d) {
    WideDiv w;
    w.d = d;
    w.df = u128_to_f64(d);
    w.dinv = 1.0 / w.df;
    return w;
}

// floor(x / D) for x < 2^160 whose quotient is below 2^100.
// Two fp64 estimate/remainder rounds, then an exact +-1 correction. The first
// estimate is within q*2^-51 of the quotient, so the remainder x - q*D stays
// below 2^126 in magnitude and is exact when taken modulo 2^128 as a signed
// value; the second estimate is then within one.
VIA_HD via_u128 div_u192(U192 x, const WideDiv &w) {
    double xf = (double)x.w2 * 340282366920938463463374607431768211456.0 +
                (double)x.w1 * 18446744073709551616.0 + (double)x.w0;
    double qf = floor(xf * w.dinv);
    if (qf < 0.0) qf = 0.0;
    via_u128 q = f64_to_u128(qf);
    via_u128 xlo = ((via_u128)x.w1 << 64) | x.w0;
    // r = x - q*d, |r| < 2^115 after the first round
    via_i128 r = (via_i128)(xlo - q * w.d);
    double rf = (double)(int64_t)(r >> 64) * 18446744073709551616.0 + (double)(uint64_t)r;
    double cf = floor(rf * w.dinv);
    via_i128 c = (via_i128)(int64_t)cf;  // |c| < 2^63 while the quotient is below 2^110
    q = (via_u128)((via_i128)q + c);
    r -= c * (via_i128)w.d;
    for (int i = 0; i < 2 && r < 0; i++) {
        q -= 1;
        r += (via_i128)w.d;
    }
    for (int i = 0; i < 2 && r >= (via_i128)w.d; i++) {
        q += 1;
        r -= (via_i128)w.d;
    }
    return q;
}

// Round-half-up scaling floor((a*s + floor(D/2)) / D) (wide.rs:145-153 and
// RescaleConsts::scale), a*s < 2^160.
VIA_HD via_u128 round_mul_div(via_u128 a, via_u128 s, const WideDiv &w) {
    U192 x = mul_u128_u128_192(a, s);
    x = add_u192_u128(x, w.d >> 1);
    return div_u192(x, w);
}

// ---- RNS (2 limbs) --------------------------------------------------------------

struct Rns2 {
    uint64_t q0, q1;
    uint64_t q0_inv_mod_q1;
    FMod m0, m1;
    via_u128 big_q;
};

// Mod switch of a two-limb residue to q2 in fp64, the value the 128-bit path
// gives, floor((x q2 + floor(Q/2)) / Q) mod q2 with x = a0 + q0 t (Garner).
// With t q2 = K q1 + R (R = t q2 mod q1 exactly; K from one fp64 quotient,
// within 2^-14 of the integer): x q2 / Q = K + F, F = R / q1 + a0 q2 / Q,
// so y = K + floor(F + 1/2 - 1/(2Q)). F is taken to within 2^-50; when
// F + 1/2 lies within 2^-30 of an integer the call declines (false) and the
// caller takes the 128-bit path. Needs q2 < q1 < 2^38 and q0 < 2^38.
struct SwitchF {
    double q2;    // the target modulus
    double inv_q1;  // 1 / q1
    double q2_Q;    // q2 / Q
};

VIA_HD bool modswitch_f64(double a0, double a1, const Rns2 &b, const SwitchF &f, double &y) {
    const double a0m = mod_reduce_small(a0, b.m1);
    const double t = mod_mul(mod_sub(a1, a0m, b.m1), (double)b.q0_inv_mod_q1, b.m1);
    const double R = mod_mul(t, f.q2, b.m1);
    const double K = rint((t * f.q2 - R) * f.inv_q1);
    const double G = R * f.inv_q1 + a0 * f.q2_Q + 0.5;
    const double fl = floor(G);
    const double fr = G - fl;
    if (fr < 0x1p-30 || fr > 1.0 - 0x1p-30) return false;
    double v = K + fl;
    if (v >= f.q2) v -= f.q2;
    y = v;
    return true;
}

// Garner reconstruction (via-rs rns/basis.rs:101-109), limbs given as doubles.
VIA_HD via_u128 rns_reconstruct(double a0, double a1, const Rns2 &b) {
    double a0m = mod_reduce_small(a0, b.m1);  // a0 < q0 < 2^38 < 2^52
    double diff = mod_sub(a1, a0m, b.m1);
    double t = mod_mul(diff, (double)b.q0_inv_mod_q1, b.m1);
    return (via_u128)(uint64_t)a0 + (via_u128)b.q0 * (uint64_t)t;
}

// ---- signed-balanced gadget digits (gadget.rs:168-190) ------------------------

// Signed magnitude value: val = neg ? -mag : mag.
struct SMag {
    via_u128 mag;
    int neg;
};

// ---- fp64 fast paths of the gadget scale --------------------------------------
//
// The gadget scale round-half-up(|c_centred| * B^L / Q) (gadget.rs:115-151,
// wide.rs:145-153) of the exact wide path above, computed in fp64 where that
// is exact or provably decided.

// Single modulus with (q/2 + 1) * B^L + q below 2^53: the numerator
// mag * B^L + floor(q/2) is an exact fp64 integer, the quotient estimate is
// within one, and the exact remainder corrects it.
VIA_HD uint64_t scale_div_f64(uint64_t mag, double bl, double q, double qinv, uint64_t half_q) {
    const double num = (double)mag * bl + (double)half_q;  // exact below 2^53
    double y = floor(num * qinv);
    const double r = fma(-y, q, num);  // exact: |r| < 2q
    if (r < 0.0) y -= 1.0;
    else if (r >= q) y += 1.0;
    return (uint64_t)y;
}

// Two-limb RNS value (a0 mod q0, a1 mod q1), Q = q0 * q1, B^L = bl < 2^40:
// with Garner's t (c = a0 + q0 * t), f = c * bl / Q = t * (bl / q1) +
// a0 * (bl / Q) in fp64 is within 2^-52 * bl of the exact value (two
// roundings of terms below bl), and f + 1/2 within 1.5 * 2^-52 * bl; the
// centring test (c > Q/2 iff f > bl/2, up to bl / 2Q) and the rounding of f
// or of bl - f (exact by Sterbenz) are decided when they lie further than
// `margin` = 2^-47 * bl (20x the error) from their boundaries; otherwise the
// caller takes the exact path.
struct ScaleF64 {
    double bl;      // B^L (exact)
    double bl_q1;   // bl / q1
    double bl_Q;    // bl / Q
    double margin;  // undecided band half-width
};

VIA_HD ScaleF64 scale_f64_make(double bl, uint64_t q1, double Q) {
    ScaleF64 k;
    k.bl = bl;
    k.bl_q1 = bl / (double)q1;
    k.bl_Q = bl / Q;
    k.margin = bl * 7.105427357601002e-15;  // 2^-47 * bl
    return k;
}

VIA_HD bool gadget_scale2_f64(double a0, double a1, const Rns2 &b, const ScaleF64 &k, SMag &v) {
    const double a0m = mod_reduce_small(a0, b.m1);
    const double diff = mod_sub(a1, a0m, b.m1);
    const double t = mod_mul(diff, (double)b.q0_inv_mod_q1, b.m1);
    const double f = fma(t, k.bl_q1, a0 * k.bl_Q);
    const double half = 0.5 * k.bl;
    int neg;
    if (f - half > k.margin) neg = 1;
    else if (half - f > k.margin) neg = 0;
    else return false;
    const double g = (neg ? k.bl - f : f) + 0.5;
    const double fl = floor(g);
    const double fr = g - fl;
    if (fr < k.margin || fr > 1.0 - k.margin) return false;
    v.mag = (via_u128)(uint64_t)fl;
    v.neg = neg && fl != 0.0;
    return true;
}

// Wide B^L (beyond 2^40, e.g. 18^18 ~ 2^75), split for the digits directly.
// With B^L = K*q1 + R (0 <= R < q1) and the centred magnitude
// mag = a0' + q0*t' (two limbs, a0' < q0; one limb: mag = t', a0' = 0):
//   mag * B^L / Q = t'*K + F,  F = t' * (R/q1) + a0' * (B^L/Q) < 2^41,
// so the scale floor((mag * B^L + floor(Q/2)) / Q) is t'*K + floor(F + 1/2 -
// 1/(2Q)). With K, q1 < 2^39 (F < 2^40) F is computed as an unevaluated sum
// s + low: R/q1 and B^L/Q are held as double-double constants, the products
// with t' and a0' split exactly with fma and the sum of the leading parts
// exactly (two-sum), so s + low is within 2^-52 of F and the rounding is
// taken when F + 1/2 lies more than 2^-40 from an integer (else the exact
// path, about one coefficient in 2^39). t'*K < 2^91 is the exact pair
// ph + pl; the split val = hi*B^h + lo takes one quotient estimate of hi
// (within one while hi < 2^40) and the fma-exact remainder. The host enables
// it only within these bounds (scale_w_make, through scale_w_consts).
struct ScaleW {
    double K;         // floor(B^L / q1)
    double r_q1;      // R / q1, leading part
    double r_q1_lo;   // its remainder: R/q1 - r_q1 to 2^-106
    double bl_Q;      // B^L / Q (two limbs), leading part
    double bl_Q_lo;   // its remainder to 2^-100
    double bh;        // B^h of the digit split
    double inv_bh;    // 1 / B^h
};

// The constants of ScaleW that depend on the gadget and the modulus (host):
// q1 the second limb (two limbs, Q = q0 * q1) or the modulus (one limb,
// Q = q1). False outside the bounds above (K, q1 < 2^39 for two limbs,
// K < 2^53 for one, Q < 2^78, B^L / Q < 2^11).
inline bool scale_w_consts(via_u128 bl, via_u128 q1, via_u128 Q, int nl, ScaleW &k) {
    const via_u128 two39 = (via_u128)1 << 39;
    const via_u128 K = bl / q1, R = bl % q1;
    if (q1 >= two39 || Q >= ((via_u128)1 << 78)) return false;
    if (nl == 2 ? K >= two39 : K >= ((via_u128)1 << 53)) return false;
    k.K = (double)(uint64_t)K;
    const double Rf = (double)(uint64_t)R, qf = (double)(uint64_t)q1;
    k.r_q1 = Rf / qf;
    k.r_q1_lo = fma(-k.r_q1, qf, Rf) / qf;  // the residual R - r_q1 q1 is exact
    k.bl_Q = k.bl_Q_lo = 0.0;
    if (nl == 2) {
        // B^L / Q = ip + f1 2^-50 + f2 2^-100 + (below 2^-100), by long division
        const via_u128 ip = bl / Q, rem = bl % Q;
        if (ip >= ((via_u128)1 << 11)) return false;
        const via_u128 f1 = (rem << 50) / Q, rem1 = (rem << 50) % Q;
        const via_u128 f2 = (rem1 << 50) / Q;
        const int64_t N = (int64_t)((ip << 50) | f1);  // below 2^61
        const double Nf = (double)N;                   // N rounded to 53 bits
        const int64_t diff = N - (int64_t)Nf;          // exact, |diff| <= 2^7
        k.bl_Q = Nf * 0x1p-50;
        k.bl_Q_lo = ((double)diff + (double)(uint64_t)f2 * 0x1p-50) * 0x1p-50;
    }
    return true;
}

// Two limbs (a0 mod q0, a1 mod q1, basis b), or one limb a0 mod q (NL = 1,
// q1 = q). Returns the signed halves of the digit split of the scale.
template <int NL>
VIA_HD bool gadget_split_w(double a0, double a1, const Rns2 &b, double q_single, const ScaleW &k, double &lo,
                           double &hi) {
    double tp, F, low;
    int neg;
    if (NL == 2) {
        const double q0 = b.m0.q, q1 = b.m1.q;
        const double a0m = mod_reduce_small(a0, b.m1);
        const double diff = mod_sub(a1, a0m, b.m1);
        const double t = mod_mul(diff, (double)b.q0_inv_mod_q1, b.m1);  // c = a0 + q0*t
        const double c1 = 0.5 * (q1 - 1.0), c0 = 0.5 * (q0 - 1.0);
        neg = t > c1 || (t == c1 && a0 > c0);  // c > floor(Q/2)
        // Q - c = (q0 - a0) + q0*(q1 - 1 - t), or q0*(q1 - t) when a0 = 0
        const double a0p = neg ? (a0 == 0.0 ? 0.0 : q0 - a0) : a0;
        tp = neg ? (a0 == 0.0 ? q1 - t : q1 - 1.0 - t) : t;
        const double p1 = tp * k.r_q1, p2 = a0p * k.bl_Q;
        const double e1 = fma(tp, k.r_q1, -p1), e2 = fma(a0p, k.bl_Q, -p2);  // exact
        F = p1 + p2;
        const double z = F - p1;
        const double es = (p1 - (F - z)) + (p2 - z);  // F + es = p1 + p2 exactly
        low = (es + e1) + (e2 + fma(tp, k.r_q1_lo, a0p * k.bl_Q_lo));
    } else {
        const double q = q_single;
        neg = a0 > 0.5 * (q - 1.0);
        tp = neg ? q - a0 : a0;
        F = tp * k.r_q1;
        low = fma(tp, k.r_q1, -F) + tp * k.r_q1_lo;
    }
    const double G = F + 0.5;  // exact: F < 2^41
    double fl = floor(G);
    double fr = (G - fl) + low;  // |low| < 2^-10
    if (fr < 0.0) {
        fl -= 1.0;
        fr += 1.0;
    } else if (fr >= 1.0) {
        fl += 1.0;
        fr -= 1.0;
    }
    if (fr < 0x1p-40 || fr > 1.0 - 0x1p-40) return false;
    const double ph = tp * k.K;
    const double pl = fma(tp, k.K, -ph);
    const double s = pl + fl;  // exact: |pl| < 2^38, fl < 2^41
    double h = floor((ph + s) * k.inv_bh);
    double l = fma(-h, k.bh, ph) + s;  // exact: |ph - h*B^h| < 2^52
    if (l < 0.0) {
        l += k.bh;
        h -= 1.0;
    } else if (l >= k.bh) {
        l -= k.bh;
        h += 1.0;
    }
    lo = neg ? -l : l;
    hi = neg ? -h : h;
    return true;
}

// One LSB digit of base B (B < 2^20), returning the digit and advancing val
// to (val - digit) / B, with the Euclidean remainder and the "d > B/2 => d -= B"
// rebalance of the reference.
VIA_HD int32_t extract_digit(SMag &v, uint32_t B) {
    uint64_t hi = (uint64_t)(v.mag >> 64), lo = (uint64_t)v.mag;
    uint64_t qh = hi / B;
    uint64_t rh = hi % B;
    uint64_t t1 = (rh << 32) | (lo >> 32);
    uint64_t q1 = t1 / B, r1 = t1 % B;
    uint64_t t0 = (r1 << 32) | (lo & 0xffffffffu);
    uint64_t q0 = t0 / B, rm = t0 % B;
    via_u128 fl = ((via_u128)qh << 64) | ((via_u128)q1 << 32) | q0;  // floor(mag / B)
    int32_t half = (int32_t)(B / 2);
    int32_t d;
    if (!v.neg) {
        d = (int32_t)rm;
        if (d > half) {
            d -= (int32_t)B;
            v.mag = fl + 1;
        } else {
            v.mag = fl;
        }
    } else {
        if (rm == 0) {
            d = 0;
            v.mag = fl;
        } else {
            int32_t re = (int32_t)(B - rm);
            if (re > half) {
                d = -(int32_t)rm;
                v.mag = fl;
            } else {
                d = re;
                v.mag = fl + 1;
            }
        }
    }
    if (v.mag == 0) v.neg = 0;
    return d;
}

// ---- digits in int64 / fp64 lanes --------------------------------------------------
//
// The scaled value |val| <= B^L/2 + 1 is split as val = hi*B^h + lo with
// B^h < 2^50, so both halves fit exact fp64/int64 arithmetic. Extracting the
// first h digits from lo leaves a carry in {-1, 0, 1} that joins hi, and the
// remaining L-h digits come from hi + carry: the same sequence as running
// the reference's rem_euclid / rebalance on the full value.

struct DigitPlan {
    uint32_t base;
    int L;
    int h;            // digits taken from the low part
    uint64_t bh;      // B^h
    double inv_base;  // 1/B
    double inv_bh;    // 1/B^h
    double base_f;    // B
    double half_f;    // floor(B/2): the largest digit
    double neg_f;     // B - 1 - floor(B/2): minus the smallest digit
    double off_f;     // neg_f + 1/2
    int exact_q;      // both halves below 2^49 (B^(L-h) < 2^49): digit_f64's quotient is exact
};

VIA_HD DigitPlan digit_plan(uint32_t base, int L) {
    DigitPlan p;
    p.base = base;
    p.L = L;
    p.h = 0;
    uint64_t bh = 1;
    while (p.h < L && (double)bh * base < 1125899906842624.0) {  // 2^50
        bh *= base;
        p.h++;
    }
    p.bh = bh;
    p.inv_base = 1.0 / (double)base;
    p.inv_bh = 1.0 / (double)bh;
    p.base_f = (double)base;
    p.half_f = (double)(base >> 1);
    p.neg_f = (double)(base - 1 - (base >> 1));
    p.off_f = p.neg_f + 0.5;
    double bhi = 1.0;  // B^(L-h), exact while below 2^53
    for (int k = p.h; k < L && bhi < 562949953421312.0; k++) bhi *= base;
    p.exact_q = bhi < 562949953421312.0;  // 2^49
    return p;
}

// One digit of an int64 value |x| < 2^52: Euclidean remainder, rebalance
// above B/2, then x <- (x - d) / B.
VIA_HD int32_t digit_i64(int64_t &x, uint32_t base, double inv_base) {
    const int64_t B = (int64_t)base;
    int64_t q = (int64_t)floor((double)x * inv_base);
    int64_t r = x - q * B;
    if (r < 0) {
        q -= 1;
        r += B;
    } else if (r >= B) {
        q += 1;
        r -= B;
    }
    if (r > (int64_t)(base >> 1)) {
        r -= B;
        q += 1;
    }
    x = q;
    return (int32_t)r;
}

// The same digit of an exact integer |x| < 2^52 held in fp64: the balanced
// digit r in [-(B-1-floor(B/2)), floor(B/2)] with x = B*q + r is
// x - B*floor((x + B-1-floor(B/2)) / B); the quotient estimate is within one,
// r = x - B*q is exact (fma), and one correction each way fixes it. Equal to
// rem_euclid then the rebalance above B/2 (gadget.rs:168-190). With both
// halves below 2^49 (p.exact_q) no correction is needed: for the integer
// N = x + neg_f, floor(N/B) = floor((N + 1/2)/B), whose argument lies at
// least 1/(2B) from an integer, and (N + 1/2) * fl(1/B), N + 1/2 exact below
// 2^51, is within |N + 1/2| 2^-52 / B < 1/(2B) of it.
VIA_HD int32_t digit_f64(double &x, const DigitPlan &p) {
    if (p.exact_q) {
        const double q = floor((x + p.off_f) * p.inv_base);
        const double r = fma(-q, p.base_f, x);
        x = q;
        return (int32_t)r;
    }
    const double q = floor((x + p.neg_f) * p.inv_base);
    const double r = fma(-q, p.base_f, x);
    const double corr = r > p.half_f ? 1.0 : (r < -p.neg_f ? -1.0 : 0.0);
    x = q + corr;
    return (int32_t)fma(-corr, p.base_f, r);
}

// Splits the scaled magnitude into signed low/high int64 parts.
VIA_HD void digit_split(const SMag &v, const DigitPlan &p, int64_t &lo, int64_t &hi) {
    via_u128 m = v.mag;
    uint64_t qh;
    if ((m >> 64) == 0 && (uint64_t)m < p.bh) {
        qh = 0;
    } else {
        double qf = floor(u128_to_f64(m) * p.inv_bh);
        via_i128 q = (via_i128)(int64_t)qf;
        via_i128 r = (via_i128)m - q * (via_i128)p.bh;
        while (r < 0) {
            q -= 1;
            r += (via_i128)p.bh;
        }
        while (r >= (via_i128)p.bh) {
            q += 1;
            r -= (via_i128)p.bh;
        }
        qh = (uint64_t)q;
    }
    const int64_t l = (int64_t)((uint64_t)m - qh * p.bh);
    lo = v.neg ? -l : l;
    hi = v.neg ? -(int64_t)qh : (int64_t)qh;
}

// digit_split into exact fp64 halves (|lo| < 2^50; |hi| < 2^52 for every
// gadget the host admits), for digit_f64.
VIA_HD void digit_split_f64(const SMag &v, const DigitPlan &p, double &lo, double &hi) {
    int64_t l, h;
    digit_split(v, p, l, h);
    lo = (double)l;
    hi = (double)h;
}
