// Host check of via-rs_amd/csrc/modarith.h against 128-bit integer arithmetic.
// Built and run by tests/test_host_arith.py; prints "OK <count>" on success.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../via-rs_amd/csrc/modarith.h"

static int fails = 0;
#define CHECK(c, ...)                                   \
    do {                                                \
        if (!(c)) {                                     \
            if (fails++ < 10) fprintf(stderr, __VA_ARGS__); \
        }                                               \
    } while (0)

static via_u128 exact_rmd(via_u128 a, via_u128 s, via_u128 d) {
    // floor((a*s + d/2)/d) with 256-bit long division in 32-bit digits.
    uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64), s0 = (uint64_t)s, s1 = (uint64_t)(s >> 64);
    via_u128 p00 = (via_u128)a0 * s0, p01 = (via_u128)a0 * s1, p10 = (via_u128)a1 * s0, p11 = (via_u128)a1 * s1;
    via_u128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;
    via_u128 lo = ((via_u128)(uint64_t)mid << 64) | (uint64_t)p00;
    via_u128 hi = p11 + (p01 >> 64) + (p10 >> 64) + (mid >> 64);
    via_u128 h = d / 2, nlo = lo + h;
    if (nlo < lo) hi++;
    uint32_t w[8];
    for (int i = 0; i < 4; i++) {
        w[i] = (uint32_t)(nlo >> (32 * i));
        w[4 + i] = (uint32_t)(hi >> (32 * i));
    }
    via_u128 rem = 0, quo = 0;
    for (int i = 7; i >= 0; i--) {
        via_u128 cur = (rem << 32) | w[i];
        via_u128 qd = cur / d;
        rem = cur - qd * d;
        if (i < 4) quo |= qd << (32 * i);
    }
    return quo;
}

int main() {
    std::mt19937_64 rng(12345);
    long count = 0;
    const uint64_t primes[] = {137438822401ULL, 274810798081ULL, 17175674881ULL, 8380417ULL,
                               268369921ULL,    536608769ULL,    65537ULL,       17ULL};
    for (uint64_t q : primes) {
        FMod m = fmod_make(q);
        for (int t = 0; t < 200000; t++) {
            uint64_t a = rng() % q, b = rng() % q;
            if (t < 4) { a = q - 1 - t; b = q - 1; }
            double r = mod_mul((double)a, (double)b, m);
            uint64_t want = (uint64_t)(((via_u128)a * b) % q);
            CHECK((uint64_t)r == want && r >= 0 && r < (double)q, "mul q=%llu a=%llu b=%llu\n",
                  (unsigned long long)q, (unsigned long long)a, (unsigned long long)b);
            double s = mod_add((double)a, (double)b, m), d = mod_sub((double)a, (double)b, m);
            CHECK((uint64_t)s == (a + b) % q && (uint64_t)d == (a + q - b) % q, "add/sub\n");
            count++;
        }
    }
    // Wide scaling: gadget scale at q1 RNS (B=18, L=18 and B=55879, L=2) and mod switch q1 -> q2.
    const via_u128 Q1 = (via_u128)137438822401ULL * 274810798081ULL;
    via_u128 bl18 = 1, bl55 = (via_u128)55879 * 55879;
    for (int i = 0; i < 18; i++) bl18 *= 18;
    WideDiv wq1 = widediv_make(Q1);
    for (int t = 0; t < 300000; t++) {
        via_u128 c = (((via_u128)rng() << 64) | rng()) % (Q1 / 2 + 1);
        if (t < 8) c = (t < 4) ? (via_u128)t : Q1 / 2 - (t - 4);
        via_u128 s = (t % 3 == 0) ? bl18 : (t % 3 == 1 ? bl55 : (via_u128)17175674881ULL);
        via_u128 got = round_mul_div(c, s, wq1), want = exact_rmd(c, s, Q1);
        CHECK(got == want, "rmd t=%d\n", t);
        count++;
    }
    // Digits against the reference rule (rem_euclid, rebalance above B/2).
    const uint32_t bases[] = {18, 55879, 81, 8, 64, 4, 2};
    for (uint32_t B : bases) {
        for (int t = 0; t < 100000; t++) {
            via_i128 val = (via_i128)((((via_u128)rng() << 64) | rng()) >> (48 + t % 40));
            if (rng() & 1) val = -val;
            if (t < 3) val = t - 1;
            SMag v;
            v.neg = val < 0;
            v.mag = (via_u128)(val < 0 ? -val : val);
            via_i128 ref = val;
            for (int k = 0; k < 4; k++) {
                via_i128 d = ref % (via_i128)B;
                if (d < 0) d += B;
                if (d > (via_i128)(B / 2)) d -= B;
                ref = (ref - d) / (via_i128)B;
                int32_t got = extract_digit(v, B);
                via_i128 cur = v.neg ? -(via_i128)v.mag : (via_i128)v.mag;
                CHECK(got == (int32_t)d && cur == ref, "digit B=%u t=%d k=%d\n", B, t, k);
            }
            count++;
        }
    }
    // int64 split digits against the u128 reference digits.
    {
        // 2^100: the high half reaches 2^51, beyond the correction-free quotient
        const uint32_t bs[] = {18, 55879, 81, 8, 64, 3, 255, 32767, 2, 7};
        const int Ls[] = {18, 2, 2, 8, 7, 40, 9, 5, 100, 26};
        for (int bi = 0; bi < 10; bi++) {
            uint32_t B = bs[bi];
            int L = Ls[bi];
            DigitPlan pl = digit_plan(B, L);
            via_u128 bl = 1;
            for (int k = 0; k < L; k++) bl *= B;
            CHECK(pl.exact_q == (B != 2), "exact quotient flag B=%u L=%d\n", B, L);
            for (int t = 0; t < 200000; t++) {
                via_u128 mag = ((((via_u128)rng() << 64) | rng()) % (bl / 2 + 2));
                if (t < 4) mag = (t < 2) ? (via_u128)t : bl / 2 + (t - 2);
                // every digit on the rounding edge: mag = sum of (B/2) B^k or
                // of -(B-1-B/2) B^k (balanced extremes), and B^k - 1
                if (t >= 4 && t < 4 + 3 * L) {
                    const int kk = (t - 4) % L, kind = (t - 4) / L;
                    via_u128 pw = 1, acc = 0;
                    for (int k = 0; k <= kk; k++) {
                        if (kind == 0) acc += (via_u128)(B / 2) * pw;
                        pw *= B;
                    }
                    mag = kind == 0 ? acc : kind == 1 ? pw - 1 : pw;
                    if (mag > bl / 2) mag = bl / 2;
                }
                SMag v;
                v.mag = mag;
                v.neg = (rng() & 1) && mag != 0;
                SMag ref = v;
                int64_t lo, hi;
                digit_split(v, pl, lo, hi);
                double flo, fhi;  // the fp64 digit path of the kernels
                digit_split_f64(v, pl, flo, fhi);
                for (int k = 0; k < L; k++) {
                    int32_t want = extract_digit(ref, B);
                    int32_t got, gotf;
                    if (k < pl.h) {
                        got = digit_i64(lo, B, pl.inv_base);
                        if (k == pl.h - 1) hi += lo;
                        gotf = digit_f64(flo, pl);
                        if (k == pl.h - 1) fhi += flo;
                    } else {
                        got = digit_i64(hi, B, pl.inv_base);
                        gotf = digit_f64(fhi, pl);
                    }
                    CHECK(got == want, "split digit B=%u t=%d k=%d\n", B, t, k);
                    CHECK(gotf == want, "fp64 split digit B=%u t=%d k=%d\n", B, t, k);
                }
                count++;
            }
        }
    }
    // Garner reconstruction.
    Rns2 b;
    b.q0 = 137438822401ULL;
    b.q1 = 274810798081ULL;
    b.m0 = fmod_make(b.q0);
    b.m1 = fmod_make(b.q1);
    {
        // q0^-1 mod q1 by extended Euclid
        long long t0 = 0, t1 = 1;
        long long r0 = (long long)b.q1, r1 = (long long)(b.q0 % b.q1);
        while (r1) {
            long long qq = r0 / r1, tmp = t0 - qq * t1;
            t0 = t1; t1 = tmp;
            tmp = r0 - qq * r1; r0 = r1; r1 = tmp;
        }
        if (t0 < 0) t0 += (long long)b.q1;
        b.q0_inv_mod_q1 = (uint64_t)t0;
    }
    b.big_q = Q1;
    for (int t = 0; t < 200000; t++) {
        via_u128 x = (((via_u128)rng() << 64) | rng()) % Q1;
        double a0 = (double)(uint64_t)(x % b.q0), a1 = (double)(uint64_t)(x % b.q1);
        CHECK(rns_reconstruct(a0, a1, b) == x, "garner\n");
        count++;
    }
    // fp64 mod switch q1 -> q2 (modswitch_f64) against floor((x q2 +
    // floor(Q/2)) / Q) mod q2: random x, x within a few units of every
    // rounding boundary (x q2 = (m + 1/2) Q), the ends of [0, Q) and
    // multiples of q0; a decided value must be exact, and values right at a
    // boundary must be declined.
    {
        const uint64_t q2 = 17175674881ULL;
        SwitchF sf;
        sf.q2 = (double)q2;
        sf.inv_q1 = 1.0 / (double)b.q1;
        sf.q2_Q = (double)q2 / u128_to_f64(Q1);
        long decided = 0, declined = 0, declined_near = 0;
        auto check = [&](via_u128 x, bool near) {
            const double a0 = (double)(uint64_t)(x % b.q0), a1 = (double)(uint64_t)(x % b.q1);
            const uint64_t want = (uint64_t)(exact_rmd(x, q2, Q1) % q2);
            double y;
            if (modswitch_f64(a0, a1, b, sf, y)) {
                decided++;
                CHECK((uint64_t)y == want, "modswitch x=%llu:%llu got %llu want %llu\n",
                      (unsigned long long)(x >> 64), (unsigned long long)x, (unsigned long long)y,
                      (unsigned long long)want);
            } else {
                declined++;
                if (near) declined_near++;
            }
            count++;
        };
        for (int t = 0; t < 400000; t++) check((((via_u128)rng() << 64) | rng()) % Q1, false);
        const long random_declined = declined;
        for (int t = 0; t < 20000; t++) {
            // x q2 within a few units of (m + 1/2) Q: x0 = floor(((2m + 1) Q) / (2 q2))
            const via_u128 m = (((via_u128)rng() << 64) | rng()) % q2;
            const via_u128 x0 = exact_rmd(2 * m + 1, Q1, 2 * (via_u128)q2) - 1;
            for (int d = -2; d <= 2; d++) {
                const via_u128 x = x0 + (via_u128)(d + 2) - 2;
                if (x < Q1) check(x, true);
            }
        }
        for (int d = 0; d < 4; d++) {
            check((via_u128)d, false);
            check(Q1 - 1 - d, false);
            check((via_u128)b.q0 * (d + 1), false);
            check((via_u128)b.q0 * (b.q1 - 1 - d), false);
        }
        CHECK(random_declined < 10, "modswitch: %ld random values declined\n", random_declined);
        CHECK(decided > 400000, "modswitch: too few decided (%ld)\n", decided);
        // x0 + 1 sits where x q2 / Q + 1/2 is within q2 / Q of an integer
        CHECK(declined_near > 0, "modswitch: no boundary value declined\n");
    }
    // fp64 two-limb gadget scale (the DMux gadget B = 55879, L = 2 at Q1, and
    // B = 81, L = 2 for the bound) against the exact centred scale: random
    // values, values within a few units of every rounding boundary
    // (c ~ (k + 1/2) Q / B^L) and of the centring boundary Q/2, on both sides.
    {
        long decided = 0, total = 0;
        const uint64_t bases[] = {55879, 81, 1024};
        for (uint64_t B : bases) {
            const via_u128 bl = (via_u128)B * B;
            const ScaleF64 k = scale_f64_make((double)(uint64_t)bl, b.q1, u128_to_f64(Q1));
            auto exact = [&](via_u128 c, SMag &v) {
                const int neg = c > (Q1 >> 1);
                const via_u128 mag = neg ? Q1 - c : c;
                v.mag = round_mul_div(mag, bl, wq1);
                v.neg = neg && v.mag != 0;
            };
            auto check = [&](via_u128 c) {
                c %= Q1;
                const double a0 = (double)(uint64_t)(c % b.q0), a1 = (double)(uint64_t)(c % b.q1);
                SMag want, got;
                exact(c, want);
                total++;
                if (gadget_scale2_f64(a0, a1, b, k, got)) {
                    decided++;
                    CHECK(got.mag == want.mag && got.neg == want.neg, "scale2 f64 B=%llu\n", (unsigned long long)B);
                }
                count++;
            };
            const long d0 = decided, t0 = total;
            for (int t = 0; t < 300000; t++) check((((via_u128)rng() << 64) | rng()) % Q1);
            // random values are decided by the fast path nearly always
            CHECK(decided - d0 > (total - t0) * 99 / 100, "scale2 f64 B=%llu decided only %ld of %ld\n",
                  (unsigned long long)B, decided - d0, total - t0);
            for (int t = 0; t < 100000; t++) {
                // (k + 1/2) Q / bl for a random k < bl, then a few units either side
                const via_u128 kk = (((via_u128)rng() << 64) | rng()) % bl;
                const via_u128 c0 = ((2 * kk + 1) * (Q1 / bl) + ((2 * kk + 1) * (Q1 % bl)) / bl) / 2;
                for (int d = -3; d <= 3; d++) {
                    check(c0 + (via_u128)(d + 3) - 3);
                    check(Q1 - (c0 + (via_u128)(d + 3) - 3));
                }
            }
            for (int d = -50; d <= 50; d++) check((Q1 >> 1) + (via_u128)(d + 50) - 50);
        }
    }
    // fp64 single-modulus scale against 64-bit integer division: the CMux
    // gadget at q2 (81^2) and the ring-switch gadget at q3 (8^8).
    {
        const uint64_t qs[] = {17175674881ULL, 8380417ULL};
        const uint64_t bls[] = {6561, 16777216};
        for (int i = 0; i < 2; i++) {
            const uint64_t q = qs[i], bl = bls[i];
            const FMod m = fmod_make(q);
            for (int t = 0; t < 300000; t++) {
                uint64_t mag = rng() % (q / 2 + 1);
                if (t < 4) mag = (t & 1) ? q / 2 : (uint64_t)t;
                const uint64_t want = (uint64_t)(((via_u128)mag * bl + (q >> 1)) / q);
                CHECK(scale_div_f64(mag, (double)bl, m.q, m.qinv, q >> 1) == want, "scale1 f64 q=%llu\n",
                      (unsigned long long)q);
                count++;
            }
        }
    }
    // fp64 wide split (gadget_split_w) against the exact scale and digit
    // split: the cascade gadget 18^18 at Q1 (two limbs) and at q2 (one limb,
    // the VIA-B repack), 16^19 and 7^26 at Q1; random values, values near
    // every rounding boundary ((k + 1/2) Q / B^L) and near Q/2, both sides.
    {
        struct Case {
            uint64_t base;
            int L;
            int nl;
        };
        const Case cases[] = {{18, 18, 2}, {16, 19, 2}, {7, 26, 2}, {18, 18, 1}};
        const uint64_t q2 = 17175674881ULL;
        for (const Case &cs : cases) {
            via_u128 bl = 1;
            for (int k = 0; k < cs.L; k++) bl *= cs.base;
            const DigitPlan dp = digit_plan((uint32_t)cs.base, cs.L);
            const via_u128 Qm = cs.nl == 2 ? Q1 : (via_u128)q2;
            const via_u128 q1 = cs.nl == 2 ? (via_u128)b.q1 : (via_u128)q2;
            const WideDiv wq = widediv_make(Qm);
            ScaleW sw;  // as scale_w_make builds it
            CHECK(scale_w_consts(bl, q1, Qm, cs.nl, sw), "wide split constants B=%llu L=%d\n",
                  (unsigned long long)cs.base, cs.L);
            sw.bh = (double)dp.bh;
            sw.inv_bh = dp.inv_bh;
            // (x * y) mod Qm for x, y < Qm < 2^78 by doubling
            auto mulmod = [&](via_u128 x, via_u128 y) {
                via_u128 acc = 0;
                for (int i = 127; i >= 0; i--) {
                    acc <<= 1;
                    if (acc >= Qm) acc -= Qm;
                    if ((y >> i) & 1) {
                        acc += x;
                        if (acc >= Qm) acc -= Qm;
                    }
                }
                return acc;
            };
            const via_u128 blm = bl % Qm;
            long decided = 0, total = 0;
            auto check = [&](via_u128 c) {
                c %= Qm;
                const int neg = c > (Qm >> 1);
                const via_u128 mg = neg ? Qm - c : c;
                SMag v;
                v.mag = round_mul_div(mg, bl, wq);
                v.neg = neg && v.mag != 0;
                int64_t lo, hi;
                digit_split(v, dp, lo, hi);
                double flo, fhi;
                const double a0 = (double)(uint64_t)(c % (cs.nl == 2 ? (via_u128)b.q0 : Qm));
                const double a1 = cs.nl == 2 ? (double)(uint64_t)(c % b.q1) : 0.0;
                const bool ok = cs.nl == 2 ? gadget_split_w<2>(a0, a1, b, 0.0, sw, flo, fhi)
                                           : gadget_split_w<1>(a0, 0.0, b, (double)q2, sw, flo, fhi);
                total++;
                if (ok) {
                    decided++;
                    CHECK(flo == (double)lo && fhi == (double)hi, "wide split B=%llu L=%d nl=%d\n",
                          (unsigned long long)cs.base, cs.L, cs.nl);
                } else {
                    // declined only within 2^-38 of a rounding boundary: the
                    // fraction of mag * B^L / Q + 1/2 is (mg * bl mod Q + Q/2) / Q
                    via_u128 fr = mulmod(mg, blm) + (Qm >> 1);
                    if (fr >= Qm) fr -= Qm;
                    const via_u128 dist = fr < Qm - fr ? fr : Qm - fr;
                    CHECK(dist <= (Qm >> 38), "wide split B=%llu L=%d nl=%d declined a decided value\n",
                          (unsigned long long)cs.base, cs.L, cs.nl);
                }
                count++;
            };
            for (int t = 0; t < 300000; t++) check((((via_u128)rng() << 64) | rng()) % Qm);
            CHECK(decided == total, "wide split B=%llu decided only %ld of %ld\n", (unsigned long long)cs.base,
                  decided, total);
            const WideDiv w2bl = widediv_make(2 * bl);
            for (int t = 0; t < 100000; t++) {
                // c near (k + 1/2) Q / bl: mag * bl / Q has fraction near 1/2
                const via_u128 kk = (((via_u128)rng() << 64) | rng()) % (bl / 2);
                const via_u128 num = (2 * kk + 1);
                // floor(num * Qm / (2 bl)) exactly, through the 192-bit product
                const via_u128 c0 = div_u192(mul_u128_u128_192(num, Qm), w2bl);
                for (int d = -3; d <= 3; d++) {
                    check(c0 + (via_u128)(d + 3) - 3);
                    check(Qm - (c0 + (via_u128)(d + 3) - 3));
                }
            }
            // magnitudes whose rounding argument lies delta / Q from a boundary:
            // mag * B^L = delta - floor(Q/2) (mod Q), through the inverse of B^L
            {
                via_i128 t0 = 0, t1 = 1, r0 = (via_i128)Qm, r1 = (via_i128)blm;
                while (r1 != 0) {
                    const via_i128 qq = r0 / r1, t2 = t0 - qq * t1, r2 = r0 - qq * r1;
                    t0 = t1;
                    t1 = t2;
                    r0 = r1;
                    r1 = r2;
                }
                CHECK(r0 == 1, "B^L not invertible mod Q\n");
                const via_u128 inv = (via_u128)(t0 < 0 ? t0 + (via_i128)Qm : t0);
                const long before = decided, tot0 = total;
                const via_u128 steps[] = {0, 1, 2, 3, 64, Qm >> 44, Qm >> 40, (Qm >> 40) * 3, Qm >> 36, Qm >> 30};
                for (via_u128 st : steps)
                    for (int sgn = 0; sgn < 2; sgn++) {
                        const via_u128 target = (Qm - (Qm >> 1) + (sgn ? Qm - st % Qm : st % Qm)) % Qm;
                        const via_u128 mg = mulmod(target, inv);
                        if (mg > (Qm >> 1)) continue;  // not a centred magnitude
                        check(mg);
                        if (mg) check(Qm - mg);
                    }
                // two limbs: the smallest distances (below 2^-40) are declined
                if (cs.nl == 2)
                    CHECK(total - tot0 > decided - before, "wide split B=%llu: no boundary value declined\n",
                          (unsigned long long)cs.base);
            }
            for (int d = -50; d <= 50; d++) check((Qm >> 1) + (via_u128)(d + 50) - 50);
            for (int d = 0; d < 4; d++) check((via_u128)d);
            if (cs.nl == 2) {
                // multiples of q0 (a0 = 0) on both sides of Q/2 and near Q
                for (int d = 0; d < 6; d++) {
                    check((via_u128)b.q0 * (b.q1 / 2 - 2 + d));
                    check((via_u128)b.q0 * (b.q1 - 1 - d));
                }
                for (int t = 0; t < 20000; t++) check((via_u128)b.q0 * (rng() % b.q1));
            }
        }
    }
    if (fails) {
        printf("FAIL %d\n", fails);
        return 1;
    }
    printf("OK %ld\n", count);
    return 0;
}
