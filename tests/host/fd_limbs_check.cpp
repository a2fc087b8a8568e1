// Host check of via-rs_amd/csrc/fd_limbs.h: the 7-bit limb split
// reconstructs every residue below 2^35, and the fold of brute-force
// limb-product sums equals the exact row sum sum_i x_i y_i mod q (128-bit
// integers) for row counts up to the stated maximum, random and all-maximal
// residues, the paper's q2 and the largest admitted modulus. Built and run by
// tests/test_host_arith.py; prints "OK <count>".
#include <cstdio>
#include <random>
#include <vector>

#include "../../via-rs_amd/csrc/fd_limbs.h"

typedef unsigned __int128 u128;

static int fails = 0;
static long count = 0;

static void check_split(uint64_t v) {
    uint32_t l[kLimbs];
    limbs7((double)v, l);
    uint64_t r = 0;
    for (int a = kLimbs - 1; a >= 0; a--) {
        if (l[a] >= 128) fails++;
        r = r * 128 + l[a];
    }
    if (r != v) {
        if (fails < 10) printf("split %llu -> %llu\n", (unsigned long long)v, (unsigned long long)r);
        fails++;
    }
    count++;
}

// rows x_i, y_i: fold of the limb sums against the exact sum mod q
static void check_fold(const std::vector<uint64_t> &x, const std::vector<uint64_t> &y, uint64_t q) {
    long long C[kLimbShifts] = {0};
    u128 exact = 0;
    for (size_t i = 0; i < x.size(); i++) {
        uint32_t a[kLimbs], b[kLimbs];
        limbs7((double)x[i], a);
        limbs7((double)y[i], b);
        for (int s = 0; s < kLimbs; s++)
            for (int t = 0; t < kLimbs; t++) C[s + t] += (long long)a[s] * b[t];
        exact = (exact + (u128)x[i] * y[i]) % q;
    }
    int32_t c32[kLimbShifts];
    for (int s = 0; s < kLimbShifts; s++) {
        if (C[s] >= (1LL << 31)) {
            printf("limb sum %d over 2^31 at %zu rows\n", s, x.size());
            fails++;
            return;
        }
        c32[s] = (int32_t)C[s];
    }
    const double got = limb_fold(c32, fmod_make(q));
    if (got != (double)(uint64_t)exact) {
        if (fails < 10)
            printf("fold rows %zu q %llu: got %.0f want %llu\n", x.size(), (unsigned long long)q, got,
                   (unsigned long long)(uint64_t)exact);
        fails++;
    }
    count++;
}

int main() {
    std::mt19937_64 rng(7);
    const uint64_t q2 = 17175674881ULL;         // paper q2
    const uint64_t qmax = (1ULL << 35) - 4095;  // a large modulus below 2^35
    const uint64_t edges[] = {0, 1, 127, 128, 16383, 16384, (1ULL << 21) - 1, 1ULL << 21, (1ULL << 28) - 1,
                              q2 - 1, qmax - 1, (1ULL << 35) - 1};
    for (uint64_t v : edges) check_split(v);
    for (int t = 0; t < 200000; t++) check_split(rng() >> 29);

    const int rows[] = {1, 2, 7, 64, 256, 1000, 4096, kLimbMaxRows};
    const uint64_t qs[] = {q2, qmax, 8380417ULL};
    for (uint64_t q : qs)
        for (int R : rows) {
            const int reps = R >= 4096 ? 3 : 40;
            for (int rep = 0; rep < reps; rep++) {
                std::vector<uint64_t> x(R), y(R);
                for (int i = 0; i < R; i++) {
                    x[i] = rng() % q;
                    y[i] = rng() % q;
                }
                check_fold(x, y, q);
            }
            // all residues maximal: the largest limb sums the fold must take
            std::vector<uint64_t> x(R, q - 1), y(R, q - 1);
            check_fold(x, y, q);
            // limbs all 127 where the modulus allows (the 2^35 - 1 pattern)
            if (q > (1ULL << 35) - 4096) {
                std::vector<uint64_t> z(R, (1ULL << 35) - 1);
                // outside [0, q) but still a valid limb input: checks the
                // largest C_s against the exact sum
                check_fold(z, z, q);
            }
        }
    if (fails) {
        printf("FAIL %d of %ld\n", fails, count);
        return 1;
    }
    printf("OK %ld\n", count);
    return 0;
}
