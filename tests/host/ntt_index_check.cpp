// Host check of via-rs_amd/csrc/ntt_index.h: radix-4 passes (plus a trailing
// radix-2 stage for odd log n) reproduce the radix-2 stage-by-stage NTT.
// Built and run by tests/test_host_arith.py; prints "OK <count>".
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../../via-rs_amd/csrc/ntt_index.h"

typedef unsigned __int128 u128;
static uint64_t mulm(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }
static uint64_t addm(uint64_t a, uint64_t b, uint64_t q) { return (a + b) % q; }
static uint64_t subm(uint64_t a, uint64_t b, uint64_t q) { return (a + q - b) % q; }

int main() {
    const uint64_t q = 137438822401ULL;
    std::mt19937_64 rng(7);
    long count = 0;
    for (int logn = 1; logn <= 11; logn++) {
        const int n = 1 << logn;
        std::vector<uint64_t> tw(n), itw(n), x(n), y, z;
        for (auto &v : tw) v = rng() % q;
        for (auto &v : itw) v = rng() % q;
        for (auto &v : x) v = rng() % q;
        // forward radix-2 reference schedule
        y = x;
        for (int lm = 0; lm < logn; lm++) {
            int m = 1 << lm, h = n >> (lm + 1);
            for (int blk = 0; blk < m; blk++)
                for (int j = 0; j < h; j++) {
                    int i0 = blk * 2 * h + j, i1 = i0 + h;
                    uint64_t t = mulm(y[i1], tw[m + blk], q);
                    uint64_t a = y[i0];
                    y[i0] = addm(a, t, q);
                    y[i1] = subm(a, t, q);
                }
        }
        // radix-4 passes
        z = x;
        int lm = 0;
        for (; lm + 1 < logn; lm += 2) {
            for (int u = 0; u < n / 4; u++) {
                R4Idx r = r4_fwd(u, lm, logn);
                uint64_t a = z[r.a], b = z[r.b], c = z[r.c], d = z[r.d];
                uint64_t t1 = mulm(c, tw[r.w1], q), t2 = mulm(d, tw[r.w1], q);
                uint64_t a1 = addm(a, t1, q), c1 = subm(a, t1, q), b1 = addm(b, t2, q), d1 = subm(b, t2, q);
                uint64_t t3 = mulm(b1, tw[r.w2], q), t4 = mulm(d1, tw[r.w3], q);
                z[r.a] = addm(a1, t3, q);
                z[r.b] = subm(a1, t3, q);
                z[r.c] = addm(c1, t4, q);
                z[r.d] = subm(c1, t4, q);
            }
        }
        if (lm < logn) {
            int m = 1 << lm, h = n >> (lm + 1);
            for (int blk = 0; blk < m; blk++)
                for (int j = 0; j < h; j++) {
                    int i0 = blk * 2 * h + j, i1 = i0 + h;
                    uint64_t t = mulm(z[i1], tw[m + blk], q), a = z[i0];
                    z[i0] = addm(a, t, q);
                    z[i1] = subm(a, t, q);
                }
        }
        if (y != z) { printf("FAIL fwd logn=%d\n", logn); return 1; }
        // inverse radix-2 reference schedule
        y = x;
        for (int lh = 0; lh < logn; lh++) {
            int h = 1 << lh, m = n >> (lh + 1);
            for (int blk = 0; blk < m; blk++)
                for (int j = 0; j < h; j++) {
                    int i0 = blk * 2 * h + j, i1 = i0 + h;
                    uint64_t a = y[i0], b = y[i1];
                    y[i0] = addm(a, b, q);
                    y[i1] = mulm(subm(a, b, q), itw[m + blk], q);
                }
        }
        z = x;
        int lh = 0;
        for (; lh + 1 < logn; lh += 2) {
            for (int u = 0; u < n / 4; u++) {
                R4Idx r = r4_inv(u, lh, logn);
                uint64_t a = z[r.a], b = z[r.b], c = z[r.c], d = z[r.d];
                uint64_t a1 = addm(a, b, q), b1 = mulm(subm(a, b, q), itw[r.w1], q);
                uint64_t c1 = addm(c, d, q), d1 = mulm(subm(c, d, q), itw[r.w2], q);
                z[r.a] = addm(a1, c1, q);
                z[r.c] = mulm(subm(a1, c1, q), itw[r.w3], q);
                z[r.b] = addm(b1, d1, q);
                z[r.d] = mulm(subm(b1, d1, q), itw[r.w3], q);
            }
        }
        if (lh < logn) {
            int h = 1 << lh, m = n >> (lh + 1);
            for (int blk = 0; blk < m; blk++)
                for (int j = 0; j < h; j++) {
                    int i0 = blk * 2 * h + j, i1 = i0 + h;
                    uint64_t a = z[i0], b = z[i1];
                    z[i0] = addm(a, b, q);
                    z[i1] = mulm(subm(a, b, q), itw[m + blk], q);
                }
        }
        if (y != z) { printf("FAIL inv logn=%d\n", logn); return 1; }
        count += 2;
    }
    printf("OK %ld\n", count);
    return 0;
}
