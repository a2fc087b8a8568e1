// Host check of via-rs_amd/csrc/ntt_lazy.h: the lazy radix-2^R passes, under
// every pass schedule the kernels use, give the radix-2 reference schedule's
// canonical result exactly, and every intermediate stays inside the bounds the
// header states (|v| < 24q forward, < 64q inverse; lazy_mul operands < 2^45);
// the fixed-shape units over the padded LDS layout (the form the kernels
// run) give the same results for every degree and schedule, two limbs and
// many segments per buffer, without touching the pad words.
// Built and run by tests/test_host_arith.py; prints "OK <count>".
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../../via-rs_amd/csrc/ntt_lazy.h"

typedef unsigned __int128 u128;
static uint64_t mulm(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }
static uint64_t addm(uint64_t a, uint64_t b, uint64_t q) { return (a + b) % q; }
static uint64_t subm(uint64_t a, uint64_t b, uint64_t q) { return (a + q - b) % q; }

static double maxabs = 0.0;

struct HostMem {
    double *p;
    double ld(int i) const { return p[i]; }
    void st(int i, double v) const {
        if (std::fabs(v) > maxabs) maxabs = std::fabs(v);
        p[i] = v;
    }
};

static uint64_t canon(double v, uint64_t q) {
    double r = mod_reduce_small(v, fmod_make(q));
    return (uint64_t)r;
}

template <int R>
static void fwd_pass(double *x, int n, int logn, int L0, const double *tw, const FMod &m) {
    for (int u = 0; u < (n >> R); u++) ntt_fwd_unit<R>(HostMem{x}, logn, L0, u, tw, m);
}
template <int R>
static void inv_pass(double *x, int n, int logn, int lh0, const double *tw, const FMod &m) {
    for (int u = 0; u < (n >> R); u++) ntt_inv_unit<R>(HostMem{x}, logn, lh0, u, tw, m);
}

static void lazy_fwd(double *x, int logn, int maxr, const double *tw, const FMod &m) {
    int rs[16];
    const int np = ntt_schedule(logn, maxr, rs);
    int L0 = 0;
    for (int i = 0; i < np; i++) {
        switch (rs[i]) {
            case 1: fwd_pass<1>(x, 1 << logn, logn, L0, tw, m); break;
            case 2: fwd_pass<2>(x, 1 << logn, logn, L0, tw, m); break;
            case 3: fwd_pass<3>(x, 1 << logn, logn, L0, tw, m); break;
            case 4: fwd_pass<4>(x, 1 << logn, logn, L0, tw, m); break;
        }
        L0 += rs[i];
    }
}

static void lazy_inv(double *x, int logn, int maxr, const double *tw, const FMod &m) {
    int rs[16];
    const int np = ntt_schedule(logn, maxr, rs);
    int lh0 = 0;
    for (int i = 0; i < np; i++) {
        switch (rs[i]) {
            case 1: inv_pass<1>(x, 1 << logn, logn, lh0, tw, m); break;
            case 2: inv_pass<2>(x, 1 << logn, logn, lh0, tw, m); break;
            case 3: inv_pass<3>(x, 1 << logn, logn, lh0, tw, m); break;
            case 4: inv_pass<4>(x, 1 << logn, logn, lh0, tw, m); break;
        }
        lh0 += rs[i];
    }
}

// Host mirror of the device pass drivers over the padded layout: every unit
// of every segment of an NL x S buffer, one pass after the other.
template <int LOGN, int MAXR, int P>
static void fwd_passes_host(double *s, int S, int NL, const double *const *tw, const FMod *m) {
    if constexpr (P < sched_np(LOGN, MAXR)) {
        constexpr int R = sched_r(LOGN, MAXR, P), L0 = sched_l0(LOGN, MAXR, P), USEG = LOGN - R;
        const int upl = S >> R;
        for (int t = 0; t < NL * upl; t++) {
            const int l = t / upl, ua = t % upl, seg = ua >> USEG, u = ua & ((1 << USEG) - 1);
            ntt_fwd_unit_p<R, LOGN, L0>(s, l * S + (seg << LOGN), u, tw[l], m[l]);
        }
        fwd_passes_host<LOGN, MAXR, P + 1>(s, S, NL, tw, m);
    }
}

template <int LOGN, int MAXR, int P>
static void inv_passes_host(double *s, int S, int NL, const double *const *tw, const FMod *m) {
    if constexpr (P < sched_np(LOGN, MAXR)) {
        constexpr int R = sched_r(LOGN, MAXR, P), LH0 = sched_l0(LOGN, MAXR, P), USEG = LOGN - R;
        const int upl = S >> R;
        for (int t = 0; t < NL * upl; t++) {
            const int l = t / upl, ua = t % upl, seg = ua >> USEG, u = ua & ((1 << USEG) - 1);
            ntt_inv_unit_p<R, LOGN, LH0>(s, l * S + (seg << LOGN), u, tw[l], m[l]);
        }
        inv_passes_host<LOGN, MAXR, P + 1>(s, S, NL, tw, m);
    }
}

// The padded fixed-shape passes against the reference schedule: two limbs
// (primes q0, q1), S = 2048 logical words per limb holding 2048 / n segments,
// every pass schedule of at most MAXR stages.
template <int LOGN, int MAXR>
static int check_padded(std::mt19937_64 &rng, uint64_t q0, uint64_t q1, long &count) {
    const int n = 1 << LOGN, S = 2048, NL = 2;
    const uint64_t qs[2] = {q0, q1};
    const FMod m[2] = {fmod_make(q0), fmod_make(q1)};
    std::vector<double> tw[2], itw[2];
    std::vector<uint64_t> twu[2], itwu[2];
    for (int l = 0; l < 2; l++) {
        twu[l].resize(n);
        itwu[l].resize(n);
        for (int i = 0; i < n; i++) {
            twu[l][i] = rng() % qs[l];
            itwu[l][i] = rng() % qs[l];
        }
        tw[l].assign(twu[l].begin(), twu[l].end());
        itw[l].assign(itwu[l].begin(), itwu[l].end());
    }
    const double *twp[2] = {tw[0].data(), tw[1].data()}, *itwp[2] = {itw[0].data(), itw[1].data()};
    std::vector<uint64_t> x(NL * S);
    std::vector<double> buf(lds_words(NL * S) + 16, 0.0);
    for (int l = 0; l < NL; l++)
        for (int i = 0; i < S; i++) {
            x[l * S + i] = rng() % qs[l];
            buf[lds_pad(l * S + i)] = (double)x[l * S + i];
        }
    std::vector<double> ibuf = buf;
    fwd_passes_host<LOGN, MAXR, 0>(buf.data(), S, NL, twp, m);
    inv_passes_host<LOGN, MAXR, 0>(ibuf.data(), S, NL, itwp, m);
    int fails = 0;
    for (int l = 0; l < NL; l++)
        for (int seg = 0; seg < S / n; seg++) {
            std::vector<uint64_t> yf(x.begin() + l * S + seg * n, x.begin() + l * S + seg * n + n), yi = yf;
            const uint64_t q = qs[l];
            for (int lm = 0; lm < LOGN; lm++) {
                const int mb = 1 << lm, h = n >> (lm + 1);
                for (int b = 0; b < mb; b++)
                    for (int j = 0; j < h; j++) {
                        const int i0 = b * 2 * h + j, i1 = i0 + h;
                        const uint64_t t = mulm(yf[i1], twu[l][mb + b], q), a = yf[i0];
                        yf[i0] = addm(a, t, q);
                        yf[i1] = subm(a, t, q);
                    }
            }
            for (int lh = 0; lh < LOGN; lh++) {
                const int h = 1 << lh, mb = n >> (lh + 1);
                for (int b = 0; b < mb; b++)
                    for (int j = 0; j < h; j++) {
                        const int i0 = b * 2 * h + j, i1 = i0 + h;
                        const uint64_t a = yi[i0], c = yi[i1];
                        yi[i0] = addm(a, c, q);
                        yi[i1] = mulm(subm(a, c, q), itwu[l][mb + b], q);
                    }
            }
            for (int i = 0; i < n; i++) {
                const int at = lds_pad(l * S + seg * n + i);
                if (canon(buf[at], q) != yf[i] || canon(ibuf[at], q) != yi[i]) {
                    if (fails++ < 5) printf("padded logn=%d maxr=%d limb=%d seg=%d i=%d\n", LOGN, MAXR, l, seg, i);
                    break;
                }
            }
            count++;
        }
    // the pad words are never written
    for (int i = 0; i < lds_words(NL * S); i++)
        if ((i % 17) == 16 && (buf[i] != 0.0 || ibuf[i] != 0.0)) fails++;
    return fails;
}

template <int LOGN>
static int check_padded_all(std::mt19937_64 &rng, long &count) {
    int f = 0;
    f += check_padded<LOGN, 2>(rng, 137438822401ULL, 274810798081ULL, count);
    f += check_padded<LOGN, 3>(rng, 137438822401ULL, 274810798081ULL, count);
    f += check_padded<LOGN, 4>(rng, 268369921ULL, 536608769ULL, count);
    f += check_padded<LOGN, 3>(rng, 17175674881ULL, 8380417ULL, count);
    return f;
}

int main() {
    const uint64_t primes[] = {137438822401ULL, 274810798081ULL, 17175674881ULL, 8380417ULL, 268369921ULL,
                               536608769ULL};
    std::mt19937_64 rng(11);
    long count = 0;
    int fails = 0;
    // schedules: every logn splits into passes of at most maxr stages, in order
    for (int logn = 1; logn <= 11; logn++)
        for (int maxr = 1; maxr <= 4; maxr++) {
            int rs[16], sum = 0;
            const int np = ntt_schedule(logn, maxr, rs);
            for (int i = 0; i < np; i++) {
                if (rs[i] < 1 || rs[i] > maxr) fails++;
                sum += rs[i];
            }
            if (sum != logn) fails++;
        }
    for (uint64_t q : primes) {
        const FMod m = fmod_make(q);
        for (int logn = 1; logn <= 11; logn++) {
            const int n = 1 << logn;
            for (int trial = 0; trial < 6; trial++) {
                // trial 0: extreme values; 1: signed digits; others uniform
                std::vector<uint64_t> tw(n), itw(n), x(n);
                std::vector<double> xin(n);
                for (int i = 0; i < n; i++) {
                    tw[i] = trial == 0 ? q - 1 - (i & 1) : rng() % q;
                    itw[i] = trial == 0 ? q - 1 : rng() % q;
                    if (trial == 0) {
                        x[i] = q - 1;
                        xin[i] = (double)(q - 1);
                    } else if (trial == 1) {
                        const int d = (int)(rng() % 19) - 9;
                        x[i] = d >= 0 ? (uint64_t)d : q - (uint64_t)(-d);
                        xin[i] = (double)d;
                    } else {
                        x[i] = rng() % q;
                        xin[i] = (double)x[i];
                    }
                }
                std::vector<double> twd(tw.begin(), tw.end()), itwd(itw.begin(), itw.end());
                // exact reference schedules (ring/ntt.rs:305-399)
                std::vector<uint64_t> yf = x, yi = x;
                for (int lm = 0; lm < logn; lm++) {
                    const int mb = 1 << lm, h = n >> (lm + 1);
                    for (int b = 0; b < mb; b++)
                        for (int j = 0; j < h; j++) {
                            const int i0 = b * 2 * h + j, i1 = i0 + h;
                            const uint64_t t = mulm(yf[i1], tw[mb + b], q), a = yf[i0];
                            yf[i0] = addm(a, t, q);
                            yf[i1] = subm(a, t, q);
                        }
                }
                for (int lh = 0; lh < logn; lh++) {
                    const int h = 1 << lh, mb = n >> (lh + 1);
                    for (int b = 0; b < mb; b++)
                        for (int j = 0; j < h; j++) {
                            const int i0 = b * 2 * h + j, i1 = i0 + h;
                            const uint64_t a = yi[i0], c = yi[i1];
                            yi[i0] = addm(a, c, q);
                            yi[i1] = mulm(subm(a, c, q), itw[mb + b], q);
                        }
                }
                for (int maxr = 1; maxr <= 4; maxr++) {
                    std::vector<double> z = xin;
                    maxabs = 0.0;
                    lazy_fwd(z.data(), logn, maxr, twd.data(), m);
                    if (maxabs >= 24.0 * (double)q) {
                        if (fails++ < 10) printf("fwd bound q=%llu logn=%d maxr=%d max=%g q\n", (unsigned long long)q,
                                                 logn, maxr, maxabs / (double)q);
                    }
                    for (int i = 0; i < n; i++)
                        if (canon(z[i], q) != yf[i]) {
                            if (fails++ < 10)
                                printf("fwd q=%llu logn=%d maxr=%d i=%d\n", (unsigned long long)q, logn, maxr, i);
                            break;
                        }
                    // inverse from canonical and from lazy (the forward output) inputs
                    for (int from_lazy = 0; from_lazy < 2; from_lazy++) {
                        std::vector<double> w(n);
                        std::vector<uint64_t> want;
                        if (from_lazy) {
                            w = z;  // lazy forward output, |v| < 24q: reduce to the pass input bound first
                            for (int i = 0; i < n; i++) w[i] = lazy_reduce(w[i], m);
                            want = yf;
                            for (int lh = 0; lh < logn; lh++) {
                                const int h = 1 << lh, mb = n >> (lh + 1);
                                for (int b = 0; b < mb; b++)
                                    for (int j = 0; j < h; j++) {
                                        const int i0 = b * 2 * h + j, i1 = i0 + h;
                                        const uint64_t a = want[i0], c = want[i1];
                                        want[i0] = addm(a, c, q);
                                        want[i1] = mulm(subm(a, c, q), itw[mb + b], q);
                                    }
                            }
                        } else {
                            for (int i = 0; i < n; i++) w[i] = (double)x[i];
                            want = yi;
                        }
                        maxabs = 0.0;
                        lazy_inv(w.data(), logn, maxr, itwd.data(), m);
                        if (maxabs >= 64.0 * (double)q) {
                            if (fails++ < 10) printf("inv bound q=%llu logn=%d maxr=%d max=%g q\n",
                                                     (unsigned long long)q, logn, maxr, maxabs / (double)q);
                        }
                        for (int i = 0; i < n; i++)
                            if (canon(w[i], q) != want[i]) {
                                if (fails++ < 10)
                                    printf("inv q=%llu logn=%d maxr=%d i=%d\n", (unsigned long long)q, logn, maxr, i);
                                break;
                            }
                        // the N^-1 scaling of the callers: mod_mul of a lazy value is canonical
                        const double r = mod_mul(w[0], (double)(q - 1), m);
                        if (!(r >= 0 && r < (double)q) || (uint64_t)r != mulm(want[0], q - 1, q)) fails++;
                        count++;
                    }
                    count++;
                }
            }
        }
        // lazy_mul / lazy_reduce ranges at the stated operand limits
        for (int t = 0; t < 200000; t++) {
            const double lim = 63.0 * (double)q;
            double a = std::floor(((double)(rng() >> 11) / 9007199254740992.0) * 2 * lim - lim);
            if (t < 4) a = (t & 1) ? lim : -lim;
            const uint64_t w = t < 8 ? q - 1 : rng() % q;
            const double r = mod_mul_lazy(a, (double)w, m);
            // exact: a*w - r must be a multiple of q and r in [-q, 2q)
            const __int128 ai = (__int128)a, ri = (__int128)r;
            const __int128 diff = ai * (__int128)w - ri;
            if (!(r >= -(double)q && r < 2.0 * (double)q) || diff % (__int128)q != 0) {
                if (fails++ < 10) printf("lazy_mul q=%llu a=%g\n", (unsigned long long)q, a);
            }
            const double v = a * 1024.0;  // |v| < 2^50 for q < 2^38
            const double rr = lazy_reduce(v, m);
            const __int128 d2 = (__int128)v - (__int128)rr;
            if (!(rr >= -(double)q && rr < 2.0 * (double)q) || d2 % (__int128)q != 0) {
                if (fails++ < 10) printf("lazy_reduce q=%llu v=%g\n", (unsigned long long)q, v);
            }
            count++;
        }
    }
    fails += check_padded_all<1>(rng, count) + check_padded_all<2>(rng, count) + check_padded_all<3>(rng, count) +
             check_padded_all<4>(rng, count) + check_padded_all<5>(rng, count) + check_padded_all<6>(rng, count) +
             check_padded_all<7>(rng, count) + check_padded_all<8>(rng, count) + check_padded_all<9>(rng, count) +
             check_padded_all<10>(rng, count) + check_padded_all<11>(rng, count);
    if (fails) {
        printf("FAIL %d\n", fails);
        return 1;
    }
    printf("OK %ld\n", count);
    return 0;
}
