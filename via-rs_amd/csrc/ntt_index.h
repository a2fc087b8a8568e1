// ntt_index.h — index maps of the radix-4 passes of the negacyclic NTT.
//
// A radix-4 pass applies two consecutive radix-2 stages of the reference
// schedule (via-rs ring/ntt.rs:305-399) to four elements held in registers.
// The butterflies and twiddles are exactly those of the two stages, so the
// result is identical to running the stages one at a time.
#pragma once

#if defined(__HIPCC__)
#define NTT_HD __host__ __device__ __forceinline__
#else
#define NTT_HD static inline
#endif

struct R4Idx {
    int a, b, c, d;     // element offsets inside one polynomial
    int w1, w2, w3;     // twiddle table indices
};

// Forward DIT, stages lm (m = 2^lm blocks, half h = n >> (lm+1)) and lm+1,
// unit u in [0, n/4). Stage lm pairs (a, c) and (b, d) with w1; stage lm+1
// pairs (a, b) with w2 and (c, d) with w3.
NTT_HD R4Idx r4_fwd(int u, int lm, int logn) {
    const int lh = logn - lm - 1;  // log2 h, h >= 2
    const int h = 1 << lh;
    const int q = h >> 1;
    const int m = 1 << lm;
    const int blk = u >> (lh - 1);
    const int j = u & (q - 1);
    const int base = (blk << (lh + 1)) + j;
    R4Idx r;
    r.a = base;
    r.b = base + q;
    r.c = base + h;
    r.d = base + h + q;
    r.w1 = m + blk;
    r.w2 = 2 * m + 2 * blk;
    r.w3 = 2 * m + 2 * blk + 1;
    return r;
}

// Inverse DIF, stages lh (half h = 2^lh, m = n >> (lh+1) blocks) and lh+1,
// unit u in [0, n/4). Stage lh pairs (a, b) with w1 and (c, d) with w2;
// stage lh+1 pairs (a, c) and (b, d) with w3.
NTT_HD R4Idx r4_inv(int u, int lh, int logn) {
    const int h = 1 << lh;
    const int m = 1 << (logn - lh - 1);
    const int B2 = u >> lh;
    const int j = u & (h - 1);
    const int base = (B2 << (lh + 2)) + j;
    R4Idx r;
    r.a = base;
    r.b = base + h;
    r.c = base + 2 * h;
    r.d = base + 3 * h;
    r.w1 = m + 2 * B2;
    r.w2 = m + 2 * B2 + 1;
    r.w3 = (m >> 1) + B2;
    return r;
}
