// keccak.h — Keccak-f[1600] and the Shake256Prg block (host and device).
//
// via-rs sampling/prg.rs:1-17 defines the PRG byte stream: block k is the
// first 136 bytes (the SHAKE-256 rate) of SHAKE-256(seed || u64_le(k)), a
// fresh hash per block, seeds of at most 64 bytes. seed || counter is at most
// 72 bytes, so each block is one absorbed rate block (padding 0x1F ... 0x80)
// and one permutation, and the 136 output bytes are the first 17 lanes of the
// state. The permutation is FIPS 202's Keccak-f[1600] (24 rounds of theta,
// rho and pi, chi, iota).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define KECCAK_HD __host__ __device__ __forceinline__
#else
#define KECCAK_HD static inline
#endif

constexpr int kPrgBlock = 136;   // bytes per PRG block
constexpr int kPrgSeedCap = 64;  // longest seed (prg.rs SEED_CAP)

KECCAK_HD uint64_t keccak_rotl(uint64_t x, int n) { return n ? (x << n) | (x >> (64 - n)) : x; }

KECCAK_HD void keccak_f1600(uint64_t a[25]) {
    // iota round constants and rho offsets (FIPS 202, 3.2.5 and 3.2.2), the
    // offsets indexed by lane x + 5y
    const uint64_t rc[24] = {0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808AULL,
                             0x8000000080008000ULL, 0x000000000000808BULL, 0x0000000080000001ULL,
                             0x8000000080008081ULL, 0x8000000000008009ULL, 0x000000000000008AULL,
                             0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000AULL,
                             0x000000008000808BULL, 0x800000000000008BULL, 0x8000000000008089ULL,
                             0x8000000000008003ULL, 0x8000000000008002ULL, 0x8000000000000080ULL,
                             0x000000000000800AULL, 0x800000008000000AULL, 0x8000000080008081ULL,
                             0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};
    const int rho[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
    for (int round = 0; round < 24; round++) {
        uint64_t c[5], d[5], b[25];
#pragma unroll
        for (int x = 0; x < 5; x++) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
#pragma unroll
        for (int x = 0; x < 5; x++) d[x] = c[(x + 4) % 5] ^ keccak_rotl(c[(x + 1) % 5], 1);
#pragma unroll
        for (int i = 0; i < 25; i++) a[i] ^= d[i % 5];
        // rho and pi: lane (x, y) moves to (y, 2x + 3y)
#pragma unroll
        for (int x = 0; x < 5; x++)
#pragma unroll
            for (int y = 0; y < 5; y++) b[y + 5 * ((2 * x + 3 * y) % 5)] = keccak_rotl(a[x + 5 * y], rho[x + 5 * y]);
        // chi
#pragma unroll
        for (int y = 0; y < 5; y++)
#pragma unroll
            for (int x = 0; x < 5; x++)
                a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
        a[0] ^= rc[round];
    }
}

// Block k of the Shake256Prg stream (prg.rs:117-128) as 17 little-endian
// lanes: seed (seed_len <= 64 bytes) || u64_le(k), SHAKE padding, one
// permutation.
KECCAK_HD void prg_block(const uint8_t *seed, int seed_len, uint64_t k, uint64_t out[17]) {
    uint64_t a[25];
#pragma unroll
    for (int i = 0; i < 25; i++) a[i] = 0;
    // absorb seed || counter byte by byte into the little-endian lanes
    for (int i = 0; i < seed_len; i++) a[i >> 3] ^= (uint64_t)seed[i] << (8 * (i & 7));
    for (int i = 0; i < 8; i++) {
        const int p = seed_len + i;
        a[p >> 3] ^= ((k >> (8 * i)) & 0xFF) << (8 * (p & 7));
    }
    const int p = seed_len + 8;
    a[p >> 3] ^= (uint64_t)0x1F << (8 * (p & 7));               // SHAKE domain bits + pad start
    a[(kPrgBlock - 1) >> 3] ^= (uint64_t)0x80 << (8 * ((kPrgBlock - 1) & 7));  // pad end
    keccak_f1600(a);
#pragma unroll
    for (int i = 0; i < 17; i++) out[i] = a[i];
}
