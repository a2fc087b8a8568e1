/*
 * prg.c — SHAKE-256 counter-mode PRG and the samplers that read it.
 * Test infrastructure (see via_oracle.h).
 *
 * The PRG follows crates/via-primitives/src/sampling/prg.rs:92-205: block k of
 * the stream is the first 136 bytes of SHAKE-256(seed || u64_le(k)), with a
 * fresh sponge per block. SHAKE-256 itself is the `sha3` 0.10.9 / `keccak`
 * 0.1.6 crate dependency (Cargo.lock), restated here from FIPS 202:
 * Keccak-f[1600], rate 136 bytes, domain padding 0x1F ... 0x80.
 */
#include <stdlib.h>
#include <string.h>

#include "via_oracle.h"

#define RATE 136

static const uint64_t KECCAK_RC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL,
    0x8000000080008000ULL, 0x000000000000808bULL, 0x0000000080000001ULL,
    0x8000000080008081ULL, 0x8000000000008009ULL, 0x000000000000008aULL,
    0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL,
    0x8000000000008003ULL, 0x8000000000008002ULL, 0x8000000000000080ULL,
    0x000000000000800aULL, 0x800000008000000aULL, 0x8000000080008081ULL,
    0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

static inline uint64_t rotl64(uint64_t x, unsigned r) {
    return r ? (x << r) | (x >> (64 - r)) : x;
}

/* Keccak-f[1600] per FIPS 202 section 3.3: theta, rho, pi, chi, iota over the
 * 5x5 lane state indexed A[x + 5y]. Rotation offsets are derived by the rho
 * walk (x,y) -> (y, 2x+3y) with offsets (t+1)(t+2)/2. */
static void keccak_f1600(uint64_t A[25]) {
    static int init = 0;
    static unsigned rho[25];
    static int piidx[25];
    if (!init) {
        memset(rho, 0, sizeof rho);
        int x = 1, y = 0;
        for (int t = 0; t < 24; t++) {
            rho[x + 5 * y] = (unsigned)(((t + 1) * (t + 2) / 2) % 64);
            int nx = y, ny = (2 * x + 3 * y) % 5;
            x = nx;
            y = ny;
        }
        /* pi: B[y, 2x+3y] = A[x, y] */
        for (int xx = 0; xx < 5; xx++)
            for (int yy = 0; yy < 5; yy++) piidx[xx + 5 * yy] = yy + 5 * ((2 * xx + 3 * yy) % 5);
        init = 1;
    }
    for (int round = 0; round < 24; round++) {
        uint64_t C[5], D[5], B[25];
        for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
        for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ rotl64(C[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) A[i] ^= D[i % 5];
        for (int i = 0; i < 25; i++) B[piidx[i]] = rotl64(A[i], rho[i]);
        for (int y = 0; y < 5; y++)
            for (int x = 0; x < 5; x++)
                A[x + 5 * y] = B[x + 5 * y] ^ ((~B[(x + 1) % 5 + 5 * y]) & B[(x + 2) % 5 + 5 * y]);
        A[0] ^= KECCAK_RC[round];
    }
}

/* One SHAKE-256 output block of RATE bytes for a message shorter than RATE. */
static void shake256_one_block(const uint8_t *msg, size_t len, uint8_t out[RATE]) {
    uint64_t A[25];
    uint8_t blk[RATE];
    memset(A, 0, sizeof A);
    memset(blk, 0, sizeof blk);
    memcpy(blk, msg, len);
    blk[len] ^= 0x1F;
    blk[RATE - 1] ^= 0x80;
    for (int i = 0; i < RATE / 8; i++) {
        uint64_t w = 0;
        for (int b = 0; b < 8; b++) w |= (uint64_t)blk[8 * i + b] << (8 * b);
        A[i] ^= w;
    }
    keccak_f1600(A);
    for (int i = 0; i < RATE / 8; i++)
        for (int b = 0; b < 8; b++) out[8 * i + b] = (uint8_t)(A[i] >> (8 * b));
}

struct vo_prg {
    uint8_t seed[64];
    size_t seed_len;
    uint64_t counter;
    uint8_t buf[RATE];
    size_t pos;
};

/* Shake256Prg::new (prg.rs:101-117): seeds up to 64 bytes. */
vo_prg *vo_prg_new(const uint8_t *seed, size_t len) {
    if (len > 64) return NULL;
    vo_prg *g = (vo_prg *)calloc(1, sizeof *g);
    memcpy(g->seed, seed, len);
    g->seed_len = len;
    g->counter = 0;
    g->pos = RATE;
    return g;
}

void vo_prg_free(vo_prg *g) {
    if (g) {
        memset(g, 0, sizeof *g);
        free(g);
    }
}

/* Shake256Prg::refill (prg.rs:124-134). */
static void prg_refill(vo_prg *g) {
    uint8_t msg[72];
    memcpy(msg, g->seed, g->seed_len);
    for (int b = 0; b < 8; b++) msg[g->seed_len + b] = (uint8_t)(g->counter >> (8 * b));
    shake256_one_block(msg, g->seed_len + 8, g->buf);
    g->counter++;
    g->pos = 0;
}

/* Shake256Prg::fill_bytes (prg.rs:143-158). */
void vo_prg_fill(vo_prg *g, uint8_t *out, size_t len) {
    size_t written = 0;
    while (written < len) {
        if (g->pos >= RATE) prg_refill(g);
        size_t take = RATE - g->pos;
        if (take > len - written) take = len - written;
        memcpy(out + written, g->buf + g->pos, take);
        g->pos += take;
        written += take;
    }
}

/* Shake256Prg::uniform_below (prg.rs:184-207) and Zq::random
 * (zq/element.rs:152-175): bitlen(bound) mask, ceil(bits/8) bytes per attempt,
 * the high buffer bytes stay zero across rejections. */
uint64_t vo_prg_uniform_below(vo_prg *g, uint64_t bound) {
    if (bound <= 1) return 0;
    unsigned bits = 64 - (unsigned)__builtin_clzll(bound);
    size_t nbytes = (bits + 7) / 8;
    uint64_t mask = bits == 64 ? ~0ULL : ((1ULL << bits) - 1);
    uint8_t buf[8] = {0};
    for (;;) {
        vo_prg_fill(g, buf, nbytes);
        uint64_t v = 0;
        for (int b = 0; b < 8; b++) v |= (uint64_t)buf[b] << (8 * b);
        v &= mask;
        if (v < bound) return v;
    }
}

/* Distribution::sample_into (sampling/distribution.rs): ternary maps
 * {0,1,2} -> {0,1,-1}; bounded uniform draws uniform_below(2b+1) - b. The
 * Gaussian arm is not restated (no e2e test of the reference uses it). */
void vo_sample(vo_prg *g, vo_dist d, int64_t *out, size_t len) {
    if (d.kind == 0) {
        for (size_t i = 0; i < len; i++) {
            uint64_t u = vo_prg_uniform_below(g, 3);
            out[i] = u == 0 ? 0 : (u == 1 ? 1 : -1);
        }
    } else {
        uint64_t range = 2ULL * d.bound + 1;
        for (size_t i = 0; i < len; i++)
            out[i] = (int64_t)vo_prg_uniform_below(g, range) - (int64_t)d.bound;
    }
}
