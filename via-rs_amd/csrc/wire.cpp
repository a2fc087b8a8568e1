// wire.cpp — wire formats (via-protocol wire/format.rs) and PRG mask
// regeneration entry points.
//
// Uncompressed (format.rs:93-132): [0x00][count u64 LE][count x u64 LE].
// PrgCompressed (format.rs:135-190): [0x01][seed, 32 bytes][count u64 LE]
// [count body coefficients u64 LE]; the masks are regenerated from the seed
// (regenerate_masks, format.rs:143-160), here on the GPU (prg.hip).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>

#include "host_internal.h"
#include "keccak.h"

using namespace via;

#define HIP_TRY(x)                                                 \
    do {                                                           \
        if ((x) != hipSuccess) {                                   \
            fprintf(stderr, "via_gpu: %s failed\n", #x);           \
            return VIA_E_DEVICE;                                   \
        }                                                          \
    } while (0)

static const uint8_t kTagUncompressed = 0x00, kTagPrg = 0x01;
static const size_t kPrgSeedLen = 32;  // PrgCompressed<32>, the reference's default

static void put_u64(uint8_t *p, uint64_t v) {
    for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i));
}
static uint64_t get_u64(const uint8_t *p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}
static via_status too_short(size_t needed, size_t got, via_err_detail *err) {
    if (err) {
        err->expected = needed;
        err->got = got;
    }
    return VIA_E_BUFFER_SHORT;
}

size_t via_wire_uncompressed_size(size_t count) { return 1 + 8 + 8 * count; }

via_status via_wire_encode_uncompressed(const uint64_t *coeffs, size_t count, uint8_t *out, size_t out_len) {
    if ((count && !coeffs) || !out) return VIA_E_INVALID;
    if (out_len < via_wire_uncompressed_size(count)) return VIA_E_INVALID;
    out[0] = kTagUncompressed;
    put_u64(out + 1, (uint64_t)count);
    for (size_t i = 0; i < count; i++) put_u64(out + 9 + 8 * i, coeffs[i]);
    return VIA_OK;
}

// The body of a packet after its header: count little-endian words at buf +
// off, BufferTooShort { needed, got } when they do not fit (decode_coeffs_u64).
static via_status decode_words(const uint8_t *buf, size_t len, size_t off, uint64_t count, uint64_t *out,
                               size_t out_cap, size_t *count_out, via_err_detail *err) {
    const uint64_t avail = (len - off) / 8;
    if (count > avail) {
        const uint64_t needed = count > (UINT64_MAX - off) / 8 ? UINT64_MAX : off + 8 * count;
        return too_short((size_t)needed, len, err);
    }
    if (count_out) *count_out = (size_t)count;
    if (!out) return VIA_OK;  // size query
    if (out_cap < count) return VIA_E_INVALID;
    for (uint64_t i = 0; i < count; i++) out[i] = get_u64(buf + off + 8 * i);
    return VIA_OK;
}

via_status via_wire_decode_uncompressed(const uint8_t *buf, size_t len, uint64_t *out, size_t out_cap,
                                        size_t *count, via_err_detail *err) {
    if (!buf && len) return VIA_E_INVALID;
    if (len == 0) return too_short(1, 0, err);
    if (buf[0] != kTagUncompressed) {
        if (err) err->got = buf[0];
        return VIA_E_FORMAT_TAG;
    }
    if (len < 9) return too_short(9, len, err);
    return decode_words(buf, len, 9, get_u64(buf + 1), out, out_cap, count, err);
}

via_status via_wire_decode_prg(const uint8_t *buf, size_t len, uint8_t *seed_out, uint64_t *out, size_t out_cap,
                               size_t *count, via_err_detail *err) {
    if (!buf && len) return VIA_E_INVALID;
    if (len == 0) return too_short(1, 0, err);
    if (buf[0] != kTagPrg) {
        if (err) err->got = buf[0];
        return VIA_E_FORMAT_TAG;
    }
    const size_t header = 1 + kPrgSeedLen + 8;
    if (len < header) return too_short(header, len, err);
    if (seed_out) memcpy(seed_out, buf + 1, kPrgSeedLen);
    return decode_words(buf, len, header, get_u64(buf + 1 + kPrgSeedLen), out, out_cap, count, err);
}

// ---- PRG draws on the device -----------------------------------------------------

static int bit_length(uint64_t q) { return q ? 64 - __builtin_clzll(q) : 0; }

// Plan of count draws at moduli q[0..nmod): attempts sized at the mean plus
// ten standard deviations of the rejections (and 64), so that a shortfall
// (reported by the kernel, then VIA_E_DEVICE) does not happen in practice.
static via_status prg_plan(const uint64_t *q, int nmod, size_t count, PrgPlan *pl) {
    memset(pl, 0, sizeof *pl);
    double pmax = 0.0;
    for (int i = 0; i < nmod; i++) {
        if (q[i] < 2) return VIA_E_INVALID;
        const int bits = bit_length(q[i]);
        const int nb = (bits + 7) / 8;
        if (i && nb != pl->nbytes) return VIA_E_INVALID;  // attempts of one width
        pl->nbytes = nb;
        pl->q[i] = q[i];
        pl->mask[i] = bits == 64 ? ~0ULL : ((1ULL << bits) - 1);
        const double span = std::ldexp(1.0, bits);
        pmax = std::max(pmax, (span - (double)q[i]) / span);
    }
    if (nmod == 1) {
        pl->q[1] = pl->q[0];
        pl->mask[1] = pl->mask[0];
    }
    pl->nmod = nmod;
    pl->count = count;
    const double mean = (double)count / (1.0 - pmax);
    const double sd = std::sqrt((double)count * pmax) / (1.0 - pmax);
    pl->attempts = (size_t)std::ceil(mean + 10.0 * sd) + 64;
    pl->blocks = (pl->attempts * pl->nbytes + kPrgBlock - 1) / kPrgBlock;
    return VIA_OK;
}

static via_status run_prg(via_ctx *c, const PrgSeeds &seeds, int streams, const PrgPlan &pl, uint64_t *d_out) {
    uint64_t *stream = nullptr;
    int *flag = nullptr;
    if (hipMalloc(&stream, sizeof(uint64_t) * 17 * pl.blocks * streams) != hipSuccess) return VIA_E_NOMEM;
    if (hipMalloc(&flag, sizeof(int)) != hipSuccess) {
        (void)hipFree(stream);
        return VIA_E_NOMEM;
    }
    int h_flag = 0;
    via_status s = VIA_OK;
    if (hipMemcpyAsync(flag, &h_flag, sizeof(int), hipMemcpyHostToDevice, c->stream) != hipSuccess) s = VIA_E_DEVICE;
    if (!s) launch_prg_draws(seeds, streams, pl, stream, d_out, flag, c->stream);
    if (!s && hipMemcpyAsync(&h_flag, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        s = VIA_E_DEVICE;
    if (!s && hipStreamSynchronize(c->stream) != hipSuccess) s = VIA_E_DEVICE;
    if (!s && hipGetLastError() != hipSuccess) s = VIA_E_DEVICE;
    (void)hipFree(stream);
    (void)hipFree(flag);
    if (!s && h_flag) s = VIA_E_DEVICE;
    return s;
}

via_status via_prg_regenerate(via_ctx *c, const uint8_t *seed, size_t seed_len, uint64_t q, size_t count,
                              uint64_t *out) {
    if (!c || (seed_len && !seed) || (count && !out) || q == 0) return VIA_E_INVALID;
    if (seed_len > (size_t)kPrgSeedCap) return VIA_E_PRG_SEED;
    if (!count) return VIA_OK;
    if (q == 1) {  // uniform_below(1) is 0 and draws no bytes (prg.rs:184-186)
        memset(out, 0, sizeof(uint64_t) * count);
        return VIA_OK;
    }
    HIP_TRY(hipSetDevice(c->device));
    PrgPlan pl;
    via_status s = prg_plan(&q, 1, count, &pl);
    if (s) return s;
    PrgSeeds seeds;
    memset(&seeds, 0, sizeof seeds);
    if (seed_len) memcpy(seeds.bytes[0], seed, seed_len);
    seeds.len[0] = (int)seed_len;
    uint64_t *d = nullptr;
    if (hipMalloc(&d, sizeof(uint64_t) * count) != hipSuccess) return VIA_E_NOMEM;
    s = run_prg(c, seeds, 1, pl, d);
    if (!s && hipMemcpy(out, d, sizeof(uint64_t) * count, hipMemcpyDeviceToHost) != hipSuccess) s = VIA_E_DEVICE;
    (void)hipFree(d);
    return s;
}

via_status via_query_from_prg(via_ctx *c, const via_params *pp, const uint8_t *seed, size_t seed_len,
                              const uint64_t *bodies, size_t n_lwe, uint64_t *d_lwes) {
    if (!c || !pp || (seed_len && !seed) || !bodies || !d_lwes || n_lwe == 0) return VIA_E_INVALID;
    if (seed_len > (size_t)kPrgSeedCap) return VIA_E_PRG_SEED;
    if (pp->q1_nl != 2) return VIA_E_INVALID;
    HIP_TRY(hipSetDevice(c->device));
    const size_t n = pp->n1, nl = 2, lw = nl * (n + 1);
    PrgPlan pl;
    via_status s = prg_plan(pp->q1, 2, n_lwe * n * nl, &pl);
    if (s) return s;
    pl.lwe = 1;
    pl.n = (int)n;
    pl.row_stride = lw;
    pl.limb_stride = n + 1;
    PrgSeeds seeds;
    memset(&seeds, 0, sizeof seeds);
    if (seed_len) memcpy(seeds.bytes[0], seed, seed_len);
    seeds.len[0] = (int)seed_len;
    // bodies [lwe][limb] go to coefficient n of each limb row
    HIP_TRY(hipMemcpy2DAsync(d_lwes + n, sizeof(uint64_t) * (n + 1), bodies, sizeof(uint64_t), sizeof(uint64_t),
                             n_lwe * nl, hipMemcpyHostToDevice, c->stream));
    return run_prg(c, seeds, 1, pl, d_lwes);
}
