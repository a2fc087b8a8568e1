// client_keys.cpp — client key generation and query encryption (via-client
// Client::setup, client.rs:89-173, and Client::query, client.rs:183-207 ->
// build_compressed_query, query.rs:75-130).
//
// Every random value comes from one Shake256Prg stream (sampling/prg.rs) in
// the reference's order, so the keys and queries are the reference's own for
// the same seed. The stream's blocks (one Keccak-f[1600] each) are computed
// on the GPU in windows of 32768 blocks and parsed on the host: which stream
// bytes a draw reads depends on every rejection before it, a chain through
// the whole key. The ring products of the encryptions (mask * secret for
// every RLWE sample, about 8 x 10^5 coefficients of products at the paper
// sizes, and S1^2) run on the GPU as batched NTT products per ring; the host
// adds errors and gadget-scaled messages.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <thread>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

#include "host_internal.h"
#include "keccak.h"

using namespace via;

typedef unsigned __int128 u128;

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "stream bytes are read as little-endian words");

struct via_prg {
    uint8_t seed[kPrgSeedCap];
    int seed_len = 0;
    uint64_t pos = 0;  // stream bytes consumed
    // stream bytes [cache_base, cache_base + cache.size()) computed by an
    // earlier call and not yet consumed (whole blocks)
    std::vector<uint8_t> cache;
    uint64_t cache_base = 0;
};

via_status via_prg_create(const uint8_t *seed, size_t seed_len, via_prg **out) {
    if ((seed_len && !seed) || !out) return VIA_E_INVALID;
    if (seed_len > (size_t)kPrgSeedCap) return VIA_E_PRG_SEED;
    via_prg *g = new (std::nothrow) via_prg();
    if (!g) return VIA_E_NOMEM;
    memset(g->seed, 0, sizeof g->seed);
    if (seed_len) memcpy(g->seed, seed, seed_len);
    g->seed_len = (int)seed_len;
    *out = g;
    return VIA_OK;
}

void via_prg_destroy(via_prg *g) {
    if (!g) return;
    memset(g->seed, 0, sizeof g->seed);
    if (!g->cache.empty()) memset(g->cache.data(), 0, g->cache.size());
    delete g;
}

uint64_t via_prg_position(const via_prg *g) { return g ? g->pos : 0; }

namespace {

constexpr size_t kWindowBlocks = 1 << 15;  // at most 4.25 MiB of stream per GPU window

// The stream from the PRG's position: windows of whole blocks computed on
// the device, parsed here; the bytes left over stay with the PRG for its
// next call. The PRG advances only when the caller commits.
struct Reader {
    via_ctx *c;
    via_prg *g;
    PrgSeeds seeds;
    std::vector<uint8_t> buf;  // stream bytes [base, base + buf.size())
    uint64_t base, pos;
    size_t next_blocks;        // size of the next window
    uint64_t *d = nullptr;
    via_status err = VIA_OK;

    // expect: about how many stream bytes the call will read (sizes the
    // first window)
    Reader(via_ctx *ctx, via_prg *prg, size_t expect) : c(ctx), g(prg) {
        memset(&seeds, 0, sizeof seeds);
        memcpy(seeds.bytes[0], g->seed, (size_t)g->seed_len);
        seeds.len[0] = g->seed_len;
        pos = g->pos;
        buf.swap(g->cache);
        base = g->cache_base;
        if (pos < base || pos > base + buf.size()) {
            buf.clear();
            base = pos / kPrgBlock * kPrgBlock;
        }
        next_blocks = std::min(kWindowBlocks, expect / kPrgBlock + 64);
    }
    ~Reader() {
        if (d) (void)hipFree(d);
        const uint64_t keep = pos / kPrgBlock * kPrgBlock;  // hand the unread whole blocks back
        if (keep > base && keep - base <= buf.size()) {
            buf.erase(buf.begin(), buf.begin() + (ptrdiff_t)(keep - base));
            base = keep;
        }
        g->cache.swap(buf);
        g->cache_base = base;
    }
    void commit() { g->pos = pos; }

    // bytes [pos, pos + k) become readable
    bool more(size_t k) {
        while (pos + k > base + buf.size()) {
            if (err) return false;
            const uint64_t keep = pos / kPrgBlock * kPrgBlock;  // drop consumed whole blocks
            if (keep > base) {
                buf.erase(buf.begin(), buf.begin() + (ptrdiff_t)(keep - base));
                base = keep;
            }
            if (!d && hipMalloc(&d, sizeof(uint64_t) * 17 * kWindowBlocks) != hipSuccess) {
                d = nullptr;
                err = VIA_E_NOMEM;
                return false;
            }
            const size_t nb = next_blocks;
            next_blocks = kWindowBlocks;
            const uint64_t first = (base + buf.size()) / kPrgBlock;
            const size_t old = buf.size();
            buf.resize(old + nb * kPrgBlock);
            launch_prg_window(seeds, first, nb, d, c->stream);
            // 17 little-endian lanes are the block's 136 bytes in stream order
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(buf.data() + old, d, nb * kPrgBlock, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                buf.resize(old);
                err = VIA_E_DEVICE;
                return false;
            }
        }
        return true;
    }

    // Shake256Prg::uniform_below (prg.rs:184-207): attempts of
    // ceil(bits/8) little-endian bytes masked to the bit length of bound,
    // rejected when >= bound; bound <= 1 draws nothing.
    uint64_t uniform_below(uint64_t bound) {
        if (bound <= 1) return 0;
        const int bits = 64 - __builtin_clzll(bound);
        const int nbytes = (bits + 7) / 8;
        const uint64_t mask = bits == 64 ? ~0ULL : ((1ULL << bits) - 1);
        for (;;) {
            uint64_t v;
            if (pos + 8 <= base + buf.size()) {
                // eight bytes at once (little-endian host): the bit mask
                // clears every byte past the attempt's nbytes
                memcpy(&v, buf.data() + (pos - base), 8);
            } else {
                if (pos + nbytes > base + buf.size() && !more((size_t)nbytes)) return 0;
                const uint8_t *p = buf.data() + (pos - base);
                v = 0;
                for (int b = 0; b < nbytes; b++) v |= (uint64_t)p[b] << (8 * b);
            }
            pos += (uint64_t)nbytes;
            v &= mask;
            if (v < bound) return v;
        }
    }

    // Runs of attempts of nbytes each until `total` are accepted: take(v)
    // masks and tests attempt v (the eight stream bytes at its position;
    // the bit mask drops the bytes past nbytes), stores it in its slot
    // whether or not it is accepted (a rejected one is overwritten by the
    // next attempt) and returns 1 on acceptance. No branch depends on the
    // stream: rejections are frequent (half the attempts for q0 just above
    // 2^37) and would mispredict.
    template <class F>
    void run(int nbytes, size_t total, F take) {
        size_t done = 0;
        while (done < total) {
            if (pos + 8 > base + buf.size() && !more(8)) return;
            const uint64_t end = base + buf.size();
            const size_t avail = (size_t)((end - pos - 8) / (uint64_t)nbytes) + 1;
            const uint8_t *p = buf.data() + (pos - base);
            size_t a = 0;
            for (; a < avail && done < total; a++) {
                uint64_t v;
                memcpy(&v, p + a * (size_t)nbytes, 8);
                done += take(v, done);
            }
            pos += (uint64_t)a * (uint64_t)nbytes;
        }
    }

    // Distribution::sample_into (sampling/distribution.rs:88-115): ternary
    // {0, 1, 2} -> {0, 1, -1}; bounded uniform uniform_below(2b + 1) - b.
    void sample(const via_dist &dist, int64_t *out, size_t n) {
        if (dist.kind == 0) {
            static const int64_t tern[4] = {0, 1, -1, 0};
            run(1, n, [&](uint64_t v, size_t d) -> size_t {
                v &= 3;
                out[d] = tern[v];
                return v < 3;
            });
            return;
        }
        const uint64_t range = 2ULL * dist.bound + 1;
        if (range <= 1) {  // uniform_below(1) draws nothing
            for (size_t i = 0; i < n; i++) out[i] = 0;
            return;
        }
        const int bits = 64 - __builtin_clzll(range);
        const uint64_t mask = bits == 64 ? ~0ULL : ((1ULL << bits) - 1);
        const int64_t shift = (int64_t)dist.bound;
        run((bits + 7) / 8, n, [&](uint64_t v, size_t d) -> size_t {
            v &= mask;
            out[d] = (int64_t)v - shift;
            return v < range;
        });
    }

    // RingPoly::random_uniform (element.rs:235, rns_element.rs:201-212): per
    // coefficient, one draw per limb, limb 0 first; out [limb][n]. Draw d is
    // coefficient d / nl of limb d % nl.
    void poly_random(const uint64_t *q, int nl, size_t n, uint64_t *out) {
        const int bits0 = 64 - __builtin_clzll(q[0]);
        const int bits1 = nl == 2 ? 64 - __builtin_clzll(q[1]) : bits0;
        if (nl > 2 || q[0] < 2 || (nl == 2 && q[1] < 2) || (bits0 + 7) / 8 != (bits1 + 7) / 8) {
            for (size_t i = 0; i < n; i++)
                for (int l = 0; l < nl; l++) out[(size_t)l * n + i] = uniform_below(q[l]);
            return;
        }
        const uint64_t m0 = bits0 == 64 ? ~0ULL : ((1ULL << bits0) - 1);
        const uint64_t m1 = bits1 == 64 ? ~0ULL : ((1ULL << bits1) - 1);
        if (nl == 1) {
            run((bits0 + 7) / 8, n, [&](uint64_t v, size_t d) -> size_t {
                v &= m0;
                out[d] = v;
                return v < q[0];
            });
            return;
        }
        const uint64_t qa = q[0], qb = q[1];
        run((bits0 + 7) / 8, 2 * n, [&](uint64_t v, size_t d) -> size_t {
            const size_t l = d & 1;
            v &= l ? m1 : m0;
            out[l * n + (d >> 1)] = v;
            return v < (l ? qb : qa);
        });
    }
};

uint64_t mod_i64(int64_t v, uint64_t q) {
    const int64_t r = v % (int64_t)q;
    return (uint64_t)(r < 0 ? r + (int64_t)q : r);
}

uint64_t addm(uint64_t a, uint64_t b, uint64_t q) {
    const uint64_t s = a + b;  // a, b < q < 2^63
    return s >= q ? s - q : s;
}

// round(Q / B^(k+1)) (gadget.rs: the gadget vector of a ring of modulus Q)
u128 gadget_g(u128 Q, uint64_t base, uint32_t k) {
    u128 div = 1;
    for (uint32_t i = 0; i <= k; i++) div *= base;
    return (Q + div / 2) / div;
}

// One ring of the key material: degree n, nl limbs, the NTT products of
// every mask with its secret gathered for one batched launch; each secret is
// stored once and products name it by index.
struct Group {
    uint32_t n = 0;
    int nl = 0;
    uint64_t q[2] = {0, 0};
    std::vector<uint64_t> a;       // masks [product][limb][n]
    std::vector<uint64_t> keys;    // secrets [key][limb][n]
    std::vector<uint32_t> kidx;    // product -> secret
    std::unordered_map<const uint64_t *, uint32_t> key_id;
    size_t count = 0;
    std::vector<uint64_t> c;       // products [product][limb][n]

    size_t add(const uint64_t *mask, const uint64_t *key) {
        const size_t w = (size_t)nl * n;
        a.insert(a.end(), mask, mask + w);
        auto it = key_id.find(key);
        if (it == key_id.end()) {
            it = key_id.emplace(key, (uint32_t)key_id.size()).first;
            keys.insert(keys.end(), key, key + w);
        }
        kidx.push_back(it->second);
        return count++;
    }
    const uint64_t *prod(size_t i) const { return c.data() + i * (size_t)nl * n; }
};

via_status run_group(via_ctx *c, Group &gr) {
    gr.c.assign(gr.a.size(), 0);
    if (!gr.count) return VIA_OK;
    RingDev r;
    via_status s = ctx_ring(c, gr.n, gr.q, gr.nl, &r);
    if (s) return s;
    const size_t words = gr.a.size(), kw = gr.keys.size();
    uint64_t *d = nullptr;
    if (hipMalloc(&d, sizeof(uint64_t) * (2 * words + kw) + sizeof(uint32_t) * gr.count) != hipSuccess)
        return VIA_E_NOMEM;
    uint32_t *dk = (uint32_t *)(d + 2 * words + kw);
    s = VIA_OK;
    if (hipMemcpyAsync(d, gr.a.data(), 8 * words, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(d + 2 * words, gr.keys.data(), 8 * kw, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(dk, gr.kidx.data(), 4 * gr.count, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        s = VIA_E_DEVICE;
    if (!s) launch_polymul_u64(r, d, d + 2 * words, d + words, gr.count * (size_t)gr.nl, 1, c->stream, dk);
    if (!s && (hipGetLastError() != hipSuccess ||
               hipMemcpyAsync(gr.c.data(), d + words, 8 * words, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
               hipStreamSynchronize(c->stream) != hipSuccess))
        s = VIA_E_DEVICE;
    (void)hipFree(d);
    return s;
}

// An RLev encryption (SecretKey::encrypt_rlev, rlev.rs:59-65): level k
// encrypts msg * g_k (g_k = round(Q / B^(k+1))) with SecretKey::encrypt
// (rlwe.rs:87-106): mask uniform, then the error, body = mask * s + e + m.
// The masks land in out at once; the bodies wait for the products.
struct Rlev {
    Group *gr;
    uint64_t *out;  // [k][part][limb][n]
    uint32_t L;
    uint64_t base;
    const uint64_t *msg;  // [limb][n], read when the bodies are assembled
    std::vector<int64_t> e;  // [k][n]
    size_t first;            // product index of level 0
};

void rlev_sample(Reader &rd, Group &gr, const uint64_t *key, const uint64_t *msg, uint64_t base, uint32_t L,
                 const via_dist &err, uint64_t *out, std::vector<Rlev> &todo) {
    const size_t n = gr.n, w = (size_t)gr.nl * n;
    Rlev rl;
    rl.gr = &gr;
    rl.out = out;
    rl.L = L;
    rl.base = base;
    rl.msg = msg;
    rl.e.resize((size_t)L * n);
    rl.first = gr.count;
    for (uint32_t k = 0; k < L; k++) {
        uint64_t *ct = out + (size_t)k * 2 * w;
        rd.poly_random(gr.q, gr.nl, n, ct);
        rd.sample(err, rl.e.data() + (size_t)k * n, n);
        gr.add(ct, key);
    }
    todo.push_back(std::move(rl));
}

// x * g mod q for a fixed g < q < 2^63 and x < q (Shoup: gs = floor(g 2^64 / q))
inline uint64_t mulmod_shoup(uint64_t x, uint64_t g, uint64_t gs, uint64_t q) {
    const uint64_t t = (uint64_t)(((u128)x * gs) >> 64);
    const uint64_t r = x * g - t * q;  // exact result in [0, 2q), computed mod 2^64
    return r >= q ? r - q : r;
}

// e mod q for the small errors (|e| < q), the general reduction otherwise
inline uint64_t err_mod(int64_t e, uint64_t q) {
    if (e >= 0 && (uint64_t)e < q) return (uint64_t)e;
    if (e < 0 && (uint64_t)(-e) < q) return q - (uint64_t)(-e);
    return mod_i64(e, q);
}

void rlev_bodies(const Rlev &rl) {
    const Group &gr = *rl.gr;
    const size_t n = gr.n, w = (size_t)gr.nl * n;
    const u128 Q = gr.nl == 2 ? (u128)gr.q[0] * gr.q[1] : (u128)gr.q[0];
    for (uint32_t k = 0; k < rl.L; k++) {
        const u128 gk = gadget_g(Q, rl.base, k);
        const uint64_t *p = gr.prod(rl.first + k);
        uint64_t *body = rl.out + (size_t)k * 2 * w + w;
        for (int l = 0; l < gr.nl; l++) {
            const uint64_t q = gr.q[l], gl = (uint64_t)(gk % q);
            const uint64_t gs = (uint64_t)(((u128)gl << 64) / q);
            const uint64_t *msg = rl.msg + (size_t)l * n;
            const int64_t *e = rl.e.data() + (size_t)k * n;
            for (size_t i = 0; i < n; i++) {
                uint64_t v = addm(p[(size_t)l * n + i], err_mod(e[i], q), q);
                if (msg[i]) v = addm(v, mulmod_shoup(msg[i], gl, gs, q), q);
                body[(size_t)l * n + i] = v;
            }
        }
    }
}

// project_at / embed_at (ring/reshape.rs:58-116), per limb: degree ns -> nd.
void project_at(int nl, size_t ns, size_t nd, const uint64_t *src, size_t slot, uint64_t *dst) {
    const size_t d = ns / nd;
    for (int l = 0; l < nl; l++)
        for (size_t i = 0; i < nd; i++) dst[(size_t)l * nd + i] = src[(size_t)l * ns + d * i + slot];
}

void embed_at(int nl, size_t ns, size_t nd, const uint64_t *src, size_t slot, uint64_t *dst) {
    const size_t d = nd / ns;
    memset(dst, 0, sizeof(uint64_t) * nl * nd);
    for (int l = 0; l < nl; l++)
        for (size_t i = 0; i < ns; i++) dst[(size_t)l * nd + d * i + slot] = src[(size_t)l * ns + i];
}

bool ntt_ring(uint64_t q, size_t n) { return q > 2 && q < kMaxModulus && (q - 1) % (2 * n) == 0; }

// B^L stays within 128 bits (the gadget vector round(Q / B^(k+1)), k < L)
bool gadget_ok(uint64_t base, uint32_t L) {
    if (base < 2 || !L) return false;
    u128 v = 1;
    for (uint32_t k = 0; k < L; k++) {
        if (v > (~(u128)0) / base) return false;
        v *= base;
    }
    return true;
}

via_status check_client(const via_params *pp) {
    if (!pp) return VIA_E_INVALID;
    if (pp->n1 != (uint32_t)kSlots || pp->q1_nl != 2 || !pp->n2 || pp->n1 % pp->n2 || !pp->n_rec ||
        pp->n1 % pp->n_rec)
        return VIA_E_INVALID;
    if (!ntt_ring(pp->q1[0], pp->n1) || !ntt_ring(pp->q1[1], pp->n1) || !ntt_ring(pp->q3, pp->n2)) return VIA_E_INVALID;
    if (!gadget_ok(pp->ck_base, pp->l_ck) || !gadget_ok(pp->b_rsk, pp->l_rsk) || !gadget_ok(pp->b1, pp->l_query) ||
        !gadget_ok(pp->b2, pp->l_query))
        return VIA_E_INVALID;
    return VIA_OK;
}

bool dist_ok(const via_dist &d) { return d.kind == 0 || d.kind == 1; }

}  // namespace

via_status via_prg_fill(via_ctx *c, via_prg *prg, uint8_t *out, size_t len) {
    if (!c || !prg || (len && !out)) return VIA_E_INVALID;
    if (!len) return VIA_OK;
    if (hipSetDevice(c->device) != hipSuccess) return VIA_E_DEVICE;
    Reader rd(c, prg, std::min(len, kWindowBlocks * kPrgBlock));
    size_t done = 0;
    while (done < len) {
        const size_t take = std::min(len - done, kWindowBlocks * kPrgBlock / 2);
        if (!rd.more(take)) return rd.err ? rd.err : VIA_E_DEVICE;
        memcpy(out + done, rd.buf.data() + (rd.pos - rd.base), take);
        rd.pos += take;
        done += take;
    }
    rd.commit();
    return VIA_OK;
}

via_status via_prg_uniform_below(via_ctx *c, via_prg *prg, uint64_t bound, uint64_t *out) {
    if (!c || !prg || !out || !bound) return VIA_E_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return VIA_E_DEVICE;
    Reader rd(c, prg, 64);
    const uint64_t v = rd.uniform_below(bound);
    if (rd.err) return rd.err;
    rd.commit();
    *out = v;
    return VIA_OK;
}

via_status via_client_key_words(const via_params *pp, size_t *cascade, size_t *conv, size_t *rsk) {
    if (!pp || !pp->n1 || !pp->n2 || !pp->q1_nl) return VIA_E_INVALID;
    const size_t n = pp->n1, nl = pp->q1_nl, logn = (size_t)__builtin_ctz(pp->n1);
    if (cascade) *cascade = logn * 2 * n * pp->l_ck * 2 * nl;
    if (conv) *conv = (size_t)pp->l_ck * 2 * nl * n;
    if (rsk) *rsk = (size_t)(pp->n1 / pp->n2) * pp->l_rsk * 2 * pp->n2;
    return VIA_OK;
}

// Client::setup of one client on context c (its stream), arguments checked
// by the caller.
static via_status client_setup_one(via_ctx *c, const via_params *pp, via_dist key1, via_dist key2, via_dist err,
                                   via_prg *prg, uint64_t *sk1, uint64_t *sk2, uint64_t *cascade_key,
                                   uint64_t *conv_key, uint64_t *rsk) {
    via_status s = VIA_OK;
    if (hipSetDevice(c->device) != hipSuccess) return VIA_E_DEVICE;
    const size_t n1 = pp->n1, n2 = pp->n2, nl = 2, L = pp->l_ck;
    const uint64_t q3 = pp->q3;
    // stream use: every mask coefficient takes a 5-byte attempt per limb
    // (about 2 for q0 just above 2^37) and every error a byte or so
    size_t words = 0;
    via_client_key_words(pp, &words, nullptr, nullptr);
    Reader rd(c, prg, words / 2 * 8);
    // 1-2. secret keys (SecretKey::keygen: centred samples lifted to each limb)
    std::vector<int64_t> c1(n1), c2(n2);
    rd.sample(key1, c1.data(), n1);
    for (size_t l = 0; l < nl; l++)
        for (size_t i = 0; i < n1; i++) sk1[l * n1 + i] = mod_i64(c1[i], pp->q1[l]);
    rd.sample(key2, c2.data(), n2);
    for (size_t i = 0; i < n2; i++) sk2[i] = mod_i64(c2[i], q3);
    // 3. cascade key (cascade.rs:90-110 via conv.rs:161-190): steps in
    // ascending output degree; key idx = m*group + j encrypts, under S
    // projected at slot j (degree nout), the slot-(m*group + j) projection of
    // S (degree nout/2) embedded at slot 0
    const int logn = __builtin_ctz((unsigned)n1);
    std::vector<Group> groups(logn + 2);
    std::vector<std::vector<uint64_t>> keys, msgs;  // stable storage of secrets and messages
    std::vector<Rlev> todo;
    keys.reserve(2 * n1 + 4);
    msgs.reserve(2 * n1 + 4);
    size_t off = 0;
    for (int st = 0; st < logn; st++) {
        const size_t nout = (size_t)2 << st, nin = nout / 2, rank_in = n1 / nin, m = rank_in / 2;
        Group &gr = groups[st];
        gr.n = (uint32_t)nout;
        gr.nl = (int)nl;
        gr.q[0] = pp->q1[0];
        gr.q[1] = pp->q1[1];
        gr.a.reserve(rank_in * L * nl * nout);
        const size_t k0 = keys.size();
        for (size_t j = 0; j < m; j++) {  // S projected at slot j (degree nout)
            keys.emplace_back(nl * nout);
            project_at((int)nl, n1, nout, sk1, j, keys.back().data());
        }
        std::vector<uint64_t> src(nl * nin);
        for (size_t idx = 0; idx < rank_in; idx++) {
            const size_t grp = idx / m, j = idx % m;
            project_at((int)nl, n1, nin, sk1, m * grp + j, src.data());
            msgs.emplace_back(nl * nout);
            embed_at((int)nl, nin, nout, src.data(), 0, msgs.back().data());
            rlev_sample(rd, gr, keys[k0 + j].data(), msgs.back().data(), pp->ck_base, (uint32_t)L, err,
                        cascade_key + off, todo);
            off += L * 2 * nl * nout;
        }
    }
    // 4. RLWE -> RGSW conversion key RLev_S(S^2) (convert.rs:46-58); S^2 is
    // one more product of the n1 ring
    Group &g1 = groups[logn];
    g1.n = (uint32_t)n1;
    g1.nl = (int)nl;
    g1.q[0] = pp->q1[0];
    g1.q[1] = pp->q1[1];
    std::vector<uint64_t> s1sq(nl * n1, 0);
    rlev_sample(rd, g1, sk1, s1sq.data(), pp->ck_base, (uint32_t)L, err, conv_key, todo);
    const size_t sq_idx = g1.add(sk1, sk1);
    // 5. ring-switch key (rekey S1 -> q3, then gen_rsk, ring_switch.rs:224-246):
    // sample j encrypts under S2 the slot-0 projection of S1 (j = 0) or of
    // -(S1 * X^(n1 - j))
    std::vector<uint64_t> s1q3(n1);
    for (size_t i = 0; i < n1; i++) s1q3[i] = mod_i64(c1[i], q3);
    Group &g3 = groups[logn + 1];
    g3.n = (uint32_t)n2;
    g3.nl = 1;
    g3.q[0] = q3;
    const size_t D = n1 / n2, Lr = pp->l_rsk;
    std::vector<uint64_t> rot(n1);
    for (size_t j = 0; j < D; j++) {
        msgs.emplace_back(n2);
        if (j == 0) {
            project_at(1, n1, n2, s1q3.data(), 0, msgs.back().data());
        } else {
            // S1 * X^(n1 - j): coefficient i moves to i - j, negated when it wraps (i < j);
            // then negated as a whole
            for (size_t i = 0; i < n1; i++) {
                const bool wrapped = i + (n1 - j) >= n1;
                const size_t p = wrapped ? i - j : i + n1 - j;
                const uint64_t v = s1q3[i];
                const uint64_t r = wrapped ? (v ? q3 - v : 0) : v;
                rot[p] = r ? q3 - r : 0;
            }
            project_at(1, n1, n2, rot.data(), 0, msgs.back().data());
        }
        rlev_sample(rd, g3, sk2, msgs.back().data(), pp->b_rsk, (uint32_t)Lr, err, rsk + j * Lr * 2 * n2, todo);
    }
    if (rd.err) return rd.err;
    // the products, one batched launch per ring
    for (Group &gr : groups)
        if ((s = run_group(c, gr))) return s;
    memcpy(s1sq.data(), g1.prod(sq_idx), sizeof(uint64_t) * nl * n1);
    for (const Rlev &rl : todo) rlev_bodies(rl);
    rd.commit();
    return VIA_OK;
}

static via_status check_setup(const via_params *pp, via_dist key1, via_dist key2, via_dist err) {
    if (!pp) return VIA_E_INVALID;
    // client.rs:125-131: the tree dimensions must be powers of two
    if (!pp->num_rows || !pp->num_cols || (pp->num_rows & (pp->num_rows - 1)) ||
        (pp->num_cols & (pp->num_cols - 1)))
        return VIA_E_DIM;
    via_status s = check_client(pp);
    if (s) return s;
    if (!dist_ok(key1) || !dist_ok(key2) || !dist_ok(err)) return VIA_E_INVALID;
    return VIA_OK;
}

via_status via_client_setup(via_ctx *c, const via_params *pp, via_dist key1, via_dist key2, via_dist err,
                            via_prg *prg, uint64_t *sk1, uint64_t *sk2, uint64_t *cascade_key, uint64_t *conv_key,
                            uint64_t *rsk) {
    if (!c || !pp || !prg || !sk1 || !sk2 || !cascade_key || !conv_key || !rsk) return VIA_E_INVALID;
    via_status s = check_setup(pp, key1, key2, err);
    if (s) return s;
    return client_setup_one(c, pp, key1, key2, err, prg, sk1, sk2, cascade_key, conv_key, rsk);
}

// Client::setup of n clients at once: the clients are independent (each its
// own PRG stream), so up to eight host threads each take whole clients in
// turn on a lane of the context (a stream of its own): one client's stream
// parse and body assembly on the host overlap other clients' PRG windows
// and ring products on the GPU.
via_status via_client_setup_many(via_ctx *c, const via_params *pp, uint32_t n, via_dist key1, via_dist key2,
                                 via_dist err, via_prg *const *prgs, uint64_t *const *sk1, uint64_t *const *sk2,
                                 uint64_t *const *cascade_key, uint64_t *const *conv_key, uint64_t *const *rsk,
                                 via_status *status) {
    if (!c || !pp || !prgs || !sk1 || !sk2 || !cascade_key || !conv_key || !rsk || !status) return VIA_E_INVALID;
    for (uint32_t i = 0; i < n; i++)
        if (!prgs[i] || !sk1[i] || !sk2[i] || !cascade_key[i] || !conv_key[i] || !rsk[i]) return VIA_E_INVALID;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t k = 0; k < i; k++)
            if (prgs[i] == prgs[k]) return VIA_E_INVALID;  // one sequential stream per client
    via_status s = check_setup(pp, key1, key2, err);
    if (s) return s;
    const uint32_t nth = std::min<uint32_t>(n, 8);
    for (uint32_t i = 0; i < n; i++) status[i] = VIA_E_DEVICE;  // until a worker sets up client i
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        via_ctx *lane = nullptr;
        if (lane_acquire(c, &lane) != VIA_OK) return;  // the other workers take its clients
        for (uint32_t i; (i = next.fetch_add(1)) < n;)
            status[i] = client_setup_one(lane, pp, key1, key2, err, prgs[i], sk1[i], sk2[i], cascade_key[i],
                                         conv_key[i], rsk[i]);
        lane_release(lane);
    };
    std::vector<std::thread> th;
    for (uint32_t t = 1; t < nth; t++) th.emplace_back(worker);
    if (nth) worker();
    for (auto &x : th) x.join();
    s = VIA_OK;
    for (uint32_t i = 0; i < n; i++)
        if (status[i] && !s) s = status[i];
    return s;
}

via_status via_client_query(via_ctx *c, const via_params *pp, const uint64_t *sk1, size_t index, via_dist err,
                            via_prg *prg, uint64_t *lwes) {
    if (!c || !pp || !sk1 || !prg || !lwes) return VIA_E_INVALID;
    via_status s = check_client(pp);
    if (s) return s;
    if (!dist_ok(err) || !pp->l_query) return VIA_E_INVALID;
    const size_t I = pp->num_rows, J = pp->num_cols, d = pp->n1 / pp->n_rec;
    // the bit decomposition needs powers of two (Client::setup refuses others)
    if (!I || !J || (I & (I - 1)) || (J & (J - 1)) || (d & (d - 1))) return VIA_E_DIM;
    if (index >= d * I * J) return VIA_E_INDEX;  // client.rs:190-195
    if (hipSetDevice(c->device) != hipSuccess) return VIA_E_DEVICE;
    // decompose_index (decompose.rs:25-54): gamma, alpha, beta; DMux bits MSB
    // first with base b1, CMux and CRot bits LSB first with base b2; each bit
    // is L_QUERY raw LWEs of bit * g_i (encrypt_lwe_raw, mlwe_ops.rs:223-248:
    // the masks, one error draw, body = <a, s> + e + m per limb)
    const size_t group = I * J, gamma = index / group, alpha = (index % group) / J, beta = index % J;
    const int nd = __builtin_ctzll(I), nc = __builtin_ctzll(J), nr = __builtin_ctzll(d);
    const size_t n = pp->n1, nl = 2, lw = nl * (n + 1);
    const u128 Q = (u128)pp->q1[0] * pp->q1[1];
    const size_t nlwe = (size_t)(nd + nc + nr) * pp->l_query;
    Reader rd(c, prg, nlwe * n * 16);
    std::vector<uint64_t> masks(nl * n);
    size_t cnt = 0;
    for (int grp = 0; grp < 3; grp++) {
        const int nb = grp == 0 ? nd : (grp == 1 ? nc : nr);
        const uint64_t base = grp == 0 ? pp->b1 : pp->b2;
        for (int i = 0; i < nb; i++) {
            int bit;
            if (grp == 0) bit = (int)((alpha >> (nd - 1 - i)) & 1);
            else if (grp == 1) bit = (int)((beta >> i) & 1);
            else bit = (int)((gamma >> i) & 1);
            for (uint32_t k = 0; k < pp->l_query; k++) {
                const u128 msg = bit ? gadget_g(Q, base, k) % Q : 0;
                rd.poly_random(pp->q1, (int)nl, n, masks.data());
                int64_t e;
                rd.sample(err, &e, 1);
                uint64_t *lwe = lwes + (cnt++) * lw;
                for (size_t l = 0; l < nl; l++) {
                    const uint64_t q = pp->q1[l];
                    u128 acc = 0;
                    for (size_t t = 0; t < n; t++) acc += (u128)masks[l * n + t] * sk1[l * n + t];
                    const uint64_t b = addm(addm((uint64_t)(acc % q), mod_i64(e, q), q), (uint64_t)(msg % q), q);
                    memcpy(lwe + l * (n + 1), masks.data() + l * n, 8 * n);
                    lwe[l * (n + 1) + n] = b;
                }
            }
        }
    }
    if (rd.err) return rd.err;
    rd.commit();
    return VIA_OK;
}
