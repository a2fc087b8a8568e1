// via_gpu.cpp — host side of the MI355X VIA-C answer path behind the C ABI of
// include/via_gpu.h. One HIP stream per context; every step of an answer is
// enqueued on it without host synchronisation.
#include "../../include/via_gpu.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <algorithm>
#include <cstring>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "host_internal.h"

using namespace via;

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            fprintf(stderr, "via_gpu: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
            return VIA_E_DEVICE;                                                       \
        }                                                                              \
    } while (0)

namespace {

typedef unsigned __int128 u128;

uint64_t pow_mod(uint64_t b, uint64_t e, uint64_t q) {
    u128 r = 1 % q, x = b % q;
    while (e) {
        if (e & 1) r = r * x % q;
        x = x * x % q;
        e >>= 1;
    }
    return (uint64_t)r;
}

uint64_t inv_mod(uint64_t a, uint64_t q) {
    __int128 t = 0, nt = 1, r = q, nr = a % q;
    while (nr) {
        __int128 qq = r / nr, tmp = t - qq * nt;
        t = nt;
        nt = tmp;
        tmp = r - qq * nr;
        r = nr;
        nr = tmp;
    }
    if (r != 1) return 0;
    if (t < 0) t += q;
    return (uint64_t)t;
}

// Primitive 2n-th root: the smallest base g = 2, 3, ... whose g^((q-1)/2n) has
// order exactly 2n, at most 1024 bases (ring/ntt.rs:155-198).
uint64_t find_psi(uint64_t q, uint32_t n) {
    uint64_t two_n = 2ULL * n;
    if (q < 3 || (q - 1) % two_n) return 0;
    uint64_t quot = (q - 1) / two_n;
    int log2n = __builtin_ctzll(two_n);
    uint64_t cap = q < 1026 ? q : 1026;
    for (uint64_t base = 2; base < cap; base++) {
        uint64_t c = pow_mod(base, quot, q), cur = c;
        bool ok = true;
        for (int i = 0; i < log2n; i++) {
            if (cur == 1) {
                ok = false;
                break;
            }
            cur = (uint64_t)((u128)cur * cur % q);
        }
        if (ok && cur == 1) return c;
    }
    return 0;
}

uint32_t bitrev(uint32_t x, int logn) {
    uint32_t r = 0;
    for (int i = 0; i < logn; i++)
        if ((x >> i) & 1) r |= 1u << (logn - 1 - i);
    return r;
}

// Constants of the direct form of the first forward pass (stages 0..R-1 of
// block 0, twiddles W_j = tw[j] of the bit-reversed table, products mod q,
// signs folded in as negative values): [W1, W2, W2W1, W3, W3W1] for R = 2
// (and W1 alone for R = 1), then for R = 3 the coefficients of x1, x5, x3,
// x7 in W_{4+k} * b_{2k+1}, k = 0..3, where b_1, b_3, b_5, b_7 are the odd
// outputs of the first two stages:
//   b1 = x1 + W1 x5 + W2 x3 + W2W1 x7    b3 = x1 + W1 x5 - W2 x3 - W2W1 x7
//   b5 = x1 - W1 x5 + W3 x3 - W3W1 x7    b7 = x1 - W1 x5 - W3 x3 + W3W1 x7
// Entries whose twiddles the table of degree n lacks stay zero.
void direct_consts(const double *tw, uint32_t n, uint64_t q, double *out) {
    auto w = [&](uint32_t j) -> uint64_t { return j < n ? (uint64_t)tw[j] : 0; };
    auto mul = [&](uint64_t a, uint64_t b) -> uint64_t { return (uint64_t)((u128)a * b % q); };
    const uint64_t w1 = w(1), w2 = w(2), w3 = w(3), w21 = mul(w2, w1), w31 = mul(w3, w1);
    out[0] = (double)w1;
    out[1] = (double)w2;
    out[2] = (double)w21;
    out[3] = (double)w3;
    out[4] = (double)w31;
    // coefficient signs of (x1, x5, x3, x7) in b1, b3, b5, b7
    const int sg[4][4] = {{1, 1, 1, 1}, {1, 1, -1, -1}, {1, -1, 1, -1}, {1, -1, -1, 1}};
    for (int k = 0; k < 4; k++) {
        const uint64_t wk = w(4 + k), wx = k < 2 ? w2 : w3, wx1 = k < 2 ? w21 : w31;
        const uint64_t v[4] = {wk, mul(wk, w1), mul(wk, wx), mul(wk, wx1)};
        for (int e = 0; e < 4; e++) out[5 + 4 * k + e] = sg[k][e] * (double)v[e];
    }
}

int ilog2(uint64_t x) { return x <= 1 ? 0 : 64 - __builtin_clzll(x - 1); }
bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }

}  // namespace

namespace {

via_status get_tab(via_ctx *c, uint64_t q, uint32_t n, DevTab *out) {
    const bool lookup_only = c->no_alloc;
    if (c->parent) c = c->parent;  // lanes share their context's tables
    std::lock_guard<std::mutex> g(c->mu);
    auto key = std::make_pair(q, n);
    auto it = c->tabs.find(key);
    if (it != c->tabs.end()) {
        *out = it->second;
        return VIA_OK;
    }
    if (lookup_only) return VIA_E_INVALID;  // tables are built by calls that may allocate
    uint64_t psi = find_psi(q, n);
    if (!psi) return VIA_E_INVALID;
    uint64_t psi_inv = inv_mod(psi, q);
    int logn = __builtin_ctz(n);
    std::vector<double> f(n), iv(n);
    uint64_t cf = 1, ci = 1;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t br = bitrev(k, logn);
        f[br] = (double)cf;
        iv[br] = (double)ci;
        cf = (uint64_t)((u128)cf * psi % q);
        ci = (uint64_t)((u128)ci * psi_inv % q);
    }
    // after the forward table: the constants of the direct first pass
    // (kDirectWords doubles, see direct_consts)
    std::vector<double> dcon(kDirectWords, 0.0);
    direct_consts(f.data(), n, q, dcon.data());
    f.insert(f.end(), dcon.begin(), dcon.end());
    DevTab t;
    t.q = q;
    t.n_inv = (double)inv_mod(n % q, q);
    HIP_TRY(hipMalloc(&t.fwd, sizeof(double) * (n + kDirectWords)));
    HIP_TRY(hipMalloc(&t.inv, sizeof(double) * n));
    HIP_TRY(hipMemcpy(t.fwd, f.data(), sizeof(double) * (n + kDirectWords), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.inv, iv.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    c->tabs[key] = t;
    *out = t;
    return VIA_OK;
}

via_status make_ring(via_ctx *c, uint32_t n, const uint64_t *q, int nl, RingDev *r) {
    memset(r, 0, sizeof *r);
    r->nl = nl;
    r->n = (int)n;
    r->logn = __builtin_ctz(n);
    for (int l = 0; l < nl; l++) {
        DevTab t;
        via_status s = get_tab(c, q[l], n, &t);
        if (s) return s;
        r->t[l].fwd = t.fwd;
        r->t[l].dfw = t.fwd + n;
        r->t[l].inv = t.inv;
        r->t[l].n_inv = t.n_inv;
        r->t[l].m = fmod_make(q[l]);
    }
    if (nl == 1) {
        r->t[1] = r->t[0];
        r->q_single = q[0];
        r->qdiv = widediv_make(q[0]);
    } else {
        r->rns.q0 = q[0];
        r->rns.q1 = q[1];
        r->rns.q0_inv_mod_q1 = inv_mod(q[0] % q[1], q[1]);
        r->rns.m0 = fmod_make(q[0]);
        r->rns.m1 = fmod_make(q[1]);
        r->rns.big_q = (u128)q[0] * q[1];
        r->qdiv = widediv_make(r->rns.big_q);
    }
    return VIA_OK;
}

Gadget make_gadget(uint64_t base, uint32_t L) {
    Gadget g;
    g.base = (uint32_t)base;
    g.L = (int)L;
    g.bl = 1;
    for (uint32_t k = 0; k < L; k++) g.bl *= base;
    g.wide = 0;
    g.f64 = 0;
    g.sf = ScaleF64();
    g.sw = ScaleW();
    return g;
}

bool fits_single(uint64_t q, uint64_t base, uint32_t L) {
    u128 bl = 1;
    for (uint32_t k = 0; k < L; k++) bl *= base;
    return (u128)(q / 2 + 1) * bl + q < ((u128)1 << 63);
}

// The high half of the digit split, |hi| <= B^(L-h), stays an exact fp64
// integer below 2^52 (digit_f64).
static bool digits_fit_f64(uint64_t base, uint32_t L) {
    const DigitPlan dp = digit_plan((uint32_t)base, (int)L);
    u128 b = 1;
    for (int k = dp.h; k < (int)L; k++) {
        b *= base;
        if (b > ((u128)1 << 52)) return false;
    }
    return true;
}

// Gadget at a single modulus q, taking the 192-bit scaling when B^L is wide.
Gadget make_gadget_q(uint64_t base, uint32_t L, uint64_t q) {
    Gadget g = make_gadget(base, L);
    g.wide = fits_single(q, base, L) ? 0 : 1;
    return g;
}

// Parameters the device path implements: the paper n1 = 2048 stack with a
// two-prime q1, NTT-friendly q1 limbs, q2 and q3 (q4 and p powers of two).
via_status check_params(const via_params *pp) {
    if (!pp) return VIA_E_INVALID;
    if (pp->n1 != (uint32_t)kSlots || pp->q1_nl != 2) return VIA_E_INVALID;
    if (!is_pow2(pp->n2) || pp->n2 > 1024 || pp->n1 % pp->n2 || !is_pow2(pp->n_rec) || pp->n1 % pp->n_rec)
        return VIA_E_INVALID;
    const uint64_t qs[4] = {pp->q1[0], pp->q1[1], pp->q2, pp->q3};
    for (uint64_t q : qs)
        if (q < 3 || q >= kMaxModulus || (q - 1) % (2ULL * pp->n1)) return VIA_E_INVALID;
    if (pp->q2 >= (1ULL << 35)) return VIA_E_INVALID;  // FirstDim limbs: five 7-bit limbs
    if (!is_pow2(pp->q4) || !is_pow2(pp->p)) return VIA_E_INVALID;
    u128 Q1 = (u128)pp->q1[0] * pp->q1[1];
    if (Q1 >> 96) return VIA_E_INVALID;
    u128 bl = 1;
    for (uint32_t k = 0; k < pp->l_ck; k++) {
        bl *= pp->ck_base;
        if (bl >> 96) return VIA_E_INVALID;
    }
    if (!fits_single(pp->q2, pp->b2, pp->l_query) || !fits_single(pp->q3, pp->b_rsk, pp->l_rsk))
        return VIA_E_INVALID;
    if (pp->b1 < 2 || pp->b1 >= (1u << 20) || pp->b2 < 2 || pp->b2 >= (1u << 20) || pp->ck_base < 2 ||
        pp->ck_base >= (1u << 20) || pp->b_rsk < 2 || pp->b_rsk >= (1u << 20))
        return VIA_E_INVALID;
    if (pp->l_query < 1 || pp->l_ck < 1 || pp->l_rsk < 1) return VIA_E_INVALID;
    if (pp->ck_base >= (1u << 15)) return VIA_E_INVALID;  // int16 digit planes
    if (!digits_fit_f64(pp->ck_base, pp->l_ck) || !digits_fit_f64(pp->b1, pp->l_query) ||
        !digits_fit_f64(pp->b2, pp->l_query) || !digits_fit_f64(pp->b_rsk, pp->l_rsk))
        return VIA_E_INVALID;
    if ((u128)pp->q2 * pp->q3 >> 63) return VIA_E_INVALID;  // q2 -> q3 switch in 64 bits
    if ((u128)pp->q3 * pp->q4 >> 63) return VIA_E_INVALID;
    return VIA_OK;
}

// Whether the cascade steps take the fixed-shape MAC (k_cascade_mac) for
// these parameters: its direct first pass sums eight products of a digit
// (at most B/2 in magnitude) with twiddle products below q, and the lazy
// stages after it add below 2q each, all within 2^45 (the input bound of the
// lazy products). The cascade keys are then stored in its slot order.
bool cascade_fixed_ok(const via_params *pp) {
    const u128 qmax = std::max(pp->q1[0], pp->q1[1]);
    return ((u128)(pp->ck_base / 2 + 1) * 8 + 14) * qmax < ((u128)1 << 45);
}

// The context a call runs on, held for the duration of the call. Device-
// buffer entry points take the context itself (waiting for a call in
// progress, since their caller synchronises on its stream); host-buffer
// entry points take it when it is free and otherwise a free lane, created on
// demand up to kMaxLanes and then waited for.
constexpr size_t kMaxLanes = 7;
struct CallCtx {
    via_ctx *c = nullptr;
    CallCtx() = default;
    CallCtx(const CallCtx &) = delete;
    CallCtx &operator=(const CallCtx &) = delete;
    ~CallCtx() {
        if (c) c->call_mu.unlock();
    }
};

void hold_ctx(via_ctx *root, CallCtx *out) {
    root->call_mu.lock();
    out->c = root;
}

via_status hold_lane(via_ctx *root, CallCtx *out) {
    if (root->call_mu.try_lock()) {
        out->c = root;
        return VIA_OK;
    }
    {
        std::lock_guard<std::mutex> g(root->lane_mu);
        for (via_ctx *l : root->lanes)
            if (l->call_mu.try_lock()) {
                out->c = l;
                return VIA_OK;
            }
        if (root->lanes.size() < kMaxLanes) {
            via_ctx *l = nullptr;
            if (via_ctx_create(root->device, &l) != VIA_OK) return VIA_E_DEVICE;
            l->parent = root;
            l->call_mu.lock();
            root->lanes.push_back(l);
            out->c = l;
            return VIA_OK;
        }
    }
    hold_ctx(root, out);
    return VIA_OK;
}

// Host-buffer calls report non-canonical query words through flag slot 1,
// cleared here and read by host_flag_status once the call has synchronised.
struct HostFlag {
    via_ctx *c;
    explicit HostFlag(via_ctx *ctx) : c(ctx) {
        c->flag_slot = 1;
        __atomic_store_n(&c->bad_h[1], 0, __ATOMIC_RELEASE);
    }
    ~HostFlag() { c->flag_slot = 0; }
    via_status status() const { return __atomic_load_n(&c->bad_h[1], __ATOMIC_ACQUIRE) ? VIA_E_INVALID : VIA_OK; }
};

template <typename T>
via_status dalloc(Workspace &w, T **p, size_t count) {
    HIP_TRY(hipMalloc((void **)p, sizeof(T) * (count ? count : 1)));
    w.all.push_back((void *)*p);
    return VIA_OK;
}

void free_ws(Workspace &w) {
    for (void *p : w.all) (void)hipFree(p);
    w = Workspace();
}

// Launch shapes the autotuner tries for `items` gadget products of depth L:
// the 256-thread kernel with radix-8 or radix-16 transform passes and the
// 512-thread kernel with radix-8 passes, each with 1, 2, 3, 4, 6, 9 or 18
// digit groups, and the wide kernel with 1, 2, 3 or 5 (group counts up to
// L), as long as a launch stays within 8192 workgroups (two tiles per item).
std::vector<StepTune> tune_candidates(int items, int L) {
    std::vector<StepTune> out;
    // variant: 0 256 threads radix 8, 1 radix 16, 2 wide (1024 threads), 3 512
    // threads, 4 and 5 the 256- and 512-thread kernels with one workgroup per limb
    for (int variant = 0; variant < 6; variant++)
        for (int g : variant == 2 ? std::initializer_list<int>{1, 2, 3, 5}
                                  : std::initializer_list<int>{1, 2, 3, 4, 6, 9, 18})
            if (g <= L && (long)items * 2 * g <= 8192) {
                StepTune t;
                t.groups = g;
                t.wide = variant == 2;
                t.radix16 = variant == 1;
                t.t512 = variant == 3 || variant == 5;
                t.lsplit = variant >= 4;
                // the items per workgroup of the fixed-shape cascade MAC follow the
                // variant's number in via_ctx_set_tuning (1 wide, 2 radix 16)
                t.ipw = variant_ipw(variant == 1 ? 2 : variant == 2 ? 1 : variant);
                out.push_back(t);
            }
    if (out.empty()) out.push_back(StepTune());
    return out;
}

int max_tune_groups(int items, int L) {
    int m = 1;
    for (const StepTune &t : tune_candidates(items, L)) m = std::max(m, t.groups);
    return m;
}

via_status ensure_ws(via_ctx *c, const via_params *pp, size_t Q, size_t B, size_t bits) {
    Workspace &w = c->ws;
    size_t I = pp->num_rows, J = pp->num_cols, l_ck = pp->l_ck;
    if (w.Q >= Q && w.B >= B && w.I >= I && w.J >= J && w.bits >= bits && w.l_ck >= l_ck && w.lwe_f) return VIA_OK;
    // grow every dimension, so that alternating shapes settle on one workspace
    Q = std::max(Q, w.Q);
    B = std::max(B, w.B);
    I = std::max(I, w.I);
    J = std::max(J, w.J);
    bits = std::max(bits, w.bits);
    l_ck = std::max(l_ck, w.l_ck);
    const uint64_t gen = w.gen;
    free_ws(w);
    w.gen = gen + 1;  // captured graphs of the old buffers never match again
    const size_t n = pp->n1, nl = 2, S = kSlots;
    w.Q = Q;
    w.B = B;
    w.I = I;
    w.J = J;
    w.bits = bits;
    w.l_ck = l_ck;
    const size_t items = Q * B;
    // digit groups per gadget product: about eight workgroups per item,
    // between 512 and 4096 in all (measured best at 36, 288 and 1152 items;
    // more groups add partial-sum traffic that small batches do not repay)
    const long wg_target = std::min<long>(4096, std::max<long>(512, (long)items * 8));
    int groups = 1;
    while (groups < (int)pp->l_ck && (long)items * 2 * groups < wg_target) groups++;
    w.groups = groups;
    // room for the autotuner's candidates (tune_candidates) at the deepest gadget held
    w.groups_max = std::max(groups, max_tune_groups((int)items, (int)l_ck));
    const FdPlan fp = fd_plan((int)I, (int)J, (int)n, (int)Q);
    via_status s;
#define A(p, cnt) \
    if ((s = dalloc(w, &(p), (cnt)))) return s
    A(w.lwe_u, items * nl * (n + 1));
    A(w.lwe_f, items * nl * (n + 1));
    A(w.masks[0], items * nl * S);
    A(w.masks[1], items * nl * S);
    A(w.body[0], items * nl * S);
    A(w.body[1], items * nl * S);
    A(w.dig, items * l_ck * S);
    A(w.pm, (size_t)w.groups_max * items * 2 * nl * S);
    A(w.pb, (size_t)w.groups_max * items * 2 * nl * S);
    // RGSW bits: bits * l_query = B LWEs, 4 polys of nl limbs per LWE
    A(w.rgsw, Q * B * 4 * nl * n);
    A(w.rgsw_eval, Q * B * 4 * nl * n);
    A(w.dm[0], Q * I * 2 * nl * n);
    A(w.dm[1], Q * I * 2 * nl * n);
    A(w.sw, Q * I * 2 * n);
    A(w.sel, Q * I * 2 * n);
    A(w.fd, Q * J * 2 * n);
    A(w.fdk, fd_out_words(fp));
    A(w.sell, fd_sel_bytes(fp));
    A(w.cm[0], Q * J * 2 * n);
    A(w.cm[1], Q * J * 2 * n);
    A(w.rgsw2, Q * B * 4 * n);
    A(w.res[0], Q * 2 * n);
    A(w.res[1], Q * 2 * n);
    A(w.ans_u, Q * 2 * 1024);
#undef A
    return VIA_OK;
}

// Workspace of the VIA-B repack for up to T queries of n_lwe LWEs.
via_status ensure_rw(via_ctx *c, const via_params *pp, int T, size_t n_lwe) {
    RepackWs &w = c->rw;
    if (w.T >= T && w.n_lwe >= n_lwe && w.l_ck >= pp->l_ck && w.n2 >= pp->n2 && w.rot) return VIA_OK;
    T = std::max(T, w.T);
    n_lwe = std::max(n_lwe, w.n_lwe);
    const size_t L = std::max<size_t>(pp->l_ck, w.l_ck), n2 = std::max<size_t>(pp->n2, w.n2);
    for (void *p : w.all) (void)hipFree(p);
    w = RepackWs();
    const size_t n = pp->n1, S = kSlots;
    const size_t pairs = (size_t)std::max(1, T / 2);
    // largest items * groups over the levels
    size_t ig = 0;
    for (size_t items = pairs;; items = std::max<size_t>(1, items / 2)) {
        size_t g = 1;
        while (g < L && items * 2 * g < 512) g++;
        ig = std::max(ig, items * g);
        if (items == 1) break;
    }
    auto al = [&](void **p, size_t bytes) -> via_status {
        if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) return VIA_E_NOMEM;
        w.all.push_back(*p);
        return VIA_OK;
    };
    via_status s;
    if ((s = al((void **)&w.rot, sizeof(double) * T * 2 * n))) return s;
    if ((s = al((void **)&w.packed, sizeof(double) * 2 * n))) return s;
    if ((s = al((void **)&w.body[0], sizeof(double) * pairs * n))) return s;
    if ((s = al((void **)&w.body[1], sizeof(double) * pairs * n))) return s;
    if ((s = al((void **)&w.pm, sizeof(double) * ig * 2 * S))) return s;
    if ((s = al((void **)&w.pb, sizeof(double) * ig * 2 * S))) return s;
    if ((s = al((void **)&w.dig, sizeof(int16_t) * pairs * L * 2 * S))) return s;
    if ((s = al((void **)&w.lwe_u, sizeof(uint64_t) * T * n_lwe * 2 * (n + 1)))) return s;
    if ((s = al((void **)&w.ans_u, sizeof(uint64_t) * 2 * n2))) return s;
    w.T = T;
    w.n_lwe = n_lwe;
    w.l_ck = L;
    w.n2 = n2;
    return VIA_OK;
}

}  // namespace

via_status via::ensure_workspace(via_ctx *c, const via_params *pp, size_t Q, size_t n_lwe) {
    const size_t bits = n_lwe / pp->l_query;
    return ensure_ws(c, pp, Q, n_lwe, bits);
}

via_status via::ensure_repack_workspace(via_ctx *c, const via_params *pp, int T, size_t n_lwe) {
    return ensure_rw(c, pp, T, n_lwe);
}

size_t via::keys_footprint(const via_params *pp) {
    const size_t n1 = pp->n1, L = pp->l_ck;
    size_t words = 0;
    for (size_t nout = 2; nout <= n1; nout <<= 1) {
        const size_t step = (2 * n1 / nout) * L * 2 * nout;  // one limb
        words += 2 * step;
        if (pp->n_rec < pp->n2 && nout >= 2 * (size_t)pp->n_rec) words += step;  // q2 repack copy
    }
    words += L * 2 * 2 * n1;                               // conversion key
    words += (n1 / pp->n2) * pp->l_rsk * 2 * pp->n2;       // ring-switch key
    return sizeof(double) * words;
}

via_status via::prepare_tables(via_ctx *c, const via_params *pp) {
    RingDev r;
    via_status s;
    for (uint32_t nout = 2; nout <= pp->n1; nout <<= 1)
        if ((s = make_ring(c, nout, pp->q1, 2, &r))) return s;
    uint64_t q2 = pp->q2;
    // q2 at n1 (FirstDim, gates) and at the VIA-B repack degrees
    for (uint32_t nout = 4; nout <= pp->n1; nout <<= 1)
        if ((s = make_ring(c, nout, &q2, 1, &r))) return s;
    DevTab t;
    return get_tab(c, pp->q3, pp->n2, &t);
}

via_status via::lane_acquire(via_ctx *root, via_ctx **lane) {
    CallCtx cc;
    via_status s = hold_lane(root, &cc);
    if (s) return s;
    *lane = cc.c;
    cc.c = nullptr;  // the hold passes to the caller
    return VIA_OK;
}

void via::lane_release(via_ctx *lane) { lane->call_mu.unlock(); }

bool via::lwes_canonical(const via_params *pp, const uint64_t *lwes, size_t n_lwe) {
    const size_t n1p = (size_t)pp->n1 + 1;
    uint64_t bad = 0;
    for (size_t i = 0; i < n_lwe * 2; i++) {
        const uint64_t q = pp->q1[i & 1], *x = lwes + i * n1p;
        for (size_t c = 0; c < n1p; c++) bad |= (uint64_t)(x[c] >= q);  // branch-free: vectorises
    }
    return bad == 0;
}

bool via::lwes_copy_canonical(const via_params *pp, uint64_t *dst, const uint64_t *lwes, size_t n_lwe) {
    const size_t n1p = (size_t)pp->n1 + 1;
    uint64_t bad = 0;
    for (size_t i = 0; i < n_lwe * 2; i++) {
        const uint64_t q = pp->q1[i & 1], *x = lwes + i * n1p;
        uint64_t *d = dst + i * n1p;
        for (size_t c = 0; c < n1p; c++) {
            const uint64_t v = x[c];
            d[c] = v;
            bad |= (uint64_t)(v >= q);
        }
    }
    return bad == 0;
}

size_t via::keys_device_bytes(const via_keys *k) {
    size_t b = 0;
    for (size_t x : k->bytes) b += x;
    return b;
}

extern "C" {

const char *via_status_str(via_status s) {
    switch (s) {
        case VIA_OK: return "ok";
        case VIA_E_DIM: return "dimension mismatch";
        case VIA_E_QUERY_LEN: return "query length mismatch";
        case VIA_E_INDEX: return "index out of range";
        case VIA_E_NOMEM: return "out of device memory";
        case VIA_E_INVALID: return "parameters not supported by the device path";
        case VIA_E_DEVICE: return "HIP runtime error";
        case VIA_E_BATCH_LEN: return "batch length mismatch";
        case VIA_E_BUFFER_SHORT: return "buffer too short";
        case VIA_E_FORMAT_TAG: return "unknown wire format tag";
        case VIA_E_PRG_SEED: return "invalid PRG seed";
    }
    return "unknown";
}

via_status via_ctx_create(int device, via_ctx **out) {
    if (!out) return VIA_E_INVALID;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return VIA_E_DEVICE;
    HIP_TRY(hipSetDevice(device));
    via_ctx *c = new via_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return VIA_E_DEVICE;
    }
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return VIA_E_DEVICE;
        }
    for (auto &e : c->ev_fd)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return VIA_E_DEVICE;
        }
    for (auto &e : c->ev_mark)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return VIA_E_DEVICE;
        }
    for (auto &e : c->ev_dec)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return VIA_E_DEVICE;
        }
    if (hipHostMalloc((void **)&c->bad_h, 3 * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostGetDevicePointer((void **)&c->bad_d, c->bad_h, 0) != hipSuccess) {
        delete c;
        return VIA_E_DEVICE;
    }
    c->bad_h[0] = c->bad_h[1] = c->bad_h[2] = 0;
    if (hipMalloc((void **)&c->tree_bar, 2 * sizeof(unsigned)) != hipSuccess ||
        hipMemset(c->tree_bar, 0, 2 * sizeof(unsigned)) != hipSuccess) {
        via_ctx_destroy(c);
        return VIA_E_NOMEM;
    }
    if (scratch_setup(device, &c->scratch) != VIA_OK) {
        via_ctx_destroy(c);
        return VIA_E_DEVICE;
    }
    *out = c;
    return VIA_OK;
}

void via_ctx_destroy(via_ctx *c) {
    if (!c) return;
    for (via_ctx *l : c->lanes) via_ctx_destroy(l);
    c->lanes.clear();
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &e : c->graph_cache) {
        (void)hipGraphExecDestroy(e.exec);
        (void)hipGraphDestroy(e.graph);
    }
    free_ws(c->ws);
    for (void *p : c->rw.all) (void)hipFree(p);
    for (auto &kv : c->tabs) {
        (void)hipFree(kv.second.fwd);
        (void)hipFree(kv.second.inv);
    }
    for (auto &e : c->ev) (void)hipEventDestroy(e);
    for (auto &e : c->ev_fd) (void)hipEventDestroy(e);
    for (auto &e : c->ev_mark) (void)hipEventDestroy(e);
    for (auto &e : c->ev_dec) (void)hipEventDestroy(e);
    if (c->bad_h) (void)hipHostFree(c->bad_h);
    if (c->tree_bar) (void)hipFree(c->tree_bar);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

// A persistent tree whose barrier wait gave up left wrong values and the
// barrier words inconsistent: clear both and report the failure.
static via_status tree_stall_status(via_ctx *c) {
    if (!__atomic_exchange_n(&c->bad_h[2], 0, __ATOMIC_ACQ_REL)) return VIA_OK;
    (void)hipMemsetAsync(c->tree_bar, 0, 2 * sizeof(unsigned), c->stream);
    (void)hipStreamSynchronize(c->stream);
    return VIA_E_DEVICE;
}

via_status via_ctx_synchronize(via_ctx *c) {
    if (!c) return VIA_E_INVALID;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (tree_stall_status(c)) return VIA_E_DEVICE;
    // a device query since the last synchronisation held a non-canonical word
    if (__atomic_exchange_n(&c->bad_h[0], 0, __ATOMIC_ACQ_REL)) return VIA_E_INVALID;
    return VIA_OK;
}

via_status via_ctx_step_times(via_ctx *c, double *t7) {
    if (!c || !t7) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    if (!c->have_times) return VIA_E_INVALID;
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int i = 0; i < 7; i++) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
        t7[i] = ms * 1e-3;
    }
    return VIA_OK;
}

via_status via_ctx_first_dim_time(via_ctx *c, double *seconds) {
    if (!c || !seconds) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    if (!c->have_fd_time) return VIA_E_INVALID;
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev_fd[0], c->ev_fd[1]));
    *seconds = ms * 1e-3;
    return VIA_OK;
}

via_status via_ctx_mark(via_ctx *c, int which) {
    if (!c || which < 0 || which > 1) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    HIP_TRY(hipEventRecord(c->ev_mark[which], c->stream));
    return VIA_OK;
}

via_status via_ctx_marked_time(via_ctx *c, double *seconds) {
    if (!c || !seconds) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    HIP_TRY(hipEventSynchronize(c->ev_mark[1]));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev_mark[0], c->ev_mark[1]));
    *seconds = ms * 1e-3;
    return VIA_OK;
}

// ---- database ---------------------------------------------------------------------

static via_status db_common(via_ctx *c, const via_params *pp, const uint64_t *records, const uint8_t *bytes,
                            size_t n_records, via_db **out) {
    if (!c || !out) return VIA_E_INVALID;
    via_status s = check_params(pp);
    if (s) return s;
    if (!is_pow2(pp->num_rows) || !is_pow2(pp->num_cols)) return VIA_E_DIM;
    CallCtx cc;
    hold_ctx(c, &cc);
    HostFlag flag(c);
    const size_t I = pp->num_rows, J = pp->num_cols, n1 = pp->n1, nr = pp->n_rec;
    if (n_records > I * J * (n1 / nr)) return VIA_E_DIM;  // answer.rs:503-508
    if (bytes && (pp->p != 16 || nr % 2)) return VIA_E_INVALID;
    HIP_TRY(hipSetDevice(c->device));
    via_db *db = new via_db();
    db->id = next_object_id();
    db->ctx = c;
    db->pp = *pp;
    const FdPlan fp = fd_plan((int)I, (int)J, (int)n1, 1);
    if (hipMalloc(&db->limbs, fd_db_bytes(fp)) != hipSuccess) {
        delete db;
        return VIA_E_NOMEM;
    }
    // setup_db and the transform run over fp64 cells, then the cells are
    // packed into the limb fragments and released
    double *cells = nullptr;
    if (hipMalloc(&cells, sizeof(double) * I * J * n1) != hipSuccess) {
        (void)hipFree(db->limbs);
        delete db;
        return VIA_E_NOMEM;
    }
    // Stream the records through a bounded staging buffer.
    const size_t rec_words = bytes ? nr / 2 : nr * 8;  // bytes per record
    size_t chunk = n_records ? n_records : 1;
    void *stage = nullptr;
    if (hipMalloc(&stage, rec_words * chunk) != hipSuccess) {
        (void)hipFree(cells);
        (void)hipFree(db->limbs);
        delete db;
        return VIA_E_NOMEM;
    }
    if (n_records) HIP_TRY(hipMemcpyAsync(stage, bytes ? (const void *)bytes : (const void *)records, rec_words * n_records,
                                          hipMemcpyHostToDevice, c->stream));
    if (bytes)
        launch_setup_db_bytes((const uint8_t *)stage, n_records, cells, (int)I, (int)J, (int)n1, (int)nr, c->stream);
    else
        launch_setup_db((const uint64_t *)stage, n_records, cells, (int)I, (int)J, (int)n1, (int)nr, pp->p,
                        c->bad_d + 1, c->stream);
    RingDev r2;
    uint64_t q2 = pp->q2;
    if ((s = make_ring(c, (uint32_t)n1, &q2, 1, &r2))) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(stage);
        (void)hipFree(cells);
        (void)hipFree(db->limbs);
        delete db;
        return s;
    }
    launch_ntt(r2, cells, I * J, 1, false, c->stream);
    launch_pack_db(cells, db->limbs, fp, c->stream);
    const hipError_t e = hipStreamSynchronize(c->stream);
    (void)hipFree(stage);
    (void)hipFree(cells);
    if (e != hipSuccess || hipGetLastError() != hipSuccess || (s = flag.status())) {
        // a record word at or above p (Poly<N_REC, p> holds canonical values only)
        (void)hipFree(db->limbs);
        delete db;
        return e != hipSuccess ? VIA_E_DEVICE : (s ? s : VIA_E_DEVICE);
    }
    *out = db;
    return VIA_OK;
}

via_status via_db_create(via_ctx *c, const via_params *pp, const uint64_t *records, size_t n_records, via_db **out) {
    if (n_records && !records) return VIA_E_INVALID;
    return db_common(c, pp, records, nullptr, n_records, out);
}

via_status via_db_create_bytes(via_ctx *c, const via_params *pp, const uint8_t *bytes, size_t n_records,
                               via_db **out) {
    if (n_records && !bytes) return VIA_E_INVALID;
    return db_common(c, pp, nullptr, bytes, n_records, out);
}

void via_db_destroy(via_db *db) {
    if (!db) return;
    CallCtx cc;
    hold_ctx(db->ctx, &cc);
    (void)hipSetDevice(db->ctx->device);
    (void)hipStreamSynchronize(db->ctx->stream);
    (void)hipFree(db->limbs);
    delete db;
}

// Incremental record updates: record m = s*I*J + i*J + j is slot s of cell
// (i, j) (setup_db.rs:58-72); each touched cell is rebuilt in place on the
// device (launch_db_update). A repeated index takes its last record.
via_status via_db_update(via_db *db, const uint64_t *indices, size_t n, const uint64_t *records) {
    if (!db || (n && (!indices || !records))) return VIA_E_INVALID;
    if (!n) return VIA_OK;
    via_ctx *c = db->ctx;
    const via_params *pp = &db->pp;
    const size_t I = pp->num_rows, J = pp->num_cols, n1 = pp->n1, nr = pp->n_rec, d = n1 / nr;
    const size_t cap = I * J * d;
    for (size_t u = 0; u < n; u++)
        if (indices[u] >= cap) return VIA_E_INDEX;
    for (size_t w = 0; w < n * nr; w++)
        if (records[w] >= pp->p) return VIA_E_INVALID;  // records hold canonical values mod p
    CallCtx cc;
    hold_ctx(c, &cc);
    // cell -> its d slots' sources (index into records, or -1)
    std::map<int64_t, std::vector<int32_t>> touched;
    for (size_t u = 0; u < n; u++) {
        const uint64_t m = indices[u];
        const int64_t cell = (int64_t)(m % (I * J));
        auto &v = touched[cell];
        if (v.empty()) v.assign(d, -1);
        v[m / (I * J)] = (int32_t)u;
    }
    std::vector<int64_t> cells;
    std::vector<int32_t> slots;
    for (auto &kv : touched) {
        cells.push_back(kv.first);
        slots.insert(slots.end(), kv.second.begin(), kv.second.end());
    }
    HIP_TRY(hipSetDevice(c->device));
    RingDev r2;
    uint64_t q2 = pp->q2;
    via_status s = make_ring(c, (uint32_t)n1, &q2, 1, &r2);
    if (s) return s;
    const FdPlan fp = fd_plan((int)I, (int)J, (int)n1, 1);
    int64_t *d_cells = nullptr;
    int32_t *d_slots = nullptr;
    uint64_t *d_recs = nullptr;
    if (hipMalloc(&d_cells, sizeof(int64_t) * cells.size()) != hipSuccess) return VIA_E_NOMEM;
    if (hipMalloc(&d_slots, sizeof(int32_t) * slots.size()) != hipSuccess) {
        (void)hipFree(d_cells);
        return VIA_E_NOMEM;
    }
    if (hipMalloc(&d_recs, sizeof(uint64_t) * n * nr) != hipSuccess) {
        (void)hipFree(d_cells);
        (void)hipFree(d_slots);
        return VIA_E_NOMEM;
    }
    hipStream_t st = c->stream;
    bool ok = hipMemcpyAsync(d_cells, cells.data(), sizeof(int64_t) * cells.size(), hipMemcpyHostToDevice, st) ==
                  hipSuccess &&
              hipMemcpyAsync(d_slots, slots.data(), sizeof(int32_t) * slots.size(), hipMemcpyHostToDevice, st) ==
                  hipSuccess &&
              hipMemcpyAsync(d_recs, records, sizeof(uint64_t) * n * nr, hipMemcpyHostToDevice, st) == hipSuccess;
    if (ok) {
        launch_db_update(db->limbs, fp, r2, d_cells, (int)cells.size(), d_slots, (int)d, d_recs, (int)nr, pp->p, st);
        ok = hipStreamSynchronize(st) == hipSuccess && hipGetLastError() == hipSuccess;
    }
    (void)hipFree(d_cells);
    (void)hipFree(d_slots);
    (void)hipFree(d_recs);
    return ok ? VIA_OK : VIA_E_DEVICE;
}

via_status via_db_read_cells(const via_db *db, uint64_t *out, size_t words) {
    if (!db || !out) return VIA_E_INVALID;
    const size_t total = (size_t)db->pp.num_rows * db->pp.num_cols * db->pp.n1;
    if (words < total) return VIA_E_INVALID;
    via_ctx *c = db->ctx;
    CallCtx cc;
    hold_ctx(c, &cc);
    uint64_t *tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, sizeof(uint64_t) * total));
    launch_unpack_db(db->limbs, tmp, fd_plan((int)db->pp.num_rows, (int)db->pp.num_cols, (int)db->pp.n1, 1), c->stream);
    HIP_TRY(hipMemcpyAsync(out, tmp, sizeof(uint64_t) * total, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree(tmp));
    return VIA_OK;
}

// ---- keys --------------------------------------------------------------------------------

// Key words are canonical residues (the reference's Zq types hold nothing
// else): the cascade and conversion keys limb by limb below q1[l], the
// ring-switch key below q3. A word outside would break the exactness bounds
// of the fp64 arithmetic, so it is refused (VIA_E_INVALID).
static bool keys_canonical(const via_params *pp, const uint64_t *cascade_key, const uint64_t *conv_key,
                           const uint64_t *rsk) {
    const size_t n1 = pp->n1, L = pp->l_ck;
    size_t off = 0;
    for (size_t nout = 2; nout <= n1; nout <<= 1) {
        // step words [idx][k][part][limb][nout]: limb = (w / nout) % 2
        const size_t words = (2 * n1 / nout) * L * 2 * 2 * nout;
        const uint64_t *x = cascade_key + off;
        for (size_t w = 0; w < words; w += nout) {
            const uint64_t q = pp->q1[(w / nout) & 1];
            for (size_t c = 0; c < nout; c++)
                if (x[w + c] >= q) return false;
        }
        off += words;
    }
    for (size_t w = 0; w < L * 2 * 2 * n1; w++)
        if (conv_key[w] >= pp->q1[(w / n1) & 1]) return false;
    const size_t rw = (size_t)(n1 / pp->n2) * pp->l_rsk * 2 * pp->n2;
    for (size_t w = 0; w < rw; w++)
        if (rsk[w] >= pp->q3) return false;
    return true;
}

// Bytes of each device buffer of a key set: cascade step s (n_out = 2^(s+1)),
// the conversion key, the ring-switch key, repack step r (n_out =
// repack_nout0 * 2^r).
static size_t key_cascade_bytes(const via_params *pp, size_t s) {
    const size_t nout = (size_t)2 << s;
    return sizeof(double) * (2 * pp->n1 / nout) * pp->l_ck * 2 * 2 * nout;
}
static size_t key_conv_bytes(const via_params *pp) { return sizeof(double) * (size_t)pp->l_ck * 2 * 2 * pp->n1; }
static size_t key_rsk_bytes(const via_params *pp) {
    return sizeof(double) * (size_t)(pp->n1 / pp->n2) * pp->l_rsk * 2 * pp->n2;
}
static size_t key_repack_bytes(const via_params *pp, uint32_t nout0, size_t r) {
    const size_t nout = (size_t)nout0 << r;
    return sizeof(double) * (2 * pp->n1 / nout) * pp->l_ck * 2 * nout;
}

// Releases every device buffer of a key set built so far (zeroizing first).
static void keys_free(via_keys *k) {
    via_ctx *c = k->ctx;
    const via_params *pp = &k->pp;
    (void)hipStreamSynchronize(c->stream);
    for (size_t i = 0; i < k->cascade.size(); i++)
        (void)hipMemsetAsync(k->cascade[i], 0, key_cascade_bytes(pp, i), c->stream);
    if (k->conv) (void)hipMemsetAsync(k->conv, 0, key_conv_bytes(pp), c->stream);
    if (k->rsk) (void)hipMemsetAsync(k->rsk, 0, key_rsk_bytes(pp), c->stream);
    for (size_t i = 0; i < k->repack.size(); i++)
        (void)hipMemsetAsync(k->repack[i], 0, key_repack_bytes(pp, k->repack_nout0, i), c->stream);
    (void)hipStreamSynchronize(c->stream);
    for (double *p : k->repack) (void)hipFree(p);
    for (double *p : k->cascade) (void)hipFree(p);
    (void)hipFree(k->conv);
    (void)hipFree(k->rsk);
    delete k;
}

via_status via_keys_create(via_ctx *c, const via_params *pp, const uint64_t *cascade_key, const uint64_t *conv_key,
                           const uint64_t *rsk, via_keys **out) {
    if (!c || !out || !cascade_key || !conv_key || !rsk) return VIA_E_INVALID;
    via_status s = check_params(pp);
    if (s) return s;
    if (!keys_canonical(pp, cascade_key, conv_key, rsk)) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    if (hipSetDevice(c->device) != hipSuccess) return VIA_E_DEVICE;
    via_keys *k = new via_keys();
    k->id = next_object_id();
    k->ctx = c;
    k->pp = *pp;
    const uint32_t n1 = pp->n1, nl = 2, L = pp->l_ck;
    const size_t max_step = (size_t)2 * n1 * L * 2 * nl;
    // VIA-B: the steps of output degree >= 2 n_rec also serve the repack at
    // q2 (a batch of T has extraction degree n2/T >= n_rec), mod-switched
    // once here (repack.rs:684-805 does it per answer; the values are equal).
    const bool vb = pp->n_rec < pp->n2;
    k->repack_nout0 = vb ? 2 * pp->n_rec : 0;
    // staging: the u64 words of the largest key piece, and for VIA-B their
    // fp64 copy and its q2 switch; every return below frees them and the
    // key buffers made so far
    uint64_t *stage = nullptr;
    double *tmp_f = nullptr, *tmp_q2 = nullptr, *tmp_perm = nullptr;
    auto fail = [&](via_status st) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(stage);
        (void)hipFree(tmp_f);
        (void)hipFree(tmp_q2);
        (void)hipFree(tmp_perm);
        keys_free(k);
        return st;
    };
    auto key_buf = [&](double **d, size_t words) {
        return hipMalloc((void **)d, sizeof(double) * words) == hipSuccess;
    };
    if (hipMalloc(&stage, sizeof(uint64_t) * max_step) != hipSuccess) return fail(VIA_E_NOMEM);
    if (vb && (hipMalloc(&tmp_f, sizeof(double) * max_step) != hipSuccess ||
               hipMalloc(&tmp_q2, sizeof(double) * max_step / 2) != hipSuccess))
        return fail(VIA_E_NOMEM);
    k->cascade_perm = cascade_fixed_ok(pp);
    if (k->cascade_perm && hipMalloc((void **)&tmp_perm, sizeof(double) * max_step) != hipSuccess)
        return fail(VIA_E_NOMEM);
    size_t off = 0;
    for (uint32_t nout = 2; nout <= n1; nout <<= 1) {
        const uint32_t rank_in = 2 * n1 / nout;
        const size_t words = (size_t)rank_in * L * 2 * nl * nout;
        double *d = nullptr;
        if (!key_buf(&d, words)) return fail(VIA_E_NOMEM);
        k->cascade.push_back(d);
        if (hipMemcpyAsync(stage, cascade_key + off, sizeof(uint64_t) * words, hipMemcpyHostToDevice, c->stream) !=
            hipSuccess)
            return fail(VIA_E_DEVICE);
        launch_key_relayout(stage, d, (int)rank_in, (int)L, (int)nl, (int)nout, c->stream);
        RingDev r;
        if ((s = make_ring(c, nout, pp->q1, 2, &r))) return fail(s);
        launch_ntt(r, d, (size_t)L * 2 * nl * rank_in, rank_in, false, c->stream);
        if (k->cascade_perm) {
            // rows of rank_in * nout = 2 n1 slots into the order of the fixed-shape MAC
            launch_cascade_key_permute(d, tmp_perm, (size_t)L * 2 * nl, 2, c->stream);
            if (hipMemcpyAsync(d, tmp_perm, sizeof(double) * words, hipMemcpyDeviceToDevice, c->stream) !=
                hipSuccess)
                return fail(VIA_E_DEVICE);
        }
        if (vb && nout >= k->repack_nout0) {
            // mod_switch_rlev (repack.rs:43-56) of every sample, q1 -> q2
            const size_t w2 = words / nl;
            double *rk = nullptr;
            if (hipMalloc((void **)&rk, sizeof(double) * w2) != hipSuccess) return fail(VIA_E_NOMEM);
            k->repack.push_back(rk);
            launch_u64_to_f64(stage, tmp_f, words, c->stream);
            launch_modswitch_q1_q2(r, pp->q2, tmp_f, tmp_q2, (size_t)rank_in * L * 2, c->stream);
            launch_key_relayout_q2(tmp_q2, rk, (int)rank_in, (int)L, (int)nout, c->stream);
            RingDev r2;
            uint64_t q2 = pp->q2;
            if ((s = make_ring(c, nout, &q2, 1, &r2))) return fail(s);
            launch_ntt(r2, rk, (size_t)L * 2 * rank_in, rank_in, false, c->stream);
        }
        // the stage is reused by the next step
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(VIA_E_DEVICE);
        off += words;
    }
    // conversion key RLev(L_CK) at (n1, q1): layout already [k][part][limb][n]
    {
        const size_t words = (size_t)L * 2 * nl * n1;
        if (!key_buf(&k->conv, words)) return fail(VIA_E_NOMEM);
        if (hipMemcpyAsync(stage, conv_key, sizeof(uint64_t) * words, hipMemcpyHostToDevice, c->stream) !=
            hipSuccess)
            return fail(VIA_E_DEVICE);
        launch_u64_to_f64(stage, k->conv, words, c->stream);
        RingDev r;
        if ((s = make_ring(c, n1, pp->q1, 2, &r))) return fail(s);
        launch_ntt(r, k->conv, (size_t)L * 2 * nl, 1, false, c->stream);
        if (k->cascade_perm) {
            // rows of n1 slots in the order of the fixed-shape conversion MAC
            launch_cascade_key_permute(k->conv, tmp_perm, (size_t)L * 2 * nl, 1, c->stream);
            if (hipMemcpyAsync(k->conv, tmp_perm, sizeof(double) * words, hipMemcpyDeviceToDevice, c->stream) !=
                hipSuccess)
                return fail(VIA_E_DEVICE);
        }
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(VIA_E_DEVICE);
    }
    // ring-switch key: d RLev(L_RSK) at (n2, q3)
    {
        const uint32_t d = n1 / pp->n2;
        const size_t words = (size_t)d * pp->l_rsk * 2 * pp->n2;
        if (!key_buf(&k->rsk, words)) return fail(VIA_E_NOMEM);
        if (hipMemcpyAsync(stage, rsk, sizeof(uint64_t) * words, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return fail(VIA_E_DEVICE);
        launch_u64_to_f64(stage, k->rsk, words, c->stream);
        RingDev r;
        uint64_t q3 = pp->q3;
        if ((s = make_ring(c, pp->n2, &q3, 1, &r))) return fail(s);
        launch_ntt(r, k->rsk, (size_t)d * pp->l_rsk * 2, 1, false, c->stream);
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(VIA_E_DEVICE);
    }
    for (size_t i = 0; i < k->cascade.size(); i++) k->bytes.push_back(key_cascade_bytes(pp, i));
    k->bytes.push_back(key_conv_bytes(pp));
    k->bytes.push_back(key_rsk_bytes(pp));
    for (size_t i = 0; i < k->repack.size(); i++) k->bytes.push_back(key_repack_bytes(pp, k->repack_nout0, i));
    (void)hipFree(stage);
    (void)hipFree(tmp_perm);
    (void)hipFree(tmp_f);
    (void)hipFree(tmp_q2);
    if (hipGetLastError() != hipSuccess) {
        stage = nullptr;
        tmp_f = tmp_q2 = nullptr;
        return fail(VIA_E_DEVICE);
    }
    *out = k;
    return VIA_OK;
}

void via_keys_destroy(via_keys *k) {
    if (!k) return;
    via_ctx *c = k->ctx;
    CallCtx cc;
    hold_ctx(c, &cc);
    (void)hipSetDevice(c->device);
    keys_free(k);
}

// ---- answer -------------------------------------------------------------------------------

}  // extern "C"

namespace {

}  // namespace

namespace {
std::mutex g_tuned_mu;
std::map<std::pair<int, int>, std::vector<StepTune>> g_tuned;  // (device, batch size)
uint64_t g_tuned_gen = 0;
}  // namespace

void via::tuned_store(int device, int Q, const std::vector<StepTune> &steps) {
    std::lock_guard<std::mutex> g(g_tuned_mu);
    g_tuned[{device, Q}] = steps;
    g_tuned_gen++;
}

via_status via::ctx_ring(via_ctx *c, uint32_t n, const uint64_t *q, int nl, RingDev *r) {
    return make_ring(c, n, q, nl, r);
}

void via::tuned_erase(int device, int Q) {
    std::lock_guard<std::mutex> g(g_tuned_mu);
    g_tuned.erase({device, Q});
    g_tuned_gen++;
}

namespace {
std::map<std::pair<int, int>, PassTune> g_pass_tuned;  // (device, batch size), under g_tuned_mu
}  // namespace

void via::pass_tuned_store(int device, int Q, const PassTune &t) {
    std::lock_guard<std::mutex> g(g_tuned_mu);
    g_pass_tuned[{device, Q}] = t;
    g_tuned_gen++;
}

bool via::pass_tuned_lookup(int device, int Q, PassTune *t) {
    std::lock_guard<std::mutex> g(g_tuned_mu);
    auto it = g_pass_tuned.upper_bound({device, Q});
    if (it == g_pass_tuned.begin()) return false;
    --it;
    if (it->first.first != device) return false;
    *t = it->second;
    return true;
}

uint64_t via::next_object_id() {
    static std::atomic<uint64_t> next{1};
    return next++;
}

uint64_t via::tuned_generation() {
    std::lock_guard<std::mutex> g(g_tuned_mu);
    return g_tuned_gen;
}

bool via::tuned_lookup(int device, int Q, std::vector<StepTune> *steps) {
    std::lock_guard<std::mutex> g(g_tuned_mu);
    auto it = g_tuned.upper_bound({device, Q});  // first entry past (device, Q)
    if (it == g_tuned.begin()) return false;
    --it;
    if (it->first.first != device) return false;
    *steps = it->second;
    return true;
}

// Client parameters agree apart from the record degree, which belongs to the
// database (the reference's N_REC is a server type parameter).
bool via::same_client_params(const via_params &a, const via_params &b) {
    return a.n1 == b.n1 && a.n2 == b.n2 && a.q1_nl == b.q1_nl && a.q1[0] == b.q1[0] && a.q1[1] == b.q1[1] &&
           a.q2 == b.q2 && a.q3 == b.q3 && a.q4 == b.q4 && a.p == b.p && a.b1 == b.b1 && a.b2 == b.b2 &&
           a.b_rsk == b.b_rsk && a.ck_base == b.ck_base && a.l_query == b.l_query && a.l_ck == b.l_ck &&
           a.l_rsk == b.l_rsk && a.num_rows == b.num_rows && a.num_cols == b.num_cols;
}

namespace {

// Step events are not recorded while a graph capture runs (the graph replays
// kernels and copies only).
#define REC_EVENT(ev, st)                                 \
    do {                                                  \
        if (!c->capturing) HIP_TRY(hipEventRecord(ev, st)); \
    } while (0)
// the step events, when the context records them
#define REC_STEP(ev, st)                                                    \
    do {                                                                    \
        if (!c->capturing && c->step_events) HIP_TRY(hipEventRecord(ev, st)); \
    } while (0)

// answer_through_crot (answer.rs:83-228) for Q queries against one database,
// every step batched across the queries; query q uses the keys of its own
// client. *res_out receives the CRot outputs [q][part][n1] at q2.
via_status prefix_impl(via_ctx *c, const via_db *db, int Q, const via_keys *const *keys,
                       const uint64_t *const *d_lwes, size_t n_lwe, via_err_detail *err, const double **res_out,
                       Workspace *carved = nullptr) {
    if (Q < 1 || Q > kMaxQ) return VIA_E_INVALID;
    via_params ep = keys[0]->pp;
    ep.n_rec = db->pp.n_rec;
    const via_params *pp = &ep;
    const size_t I = pp->num_rows, J = pp->num_cols, n = pp->n1, nl = 2, S = kSlots;
    // guards of answer_through_crot (answer.rs:131-157)
    if (!is_pow2(I) || !is_pow2(J)) return VIA_E_DIM;
    for (int q = 0; q < Q; q++)
        if (!same_client_params(keys[q]->pp, keys[0]->pp)) return VIA_E_DIM;
    for (int q = 0; q < Q; q++)
        if (!d_lwes[q] || !keys[q]) return VIA_E_INVALID;
    if (db->pp.num_rows != pp->num_rows || db->pp.num_cols != pp->num_cols || db->pp.n1 != pp->n1 ||
        db->pp.q2 != pp->q2)
        return VIA_E_DIM;
    const int nd = ilog2(I), nc = ilog2(J), nr = ilog2(pp->n1 / pp->n_rec);
    const size_t bits = (size_t)(nd + nc + nr), lq = pp->l_query;
    const size_t expected = bits * lq;
    if (n_lwe != expected) {
        if (err) {
            err->expected = expected;
            err->got = n_lwe;
            snprintf(err->msg, sizeof err->msg, "query length mismatch: expected %zu LWEs, got %zu", expected, n_lwe);
        }
        return VIA_E_QUERY_LEN;
    }
    via_status s = VIA_OK;
    if (!carved && (s = ensure_ws(c, pp, (size_t)Q, n_lwe, bits))) return s;
    Workspace &w = carved ? *carved : c->ws;
    hipStream_t st = c->stream;
    const size_t B = n_lwe, items = (size_t)Q * B;
    const size_t lw = nl * (n + 1);

    RingDev r1n;  // q1 at n1
    if ((s = make_ring(c, (uint32_t)n, pp->q1, 2, &r1n))) return s;
    RingDev r2;
    uint64_t q2 = pp->q2;
    if ((s = make_ring(c, (uint32_t)n, &q2, 1, &r2))) return s;

    REC_STEP(c->ev[0], st);
    // ---- step 1: query_decomp (query_decomp.rs:39-95) ----
    {
        U64Srcs srcs;
        memset(&srcs, 0, sizeof srcs);
        for (int q = 0; q < Q; q++) srcs.p[q] = d_lwes[q];
        launch_u64_to_f64_multi(srcs, Q, w.lwe_f, B * lw, pp->q1, (int)n, c->bad_d + c->flag_slot, st);
    }
    const Gadget gck = make_gadget(pp->ck_base, pp->l_ck);
    const DigitPlan dpk = digit_plan((uint32_t)pp->ck_base, (int)pp->l_ck);
    launch_lwe_digits(r1n, gck, dpk, w.lwe_f, (int)items, w.dig, st);
    // digit groups of each decompression step: the autotuner's measurement,
    // its stored choice for this batch size, or the heuristic
    StepTune step_tune[kDecompSteps];
    {
        const std::vector<StepTune> *tg = nullptr;
        std::vector<StepTune> found;
        if (!c->tune_override.empty()) tg = &c->tune_override;
        else if (tuned_lookup(c->device, Q, &found)) tg = &found;
        for (int i = 0; i < kDecompSteps; i++) {
            if (tg && (int)tg->size() == kDecompSteps) {
                step_tune[i] = (*tg)[i];
                step_tune[i].groups = std::min(step_tune[i].groups, w.groups_max);
            } else {
                // the heuristic of the workspace's sizing, within this gadget's depth
                step_tune[i].groups = std::min(w.groups, (int)pp->l_ck);
            }
        }
    }
    if (c->time_decomp) REC_EVENT(c->ev_dec[0], st);
    int cur = 0;
    for (uint32_t nout = 2, step = 0; nout <= n; nout <<= 1, step++) {
        RingDev r;
        if ((s = make_ring(c, nout, pp->q1, 2, &r))) return s;
        const uint32_t rank_in = (uint32_t)(2 * n / nout);
        const bool first = (nout == 2);
        MacArgs a;
        memset(&a, 0, sizeof a);
        a.dig = w.dig;
        a.dig_coeffs = S;
        a.embed = 1;
        for (int q = 0; q < Q; q++) a.key[q] = keys[q]->cascade[step];
        a.items_per_key = (int)B;
        a.key_slots = (size_t)rank_in * nout;
        a.tiles = 2;
        a.groups = step_tune[step].groups;
        a.wide = step_tune[step].wide;
        a.radix16 = step_tune[step].radix16;
        a.t512 = step_tune[step].t512;
        a.lsplit = step_tune[step].lsplit;
        a.ipw = step_tune[step].ipw;
        a.out_mask = w.pm;
        a.out_body = w.pb;
        a.reduce_body = 1;
        a.items = (int)items;
        if (keys[0]->cascade_perm) {
            if (!cascade_mac_fits(r, gck, a)) return VIA_E_INVALID;
            launch_cascade_mac(r, gck, a, st);
        } else {
            launch_gadget_mac(r, gck, a, st);
        }
        CombineArgs cb;
        cb.pm = w.pm;
        cb.pb = w.pb;
        cb.groups = step_tune[step].groups;
        cb.items = (int)items;
        const uint32_t nin = nout / 2;
        cb.body_prev = first ? w.lwe_f + n : w.body[cur];
        cb.body_item_stride = first ? lw : nl * nin;
        cb.body_limb_stride = first ? (n + 1) : nin;
        cb.masks_next = w.masks[cur ^ 1];
        cb.body_next = w.body[cur ^ 1];
        cb.dig_next = w.dig;
        launch_cascade_combine(r, gck, dpk, r1n, cb, st);
        if (c->time_decomp) REC_EVENT(c->ev_dec[step + 1], st);
        cur ^= 1;
    }
    // rlwe_to_rgsw_eval (convert.rs:162-180) with each client's conversion
    // key; the last combine left the digits of the RLWE masks in w.dig
    {
        MacArgs a;
        memset(&a, 0, sizeof a);
        a.dig = w.dig;
        a.dig_coeffs = S;
        a.embed = 0;
        for (int q = 0; q < Q; q++) a.key[q] = keys[q]->conv;
        a.items_per_key = (int)B;
        a.key_slots = n;
        a.tiles = 1;
        a.groups = step_tune[kDecompSteps - 1].groups;
        a.wide = step_tune[kDecompSteps - 1].wide;
        a.radix16 = step_tune[kDecompSteps - 1].radix16;
        a.t512 = step_tune[kDecompSteps - 1].t512;
        a.lsplit = step_tune[kDecompSteps - 1].lsplit;
        a.ipw = step_tune[kDecompSteps - 1].ipw;
        a.out_mask = w.pm;
        a.out_body = w.pb;
        a.reduce_body = 0;
        a.items = (int)items;
        if (keys[0]->cascade_perm) {
            if (!cascade_mac_fits(r1n, gck, a)) return VIA_E_INVALID;
            launch_cascade_mac(r1n, gck, a, st);
        } else {
            launch_gadget_mac(r1n, gck, a, st);
        }
        launch_r2g_combine(r1n, w.pm, w.pb, a.groups, w.masks[cur], w.body[cur], w.rgsw, (int)(Q * bits), (int)lq,
                           st);
        if (c->time_decomp) REC_EVENT(c->ev_dec[kDecompSteps], st);
    }
    const size_t rgw = 4 * lq * nl * n;  // words per RGSW at q1
    c->stage_ptr[0] = w.rgsw;
    c->stage_words[0] = bits * rgw;
    REC_STEP(c->ev[1], st);

    // ---- step 2: dmux_tree from trivial(Delta) (answer.rs:181-191, mux.rs:222-251) ----
    const Gadget g1 = make_gadget(pp->b1, pp->l_query);
    const size_t ctw1 = 2 * nl * n;  // words per ciphertext at q1
    if (nd > 1) {
        // eval form of the DMux bits 1..nd-1 of every query, one launch
        launch_ntt_groups(r1n, w.rgsw + rgw, w.rgsw_eval + rgw, (size_t)(nd - 1) * 4 * lq * nl, (size_t)Q, bits * rgw,
                          bits * rgw, 1, st);
    }
    int dc = 0;
    const u128 Q1 = (u128)pp->q1[0] * pp->q1[1];
    const u128 delta = (Q1 + pp->p - 1) / pp->p;
    if (nd > 0) {
        // level 0 acts on trivial(Delta): digits of Delta are constants
        DeltaDigits dd;
        memset(&dd, 0, sizeof dd);
        if (pp->l_query > 8) return VIA_E_INVALID;
        SMag v;
        v.neg = delta > (Q1 >> 1);
        const u128 mag = v.neg ? Q1 - delta : delta;
        v.mag = round_mul_div(mag, g1.bl, widediv_make(Q1));
        if (v.mag == 0) v.neg = 0;
        for (uint32_t k = 0; k < pp->l_query; k++) dd.d[k] = extract_digit(v, g1.base);
        dd.delta[0] = (double)(uint64_t)(delta % pp->q1[0]);
        dd.delta[1] = (double)(uint64_t)(delta % pp->q1[1]);
        launch_dmux_trivial(r1n, w.rgsw, bits * rgw, (int)pp->l_query, dd, w.dm[1], I * ctw1, Q, st);
        dc = 1;
    } else {
        std::vector<double> triv(ctw1, 0.0);
        triv[nl * n + 0] = (double)(uint64_t)(delta % pp->q1[0]);
        triv[nl * n + n] = (double)(uint64_t)(delta % pp->q1[1]);
        for (int q = 0; q < Q; q++)
            HIP_TRY(hipMemcpyAsync(w.dm[0] + (size_t)q * I * ctw1, triv.data(), sizeof(double) * ctw1,
                                   hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));  // triv is a host temporary
    }
    // trees in one persistent launch each: forced by the context's mode, else
    // the autotuner's choice for this batch size (never while capturing a
    // graph, nor on the serving loop's slots, whose streams run side by side)
    PassTune pt;
    const bool have_pt = pass_tuned_lookup(c->device, Q, &pt);
    bool pers_d = false, pers_c = false;
    if (!c->capturing && c->allow_persistent) {
        if (c->tree_override >= 0) {
            pers_d = pers_c = c->tree_override == 1;
        } else if (have_pt) {
            pers_d = pt.dmux != 0;
            pers_c = pt.cmux != 0;
        }
    }
    // the gates' limb split and FirstDim's column-block groups: the
    // autotuner's measurement, its stored choice, or the launchers' heuristics
    const int gate_split = c->split_override > -2 ? c->split_override : (have_pt ? pt.gate_split : -1);
    const int fd_groups = c->fd_groups_override > 0 ? c->fd_groups_override : (have_pt ? pt.fd_groups : 0);
    int dmux_from = 1;
    if (pers_d && nd > 2 && nd - 1 <= kMaxTreeLevels) {
        TreeArgs ta;
        memset(&ta, 0, sizeof ta);
        ta.gate.mode = GATE_DMUX;
        ta.gate.rgsw_qstride = bits * rgw;
        ta.gate.in_qstride = I * ctw1;
        ta.gate.out_qstride = I * ctw1;
        ta.gate.queries = Q;
        int x = dc;
        for (int i = 1; i < nd; i++, x ^= 1)
            ta.lev[i - 1] = TreeLevel{w.rgsw_eval + (size_t)i * rgw, w.dm[x], w.dm[x ^ 1], 1 << i};
        ta.nlev = nd - 1;
        ta.bar = c->tree_bar;
        ta.fail = c->bad_d + 2;
        ta.spin = c->tree_spin;
        if (launch_tree_persistent(r1n, g1, ta, st)) {
            dc = x;
            dmux_from = nd;
        }
    }
    for (int i = dmux_from; i < nd; i++) {
        GateArgs ga;
        ga.split = gate_split;
        ga.mode = GATE_DMUX;
        ga.rgsw = w.rgsw_eval + (size_t)i * rgw;
        ga.rgsw_qstride = bits * rgw;
        ga.in = w.dm[dc];
        ga.in_qstride = I * ctw1;
        ga.out = w.dm[dc ^ 1];
        ga.out_qstride = I * ctw1;
        ga.items_per_query = 1 << i;
        ga.queries = Q;
        ga.shift = 0;
        // the selection array is free until step 3: the split gates' partials
        ga.partials = w.sw;
        ga.partials_words = (size_t)Q * I * 2 * n;
        launch_gate(r1n, g1, ga, st);
        dc ^= 1;
    }
    c->stage_ptr[1] = w.dm[dc];
    c->stage_words[1] = I * ctw1;
    REC_STEP(c->ev[2], st);

    // ---- step 3: mod_switch_sym q1 -> q2 per row, fused with step 4's forward
    // transform of the selection vector where the degree allows ----
    const bool sw_fused = ntt_switch_fits(r1n, r2);
    if (sw_fused) {
        // the switched coefficient form is recomputed from stage 2 when asked for
        c->stage_ptr[2] = w.dm[dc];
        c->stage3_ring = r1n;
        c->stage3_q2 = pp->q2;
    } else {
        launch_modswitch_q1_q2(r1n, pp->q2, w.dm[dc], w.sw, (size_t)Q * I * 2, st);
        c->stage_ptr[2] = w.sw;
    }
    c->stage3_switch = sw_fused;
    c->resp_partials = sw_fused ? w.sw : nullptr;
    c->resp_partials_words = sw_fused ? (size_t)Q * I * 2 * n : 0;
    c->stage_words[2] = I * 2 * n;
    REC_STEP(c->ev[3], st);

    // ---- step 4: first_dim (first_dim.rs:94-184), one database pass for all queries ----
    if (sw_fused)
        launch_ntt_switch(r1n, r2, pp->q2, w.dm[dc], w.sel, (size_t)Q * I * 2, st);
    else
        launch_ntt_to(r2, w.sw, w.sel, (size_t)Q * I * 2, 1, st);
    FdPlan fp = fd_plan((int)I, (int)J, (int)n, Q);
    fp.G = fd_groups;
    const FMod m2 = fmod_make(pp->q2);
    launch_sel_pack(w.sel, w.sell, fp, st);
    REC_EVENT(c->ev_fd[0], st);
    launch_fd_mfma(db->limbs, w.sell, w.fdk, fp, m2, st);
    REC_EVENT(c->ev_fd[1], st);
    if (!c->capturing) c->have_fd_time = true;
    // segment sum, transpose to [q][j][part][k] and inverse transform, one
    // pass; in pair form when the first CMux level runs the fixed-shape gate
    // level by level (one inverse transform per pair of columns)
    const Gadget g2 = make_gadget(pp->b2, pp->l_query);
    const bool pers_tree_c = pers_c && nc > 1 && nc <= kMaxTreeLevels;
    const bool fd_pairs = nc >= 1 && !pers_tree_c && fd_pairs_fits(fp) && gate_takes_fx(r2, g2, (int)(Q * (J / 2)));
    if (fd_pairs)
        launch_fd_intt_pairs(w.fdk, w.fd, fp, r2, st);
    else
        launch_fd_intt(w.fdk, w.fd, fp, r2, st);
    c->stage_ptr[3] = w.fd;
    c->stage_words[3] = J * 2 * n;
    c->stage4_pairs = fd_pairs;
    c->stage4_ring = r2;
    REC_STEP(c->ev[4], st);

    // ---- step 5: cmux_tree over mod-switched RGSWs (answer.rs:194-203) ----
    const size_t rgw2 = 4 * lq * n;
    const size_t ctw2 = 2 * n;
    // the CMux and CRot bits of query q, mod-switched: w.rgsw2 + q * rq2
    const size_t rq2 = (size_t)(nc + nr) * rgw2;
    if (nc + nr > 0) {
        if (sw_fused) {
            launch_ntt_switch_groups(r1n, r2, pp->q2, w.rgsw + (size_t)nd * rgw, bits * rgw, w.rgsw2, rq2,
                                     (size_t)(nc + nr) * 4 * lq, Q, st);
        } else {
            launch_modswitch_q1_q2_groups(r1n, pp->q2, w.rgsw + (size_t)nd * rgw, bits * rgw, w.rgsw2, rq2,
                                          (size_t)(nc + nr) * 4 * lq, Q, st);
            launch_ntt(r2, w.rgsw2, (size_t)Q * (nc + nr) * 4 * lq, 1, false, st);
        }
    }
    const double *cm_in = w.fd;
    int cc = 0;
    int cmux_from = 0;
    if (pers_tree_c) {
        TreeArgs ta;
        memset(&ta, 0, sizeof ta);
        ta.gate.mode = GATE_CMUX;
        ta.gate.rgsw_qstride = rq2;
        ta.gate.in_qstride = J * ctw2;
        ta.gate.out_qstride = J * ctw2;
        ta.gate.queries = Q;
        const double *in = cm_in;
        int x = cc;
        for (int b = 0; b < nc; b++, x ^= 1) {
            ta.lev[b] = TreeLevel{w.rgsw2 + (size_t)b * rgw2, in, w.cm[x], (int)(J >> (b + 1))};
            in = w.cm[x];
        }
        ta.nlev = nc;
        ta.bar = c->tree_bar;
        ta.fail = c->bad_d + 2;
        ta.spin = c->tree_spin;
        if (launch_tree_persistent(r2, g2, ta, st)) {
            cm_in = in;
            cc = x;
            cmux_from = nc;
        }
    }
    for (int b = cmux_from; b < nc; b++) {
        GateArgs ga;
        ga.split = gate_split;
        ga.mode = b == 0 && fd_pairs ? GATE_CMUX_EV : GATE_CMUX;
        ga.rgsw = w.rgsw2 + (size_t)b * rgw2;
        ga.rgsw_qstride = rq2;
        ga.in = cm_in;
        ga.in_qstride = J * ctw2;
        ga.out = w.cm[cc];
        ga.out_qstride = J * ctw2;
        ga.items_per_query = (int)(J >> (b + 1));
        ga.queries = Q;
        ga.shift = 0;
        if (sw_fused) {  // the selection array held no stage for debugging
            ga.partials = w.sw;
            ga.partials_words = (size_t)Q * I * 2 * n;
        }
        launch_gate(r2, g2, ga, st);
        cm_in = w.cm[cc];
        cc ^= 1;
    }
    // gather each query's tree root into res[0] (stride 2n)
    const double *res = cm_in;
    size_t res_qstride = J * ctw2;
    c->stage_ptr[4] = cm_in;
    c->stage_words[4] = ctw2;
    REC_STEP(c->ev[5], st);

    // ---- step 6: crot SlotExtract (rotate.rs:128-149) ----
    int rc = 0;
    for (int b = 0; b < nr; b++) {
        GateArgs ga;
        ga.split = gate_split;
        ga.mode = GATE_CROT;
        ga.rgsw = w.rgsw2 + (size_t)(nc + b) * rgw2;
        ga.rgsw_qstride = rq2;
        ga.in = res;
        ga.in_qstride = res_qstride;
        ga.out = w.res[rc];
        ga.out_qstride = ctw2;
        ga.items_per_query = 1;
        ga.queries = Q;
        ga.shift = (int)(n - (1u << b));
        if (sw_fused) {
            ga.partials = w.sw;
            ga.partials_words = (size_t)Q * I * 2 * n;
        }
        launch_gate(r2, g2, ga, st);
        res = w.res[rc];
        res_qstride = ctw2;
        rc ^= 1;
    }
    if (res_qstride != ctw2) {
        for (int q = 0; q < Q; q++)
            HIP_TRY(hipMemcpyAsync(w.res[rc] + (size_t)q * ctw2, res + (size_t)q * res_qstride, sizeof(double) * ctw2,
                                   hipMemcpyDeviceToDevice, st));
        res = w.res[rc];
    }
    c->stage_ptr[5] = res;
    c->stage_words[5] = ctw2;
    REC_STEP(c->ev[6], st);

    *res_out = res;
    return VIA_OK;
}

// resp_comp (resp_comp.rs:46-76) of Q ciphertexts ct [q][part][n1] at q2,
// query q with the ring-switch key of keys[q].
via_status resp_comp_impl(via_ctx *c, const via_params *pp, int Q, const via_keys *const *keys, const double *ct,
                          uint64_t *const *d_mask, uint64_t *const *d_body) {
    DevTab t3;
    via_status s;
    if ((s = get_tab(c, pp->q3, pp->n2, &t3))) return s;
    RespArgs ra;
    memset(&ra, 0, sizeof ra);
    ra.ct = ct;
    for (int q = 0; q < Q; q++) {
        ra.rsk_eval[q] = keys[q]->rsk;
        ra.out_mask[q] = d_mask[q];
        ra.out_body[q] = d_body[q];
    }
    ra.n1 = (int)pp->n1;
    ra.n2 = (int)pp->n2;
    ra.d = (int)(pp->n1 / pp->n2);
    ra.l_rsk = (int)pp->l_rsk;
    ra.b_rsk = (uint32_t)pp->b_rsk;
    ra.q2 = pp->q2;
    ra.q3 = pp->q3;
    ra.q4 = pp->q4;
    ra.t3.fwd = t3.fwd;
    ra.t3.inv = t3.inv;
    ra.t3.n_inv = t3.n_inv;
    ra.t3.m = fmod_make(pp->q3);
    ra.queries = Q;
    ra.partials = c->resp_partials;
    ra.partials_words = c->resp_partials_words;
    launch_resp_comp(ra, c->stream);
    return VIA_OK;
}

}  // namespace

// Answers Q queries against one database in one pass over the pipeline:
// answer_through_crot then resp_comp for each (answer_one_query,
// answer.rs:289-354), every step batched across the queries.
via_status via::answer_batch_impl(via_ctx *c, const via_db *db, int Q, const via_keys *const *keys,
                                  const uint64_t *const *d_lwes, size_t n_lwe, uint64_t *const *d_mask,
                                  uint64_t *const *d_body, via_err_detail *err, Workspace *carved) {
    const double *res = nullptr;
    via_status s = prefix_impl(c, db, Q, keys, d_lwes, n_lwe, err, &res, carved);
    if (s) return s;
    if ((s = resp_comp_impl(c, &keys[0]->pp, Q, keys, res, d_mask, d_body))) return s;
    REC_STEP(c->ev[7], c->stream);
    HIP_TRY(hipGetLastError());
    c->have_times = c->step_events;
    return VIA_OK;
}

via_status via::check_vb(const via_db *db, const via_keys *keys, uint32_t T) {
    const via_params *pp = &db->pp;
    if (!is_pow2(T) || pp->n2 % T || (uint64_t)T * pp->n_rec > pp->n2) return VIA_E_INVALID;
    const uint32_t G = pp->n2 / T;
    if (G < 2 || !keys->repack_nout0 || keys->repack_nout0 > 2 * G) return VIA_E_INVALID;
    if (T > 256) return VIA_E_INVALID;
    return VIA_OK;
}

// answer_batch (batch.rs:72-163): answer_through_crot for each of the T
// queries (record degree n_rec of the database), Repack_{n2} of the T
// outputs with the client's q2 repack keys (repack.rs:331-369), resp_comp
// once.
via_status via::answer_vb_impl(via_ctx *c, const via_db *db, const via_keys *keys, int T, const uint64_t *d_lwes,
                               size_t n_lwe, uint64_t *d_mask, uint64_t *d_body, via_err_detail *err) {
    via_status s = check_vb(db, keys, (uint32_t)T);
    if (s) return s;
    via_params ep = keys->pp;
    ep.n_rec = db->pp.n_rec;
    const via_params *pp = &ep;
    const size_t n = pp->n1, S = kSlots, lw = 2 * (n + 1);
    if ((s = ensure_rw(c, pp, T, n_lwe))) return s;
    RepackWs &w = c->rw;
    hipStream_t st = c->stream;
    // steps 1-6 for every query, up to kMaxQ per batched pass
    for (int q0 = 0; q0 < T; q0 += kMaxQ) {
        const int Q = std::min(kMaxQ, T - q0);
        std::vector<const via_keys *> ks(Q, keys);
        std::vector<const uint64_t *> dl(Q);
        for (int q = 0; q < Q; q++) dl[q] = d_lwes + (size_t)(q0 + q) * n_lwe * lw;
        const double *res = nullptr;
        if ((s = prefix_impl(c, db, Q, ks.data(), dl.data(), n_lwe, err, &res))) return s;
        HIP_TRY(hipMemcpyAsync(w.rot + (size_t)q0 * 2 * n, res, sizeof(double) * Q * 2 * n, hipMemcpyDeviceToDevice,
                               st));
    }
    HIP_TRY(hipEventRecord(c->ev[7], st));
    // repack at q2 with the base-ck_base gadget of the cascade keys
    const uint32_t G = pp->n2 / (uint32_t)T;
    const Gadget g = make_gadget_q(pp->ck_base, pp->l_ck, pp->q2);
    const DigitPlan dp = digit_plan((uint32_t)pp->ck_base, (int)pp->l_ck);
    uint64_t q2 = pp->q2;
    RingDev r2n;
    if ((s = make_ring(c, (uint32_t)n, &q2, 1, &r2n))) return s;
    const int pairs0 = std::max(1, T / 2);
    launch_repack_extract(r2n, g, dp, w.rot, T, (int)G, pairs0, w.body[0], n, w.dig, st);
    const int levels = ilog2(n / G);
    int cur = 0;
    for (int lev = 0; lev < levels; lev++) {
        const uint32_t nout = (2 * G) << lev;
        const int items = std::max(1, T >> (lev + 1));
        RingDev r;
        if ((s = make_ring(c, nout, &q2, 1, &r))) return s;
        int groups = 1;
        while (groups < (int)pp->l_ck && items * 2 * groups < 512) groups++;
        MacArgs a;
        memset(&a, 0, sizeof a);
        a.dig = w.dig;
        a.dig_coeffs = 2 * S;
        a.embed = 0;
        a.key[0] = keys->repack[ilog2(nout) - ilog2(keys->repack_nout0)];
        a.items_per_key = items;
        a.key_slots = 2 * S;
        a.tiles = 2;
        a.groups = groups;
        a.out_mask = w.pm;
        a.out_body = w.pb;
        a.reduce_body = 1;
        a.items = items;
        launch_gadget_mac(r, g, a, st);
        RepackCombineArgs cb;
        memset(&cb, 0, sizeof cb);
        cb.pm = w.pm;
        cb.pb = w.pb;
        cb.groups = groups;
        cb.items = items;
        cb.body_cur = w.body[cur];
        cb.body_stride = n;
        cb.final_level = lev == levels - 1;
        cb.out = w.packed;
        cb.dig_next = w.dig;
        cb.body_next = w.body[cur ^ 1];
        cb.zero_partner = items == 1;
        launch_repack_combine(r, g, dp, cb, st);
        cur ^= 1;
    }
    c->stage_ptr[6] = w.packed;
    c->stage_words[6] = 2 * n;
    const via_keys *k1[1] = {keys};
    if ((s = resp_comp_impl(c, pp, 1, k1, w.packed, &d_mask, &d_body))) return s;
    HIP_TRY(hipGetLastError());
    c->have_times = false;
    return VIA_OK;
}

extern "C" {

// answer_batch_impl, or in graph mode the replay of its captured launches:
// a call whose database, batch size, keys, query and answer buffers, query
// length, launch shapes and workspace match a cached capture replays it; any
// other call runs eagerly (which also grows the workspace and loads the
// tables a capture cannot allocate) and then captures the same sequence for
// the next identical call. Databases of one row (the DMux root is built on
// the host) always run eagerly.
static via_status answer_graph_or_eager(via_ctx *c, const via_db *db, int Q, const via_keys *const *keys,
                                        const uint64_t *const *d_lwes, size_t n_lwe, uint64_t *const *d_mask,
                                        uint64_t *const *d_body, via_err_detail *err) {
    if (!c->graphs || db->pp.num_rows < 2) return answer_batch_impl(c, db, Q, keys, d_lwes, n_lwe, d_mask, d_body, err);
    std::vector<uintptr_t> key = {(uintptr_t)db->id, (uintptr_t)Q, (uintptr_t)n_lwe, (uintptr_t)tuned_generation(),
                                  (uintptr_t)c->ws.gen};
    for (int q = 0; q < Q; q++) {
        key.push_back((uintptr_t)keys[q]->id);
        key.push_back((uintptr_t)d_lwes[q]);
        key.push_back((uintptr_t)d_mask[q]);
        key.push_back((uintptr_t)d_body[q]);
    }
    for (auto &e : c->graph_cache)
        if (e.key == key) {
            HIP_TRY(hipGraphLaunch(e.exec, c->stream));
            c->have_times = false;
            return VIA_OK;
        }
    via_status s = answer_batch_impl(c, db, Q, keys, d_lwes, n_lwe, d_mask, d_body, err);
    if (s) return s;
    // the eager run may have grown the workspace: key the capture on the current one
    key[4] = (uintptr_t)c->ws.gen;
    via_ctx::GraphEntry e;
    e.key = key;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    s = answer_batch_impl(c, db, Q, keys, d_lwes, n_lwe, d_mask, d_body, err);
    c->capturing = false;
    const hipError_t ec = hipStreamEndCapture(c->stream, &e.graph);
    if (s || ec != hipSuccess) {
        if (e.graph) (void)hipGraphDestroy(e.graph);
        return s ? s : VIA_E_DEVICE;
    }
    if (hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(e.graph);
        return VIA_E_DEVICE;
    }
    if (c->graph_cache.size() >= 8) {  // keep the eight most recent captures
        (void)hipGraphExecDestroy(c->graph_cache.front().exec);
        (void)hipGraphDestroy(c->graph_cache.front().graph);
        c->graph_cache.erase(c->graph_cache.begin());
    }
    c->graph_cache.push_back(e);
    c->have_times = c->step_events;
    return VIA_OK;
}

static void drop_graphs(via_ctx *c) {
    for (auto &e : c->graph_cache) {
        (void)hipGraphExecDestroy(e.exec);
        (void)hipGraphDestroy(e.graph);
    }
    c->graph_cache.clear();
}

via_status via_ctx_set_step_timing(via_ctx *c, int on) {
    if (!c) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    c->step_events = on != 0;
    c->have_times = false;
    return VIA_OK;
}

via_status via_ctx_set_graphs(via_ctx *c, int on) {
    if (!c) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!on) drop_graphs(c);
    c->graphs = on != 0;
    return VIA_OK;
}

via_status via_answer_device(via_ctx *c, const via_db *db, const via_keys *keys, const uint64_t *d_lwes,
                             size_t n_lwe, uint64_t *d_mask, uint64_t *d_body, via_err_detail *err) {
    if (!c || !db || !keys || !d_mask || !d_body || (n_lwe && !d_lwes)) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    HIP_TRY(hipSetDevice(c->device));
    return answer_graph_or_eager(c, db, 1, &keys, &d_lwes, n_lwe, &d_mask, &d_body, err);
}

// Workspace of one query (B LWEs of `bits` bits) against an I x J database,
// carved from `base`. Lifetimes let buffers share space: the query-
// decomposition buffers (region A) are dead once the RGSW bits exist, so the
// DMux ping-pong (steps 2-3) and then the CMux ping-pong (step 5) start at
// region A's base; region C holds the FirstDim pass (steps 3-4); region B the
// RGSW bits and the CRot outputs, live to the end. Every kernel of one answer
// runs in stream order, so no two overlapping buffers are live at once.
// Returns the bytes the layout needs; writes w only when base is not null.
static size_t carve_ws(const via_params *pp, size_t B, size_t bits, int groups, uint8_t *base, Workspace *w) {
    const size_t n = pp->n1, S = kSlots, I = pp->num_rows, J = pp->num_cols, L = pp->l_ck;
    const FdPlan fp = fd_plan((int)I, (int)J, (int)n, 1);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    // region A: query decomposition
    const size_t a0 = off;
    const size_t o_lwe = take(8 * B * 2 * (n + 1)), o_dig = take(2 * B * L * S);
    const size_t o_pm = take(8 * (size_t)groups * B * 2 * 2 * S), o_pb = take(8 * (size_t)groups * B * 2 * 2 * S);
    size_t o_mb[4];
    for (size_t &o : o_mb) o = take(8 * B * 2 * S);
    const size_t end_dec = off;
    off = a0;  // DMux ping-pong over the dead decomposition buffers
    const size_t o_dm0 = take(8 * I * 2 * 2 * n), o_dm1 = take(8 * I * 2 * 2 * n);
    const size_t end_dm = off;
    off = a0;  // CMux ping-pong, after the DMux outputs were switched to q2
    const size_t o_cm0 = take(8 * J * 2 * n), o_cm1 = take(8 * J * 2 * n);
    off = std::max(std::max(end_dec, end_dm), off);
    // region C: mod switch and FirstDim
    const size_t o_sw = take(8 * I * 2 * n), o_sel = take(8 * I * 2 * n), o_sell = take(fd_sel_bytes(fp));
    const size_t o_fdk = take(8 * fd_out_words(fp)), o_fd = take(8 * J * 2 * n);
    // region B: RGSW bits (q1, eval form, q2) and the CRot outputs
    const size_t o_rg = take(8 * B * 4 * 2 * n), o_rge = take(8 * B * 4 * 2 * n), o_rg2 = take(8 * B * 4 * n);
    const size_t o_res0 = take(8 * 2 * n), o_res1 = take(8 * 2 * n);
    if (base) {
        *w = Workspace();
        auto D = [&](size_t o) { return (double *)(base + o); };
        w->Q = 1;
        w->B = B;
        w->I = I;
        w->J = J;
        w->bits = bits;
        w->l_ck = L;
        w->lwe_f = D(o_lwe);
        w->dig = (int16_t *)(base + o_dig);
        w->pm = D(o_pm);
        w->pb = D(o_pb);
        w->masks[0] = D(o_mb[0]);
        w->masks[1] = D(o_mb[1]);
        w->body[0] = D(o_mb[2]);
        w->body[1] = D(o_mb[3]);
        w->dm[0] = D(o_dm0);
        w->dm[1] = D(o_dm1);
        w->cm[0] = D(o_cm0);
        w->cm[1] = D(o_cm1);
        w->sw = D(o_sw);
        w->sel = D(o_sel);
        w->sell = base + o_sell;
        w->fdk = D(o_fdk);
        w->fd = D(o_fd);
        w->rgsw = D(o_rg);
        w->rgsw_eval = D(o_rge);
        w->rgsw2 = D(o_rg2);
        w->res[0] = D(o_res0);
        w->res[1] = D(o_res1);
        w->groups = w->groups_max = groups;
    }
    return off;
}

// via_answer_device for one query with no device allocation: the workspace
// is carved from this GPU's scratch array (kScratchBytes), the twiddle
// tables are looked up (never built), and the launch sequence is the one of
// via_answer_device. Calls on one GPU share the array: they must not
// overlap on the device.
via_status via_answer_device_noalloc(via_ctx *c, const via_db *db, const via_keys *keys, const uint64_t *d_lwes,
                                     size_t n_lwe, uint64_t *d_mask, uint64_t *d_body, via_err_detail *err) {
    if (!c || !db || !keys || !d_lwes || !d_mask || !d_body) return VIA_E_INVALID;
    if (!c->scratch) return VIA_E_DEVICE;
    CallCtx cc;
    hold_ctx(c, &cc);
    HIP_TRY(hipSetDevice(c->device));
    via_params ep = keys->pp;
    ep.n_rec = db->pp.n_rec;
    if (!is_pow2(ep.num_rows) || !is_pow2(ep.num_cols)) return VIA_E_DIM;
    const size_t bits = (size_t)(ilog2(ep.num_rows) + ilog2(ep.num_cols) + ilog2(ep.n1 / ep.n_rec));
    const size_t expected = bits * ep.l_query;
    if (n_lwe != expected) {  // answer.rs:132-157
        if (err) {
            err->expected = expected;
            err->got = n_lwe;
            snprintf(err->msg, sizeof err->msg, "query length mismatch: expected %zu LWEs, got %zu", expected, n_lwe);
        }
        return VIA_E_QUERY_LEN;
    }
    if (ep.num_rows < 2) {  // the trivial DMux root of a one-row grid is built from host memory
        if (err) snprintf(err->msg, sizeof err->msg, "no-allocation answers need at least two rows");
        return VIA_E_INVALID;
    }
    // the digit groups of the heuristic, fewer while the layout does not fit
    int groups = 1;
    while (groups < (int)ep.l_ck && (long)n_lwe * 2 * groups < 512) groups++;
    while (groups > 1 && carve_ws(&ep, n_lwe, bits, groups, nullptr, nullptr) > kScratchBytes) groups--;
    const size_t need = carve_ws(&ep, n_lwe, bits, groups, nullptr, nullptr);
    if (need > kScratchBytes) {
        if (err) {
            err->expected = need;
            err->got = kScratchBytes;
            snprintf(err->msg, sizeof err->msg, "workspace of %zu bytes exceeds the %zu-byte scratch array", need,
                     kScratchBytes);
        }
        return VIA_E_NOMEM;
    }
    {
        // the rings of the answer: q1 at every cascade degree, q2 at n1, q3 at
        // n2 (built by via_db_create and via_keys_create on this context)
        std::lock_guard<std::mutex> g(c->mu);
        bool ready = c->tabs.count({ep.q2, ep.n1}) && c->tabs.count({ep.q3, ep.n2});
        for (uint32_t nout = 2; nout <= ep.n1; nout <<= 1)
            ready = ready && c->tabs.count({ep.q1[0], nout}) && c->tabs.count({ep.q1[1], nout});
        if (!ready) {
            if (err) snprintf(err->msg, sizeof err->msg, "twiddle tables of these parameters are not built on this "
                                                         "context (create the database and keys on it first)");
            return VIA_E_INVALID;
        }
    }
    Workspace w;
    carve_ws(&ep, n_lwe, bits, groups, (uint8_t *)c->scratch, &w);
    c->no_alloc = true;
    via_status s = answer_batch_impl(c, db, 1, &keys, &d_lwes, n_lwe, &d_mask, &d_body, err, &w);
    c->no_alloc = false;
    // stage buffers overlap in the scratch array: not offered for debugging
    for (auto &p : c->stage_ptr) p = nullptr;
    return s;
}

via_status via_answer_many_device(via_ctx *c, const via_db *db, uint32_t nq, const via_keys *const *keys,
                                   const uint64_t *const *d_lwes, size_t n_lwe, uint64_t *const *d_masks,
                                   uint64_t *const *d_bodies, via_err_detail *err) {
    if (!c || !db || !keys || !d_lwes || !d_masks || !d_bodies || nq < 1) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    HIP_TRY(hipSetDevice(c->device));
    for (uint32_t q0 = 0; q0 < nq; q0 += kMaxQ) {
        const int Q = (int)std::min<uint32_t>(kMaxQ, nq - q0);
        via_status s = answer_graph_or_eager(c, db, Q, keys + q0, d_lwes + q0, n_lwe, d_masks + q0, d_bodies + q0, err);
        if (s) return s;
    }
    return VIA_OK;
}

// Autotuner: runs the batched answer with each candidate launch shape
// (tune_candidates) for every decompression step, `reps` timed runs each
// after one untimed run, timing each step with HIP events on the context
// stream, and keeps the fastest shape per step for this batch size.
// Stores (or, with groups null, erases) the launch shapes of batches of nq
// queries, as via_ctx_autotune would: a shape the tuner could choose.
via_status via_ctx_set_tuning(via_ctx *c, uint32_t nq, const int *groups, const int *variants) {
    if (!c || nq < 1 || nq > (uint32_t)kMaxQ) return VIA_E_INVALID;
    if (!groups) {
        tuned_erase(c->device, (int)nq);
        return VIA_OK;
    }
    if (!variants) return VIA_E_INVALID;
    std::vector<StepTune> steps(kDecompSteps);
    for (int i = 0; i < kDecompSteps; i++) {
        const int v = variants[i], gr = groups[i];
        if (v < 0 || v > 5 || gr < 1 || gr > 18) return VIA_E_INVALID;
        if (v == 1 && gr != 1 && gr != 2 && gr != 3 && gr != 5) return VIA_E_INVALID;
        if (v != 1 && gr != 1 && gr != 2 && gr != 3 && gr != 4 && gr != 6 && gr != 9 && gr != 18) return VIA_E_INVALID;
        steps[i].groups = gr;
        steps[i].wide = v == 1;
        steps[i].radix16 = v == 2;
        steps[i].t512 = v == 3 || v == 5;
        steps[i].lsplit = v >= 4;
        steps[i].ipw = variant_ipw(v);
    }
    tuned_store(c->device, (int)nq, steps);
    return VIA_OK;
}

via_status via_ctx_autotune(via_ctx *c, const via_db *db, uint32_t nq, const via_keys *const *keys,
                            const uint64_t *const *d_lwes, size_t n_lwe, uint64_t *const *d_mask,
                            uint64_t *const *d_body, int reps, int *groups_out, int *variant_out) {
    if (!c || !db || !keys || !d_lwes || !d_mask || !d_body || nq < 1 || nq > (uint32_t)kMaxQ || reps < 1)
        return VIA_E_INVALID;
    for (uint32_t q = 0; q < nq; q++)
        if (!keys[q] || !d_lwes[q] || !d_mask[q] || !d_body[q]) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    HIP_TRY(hipSetDevice(c->device));
    struct Guard {
        via_ctx *c;
        bool step_events;
        ~Guard() {
            c->time_decomp = false;
            c->tune_override.clear();
            c->tree_override = -1;
            c->fd_groups_override = 0;
            c->split_override = -2;
            c->step_events = step_events;
        }
    } guard{c, c->step_events};
    c->step_events = true;  // the candidates are timed by the step events
    const int Q = (int)nq;
    via_status s = answer_batch_impl(c, db, Q, keys, d_lwes, n_lwe, d_mask, d_body, nullptr);
    if (s) return s;
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int L = (int)keys[0]->pp.l_ck;
    std::vector<StepTune> cands;
    for (const StepTune &t : tune_candidates((int)(Q * n_lwe), L))
        if (t.groups <= c->ws.groups_max) cands.push_back(t);
    std::vector<std::vector<double>> t(cands.size(), std::vector<double>(kDecompSteps, 0.0));
    c->time_decomp = true;
    for (size_t ci = 0; ci < cands.size(); ci++) {
        c->tune_override.assign(kDecompSteps, cands[ci]);
        for (int r = 0; r <= reps; r++) {
            if ((s = answer_batch_impl(c, db, Q, keys, d_lwes, n_lwe, d_mask, d_body, nullptr))) return s;
            HIP_TRY(hipStreamSynchronize(c->stream));
            if (r == 0) continue;
            for (int i = 0; i < kDecompSteps; i++) {
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, c->ev_dec[i], c->ev_dec[i + 1]));
                t[ci][i] += ms;
            }
        }
    }
    std::vector<StepTune> best(kDecompSteps);
    for (int i = 0; i < kDecompSteps; i++) {
        size_t bi = 0;
        for (size_t ci = 1; ci < cands.size(); ci++)
            if (t[ci][i] < t[bi][i]) bi = ci;
        best[i] = cands[bi];
        if (groups_out) groups_out[i] = best[i].groups;
        if (variant_out)
            variant_out[i] = best[i].lsplit ? (best[i].t512 ? 5 : 4)
                                            : best[i].wide ? 1 : (best[i].radix16 ? 2 : (best[i].t512 ? 3 : 0));
    }
    tuned_store(c->device, Q, best);
    // the tree launch modes with the stored shapes: level by level against
    // one persistent launch per tree, by the DMux and CMux step events
    PassTune tt;
    if (c->allow_persistent) {
        c->time_decomp = false;
        c->tune_override.clear();
        double td[2] = {0, 0}, tc[2] = {0, 0};
        for (int mode = 0; mode < 2; mode++) {
            c->tree_override = mode;
            for (int r = 0; r <= reps; r++) {
                if ((s = answer_batch_impl(c, db, Q, keys, d_lwes, n_lwe, d_mask, d_body, nullptr))) return s;
                HIP_TRY(hipStreamSynchronize(c->stream));
                if (r == 0) continue;
                float a = 0, b = 0;
                HIP_TRY(hipEventElapsedTime(&a, c->ev[1], c->ev[2]));
                HIP_TRY(hipEventElapsedTime(&b, c->ev[4], c->ev[5]));
                td[mode] += a;
                tc[mode] += b;
            }
        }
        c->tree_override = -1;
        if (tree_stall_status(c) == VIA_OK) {
            tt.dmux = td[1] < td[0];
            tt.cmux = tc[1] < tc[0];
        }
    }
    pass_tuned_store(c->device, Q, tt);
    // FirstDim's column-block groups per slot, by the k_fd_mfma events
    {
        const FdPlan fp = fd_plan((int)db->pp.num_rows, (int)db->pp.num_cols, (int)db->pp.n1, Q);
        double best_t = 0;
        for (int G : {1, 2, 4, 8, 16, 32, 64}) {
            // up to one group per 4 column blocks (one per wavefront)
            if (G > 1 && 4 * G > fp.MB) continue;
            c->fd_groups_override = G;
            double t = 0;
            // one launch per answer, a few percent apart: more runs than the
            // decompression shapes get, so noise does not decide
            for (int r = 0; r <= 4 * reps; r++) {
                if ((s = answer_batch_impl(c, db, Q, keys, d_lwes, n_lwe, d_mask, d_body, nullptr))) return s;
                HIP_TRY(hipStreamSynchronize(c->stream));
                if (r == 0) continue;
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, c->ev_fd[0], c->ev_fd[1]));
                t += ms;
            }
            if (!tt.fd_groups || t < best_t) {
                tt.fd_groups = G;
                best_t = t;
            }
        }
        c->fd_groups_override = 0;
    }
    // the gates' limb split, by the DMux, CMux and CRot step events
    {
        double best_t = 0;
        int best = -1;
        for (int sp : {-1, 0, 1}) {
            c->split_override = sp;
            double t = 0;
            for (int r = 0; r <= reps; r++) {
                if ((s = answer_batch_impl(c, db, Q, keys, d_lwes, n_lwe, d_mask, d_body, nullptr))) return s;
                HIP_TRY(hipStreamSynchronize(c->stream));
                if (r == 0) continue;
                float a = 0, b = 0;
                HIP_TRY(hipEventElapsedTime(&a, c->ev[1], c->ev[2]));
                HIP_TRY(hipEventElapsedTime(&b, c->ev[4], c->ev[6]));
                t += a + b;
            }
            if (sp == -1 || t < best_t) {
                best = sp;
                best_t = t;
            }
        }
        c->split_override = -2;
        tt.gate_split = best;
    }
    pass_tuned_store(c->device, Q, tt);
    return VIA_OK;
}

via_status via_ctx_set_tree_mode(via_ctx *c, int mode) {
    if (!c || mode < -1 || mode > 1) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    c->tree_override = mode;
    return VIA_OK;
}

via_status via_ctx_tree_tuning(via_ctx *c, uint32_t nq, int *dmux, int *cmux) {
    if (!c || !dmux || !cmux) return VIA_E_INVALID;
    PassTune tt;
    if (!pass_tuned_lookup(c->device, (int)nq, &tt)) return VIA_E_INVALID;
    *dmux = tt.dmux;
    *cmux = tt.cmux;
    return VIA_OK;
}

via_status via_ctx_pass_tuning(via_ctx *c, uint32_t nq, int *fd_groups, int *gate_split) {
    if (!c || !fd_groups || !gate_split) return VIA_E_INVALID;
    PassTune tt;
    if (!pass_tuned_lookup(c->device, (int)nq, &tt)) return VIA_E_INVALID;
    *fd_groups = tt.fd_groups;
    *gate_split = tt.gate_split;
    return VIA_OK;
}

via_status via_debug_set_tree_spin(via_ctx *c, long spin) {
    if (!c || spin < 0) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    c->tree_spin = spin;
    return VIA_OK;
}

via_status via_answer_many(via_ctx *root, const via_db *db, uint32_t nq, const via_keys *const *keys,
                            const uint64_t *const *lwes, size_t n_lwe, uint64_t *const *masks, uint64_t *const *bodies,
                            via_err_detail *err) {
    if (!root || !db || !keys || !lwes || !masks || !bodies || nq < 1) return VIA_E_INVALID;
    for (uint32_t q = 0; q < nq; q++)
        if (!keys[q] || !lwes[q] || !masks[q] || !bodies[q]) return VIA_E_INVALID;
    CallCtx cc;
    via_status hs = hold_lane(root, &cc);
    if (hs) return hs;
    via_ctx *c = cc.c;
    HostFlag flag(c);
    HIP_TRY(hipSetDevice(c->device));
    // the query length follows the database's record degree
    via_params ep = keys[0]->pp;
    ep.n_rec = db->pp.n_rec;
    const via_params *pp = &ep;
    if (!is_pow2(pp->num_rows) || !is_pow2(pp->num_cols)) return VIA_E_DIM;
    const int nd = ilog2(pp->num_rows), nc = ilog2(pp->num_cols), nr = ilog2(pp->n1 / pp->n_rec);
    const size_t bits = (size_t)(nd + nc + nr);
    const size_t expected = bits * pp->l_query;
    if (n_lwe != expected) {  // answer.rs:132-157
        if (err) {
            err->expected = expected;
            err->got = n_lwe;
            snprintf(err->msg, sizeof err->msg, "query length mismatch: expected %zu LWEs, got %zu", expected, n_lwe);
        }
        return VIA_E_QUERY_LEN;
    }
    const size_t lw = (size_t)2 * (pp->n1 + 1);
    for (uint32_t q0 = 0; q0 < nq; q0 += kMaxQ) {
        const int Q = (int)std::min<uint32_t>(kMaxQ, nq - q0);
        via_status s = ensure_ws(c, pp, (size_t)Q, n_lwe, bits);
        if (s) return s;
        Workspace &w = c->ws;
        std::vector<const uint64_t *> dl(Q);
        std::vector<uint64_t *> dm(Q), dbp(Q);
        for (int q = 0; q < Q; q++) {
            HIP_TRY(hipMemcpyAsync(w.lwe_u + (size_t)q * lw * n_lwe, lwes[q0 + q], sizeof(uint64_t) * lw * n_lwe,
                                   hipMemcpyHostToDevice, c->stream));
            dl[q] = w.lwe_u + (size_t)q * lw * n_lwe;
            dm[q] = w.ans_u + (size_t)q * 2048;
            dbp[q] = w.ans_u + (size_t)q * 2048 + 1024;
        }
        s = answer_batch_impl(c, db, Q, keys + q0, dl.data(), n_lwe, dm.data(), dbp.data(), err);
        if (s) return s;
        for (int q = 0; q < Q; q++) {
            HIP_TRY(hipMemcpyAsync(masks[q0 + q], dm[q], sizeof(uint64_t) * pp->n2, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipMemcpyAsync(bodies[q0 + q], dbp[q], sizeof(uint64_t) * pp->n2, hipMemcpyDeviceToHost, c->stream));
        }
        HIP_TRY(hipStreamSynchronize(c->stream));
        if ((s = flag.status())) {
            if (err) snprintf(err->msg, sizeof err->msg, "query word not a canonical residue mod q1");
            return s;
        }
        if ((s = tree_stall_status(c))) {
            if (err) snprintf(err->msg, sizeof err->msg, "persistent tree barrier gave up waiting");
            return s;
        }
    }
    return VIA_OK;
}

via_status via_answer(via_ctx *c, const via_db *db, const via_keys *keys, const uint64_t *lwes, size_t n_lwe,
                      uint64_t *ans_mask, uint64_t *ans_body, via_err_detail *err) {
    if (!c || !db || !keys || !ans_mask || !ans_body || (n_lwe && !lwes)) return VIA_E_INVALID;
    return via_answer_many(c, db, 1, &keys, &lwes, n_lwe, &ans_mask, &ans_body, err);
}

via_status via_answer_batch_device(via_ctx *c, const via_db *db, const via_keys *keys, uint32_t T,
                                   const uint64_t *d_lwes, size_t n_lwe, uint64_t *d_ans_mask, uint64_t *d_ans_body,
                                   via_err_detail *err) {
    if (!c || !db || !keys || !d_lwes || !d_ans_mask || !d_ans_body) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    HIP_TRY(hipSetDevice(c->device));
    return answer_vb_impl(c, db, keys, (int)T, d_lwes, n_lwe, d_ans_mask, d_ans_body, err);
}

via_status via_answer_batch(via_ctx *root, const via_db *db, const via_keys *keys, uint32_t T,
                            const uint64_t *const *lwes, const size_t *n_lwe, size_t n_queries, uint64_t *ans_mask,
                            uint64_t *ans_body, via_err_detail *err) {
    if (!root || !db || !keys || !ans_mask || !ans_body || (n_queries && (!lwes || !n_lwe))) return VIA_E_INVALID;
    for (size_t t = 0; t < n_queries; t++)
        if (!lwes[t]) return VIA_E_INVALID;
    via_status s = check_vb(db, keys, T);
    if (s) return s;
    if (n_queries != T) {  // batch.rs:124-128
        if (err) {
            err->expected = T;
            err->got = n_queries;
            snprintf(err->msg, sizeof err->msg, "answer_batch: batch.len() must equal T");
        }
        return VIA_E_DIM;
    }
    via_params ep = keys->pp;
    ep.n_rec = db->pp.n_rec;
    if (!is_pow2(ep.num_rows) || !is_pow2(ep.num_cols)) return VIA_E_DIM;
    const size_t bits = (size_t)(ilog2(ep.num_rows) + ilog2(ep.num_cols) + ilog2(ep.n1 / ep.n_rec));
    const size_t expected = bits * ep.l_query;
    for (uint32_t t = 0; t < T; t++)
        if (n_lwe[t] != expected) {  // answer.rs:132-157, per inner query
            if (err) {
                err->expected = expected;
                err->got = n_lwe[t];
                snprintf(err->msg, sizeof err->msg, "query %u: expected %zu LWEs, got %zu", t, expected, n_lwe[t]);
            }
            return VIA_E_QUERY_LEN;
        }
    CallCtx cc;
    if ((s = hold_lane(root, &cc))) return s;
    via_ctx *c = cc.c;
    HostFlag flag(c);
    HIP_TRY(hipSetDevice(c->device));
    if ((s = ensure_rw(c, &ep, (int)T, expected))) return s;
    RepackWs &w = c->rw;
    const size_t lw = (size_t)2 * (ep.n1 + 1);
    for (uint32_t t = 0; t < T; t++)
        HIP_TRY(hipMemcpyAsync(w.lwe_u + (size_t)t * expected * lw, lwes[t], sizeof(uint64_t) * expected * lw,
                               hipMemcpyHostToDevice, c->stream));
    if ((s = answer_vb_impl(c, db, keys, (int)T, w.lwe_u, expected, w.ans_u, w.ans_u + ep.n2, err))) return s;
    HIP_TRY(hipMemcpyAsync(ans_mask, w.ans_u, sizeof(uint64_t) * ep.n2, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(ans_body, w.ans_u + ep.n2, sizeof(uint64_t) * ep.n2, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if ((s = flag.status())) {
        if (err) snprintf(err->msg, sizeof err->msg, "query word not a canonical residue mod q1");
        return s;
    }
    if ((s = tree_stall_status(c))) {
        if (err) snprintf(err->msg, sizeof err->msg, "persistent tree barrier gave up waiting");
        return s;
    }
    return VIA_OK;
}

via_status via_debug_stage(via_ctx *c, int stage, uint64_t *out, size_t words) {
    if (!c || stage < 1 || stage > 7 || !out) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    const double *p = c->stage_ptr[stage - 1];
    const size_t total = c->stage_words[stage - 1];
    if (!p || words < total) return VIA_E_INVALID;
    uint64_t *tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, sizeof(uint64_t) * total));
    double *coef = nullptr;
    if (stage == 3 && c->stage3_switch) {
        // the answer switched and transformed in one pass: switch stage 2 again
        if (hipMalloc(&coef, sizeof(double) * total) != hipSuccess) {
            (void)hipFree(tmp);
            return VIA_E_NOMEM;
        }
        launch_modswitch_q1_q2(c->stage3_ring, c->stage3_q2, p, coef, total / kSlots, c->stream);
        p = coef;
    }
    if (stage == 4 && c->stage4_pairs) {
        // FirstDim outputs kept in pair form: back to coefficient form
        const int J = (int)(total / (2 * kSlots));
        if (hipMalloc(&coef, sizeof(double) * total) != hipSuccess) {
            (void)hipFree(tmp);
            return VIA_E_NOMEM;
        }
        launch_fd_unpair(p, coef, J, c->stage4_ring, c->stream);
        p = coef;
    }
    launch_f64_to_u64(p, tmp, total, c->stream);
    HIP_TRY(hipMemcpyAsync(out, tmp, sizeof(uint64_t) * total, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree(tmp));
    if (coef) HIP_TRY(hipFree(coef));
    return VIA_OK;
}

via_status via_debug_set_exact_paths(int on) {
    set_exact_paths_only(on != 0);
    return VIA_OK;
}

// ---- ring-level entry points -----------------------------------------------------------

static via_status ring_for(via_ctx *c, uint32_t n, const uint64_t *moduli, uint32_t nmod, RingDev *r) {
    if (!is_pow2(n) || n < 2 || n > (uint32_t)kSlots || nmod < 1 || nmod > 2) return VIA_E_INVALID;
    for (uint32_t i = 0; i < nmod; i++)
        if (moduli[i] >= kMaxModulus || (moduli[i] - 1) % (2ULL * n)) return VIA_E_INVALID;
    return make_ring(c, n, moduli, (int)nmod, r);
}

via_status via_poly_mul_device(via_ctx *c, uint32_t n, const uint64_t *moduli, uint32_t nmod, size_t group,
                               const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_c, size_t count) {
    if (!c || !moduli || !group) return VIA_E_INVALID;
    CallCtx cc;
    hold_ctx(c, &cc);
    HIP_TRY(hipSetDevice(c->device));
    RingDev r;
    via_status s = ring_for(c, n, moduli, nmod, &r);
    if (s) return s;
    if (n > (uint32_t)kSlots) return VIA_E_INVALID;
    launch_polymul_u64(r, d_a, d_b, d_c, count, group, c->stream);
    HIP_TRY(hipGetLastError());
    return VIA_OK;
}

via_status via_ntt(via_ctx *root, uint32_t n, const uint64_t *moduli, uint32_t nmod, size_t group, uint64_t *polys,
                   size_t count, int inverse) {
    if (!root || !moduli || !polys || !group) return VIA_E_INVALID;
    CallCtx cc;
    if (hold_lane(root, &cc)) return VIA_E_DEVICE;
    via_ctx *c = cc.c;
    HIP_TRY(hipSetDevice(c->device));
    RingDev r;
    via_status s = ring_for(c, n, moduli, nmod, &r);
    if (s) return s;
    const size_t P = kSlots / n;
    if (nmod == 2 && group % P) return VIA_E_INVALID;
    const size_t words = count * n;
    uint64_t *u = nullptr;
    double *d = nullptr;
    HIP_TRY(hipMalloc(&u, sizeof(uint64_t) * words));
    HIP_TRY(hipMalloc(&d, sizeof(double) * words));
    HIP_TRY(hipMemcpyAsync(u, polys, sizeof(uint64_t) * words, hipMemcpyHostToDevice, c->stream));
    launch_u64_to_f64(u, d, words, c->stream);
    launch_ntt(r, d, count, group, inverse != 0, c->stream);
    launch_f64_to_u64(d, u, words, c->stream);
    HIP_TRY(hipMemcpyAsync(polys, u, sizeof(uint64_t) * words, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree(u));
    HIP_TRY(hipFree(d));
    return VIA_OK;
}

// FirstDim alone over host buffers (first_dim.rs:94-184 batched over
// queries): switched [q][I][part][n1] -> out [q][J][part][n1], both
// coefficient form at q2; the same kernels as the answer path.
via_status via_first_dim(via_ctx *root, const via_db *db, uint32_t nq, const uint64_t *switched, uint64_t *out) {
    if (!root || !db || !switched || !out || nq < 1 || nq > (uint32_t)kMaxQ) return VIA_E_INVALID;
    if (db->ctx != root) return VIA_E_INVALID;
    CallCtx cc;
    if (hold_lane(root, &cc)) return VIA_E_DEVICE;
    via_ctx *c = cc.c;
    HIP_TRY(hipSetDevice(c->device));
    const via_params *pp = &db->pp;
    via_status s = ensure_ws(c, pp, nq, 1, 1);
    if (s) return s;
    Workspace &w = c->ws;
    const size_t I = pp->num_rows, J = pp->num_cols, n = pp->n1;
    const size_t in_words = (size_t)nq * I * 2 * n, out_words = (size_t)nq * J * 2 * n;
    RingDev r2;
    uint64_t q2 = pp->q2;
    if ((s = make_ring(c, (uint32_t)n, &q2, 1, &r2))) return s;
    uint64_t *u = nullptr;
    HIP_TRY(hipMalloc(&u, sizeof(uint64_t) * std::max(in_words, out_words)));
    hipStream_t st = c->stream;
    HIP_TRY(hipMemcpyAsync(u, switched, sizeof(uint64_t) * in_words, hipMemcpyHostToDevice, st));
    launch_u64_to_f64(u, w.sw, in_words, st);
    launch_ntt_to(r2, w.sw, w.sel, (size_t)nq * I * 2, 1, st);
    const FdPlan fp = fd_plan((int)I, (int)J, (int)n, (int)nq);
    const FMod m2 = fmod_make(pp->q2);
    launch_sel_pack(w.sel, w.sell, fp, st);
    launch_fd_mfma(db->limbs, w.sell, w.fdk, fp, m2, st);
    launch_fd_intt(w.fdk, w.fd, fp, r2, st);
    launch_f64_to_u64(w.fd, u, out_words, st);
    HIP_TRY(hipMemcpyAsync(out, u, sizeof(uint64_t) * out_words, hipMemcpyDeviceToHost, st));
    const hipError_t e = hipStreamSynchronize(st);
    (void)hipFree(u);
    if (e != hipSuccess) return VIA_E_DEVICE;
    HIP_TRY(hipGetLastError());
    return VIA_OK;
}

via_status via_device_alloc(via_ctx *c, size_t bytes, void **d) {
    if (!c || !d) return VIA_E_INVALID;
    HIP_TRY(hipSetDevice(c->device));
    if (hipMalloc(d, bytes ? bytes : 1) != hipSuccess) return VIA_E_NOMEM;
    return VIA_OK;
}

via_status via_device_free(via_ctx *c, void *d) {
    if (!c) return VIA_E_INVALID;
    HIP_TRY(hipFree(d));
    return VIA_OK;
}

via_status via_copy_to_device(via_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c) return VIA_E_INVALID;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VIA_OK;
}

via_status via_copy_device(via_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c) return VIA_E_INVALID;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return VIA_OK;
}

via_status via_copy_to_host(via_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c) return VIA_E_INVALID;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VIA_OK;
}

}  // extern "C"
