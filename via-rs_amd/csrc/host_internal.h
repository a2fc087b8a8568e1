// host_internal.h — host-side objects shared by the answer path (via_gpu.cpp)
// and the serving loop (serve.cpp). Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <condition_variable>
#include <utility>
#include <vector>

#include "../../include/via_gpu.h"
#include "internal.h"

struct DevTab {
    double *fwd = nullptr, *inv = nullptr;
    double n_inv = 0;
    uint64_t q = 0;
};

struct Workspace {
    // sizes the buffers were allocated for; l_ck sizes the digit planes and
    // partial-sum planes, so a parameter set with a deeper cascade gadget
    // regrows them
    size_t Q = 0, B = 0, I = 0, J = 0, bits = 0, l_ck = 0;
    uint64_t gen = 0;  // incremented whenever the buffers are reallocated
    double *lwe_f = nullptr, *masks[2] = {nullptr, nullptr}, *body[2] = {nullptr, nullptr};
    double *pm = nullptr, *pb = nullptr, *rgsw = nullptr, *rgsw_eval = nullptr;
    double *dm[2] = {nullptr, nullptr}, *sw = nullptr, *sel = nullptr, *fd = nullptr, *fdk = nullptr;
    double *cm[2] = {nullptr, nullptr}, *rgsw2 = nullptr, *res[2] = {nullptr, nullptr};
    uint64_t *lwe_u = nullptr, *ans_u = nullptr;
    int16_t *dig = nullptr;
    uint8_t *sell = nullptr;  // packed selection limbs of a FirstDim pass
    int groups = 1;      // heuristic digit groups per gadget product
    int groups_max = 1;  // the partial-sum planes pm / pb hold this many groups
    std::vector<void *> all;
};

// Gadget-product steps of query decompression: the eleven cascade steps and
// the RLWE -> RGSW conversion.
constexpr int kDecompSteps = 12;

// Launch shape of one decompression step: digit groups (workgroups per item
// and tile along the digits) and the kernel variant (wide: the digit
// transforms of a group side by side; radix16: the 256-thread kernel with
// radix-16 rather than radix-8 transform passes).
struct StepTune {
    int groups = 1;
    int wide = 0;
    int radix16 = 0;
    int t512 = 0;  // the 512-thread kernel (radix-8 passes)
    int lsplit = 0;  // one workgroup per RNS limb (with the 256- or 512-thread kernel)
    int ipw = 2;     // fixed-shape cascade MAC: items (LWEs of one key) per workgroup, 1, 2 or 4
};

// The fixed-shape cascade MAC's items per workgroup that a launch-shape
// variant (0..5) selects: 1, 2 or 4 for variant mod 3 = 0, 1, 2.
inline int variant_ipw(int variant) { return variant % 3 == 0 ? 1 : variant % 3 == 1 ? 2 : 4; }

// Buffers of the VIA-B repack, sized for batches of up to T queries.
struct RepackWs {
    int T = 0;
    size_t n_lwe = 0;
    size_t l_ck = 0, n2 = 0;  // gadget depth and answer degree the buffers hold
    double *rot = nullptr;      // [T][part][n1] post-CRot ciphertexts at q2
    double *packed = nullptr;   // [part][n1] repacked RLWE
    double *body[2] = {nullptr, nullptr};  // [pair][n1] interleaved bodies
    double *pm = nullptr, *pb = nullptr;   // gadget-product partials
    int16_t *dig = nullptr;     // [pair][k][2*n1]
    uint64_t *lwe_u = nullptr;  // staging for host queries [T][n_lwe][limb][n1+1]
    uint64_t *ans_u = nullptr;  // [part][n2] staging for the host answer
    std::vector<void *> all;
};

struct via_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<std::pair<uint64_t, uint32_t>, DevTab> tabs;
    std::mutex mu;
    // One call at a time uses this context's stream and workspace. Server::
    // answer(&self) is re-entrant in the reference (answer.rs:534-554): the
    // host-buffer entry points borrow a lane (a child context with its own
    // stream and workspace, sharing the parent's twiddle tables) when the
    // context is busy, so concurrent callers overlap on the device; the
    // device-buffer entry points wait for the context itself, whose stream
    // their caller synchronises on.
    std::recursive_mutex call_mu;
    via_ctx *parent = nullptr;
    std::mutex lane_mu;
    std::vector<via_ctx *> lanes;
    // Mapped host flag set by the query conversion when a word is not a
    // canonical residue (the word is taken as 0); read after the call's
    // synchronisation, or by via_ctx_synchronize for device entry points.
    // Slot 0 serves the device-buffer entry points, slot 1 the host-buffer
    // ones (which clear and read it within the call); flag_slot names the
    // slot of the call in progress.
    int *bad_h = nullptr;
    int *bad_d = nullptr;
    int flag_slot = 0;
    // the device address of this GPU's no-allocation scratch array, and
    // whether the call in progress must not allocate (twiddle tables are
    // then looked up, never built)
    void *scratch = nullptr;
    bool no_alloc = false;
    // persistent trees (k_tree_persistent): the barrier words, whether this
    // context may use them (not the serving loop's slots, whose streams run
    // side by side), and a forced mode (-1: the tuned choice, 0: one launch
    // per level, 1: one launch per tree). bad_h[2] is the barrier-stall flag.
    unsigned *tree_bar = nullptr;
    bool allow_persistent = true;
    int tree_override = -1;
    int fd_groups_override = 0;  // autotuner measurement: FirstDim column-block groups (0: none)
    int split_override = -2;     // autotuner measurement: gate limb split (-2: none)
    long tree_spin = 1L << 24;  // polls before a barrier wait gives up
    Workspace ws;
    RepackWs rw;
    hipEvent_t ev[8] = {};
    hipEvent_t ev_fd[2] = {};
    hipEvent_t ev_mark[2] = {};  // via_ctx_mark start / stop
    hipEvent_t ev_dec[kDecompSteps + 1] = {};  // around each decompression step (autotuner)
    bool time_decomp = false;
    // the step events ev[] of every answer (via_ctx_set_step_timing); each
    // record leaves the device idle for a few microseconds between kernels
    bool step_events = true;
    std::vector<StepTune> tune_override;  // set while the autotuner measures
    // hipGraph mode (via_ctx_set_graphs): captured batched answers keyed by
    // every pointer and size they bake in; `capturing` suppresses the step
    // events while a capture runs
    bool graphs = false;
    bool capturing = false;
    struct GraphEntry {
        std::vector<uintptr_t> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<GraphEntry> graph_cache;
    bool have_times = false;     // step events of the last answer are valid
    bool have_fd_time = false;   // the FirstDim events of the last database pass are valid
    // pointers to the last answer's stage buffers; the FirstDim outputs
    // (stage 4) may be in the pair form of launch_fd_intt_pairs, over the
    // ring recorded beside them
    const double *stage_ptr[7] = {};
    size_t stage_words[7] = {};
    bool stage4_pairs = false;
    via::RingDev stage4_ring{};
    // stage 3 (the q1 -> q2 switch) may be fused into the next transform:
    // then its pointer is stage 2's array, switched again when asked for
    bool stage3_switch = false;
    // the selection-vector array of the last prefix when nothing else holds
    // it after step 4 (switch fused into the transform): RespComp's partials
    double *resp_partials = nullptr;
    size_t resp_partials_words = 0;
    via::RingDev stage3_ring{};
    uint64_t stage3_q2 = 0;
};

struct via_db {
    uint64_t id = 0;  // unique per process (captured graphs key on it)
    via_ctx *ctx = nullptr;
    via_params pp;
    // eval form at q2 as 7-bit limb fragments of the matrix-core FirstDim
    // (first_dim.hip), 5 bytes per residue
    uint8_t *limbs = nullptr;
};

struct via_keys {
    uint64_t id = 0;  // unique per process (captured graphs key on it)
    via_ctx *ctx = nullptr;
    via_params pp;
    std::vector<double *> cascade;  // per step, [k][part][limb][rank_in*n_out]
    // cascade and conversion key rows in the slot order of k_cascade_mac
    // (cascade_key_pos within each 2048-slot tile), set when its exactness
    // bound holds for these parameters (cascade_fixed_ok); otherwise slot order
    bool cascade_perm = false;
    double *conv = nullptr;         // [k][part][limb][n1]
    double *rsk = nullptr;          // [j][k][part][n2]
    // VIA-B repack keys: the cascade steps of output degree repack_nout0..n1
    // mod-switched to q2, eval form, [k][part][rank_in*n_out] per step
    std::vector<double *> repack;
    uint32_t repack_nout0 = 0;
    std::vector<size_t> bytes;
};


namespace via {
// Launch shapes chosen by via_ctx_autotune, shared by every context of the
// process on the same device (the serving loop's slots included): per
// device and batch size, the twelve decompression steps. A batch size that
// was not tuned takes the largest tuned size below it.
void tuned_erase(int device, int Q);
// Device ring of degree n over nl NTT-friendly moduli (twiddle tables cached
// on the context).
via_status ctx_ring(via_ctx *c, uint32_t n, const uint64_t *q, int nl, RingDev *r);
void tuned_store(int device, int Q, const std::vector<StepTune> &steps);
bool tuned_lookup(int device, int Q, std::vector<StepTune> *steps);
// Launch choices via_ctx_autotune makes per device and batch size beyond
// the decompression shapes (the same lookup rule): persistent DMux and CMux
// trees (one launch per tree) or level by level; FirstDim's column-block
// groups per slot (0: the heuristic); the gates' limb split (-1: split the
// two-limb levels narrower than the CU count, 0: never, 1: always).
struct PassTune {
    int dmux = 0, cmux = 0;
    int fd_groups = 0;
    int gate_split = -1;
};
void pass_tuned_store(int device, int Q, const PassTune &t);
bool pass_tuned_lookup(int device, int Q, PassTune *t);
// Incremented by every tuned_store: captured graphs bake in launch shapes.
uint64_t tuned_generation();
// Process-wide unique ids for databases and key sets.
uint64_t next_object_id();
// Answers Q <= kMaxQ queries (device buffers) against one database on the
// context's stream, asynchronously.
// carved: a workspace laid out by the caller (the no-allocation entry point);
// null uses the context's own, grown on demand.
via_status answer_batch_impl(via_ctx *c, const via_db *db, int Q, const via_keys *const *keys,
                             const uint64_t *const *d_lwes, size_t n_lwe, uint64_t *const *d_mask,
                             uint64_t *const *d_body, via_err_detail *err, Workspace *carved = nullptr);
// VIA-B: T queries of one client (d_lwes holds them back to back), repacked
// into one answer, on the context's stream, asynchronously.
via_status answer_vb_impl(via_ctx *c, const via_db *db, const via_keys *keys, int T, const uint64_t *d_lwes,
                          size_t n_lwe, uint64_t *d_mask, uint64_t *d_body, via_err_detail *err);
// Checks a VIA-B batch size against the database and keys (VIA_E_INVALID).
via_status check_vb(const via_db *db, const via_keys *keys, uint32_t T);
// Grows the context workspace for Q queries of n_lwe LWEs.
via_status ensure_workspace(via_ctx *c, const via_params *pp, size_t Q, size_t n_lwe);
// Grows the context's VIA-B repack workspace for batches of T queries.
via_status ensure_repack_workspace(via_ctx *c, const via_params *pp, int T, size_t n_lwe);
size_t keys_device_bytes(const via_keys *k);
// Client parameters agree apart from the record degree.
bool same_client_params(const via_params &a, const via_params &b);
// Device bytes via_keys_create allocates for keys of these parameters.
size_t keys_footprint(const via_params *pp);
// Bytes of the per-device scratch array of via_answer_device_noalloc.
constexpr size_t kScratchBytes = (size_t)96 << 20;
// The device address of this GPU's scratch array (noalloc.hip), looked up
// once per device; `device` must be current.
via_status scratch_setup(int device, void **ptr);
// Largest VIA-B batch the serving loop stages.
constexpr uint32_t kServeMaxT = 64;
// True when every word of n_lwe LWEs ([lwe][limb][n1 + 1]) is a canonical
// residue of its q1 limb.
bool lwes_canonical(const via_params *pp, const uint64_t *lwes, size_t n_lwe);
// The same check fused with a copy of the words to dst (one pass).
bool lwes_copy_canonical(const via_params *pp, uint64_t *dst, const uint64_t *lwes, size_t n_lwe);
// Holds the context, or a lane of it when the context is busy (created on
// demand, up to seven), for a host-buffer call on the calling thread;
// lane_release (same thread) lets it go.
via_status lane_acquire(via_ctx *root, via_ctx **lane);
void lane_release(via_ctx *lane);
// Builds the twiddle tables of every ring the answer path uses on this context.
via_status prepare_tables(via_ctx *c, const via_params *pp);
}  // namespace via
