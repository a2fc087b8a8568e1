// prg.hip — PrgCompressed mask regeneration (via-protocol wire/format.rs:
// 143-160) on the GPU.
//
// A stream is Shake256Prg(seed) (sampling/prg.rs): block k = SHAKE-256(seed
// || u64_le(k)), 136 bytes each (keccak.h). Draw d is uniform_below(q_d)
// (prg.rs:153-200) with q_d = q[d % nmod] (one modulus, or the two limbs of
// an RNS coefficient, limb 0 first): attempts of nbytes little-endian bytes,
// masked to the bit length of q_d, rejected when >= q_d. With equal nbytes
// for every modulus, attempt a always reads bytes [a*nbytes, a*nbytes +
// nbytes); which draw an accepted attempt serves depends on the acceptances
// before it. That is a two-state automaton (the modulus of the pending
// draw) counting acceptances, so a workgroup per stream scans it: each
// thread summarises its run of attempts as a map start state -> (end state,
// acceptances), a Hillis-Steele scan composes the maps across the 1024
// threads, and each thread re-walks its run from its exact start, writing
// the draws it accepts.
#include "internal.h"
#include "keccak.h"

namespace via {

// Stream s's blocks b (grid: blocks x streams) into stream + (s*blocks + b)*17.
__global__ __launch_bounds__(256) void k_prg_blocks(PrgSeeds seeds, uint64_t *stream, size_t blocks) {
    const int s = blockIdx.y;
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= blocks) return;
    uint64_t lanes[17];
    prg_block(seeds.bytes[s], seeds.len[s], (uint64_t)b, lanes);
    uint64_t *o = stream + ((size_t)s * blocks + b) * 17;
#pragma unroll
    for (int i = 0; i < 17; i++) o[i] = lanes[i];
}

__device__ __forceinline__ uint64_t prg_attempt(const uint64_t *st, size_t a, int nbytes) {
    uint64_t v = 0;
    for (int j = 0; j < nbytes; j++) {
        const size_t i = a * nbytes + j;  // byte index in the stream
        const size_t blk = i / kPrgBlock, off = i - blk * kPrgBlock;
        v |= ((st[blk * 17 + (off >> 3)] >> (8 * (off & 7))) & 0xFF) << (8 * j);
    }
    return v;
}

__device__ __forceinline__ bool prg_accept(uint64_t v, int p, const PrgPlan &pl) {
    return (v & pl.mask[p]) < pl.q[p];
}

// One workgroup (1024 threads) per stream.
__global__ __launch_bounds__(1024) void k_prg_draws(const uint64_t *stream, PrgPlan pl, uint64_t *out, int *short_flag) {
    constexpr int TPB = 1024;
    __shared__ int se[2][2][TPB];       // [buffer][start state] -> end state
    __shared__ unsigned sc[2][2][TPB];  // acceptances
    const int s = blockIdx.x, t = threadIdx.x;
    const uint64_t *st = stream + (size_t)s * pl.blocks * 17;
    const size_t per = (pl.attempts + TPB - 1) / TPB;
    const size_t a0 = min(pl.attempts, (size_t)t * per), a1 = min(pl.attempts, a0 + per);
    // phase 1: this thread's run as a map of the start state
    int e[2] = {0, 1 % pl.nmod};
    unsigned c[2] = {0, 0};
    for (size_t a = a0; a < a1; a++) {
        const uint64_t v = prg_attempt(st, a, pl.nbytes);
#pragma unroll
        for (int p = 0; p < 2; p++) {
            if (p >= pl.nmod) continue;
            if (prg_accept(v, e[p], pl)) {
                c[p]++;
                e[p] = (e[p] + 1) % pl.nmod;
            }
        }
    }
    se[0][0][t] = e[0];
    se[0][1][t] = e[1];
    sc[0][0][t] = c[0];
    sc[0][1][t] = c[1];
    __syncthreads();
    // phase 2: inclusive scan of the maps (earlier runs first)
    int cur = 0;
    for (int off = 1; off < TPB; off <<= 1) {
        const int nxt = cur ^ 1;
#pragma unroll
        for (int p = 0; p < 2; p++) {
            if (t >= off) {
                const int mid = se[cur][p][t - off];
                se[nxt][p][t] = se[cur][mid][t];
                sc[nxt][p][t] = sc[cur][p][t - off] + sc[cur][mid][t];
            } else {
                se[nxt][p][t] = se[cur][p][t];
                sc[nxt][p][t] = sc[cur][p][t];
            }
        }
        __syncthreads();
        cur = nxt;
    }
    // every draw has an attempt (the host sized the stream with slack)
    if (t == TPB - 1 && sc[cur][0][t] < pl.count) atomicOr(short_flag, 1);
    // phase 3: re-walk the run from its exact start, writing the accepted draws
    int p = t ? se[cur][0][t - 1] : 0;
    size_t d = t ? sc[cur][0][t - 1] : 0;
    uint64_t *o = out + (size_t)s * pl.out_sstride;
    for (size_t a = a0; a < a1 && d < pl.count; a++) {
        const uint64_t v = prg_attempt(st, a, pl.nbytes) & pl.mask[p];
        if (v >= pl.q[p]) continue;
        size_t idx = d;
        if (pl.lwe) {
            const size_t rowlen = (size_t)pl.n * pl.nmod;
            const size_t row = d / rowlen, rem = d - row * rowlen;
            idx = row * pl.row_stride + (rem % pl.nmod) * pl.limb_stride + rem / pl.nmod;
        }
        o[idx] = v;
        d++;
        p = (p + 1) % pl.nmod;
    }
}

// Blocks first .. first + count - 1 of one stream (seed index 0) into
// out[(b - first) * 17 ...]: a window of the byte stream for host-side
// parsing (the client key generation).
__global__ __launch_bounds__(256) void k_prg_window(PrgSeeds seeds, uint64_t first, size_t count, uint64_t *out) {
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= count) return;
    uint64_t lanes[17];
    prg_block(seeds.bytes[0], seeds.len[0], first + b, lanes);
    uint64_t *o = out + b * 17;
#pragma unroll
    for (int i = 0; i < 17; i++) o[i] = lanes[i];
}

void launch_prg_window(const PrgSeeds &seeds, uint64_t first, size_t count, uint64_t *out, hipStream_t st) {
    if (count)
        hipLaunchKernelGGL(k_prg_window, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, seeds, first, count,
                           out);
}

void launch_prg_draws(const PrgSeeds &seeds, int streams, const PrgPlan &pl, uint64_t *stream, uint64_t *out,
                      int *short_flag, hipStream_t st) {
    const dim3 g1((unsigned)((pl.blocks + 255) / 256), streams);
    hipLaunchKernelGGL(k_prg_blocks, g1, dim3(256), 0, st, seeds, stream, pl.blocks);
    hipLaunchKernelGGL(k_prg_draws, dim3(streams), dim3(1024), 0, st, stream, pl, out, short_flag);
}

}  // namespace via
