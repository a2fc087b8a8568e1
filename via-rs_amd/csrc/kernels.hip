// kernels.hip — the VIA-C server answer path on gfx950.
//
// Every kernel computes exactly the integers the reference computes; the
// comments name the via-rs function each one restates.
#include "ntt_device.h"

namespace via {

// ---- gadget digits ---------------------------------------------------------------

// Centred lift and gadget scale of one coefficient (gadget.rs:115-151):
// val = round-half-away(c_centred * B^L / Q), returned as sign and magnitude.
template <int NL>
__device__ __forceinline__ SMag gadget_scale(double a0, double a1, const RingDev &r, const Gadget &g) {
    SMag v;
    if (NL == 2) {
        via_u128 x = rns_reconstruct(a0, a1, r.rns);
        via_u128 Q = r.rns.big_q;
        int neg = x > (Q >> 1);
        via_u128 mag = neg ? Q - x : x;
        v.mag = round_mul_div(mag, g.bl, r.qdiv);
        v.neg = neg && v.mag != 0;
    } else {
        uint64_t q = r.q_single;
        uint64_t x = (uint64_t)a0;
        int neg = x > (q >> 1);
        uint64_t mag = neg ? q - x : x;
        // (mag * B^L + floor(q/2)) / q fits in 64 bits (checked by the host).
        uint64_t bl = (uint64_t)g.bl;
        v.mag = (via_u128)((mag * bl + (q >> 1)) / q);
        v.neg = neg && v.mag != 0;
    }
    return v;
}

__device__ __forceinline__ double digit_residue(int32_t d, const FMod &m) { return d >= 0 ? (double)d : (double)d + m.q; }

// ---- gadget multiply-accumulate ------------------------------------------------------

// RLevEval::gadget_product (rlev.rs:259-283) without the final inverse:
// acc = sum_k NTT(digit_k) (.) key[L-1-k] for mask and body. One workgroup
// handles kSlots coefficient slots of one item (a tile); the digit of every
// slot is extracted in registers, LSB first, one level per NTT round.
template <int NL>
__global__ __launch_bounds__(kThreads) void k_gadget_mac(RingDev r, Gadget g, MacArgs a) {
    __shared__ double s[NL * kSlots];
    const int tile = blockIdx.x, item = blockIdx.y;
    const int S = kSlots;
    SMag v[kPer];
    const double *in = a.in + (size_t)item * a.in_item_stride;
#pragma unroll
    for (int e = 0; e < kPer; e++) {
        const int sg = threadIdx.x + e * kThreads;
        double a0 = 0.0, a1 = 0.0;
        bool zero = false;
        size_t idx;
        if (a.embed) {
            zero = (sg & 1) != 0;
            idx = (size_t)tile * (S / 2) + (sg >> 1);
        } else {
            idx = (size_t)tile * S + sg;
        }
        if (!zero) {
            a0 = in[idx];
            if (NL == 2) a1 = in[a.in_limb_stride + idx];
        }
        v[e] = gadget_scale<NL>(a0, a1, r, g);
    }
    double am[NL][kPer], ab[NL][kPer];
#pragma unroll
    for (int l = 0; l < NL; l++)
#pragma unroll
        for (int e = 0; e < kPer; e++) am[l][e] = ab[l][e] = 0.0;
    const double *key = a.key + (size_t)item * a.key_item_stride + (size_t)tile * S;
    const FMod m0 = r.t[0].m, m1 = r.t[1].m;
    for (int k = 0; k < g.L; k++) {
#pragma unroll
        for (int e = 0; e < kPer; e++) {
            const int sg = threadIdx.x + e * kThreads;
            const int32_t d = extract_digit(v[e], g.base);
            s[sg] = digit_residue(d, m0);
            if (NL == 2) s[S + sg] = digit_residue(d, m1);
        }
        __syncthreads();
        lds_ntt_fwd<NL>(s, S, r);
        const int ks = g.L - 1 - k;
#pragma unroll
        for (int l = 0; l < NL; l++) {
            const FMod md = l ? m1 : m0;
            const double *km = key + ((size_t)(ks * 2 + 0) * NL + l) * a.key_slots;
            const double *kb = key + ((size_t)(ks * 2 + 1) * NL + l) * a.key_slots;
#pragma unroll
            for (int e = 0; e < kPer; e++) {
                const int sg = threadIdx.x + e * kThreads;
                const double x = s[l * S + sg];
                am[l][e] = mod_add(am[l][e], mod_mul(x, km[sg], md), md);
                ab[l][e] = mod_add(ab[l][e], mod_mul(x, kb[sg], md), md);
            }
        }
        __syncthreads();
    }
    const size_t ot = (size_t)item * a.tiles + tile;
#pragma unroll
    for (int l = 0; l < NL; l++)
#pragma unroll
        for (int e = 0; e < kPer; e++) {
            const int sg = threadIdx.x + e * kThreads;
            a.out_mask[(ot * NL + l) * S + sg] = am[l][e];
            if (!a.reduce_body) a.out_body[(ot * NL + l) * S + sg] = ab[l][e];
            else s[l * S + sg] = ab[l][e];
        }
    if (!a.reduce_body) return;
    __syncthreads();
    // Sum the S/n segment bodies of the tile (exact modular additions).
    const int n = r.n;
    for (int half = (S / n) >> 1; half >= 1; half >>= 1) {
        for (int t = threadIdx.x; t < NL * half * n; t += blockDim.x) {
            const int l = t / (half * n);
            const int rem = t - l * half * n;
            const FMod md = l ? m1 : m0;
            s[l * S + rem] = mod_add(s[l * S + rem], s[l * S + half * n + rem], md);
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < NL * n; t += blockDim.x) {
        const int l = t / n, c = t - l * n;
        a.out_body[(ot * NL + l) * n + c] = s[l * S + c];
    }
}

void launch_gadget_mac(const RingDev &r, const Gadget &g, const MacArgs &a, hipStream_t st) {
    dim3 grid(a.tiles, a.items);
    if (r.nl == 2)
        hipLaunchKernelGGL(k_gadget_mac<2>, grid, dim3(kThreads), 0, st, r, g, a);
    else
        hipLaunchKernelGGL(k_gadget_mac<1>, grid, dim3(kThreads), 0, st, r, g, a);
}

// ---- cascade combine (conv.rs:62-95) --------------------------------------------

template <int NL>
__global__ __launch_bounds__(kThreads) void k_cascade_combine(RingDev r, CombineArgs a) {
    __shared__ double s[NL * kSlots];
    const int item = blockIdx.x;
    const int S = kSlots, n = r.n;
    const FMod m0 = r.t[0].m, m1 = r.t[1].m;
    // masks: -iNTT(pm[tile 0] + pm[tile 1])
    const double *pm0 = a.pm + (size_t)item * 2 * NL * S;
    const double *pm1 = pm0 + (size_t)NL * S;
    for (int t = threadIdx.x; t < NL * S; t += blockDim.x) {
        const FMod md = (t >= S) ? m1 : m0;
        s[t] = mod_add(pm0[t], pm1[t], md);
    }
    __syncthreads();
    lds_ntt_inv_noscale<NL>(s, S, r);
    for (int t = threadIdx.x; t < NL * S; t += blockDim.x) {
        const int l = t >= S ? 1 : 0;
        const FMod md = l ? m1 : m0;
        a.masks_next[(size_t)item * NL * S + t] = mod_neg(mod_mul(s[t], l ? r.t[1].n_inv : r.t[0].n_inv, md), md);
    }
    __syncthreads();
    // body: embed(body_prev) - iNTT(pb[tile 0] + pb[tile 1]); the limbs are
    // packed at stride n so the NL-limb routine runs on one poly per limb.
    const double *pb0 = a.pb + (size_t)item * 2 * NL * n;
    const double *pb1 = pb0 + (size_t)NL * n;
    for (int t = threadIdx.x; t < NL * n; t += blockDim.x) {
        const FMod md = (t >= n) ? m1 : m0;
        s[t] = mod_add(pb0[t], pb1[t], md);
    }
    __syncthreads();
    lds_ntt_inv_noscale<NL>(s, n, r);
    for (int t = threadIdx.x; t < NL * n; t += blockDim.x) {
        const int l = t >= n ? 1 : 0, c = t - l * n;
        const FMod md = l ? m1 : m0;
        const double x = mod_mul(s[t], l ? r.t[1].n_inv : r.t[0].n_inv, md);
        double e = 0.0;
        if ((c & 1) == 0) e = a.body_prev[(size_t)item * a.body_item_stride + (size_t)l * a.body_limb_stride + (c >> 1)];
        a.body_next[((size_t)item * NL + l) * n + c] = mod_sub(e, x, md);
    }
}

void launch_cascade_combine(const RingDev &r, const CombineArgs &a, hipStream_t st) {
    if (r.nl == 2)
        hipLaunchKernelGGL(k_cascade_combine<2>, dim3(a.items), dim3(kThreads), 0, st, r, a);
    else
        hipLaunchKernelGGL(k_cascade_combine<1>, dim3(a.items), dim3(kThreads), 0, st, r, a);
}

// ---- rlwe_to_rgsw combine (convert.rs:162-180) -----------------------------------

template <int NL>
__global__ __launch_bounds__(kThreads) void k_r2g_combine(RingDev r, const double *pm, const double *pb,
                                                          const double *rlwe_masks, const double *rlwe_bodies,
                                                          double *rgsw, int lq) {
    __shared__ double s[NL * kSlots];
    const int item = blockIdx.x;  // bit * lq + i
    const int bit = item / lq, i = item - bit * lq;
    const int n = r.n;  // == kSlots
    const FMod m0 = r.t[0].m, m1 = r.t[1].m;
    const size_t pw = (size_t)NL * n;
    double *g = rgsw + (size_t)bit * 4 * lq * pw;
    double *nsm = g + (size_t)i * 2 * pw;            // neg_s_m[i]
    double *mm = g + (size_t)(lq + i) * 2 * pw;      // m[i]
    const double *rm = rlwe_masks + (size_t)item * pw, *rb = rlwe_bodies + (size_t)item * pw;
    for (int part = 0; part < 2; part++) {
        const double *src = (part ? pb : pm) + (size_t)item * pw;
        for (int t = threadIdx.x; t < (int)pw; t += blockDim.x) s[t] = src[t];
        __syncthreads();
        lds_ntt_inv_noscale<NL>(s, n, r);
        for (int t = threadIdx.x; t < (int)pw; t += blockDim.x) {
            const int l = t >= n ? 1 : 0;
            const FMod md = l ? m1 : m0;
            const double x = mod_mul(s[t], l ? r.t[1].n_inv : r.t[0].n_inv, md);
            if (part == 0) nsm[t] = mod_add(rb[t], x, md);
            else nsm[pw + t] = x;
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < (int)pw; t += blockDim.x) {
        mm[t] = rm[t];
        mm[pw + t] = rb[t];
    }
}

void launch_r2g_combine(const RingDev &r, const double *pm, const double *pb, const double *rlwe_masks,
                        const double *rlwe_bodies, double *rgsw, int bits, int lq, hipStream_t st) {
    if (r.nl == 2)
        hipLaunchKernelGGL(k_r2g_combine<2>, dim3(bits * lq), dim3(kThreads), 0, st, r, pm, pb, rlwe_masks,
                           rlwe_bodies, rgsw, lq);
    else
        hipLaunchKernelGGL(k_r2g_combine<1>, dim3(bits * lq), dim3(kThreads), 0, st, r, pm, pb, rlwe_masks,
                           rlwe_bodies, rgsw, lq);
}

// ---- gates: external product with DMux / CMux / CRot combine -------------------

// Coefficient sg of the gate input ct for item j (see launch_gate).
template <int NL>
__device__ __forceinline__ void gate_input(int mode, const double *in, int j, int part, int l, int sg, int n,
                                           int shift, const FMod &md, double &base_v, double &ct_v) {
    const size_t pw = (size_t)NL * n;
    if (mode == GATE_DMUX) {
        const double x = in[(size_t)j * 2 * pw + part * pw + (size_t)l * n + sg];
        base_v = x;
        ct_v = x;
    } else if (mode == GATE_CMUX) {
        const double x0 = in[(size_t)(2 * j) * 2 * pw + part * pw + (size_t)l * n + sg];
        const double x1 = in[(size_t)(2 * j + 1) * 2 * pw + part * pw + (size_t)l * n + sg];
        base_v = x0;
        ct_v = mod_sub(x1, x0, md);
    } else {
        // CRot SlotExtract (rotate.rs:128-149): rot = -(res * X^shift),
        // ct = rot - res, out = res + RGSW (x) ct.
        const double *p = in + part * pw + (size_t)l * n;
        const double x = p[sg];
        double rv;
        if (sg >= shift) rv = p[sg - shift];
        else rv = mod_neg(p[sg - shift + n], md);
        const double rneg = mod_neg(rv, md);
        base_v = x;
        ct_v = mod_sub(rneg, x, md);
    }
}

// RGSWCiphertext::external_product (rgsw.rs:108-120) over an eval-form RGSW
// [half][i][part][limb][n], then dmux (mux.rs:129-137), cmux (mux.rs:75-84)
// or the SlotExtract step of crot. One workgroup per output item, n == kSlots.
template <int NL>
__global__ __launch_bounds__(kThreads) void k_gate(RingDev r, Gadget g, int mode, const double *rgsw,
                                                   const double *in, double *out, int shift) {
    __shared__ double s[NL * kSlots];
    const int j = blockIdx.x;
    const int n = r.n, S = kSlots;
    const size_t pw = (size_t)NL * n;
    const FMod m0 = r.t[0].m, m1 = r.t[1].m;
    double am[NL][kPer], ab[NL][kPer];
#pragma unroll
    for (int l = 0; l < NL; l++)
#pragma unroll
        for (int e = 0; e < kPer; e++) am[l][e] = ab[l][e] = 0.0;
    for (int half = 0; half < 2; half++) {  // 0: ct.mask with -s*m, 1: ct.body with m
        SMag v[kPer];
#pragma unroll
        for (int e = 0; e < kPer; e++) {
            const int sg = threadIdx.x + e * kThreads;
            double b0, c0, b1 = 0.0, c1 = 0.0;
            gate_input<NL>(mode, in, j, half, 0, sg, n, shift, m0, b0, c0);
            if (NL == 2) gate_input<NL>(mode, in, j, half, 1, sg, n, shift, m1, b1, c1);
            v[e] = gadget_scale<NL>(c0, c1, r, g);
        }
        const double *key = rgsw + (size_t)half * g.L * 2 * pw;
        for (int k = 0; k < g.L; k++) {
#pragma unroll
            for (int e = 0; e < kPer; e++) {
                const int sg = threadIdx.x + e * kThreads;
                const int32_t d = extract_digit(v[e], g.base);
                s[sg] = digit_residue(d, m0);
                if (NL == 2) s[S + sg] = digit_residue(d, m1);
            }
            __syncthreads();
            lds_ntt_fwd<NL>(s, S, r);
            const int ks = g.L - 1 - k;
#pragma unroll
            for (int l = 0; l < NL; l++) {
                const FMod md = l ? m1 : m0;
                const double *km = key + (size_t)ks * 2 * pw + (size_t)l * n;
                const double *kb = km + pw;
#pragma unroll
                for (int e = 0; e < kPer; e++) {
                    const int sg = threadIdx.x + e * kThreads;
                    const double x = s[l * S + sg];
                    am[l][e] = mod_add(am[l][e], mod_mul(x, km[sg], md), md);
                    ab[l][e] = mod_add(ab[l][e], mod_mul(x, kb[sg], md), md);
                }
            }
            __syncthreads();
        }
    }
    // inverse transforms of the product, then the gate combine
    for (int part = 0; part < 2; part++) {
#pragma unroll
        for (int l = 0; l < NL; l++)
#pragma unroll
            for (int e = 0; e < kPer; e++) s[l * S + threadIdx.x + e * kThreads] = part ? ab[l][e] : am[l][e];
        __syncthreads();
        lds_ntt_inv_noscale<NL>(s, S, r);
#pragma unroll
        for (int l = 0; l < NL; l++) {
            const FMod md = l ? m1 : m0;
            const double ni = l ? r.t[1].n_inv : r.t[0].n_inv;
#pragma unroll
            for (int e = 0; e < kPer; e++) {
                const int sg = threadIdx.x + e * kThreads;
                const double p = mod_mul(s[l * S + sg], ni, md);
                double bv, cv;
                gate_input<NL>(mode, in, j, part, l, sg, n, shift, md, bv, cv);
                const size_t o = (size_t)part * pw + (size_t)l * n + sg;
                if (mode == GATE_DMUX) {
                    out[(size_t)(2 * j) * 2 * pw + o] = mod_sub(cv, p, md);
                    out[(size_t)(2 * j + 1) * 2 * pw + o] = p;
                } else {
                    out[(size_t)j * 2 * pw + o] = mod_add(bv, p, md);
                }
            }
        }
        __syncthreads();
    }
}

void launch_gate(const RingDev &r, const Gadget &g, int mode, const double *rgsw_eval, const double *in,
                 double *out, int items, int rot_shift, hipStream_t st) {
    if (r.nl == 2)
        hipLaunchKernelGGL(k_gate<2>, dim3(items), dim3(kThreads), 0, st, r, g, mode, rgsw_eval, in, out, rot_shift);
    else
        hipLaunchKernelGGL(k_gate<1>, dim3(items), dim3(kThreads), 0, st, r, g, mode, rgsw_eval, in, out, rot_shift);
}

// ---- mod switch q1 -> q2 (mod_switch.rs:57-86) ----------------------------------

__global__ void k_modswitch_q1_q2(RingDev r1, uint64_t q2, WideDiv qd, const double *in, double *out,
                                  size_t polys) {
    const int n = r1.n;
    const size_t total = polys * n;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t p = t / n, c = t - p * n;
        const double a0 = in[p * 2 * n + c], a1 = in[p * 2 * n + n + c];
        const via_u128 x = rns_reconstruct(a0, a1, r1.rns);
        uint64_t y = (uint64_t)round_mul_div(x, (via_u128)q2, qd);
        if (y >= q2) y -= q2;
        out[t] = (double)y;
    }
}

void launch_modswitch_q1_q2(const RingDev &r1, uint64_t q2, const double *in, double *out, size_t polys,
                            hipStream_t st) {
    const size_t total = polys * r1.n;
    if (!total) return;
    unsigned blocks = (unsigned)((total + 255) / 256);
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_modswitch_q1_q2, dim3(blocks), dim3(256), 0, st, r1, q2, r1.qdiv, in, out, polys);
}

// ---- FirstDim (first_dim.rs:94-184) -------------------------------------------------

constexpr int kFdCols = 4;

__global__ __launch_bounds__(256) void k_first_dim(const double *db, const double *sel, double *out, int I, int J,
                                                   int n, FMod m) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j0 = blockIdx.y * kFdCols;
    if (k >= n) return;
    double am[kFdCols], ab[kFdCols];
#pragma unroll
    for (int t = 0; t < kFdCols; t++) am[t] = ab[t] = 0.0;
    for (int i = 0; i < I; i++) {
        const double sm = sel[(size_t)(2 * i) * n + k];
        const double sb = sel[(size_t)(2 * i + 1) * n + k];
        const double *row = db + ((size_t)i * J + j0) * n + k;
#pragma unroll
        for (int t = 0; t < kFdCols; t++) {
            if (j0 + t < J) {
                const double d = row[(size_t)t * n];
                am[t] = mod_add(am[t], mod_mul(d, sm, m), m);
                ab[t] = mod_add(ab[t], mod_mul(d, sb, m), m);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < kFdCols; t++) {
        if (j0 + t < J) {
            out[((size_t)(j0 + t) * 2 + 0) * n + k] = am[t];
            out[((size_t)(j0 + t) * 2 + 1) * n + k] = ab[t];
        }
    }
}

void launch_first_dim(const double *db, const double *sel_eval, double *out_eval, int I, int J, int n, FMod m,
                      hipStream_t st) {
    dim3 grid((n + 255) / 256, (J + kFdCols - 1) / kFdCols);
    hipLaunchKernelGGL(k_first_dim, grid, dim3(256), 0, st, db, sel_eval, out_eval, I, J, n, m);
}

// ---- RespComp (resp_comp.rs:46-76) -----------------------------------------------------

__global__ __launch_bounds__(1024) void k_resp_comp(RespArgs a) {
    extern __shared__ double sm[];
    double *c3m = sm;           // [n1] mask at q3
    double *buf = sm + a.n1;    // [n2] NTT buffer
    const int n1 = a.n1, n2 = a.n2, d = a.d;
    const int i = threadIdx.x;  // one slot of the n2 ring per thread
    const uint64_t q2 = a.q2, q3 = a.q3, q4 = a.q4;
    const FMod m3 = a.t3.m;
    // mod_switch_sym q2 -> q3 at n1
    for (int c = threadIdx.x; c < n1; c += blockDim.x) {
        const uint64_t x = (uint64_t)a.ct[c];
        c3m[c] = (double)(((x * q3 + (q2 >> 1)) / q2) % q3);
    }
    const uint64_t xb = (uint64_t)a.ct[n1 + d * i];
    const double body_proj = (double)(((xb * q3 + (q2 >> 1)) / q2) % q3);
    __syncthreads();
    RingDev r;
    r.nl = 1;
    r.n = n2;
    r.logn = 31 - __clz(n2);
    r.t[0] = a.t3;
    r.t[1] = a.t3;
    uint64_t bl = 1;
    for (int k = 0; k < a.l_rsk; k++) bl *= a.b_rsk;
    double am = 0.0, ab = 0.0;
    const size_t kw = (size_t)a.l_rsk * 2 * n2;
    for (int j = 0; j < d; j++) {
        // (mask * X^j).project_at(0) (ring_switch.rs:358-385)
        const int idx = d * i - j;
        const double pv = idx >= 0 ? c3m[idx] : mod_neg(c3m[idx + n1], m3);
        const uint64_t x = (uint64_t)pv;
        const int neg = x > (q3 >> 1);
        const uint64_t mag = neg ? q3 - x : x;
        SMag v;
        v.mag = (via_u128)((mag * bl + (q3 >> 1)) / q3);
        v.neg = neg && v.mag != 0;
        const double *key = a.rsk_eval + (size_t)j * kw;
        for (int k = 0; k < a.l_rsk; k++) {
            const int32_t dg = extract_digit(v, a.b_rsk);
            buf[i] = digit_residue(dg, m3);
            __syncthreads();
            lds_ntt_fwd<1>(buf, n2, r);
            const int ks = a.l_rsk - 1 - k;
            const double y = buf[i];
            am = mod_add(am, mod_mul(y, key[(size_t)ks * 2 * n2 + i], m3), m3);
            ab = mod_add(ab, mod_mul(y, key[(size_t)ks * 2 * n2 + n2 + i], m3), m3);
            __syncthreads();
        }
    }
    double res[2];
    for (int part = 0; part < 2; part++) {
        buf[i] = part ? ab : am;
        __syncthreads();
        lds_ntt_inv_noscale<1>(buf, n2, r);
        res[part] = mod_mul(buf[i], a.t3.n_inv, m3);
        __syncthreads();
    }
    const uint64_t mask = (uint64_t)mod_neg(res[0], m3);
    const uint64_t body = (uint64_t)mod_sub(body_proj, res[1], m3);
    // mod_switch_asym (mod_switch.rs:124-154): mask q3 -> q3, body q3 -> q4
    a.out_mask[i] = ((mask * q3 + (q3 >> 1)) / q3) % q3;
    a.out_body[i] = ((body * q4 + (q3 >> 1)) / q3) % q4;
}

void launch_resp_comp(const RespArgs &a, hipStream_t st) {
    const size_t shmem = (size_t)(a.n1 + a.n2) * sizeof(double);
    hipLaunchKernelGGL(k_resp_comp, dim3(1), dim3(a.n2), shmem, st, a);
}

// ---- setup: key relayout and DB packing ----------------------------------------------

__global__ void k_key_relayout(const uint64_t *src, double *dst, int rank_in, int L, int nl, int n) {
    const size_t total = (size_t)rank_in * L * 2 * nl * n;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        // src index t = (((idx*L + k)*2 + part)*nl + l)*n + c
        size_t x = t;
        const int c = (int)(x % n);
        x /= n;
        const int l = (int)(x % nl);
        x /= nl;
        const int part = (int)(x % 2);
        x /= 2;
        const int k = (int)(x % L);
        const size_t idx = x / L;
        const size_t o = (((size_t)k * 2 + part) * nl + l) * ((size_t)rank_in * n) + idx * n + c;
        dst[o] = (double)src[t];
    }
}

void launch_key_relayout(const uint64_t *src, double *dst, int rank_in, int L, int nl, int n, hipStream_t st) {
    const size_t total = (size_t)rank_in * L * 2 * nl * n;
    unsigned blocks = (unsigned)((total + 255) / 256);
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_key_relayout, dim3(blocks), dim3(256), 0, st, src, dst, rank_in, L, nl, n);
}

__global__ void k_setup_db(const uint64_t *records, const uint8_t *bytes, size_t n_records, double *cells, int I,
                           int J, int n1, int n_rec, uint64_t p) {
    const int d = n1 / n_rec;
    const size_t total = (size_t)I * J * n1;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t cell = t / n1;
        const int pos = (int)(t - cell * n1);
        const int c = pos / d, k = pos - c * d;
        const size_t i = cell / J, j = cell - i * J;
        const size_t flat = (size_t)k * I * J + i * J + j;
        double v = 0.0;
        if (flat < n_records) {
            if (records) {
                v = (double)(records[flat * n_rec + c] % p);
            } else {
                const uint8_t b = bytes[flat * (n_rec / 2) + (c >> 1)];
                v = (double)((c & 1) ? (b >> 4) : (b & 0xF));
            }
        }
        cells[t] = v;
    }
}

void launch_setup_db(const uint64_t *records, size_t n_records, double *cells, int I, int J, int n1, int n_rec,
                     uint64_t p, hipStream_t st) {
    const size_t total = (size_t)I * J * n1;
    unsigned blocks = (unsigned)((total + 255) / 256);
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_setup_db, dim3(blocks), dim3(256), 0, st, records, (const uint8_t *)nullptr, n_records,
                       cells, I, J, n1, n_rec, p);
}

void launch_setup_db_bytes(const uint8_t *bytes, size_t n_records, double *cells, int I, int J, int n1, int n_rec,
                           hipStream_t st) {
    const size_t total = (size_t)I * J * n1;
    unsigned blocks = (unsigned)((total + 255) / 256);
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_setup_db, dim3(blocks), dim3(256), 0, st, (const uint64_t *)nullptr, bytes, n_records,
                       cells, I, J, n1, n_rec, (uint64_t)16);
}

}  // namespace via
