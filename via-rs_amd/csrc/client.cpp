// client.cpp — client-side recovery (via-client), host code.
//
// Client::recover (client.rs:222-239) is decrypt_asymmetric (via-primitives
// encryption/rlwe.rs:187-224) with the client's S2 at (n2, q3): the mask A'
// times S2 in Z_q3[X]/(X^n2 + 1), each coefficient rescaled to q4 as
// floor((v * q4 + floor(q3/2)) / q3), B' minus that at q4, then decode
// (rlwe.rs:296-330): centre c to (-q4/2, q4/2], m = floor((p*c + floor(q4/2))
// / q4) mod p (Euclidean). deinterleave_batch (batch.rs:31-52) then reads
// record t of a VIA-B plaintext at coefficients d*c + t, d = n2 / n_rec.
#include <cstring>
#include <vector>

#include "../../include/via_gpu.h"

typedef unsigned __int128 u128;
typedef __int128 i128;

// c = A' * S2 in Z_q[X]/(X^n + 1), schoolbook, accumulated in Acc: the sums
// of the products landing at k (pos) and wrapping past X^n (neg), reduced
// once per coefficient.
template <typename Acc>
static void negacyclic_mul(size_t n, uint64_t q, const uint64_t *a, const uint64_t *s, uint64_t *c) {
    std::vector<Acc> pos(n, 0), neg(n, 0);
    for (size_t i = 0; i < n; i++) {
        const Acc ai = a[i];
        if (!ai) continue;
        for (size_t j = 0; j < n - i; j++) pos[i + j] += ai * s[j];
        for (size_t j = n - i; j < n; j++) neg[i + j - n] += ai * s[j];
    }
    for (size_t k = 0; k < n; k++) {
        const uint64_t x = (uint64_t)(pos[k] % q), y = (uint64_t)(neg[k] % q);
        c[k] = x >= y ? x - y : x + (q - y);
    }
}

via_status via_client_recover(const via_params *pp, const uint64_t *sk2, const uint64_t *mask, const uint64_t *body,
                              uint64_t *plain) {
    if (!pp || !sk2 || !mask || !body || !plain) return VIA_E_INVALID;
    const size_t n = pp->n2;
    const uint64_t q3 = pp->q3, q4 = pp->q4, p = pp->p;
    if (n == 0 || (n & (n - 1)) || q3 < 2 || q4 < 2 || p < 2 || p > q4) return VIA_E_INVALID;
    for (size_t i = 0; i < n; i++)
        if (sk2[i] >= q3 || mask[i] >= q3 || body[i] >= q4) return VIA_E_INVALID;
    // A' * S2 (the centred lift of S2 at q3, taken back to q3, is S2 itself);
    // u64 sums when n (q3 - 1)^2 fits (the paper q3 = 8380417: 2^55 at n = 512)
    std::vector<uint64_t> prod(n);
    const u128 bound = (u128)n * (q3 - 1) * (q3 - 1);
    if (bound < ((u128)1 << 64)) negacyclic_mul<uint64_t>(n, q3, mask, sk2, prod.data());
    else negacyclic_mul<u128>(n, q3, mask, sk2, prod.data());
    for (size_t k = 0; k < n; k++) {
        // rescale to q4: floor((v q4 + floor(q3/2)) / q3) mod q4
        const uint64_t sw = (uint64_t)((((u128)prod[k] * q4 + q3 / 2) / q3) % q4);
        const uint64_t noisy = body[k] >= sw ? body[k] - sw : body[k] + (q4 - sw);
        // decode: centre at q4, floor((p c + floor(q4/2)) / q4) mod p, Euclidean
        const i128 c = noisy > q4 / 2 ? (i128)noisy - (i128)q4 : (i128)noisy;
        const i128 num = (i128)p * c + (i128)(q4 / 2);
        i128 rounded = num / (i128)q4;
        if (num % (i128)q4 < 0) rounded -= 1;
        i128 m = rounded % (i128)p;
        if (m < 0) m += p;
        plain[k] = (uint64_t)m;
    }
    return VIA_OK;
}

via_status via_client_deinterleave(const via_params *pp, uint32_t T, const uint64_t *plain, uint64_t *records) {
    if (!pp || !plain || !records || T == 0) return VIA_E_INVALID;
    const size_t n2 = pp->n2, nr = pp->n_rec;
    if (nr == 0 || n2 % nr || (size_t)T * nr > n2) return VIA_E_INVALID;
    const size_t d = n2 / nr;
    for (size_t t = 0; t < T; t++)
        for (size_t c = 0; c < nr; c++) records[t * nr + c] = plain[d * c + t];
    return VIA_OK;
}
