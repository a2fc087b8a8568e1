"""VIA-B pieces of the CPU restatement, pinned on the reference's own tests.

* embed_d / mlwes_insert interleave: repack.rs:823-870 (hand vector).
* deinterleave_batch strided map: via-client/src/batch.rs:63-77 (hand vector).
* extr round trips: extract.rs:119-176 (D = 1 and D = 2 recover M_0, (M_0, M_4)).
* bit_reverse_index: repack.rs:222-230 (T = 8 permutation, an involution).
* Toy VIA-B batch e2e: batch_e2e_toy.rs:40-225 (n1 = 64, n3 = 2, T = 8, both
  index sets of the reference), every record recovered in batch order.
* Paper VIA-B batch e2e: batch_e2e_paper.rs:38-174 (T = 8, n3 = 64, I = J = 2,
  seed b"via-b-paper-batch-e2e", the reference's eight indices).

The VIA-B answer bytes themselves have no reference vector (the reference
has none for any VIA-B configuration): they are pinned by these round trips
and by the KAT-pinned pieces they are built from (conv step, key switch,
mod switch, resp_comp).
"""
import ctypes

import numpy as np
import pytest

from oracle.vo import TERNARY, Ring, paper_b_params, toy_b_params, u64p


def P(a):
    return a.ctypes.data_as(u64p)


def arr(x):
    return np.ascontiguousarray(np.array(x, dtype=np.uint64))


def test_embed_d_interleaves_pair(oracle):
    # repack.rs:823-848: b0 = [10, 20], b1 = [30, 40] -> [10, 30, 20, 40]
    out0, out1 = arr([0] * 4), arr([0] * 4)
    oracle.L.vo_embed_at(1, 2, 4, P(arr([10, 20])), 0, P(out0))
    oracle.L.vo_embed_at(1, 2, 4, P(arr([30, 40])), 1, P(out1))
    assert (out0 + out1).tolist() == [10, 30, 20, 40]
    oracle.L.vo_embed_at(1, 2, 4, P(arr([1, 2])), 0, P(out0))
    oracle.L.vo_embed_at(1, 2, 4, P(arr([3, 4])), 1, P(out1))
    assert (out0 + out1).tolist() == [1, 3, 2, 4]


def test_deinterleave_strided(oracle):
    # via-client/src/batch.rs:63-77: [1, 2, 3, 4], N3 = 2 -> [1, 3], [2, 4]
    poly = arr([1, 2, 3, 4])
    got = []
    for t in range(2):
        r = arr([0, 0])
        oracle.L.vo_project_at(1, 4, 2, P(poly), t, P(r))
        got.append(r.tolist())
    assert got == [[1, 3], [2, 4]]


def _extr_roundtrip(oracle, seed, D, m):
    q, n, p = 65537, 8, 16
    r = Ring(n, 1, (q, 0))
    prg = oracle.prg(seed)
    sk = arr([0] * n)
    oracle.keygen(r, TERNARY, prg.h, P(sk))
    enc = arr([0] * n)
    oracle.encode(r, p, P(arr(m)), P(enc))
    ct = arr([0] * (2 * n))
    oracle.encrypt(r, P(sk), P(enc), TERNARY, prg.h, P(ct))
    rank = n // D
    masks, body = arr([0] * (rank * D)), arr([0] * D)
    oracle.extr(r, P(ct), D, P(masks), P(body))
    rd = Ring(D, 1, (q, 0))
    acc = body.astype(object)
    for j in range(rank):
        kj = arr([0] * D)
        oracle.L.vo_project_at(1, n, D, P(sk), j, P(kj))
        prod = arr([0] * D)
        oracle.poly_mul(rd, P(prod), P(np.ascontiguousarray(masks[j * D:(j + 1) * D])), P(kj))
        acc = (acc - prod.astype(object)) % q
    pt = arr([0] * D)
    oracle.decode(rd, p, P(arr([int(x) for x in acc])), P(pt))
    return pt.tolist()


def test_extr_d1_sample_extraction(oracle):
    # extract.rs:119-141
    m = [3, 1, 4, 1, 5, 9, 2, 6]
    assert _extr_roundtrip(oracle, b"extr-d1", 1, m) == [m[0]]


def test_extr_d2(oracle):
    # extract.rs:143-168: pi_0^{8->2}(M) = (M_0, M_4)
    m = [3, 1, 4, 1, 5, 9, 2, 6]
    assert _extr_roundtrip(oracle, b"extr-d2", 2, m) == [m[0], m[4]]


class Batch:
    """Client keys, a VIA-B server and the round trip, all on the oracle."""

    def __init__(self, oracle, pp, seed, records):
        self.o, self.pp = oracle, pp
        self.prg = oracle.prg(seed)
        n1, n2 = pp.n1, pp.n2
        self.sk1, self.sk2 = arr([0] * (pp.q1_nl * n1)), arr([0] * n2)
        self.ck = arr([0] * oracle.words_cascade_key(pp))
        self.cv = arr([0] * oracle.words_conv_key(pp))
        self.rsk = arr([0] * oracle.words_rsk(pp))
        assert oracle.client_setup(pp, TERNARY, TERNARY, TERNARY, self.prg.h, P(self.sk1), P(self.sk2),
                                   P(self.ck), P(self.cv), P(self.rsk)) == 0
        self.records = records
        flat = arr(np.array(records, dtype=np.uint64).reshape(-1))
        h = ctypes.c_void_p()
        assert oracle.server_new(pp, P(flat), len(records), P(self.ck), P(self.cv), P(self.rsk),
                                 ctypes.byref(h)) == 0
        self.h = h

    def close(self):
        self.o.server_free(self.h)

    def query(self, idxs):
        """Client::batch_query (client.rs:256-300): T compressed queries at the
        record degree n3, drawn from the one PRG stream in batch order."""
        o, pp = self.o, self.pp
        nl = o.query_lwes(pp) * o.words_lwe(pp)
        lwes = arr([0] * (len(idxs) * nl))
        for t, idx in enumerate(idxs):
            part = arr([0] * nl)
            assert o.client_query(pp, P(self.sk1), idx, TERNARY, self.prg.h, P(part)) == 0
            lwes[t * nl:(t + 1) * nl] = part
        return lwes

    def answer(self, T, lwes):
        o, pp = self.o, self.pp
        am, ab = arr([0] * pp.n2), arr([0] * pp.n2)
        e, g = ctypes.c_size_t(), ctypes.c_size_t()
        rc = o.server_answer_batch(self.h, T, P(lwes), o.query_lwes(pp), P(am), P(ab), ctypes.byref(e),
                                   ctypes.byref(g))
        assert rc == 0
        return am, ab

    def recover(self, T, am, ab):
        """Client::recover_batch (client.rs:316-333) + deinterleave_batch."""
        o, pp = self.o, self.pp
        pt = arr([0] * pp.n2)
        o.client_recover(pp, P(self.sk2), P(am), P(ab), P(pt))
        out = []
        for t in range(T):
            rec = arr([0] * pp.n_rec)
            o.L.vo_project_at(1, pp.n2, pp.n_rec, P(pt), t, P(rec))
            out.append(rec.tolist())
        return out


def toy_record(m):
    # batch_e2e_toy.rs:79-83
    return [(m + 1) % 16, (2 * m + 3) % 16]


def test_bit_reverse_order_t8():
    # repack.rs:222-230 at 3 bits; the repack reorders by this involution
    def br(i, bits):
        r = 0
        for _ in range(bits):
            r, i = (r << 1) | (i & 1), i >> 1
        return r
    perm = [br(i, 3) for i in range(8)]
    assert perm == [0, 4, 2, 6, 1, 5, 3, 7]
    assert [perm[x] for x in perm] == list(range(8))


def test_toy_batch_e2e(oracle):
    # batch_e2e_toy.rs:190-225: both tests share one set-up; the second test's
    # two index sets continue the PRG stream
    pp = toy_b_params(2, 2)
    b = Batch(oracle, pp, b"via-b-batch-e2e-toy", [toy_record(m) for m in range(32 * 4)])
    try:
        for idxs in ([3, 11, 20, 38, 72, 89, 106, 127], [0, 17, 34, 51, 68, 85, 102, 119],
                     [8, 25, 42, 59, 76, 93, 110, 127]):
            am, ab = b.answer(8, b.query(idxs))
            assert b.recover(8, am, ab) == [toy_record(i) for i in idxs]
    finally:
        b.close()


def test_toy_batch_errors(oracle):
    pp = toy_b_params(2, 2)
    b = Batch(oracle, pp, b"via-b-batch-errors", [toy_record(m) for m in range(8)])
    try:
        lwes = b.query([1, 2, 3, 4, 5, 6, 7, 0])
        e, g = ctypes.c_size_t(), ctypes.c_size_t()
        am, ab = arr([0] * pp.n2), arr([0] * pp.n2)
        n = oracle.query_lwes(pp)
        assert oracle.server_answer_batch(b.h, 8, P(lwes), n - 1, P(am), P(ab), ctypes.byref(e),
                                          ctypes.byref(g)) == 2
        assert (e.value, g.value) == (n, n - 1)
        # T = 3 is not a power of two; T = 16 breaks the record fit T * n3 <= n2
        assert oracle.server_answer_batch(b.h, 3, P(lwes), n, P(am), P(ab), None, None) == 1
        assert oracle.server_answer_batch(b.h, 16, P(lwes), n, P(am), P(ab), None, None) == 1
    finally:
        b.close()


def paper_b_record(m):
    # batch_e2e_paper.rs:66-71
    return [((m + 1) * (j + 1)) % 16 for j in range(64)]


@pytest.mark.slow
def test_paper_batch_e2e(oracle):
    # batch_e2e_paper.rs:103-174
    pp = paper_b_params(2, 2, T=8)
    b = Batch(oracle, pp, b"via-b-paper-batch-e2e", [paper_b_record(m) for m in range(32 * 4)])
    try:
        idxs = [3, 17, 42, 5, 88, 100, 7, 120]
        am, ab = b.answer(8, b.query(idxs))
        assert b.recover(8, am, ab) == [paper_b_record(i) for i in idxs]
    finally:
        b.close()
