"""via_answer_device_noalloc: the one-query answer with its workspace in the
library's per-device scratch array, through the C ABI.

Two calls in a row, the second on a smaller database and query, each equal
word for word to via_answer_device on the same inputs (which the parity
tests pin to the oracle); then the refusals: a workspace beyond the scratch
array, a wrong query length, a one-row grid.
"""
import ctypes as C
import hashlib

import numpy as np
import pytest

from oracle.vo import TERNARY, paper_params, u64p

pytestmark = pytest.mark.gpu


def P(a):
    return a.ctypes.data_as(u64p)


def _gpu_params(opp):
    import via_gpu

    g = via_gpu.Params()
    for name, _ in opp._fields_:
        v = getattr(opp, name)
        if name == "q1":
            g.q1[0], g.q1[1] = v[0], v[1]
        else:
            setattr(g, name, v)
    return g


def _client(oracle, rows, cols, seed, indices):
    opp = paper_params(rows, cols)
    prg = oracle.prg(seed)
    n1, n2 = opp.n1, opp.n2
    sk1, sk2 = np.zeros(2 * n1, np.uint64), np.zeros(n2, np.uint64)
    ck = np.zeros(oracle.words_cascade_key(opp), np.uint64)
    cv = np.zeros(oracle.words_conv_key(opp), np.uint64)
    rsk = np.zeros(oracle.words_rsk(opp), np.uint64)
    assert oracle.client_setup(opp, TERNARY, TERNARY, TERNARY, prg.h, P(sk1), P(sk2), P(ck), P(cv), P(rsk)) == 0
    qs = []
    for idx in indices:
        q = np.zeros(oracle.query_lwes(opp) * oracle.words_lwe(opp), np.uint64)
        assert oracle.client_query(opp, P(sk1), idx, TERNARY, prg.h, P(q)) == 0
        qs.append(q)
    return opp, ck, cv, rsk, qs


def _records(opp, seed):
    rng = np.random.default_rng(int.from_bytes(hashlib.sha256(seed).digest()[:8], "little"))
    cap = 4 * opp.num_rows * opp.num_cols
    return np.ascontiguousarray(rng.integers(0, 16, size=cap * opp.n_rec, dtype=np.uint64))


def _answer_both(ctx, pp, db, keys, q):
    import via_gpu

    L = via_gpu.lib()
    err = via_gpu.ErrDetail()
    dq = ctx.alloc(q.nbytes)
    dm = [ctx.alloc(8 * pp.n2) for _ in range(4)]
    try:
        ctx.to_device(dq, q)
        n_lwe = q.size // pp.lwe_words
        st = L.via_answer_device(ctx.h, db.h, keys.h, dq, n_lwe, dm[0], dm[1], C.byref(err))
        assert st == 0, st
        assert L.via_ctx_synchronize(ctx.h) == 0
        st = L.via_answer_device_noalloc(ctx.h, db.h, keys.h, dq, n_lwe, dm[2], dm[3], C.byref(err))
        assert st == 0, (st, err.msg)
        assert L.via_ctx_synchronize(ctx.h) == 0
        return [ctx.to_host(np.zeros(pp.n2, np.uint64), d) for d in dm]
    finally:
        for d in [dq] + dm:
            ctx.free(d)


def test_noalloc_matches_answer_device_twice(oracle):
    import via_gpu

    ctx = via_gpu.Context(0)
    try:
        # first call: 16 x 16 (20 LWEs); second: 4 x 4 (12 LWEs), smaller in
        # every buffer of the carved workspace
        for rows, cols, idx in [(16, 16, 777), (4, 4, 21)]:
            opp, ck, cv, rsk, qs = _client(oracle, rows, cols, b"via-gpu/noalloc/%d" % rows, [idx])
            pp = _gpu_params(opp)
            db = via_gpu.PreparedDb(ctx, pp, records=_records(opp, b"via-gpu/noalloc/db%d" % rows))
            keys = via_gpu.PreparedKeys(ctx, pp, ck, cv, rsk)
            print("case: noalloc vs via_answer_device, %dx%d index %d" % (rows, cols, idx), flush=True)
            m, b, m_na, b_na = _answer_both(ctx, pp, db, keys, qs[0])
            assert np.array_equal(m_na, m) and np.array_equal(b_na, b), (rows, cols)
            keys.close()
            db.close()
    finally:
        ctx.close()


def test_noalloc_refusals(oracle):
    import via_gpu

    L = via_gpu.lib()
    ctx = via_gpu.Context(0)
    try:
        # 512 x 512: the carved workspace needs more than the scratch array
        opp, ck, cv, rsk, qs = _client(oracle, 512, 512, b"via-gpu/noalloc/big", [5])
        pp = _gpu_params(opp)
        db = via_gpu.PreparedDb(ctx, pp, records=np.zeros(0, np.uint64))
        keys = via_gpu.PreparedKeys(ctx, pp, ck, cv, rsk)
        n_lwe = pp.query_lwes
        dq = ctx.alloc(8 * n_lwe * pp.lwe_words)
        dm, dbd = ctx.alloc(8 * pp.n2), ctx.alloc(8 * pp.n2)
        err = via_gpu.ErrDetail()
        st = L.via_answer_device_noalloc(ctx.h, db.h, keys.h, dq, n_lwe, dm, dbd, C.byref(err))
        assert st == via_gpu.VIA_E_NOMEM
        assert err.got == 96 << 20 and err.expected > err.got
        st = L.via_answer_device_noalloc(ctx.h, db.h, keys.h, dq, n_lwe - 1, dm, dbd, C.byref(err))
        assert st == via_gpu.VIA_E_QUERY_LEN and (err.expected, err.got) == (n_lwe, n_lwe - 1)
        assert L.via_ctx_synchronize(ctx.h) == 0
        for d in (dq, dm, dbd):
            ctx.free(d)
        keys.close()
        db.close()
        # a one-row grid keeps its host-built DMux root: refused
        opp, ck, cv, rsk, qs = _client(oracle, 1, 4, b"via-gpu/noalloc/row", [2])
        pp = _gpu_params(opp)
        db = via_gpu.PreparedDb(ctx, pp, records=_records(opp, b"via-gpu/noalloc/row-db"))
        keys = via_gpu.PreparedKeys(ctx, pp, ck, cv, rsk)
        dq = ctx.alloc(qs[0].nbytes)
        dm, dbd = ctx.alloc(8 * pp.n2), ctx.alloc(8 * pp.n2)
        ctx.to_device(dq, qs[0])
        st = L.via_answer_device_noalloc(ctx.h, db.h, keys.h, dq, pp.query_lwes, dm, dbd, C.byref(err))
        assert st == via_gpu.VIA_E_INVALID
        for d in (dq, dm, dbd):
            ctx.free(d)
    finally:
        ctx.close()
