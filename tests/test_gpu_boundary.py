"""The C-ABI contract of the answer path, on the GPU.

* Server::answer(&self) is re-entrant in the reference (via-server
  answer.rs:534-554): several host threads answer on one shared context,
  database and key set at once, and every answer equals the oracle's.
* The reference's Zq types hold canonical residues only: query, key and
  record words at or above their modulus are refused with VIA_E_INVALID
  (host entry points at once, device entry points at the next
  via_ctx_synchronize), and the context keeps answering correctly after.
* One context serves parameter sets of different cascade-gadget depth in
  turn: the workspace regrows for the deeper gadget, and each answer is
  word-exact against the oracle.
"""
import ctypes as C
import hashlib
import threading

import numpy as np
import pytest

from oracle.vo import TERNARY, make_params, paper_params, u64p

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


class Client:
    """One client's keys (Client::setup order) and its queries, with the
    oracle's answer of every query against one random database."""

    def __init__(self, oracle, opp, seed, rec_seed, indices):
        o = oracle
        n1, n2 = opp.n1, opp.n2
        prg = o.prg(seed)
        self.sk1, self.sk2 = np.zeros(2 * n1, np.uint64), np.zeros(n2, np.uint64)
        self.ck = np.zeros(o.words_cascade_key(opp), np.uint64)
        self.cv = np.zeros(o.words_conv_key(opp), np.uint64)
        self.rsk = np.zeros(o.words_rsk(opp), np.uint64)
        assert o.client_setup(opp, TERNARY, TERNARY, TERNARY, prg.h, P(self.sk1), P(self.sk2), P(self.ck),
                              P(self.cv), P(self.rsk)) == 0
        cap = (n1 // opp.n_rec) * opp.num_rows * opp.num_cols
        rng = np.random.default_rng(int.from_bytes(hashlib.sha256(rec_seed).digest()[:8], "little"))
        self.recs = rng.integers(0, opp.p, size=(cap, opp.n_rec), dtype=np.uint64)
        self.records = np.ascontiguousarray(self.recs.reshape(-1))
        h = C.c_void_p()
        assert o.server_new(opp, P(self.records), cap, P(self.ck), P(self.cv), P(self.rsk), C.byref(h)) == 0
        self.queries, self.answers = [], []
        nl = o.query_lwes(opp)
        for idx in indices:
            q = np.zeros(nl * o.words_lwe(opp), np.uint64)
            assert o.client_query(opp, P(self.sk1), idx, TERNARY, prg.h, P(q)) == 0
            m, b = np.zeros(n2, np.uint64), np.zeros(n2, np.uint64)
            assert o.server_answer(h, P(q), nl, P(m), P(b), None, None) == 0
            self.queries.append(q)
            self.answers.append((m, b))
        o.server_free(h)
        self.indices = list(indices)
        self.opp = opp
        self.oracle = o

    def recover(self, mask, body):
        rec = np.zeros(self.opp.n2, np.uint64)
        self.oracle.client_recover(self.opp, P(self.sk2), P(np.ascontiguousarray(mask)),
                                   P(np.ascontiguousarray(body)), P(rec))
        return rec


@pytest.fixture(scope="module")
def shared(oracle):
    import via_gpu

    opp = paper_params(2, 4)
    cl = Client(oracle, opp, b"via-gpu/boundary/client", b"via-gpu/boundary/db", [3, 17, 22, 30])
    ctx = via_gpu.Context(0)
    pp = _gpu_params(opp)
    db = via_gpu.PreparedDb(ctx, pp, records=cl.records)
    keys = via_gpu.PreparedKeys(ctx, pp, cl.ck, cl.cv, cl.rsk)
    yield cl, ctx, pp, db, keys
    ctx.close()


def test_concurrent_answers_on_one_context(shared):
    # four host threads, each answering its own query repeatedly through the
    # one shared context, database and keys (via_answer, host buffers)
    import via_gpu

    cl, ctx, pp, db, keys = shared
    server = via_gpu.Server(ctx, pp, db, keys)
    results = {}
    errors = []
    start = threading.Barrier(4)

    def run(t):
        try:
            start.wait()
            out = []
            for rep in range(6):
                out.append(server.answer(cl.queries[t]))
            results[t] = out
        except Exception as e:  # reported below, the test thread raises
            errors.append((t, repr(e)))

    threads = [threading.Thread(target=run, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads), "an answering thread did not finish"
    assert not errors, errors
    for t in range(4):
        want_m, want_b = cl.answers[t]
        for rep, (m, b) in enumerate(results[t]):
            assert np.array_equal(m, want_m) and np.array_equal(b, want_b), (t, rep)
        assert np.array_equal(cl.recover(*results[t][0]), cl.recs[cl.indices[t]])


def test_concurrent_host_and_device_calls(shared):
    # host-buffer answers on borrowed lanes while a device-buffer answer runs
    # on the context's own stream: every answer equals the oracle's
    import via_gpu

    cl, ctx, pp, db, keys = shared
    server = via_gpu.Server(ctx, pp, db, keys)
    L = via_gpu.lib()
    dq = ctx.alloc(cl.queries[0].nbytes)
    dm, dbd = ctx.alloc(8 * pp.n2), ctx.alloc(8 * pp.n2)
    ctx.to_device(dq, cl.queries[0])
    got = {}
    errors = []

    def host(t):
        try:
            got[t] = [server.answer(cl.queries[t]) for _ in range(4)]
        except Exception as e:
            errors.append((t, repr(e)))

    threads = [threading.Thread(target=host, args=(t,)) for t in (1, 2, 3)]
    for th in threads:
        th.start()
    err = via_gpu.ErrDetail()
    for _ in range(4):
        assert L.via_answer_device(ctx.h, db.h, keys.h, dq, pp.query_lwes, dm, dbd, C.byref(err)) == 0
        ctx.synchronize()
        m = ctx.to_host(np.zeros(pp.n2, np.uint64), dm)
        b = ctx.to_host(np.zeros(pp.n2, np.uint64), dbd)
        assert np.array_equal(m, cl.answers[0][0]) and np.array_equal(b, cl.answers[0][1])
    for th in threads:
        th.join(timeout=120)
    assert not errors, errors
    for t in (1, 2, 3):
        for m, b in got[t]:
            assert np.array_equal(m, cl.answers[t][0]) and np.array_equal(b, cl.answers[t][1]), t
    for d in (dq, dm, dbd):
        ctx.free(d)


def test_noncanonical_query_words_refused(shared):
    import via_gpu

    cl, ctx, pp, db, keys = shared
    server = via_gpu.Server(ctx, pp, db, keys)
    n1p = pp.n1 + 1
    for pos, val in [(5, pp.q1[0]), (n1p + 7, pp.q1[1]), (n1p - 1, 1 << 53), (3 * n1p + 2, (1 << 64) - 1)]:
        bad = cl.queries[1].copy()
        bad[pos] = val
        with pytest.raises(via_gpu.ViaError) as e:
            server.answer(bad)
        assert e.value.status == via_gpu.VIA_E_INVALID, (pos, val)
    # q1[1] - 1 in a limb-0 slot is above q1[0]: refused; in a limb-1 slot it is canonical
    bad = cl.queries[1].copy()
    bad[9] = pp.q1[1] - 1
    with pytest.raises(via_gpu.ViaError):
        server.answer(bad)
    # the device entry point reports at the next synchronisation
    L = via_gpu.lib()
    bad = cl.queries[1].copy()
    bad[11] = pp.q1[0] + 3
    dq = ctx.alloc(bad.nbytes)
    dm, dbd = ctx.alloc(8 * pp.n2), ctx.alloc(8 * pp.n2)
    ctx.to_device(dq, bad)
    err = via_gpu.ErrDetail()
    assert L.via_answer_device(ctx.h, db.h, keys.h, dq, pp.query_lwes, dm, dbd, C.byref(err)) == 0
    assert L.via_ctx_synchronize(ctx.h) == via_gpu.VIA_E_INVALID
    assert L.via_ctx_synchronize(ctx.h) == 0  # reported once
    # a canonical query afterwards, device and host paths, answers exactly
    ctx.to_device(dq, cl.queries[1])
    assert L.via_answer_device(ctx.h, db.h, keys.h, dq, pp.query_lwes, dm, dbd, C.byref(err)) == 0
    assert L.via_ctx_synchronize(ctx.h) == 0
    assert np.array_equal(ctx.to_host(np.zeros(pp.n2, np.uint64), dm), cl.answers[1][0])
    m, b = server.answer(cl.queries[1])
    assert np.array_equal(m, cl.answers[1][0]) and np.array_equal(b, cl.answers[1][1])
    for d in (dq, dm, dbd):
        ctx.free(d)


def test_noncanonical_keys_and_records_refused(shared):
    import via_gpu

    cl, ctx, pp, db, keys = shared
    n1 = pp.n1
    # cascade key: limb 1 of the first sample of step 0 (n_out = 2) at q1[1]
    for arr_name, pos, val in [("ck", 2, pp.q1[1]), ("ck", 0, pp.q1[0]), ("cv", n1 + 4, pp.q1[1]),
                               ("rsk", 10, pp.q3)]:
        words = {"ck": cl.ck.copy(), "cv": cl.cv.copy(), "rsk": cl.rsk.copy()}
        words[arr_name][pos] = val
        with pytest.raises(via_gpu.ViaError) as e:
            via_gpu.PreparedKeys(ctx, pp, words["ck"], words["cv"], words["rsk"])
        assert e.value.status == via_gpu.VIA_E_INVALID, (arr_name, pos)
    recs = cl.records.copy()
    recs[1000] = pp.p
    with pytest.raises(via_gpu.ViaError) as e:
        via_gpu.PreparedDb(ctx, pp, records=recs)
    assert e.value.status == via_gpu.VIA_E_INVALID
    with pytest.raises(via_gpu.ViaError) as e:
        db.update([4], np.full(pp.n_rec, pp.p + 1, np.uint64))
    assert e.value.status == via_gpu.VIA_E_INVALID
    # nothing changed: the database still answers as the oracle's
    server = via_gpu.Server(ctx, pp, db, keys)
    m, b = server.answer(cl.queries[2])
    assert np.array_equal(m, cl.answers[2][0]) and np.array_equal(b, cl.answers[2][1])


def test_one_context_two_gadget_depths(oracle):
    # the paper cascade gadget (base 18, depth 18), then a deeper one (base
    # 14, depth 20), then the paper one again, all on one context: the
    # workspace regrows for depth 20 and every answer is word-exact
    import via_gpu

    opp_a = paper_params(2, 2)
    opp_b = make_params(2048, 512, (137438822401, 274810798081), 17175674881, 8380417, 4096, 16, 55879, 81, 8, 14,
                        2, 20, 8, 2, 2)
    ca = Client(oracle, opp_a, b"via-gpu/depth/a", b"via-gpu/depth/db", [6])
    cb = Client(oracle, opp_b, b"via-gpu/depth/b", b"via-gpu/depth/db", [9])
    ctx = via_gpu.Context(0)
    try:
        for cl in (ca, cb, ca):
            pp = _gpu_params(cl.opp)
            db = via_gpu.PreparedDb(ctx, pp, records=cl.records)
            keys = via_gpu.PreparedKeys(ctx, pp, cl.ck, cl.cv, cl.rsk)
            m, b = via_gpu.Server(ctx, pp, db, keys).answer(cl.queries[0])
            assert np.array_equal(m, cl.answers[0][0]) and np.array_equal(b, cl.answers[0][1]), cl.opp.l_ck
            assert np.array_equal(cl.recover(m, b), cl.recs[cl.indices[0]])
            keys.close()
            db.close()
    finally:
        ctx.close()
