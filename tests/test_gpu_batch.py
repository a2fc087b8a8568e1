"""VIA-B on the HIP path: Server::answer_batch::<T> (answer.rs:561-602).

* Paper T = 8 batch of batch_e2e_paper.rs (I = J = 2, n3 = 64, the
  reference's seed and indices): the repacked ciphertext and the answer are
  bit-identical to the CPU restatement, and every record recovers.
* A random-record T = 8 batch at I = J = 4, compared the same way.
* T = 256 (the production preset, n3 = 2): every one of the 256 records
  recovers in batch order (a size-independent property; the CPU restatement
  takes minutes for it).
* Errors: batch length, query length, unsupported T.
"""
import ctypes

import numpy as np
import pytest

from tests.paper_setup import BatchClient as Client

pytestmark = pytest.mark.gpu


def P(a):
    from oracle.vo import u64p

    return a.ctypes.data_as(u64p)


def arr(x):
    return np.ascontiguousarray(np.array(x, dtype=np.uint64))


def oracle_batch(oracle, pp, client, records, T, queries):
    """The restatement's answer_batch, plus its repacked ciphertext."""
    h = ctypes.c_void_p()
    flat = arr(np.array(records, dtype=np.uint64).reshape(-1))
    assert oracle.server_new(pp, P(flat), len(records), P(client.ck), P(client.cv), P(client.rsk),
                             ctypes.byref(h)) == 0
    try:
        lw = np.concatenate(queries)
        am, ab = arr([0] * pp.n2), arr([0] * pp.n2)
        assert oracle.server_answer_batch(h, T, P(lw), oracle.query_lwes(pp), P(am), P(ab), None, None) == 0
        # repacked RLWE from the restatement's pieces
        n1 = pp.n1
        rot = arr([0] * (T * 2 * n1))
        for t in range(T):
            seg = arr([0] * (2 * n1))
            oracle.server_answer_stages(h, P(queries[t]), oracle.query_lwes(pp), None, None, None, None, None,
                                        P(seg), None, None)
            rot[t * 2 * n1:(t + 1) * 2 * n1] = seg
    finally:
        oracle.server_free(h)
    from oracle.vo import Ring

    r1 = Ring.make(n1, pp.q1[0], pp.q1[1])
    r2 = Ring.make(n1, pp.q2)
    G = pp.n2 // T
    key = arr([0] * oracle.repack_key_words(n1, G, pp.l_ck))
    oracle.repack_key_from_cascade(r1, P(client.ck), pp.q2, G, pp.l_ck, P(key))
    packed = arr([0] * (2 * n1))
    assert oracle.repack(r2, T, G, P(rot), P(key), pp.ck_base, pp.l_ck, P(packed)) == 0
    return am, ab, packed


def gpu_server(client, pp_gpu, records):
    import via_gpu

    return via_gpu.Server.setup(via_gpu.ServerConfig(pp_gpu, client.ck, client.cv, client.rsk),
                                np.array(records, dtype=np.uint64))


def paper_b_record(m):
    # batch_e2e_paper.rs:66-71
    return [((m + 1) * (j + 1)) % 16 for j in range(64)]


@pytest.fixture(scope="module")
def paper_b(oracle):
    import via_gpu
    from oracle.vo import paper_b_params

    pp = paper_b_params(2, 2, T=8)
    client = Client(oracle, pp, b"via-b-paper-batch-e2e")
    records = [paper_b_record(m) for m in range(32 * 4)]
    server = gpu_server(client, via_gpu.paper_b_params(2, 2, T=8), records)
    yield pp, client, records, server
    server.ctx.close()


@pytest.fixture(scope="module")
def paper_b_t8(oracle, paper_b):
    pp, client, records, server = paper_b
    idxs = [3, 17, 42, 5, 88, 100, 7, 120]
    qs = client.queries(idxs)
    return idxs, qs, oracle_batch(oracle, pp, client, records, 8, qs)


def test_paper_t8_batch_matches_oracle(paper_b, paper_b_t8):
    pp, client, records, server = paper_b
    idxs, qs, (wm, wb, wpacked) = paper_b_t8
    print("case: paper VIA-B T=8 I=J=2", flush=True)
    am, ab = server.answer_batch(8, qs)
    packed = server.ctx.stage(7, 2 * pp.n1)
    assert np.array_equal(packed, wpacked)
    assert np.array_equal(am, wm) and np.array_equal(ab, wb)
    assert client.recover(8, am, ab) == [paper_b_record(i) for i in idxs]


def test_paper_t8_exact_paths_match_oracle(paper_b, paper_b_t8):
    # every coefficient through the exact paths the fp64 fast paths fall back
    # to, in the VIA-C prefix of the eight queries and in the repack at q2
    import via_gpu

    pp, client, records, server = paper_b
    idxs, qs, (wm, wb, wpacked) = paper_b_t8
    L = via_gpu.lib()
    assert L.via_debug_set_exact_paths(1) == 0
    try:
        print("case: paper VIA-B T=8 I=J=2, exact paths", flush=True)
        am, ab = server.answer_batch(8, qs)
        assert np.array_equal(server.ctx.stage(7, 2 * pp.n1), wpacked)
        assert np.array_equal(am, wm) and np.array_equal(ab, wb)
    finally:
        L.via_debug_set_exact_paths(0)


def test_batch_errors(paper_b):
    import via_gpu

    pp, client, records, server = paper_b
    qs = client.queries([0, 1, 2, 3, 4, 5, 6, 7])
    with pytest.raises(via_gpu.ViaError) as e:
        server.answer_batch(8, qs[:7])
    assert e.value.status == via_gpu.VIA_E_DIM
    bad = list(qs)
    bad[3] = bad[3][: (pp.l_query * 6) * pp.q1_nl * (pp.n1 + 1)]
    with pytest.raises(via_gpu.ViaError) as e:
        server.answer_batch(8, bad)
    assert e.value.status == via_gpu.VIA_E_QUERY_LEN
    assert (e.value.expected, e.value.got) == (14, 12)
    for T in (3, 16):
        with pytest.raises(via_gpu.ViaError) as e:
            server.answer_batch(T, qs[:T] if T < 8 else qs * 2)
        assert e.value.status == via_gpu.VIA_E_INVALID


def test_random_t8_batch_4x4(oracle):
    import via_gpu
    from oracle.vo import paper_b_params

    pp = paper_b_params(4, 4, T=8)
    client = Client(oracle, pp, b"via-gpu/vb-4x4/setup")
    rng = np.random.default_rng(44)
    records = rng.integers(0, 16, size=(32 * 16 - 5, 64), dtype=np.uint64).tolist()
    server = gpu_server(client, via_gpu.paper_b_params(4, 4, T=8), records)
    try:
        idxs = [int(x) for x in rng.choice(len(records), size=8, replace=False)]
        qs = client.queries(idxs)
        print("case: random VIA-B T=8 I=J=4", idxs, flush=True)
        am, ab = server.answer_batch(8, qs)
        wm, wb, _ = oracle_batch(oracle, pp, client, records, 8, qs)
        assert np.array_equal(am, wm) and np.array_equal(ab, wb)
        assert client.recover(8, am, ab) == [records[i] for i in idxs]
        # deterministic on a second call
        am2, ab2 = server.answer_batch(8, qs)
        assert np.array_equal(am2, am) and np.array_equal(ab2, ab)
    finally:
        server.ctx.close()


def test_t256_batch_recovers_all(oracle):
    import via_gpu
    from oracle.vo import paper_b_params

    T = 256
    pp = paper_b_params(2, 2, T=T)
    client = Client(oracle, pp, b"via-gpu/vb-t256/setup")
    rng = np.random.default_rng(256)
    n_records = (pp.n1 // pp.n_rec) * 4
    records = rng.integers(0, 16, size=(n_records, pp.n_rec), dtype=np.uint64).tolist()
    server = gpu_server(client, via_gpu.paper_b_params(2, 2, T=T), records)
    try:
        idxs = [int(x) for x in rng.choice(n_records, size=T, replace=False)]
        qs = client.queries(idxs)
        print("case: VIA-B T=256 I=J=2", flush=True)
        am, ab = server.answer_batch(T, qs)
        assert client.recover(T, am, ab) == [records[i] for i in idxs]
    finally:
        server.ctx.close()


def test_query_length_follows_database_record_degree(oracle, paper_b):
    """Keys made at the VIA-B record degree serve a VIA-C database: the
    expected query length is the database's (8 LWEs at I = J = 2, n_rec =
    512), and a query of the VIA-B length is refused before any launch."""
    import via_gpu

    pp, client, records, server = paper_b
    g_c = via_gpu.paper_params(2, 2)
    db_c = via_gpu.PreparedDb(server.ctx, g_c, records=np.zeros((16, 512), dtype=np.uint64))
    s = via_gpu.Server(server.ctx, g_c, db_c, server.keys)
    q = client.queries([5])[0]  # 14 LWEs: the VIA-B length
    with pytest.raises(via_gpu.ViaError) as e:
        s.answer(q)
    assert e.value.status == via_gpu.VIA_E_QUERY_LEN
    assert (e.value.expected, e.value.got) == (8, 14)
    db_c.close()
