"""The serving loop: each answer in a mix of clients equals its solo answer
(and the oracle's), clients are admitted under an HBM key budget."""
import numpy as np
import pytest

from tests.paper_setup import paper_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import via_gpu

    c = via_gpu.Context(0)
    yield c
    c.close()


def test_serving_mix_matches_solo(ctx):
    import via_gpu

    cases = [paper_case(8, 8, seed=b"via-gpu/serve/client%d" % q, index=(53 * q + 11) % 256, records="random",
                        rec_seed=b"via-gpu/serve-db") for q in range(6)]
    pp = cases[0].pp_gpu
    db = via_gpu.PreparedDb(ctx, pp, records=cases[0].records)
    loop = via_gpu.ServeLoop(pp, db, key_budget_bytes=1 << 34, max_batch=4, slots=2)
    for cid, c in enumerate(cases):
        loop.admit(cid, c.cascade_key, c.conv_key, c.rsk)
    # a burst: every client submits three times, interleaved
    tickets = []
    for rep in range(3):
        for cid, c in enumerate(cases):
            tickets.append((cid, loop.submit(cid, c.lwes)))
    for cid, t in tickets:
        m, b, lat = loop.wait(t)
        c = cases[cid]
        assert np.array_equal(m, c.ans_mask) and np.array_equal(b, c.ans_body), cid
        assert np.array_equal(c.recover(m, b), c.record)
        assert lat > 0
    st = loop.stats()
    assert st["answered"] == 18 and st["max_batch_seen"] > 1
    loop.close()
    db.close()


def test_admission_by_key_budget(ctx):
    import via_gpu

    case = paper_case(2, 2, seed=b"via-gpu/admit", index=3)
    pp = case.pp_gpu
    db = via_gpu.PreparedDb(ctx, pp, records=case.records)
    per_client = 27_400_000  # one client's eval keys: 27,394,048 bytes of fp64
    loop = via_gpu.ServeLoop(pp, db, key_budget_bytes=2 * per_client, max_batch=2, slots=1)
    loop.admit(1, case.cascade_key, case.conv_key, case.rsk)
    loop.admit(2, case.cascade_key, case.conv_key, case.rsk)
    with pytest.raises(via_gpu.ViaError) as e:
        loop.admit(3, case.cascade_key, case.conv_key, case.rsk)
    assert e.value.status == via_gpu.VIA_E_NOMEM
    t = loop.submit(2, case.lwes)
    loop.evict(1)
    loop.admit(3, case.cascade_key, case.conv_key, case.rsk)
    m, b, _ = loop.wait(t)
    assert np.array_equal(m, case.ans_mask) and np.array_equal(b, case.ans_body)
    t3 = loop.submit(3, case.lwes)
    m3, b3, _ = loop.wait(t3)
    assert np.array_equal(m3, case.ans_mask) and np.array_equal(case.recover(m3, b3), case.record)
    loop.close()
    db.close()


def test_serving_mixes_via_c_and_via_b(ctx, oracle):
    """VIA-C queries and VIA-B batches (T = 8) of three clients through one
    loop with two slots: every answer equals the one the same request gets
    alone (via_answer / via_answer_batch on a fresh context), and recovers."""
    import via_gpu
    from oracle.vo import paper_b_params, paper_params
    from tests.paper_setup import BatchClient as Client

    pp_c, pp_b = paper_params(2, 2), paper_b_params(2, 2, T=8)
    g_c, g_b = via_gpu.paper_params(2, 2), via_gpu.paper_b_params(2, 2, T=8)
    rng = np.random.default_rng(77)
    rec_c = rng.integers(0, 16, size=(16, 512), dtype=np.uint64)
    rec_b = rng.integers(0, 16, size=(128, 64), dtype=np.uint64)
    db_c = via_gpu.PreparedDb(ctx, g_c, records=rec_c)
    db_b = via_gpu.PreparedDb(ctx, g_b, records=rec_b)
    clients = [Client(oracle, pp_b, b"via-gpu/serve-mix/client%d" % i) for i in range(3)]
    reqs = []  # (client, kind, payload, expected records)
    for i, cl in enumerate(clients):
        for rep in range(2):
            idx = int(rng.integers(0, 16))
            cl.pp = pp_c
            q = cl.queries([idx])[0]
            reqs.append((i, "c", q, rec_c[idx].tolist()))
        idxs = [int(x) for x in rng.choice(128, size=8, replace=False)]
        cl.pp = pp_b
        reqs.append((i, "b", cl.queries(idxs), [rec_b[j].tolist() for j in idxs]))
    # solo answers, one request per call
    solo = []
    sctx = via_gpu.Context(0)
    keys = [via_gpu.PreparedKeys(sctx, g_b, cl.ck, cl.cv, cl.rsk) for cl in clients]
    for i, kind, payload, _ in reqs:
        if kind == "c":
            solo.append(via_gpu.Server(sctx, g_c, db_c, keys[i]).answer(payload))
        else:
            solo.append(via_gpu.Server(sctx, g_b, db_b, keys[i]).answer_batch(8, payload))
    sctx.close()
    loop = via_gpu.ServeLoop(g_c, db_c, key_budget_bytes=1 << 34, max_batch=4, slots=2, db_b=db_b)
    for i, cl in enumerate(clients):
        loop.admit(i, cl.ck, cl.cv, cl.rsk)
    tickets = []
    for rep in range(2):
        for i, kind, payload, _ in reqs:
            tickets.append(loop.submit(i, payload) if kind == "c" else loop.submit_batch(i, 8, payload))
    for k, t in enumerate(tickets):
        i, kind, payload, want = reqs[k % len(reqs)]
        m, b, _ = loop.wait(t)
        sm, sb = solo[k % len(reqs)]
        assert np.array_equal(m, sm) and np.array_equal(b, sb), (k, kind)
        cl = clients[i]
        if kind == "c":
            cl.pp = pp_c
            assert cl.recover(1, m, b)[0] == want
        else:
            cl.pp = pp_b
            assert cl.recover(8, m, b) == want
    st = loop.stats()
    assert st["answered"] == len(tickets) and st["answered_b"] == 6 and st["records"] == 12 + 6 * 8
    loop.close()
    db_c.close()
    db_b.close()
