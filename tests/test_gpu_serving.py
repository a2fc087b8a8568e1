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
