"""Incremental database updates against a database rebuilt by the oracle.

via_db_update rewrites records of the HBM-resident prepared database in
place (record m is slot m / (I*J) of cell m mod I*J, setup_db.rs:58-72); the
result must equal, residue for residue, the oracle's PreparedDb built from
the updated record list: present and padded records, a repeated index (the
last record wins), both fragment heights (32- and 64-row databases).
"""
import numpy as np
import pytest

from tests.paper_setup import oracle
from tests.test_gpu_first_dim import oracle_db

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import via_gpu

    c = via_gpu.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("I,J", [(8, 4), (64, 4)])
def test_update_matches_rebuilt_database(ctx, I, J):
    import via_gpu
    from oracle.vo import paper_params

    print(f"db update I={I} J={J}", flush=True)
    o = oracle()
    cap = 4 * I * J
    nrec = cap - 10
    rng = np.random.default_rng(I * 100 + J)
    recs = rng.integers(0, 16, size=(nrec, 512), dtype=np.uint64)
    db = via_gpu.PreparedDb(ctx, via_gpu.paper_params(I, J), records=recs.reshape(-1))
    idx = [int(x) for x in rng.choice(cap, size=12, replace=False)] + [cap - 1, 3, 3]
    new = rng.integers(0, 16, size=(len(idx), 512), dtype=np.uint64)
    db.update(np.array(idx, dtype=np.uint64), new.reshape(-1))
    full = np.zeros((cap, 512), dtype=np.uint64)
    full[:nrec] = recs
    for u, m in enumerate(idx):
        full[m] = new[u]
    want = oracle_db(o, paper_params(I, J), np.ascontiguousarray(full.reshape(-1)), cap)
    assert np.array_equal(db.cells(), want)
    with pytest.raises(via_gpu.ViaError) as e:
        db.update(np.array([cap], dtype=np.uint64), new[:1].reshape(-1))
    assert e.value.status == via_gpu.VIA_E_INDEX
    db.close()
