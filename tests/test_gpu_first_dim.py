"""FirstDim on the int8 matrix cores against the CPU restatement (bit-exact).

via_first_dim runs the answer path's FirstDim kernels (selection pack, the
k_fd_mfma pass, segment gather) on given mod-switched selection ciphertexts;
the oracle's vo_first_dim (first_dim.rs:94-184) computes the same J output
ciphertexts per query from the oracle's own prepared database. The shapes
cover both fragment heights (32 and 64 rows), several row chunks and two row
segments, databases narrower than one 16-column block, and 1 to 32 queries
per pass (one to four 16-column selection blocks, and column counts that are
not multiples of 16).
"""
import ctypes

import numpy as np
import pytest

from tests.paper_setup import oracle

pytestmark = pytest.mark.gpu

Q2 = 17175674881


@pytest.fixture(scope="module")
def ctx():
    import via_gpu

    c = via_gpu.Context(0)
    yield c
    c.close()


def P(a):
    from oracle.vo import u64p

    return a.ctypes.data_as(u64p)


def oracle_db(o, pp, records, nrec):
    """The oracle's PreparedDb (setup_db + eval form at q2), [I][J][n1]."""
    ck = np.zeros(o.words_cascade_key(pp), np.uint64)
    cv = np.zeros(o.words_conv_key(pp), np.uint64)
    rsk = np.zeros(o.words_rsk(pp), np.uint64)
    h = ctypes.c_void_p()
    assert o.server_new(pp, P(records), nrec, P(ck), P(cv), P(rsk), ctypes.byref(h)) == 0
    n = pp.num_rows * pp.num_cols * pp.n1
    db = np.ctypeslib.as_array(o.server_db(h), shape=(n,)).copy()
    o.server_free(h)
    return db


def check_first_dim(ctx, I, J, Q, seed):
    import via_gpu
    from oracle.vo import paper_params

    o = oracle()
    pp = paper_params(I, J)
    n1 = pp.n1
    rng = np.random.default_rng(seed)
    nrec = 4 * I * J
    records = np.ascontiguousarray(rng.integers(0, 16, size=nrec * 512, dtype=np.uint64))
    want_db = oracle_db(o, pp, records, nrec)
    db = via_gpu.PreparedDb(ctx, via_gpu.paper_params(I, J), records=records)
    assert np.array_equal(db.cells(), want_db), "prepared database differs"
    switched = rng.integers(0, Q2, size=(Q, I, 2, n1), dtype=np.uint64)
    got = db.first_dim(switched, Q).reshape(Q, J, 2, n1)
    for q in range(Q):
        want = np.zeros(J * 2 * n1, np.uint64)
        o.first_dim(pp, P(np.ascontiguousarray(switched[q]).reshape(-1)), P(want_db), 0, P(want))
        assert np.array_equal(got[q].reshape(-1), want), (I, J, Q, q)
    db.close()


@pytest.mark.parametrize("I,J,Q", [
    (2, 2, 1),      # two real rows in a 32-row fragment, two of 16 columns
    (32, 8, 9),     # 32-row fragments (16x16x32), two selection blocks
    (16, 32, 17),   # two column blocks, three selection blocks
    (64, 4, 1),     # 64-row fragments, one chunk
    (128, 16, 32),  # two chunks, four full selection blocks
    (512, 2, 3),    # two row segments summed by the gather
])
def test_first_dim_matches_oracle(ctx, I, J, Q):
    print(f"first_dim I={I} J={J} Q={Q}", flush=True)
    check_first_dim(ctx, I, J, Q, seed=I * 1000 + J * 10 + Q)


def test_first_dim_paper_size(ctx):
    # the paper database shape (I = J = 256, 1 GiB as u64 residues), two
    # queries in one pass
    print("first_dim I=256 J=256 Q=2", flush=True)
    check_first_dim(ctx, 256, 256, 2, seed=256)


def test_first_dim_rejects_bad_batches(ctx):
    import via_gpu

    rng = np.random.default_rng(3)
    records = np.ascontiguousarray(rng.integers(0, 16, size=4 * 4 * 4 * 512, dtype=np.uint64))
    db = via_gpu.PreparedDb(ctx, via_gpu.paper_params(4, 4), records=records)
    sw = np.zeros(33 * 4 * 2 * 2048, np.uint64)
    with pytest.raises(via_gpu.ViaError):
        db.first_dim(sw, 33)  # more queries than one pass takes
    with pytest.raises(ValueError):
        db.first_dim(sw[:100], 1)
    db.close()
