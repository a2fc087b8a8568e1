"""Full-size properties of the HIP path (the oracle does not finish an answer
at these sizes in test time): library keys and queries (Client::setup /
Client::query over Shake256Prg), records from the stream's fill_bytes, the
GPU answer, the library's recovery. recover(answer(query(i))) == record(i)
is the reference's own e2e property (client_server_e2e_paper.rs:316-349,
batch_e2e_paper.rs:130-174); the smaller shapes are pinned word for word
against the oracle in test_gpu_parity.py and test_gpu_batch.py.
"""
import numpy as np
import pytest

import via_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = via_gpu.Context(0)
    yield c
    c.close()


def record_of(pp, record_bytes, i):
    """Record i of the byte stream: coefficient 2k = byte k & 0xF, 2k+1 =
    byte k >> 4 (the layout via_db_create_bytes reads)."""
    b = record_bytes[i * pp.n_rec // 2:(i + 1) * pp.n_rec // 2].astype(np.uint64)
    rec = np.empty(pp.n_rec, np.uint64)
    rec[0::2] = b & 0xF
    rec[1::2] = b >> 4
    return rec


@pytest.mark.parametrize("rows,cols", [(256, 256), (512, 512)])
def test_via_c_full_size_recovers(ctx, rows, cols):
    # 2^18 records (0.625 GiB prepared) and 2^20 records (2.5 GiB prepared)
    pp = via_gpu.paper_params(rows, cols)
    n_rec = 4 * rows * cols
    tag = b"via-gpu/full/%dx%d" % (rows, cols)
    g_db = via_gpu.Prg(tag + b"/db")
    rb = g_db.fill(ctx, n_rec * pp.n_rec // 2)
    g = via_gpu.Prg(tag + b"/setup")
    keys = via_gpu.client_setup(ctx, pp, g)
    db = via_gpu.PreparedDb(ctx, pp, record_bytes=rb.tobytes(), n_records=n_rec)
    pk = via_gpu.PreparedKeys(ctx, pp, keys["cascade_key"], keys["conv_key"], keys["rsk"])
    server = via_gpu.Server(ctx, pp, db, pk)
    g_q = via_gpu.Prg(tag + b"/query")
    for idx in (0, n_rec - 1, g_q.uniform_below(ctx, n_rec), g_q.uniform_below(ctx, n_rec)):
        mask, body = server.answer(via_gpu.client_query(ctx, pp, keys["sk1"], idx, g))
        assert np.array_equal(via_gpu.client_recover(pp, keys["sk2"], mask, body), record_of(pp, rb, idx)), idx
    pk.close()
    db.close()


@pytest.mark.parametrize("T", [8, 32])
def test_via_b_full_size_recovers(ctx, T):
    # I = 32, J = 4096 (2^17 cells, 1.25 GiB prepared), T records per answer
    pp = via_gpu.paper_b_params(32, 4096, T=T)
    d = pp.n1 // pp.n_rec
    n_rec = d * 32 * 4096
    tag = b"via-gpu/full/vb%d" % T
    g_db = via_gpu.Prg(tag + b"/db")
    rb = g_db.fill(ctx, n_rec * pp.n_rec // 2)
    g = via_gpu.Prg(tag + b"/setup")
    keys = via_gpu.client_setup(ctx, pp, g)
    db = via_gpu.PreparedDb(ctx, pp, record_bytes=rb.tobytes(), n_records=n_rec)
    pk = via_gpu.PreparedKeys(ctx, pp, keys["cascade_key"], keys["conv_key"], keys["rsk"])
    server = via_gpu.Server(ctx, pp, db, pk)
    g_q = via_gpu.Prg(tag + b"/query")
    idxs = [0, n_rec - 1] + [g_q.uniform_below(ctx, n_rec) for _ in range(T - 2)]
    queries = [via_gpu.client_query(ctx, pp, keys["sk1"], i, g) for i in idxs]
    mask, body = server.answer_batch(T, queries)
    plain = via_gpu.client_recover(pp, keys["sk2"], mask, body)
    got = via_gpu.client_deinterleave(pp, T, plain)
    for t, i in enumerate(idxs):
        assert np.array_equal(got[t], record_of(pp, rb, i)), (t, i)
    pk.close()
    db.close()
