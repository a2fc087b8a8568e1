"""Full-size answers of the HIP path, byte for byte.

Inputs come from the library alone: library keys and queries
(Client::setup / Client::query over Shake256Prg), records from the stream's
fill_bytes. Each GPU answer must (1) hash to the SHA-256 of the oracle's
answer for the same seeded inputs, committed as a self-golden fixture by
tests/golden/make_answer_digests.py (the oracle takes minutes at these
sizes), and (2) recover the record, the reference's own e2e property
(client_server_e2e_paper.rs:316-349, batch_e2e_paper.rs:130-174).

Shapes: VIA-C at I = J = 128 (2^16 records, the "2^16 records" reading of
the paper configuration), 256 x 256 (2^18 records, 1 GiB as u64 residues)
and 512 x 512 (2^20 records); VIA-B at I = 32, J = 4096 (2^17 cells) with
T = 8 (the reference's n1 = 2048 batch size), 16, 32 and 64.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import via_gpu

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "self_golden_answers.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def digest(mask, body):
    return hashlib.sha256(np.ascontiguousarray(mask, "<u8").tobytes() +
                          np.ascontiguousarray(body, "<u8").tobytes()).hexdigest()


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


@pytest.mark.parametrize("rows,cols", [(128, 128), (256, 256), (512, 512)])
def test_via_c_full_size_bytes(ctx, golden, rows, cols):
    # 2^16, 2^18 and 2^20 records (0.156, 0.625 and 2.5 GiB prepared)
    want = golden["paper_%dx%d" % (rows, cols)]
    pp = via_gpu.paper_params(rows, cols)
    n_rec = 4 * rows * cols
    tag = b"via-gpu/full/%dx%d" % (rows, cols)
    assert want["tag"] == tag.decode()
    g_db = via_gpu.Prg(tag + b"/db")
    rb = g_db.fill(ctx, n_rec * pp.n_rec // 2)
    g = via_gpu.Prg(tag + b"/setup")
    keys = via_gpu.client_setup(ctx, pp, g)
    db = via_gpu.PreparedDb(ctx, pp, record_bytes=rb.tobytes(), n_records=n_rec)
    pk = via_gpu.PreparedKeys(ctx, pp, keys["cascade_key"], keys["conv_key"], keys["rsk"])
    server = via_gpu.Server(ctx, pp, db, pk)
    g_q = via_gpu.Prg(tag + b"/query")
    idxs = [0, n_rec - 1] + [g_q.uniform_below(ctx, n_rec) for _ in range(len(want["indices"]) - 2)]
    assert idxs == want["indices"]
    print("case: VIA-C %dx%d indices %s" % (rows, cols, idxs), flush=True)
    for idx, h in zip(idxs, want["sha256"]):
        mask, body = server.answer(via_gpu.client_query(ctx, pp, keys["sk1"], idx, g))
        assert digest(mask, body) == h, idx
        assert np.array_equal(via_gpu.client_recover(pp, keys["sk2"], mask, body), record_of(pp, rb, idx)), idx
    pk.close()
    db.close()


@pytest.mark.parametrize("T", [8, 16, 32, 64])
def test_via_b_full_size_bytes(ctx, golden, T):
    # I = 32, J = 4096 (2^17 cells, 1.25 GiB prepared), T records per answer
    want = golden["vb_t%d_32x4096" % T]
    pp = via_gpu.paper_b_params(32, 4096, T=T)
    d = pp.n1 // pp.n_rec
    n_rec = d * 32 * 4096
    tag = b"via-gpu/full/vb%d" % T
    assert want["tag"] == tag.decode()
    g_db = via_gpu.Prg(tag + b"/db")
    rb = g_db.fill(ctx, n_rec * pp.n_rec // 2)
    g = via_gpu.Prg(tag + b"/setup")
    keys = via_gpu.client_setup(ctx, pp, g)
    db = via_gpu.PreparedDb(ctx, pp, record_bytes=rb.tobytes(), n_records=n_rec)
    pk = via_gpu.PreparedKeys(ctx, pp, keys["cascade_key"], keys["conv_key"], keys["rsk"])
    server = via_gpu.Server(ctx, pp, db, pk)
    g_q = via_gpu.Prg(tag + b"/query")
    idxs = [0, n_rec - 1] + [g_q.uniform_below(ctx, n_rec) for _ in range(T - 2)]
    assert idxs == want["indices"]
    print("case: VIA-B T=%d 32x4096" % T, flush=True)
    queries = [via_gpu.client_query(ctx, pp, keys["sk1"], i, g) for i in idxs]
    mask, body = server.answer_batch(T, queries)
    assert digest(mask, body) == want["sha256"][0]
    plain = via_gpu.client_recover(pp, keys["sk2"], mask, body)
    got = via_gpu.client_deinterleave(pp, T, plain)
    for t, i in enumerate(idxs):
        assert np.array_equal(got[t], record_of(pp, rb, i)), (t, i)
    pk.close()
    db.close()


def test_paper_2x2_index15_bytes_from_library_inputs(ctx, golden):
    # client_server_e2e_paper.rs:287-349 with every input made by the library
    # (keys and query over Shake256Prg(b"via-c-paper-scale-e2e")): the answer
    # equals the oracle's word for word
    want = golden["paper_2x2_index15"]["answers"][0]
    pp = via_gpu.paper_params(2, 2)
    g = via_gpu.Prg(b"via-c-paper-scale-e2e")
    keys = via_gpu.client_setup(ctx, pp, g)
    recs = np.array([[(m + 1 + j) % 16 if j < 4 else 0 for j in range(512)] for m in range(16)], np.uint64)
    server = via_gpu.Server(ctx, pp, via_gpu.PreparedDb(ctx, pp, records=recs),
                            via_gpu.PreparedKeys(ctx, pp, keys["cascade_key"], keys["conv_key"], keys["rsk"]))
    mask, body = server.answer(via_gpu.client_query(ctx, pp, keys["sk1"], 15, g))
    assert mask.tolist() == want["mask"] and body.tolist() == want["body"]
    server.keys.close()
    server.db.close()
