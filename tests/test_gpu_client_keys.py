"""Client key generation and query encryption in the library (Client::setup,
client.rs:89-173; Client::query, client.rs:183-207) against the oracle.

The oracle's client_setup and client_query are pinned by the reference's
e2e tests (the paper and toy recoveries under the reference's seeds) and its
layer KATs; the library must draw the same PRG stream in the same order, so
every key word and every query word is compared. The keys then drive the
HIP answer path and the library's recovery end to end.
"""
import ctypes

import numpy as np
import pytest

import via_gpu
from oracle.vo import Dist as ODist
from oracle.vo import paper_b_params, paper_params, u64p

pytestmark = pytest.mark.gpu


def P(a):
    return a.ctypes.data_as(u64p)


def arr(x):
    return np.ascontiguousarray(np.array(x, dtype=np.uint64))


def lib_params(opp):
    pp = via_gpu.Params()
    for name, _ in via_gpu.Params._fields_:
        v = getattr(opp, name)
        if name == "q1":
            pp.q1[0], pp.q1[1] = v[0], v[1]
        else:
            setattr(pp, name, v)
    return pp


@pytest.fixture(scope="module")
def ctx():
    c = via_gpu.Context(0)
    yield c
    c.close()


def oracle_setup(oracle, opp, prg, d1, d2, de):
    n1, n2 = opp.n1, opp.n2
    out = {"sk1": arr([0] * (2 * n1)), "sk2": arr([0] * n2),
           "cascade_key": arr([0] * oracle.words_cascade_key(opp)),
           "conv_key": arr([0] * oracle.words_conv_key(opp)), "rsk": arr([0] * oracle.words_rsk(opp))}
    assert oracle.client_setup(opp, d1, d2, de, prg.h, P(out["sk1"]), P(out["sk2"]), P(out["cascade_key"]),
                               P(out["conv_key"]), P(out["rsk"])) == 0
    return out


def oracle_query(oracle, opp, sk1, index, de, prg):
    q = arr([0] * (oracle.query_lwes(opp) * oracle.words_lwe(opp)))
    assert oracle.client_query(opp, P(sk1), index, de, prg.h, P(q)) == 0
    return q


def compare_setup(got, want):
    for k in ("sk1", "sk2", "cascade_key", "conv_key", "rsk"):
        assert got[k].shape == want[k].shape, k
        bad = np.flatnonzero(got[k] != want[k])
        assert bad.size == 0, f"{k}: {bad.size} words differ, first at {bad[:4]}"


def test_setup_and_queries_match_oracle_paper(ctx, oracle):
    # client_server_e2e_paper.rs:287-349: seed b"via-c-paper-scale-e2e", ternary
    # keys and errors; the query of index 15 continues the setup's stream, a
    # second query continues the first
    seed = b"via-c-paper-scale-e2e"
    opp = paper_params(2, 2)
    pp = lib_params(opp)
    og = oracle.prg(seed)
    want = oracle_setup(oracle, opp, og, ODist(0, 0), ODist(0, 0), ODist(0, 0))
    g = via_gpu.Prg(seed)
    got = via_gpu.client_setup(ctx, pp, g)
    compare_setup(got, want)
    for index in (15, 2):
        wq = oracle_query(oracle, opp, want["sk1"], index, ODist(0, 0), og)
        gq = via_gpu.client_query(ctx, pp, got["sk1"], index, g)
        assert np.array_equal(gq, wq), index
    g.close()


def test_bounded_uniform_distributions_match(ctx, oracle):
    # BoundedUniform keys and errors (distribution.rs:101-109), bound 0
    # drawing no bytes at all
    seed = b"via-gpu/keys/bounded"
    opp = paper_params(4, 2)
    pp = lib_params(opp)
    og = oracle.prg(seed)
    want = oracle_setup(oracle, opp, og, ODist(1, 2), ODist(1, 0), ODist(1, 7))
    g = via_gpu.Prg(seed)
    got = via_gpu.client_setup(ctx, pp, g, via_gpu.Dist(1, 2), via_gpu.Dist(1, 0), via_gpu.Dist(1, 7))
    compare_setup(got, want)
    wq = oracle_query(oracle, opp, want["sk1"], 5, ODist(1, 7), og)
    assert np.array_equal(via_gpu.client_query(ctx, pp, got["sk1"], 5, g, via_gpu.Dist(1, 7)), wq)


def test_via_b_batch_queries_match(ctx, oracle):
    # batch_e2e_paper.rs:38-174: T = 8 queries at the record degree n3 = 64,
    # in batch order from the one stream
    seed = b"via-b-paper-batch-e2e"
    opp = paper_b_params(2, 2, T=8)
    pp = lib_params(opp)
    og = oracle.prg(seed)
    want = oracle_setup(oracle, opp, og, ODist(0, 0), ODist(0, 0), ODist(0, 0))
    g = via_gpu.Prg(seed)
    got = via_gpu.client_setup(ctx, pp, g)
    compare_setup(got, want)
    for idx in [3, 17, 42, 5, 88, 100, 7, 120]:
        assert np.array_equal(via_gpu.client_query(ctx, pp, got["sk1"], idx, g),
                              oracle_query(oracle, opp, want["sk1"], idx, ODist(0, 0), og)), idx


def test_prg_fill_and_uniform_below_match(ctx, oracle):
    # Shake256Prg::fill_bytes and uniform_below (prg.rs:143-207) interleaved,
    # across GPU stream windows (5 MB > one 4.25 MiB window)
    seed = b"via-gpu/prg/fill"
    og = oracle.prg(seed)
    g = via_gpu.Prg(seed)
    for step, (kind, arg) in enumerate([("fill", 100), ("draw", (1 << 37) + 12345), ("fill", 5_000_000),
                                        ("draw", 3), ("draw", 1), ("fill", 7), ("draw", 8380417)]):
        if kind == "fill":
            want = ctypes.create_string_buffer(arg)
            oracle.L.vo_prg_fill(og.h, want, arg)
            assert g.fill(ctx, arg).tobytes() == want.raw, step
        else:
            assert g.uniform_below(ctx, arg) == oracle.L.vo_prg_uniform_below(og.h, arg), step


def test_library_client_end_to_end(ctx):
    # keys and queries from the library, answered on the GPU, recovered by
    # the library: every queried record comes back
    pp = via_gpu.paper_params(8, 4)
    rng = np.random.default_rng(99)
    n_rec = 4 * 8 * 4
    records = rng.integers(0, 16, size=(n_rec, 512), dtype=np.uint64)
    g = via_gpu.Prg(b"via-gpu/keys/e2e")
    keys = via_gpu.client_setup(ctx, pp, g)
    server = via_gpu.Server.setup(via_gpu.ServerConfig(pp, keys["cascade_key"], keys["conv_key"], keys["rsk"]),
                                  records.reshape(-1), ctx=ctx)
    for idx in (0, 31, 77, n_rec - 1):
        q = via_gpu.client_query(ctx, pp, keys["sk1"], idx, g)
        mask, body = server.answer(q)
        assert np.array_equal(via_gpu.client_recover(pp, keys["sk2"], mask, body), records[idx]), idx


def test_client_errors(ctx):
    pp = via_gpu.paper_params(2, 2)
    g = via_gpu.Prg(b"via-gpu/keys/errors")
    sk1 = np.zeros(2 * pp.n1, np.uint64)
    with pytest.raises(via_gpu.ViaError) as e:  # index beyond D * I * J
        via_gpu.client_query(ctx, pp, sk1, 4 * 2 * 2, g)
    assert e.value.status == via_gpu.VIA_E_INDEX
    bad = via_gpu.paper_params(3, 2)
    with pytest.raises(via_gpu.ViaError) as e:  # client.rs:125-131
        via_gpu.client_setup(ctx, bad, g)
    assert e.value.status == via_gpu.VIA_E_DIM
    with pytest.raises(via_gpu.ViaError) as e:  # no bit decomposition of 3 rows
        via_gpu.client_query(ctx, bad, sk1, 0, g)
    assert e.value.status == via_gpu.VIA_E_DIM
    with pytest.raises(via_gpu.ViaError) as e:  # the Gaussian arm
        via_gpu.client_setup(ctx, pp, g, err=via_gpu.Dist(2, 0))
    assert e.value.status == via_gpu.VIA_E_INVALID
    assert g.position == 0  # failed calls leave the stream where it was


def test_setup_many_matches_one_at_a_time():
    # via_client_setup_many over six clients (host threads on lanes of one
    # context): every word of every client's keys equals via_client_setup on
    # the same stream, the streams end at the same positions, and the
    # queries that continue them are the same
    pp = via_gpu.paper_params(8, 8)
    ctx = via_gpu.Context(0)
    try:
        seeds = [b"via-gpu/setup-many/%d" % i for i in range(6)]
        one = []
        pos = []
        for sd in seeds:
            g = via_gpu.Prg(sd)
            one.append(via_gpu.client_setup(ctx, pp, g))
            pos.append(g.position)
            one[-1]["q"] = via_gpu.client_query(ctx, pp, one[-1]["sk1"], 17, g)
            g.close()
        prgs = [via_gpu.Prg(sd) for sd in seeds]
        many = via_gpu.client_setup_many(ctx, pp, prgs)
        for i, (a, b, g) in enumerate(zip(one, many, prgs)):
            for k in ("sk1", "sk2", "cascade_key", "conv_key", "rsk"):
                assert np.array_equal(a[k], b[k]), (i, k)
            assert g.position == pos[i]
            assert np.array_equal(via_gpu.client_query(ctx, pp, b["sk1"], 17, g), a["q"]), i
        for g in prgs:
            g.close()
        # the same stream twice is refused (one sequential stream per client)
        g = via_gpu.Prg(b"via-gpu/setup-many/dup")
        with pytest.raises(via_gpu.ViaError) as e:
            via_gpu.client_setup_many(ctx, pp, [g, g])
        assert e.value.status == via_gpu.VIA_E_INVALID
        g.close()
    finally:
        ctx.close()
