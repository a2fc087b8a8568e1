"""PrgCompressed mask regeneration on the GPU against the CPU restatement.

via_prg_regenerate must equal Shake256Prg(seed).uniform_below(q) drawn count
times (format.rs:143-160, the reference's own test at format.rs:232-245 with
seed b"mask-seed" and q3), here from the oracle's PRG (pinned by the PRG
KATs), for moduli whose attempts take 1 to 8 bytes and reject from 0% to
~50% of draws, seeds of 0 to 64 bytes; query_from_prg must lay the limb-0
first RNS draws and the given bodies out as the answer path's LWE array, and
a query so expanded must answer exactly as the same LWEs sent whole.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Q1 = (137438822401, 274810798081)
Q2, Q3 = 17175674881, 8380417


@pytest.fixture(scope="module")
def ctx():
    import via_gpu

    c = via_gpu.Context(0)
    yield c
    c.close()


def oracle_draws(oracle, seed, qs, count):
    g = oracle.prg(seed)
    return np.array([g.uniform_below(qs[d % len(qs)]) for d in range(count)], dtype=np.uint64)


def test_regenerate_matches_reference_case(ctx, oracle):
    got = ctx.prg_regenerate(b"mask-seed", Q3, 4)
    assert np.array_equal(got, oracle_draws(oracle, b"mask-seed", [Q3], 4))


@pytest.mark.parametrize("q", [Q3, Q2, Q1[0], Q1[1], 2, 3, 2**40 + 1, 2**64 - 59])
def test_regenerate_matches_oracle(ctx, oracle, q):
    print(f"prg q={q}", flush=True)
    seed = b"via-gpu/prg/%d" % q
    got = ctx.prg_regenerate(seed, q, 60000)
    want = oracle_draws(oracle, seed, [q], 60000)
    assert np.array_equal(got, want)
    assert int(got.max()) < q


def test_regenerate_seed_lengths(ctx, oracle):
    for seed in (b"", bytes(range(32)), bytes(range(64))):
        assert np.array_equal(ctx.prg_regenerate(seed, Q2, 5000), oracle_draws(oracle, seed, [Q2], 5000))
    assert np.array_equal(ctx.prg_regenerate(b"x", 1, 7), np.zeros(7, np.uint64))  # uniform_below(1) = 0


def test_regenerate_rejects_long_seed(ctx):
    import via_gpu

    with pytest.raises(via_gpu.ViaError) as e:
        ctx.prg_regenerate(bytes(65), Q3, 4)
    assert e.value.status == via_gpu.VIA_E_PRG_SEED


def test_query_from_prg_layout_and_answer(ctx, oracle):
    import via_gpu
    from tests.paper_setup import paper_case

    case = paper_case(2, 2, seed=b"via-gpu/prg-query", index=3)
    pp = case.pp_gpu
    n, n_lwe = pp.n1, pp.query_lwes
    seed = bytes(range(100, 132))
    rng = np.random.default_rng(11)
    bodies = np.stack([rng.integers(0, Q1[0], n_lwe, dtype=np.uint64),
                       rng.integers(0, Q1[1], n_lwe, dtype=np.uint64)], axis=1).reshape(-1)
    d = ctx.query_from_prg(pp, seed, bodies)
    got = ctx.to_host(np.zeros(n_lwe * 2 * (n + 1), np.uint64), d).reshape(n_lwe, 2, n + 1)
    draws = oracle_draws(oracle, seed, list(Q1), n_lwe * n * 2).reshape(n_lwe, n, 2)
    want = np.zeros((n_lwe, 2, n + 1), np.uint64)
    want[:, :, :n] = draws.transpose(0, 2, 1)
    want[:, :, n] = bodies.reshape(n_lwe, 2)
    assert np.array_equal(got, want)
    # the expanded query answers exactly as the same LWEs sent whole
    server = via_gpu.Server.setup(
        via_gpu.ServerConfig(pp, case.cascade_key, case.conv_key, case.rsk), case.records, ctx=ctx)
    m_host, b_host = server.answer(want.reshape(-1))
    d_m, d_b = ctx.alloc(8 * pp.n2), ctx.alloc(8 * pp.n2)
    err = via_gpu.ErrDetail()
    st = via_gpu.lib().via_answer_device(ctx.h, server.db.h, server.keys.h, d, n_lwe, d_m, d_b, C.byref(err))
    assert st == 0
    ctx.synchronize()
    assert np.array_equal(ctx.to_host(np.zeros(pp.n2, np.uint64), d_m), m_host)
    assert np.array_equal(ctx.to_host(np.zeros(pp.n2, np.uint64), d_b), b_host)
    for x in (d, d_m, d_b):
        ctx.free(x)
