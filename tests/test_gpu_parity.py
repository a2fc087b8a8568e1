"""Parity of the HIP path with the CPU restatement (bit-exact, integer work).

Run on an MI355X: pytest -m gpu. Every comparison is exact equality of u64
residues; the paper-stack cases also check client recovery of the record.
"""
import ctypes as C

import numpy as np
import pytest

from tests.paper_setup import paper_case

pytestmark = pytest.mark.gpu

PRIMES = [137438822401, 274810798081, 17175674881, 8380417, 268369921, 536608769]


@pytest.fixture(scope="module")
def ctx():
    import via_gpu

    c = via_gpu.Context(0)
    yield c
    c.close()


def P(a):
    from oracle.vo import u64p

    return a.ctypes.data_as(u64p)


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048])
def test_ntt_matches_oracle(ctx, oracle, n):
    from oracle.vo import Ring

    rng = np.random.default_rng(n)
    for q in PRIMES:
        if (q - 1) % (2 * n):
            continue
        count = max(1, 4096 // n)
        a = rng.integers(0, q, size=count * n, dtype=np.uint64)
        got = ctx.ntt(n, [q], a)
        want = a.copy()
        r = Ring.make(n, q)
        for p in range(count):
            seg = np.ascontiguousarray(want[p * n:(p + 1) * n])
            oracle.ntt_fwd(r, P(seg))
            want[p * n:(p + 1) * n] = seg
        assert np.array_equal(got, want), (n, q)
        back = ctx.ntt(n, [q], got, inverse=True)
        assert np.array_equal(back, a), (n, q)


def test_poly_mul_microbench_shape(ctx, oracle):
    # negacyclic NTT -> pointwise -> iNTT for the two VIA q1 primes at n=2048
    from oracle.vo import Ring

    n, count, qs = 2048, 16, [268369921, 536608769]
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.integers(0, qs[p // 8], size=n, dtype=np.uint64) for p in range(count)])
    b = np.concatenate([rng.integers(0, qs[p // 8], size=n, dtype=np.uint64) for p in range(count)])
    da, db, dc = ctx.alloc(a.nbytes), ctx.alloc(b.nbytes), ctx.alloc(a.nbytes)
    ctx.to_device(da, a)
    ctx.to_device(db, b)
    ctx.poly_mul_device(n, qs, 8, da, db, dc, count)
    ctx.synchronize()
    c = ctx.to_host(np.zeros_like(a), dc)
    for p in range(count):
        r = Ring.make(n, qs[p // 8])
        want = np.zeros(n, np.uint64)
        oracle.poly_mul(r, P(want), P(np.ascontiguousarray(a[p * n:(p + 1) * n])),
                        P(np.ascontiguousarray(b[p * n:(p + 1) * n])))
        assert np.array_equal(c[p * n:(p + 1) * n], want), p
    for d in (da, db, dc):
        ctx.free(d)


def run_case(ctx, case):
    import via_gpu

    server = via_gpu.Server.setup(
        via_gpu.ServerConfig(case.pp_gpu, case.cascade_key, case.conv_key, case.rsk), case.records, ctx=ctx)
    assert np.array_equal(server.db.cells(), case.oracle_db), "prepared database differs"
    mask, body = server.answer(case.lwes)
    return server, mask, body


@pytest.fixture(scope="module")
def case_2x2_15():
    return paper_case(2, 2, seed=b"via-c-paper-scale-e2e", index=15, stages=True)


def check_stages(ctx, case, mask, body):
    n1 = 2048
    for k, words in [(1, 4 * 2 * 2 * 2 * 2 * n1), (2, 2 * 2 * 2 * n1), (3, 2 * 2 * n1), (4, 2 * 2 * n1),
                     (5, 2 * n1), (6, 2 * n1)]:
        got = ctx.stage(k, words)
        assert np.array_equal(got, case.st[k]), f"stage {k} differs"
    assert np.array_equal(mask, case.ans_mask) and np.array_equal(body, case.ans_body)
    assert np.array_equal(case.recover(mask, body), case.record)


def test_answer_paper_2x2_index15(ctx, case_2x2_15):
    # client_server_e2e_paper.rs:316-349, answered on the GPU.
    server, mask, body = run_case(ctx, case_2x2_15)
    check_stages(ctx, case_2x2_15, mask, body)


def test_step_timing_off(ctx, case_2x2_15):
    # answers that record only the FirstDim events: the same words and
    # stages; the step split is refused until the events are on again
    import via_gpu

    ctx.set_step_timing(False)
    try:
        server, mask, body = run_case(ctx, case_2x2_15)
        check_stages(ctx, case_2x2_15, mask, body)
        with pytest.raises(via_gpu.ViaError):
            ctx.step_times()
        assert ctx.first_dim_time() > 0
    finally:
        ctx.set_step_timing(True)
    server, mask, body = run_case(ctx, case_2x2_15)
    assert np.array_equal(mask, case_2x2_15.ans_mask) and np.array_equal(body, case_2x2_15.ans_body)
    assert len(ctx.step_times()) == 7


def test_exact_paths_match_oracle(ctx, case_2x2_15):
    # the fp64 fast paths of the gadget scale (one limb, two limbs, the wide
    # cascade split) and of the q1 -> q2 mod switch decline values next to a
    # rounding boundary and hand them to the exact 128/192-bit paths; random
    # values almost never reach those, so the switch sends every coefficient
    # through them: every stage stays word-exact with the oracle
    import via_gpu

    L = via_gpu.lib()
    assert L.via_debug_set_exact_paths(1) == 0
    try:
        server, mask, body = run_case(ctx, case_2x2_15)
        check_stages(ctx, case_2x2_15, mask, body)
    finally:
        L.via_debug_set_exact_paths(0)


@pytest.mark.parametrize("rows,cols,index", [(4, 8, 77), (8, 4, 3), (16, 16, 1000)])
def test_answer_paper_random_db(ctx, rows, cols, index):
    case = paper_case(rows, cols, seed=b"via-gpu/parity/%d" % index, index=index, records="random",
                      rec_seed=b"via-gpu/db/%d" % index)
    server, mask, body = run_case(ctx, case)
    assert np.array_equal(mask, case.ans_mask) and np.array_equal(body, case.ans_body)
    assert np.array_equal(case.recover(mask, body), case.record)


def test_answer_ragged_db(ctx):
    # fewer records than cells: setup_db zero-pads (setup_db.rs:39-75); query
    # a present record and a padded slot.
    for index in (5, 30):
        case = paper_case(4, 4, seed=b"via-gpu/ragged", index=index, records="random", n_records=21,
                          rec_seed=b"via-gpu/ragged-db")
        server, mask, body = run_case(ctx, case)
        assert np.array_equal(mask, case.ans_mask) and np.array_equal(body, case.ans_body)
        assert np.array_equal(case.recover(mask, body), case.record)


def test_answer_query_length_error(ctx):
    import via_gpu

    case = paper_case(2, 2, seed=b"via-gpu/err", index=0)
    server, _, _ = run_case(ctx, case)
    with pytest.raises(via_gpu.ViaError) as e:
        server.answer(case.lwes[: -case.pp_gpu.lwe_words])
    assert e.value.status == via_gpu.VIA_E_QUERY_LEN
    assert (e.value.expected, e.value.got) == (8, 7)


def test_answer_batch_mixed_clients(ctx):
    # Q queries from different clients answered in one batched pass: each
    # answer equals the oracle's answer for that query alone.
    import via_gpu

    cases = [paper_case(8, 8, seed=b"via-gpu/batch/client%d" % q, index=(37 * q + 5) % 256, records="random",
                        rec_seed=b"via-gpu/batch-db") for q in range(5)]
    pp = cases[0].pp_gpu
    db = via_gpu.PreparedDb(ctx, pp, records=cases[0].records)
    keys = [via_gpu.PreparedKeys(ctx, pp, c.cascade_key, c.conv_key, c.rsk) for c in cases]
    masks, bodies = via_gpu.answer_many(ctx, db, keys, [c.lwes for c in cases], pp)
    for c, m, b in zip(cases, masks, bodies):
        assert np.array_equal(m, c.ans_mask) and np.array_equal(b, c.ans_body)
        assert np.array_equal(c.recover(m, b), c.record)
    # the same query alone gives the same bytes
    server = via_gpu.Server(ctx, pp, db, keys[2])
    m2, b2 = server.answer(cases[2].lwes)
    assert np.array_equal(m2, masks[2]) and np.array_equal(b2, bodies[2])


def test_autotuned_answers_match(ctx):
    # the autotuner changes only how the digits of each gadget product are
    # split over workgroups: after tuning for a batch size, the batched
    # answers still equal the oracle's and the per-step choices are counts
    # the tuner offers
    import via_gpu

    cases = [paper_case(4, 4, seed=b"via-gpu/tune/client%d" % q, index=(11 * q + 3) % 64, records="random",
                        rec_seed=b"via-gpu/tune-db") for q in range(3)]
    pp = cases[0].pp_gpu
    tctx = via_gpu.Context(0)
    db = via_gpu.PreparedDb(tctx, pp, records=cases[0].records)
    keys = [via_gpu.PreparedKeys(tctx, pp, c.cascade_key, c.conv_key, c.rsk) for c in cases]
    n_lwe = pp.query_lwes
    d_q = [tctx.alloc(c.lwes.nbytes) for c in cases]
    for d, c in zip(d_q, cases):
        tctx.to_device(d, c.lwes)
    d_m = [tctx.alloc(8 * pp.n2) for _ in cases]
    d_b = [tctx.alloc(8 * pp.n2) for _ in cases]
    shapes = tctx.autotune(db, keys, d_q, n_lwe, d_m, d_b, reps=1)
    assert len(shapes) == 12
    assert all((v in (0, 2, 3, 4, 5) and g in (1, 2, 3, 4, 6, 9, 18)) or (v == 1 and g in (1, 2, 3, 5))
               for g, v in shapes)
    masks, bodies = via_gpu.answer_many(tctx, db, keys, [c.lwes for c in cases], pp)
    for c, m, b in zip(cases, masks, bodies):
        assert np.array_equal(m, c.ans_mask) and np.array_equal(b, c.ans_body)
    for d in d_q + d_m + d_b:
        tctx.free(d)
    tctx.close()


def test_every_mac_variant_matches(ctx):
    # each gadget-MAC kernel (256 threads radix 8 / radix 16, 512 threads,
    # 1024 threads, 256 and 512 threads with one workgroup per limb) pinned
    # for every decompression step gives the oracle's answers; shapes the
    # tuner never offers are refused
    import via_gpu

    cases = [paper_case(4, 4, seed=b"via-gpu/variant/client%d" % q, index=(5 * q + 9) % 64, records="random",
                        rec_seed=b"via-gpu/variant-db") for q in range(2)]
    pp = cases[0].pp_gpu
    vctx = via_gpu.Context(0)
    db = via_gpu.PreparedDb(vctx, pp, records=cases[0].records)
    keys = [via_gpu.PreparedKeys(vctx, pp, c.cascade_key, c.conv_key, c.rsk) for c in cases]
    try:
        for variant, groups in [(0, 2), (1, 2), (2, 3), (3, 1), (3, 6), (4, 1), (4, 3), (5, 2), (5, 18)]:
            vctx.set_tuning(2, [(groups, variant)] * 12)
            masks, bodies = via_gpu.answer_many(vctx, db, keys, [c.lwes for c in cases], pp)
            for c, m, b in zip(cases, masks, bodies):
                assert np.array_equal(m, c.ans_mask) and np.array_equal(b, c.ans_body), (variant, groups)
        for bad in [(7, 0), (4, 1), (2, 6)]:
            with pytest.raises(via_gpu.ViaError):
                vctx.set_tuning(2, [bad] * 12)
    finally:
        vctx.set_tuning(2, None)
        vctx.close()


def test_graph_replays_match_eager():
    # hipGraph mode: the first call of a shape runs eagerly and is captured;
    # later identical calls replay the capture. Every replay, with eager work
    # of another shape and host round trips between the replays, must equal
    # the eager answers (and the oracle's).
    import via_gpu

    cases = [paper_case(4, 4, seed=b"via-gpu/graph/client%d" % q, index=(7 * q + 2) % 64, records="random",
                        rec_seed=b"via-gpu/graph-db") for q in range(3)]
    pp = cases[0].pp_gpu
    g = via_gpu.Context(0)
    db = via_gpu.PreparedDb(g, pp, records=cases[0].records)
    keys = [via_gpu.PreparedKeys(g, pp, c.cascade_key, c.conv_key, c.rsk) for c in cases]
    d_q = [g.alloc(c.lwes.nbytes) for c in cases]
    for d, c in zip(d_q, cases):
        g.to_device(d, c.lwes)
    d_m = [g.alloc(8 * pp.n2) for _ in cases]
    d_b = [g.alloc(8 * pp.n2) for _ in cases]
    L = via_gpu.lib()
    err = via_gpu.ErrDetail()
    vp = C.c_void_p

    def many(idx):
        n = len(idx)
        K = (vp * n)(*[keys[i].h.value for i in idx])
        Qa = (vp * n)(*[d_q[i].value for i in idx])
        M = (vp * n)(*[d_m[i].value for i in idx])
        B = (vp * n)(*[d_b[i].value for i in idx])
        assert L.via_answer_many_device(g.h, db.h, n, K, Qa, pp.query_lwes, M, B, C.byref(err)) == 0
        g.synchronize()

    def read(i):
        return (g.to_host(np.zeros(pp.n2, np.uint64), d_m[i]), g.to_host(np.zeros(pp.n2, np.uint64), d_b[i]))

    def clear(i):
        g.to_device(d_m[i], np.zeros(pp.n2, np.uint64))
        g.to_device(d_b[i], np.zeros(pp.n2, np.uint64))

    many([0, 1])
    eager = [read(0), read(1)]
    many([2])
    eager.append(read(2))
    for c, (m, b) in zip(cases, eager):
        assert np.array_equal(m, c.ans_mask) and np.array_equal(b, c.ans_body)
    g.set_graphs(True)
    for rep in range(4):
        for i in range(3):
            clear(i)
        many([0, 1])  # rep 0: eager and captured; later: replayed
        many([2])     # a second shape between the replays
        for i in range(3):
            m, b = read(i)
            assert np.array_equal(m, eager[i][0]) and np.array_equal(b, eager[i][1]), (rep, i)
    g.set_graphs(False)
    for d in d_q + d_m + d_b:
        g.free(d)
    g.close()
