"""Persistent trees on the GPU: every DMux level, and every CMux level, in
one cooperative launch (mux.rs:167-251 as one kernel each).

* Forced on, the answers (one query, and five clients' queries in one pass)
  equal the oracle's word for word, and equal the level-by-level launches.
* A barrier whose waits may not poll at all (spin 0) gives up as soon as a
  level is unfinished: the call reports VIA_E_DEVICE, every kernel still
  exits, and the next answer on the context is exact again.
* The autotuner records a tree mode for the batch size it measured.
"""
import ctypes as C

import numpy as np
import pytest

from tests.paper_setup import paper_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case16():
    return paper_case(16, 16, seed=b"via-gpu/tree/16", index=900, records="random", rec_seed=b"via-gpu/tree/db")


def test_persistent_trees_match_oracle(case16):
    import via_gpu

    ctx = via_gpu.Context(0)
    try:
        c = case16
        server = via_gpu.Server.setup(via_gpu.ServerConfig(c.pp_gpu, c.cascade_key, c.conv_key, c.rsk), c.records,
                                      ctx=ctx)
        for mode in (1, 0, 1):
            ctx.set_tree_mode(mode)
            print("case: 16x16 index 900, tree mode %d" % mode, flush=True)
            m, b = server.answer(c.lwes)
            assert np.array_equal(m, c.ans_mask) and np.array_equal(b, c.ans_body), mode
        assert np.array_equal(c.recover(m, b), c.record)
    finally:
        ctx.close()


def test_persistent_trees_batched_clients():
    import via_gpu

    cases = [paper_case(8, 8, seed=b"via-gpu/tree/client%d" % q, index=(29 * q + 7) % 256, records="random",
                        rec_seed=b"via-gpu/tree/batch-db") for q in range(5)]
    pp = cases[0].pp_gpu
    ctx = via_gpu.Context(0)
    try:
        db = via_gpu.PreparedDb(ctx, pp, records=cases[0].records)
        keys = [via_gpu.PreparedKeys(ctx, pp, c.cascade_key, c.conv_key, c.rsk) for c in cases]
        ctx.set_tree_mode(1)
        print("case: 8x8, five clients in one pass, persistent trees", flush=True)
        masks, bodies = via_gpu.answer_many(ctx, db, keys, [c.lwes for c in cases], pp)
        for c, m, b in zip(cases, masks, bodies):
            assert np.array_equal(m, c.ans_mask) and np.array_equal(b, c.ans_body)
    finally:
        ctx.close()


def test_barrier_that_gives_up_is_reported(case16):
    import via_gpu

    L = via_gpu.lib()
    ctx = via_gpu.Context(0)
    try:
        c = case16
        server = via_gpu.Server.setup(via_gpu.ServerConfig(c.pp_gpu, c.cascade_key, c.conv_key, c.rsk), c.records,
                                      ctx=ctx)
        ctx.set_tree_mode(1)
        assert L.via_debug_set_tree_spin(ctx.h, 0) == 0
        print("case: 16x16, persistent trees, barrier spin 0", flush=True)
        with pytest.raises(via_gpu.ViaError) as e:
            server.answer(c.lwes)
        assert e.value.status == via_gpu.VIA_E_DEVICE
        assert L.via_debug_set_tree_spin(ctx.h, 1 << 24) == 0
        m, b = server.answer(c.lwes)
        assert np.array_equal(m, c.ans_mask) and np.array_equal(b, c.ans_body)
    finally:
        ctx.close()


def test_autotune_records_tree_mode(case16):
    import via_gpu

    ctx = via_gpu.Context(0)
    try:
        c = case16
        pp = c.pp_gpu
        db = via_gpu.PreparedDb(ctx, pp, records=c.records)
        keys = via_gpu.PreparedKeys(ctx, pp, c.cascade_key, c.conv_key, c.rsk)
        dq = ctx.alloc(c.lwes.nbytes)
        ctx.to_device(dq, c.lwes)
        dm, dbd = ctx.alloc(8 * pp.n2), ctx.alloc(8 * pp.n2)
        ctx.autotune(db, [keys], [dq], pp.query_lwes, [dm], [dbd], reps=1)
        t = ctx.tree_tuning(1)
        assert t is not None and t["dmux"] in (0, 1) and t["cmux"] in (0, 1)
        # answers with the stored choice are exact
        m, b = via_gpu.Server(ctx, pp, db, keys).answer(c.lwes)
        assert np.array_equal(m, c.ans_mask) and np.array_equal(b, c.ans_body)
        for d in (dq, dm, dbd):
            ctx.free(d)
    finally:
        ctx.close()
