"""FirstDim outputs in pair form (k_fd_intt_pairs) feeding the first CMux
level (GATE_CMUX_EV gates): one inverse transform per column pair, c0 kept
in eval form and added by the gate before its own inverse transform.

The pass takes that form when the first CMux level runs the fixed-shape gate
level by level (at least 256 gates: four clients x J/2 = 64 here, J a
multiple of 16). Four clients' queries at I=8, J=128 in one pass:

* every answer equals the oracle's (server_answer) word for word;
* query 0's FirstDim outputs, brought back to coefficient form from the
  pair form (via_debug_stage 4), equal the oracle's stage 4 word for word;
* with persistent trees (no pair form) the same pass gives the same answers.
"""
import numpy as np
import pytest

from tests.paper_setup import paper_case

pytestmark = pytest.mark.gpu


def test_first_dim_pairs_feed_cmux():
    import via_gpu

    Q, I, J = 4, 8, 128
    cases = []
    for q in range(Q):
        print("oracle: client %d of %d" % (q + 1, Q), flush=True)
        cases.append(paper_case(I, J, seed=b"via-gpu/fdpairs/client%d" % q, index=(37 * q + 11) % (4 * I * J),
                                records="random", rec_seed=b"via-gpu/fdpairs/db", stages=(q == 0)))
    pp = cases[0].pp_gpu
    n1 = pp.n1
    ctx = via_gpu.Context(0)
    try:
        db = via_gpu.PreparedDb(ctx, pp, records=cases[0].records)
        keys = [via_gpu.PreparedKeys(ctx, pp, c.cascade_key, c.conv_key, c.rsk) for c in cases]
        for mode in (0, 1):
            ctx.set_tree_mode(mode)
            print("case: %dx%d, %d clients in one pass, tree mode %d" % (I, J, Q, mode), flush=True)
            masks, bodies = via_gpu.answer_many(ctx, db, keys, [c.lwes for c in cases], pp)
            for q, (c, m, b) in enumerate(zip(cases, masks, bodies)):
                assert np.array_equal(m, c.ans_mask) and np.array_equal(b, c.ans_body), (mode, q)
            st4 = ctx.stage(4, J * 2 * n1)
            assert np.array_equal(st4, cases[0].st[4]), mode
        assert np.array_equal(cases[0].recover(masks[0], bodies[0]), cases[0].record)
    finally:
        ctx.close()
