"""Client recovery in the library (host code, no GPU) against the oracle.

* via_client_recover is Client::recover (client.rs:222-239): decrypt_asymmetric
  (rlwe.rs:187-224) and decode (rlwe.rs:296-330) at (n2, q3, q4, p). The
  oracle's client_recover is pinned by the reference's layer KATs and its e2e
  recoveries; here the library's function must agree with it on every
  coefficient for random canonical inputs, and recover the records of toy
  VIA-C and VIA-B answers end to end.
* via_client_deinterleave is deinterleave_batch (batch.rs:31-52), pinned on
  the reference's hand vector (batch.rs:63-77) and on the oracle's project_at.
"""
import ctypes

import numpy as np
import pytest

import via_gpu
from oracle.vo import TERNARY, paper_b_params, paper_params, toy_b_params, toy_params, u64p


def P(a):
    return a.ctypes.data_as(u64p)


def arr(x):
    return np.ascontiguousarray(np.array(x, dtype=np.uint64))


def lib_params(opp):
    """The oracle's params struct as the library's (same field names)."""
    pp = via_gpu.Params()
    for name, _ in via_gpu.Params._fields_:
        v = getattr(opp, name)
        if name == "q1":
            pp.q1[0], pp.q1[1] = v[0], v[1]
        else:
            setattr(pp, name, v)
    return pp


def oracle_recover(oracle, opp, sk2, mask, body):
    out = arr([0] * opp.n2)
    oracle.client_recover(opp, P(arr(sk2)), P(arr(mask)), P(arr(body)), P(out))
    return out


def wide_params(q3):
    """The paper n2 = 512 with a q3 for which n2 (q3 - 1)^2 exceeds 2^64
    (the library's 128-bit accumulation)."""
    opp = paper_params(2, 2)
    opp.q3 = q3
    return opp


@pytest.mark.parametrize("which", ["toy", "paper", "wide_pow2", "wide_odd"])
def test_recover_matches_oracle_on_random_inputs(oracle, which):
    opp = {"toy": lambda: toy_params(), "paper": lambda: paper_params(2, 2),
           "wide_pow2": lambda: wide_params(1 << 36), "wide_odd": lambda: wide_params((1 << 35) + 13)}[which]()
    pp = lib_params(opp)
    rng = np.random.default_rng(20261203)
    n, q3, q4 = opp.n2, opp.q3, opp.q4
    for trial in range(3):
        if trial == 0:  # ternary S2, as the client draws it
            sk2 = rng.choice(np.array([0, 1, q3 - 1], dtype=np.uint64), size=n)
        else:  # any canonical S2 (the centred lift is taken at q3)
            sk2 = rng.integers(0, q3, size=n, dtype=np.uint64)
        mask = rng.integers(0, q3, size=n, dtype=np.uint64)
        body = rng.integers(0, q4, size=n, dtype=np.uint64)
        got = via_gpu.client_recover(pp, sk2, mask, body)
        want = oracle_recover(oracle, opp, sk2, mask, body)
        assert np.array_equal(got, want), (which, trial)
        assert int(got.max()) < opp.p


def test_recover_edge_values(oracle):
    """Extremes of every input: zero and q-1 masks and bodies, the rounding
    boundaries of the q3 -> q4 switch and of decode."""
    opp = paper_params(2, 2)
    pp = lib_params(opp)
    n, q3, q4 = opp.n2, opp.q3, opp.q4
    sk2 = np.zeros(n, np.uint64)
    sk2[0] = 1  # A' * S2 = A'
    mask = np.array([(i * 7919) % q3 for i in range(n)], dtype=np.uint64)
    mask[:4] = [0, q3 - 1, q3 // 2, q3 // 2 + 1]
    body = np.array([(i * 127) % q4 for i in range(n)], dtype=np.uint64)
    body[:6] = [0, q4 - 1, q4 // 2, q4 // 2 + 1, q4 // 32, q4 // 32 - 1]
    got = via_gpu.client_recover(pp, sk2, mask, body)
    assert np.array_equal(got, oracle_recover(oracle, opp, sk2, mask, body))
    # S2 = X^(n-1): the product wraps with a sign (X^n = -1)
    sk2[0], sk2[-1] = 0, q3 - 1
    got = via_gpu.client_recover(pp, sk2, mask, body)
    assert np.array_equal(got, oracle_recover(oracle, opp, sk2, mask, body))


def test_recover_rejects_non_canonical():
    pp = lib_params(toy_params())
    n = pp.n2
    z = np.zeros(n, np.uint64)
    for sk2, mask, body in [(z + pp.q3, z, z), (z, z + pp.q3, z), (z, z, z + pp.q4)]:
        with pytest.raises(via_gpu.ViaError) as e:
            via_gpu.client_recover(pp, sk2, mask, body)
        assert e.value.status == via_gpu.VIA_E_INVALID


def toy_record(m):
    # client_server_e2e.rs:56-60
    return [(m + 1) % 16, (2 * m + 1) % 16, (3 * m + 1) % 16, (m + 5) % 16]


def _setup(oracle, opp, seed):
    prg = oracle.prg(seed)
    sk1, sk2 = arr([0] * (opp.q1_nl * opp.n1)), arr([0] * opp.n2)
    ck = arr([0] * oracle.words_cascade_key(opp))
    cv = arr([0] * oracle.words_conv_key(opp))
    rsk = arr([0] * oracle.words_rsk(opp))
    assert oracle.client_setup(opp, TERNARY, TERNARY, TERNARY, prg.h, P(sk1), P(sk2), P(ck), P(cv), P(rsk)) == 0
    return prg, sk1, sk2, ck, cv, rsk


def _query(oracle, opp, sk1, index, prg):
    lwes = arr([0] * (oracle.query_lwes(opp) * oracle.words_lwe(opp)))
    assert oracle.client_query(opp, P(sk1), index, TERNARY, prg.h, P(lwes)) == 0
    return lwes


@pytest.mark.parametrize("index", [0, 3, 7])
def test_recover_toy_via_c_answers(oracle, index):
    """client_server_e2e.rs:92-172: the oracle's server answers, the library
    recovers the record."""
    opp = toy_params()
    pp = lib_params(opp)
    prg, sk1, sk2, ck, cv, rsk = _setup(oracle, opp, b"via-c-client-server-e2e")
    records = arr(sum([toy_record(m) for m in range(8)], []))
    h = ctypes.c_void_p()
    assert oracle.server_new(opp, P(records), 8, P(ck), P(cv), P(rsk), ctypes.byref(h)) == 0
    lwes = _query(oracle, opp, sk1, index, prg)
    am, ab = arr([0] * opp.n2), arr([0] * opp.n2)
    assert oracle.server_answer(h, P(lwes), oracle.query_lwes(opp), P(am), P(ab), None, None) == 0
    oracle.server_free(h)
    assert via_gpu.client_recover(pp, sk2, am, ab).tolist() == toy_record(index)


def test_deinterleave_hand_vector():
    # batch.rs:63-77: [1, 2, 3, 4] with N3 = 2 -> [1, 3], [2, 4]
    pp = via_gpu.Params()
    pp.n2, pp.n_rec = 4, 2
    got = via_gpu.client_deinterleave(pp, 2, arr([1, 2, 3, 4]))
    assert got.tolist() == [[1, 3], [2, 4]]


@pytest.mark.parametrize("T", [1, 3, 8])
def test_deinterleave_matches_project_at(oracle, T):
    opp = paper_b_params(2, 2, T=8)
    pp = lib_params(opp)
    pt = np.random.default_rng(T).integers(0, 16, size=opp.n2, dtype=np.uint64)
    got = via_gpu.client_deinterleave(pp, T, pt)
    for t in range(T):
        rec = arr([0] * opp.n_rec)
        oracle.L.vo_project_at(1, opp.n2, opp.n_rec, P(pt), t, P(rec))
        assert got[t].tolist() == rec.tolist()


def test_deinterleave_rejects_too_many_records():
    pp = lib_params(paper_b_params(2, 2, T=8))
    with pytest.raises(via_gpu.ViaError) as e:
        via_gpu.client_deinterleave(pp, 9, np.zeros(pp.n2, np.uint64))
    assert e.value.status == via_gpu.VIA_E_INVALID


def test_recover_toy_via_b_batch(oracle):
    """batch_e2e_toy.rs:40-225 (n1 = 64, n3 = 2, T = 8): the oracle's batch
    answer, recovered and deinterleaved by the library, in batch order."""
    opp = toy_b_params()
    pp = lib_params(opp)
    prg, sk1, sk2, ck, cv, rsk = _setup(oracle, opp, b"via-b-batch-e2e-toy")
    n_recs = opp.num_rows * opp.num_cols * (opp.n1 // opp.n_rec)
    recs = [[(m + 1) % 16, (2 * m + 3) % 16] for m in range(n_recs)]
    h = ctypes.c_void_p()
    assert oracle.server_new(opp, P(arr(sum(recs, []))), n_recs, P(ck), P(cv), P(rsk), ctypes.byref(h)) == 0
    idxs = [3, 11, 20, 38, 72, 89, 106, 127]
    nl = oracle.query_lwes(opp) * oracle.words_lwe(opp)
    lwes = arr([0] * (len(idxs) * nl))
    for t, idx in enumerate(idxs):
        lwes[t * nl:(t + 1) * nl] = _query(oracle, opp, sk1, idx, prg)
    am, ab = arr([0] * opp.n2), arr([0] * opp.n2)
    e, g = ctypes.c_size_t(), ctypes.c_size_t()
    assert oracle.server_answer_batch(h, 8, P(lwes), oracle.query_lwes(opp), P(am), P(ab), ctypes.byref(e),
                                      ctypes.byref(g)) == 0
    oracle.server_free(h)
    plain = via_gpu.client_recover(pp, sk2, am, ab)
    got = via_gpu.client_deinterleave(pp, 8, plain)
    assert got.tolist() == [recs[i] for i in idxs]


def test_prg_handles():
    # Shake256Prg seeds of 0 to 64 bytes; longer ones are InvalidPrgSeed
    for n in (0, 32, 64):
        g = via_gpu.Prg(bytes(range(n)))
        assert g.position == 0
        g.close()
    with pytest.raises(via_gpu.ViaError) as e:
        via_gpu.Prg(bytes(65))
    assert e.value.status == via_gpu.VIA_E_PRG_SEED


def test_client_key_words_match_oracle(oracle):
    for opp in (paper_params(2, 2), paper_b_params(4, 4, T=8)):
        assert via_gpu.client_key_words(lib_params(opp)) == (
            oracle.words_cascade_key(opp), oracle.words_conv_key(opp), oracle.words_rsk(opp))
