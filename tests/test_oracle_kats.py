"""Pin the CPU restatement (oracle/) against the reference's own known answers.

Each test mirrors a via-rs test (cited) with the same seeds and draw order and
compares against the constants of tests/golden/reference_kats.json, which
tests/golden/make_kat_fixtures.py extracted from the reference's data files.
Where the reference has no constant, the check is a closed form or a brute
force computed here independently of the oracle (Python big integers).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle.vo import TERNARY, Ring, zeros

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def u64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def i64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


# ---------------------------------------------------------------- PRG -------


def test_prg_byte_stream(oracle):
    # prg.rs:286-329
    g = oracle.prg(b"")
    assert list(g.fill(32)) == GOLD["prg"]["EMPTY_SEED_FIRST_32"]
    g = oracle.prg(b"test")
    assert list(g.fill(32)) == GOLD["prg"]["TEST_SEED_FIRST_32"]
    g = oracle.prg(b"test")
    g.fill(134)
    assert list(g.fill(4)) == GOLD["prg"]["TEST_SEED_BOUNDARY_134_138"]


def test_prg_matches_hashlib_shake256(oracle):
    # Independent SHAKE-256 (Python hashlib) over several blocks and split reads.
    seed = b"split-test"
    want = b"".join(hashlib.shake_256(seed + k.to_bytes(8, "little")).digest(136) for k in range(5))
    g = oracle.prg(seed)
    got = b"".join(g.fill(c) for c in [1, 134, 1, 5, 130, 1, 28, 380])
    assert got == want[: len(got)]


@pytest.mark.parametrize(
    "bound,key",
    [
        (3, "TEST_SEED_RANDBELOW_3_FIRST_8"),
        (256, "TEST_SEED_RANDBELOW_256_FIRST_8"),
        (8380417, "TEST_SEED_RANDBELOW_Q3_FIRST_4"),
        (1 << 32, "TEST_SEED_RANDBELOW_2P32_FIRST_4"),
        ((1 << 32) + 1, "TEST_SEED_RANDBELOW_2P32_PLUS_1_FIRST_4"),
    ],
)
def test_uniform_below(oracle, bound, key):
    g = oracle.prg(b"test")
    want = GOLD["prg"][key]
    assert [g.uniform_below(bound) for _ in want] == want


def test_uniform_below_one_draws_nothing(oracle):
    a, b = oracle.prg(b"check-state"), oracle.prg(b"check-state")
    for _ in range(10):
        assert a.uniform_below(1) == 0
    assert a.fill(16) == b.fill(16)


@pytest.mark.parametrize("seed,key", [(b"test", "TEST_SEED_TP_N16"), (b"", "EMPTY_SEED_TP_N16")])
def test_ternary(oracle, seed, key):
    # ternary.rs:61-64
    g = oracle.prg(seed)
    out = i64(np.zeros(16))
    oracle.sample(g.h, TERNARY, out.ctypes.data_as(oracle_i64p()), 16)
    assert list(out) == GOLD["ternary"][key]


def oracle_i64p():
    from oracle.vo import i64p

    return i64p


# ---------------------------------------------------------------- NTT -------


def test_psi_closed_forms(oracle):
    # ntt.rs:492-527: psi(17, N=4) = 9; every paper prime admits psi of order 2N.
    assert oracle.ntt_psi(17, 4) == 9
    for q in [268369921, 536608769, 137438822401, 274810798081, 17175674881, 8380417]:
        for n in [4, 512, 2048]:
            psi = oracle.ntt_psi(q, n)
            assert pow(psi, 2 * n, q) == 1 and pow(psi, n, q) != 1


def _bitrev(i, logn):
    return int(format(i, f"0{logn}b")[::-1], 2) if logn else 0


@pytest.mark.parametrize("q,n", [(17, 4), (17, 8), (8380417, 8), (137438822401, 16), (17175674881, 64)])
def test_ntt_bit_reversed_evaluations(oracle, q, n):
    # ntt.rs:604-702: lane i holds f(psi^(2*br(i)+1)), checked by direct evaluation.
    rng = np.random.default_rng(q + n)
    coeffs = [int(x) for x in rng.integers(0, q, n, dtype=np.uint64)]
    r = Ring.make(n, q)
    buf = u64(coeffs)
    oracle.ntt_fwd(r, buf.ctypes.data_as(_u64p()))
    psi = oracle.ntt_psi(q, n)
    logn = n.bit_length() - 1
    for i in range(n):
        x = pow(psi, 2 * _bitrev(i, logn) + 1, q)
        assert int(buf[i]) == sum(c * pow(x, k, q) for k, c in enumerate(coeffs)) % q
    oracle.ntt_inv(r, buf.ctypes.data_as(_u64p()))
    assert [int(v) for v in buf] == coeffs


def _u64p():
    from oracle.vo import u64p

    return u64p


def schoolbook(a, b, q, n):
    out = [0] * n
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            k = i + j
            if k < n:
                out[k] += x * y
            else:
                out[k - n] -= x * y
    return [v % q for v in out]


@pytest.mark.parametrize(
    "qs,n",
    [((17,), 8), ((65537,), 64), ((1 << 36,), 8), ((137438822401, 274810798081), 32), ((8380417,), 512)],
)
def test_poly_mul_matches_schoolbook(oracle, qs, n):
    # ntt.rs:817-842 (convolution homomorphism); RNS limbs checked per prime.
    r = Ring.make(n, *qs)
    rng = np.random.default_rng(n)
    a = [[int(v) for v in rng.integers(0, q, n, dtype=np.uint64)] for q in qs]
    b = [[int(v) for v in rng.integers(0, q, n, dtype=np.uint64)] for q in qs]
    A, B, O = u64(sum(a, [])), u64(sum(b, [])), zeros(len(qs) * n)
    oracle.poly_mul(r, O.ctypes.data_as(_u64p()), A.ctypes.data_as(_u64p()), B.ctypes.data_as(_u64p()))
    for l, q in enumerate(qs):
        assert [int(v) for v in O[l * n:(l + 1) * n]] == schoolbook(a[l], b[l], q, n)


# ---------------------------------------------------------------- gadget ----


def gadget_vector(oracle, r, base, L):
    out = zeros(2 * L)
    oracle.gadget_vector(r, base, L, out.ctypes.data_as(_u64p()))
    return [int(out[2 * i]) | (int(out[2 * i + 1]) << 64) for i in range(L)]


def test_gadget_vector_hand_values(oracle):
    # gadget.rs:267-300
    assert gadget_vector(oracle, Ring.make(4, 17), 2, 4) == [9, 4, 2, 1]
    assert gadget_vector(oracle, Ring.make(4, 1024), 2, 10) == [512, 256, 128, 64, 32, 16, 8, 4, 2, 1]
    Q = 137438822401 * 274810798081
    g = gadget_vector(oracle, Ring.make(16, 137438822401, 274810798081), 18, 18)
    assert g == [(Q + 18 ** (i + 1) // 2) // 18 ** (i + 1) for i in range(18)]


def decompose(oracle, r, coeffs_limbs, base, L):
    inp = u64(coeffs_limbs)
    out = i64(np.zeros(L * r.n))
    from oracle.vo import i64p

    oracle.gadget_decompose(r, inp.ctypes.data_as(_u64p()), base, L, out.ctypes.data_as(i64p))
    return out.reshape(L, r.n)


def test_gadget_decompose_hand_vector(oracle):
    # gadget.rs:481-510
    d = decompose(oracle, Ring.make(4, 17), [5, 12, 0, 8], 2, 4)
    assert d.tolist() == [[0, 1, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0], [1, 1, 0, 0]]


def py_digits(c_centered, Q, base, L):
    """Independent restatement with Python integers: round-half-away scale,
    Euclidean digit, rebalance above B/2 (gadget.rs:115-190)."""
    mag = abs(c_centered)
    v = (mag * base ** L + Q // 2) // Q
    val = v if c_centered >= 0 else -v
    out = []
    for _ in range(L):
        d = val % base
        if d > base // 2:
            d -= base
        out.append(d)
        val = (val - d) // base
    return out[::-1]


@pytest.mark.parametrize(
    "qs,base,L", [((137438822401, 274810798081), 18, 18), ((137438822401, 274810798081), 55879, 2),
                  ((17175674881,), 81, 2), ((8380417,), 8, 8), ((1 << 36,), 64, 7)])
def test_gadget_decompose_wide_path(oracle, qs, base, L):
    # The 256-bit path (wide.rs:145-153) against Python big integers on random
    # and boundary coefficients, plus the reconstruction bound (gadget.rs:388-421).
    Q = 1
    for q in qs:
        Q *= q
    n = 64
    rng = np.random.default_rng(base)
    vals = [0, 1, Q - 1, Q // 2, Q // 2 + 1, (Q + 1) // 2] + [int(x) for x in rng.integers(0, 2 ** 62, n - 6)]
    vals = [(v * 7919 + (i << 40)) % Q if i >= 6 else v for i, v in enumerate(vals)]
    limbs = sum([[v % q for v in vals] for q in qs], [])
    d = decompose(oracle, Ring.make(n, *qs), limbs, base, L)
    g = [(Q + base ** (i + 1) // 2) // base ** (i + 1) for i in range(L)]
    for i, v in enumerate(vals):
        c = v if v <= Q // 2 else v - Q
        assert d[:, i].tolist() == py_digits(c, Q, base, L), i
        rec = sum(int(d[k, i]) * g[k] for k in range(L))
        err = (rec - c) % Q
        err = min(err, Q - err)
        assert err <= g[-1] // 2 + L * base, (i, err)


def test_round_mul_div_against_python(oracle):
    from oracle.vo import C

    rng = np.random.default_rng(7)
    Q = 137438822401 * 274810798081
    for _ in range(200):
        a = int(rng.integers(0, 2 ** 62)) * int(rng.integers(1, 2 ** 12)) % Q
        s = 18 ** 18
        lo, hi = C.c_uint64(), C.c_uint64()
        oracle.round_mul_div(a & (2 ** 64 - 1), a >> 64, s & (2 ** 64 - 1), s >> 64, Q & (2 ** 64 - 1), Q >> 64,
                             C.byref(lo), C.byref(hi))
        assert lo.value | (hi.value << 64) == (a * s + Q // 2) // Q
