"""Layer 3-6 known-answer tests and end-to-end recovery for the CPU restatement.

Each test restates a via-rs test (cited) with its seeds and draw order and
compares to the reference's generated constants (tests/golden/reference_kats.json).
"""
import json
import os

import numpy as np
import pytest

from oracle.vo import TERNARY, Ring, i64p, paper_params, toy_params, u64p, zeros

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
L3, L4, L5, L6 = GOLD["layer3"], GOLD["layer4"], GOLD["layer5"], GOLD["layer6"]


def P(a):
    return a.ctypes.data_as(u64p)


def arr(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.uint64))


# ------------------------------------------------------------ helpers -------


class Ctx:
    def __init__(self, o, seed):
        self.o = o
        self.g = o.prg(seed)

    def keygen(self, r):
        sk = zeros(r.pw)
        self.o.keygen(r, TERNARY, self.g.h, P(sk))
        return sk

    def random(self, r):
        out = zeros(r.pw)
        self.o.poly_random(r, self.g.h, P(out))
        return out

    def encrypt(self, r, sk, m):
        ct = zeros(2 * r.pw)
        self.o.encrypt(r, P(sk), P(arr(m)), TERNARY, self.g.h, P(ct))
        return ct

    def encode(self, r, p, pt):
        out = zeros(r.pw)
        self.o.encode(r, p, P(arr(pt)), P(out))
        return out

    def rlev(self, r, sk, m, base, L):
        out = zeros(2 * L * r.pw)
        self.o.encrypt_rlev(r, P(sk), P(arr(m)), base, L, TERNARY, self.g.h, P(out))
        return out

    def rgsw(self, r, sk, m, base, L):
        out = zeros(4 * L * r.pw)
        self.o.encrypt_rgsw(r, P(sk), P(arr(m)), base, base, L, L, TERNARY, self.g.h, P(out))
        return out


def const_poly(n, v):
    a = np.zeros(n, dtype=np.uint64)
    a[0] = v
    return a


# ------------------------------------------------------------ layer 3 -------

N1, N2, BASE3, DEPTH3, D3 = 64, 16, 4, 8, 4
RQ2, RQ3, RQ3N2 = Ring.make(N1, 1 << 32), Ring.make(N1, 1 << 16), Ring.make(N2, 1 << 16)


def test_l3_mod_switch_sym(oracle):
    # layer3_kats.rs:53-63
    c = Ctx(oracle, b"layer3-kat-mod-switch-sym")
    ct = np.concatenate([c.random(RQ2), c.random(RQ2)])
    out = zeros(2 * N1)
    oracle.mod_switch_sym(RQ2, P(ct), RQ3, P(out))
    assert out[:N1].tolist() == L3["MOD_SWITCH_SYM_MASK_AT_Q3"]
    assert out[N1:].tolist() == L3["MOD_SWITCH_SYM_BODY_AT_Q3"]


def test_l3_mod_switch_asym(oracle):
    c = Ctx(oracle, b"layer3-kat-mod-switch-asym")
    ct = np.concatenate([c.random(RQ2), c.random(RQ2)])
    m, b = zeros(N1), zeros(N1)
    oracle.mod_switch_asym(RQ2, P(ct), 1 << 16, 1 << 12, P(m), P(b))
    assert m.tolist() == L3["MOD_SWITCH_ASYM_MASK_AT_Q3"]
    assert b.tolist() == L3["MOD_SWITCH_ASYM_BODY_AT_Q4"]


def gen_rsk(oracle, c, s1, s2):
    rsk = zeros(D3 * DEPTH3 * 2 * N2)
    oracle.gen_rsk(RQ3, P(s1), RQ3N2, P(s2), BASE3, DEPTH3, TERNARY, c.g.h, P(rsk))
    return rsk


def test_l3_gen_rsk(oracle):
    c = Ctx(oracle, b"layer3-kat-gen-rsk-rsk")
    s1, s2 = c.keygen(RQ3), c.keygen(RQ3N2)
    rsk = gen_rsk(oracle, c, s1, s2).reshape(D3 * DEPTH3, 2, N2)
    assert rsk[0, 0].tolist() == L3["GEN_RSK_J0_L0_MASK"]
    assert rsk[0, 1].tolist() == L3["GEN_RSK_J0_L0_BODY"]
    assert rsk[:, 0].reshape(-1).tolist() == L3["GEN_RSK_FULL_MASK_STREAM"]
    assert rsk[:, 1].reshape(-1).tolist() == L3["GEN_RSK_FULL_BODY_STREAM"]


def test_l3_ring_switch(oracle):
    c = Ctx(oracle, b"layer3-kat-ring-switch")
    s1, s2 = c.keygen(RQ3), c.keygen(RQ3N2)
    rsk = gen_rsk(oracle, c, s1, s2)
    enc = c.encode(RQ3, 16, [i % 16 for i in range(N1)])
    ct = c.encrypt(RQ3, s1, enc)
    out = zeros(2 * N2)
    oracle.ring_switch(RQ3, P(ct), RQ3N2, P(rsk), BASE3, DEPTH3, P(out))
    assert out[:N2].tolist() == L3["RING_SWITCH_OUT_MASK"]
    assert out[N2:].tolist() == L3["RING_SWITCH_OUT_BODY"]


def test_l3_rekey(oracle):
    c = Ctx(oracle, b"layer3-kat-rekey")
    sk = c.keygen(RQ2)
    out = zeros(N1)
    oracle.rekey(RQ2, P(sk), RQ3, P(out))
    assert out.tolist() == L3["REKEY_SK_Q3_COEFFS"]


def test_l3_decrypt_asym(oracle):
    c = Ctx(oracle, b"layer3-kat-decrypt-asym")
    s2 = c.keygen(RQ3N2)
    m = [(3 * i + 1) % 16 for i in range(N2)]
    ct = c.encrypt(RQ3N2, s2, c.encode(RQ3N2, 16, m))
    mm, mb, rec = zeros(N2), zeros(N2), zeros(N2)
    oracle.mod_switch_asym(RQ3N2, P(ct), 1 << 16, 1 << 12, P(mm), P(mb))
    oracle.decrypt_asym(RQ3N2, P(s2), P(mm), P(mb), 1 << 12, 16, P(rec))
    assert rec.tolist() == L3["DECRYPT_ASYM_RECOVERED"] == m


# ------------------------------------------------------------ layer 4 -------

N4, Q4R, BASE4, DEPTH4 = 64, 65537, 4, 3
R4 = Ring.make(N4, Q4R)


def enc4(c, sk, coeffs):
    pt = np.zeros(N4, dtype=np.uint64)
    pt[: len(coeffs)] = coeffs
    return c.encrypt(R4, sk, c.encode(R4, 2, pt))


def rgsw_bit(c, sk, b):
    return c.rgsw(R4, sk, const_poly(N4, b), BASE4, DEPTH4)


def check4(ct, km, kb):
    assert ct[:N4].tolist() == L4[km]
    assert ct[N4:].tolist() == L4[kb]


def test_l4_rotate(oracle):
    c = Ctx(oracle, b"layer4-kat-rotate")
    sk = c.keygen(R4)
    ct = enc4(c, sk, [1])
    out = zeros(2 * N4)
    oracle.rotate(R4, P(ct), 3, P(out))
    check4(out, "ROTATE_K3_MASK", "ROTATE_K3_BODY")


def test_l4_cmux(oracle):
    c = Ctx(oracle, b"layer4-kat-cmux")
    sk = c.keygen(R4)
    g1 = rgsw_bit(c, sk, 1)
    ct0, ct1 = enc4(c, sk, [1]), enc4(c, sk, [0, 1])
    out = zeros(2 * N4)
    oracle.cmux(R4, P(g1), DEPTH4, BASE4, P(ct0), P(ct1), P(out))
    check4(out, "CMUX_SELECT1_MASK", "CMUX_SELECT1_BODY")


def test_l4_dmux(oracle):
    c = Ctx(oracle, b"layer4-kat-dmux")
    sk = c.keygen(R4)
    g1 = rgsw_bit(c, sk, 1)
    ct = enc4(c, sk, [1, 1])
    r0, r1 = zeros(2 * N4), zeros(2 * N4)
    oracle.dmux(R4, P(g1), DEPTH4, BASE4, P(ct), P(r0), P(r1))
    check4(r0, "DMUX_BIT1_RESULT0_MASK", "DMUX_BIT1_RESULT0_BODY")
    check4(r1, "DMUX_BIT1_RESULT1_MASK", "DMUX_BIT1_RESULT1_BODY")


def test_l4_cmux_tree(oracle):
    c = Ctx(oracle, b"layer4-kat-cmux-tree")
    sk = c.keygen(R4)
    bits = np.concatenate([rgsw_bit(c, sk, 1) for _ in range(2)])
    inputs = np.concatenate([enc4(c, sk, [0] * i + [1]) for i in range(4)])
    out = zeros(2 * N4)
    oracle.cmux_tree(R4, P(bits), 2, DEPTH4, BASE4, P(inputs), P(out))
    check4(out, "CMUX_TREE_IDX3_MASK", "CMUX_TREE_IDX3_BODY")


def test_l4_dmux_tree(oracle):
    c = Ctx(oracle, b"layer4-kat-dmux-tree")
    sk = c.keygen(R4)
    b0, b1 = rgsw_bit(c, sk, 1), rgsw_bit(c, sk, 0)
    ct = enc4(c, sk, [1, 1])
    out = zeros(4 * 2 * N4)
    oracle.dmux_tree(R4, P(np.concatenate([b0, b1])), 2, DEPTH4, BASE4, P(ct), P(out))
    for k in range(4):
        check4(out[k * 2 * N4:(k + 1) * 2 * N4], f"DMUX_TREE_IDX2_OUT{k}_MASK", f"DMUX_TREE_IDX2_OUT{k}_BODY")


@pytest.mark.parametrize("slot,seed,key", [(0, b"layer4-kat-crot-forward", "CROT_FORWARD_ROT3"),
                                           (1, b"layer4-kat-crot-slot-extract", "CROT_SLOT_EXTRACT_ROT3")])
def test_l4_crot(oracle, slot, seed, key):
    c = Ctx(oracle, seed)
    sk = c.keygen(R4)
    bits = np.concatenate([rgsw_bit(c, sk, 1) for _ in range(2)])
    ct = enc4(c, sk, [1])
    out = zeros(2 * N4)
    oracle.crot(R4, slot, P(bits), 2, DEPTH4, BASE4, P(ct), P(out))
    check4(out, key + "_MASK", key + "_BODY")


def test_l4_rlwe_to_rgsw(oracle):
    c = Ctx(oracle, b"layer4-kat-rlwe-to-rgsw")
    sk = c.keygen(R4)
    m = const_poly(N4, 1)
    g = [(Q4R + BASE4 ** (i + 1) // 2) // BASE4 ** (i + 1) for i in range(DEPTH4)]
    levels = np.concatenate([c.encrypt(R4, sk, const_poly(N4, g[i])) for i in range(DEPTH4)])
    s2 = zeros(N4)
    oracle.poly_mul(R4, P(s2), P(sk), P(sk))
    conv_key = c.rlev(R4, sk, s2, 2, 16)
    m_rlev = c.rlev(R4, sk, m, BASE4, DEPTH4)
    ct_prime = enc4(c, sk, [1, 1])
    rgsw = zeros(4 * DEPTH4 * N4)
    oracle.rlwe_to_rgsw(R4, P(levels), P(conv_key), P(m_rlev), 2, DEPTH4, 16, P(rgsw))
    check4(rgsw[: 2 * N4], "RLWE_TO_RGSW_NEG_S_M_L0_MASK", "RLWE_TO_RGSW_NEG_S_M_L0_BODY")
    ep = zeros(2 * N4)
    oracle.external_product(R4, P(rgsw), DEPTH4, DEPTH4, BASE4, BASE4, P(ct_prime), P(ep))
    check4(ep, "RLWE_TO_RGSW_EXT_PRODUCT_MASK", "RLWE_TO_RGSW_EXT_PRODUCT_BODY")


def test_l4_mod_switch_rgsw(oracle):
    c = Ctx(oracle, b"layer4-kat-mod-switch-rgsw")
    sk = c.keygen(R4)
    g = rgsw_bit(c, sk, 1)
    out = zeros(4 * DEPTH4 * N4)
    oracle.mod_switch_rgsw(R4, P(g), DEPTH4, DEPTH4, Ring.make(N4, 4096), P(out))
    check4(out[: 2 * N4], "MOD_SWITCH_RGSW_NEG_S_M_L0_MASK", "MOD_SWITCH_RGSW_NEG_S_M_L0_BODY")
    m0 = out[2 * DEPTH4 * N4:(2 * DEPTH4 + 2) * N4]
    check4(m0, "MOD_SWITCH_RGSW_M_L0_MASK", "MOD_SWITCH_RGSW_M_L0_BODY")


# ------------------------------------------------------------ layer 5 -------

Q5, BASE5, DEPTH5 = 65537, 8, 6
R5_8, R5_4 = Ring.make(8, Q5), Ring.make(4, Q5)


def test_l5_embed_mlwe(oracle):
    k = Ctx(oracle, b"layer5-kat-embed-key")
    sk = k.keygen(R5_4)
    e = Ctx(oracle, b"layer5-kat-embed-enc")
    ct = e.encrypt(R5_4, sk, e.encode(R5_4, 16, [1, 2, 3, 4]))
    # embed_mlwe (mlwe_ops.rs:71-80): coefficient i -> 2i
    mask = np.zeros(8, dtype=np.uint64)
    body = np.zeros(8, dtype=np.uint64)
    mask[0::2], body[0::2] = ct[:4], ct[4:]
    assert mask.tolist() == L5["EMBED_MLWE_MASK0"]
    assert body.tolist() == L5["EMBED_MLWE_BODY"]


def lwe8(oracle, key_seed, enc_seed):
    k = Ctx(oracle, key_seed)
    sk = k.keygen(R5_8)
    e = Ctx(oracle, enc_seed)
    lwe = zeros(9)
    oracle.encrypt_lwe(R5_8, P(sk), 5, 16, TERNARY, e.g.h, P(lwe))
    return sk, lwe


def cascade_key8(oracle, sk, seed):
    c = Ctx(oracle, seed)
    key = zeros(oracle.cascade_key_words(R5_8, DEPTH5))
    oracle.gen_cascade_key(R5_8, P(sk), BASE5, DEPTH5, TERNARY, c.g.h, P(key))
    return key


def test_l5_encrypt_lwe(oracle):
    _, lwe = lwe8(oracle, b"layer5-kat-encrypt-lwe-key", b"layer5-kat-encrypt-lwe-enc")
    assert lwe[:8].tolist() == L5["ENCRYPT_LWE_MASKS"]
    assert [int(lwe[8])] == L5["ENCRYPT_LWE_BODY"]


def test_l5_conv_step(oracle):
    sk, lwe = lwe8(oracle, b"layer5-kat-conv-key", b"layer5-kat-conv-enc")
    key = cascade_key8(oracle, sk, b"layer5-kat-conv-keygen")
    om, ob = zeros(4 * 2), zeros(2)
    oracle.conv_step(R5_8, 2, P(arr(lwe[:8])), P(arr(lwe[8:])), P(key), BASE5, DEPTH5, P(om), P(ob))
    assert om[:2].tolist() == L5["CONV_D_OUT_MASK0"]
    assert ob.tolist() == L5["CONV_D_OUT_BODY"]


def test_l5_gen_cascade_key(oracle):
    c = Ctx(oracle, b"layer5-kat-keygen-key")
    sk = c.keygen(R5_8)
    key = cascade_key8(oracle, sk, b"layer5-kat-keygen-gen")
    assert key[:2].tolist() == L5["LWE2RLWE_KEY_STEP0_S0_L0_MASK"]


def test_l5_cascade(oracle):
    sk, lwe = lwe8(oracle, b"layer5-kat-cascade-key", b"layer5-kat-cascade-enc")
    key = cascade_key8(oracle, sk, b"layer5-kat-cascade-keygen")
    ct = zeros(16)
    oracle.cascade(R5_8, P(lwe), P(key), BASE5, DEPTH5, P(ct))
    assert ct[:8].tolist() == L5["LWE2RLWE_OUT_MASK"]
    assert ct[8:].tolist() == L5["LWE2RLWE_OUT_BODY"]


# ------------------------------------------------------------ layer 6 -------

R6 = Ring.make(8, 1 << 36)


def test_l6_encrypt_lwe_raw(oracle):
    k = Ctx(oracle, b"layer6-kat-elr-key")
    sk = k.keygen(R6)
    e = Ctx(oracle, b"layer6-kat-elr-enc")
    lwe = zeros(9)
    oracle.encrypt_lwe_raw(R6, P(sk), 0xDEADBEEF % (1 << 36), 0, TERNARY, e.g.h, P(lwe))
    assert lwe[:8].tolist() == L6["ENCRYPT_LWE_RAW_MASKS"]
    assert [int(lwe[8])] == L6["ENCRYPT_LWE_RAW_BODY"]


def test_l6_gen_rlwe_to_rgsw_key(oracle):
    k = Ctx(oracle, b"layer6-kat-ck-key")
    sk = k.keygen(R6)
    g = Ctx(oracle, b"layer6-kat-ck-gen")
    s2 = zeros(8)
    oracle.poly_mul(R6, P(s2), P(sk), P(sk))
    key = g.rlev(R6, sk, s2, 64, 7)
    assert key[:8].tolist() == L6["RLWE_TO_RGSW_KEY_S0_MASK"]
    assert key[8:16].tolist() == L6["RLWE_TO_RGSW_KEY_S0_BODY"]


def client_setup(oracle, pp, prg):
    n1, n2 = pp.n1, pp.n2
    sk1, sk2 = zeros(pp.q1_nl * n1), zeros(n2)
    ck = zeros(oracle.words_cascade_key(pp))
    cv = zeros(oracle.words_conv_key(pp))
    rsk = zeros(oracle.words_rsk(pp))
    rc = oracle.client_setup(pp, TERNARY, TERNARY, TERNARY, prg.h, P(sk1), P(sk2), P(ck), P(cv), P(rsk))
    assert rc == 0
    return sk1, sk2, ck, cv, rsk


def client_query(oracle, pp, sk1, index, prg):
    lwes = zeros(oracle.query_lwes(pp) * oracle.words_lwe(pp))
    assert oracle.client_query(pp, P(sk1), index, TERNARY, prg.h, P(lwes)) == 0
    return lwes


def test_l6_query_comp(oracle):
    pp = toy_params()
    sk1, *_ = client_setup(oracle, pp, oracle.prg(b"layer6-kat-qc-setup"))
    lwes = client_query(oracle, pp, sk1, 5, oracle.prg(b"layer6-kat-qc-query")).reshape(21, 9)
    assert lwes[:, :8].reshape(-1).tolist() == L6["QUERY_COMP_MASKS"]
    assert lwes[:, 8].tolist() == L6["QUERY_COMP_BODIES"]


def test_l6_resp_comp(oracle):
    rq2, rq3n1, rq3n2 = Ring.make(8, 1 << 28), Ring.make(8, 1 << 20), Ring.make(4, 1 << 20)
    sk1 = Ctx(oracle, b"layer6-kat-rc-s1").keygen(rq2)
    sk2 = Ctx(oracle, b"layer6-kat-rc-s2").keygen(rq3n2)
    e = Ctx(oracle, b"layer6-kat-rc-enc")
    ct = e.encrypt(rq2, sk1, e.encode(rq2, 16, [3, 0, 0, 0, 0, 0, 0, 0]))
    s1q3 = zeros(8)
    oracle.rekey(rq2, P(sk1), rq3n1, P(s1q3))
    r = Ctx(oracle, b"layer6-kat-rc-rsk")
    rsk = zeros(2 * 8 * 2 * 4)
    oracle.gen_rsk(rq3n1, P(s1q3), rq3n2, P(sk2), 8, 8, TERNARY, r.g.h, P(rsk))
    # resp_comp (resp_comp.rs:46-76)
    c3, rs, am, ab = zeros(16), zeros(8), zeros(4), zeros(4)
    oracle.mod_switch_sym(rq2, P(ct), rq3n1, P(c3))
    oracle.ring_switch(rq3n1, P(c3), rq3n2, P(rsk), 8, 8, P(rs))
    oracle.mod_switch_asym(rq3n2, P(rs), 1 << 20, 1 << 12, P(am), P(ab))
    assert am.tolist() == L6["RESP_COMP_MASK_Q3"]
    assert ab.tolist() == L6["RESP_COMP_BODY_Q4"]


# ------------------------------------------------------------ end to end ----


def server(oracle, pp, records, ck, cv, rsk):
    import ctypes

    h = ctypes.c_void_p()
    rc = oracle.server_new(pp, P(records), len(records) // pp.n_rec, P(ck), P(cv), P(rsk), ctypes.byref(h))
    assert rc == 0
    return h


def answer(oracle, pp, h, lwes):
    am, ab = zeros(pp.n2), zeros(pp.n2)
    rc = oracle.server_answer(h, P(lwes), oracle.query_lwes(pp), P(am), P(ab), None, None)
    assert rc == 0
    return am, ab


def recover(oracle, pp, sk2, am, ab):
    rec = zeros(pp.n2)
    oracle.client_recover(pp, P(sk2), P(am), P(ab), P(rec))
    return rec


def toy_record(m):
    return [(m + 1) % 16, (2 * m + 1) % 16, (3 * m + 1) % 16, (m + 5) % 16]


@pytest.mark.parametrize("index", range(8))
def test_e2e_toy_n8(oracle, index):
    # client_server_e2e.rs:92-172
    pp = toy_params()
    prg = oracle.prg(b"via-c-client-server-e2e")
    sk1, sk2, ck, cv, rsk = client_setup(oracle, pp, prg)
    records = arr(sum([toy_record(m) for m in range(8)], []))
    h = server(oracle, pp, records, ck, cv, rsk)
    lwes = client_query(oracle, pp, sk1, index, prg)
    am, ab = answer(oracle, pp, h, lwes)
    assert recover(oracle, pp, sk2, am, ab).tolist() == toy_record(index)
    oracle.server_free(h)


def paper_record(m):
    return [(m + 1 + j) % 16 if j < 4 else 0 for j in range(512)]


@pytest.mark.slow
def test_e2e_paper_2x2_index15(oracle):
    # client_server_e2e_paper.rs:287-349
    pp = paper_params(2, 2)
    prg = oracle.prg(b"via-c-paper-scale-e2e")
    sk1, sk2, ck, cv, rsk = client_setup(oracle, pp, prg)
    records = arr(sum([paper_record(m) for m in range(16)], []))
    h = server(oracle, pp, records, ck, cv, rsk)
    lwes = client_query(oracle, pp, sk1, 15, prg)
    am, ab = answer(oracle, pp, h, lwes)
    assert recover(oracle, pp, sk2, am, ab).tolist() == paper_record(15)
    oracle.server_free(h)


def test_answer_errors(oracle):
    # answer.rs:131-157: wrong query length -> QueryLengthMismatch.
    import ctypes

    pp = toy_params()
    prg = oracle.prg(b"err")
    sk1, sk2, ck, cv, rsk = client_setup(oracle, pp, prg)
    h = server(oracle, pp, arr([0] * 4), ck, cv, rsk)
    lwes = client_query(oracle, pp, sk1, 0, prg)
    e, g = ctypes.c_size_t(), ctypes.c_size_t()
    am, ab = zeros(4), zeros(4)
    assert oracle.server_answer(h, P(lwes), 20, P(am), P(ab), ctypes.byref(e), ctypes.byref(g)) == 2
    assert (e.value, g.value) == (21, 20)
    # Too many records -> DimMismatch at setup (answer.rs:503-508).
    h2 = ctypes.c_void_p()
    assert oracle.server_new(pp, P(arr([0] * 40)), 10, P(ck), P(cv), P(rsk), ctypes.byref(h2)) == 1
    oracle.server_free(h)
