"""Seeded paper-stack VIA-C cases built with the CPU restatement (the checker).

A case holds the client keys, the public keys, the records, one query and the
oracle's answer and stage outputs, ready to feed the HIP path.
"""
import ctypes
import hashlib

import numpy as np

from oracle.vo import TERNARY, Oracle, paper_params, u64p


def P(a):
    return a.ctypes.data_as(u64p)


class Case:
    pass


_ORACLE = None


def oracle():
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = Oracle()
        # the checker's query decomposition over every host CPU (same values)
        _ORACLE.set_decomp_threads(_ORACLE.host_threads())
    return _ORACLE


def paper_record(m):
    return [(m + 1 + j) % 16 if j < 4 else 0 for j in range(512)]


def paper_case(num_rows, num_cols, seed, index, records="paper", n_records=None, stages=False, rec_seed=b"recs"):
    o = oracle()
    pp = paper_params(num_rows, num_cols)
    prg = o.prg(seed)
    c = Case()
    c.pp = pp
    n1, n2 = pp.n1, pp.n2
    c.sk1, c.sk2 = np.zeros(2 * n1, np.uint64), np.zeros(n2, np.uint64)
    c.cascade_key = np.zeros(o.words_cascade_key(pp), np.uint64)
    c.conv_key = np.zeros(o.words_conv_key(pp), np.uint64)
    c.rsk = np.zeros(o.words_rsk(pp), np.uint64)
    assert o.client_setup(pp, TERNARY, TERNARY, TERNARY, prg.h, P(c.sk1), P(c.sk2), P(c.cascade_key),
                          P(c.conv_key), P(c.rsk)) == 0
    cap = 4 * num_rows * num_cols
    nrec = cap if n_records is None else n_records
    if records == "paper":
        recs = np.array([paper_record(m) for m in range(nrec)], dtype=np.uint64)
    else:
        rng = np.random.default_rng(int.from_bytes(hashlib.sha256(rec_seed).digest()[:8], 'little'))
        recs = rng.integers(0, 16, size=(nrec, 512), dtype=np.uint64)
    c.records = np.ascontiguousarray(recs.reshape(-1))
    c.index = index
    c.record = recs[index] if index < nrec else np.zeros(512, np.uint64)
    c.lwes = np.zeros(o.query_lwes(pp) * o.words_lwe(pp), np.uint64)
    assert o.client_query(pp, P(c.sk1), index, TERNARY, prg.h, P(c.lwes)) == 0
    h = ctypes.c_void_p()
    assert o.server_new(pp, P(c.records), nrec, P(c.cascade_key), P(c.conv_key), P(c.rsk), ctypes.byref(h)) == 0
    c.ans_mask, c.ans_body = np.zeros(n2, np.uint64), np.zeros(n2, np.uint64)
    I, J = num_rows, num_cols
    bits = o.query_lwes(pp) // pp.l_query
    if stages:
        c.st = {
            1: np.zeros(bits * 4 * pp.l_query * 2 * n1, np.uint64),
            2: np.zeros(I * 2 * 2 * n1, np.uint64),
            3: np.zeros(I * 2 * n1, np.uint64),
            4: np.zeros(J * 2 * n1, np.uint64),
            5: np.zeros(2 * n1, np.uint64),
            6: np.zeros(2 * n1, np.uint64),
        }
        o.server_answer_stages(h, P(c.lwes), o.query_lwes(pp), *[P(c.st[k]) for k in range(1, 7)],
                               P(c.ans_mask), P(c.ans_body))
    else:
        assert o.server_answer(h, P(c.lwes), o.query_lwes(pp), P(c.ans_mask), P(c.ans_body), None, None) == 0
    c.oracle_db = np.ctypeslib.as_array(o.server_db(h), shape=(I * J * n1,)).copy()
    o.server_free(h)

    def recover(mask, body):
        rec = np.zeros(n2, np.uint64)
        o.client_recover(pp, P(c.sk2), P(np.ascontiguousarray(mask, dtype=np.uint64)),
                         P(np.ascontiguousarray(body, dtype=np.uint64)), P(rec))
        return rec

    c.recover = recover
    import via_gpu

    g = via_gpu.paper_params(num_rows, num_cols)
    c.pp_gpu = g
    return c


def arr(x):
    return np.ascontiguousarray(np.array(x, dtype=np.uint64))


class BatchClient:
    """Client keys and batch queries from the CPU restatement (the reference's
    Client::setup / batch_query order)."""

    def __init__(self, oracle, pp, seed):
        from oracle.vo import TERNARY

        self.o, self.pp, self.T_ = oracle, pp, TERNARY
        self.prg = oracle.prg(seed)
        n1, n2 = pp.n1, pp.n2
        self.sk1, self.sk2 = arr([0] * (2 * n1)), arr([0] * n2)
        self.ck = arr([0] * oracle.words_cascade_key(pp))
        self.cv = arr([0] * oracle.words_conv_key(pp))
        self.rsk = arr([0] * oracle.words_rsk(pp))
        assert oracle.client_setup(pp, TERNARY, TERNARY, TERNARY, self.prg.h, P(self.sk1), P(self.sk2),
                                   P(self.ck), P(self.cv), P(self.rsk)) == 0

    def queries(self, idxs):
        o, pp = self.o, self.pp
        out = []
        for idx in idxs:
            q = arr([0] * (o.query_lwes(pp) * o.words_lwe(pp)))
            assert o.client_query(pp, P(self.sk1), idx, self.T_, self.prg.h, P(q)) == 0
            out.append(q)
        return out

    def recover(self, T, am, ab):
        o, pp = self.o, self.pp
        pt = arr([0] * pp.n2)
        o.client_recover(pp, P(self.sk2), P(arr(am)), P(arr(ab)), P(pt))
        return [pt[t::pp.n2 // pp.n_rec][:pp.n_rec].tolist() for t in range(T)]
