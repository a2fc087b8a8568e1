"""Self-golden answers of the oracle for the GPU parity tests.

The reference holds no Answer bytes beyond its toy layer KATs: its paper-scale
tests check only that the client recovers the record
(client_server_e2e_paper.rs:316-349, batch_e2e_paper.rs:130-174). These
fixtures are therefore self-golden: the answers of oracle/ (the C
restatement, itself pinned by the reference's layer 3-6 KATs, PRG vectors and
e2e recoveries), recorded so that the GPU tests compare bytes at sizes where
running the oracle inside a test would take minutes.

Each case rebuilds, on the CPU, exactly the inputs its GPU test builds with
the library: records from fill_bytes of Shake256Prg(<tag>/db) (coefficient 2k
= byte k & 0xF, 2k+1 = byte k >> 4), keys from Client::setup over
Shake256Prg(<tag>/setup) with ternary keys and errors, queries from
Client::query continuing that stream, indices 0, n_records - 1, then
uniform_below(n_records) draws of Shake256Prg(<tag>/query). Small cases keep
the whole answer (mask mod q3 then body mod q4); full-size cases keep
SHA-256(mask u64 LE || body u64 LE) per answer.

    python tests/golden/make_answer_digests.py            # every case
    python tests/golden/make_answer_digests.py --only NAME [--only NAME ...]
    python tests/golden/make_answer_digests.py --list

Writes tests/golden/self_golden_answers.json, merging into what is there.
"""
import argparse
import ctypes as C
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.vo import TERNARY, Oracle, make_params, paper_b_params, paper_params, toy_b_params, toy_params, u64p  # noqa: E402

OUT = os.path.join(HERE, "self_golden_answers.json")


def P(a):
    return a.ctypes.data_as(u64p)


def z(n):
    return np.zeros(n, np.uint64)


def answer_digest(mask, body):
    return hashlib.sha256(np.ascontiguousarray(mask, "<u8").tobytes() +
                          np.ascontiguousarray(body, "<u8").tobytes()).hexdigest()


def params_dict(pp):
    d = {}
    for name, _ in pp._fields_:
        v = getattr(pp, name)
        d[name] = [v[0], v[1]] if name == "q1" else int(v)
    return d


def toy_n64_params(num_rows, num_cols):
    """The n1 = 64 single-prime stack of batch_e2e_toy.rs:40-68 at the VIA-C
    record degree n_rec = n2 = 16 (records of 16 nibbles, 8 bytes)."""
    return make_params(64, 16, 1 << 36, 1 << 28, 1 << 20, 1 << 12, 16, 64, 64, 8, 8, 7, 12, 8,
                       num_rows, num_cols, n_rec=16)


class Seeded:
    """The inputs the GPU full-size tests build with the library."""

    def __init__(self, o, pp, tag, n_records):
        self.o, self.pp = o, pp
        self.n_records = n_records
        nb = n_records * pp.n_rec // 2
        g_db = o.prg(tag + b"/db")
        rb = np.frombuffer(g_db.fill(nb), np.uint8).astype(np.uint64)
        recs = np.empty(n_records * pp.n_rec, np.uint64)
        recs[0::2] = rb & 0xF
        recs[1::2] = rb >> 4
        self.records = recs
        self.g = o.prg(tag + b"/setup")
        n1, n2 = pp.n1, pp.n2
        self.sk1, self.sk2 = z(pp.q1_nl * n1), z(n2)
        self.ck, self.cv = z(o.words_cascade_key(pp)), z(o.words_conv_key(pp))
        self.rsk = z(o.words_rsk(pp))
        assert o.client_setup(pp, TERNARY, TERNARY, TERNARY, self.g.h, P(self.sk1), P(self.sk2), P(self.ck),
                              P(self.cv), P(self.rsk)) == 0
        self.g_q = o.prg(tag + b"/query")

    def indices(self, count):
        return [0, self.n_records - 1] + [int(self.g_q.uniform_below(self.n_records)) for _ in range(count - 2)]

    def query(self, idx):
        o, pp = self.o, self.pp
        q = z(o.query_lwes(pp) * o.words_lwe(pp))
        assert o.client_query(pp, P(self.sk1), idx, TERNARY, self.g.h, P(q)) == 0
        return q

    def server(self):
        h = C.c_void_p()
        o = self.o
        assert o.server_new(self.pp, P(self.records), self.n_records, P(self.ck), P(self.cv), P(self.rsk),
                            C.byref(h)) == 0
        return h

    def recover(self, mask, body):
        pt = z(self.pp.n2)
        self.o.client_recover(self.pp, P(self.sk2), P(mask), P(body), P(pt))
        return pt

    def record(self, idx):
        n = self.pp.n_rec
        return self.records[idx * n:(idx + 1) * n]


def via_c_full(o, rows, cols, count=4):
    pp = paper_params(rows, cols)
    tag = b"via-gpu/full/%dx%d" % (rows, cols)
    s = Seeded(o, pp, tag, 4 * rows * cols)
    idxs = s.indices(count)
    h = s.server()
    digests = []
    try:
        for idx in idxs:
            m, b = z(pp.n2), z(pp.n2)
            assert o.server_answer(h, P(s.query(idx)), o.query_lwes(pp), P(m), P(b), None, None) == 0
            assert np.array_equal(s.recover(m, b), s.record(idx)), idx
            digests.append(answer_digest(m, b))
    finally:
        o.server_free(h)
    return {"kind": "via-c digests", "tag": tag.decode(), "params": params_dict(pp), "indices": idxs,
            "sha256": digests}


def via_b_full(o, T, rows=32, cols=4096):
    pp = paper_b_params(rows, cols, T=T)
    tag = b"via-gpu/full/vb%d" % T
    s = Seeded(o, pp, tag, (pp.n1 // pp.n_rec) * rows * cols)
    idxs = s.indices(T)
    qs = np.concatenate([s.query(i) for i in idxs])
    h = s.server()
    try:
        m, b = z(pp.n2), z(pp.n2)
        assert o.server_answer_batch(h, T, P(qs), o.query_lwes(pp), P(m), P(b), None, None) == 0
    finally:
        o.server_free(h)
    pt = s.recover(m, b)
    d = pp.n2 // pp.n_rec
    for t, idx in enumerate(idxs):
        assert np.array_equal(pt[t::d][:pp.n_rec], s.record(idx)), (t, idx)
    return {"kind": "via-b digest", "tag": tag.decode(), "params": params_dict(pp), "T": T, "indices": idxs,
            "sha256": [answer_digest(m, b)]}


def toy_n8_all(o):
    # client_server_e2e.rs:92-172: a fresh setup per index, records 4 nibbles
    pp = toy_params()

    def rec(m):
        return [(m + 1) % 16, (2 * m + 1) % 16, (3 * m + 1) % 16, (m + 5) % 16]

    records = np.array(sum([rec(m) for m in range(8)], []), np.uint64)
    answers = []
    for index in range(8):
        g = o.prg(b"via-c-client-server-e2e")
        sk1, sk2 = z(pp.q1_nl * pp.n1), z(pp.n2)
        ck, cv, rsk = z(o.words_cascade_key(pp)), z(o.words_conv_key(pp)), z(o.words_rsk(pp))
        assert o.client_setup(pp, TERNARY, TERNARY, TERNARY, g.h, P(sk1), P(sk2), P(ck), P(cv), P(rsk)) == 0
        h = C.c_void_p()
        assert o.server_new(pp, P(records), 8, P(ck), P(cv), P(rsk), C.byref(h)) == 0
        q = z(o.query_lwes(pp) * o.words_lwe(pp))
        assert o.client_query(pp, P(sk1), index, TERNARY, g.h, P(q)) == 0
        m, b = z(pp.n2), z(pp.n2)
        assert o.server_answer(h, P(q), o.query_lwes(pp), P(m), P(b), None, None) == 0
        o.server_free(h)
        pt = z(pp.n2)
        o.client_recover(pp, P(sk2), P(m), P(b), P(pt))
        assert pt.tolist() == rec(index)
        answers.append({"index": index, "mask": m.tolist(), "body": b.tolist()})
    return {"kind": "via-c answers", "seed": "via-c-client-server-e2e", "params": params_dict(pp),
            "answers": answers}


def paper_2x2_15(o):
    # client_server_e2e_paper.rs:287-349
    pp = paper_params(2, 2)

    def rec(m):
        return [(m + 1 + j) % 16 if j < 4 else 0 for j in range(512)]

    g = o.prg(b"via-c-paper-scale-e2e")
    sk1, sk2 = z(pp.q1_nl * pp.n1), z(pp.n2)
    ck, cv, rsk = z(o.words_cascade_key(pp)), z(o.words_conv_key(pp)), z(o.words_rsk(pp))
    assert o.client_setup(pp, TERNARY, TERNARY, TERNARY, g.h, P(sk1), P(sk2), P(ck), P(cv), P(rsk)) == 0
    records = np.array(sum([rec(m) for m in range(16)], []), np.uint64)
    h = C.c_void_p()
    assert o.server_new(pp, P(records), 16, P(ck), P(cv), P(rsk), C.byref(h)) == 0
    q = z(o.query_lwes(pp) * o.words_lwe(pp))
    assert o.client_query(pp, P(sk1), 15, TERNARY, g.h, P(q)) == 0
    m, b = z(pp.n2), z(pp.n2)
    assert o.server_answer(h, P(q), o.query_lwes(pp), P(m), P(b), None, None) == 0
    o.server_free(h)
    pt = z(pp.n2)
    o.client_recover(pp, P(sk2), P(m), P(b), P(pt))
    assert pt.tolist() == rec(15)
    return {"kind": "via-c answers", "seed": "via-c-paper-scale-e2e", "params": params_dict(pp),
            "answers": [{"index": 15, "mask": m.tolist(), "body": b.tolist(), "sha256": answer_digest(m, b)}]}


def toy_b_n64(o):
    # batch_e2e_toy.rs:190-225, the first index set of the shared set-up
    pp = toy_b_params(2, 2)

    def rec(m):
        return [(m + 1) % 16, (2 * m + 3) % 16]

    g = o.prg(b"via-b-batch-e2e-toy")
    sk1, sk2 = z(pp.q1_nl * pp.n1), z(pp.n2)
    ck, cv, rsk = z(o.words_cascade_key(pp)), z(o.words_conv_key(pp)), z(o.words_rsk(pp))
    assert o.client_setup(pp, TERNARY, TERNARY, TERNARY, g.h, P(sk1), P(sk2), P(ck), P(cv), P(rsk)) == 0
    records = np.array(sum([rec(m) for m in range(32 * 4)], []), np.uint64)
    h = C.c_void_p()
    assert o.server_new(pp, P(records), 128, P(ck), P(cv), P(rsk), C.byref(h)) == 0
    idxs = [3, 11, 20, 38, 72, 89, 106, 127]
    nl = o.query_lwes(pp) * o.words_lwe(pp)
    qs = z(8 * nl)
    for t, idx in enumerate(idxs):
        part = z(nl)
        assert o.client_query(pp, P(sk1), idx, TERNARY, g.h, P(part)) == 0
        qs[t * nl:(t + 1) * nl] = part
    m, b = z(pp.n2), z(pp.n2)
    assert o.server_answer_batch(h, 8, P(qs), o.query_lwes(pp), P(m), P(b), None, None) == 0
    o.server_free(h)
    pt = z(pp.n2)
    o.client_recover(pp, P(sk2), P(m), P(b), P(pt))
    assert [pt[t::pp.n2 // pp.n_rec][:pp.n_rec].tolist() for t in range(8)] == [rec(i) for i in idxs]
    return {"kind": "via-b answer", "seed": "via-b-batch-e2e-toy", "params": params_dict(pp), "T": 8,
            "indices": idxs, "mask": m.tolist(), "body": b.tolist()}


def toy_n64_1k(o):
    # 2^10 records of 8 uniform random bytes on the n1 = 64 stack (I = J = 16,
    # four records per cell, 10 query bits x 7 LWEs): the answer of one query
    # at a seeded index and the recovered record
    pp = toy_n64_params(16, 16)
    s = Seeded(o, pp, b"via-gpu/toy-n64", 1024)
    idx = int(s.g_q.uniform_below(1024))
    h = s.server()
    m, b = z(pp.n2), z(pp.n2)
    q = s.query(idx)
    assert o.server_answer(h, P(q), o.query_lwes(pp), P(m), P(b), None, None) == 0
    o.server_free(h)
    pt = s.recover(m, b)
    return {"kind": "via-c answers", "tag": "via-gpu/toy-n64", "params": params_dict(pp), "query_lwes":
            int(o.query_lwes(pp)), "answers": [{"index": idx, "mask": m.tolist(), "body": b.tolist(),
                                                "record": s.record(idx).tolist(), "recovered": pt.tolist()}]}


CASES = {
    "toy_n8_all": toy_n8_all,
    "paper_2x2_index15": paper_2x2_15,
    "toy_b_n64_t8": toy_b_n64,
    "toy_n64_1k_records": toy_n64_1k,
    "paper_128x128": lambda o: via_c_full(o, 128, 128),
    "paper_256x256": lambda o: via_c_full(o, 256, 256),
    "paper_512x512": lambda o: via_c_full(o, 512, 512, count=2),
    "vb_t8_32x4096": lambda o: via_b_full(o, 8),
    "vb_t16_32x4096": lambda o: via_b_full(o, 16),
    "vb_t32_32x4096": lambda o: via_b_full(o, 32),
    "vb_t64_32x4096": lambda o: via_b_full(o, 64),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", action="append")
    ap.add_argument("--list", action="store_true")
    a = ap.parse_args()
    if a.list:
        print("\n".join(CASES))
        return
    o = Oracle()
    o.set_decomp_threads(o.host_threads())
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    data.setdefault("cases", {})
    data["about"] = ("Self-golden answers of the oracle (tests/golden/make_answer_digests.py); the reference "
                     "holds no answer bytes at these sizes. Answer words: mask mod q3 then body mod q4; "
                     "sha256 over mask u64 LE || body u64 LE.")
    for name in a.only or list(CASES):
        t0 = time.time()
        print(f"case {name} ...", flush=True)
        data["cases"][name] = CASES[name](o)
        print(f"case {name} done in {time.time() - t0:.1f} s", flush=True)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
