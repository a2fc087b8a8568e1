"""VIA-C server answer benchmark on one MI355X.

A step is one Server::answer of the paper VIA-C stack (n1=2048, two-prime
q1, q2~2^34, q3~2^23) over a database of 256-byte records drawn from a
Shake256Prg stream, with keys from Client::setup and the query from
Client::query (the library's own client path), query and answer buffers
resident in HBM. Prints one JSON line: answers/s, p50/p99 answer latency,
per-step times and rates, the FirstDim roofline (HBM bytes per launch from
live HIP events), whether the answer recovers the record, and the CPU
restatement timed on one answer of the same workload, whose answer must
equal the GPU's bit for bit.

    python bench.py [--steps K] [--warmup W] [--rows I] [--cols J] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM traffic of one FirstDim launch (k_fd_mfma) per (I, J, queries per
# pass, column-block groups per slot), measured with rocprofv3 --pmc:
# FETCH_SIZE and WRITE_SIZE per dispatch, scaled by what the same counters
# read for a 256 MiB device copy (fetched bytes = 2.000 x FETCH_SIZE KiB,
# written bytes = 1.000 x WRITE_SIZE KiB on this part).
FD_TRAFFIC = {(256, 256, 1, 1): (2.0 * 330305.4 + 8192.0) * 1024,
              (256, 256, 1, 2): (2.0 * 332864.0 + 8192.0) * 1024,
              (256, 256, 1, 4): (2.0 * 337982.2 + 8192.0) * 1024,
              (256, 256, 1, 8): (2.0 * 348222.0 + 8192.0) * 1024}

# Modular multiplications (NTT butterflies, pointwise and scaling products
# counted alike) of one answer's steps at the paper stack, counted from the
# reference's loops (cascade + RLWE -> RGSW, the I - 1 DMux external
# products at q1, the CMux and CRot trees at q2, RespComp), per (I, J).
STEP_MODMUL = {
    (128, 128): {"query_decomp": 296e6, "dmux": 49.9e6, "cmux_crot": 25.4e6, "resp_comp": 0.12e6},
    (256, 256): {"query_decomp": 333e6, "dmux": 100.3e6, "cmux_crot": 50.5e6, "resp_comp": 0.12e6},
}


def step_work(counts, step_s, Q):
    """Per-step modmul-equivalents of the pass (Q queries) and the rate the
    step's measured time gives them (G/s), where the counts are known."""
    if not counts:
        return None
    t = dict(zip(["query_decomp", "dmux", "mod_switch", "first_dim", "cmux", "crot", "resp_comp"], step_s))
    t["cmux_crot"] = t["cmux"] + t["crot"]
    return {k: {"modmul_eq": c * Q, "ms": round(t[k] * 1e3, 3), "G_per_s": round(c * Q / t[k] / 1e9, 1)}
            for k, c in counts.items() if t[k] > 0}


def fd_plan_bytes(I, J, n, Q):
    """Algorithmic bytes of one k_fd_mfma launch: the prepared database as
    5-byte limb fragments (rows padded to the fragment height, columns to 16),
    the packed selection (5 bytes per residue) and the fp64 segment outputs."""
    ipad = 32 if I <= 32 else -(-I // 64) * 64
    jpad = -(-J // 16) * 16
    kc = 64 if ipad % 64 == 0 else 32
    nseg = -(-(ipad // kc) // (256 // kc))
    return ipad * jpad * n * 5 + n * ipad * 5 * 2 * Q + nseg * n * Q * J * 2 * 8


def synthetic_query(pp, rng):
    q0, q1 = pp.q1[0], pp.q1[1]
    words = pp.query_lwes * 2 * (pp.n1 + 1)
    idx = (np.arange(words) // (pp.n1 + 1)) % 2
    out = rng.integers(0, q0, size=words, dtype=np.uint64)
    out[idx == 1] = rng.integers(0, q1, size=int((idx == 1).sum()), dtype=np.uint64)
    return out


def seeded_inputs(via_gpu, ctx, pp, tag):
    """The workload's inputs from Shake256Prg streams, all drawn by the
    library: the records are fill_bytes of b"via-gpu/<tag>/db" (record
    coefficient 2k = byte k & 0xF, 2k+1 = byte k >> 4), the keys Client::setup
    over b"via-gpu/<tag>/setup" (ternary keys and errors), the index
    uniform_below(n_records) of b"via-gpu/<tag>/query" and the query
    Client::query continuing the setup's stream."""
    n_rec = (pp.n1 // pp.n_rec) * pp.num_rows * pp.num_cols
    g_db = via_gpu.Prg(b"via-gpu/%s/db" % tag.encode())
    record_bytes = g_db.fill(ctx, n_rec * pp.n_rec // 2)
    g_db.close()
    g_setup = via_gpu.Prg(b"via-gpu/%s/setup" % tag.encode())
    keys = via_gpu.client_setup(ctx, pp, g_setup)
    g_q = via_gpu.Prg(b"via-gpu/%s/query" % tag.encode())
    index = g_q.uniform_below(ctx, n_rec)
    lwes = via_gpu.client_query(ctx, pp, keys["sk1"], index, g_setup)
    return keys, lwes, record_bytes, n_rec, index, (g_setup, g_q)


def record_of(record_bytes, pp, index):
    b = record_bytes[index * pp.n_rec // 2:(index + 1) * pp.n_rec // 2].astype(np.uint64)
    rec = np.empty(pp.n_rec, np.uint64)
    rec[0::2] = b & 0xF
    rec[1::2] = b >> 4
    return rec


def cpu_baseline(pp, cascade, conv, rsk, lwes, record_bytes, n_rec):
    """The C restatement (oracle/) on one answer of the same workload, with
    FirstDim fanned out over the host's cores as the reference does."""
    from oracle.vo import Oracle, make_params, u64p

    o = Oracle()
    op = make_params(pp.n1, pp.n2, (pp.q1[0], pp.q1[1]), pp.q2, pp.q3, pp.q4, pp.p, pp.b1, pp.b2, pp.b_rsk,
                     pp.ck_base, pp.l_query, pp.l_ck, pp.l_rsk, pp.num_rows, pp.num_cols)
    b = record_bytes
    recs = np.empty(b.size * 2, dtype=np.uint64)
    recs[0::2] = b & 0xF
    recs[1::2] = b >> 4
    P = lambda a: a.ctypes.data_as(u64p)
    h = C.c_void_p()
    t0 = time.perf_counter()
    assert o.server_new(op, P(recs), n_rec, P(cascade), P(conv), P(rsk), C.byref(h)) == 0
    t_setup = time.perf_counter() - t0
    del recs
    am, ab = np.zeros(pp.n2, np.uint64), np.zeros(pp.n2, np.uint64)
    t0 = time.perf_counter()
    assert o.server_answer(h, P(lwes), pp.query_lwes, P(am), P(ab), None, None) == 0
    t_ans = time.perf_counter() - t0
    steps = (C.c_double * 7)()
    o.server_step_times(h, steps)
    o.server_free(h)
    return am, ab, t_ans, t_setup, list(steps)


def tune_batches(via_gpu, ctx, pp, db, keys, sizes, seed=11):
    """Autotune the query-decompression launch shapes for each batch size in
    `sizes` (the choice is shared by every context on the device, the serving
    loop's included). Returns {size: [(groups, wide) x 12]}."""
    rng = np.random.default_rng(seed)
    out = {}
    for Q in sizes:
        qs = [synthetic_query(pp, rng) for _ in range(Q)]
        d_q = [ctx.alloc(q.nbytes) for q in qs]
        for d, q in zip(d_q, qs):
            ctx.to_device(d, q)
        d_m = [ctx.alloc(8 * pp.n2) for _ in range(Q)]
        d_b = [ctx.alloc(8 * pp.n2) for _ in range(Q)]
        out[Q] = ctx.autotune(db, [keys] * Q, d_q, pp.query_lwes, d_m, d_b, reps=2)
        for d in d_q + d_m + d_b:
            ctx.free(d)
    return out


def serve_setup(via_gpu, ctx, pp, db, n_clients, max_batch, slots, db_b=None):
    """A serving loop with n_clients admitted clients (their eval keys
    resident in HBM under the loop's key budget) and one VIA-C query (and,
    with db_b, one VIA-B batch of T = n2 / n_rec(db_b) queries) per client.
    Client c's keys are Client::setup over Shake256Prg(b"via-gpu/serve/setup/<c>")
    and its queries Client::query continuing that stream, at indices drawn
    from b"via-gpu/serve/query"."""
    kpp = db_b.pp if db_b is not None else pp
    T = pp.n2 // kpp.n_rec if db_b is not None else 0
    loop = via_gpu.ServeLoop(pp, db, key_budget_bytes=64 << 30, max_batch=max_batch, slots=slots, db_b=db_b)
    n_c = (pp.n1 // pp.n_rec) * pp.num_rows * pp.num_cols
    n_b = (kpp.n1 // kpp.n_rec) * kpp.num_rows * kpp.num_cols
    g_q = via_gpu.Prg(b"via-gpu/serve/query")
    queries, bqueries = [], []
    for c in range(n_clients):
        g = via_gpu.Prg(b"via-gpu/serve/setup/%d" % c)
        k = via_gpu.client_setup(ctx, kpp, g)  # the keys do not depend on the record degree
        loop.admit(c, k["cascade_key"], k["conv_key"], k["rsk"])
        queries.append(via_gpu.client_query(ctx, pp, k["sk1"], g_q.uniform_below(ctx, n_c), g))
        if db_b is not None:
            bqueries.append([via_gpu.client_query(ctx, kpp, k["sk1"], g_q.uniform_below(ctx, n_b), g)
                             for _ in range(T)])
        g.close()
    g_q.close()
    # warm every slot once
    warm = [loop.submit(c % n_clients, queries[c % n_clients]) for c in range(2 * max_batch * slots)]
    if db_b is not None:
        warm += [loop.submit_batch(c % n_clients, T, bqueries[c % n_clients]) for c in range(2 * slots)]
    for t in warm:
        loop.wait(t)
    return loop, queries, bqueries, T


def serve_run(loop, queries, bqueries, T, qps, n_req, seed=7, frac_b=0.0, b_sizes=(4, 64)):
    """Poisson arrivals at `qps` through the continuous-batching loop. A
    fraction frac_b of the requests are VIA-B requests for a number of
    records uniform in b_sizes, which the client splits into batches of T
    queries, padding the last one with index-0 queries. Returns answers/s,
    records/s and the latency percentiles of the requests (a VIA-B request
    completes with its last batch)."""
    rng = np.random.default_rng(seed)
    n_clients = len(queries)
    loop.reset_stats()
    gaps = rng.exponential(1.0 / qps, size=n_req)
    who = rng.integers(0, n_clients, size=n_req)
    is_b = (rng.random(n_req) < frac_b) if bqueries else np.zeros(n_req, bool)
    sizes = rng.integers(b_sizes[0], b_sizes[1] + 1, size=n_req)
    reqs = []  # (kind, records, [tickets])
    t0 = time.perf_counter()
    t_next = t0
    for g, c, b, sz in zip(gaps, who, is_b, sizes):
        t_next += g
        while True:
            now = time.perf_counter()
            if now >= t_next:
                break
            if t_next - now > 0.002:
                time.sleep(0.001)
        c = int(c)
        if b:
            nb = -(-int(sz) // T)
            reqs.append(("b", int(sz), [loop.submit_batch(c, T, bqueries[c]) for _ in range(nb)]))
        else:
            reqs.append(("c", 1, [loop.submit(c, queries[c])]))
    lat_c, lat_b = [], []
    for kind, _, ts in reqs:
        worst = max(loop.wait(t)[2] for t in ts)
        (lat_b if kind == "b" else lat_c).append(worst)
    elapsed = time.perf_counter() - t0
    st = loop.stats()
    pct = lambda v, q: round(float(np.percentile(np.array(v) * 1e3, q)), 3) if v else None
    allv = lat_c + lat_b
    out = {"offered_qps": qps, "answers_per_s": round(st["answered"] / elapsed, 2),
           "requests_per_s": round(n_req / elapsed, 2),
           "records_per_s": round(sum(r for _, r, _ in reqs) / elapsed, 2),
           "p50_ms": pct(allv, 50), "p99_ms": pct(allv, 99), "batches": st["batches"],
           "max_batch_seen": st["max_batch_seen"], "requests": n_req}
    if bqueries:
        out.update({"via_b_requests": len(lat_b), "via_b_T": T, "p50_c_ms": pct(lat_c, 50), "p99_c_ms": pct(lat_c, 99),
                    "p50_b_ms": pct(lat_b, 50), "p99_b_ms": pct(lat_b, 99)})
    return out


def serving_measure(via_gpu, ctx, pp, db, args, clients=64, slots=2, max_batch=32):
    """The mixed serving trace of the default line: `clients` clients with
    their own Client::setup keys resident in HBM, Poisson arrivals at
    --serve-qps, 70% VIA-C queries against `db` and 30% VIA-B requests of a
    number of records uniform in [4, 64] as T = 8 batches against a second
    database of the same grid at record degree 64 (records from
    b"via-gpu/serve/db-b"). Every answer is the library's answer of that
    request alone (tests/test_gpu_serving.py)."""
    pb = via_gpu.paper_b_params(pp.num_rows, pp.num_cols, T=8)
    nb_rec = (pb.n1 // pb.n_rec) * pb.num_rows * pb.num_cols
    g_b = via_gpu.Prg(b"via-gpu/serve/db-b")
    rb = g_b.fill(ctx, nb_rec * pb.n_rec // 2)
    g_b.close()
    db_b = via_gpu.PreparedDb(ctx, pb, record_bytes=rb.tobytes(), n_records=nb_rec)
    del rb
    t0 = time.perf_counter()
    loop, queries, bqueries, T = serve_setup(via_gpu, ctx, pp, db, clients, max_batch=max_batch, slots=slots,
                                             db_b=db_b)
    t_setup = time.perf_counter() - t0
    r = serve_run(loop, queries, bqueries, T, args.serve_qps, args.serve_requests, frac_b=0.3)
    st = loop.stats()
    rc = None
    if args.serve_c_qps > 0:
        # the same loop and clients under VIA-C queries only, at a rate near
        # the loop's saturation: continuous batching against the one-query path
        rc = serve_run(loop, queries, None, T, args.serve_c_qps, args.serve_c_requests, frac_b=0.0)
        rc.update({"workload": f"serving loop, Poisson arrivals of VIA-C queries from {clients} clients, "
                               f"batches <= {max_batch}, {slots} slots, I={pp.num_rows} J={pp.num_cols}"})
    loop.close()
    db_b.close()
    r.update({"workload": f"serving loop, Poisson arrivals from {clients} clients (keys resident in HBM), "
                          f"70% VIA-C queries / 30% VIA-B requests of 4..64 records in T=8 batches, "
                          f"VIA-C batches <= {max_batch}, {slots} slots, I={pp.num_rows} J={pp.num_cols}",
              "clients": clients, "key_bytes": st["key_bytes"], "setup_s": round(t_setup, 2),
              "data": "seeded: client c's keys Client::setup of b'via-gpu/serve/setup/<c>' and its queries "
                      "Client::query, indices uniform_below of b'via-gpu/serve/query', arrivals from numpy seed 7"})
    if rc is not None:
        r["via_c_saturation"] = rc
    return r


def vb_measure(via_gpu, ctx, T, rows, cols, indices, steps, warmup):
    """VIA-B: `indices` seeded distinct indices answered as batched queries of
    T (answer_batch::<T>, device buffers) against a database at record degree
    n2/T; one step answers all of them. Returns records/s, per-batch latency,
    the FirstDim pass's roofline (HIP events around k_fd_mfma of the last
    pass) and whether every batch recovers its T records."""
    pp = via_gpu.paper_b_params(rows, cols, T=T)
    I, J, n = pp.num_rows, pp.num_cols, pp.n1
    n_rec = (n // pp.n_rec) * I * J
    tag = f"vb{T}-{I}x{J}".encode()
    # records, keys, distinct indices and batch queries from seeded streams
    g_db = via_gpu.Prg(b"via-gpu/%s/db" % tag)
    record_bytes = g_db.fill(ctx, n_rec * pp.n_rec // 2)
    g_db.close()
    g = via_gpu.Prg(b"via-gpu/%s/setup" % tag)
    ck = via_gpu.client_setup(ctx, pp, g)
    t0 = time.perf_counter()
    db = via_gpu.PreparedDb(ctx, pp, record_bytes=record_bytes.tobytes(), n_records=n_rec)
    keys = via_gpu.PreparedKeys(ctx, pp, ck["cascade_key"], ck["conv_key"], ck["rsk"])
    t_setup = time.perf_counter() - t0
    nb = max(1, indices // T)
    g_q = via_gpu.Prg(b"via-gpu/%s/query" % tag)
    idx = []
    while len(idx) < nb * T:
        i = g_q.uniform_below(ctx, n_rec)
        if i not in idx:
            idx.append(i)
    g_q.close()
    d_q, d_m, d_b = [], [], []
    for b in range(nb):
        qs = np.concatenate([via_gpu.client_query(ctx, pp, ck["sk1"], i, g) for i in idx[b * T:(b + 1) * T]])
        d = ctx.alloc(qs.nbytes)
        ctx.to_device(d, qs)
        d_q.append(d)
        d_m.append(ctx.alloc(8 * pp.n2))
        d_b.append(ctx.alloc(8 * pp.n2))
    g.close()
    L = via_gpu.lib()
    err = via_gpu.ErrDetail()

    def one(b):
        st = L.via_answer_batch_device(ctx.h, db.h, keys.h, T, d_q[b], pp.query_lwes, d_m[b], d_b[b], C.byref(err))
        if st:
            raise via_gpu.ViaError(st, msg=L.via_status_str(st).decode())

    for b in range(min(nb, max(1, warmup))):
        one(b)
    ctx.synchronize()
    lat, fd_ms = [], []
    t_start = time.perf_counter()
    for _ in range(steps):
        for b in range(nb):
            t0 = time.perf_counter()
            one(b)
            ctx.synchronize()
            lat.append(time.perf_counter() - t0)
            fd_ms.append(ctx.first_dim_time() * 1e3)
    elapsed = time.perf_counter() - t_start
    lat_ms = np.array(lat) * 1e3
    recs = steps * nb * T
    # every batch's answer recovers its T records (recover + deinterleave_batch)
    ok = True
    for b in range(nb):
        m = ctx.to_host(np.zeros(pp.n2, np.uint64), d_m[b])
        bd = ctx.to_host(np.zeros(pp.n2, np.uint64), d_b[b])
        got = via_gpu.client_deinterleave(pp, T, via_gpu.client_recover(pp, ck["sk2"], m, bd))
        ok = ok and all(np.array_equal(got[t], record_of(record_bytes, pp, i))
                        for t, i in enumerate(idx[b * T:(b + 1) * T]))
    for d in d_q + d_m + d_b:
        ctx.free(d)
    keys.close()
    db.close()
    # the last pass of a batch holds min(T, 32) queries
    Qp = min(T, 32)
    fd_bytes = fd_plan_bytes(I, J, n, Qp)
    fd_s = float(np.median(fd_ms)) * 1e-3
    return {
        "T": T, "records_per_s": round(recs / elapsed, 2), "batch_answers_per_s": round(steps * nb / elapsed, 3),
        "batches_per_step": nb, "steps": steps,
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3), "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "latency_samples": len(lat), "recovered": bool(ok), "setup_s": round(t_setup, 3),
        "workload": f"VIA-B, {nb * T} distinct indices per step as {nb} batch(es) of T={T}, record degree "
                    f"{pp.n_rec} ({pp.n_rec // 2} B records), I={I} J={J}, prepared DB "
                    f"{I * J * n * 5 / 2**30:.2f} GiB as 7-bit limb fragments "
                    f"({I * J * n * 8 / 2**30:.2f} GiB as u64 residues)",
        "roofline": {"kernel": f"k_fd_mfma, one database pass for {Qp} queries", "bound": "hbm",
                     "achieved": round(fd_bytes / fd_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(fd_bytes / fd_s / 1e9 / HBM_PEAK_GBS, 4), "algorithmic_bytes": fd_bytes,
                     "ms": round(fd_s * 1e3, 4)},
        "data": f"seeded: records fill_bytes of Shake256Prg(b'via-gpu/{tag.decode()}/db'), keys Client::setup and "
                f"queries Client::query of b'via-gpu/{tag.decode()}/setup', distinct indices uniform_below of "
                f"b'via-gpu/{tag.decode()}/query'",
    }


def vb_run(via_gpu, args):
    """VIA-B mode (--vb T): one JSON line for batch size T."""
    ctx = via_gpu.Context(0)
    r = vb_measure(via_gpu, ctx, args.vb, args.rows, args.cols, args.indices, args.steps, args.warmup)
    ctx.close()
    out = {
        "metric": "VIA-B records/sec (answer_batch of T seeded indices) + p50/p99 batch answer latency",
        "value": r["records_per_s"], "unit": "records/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * r["T"] * r["batches_per_step"] / r["records_per_s"], 3),
        "batch_answers_per_s": r["batch_answers_per_s"], "T": r["T"], "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int (exact modular, residues < 2^38 in fp64 lanes)", "data": r["data"],
        "recovered": r["recovered"], "roofline": r["roofline"],
        "config": {"workload": r["workload"], "num_rows": args.rows, "num_cols": args.cols, "T": r["T"]},
        "setup_s": r["setup_s"],
    }
    print(json.dumps(out))


NTT_PRIMES = (268369921, 536608769)


def ntt_run(via_gpu, args):
    """Batched negacyclic NTT -> pointwise -> iNTT: `--polys` RNS polynomial
    pairs at n=2048 over the two primes 268369921 and 536608769 (both = 1 mod
    4096), residues uniform below each prime from a seed, u64 storage. One step
    is one launch computing every product c = iNTT(NTT(a) * NTT(b)); its time
    comes from HIP events on the library stream around each launch."""
    n, P = 2048, args.polys
    rng = np.random.default_rng(2048)
    a = np.empty((P, 2, n), np.uint64)
    b = np.empty((P, 2, n), np.uint64)
    for l, q in enumerate(NTT_PRIMES):
        a[:, l, :] = rng.integers(0, q, size=(P, n), dtype=np.uint64)
        b[:, l, :] = rng.integers(0, q, size=(P, n), dtype=np.uint64)
    ctx = via_gpu.Context(0)
    nb = a.nbytes
    da, db, dc = ctx.alloc(nb), ctx.alloc(nb), ctx.alloc(nb)
    ctx.to_device(da, a)
    ctx.to_device(db, b)
    count = 2 * P  # limb polys; poly p is limb p % 2 (one RNS poly = [limb0][limb1])
    for _ in range(args.warmup):
        ctx.poly_mul_device(n, NTT_PRIMES, 1, da, db, dc, count)
    ctx.synchronize()
    kt = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ctx.mark(0)
        ctx.poly_mul_device(n, NTT_PRIMES, 1, da, db, dc, count)
        ctx.mark(1)
        kt.append(ctx.marked_time())
    elapsed = time.perf_counter() - t_start
    c = ctx.to_host(np.zeros_like(a), dc)
    if args.calib_copy:
        ctx.copy_device(dc, da, nb)
        ctx.synchronize()
    k_s = float(np.mean(kt))
    bytes_ = 3 * nb
    logn = 11
    bfly = 3 * count * (n // 2) * logn
    modmul = bfly + 2 * count * n  # butterflies + pointwise + N^-1 scaling
    out = {
        "metric": "negacyclic NTT->pointwise->iNTT products/sec (n=2048, two-prime RNS), bit-exact vs CPU",
        "value": round(P / k_s, 1), "unit": "RNS products/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "kernel_ms": round(k_s * 1e3, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int (exact modular, residues < 2^30 in fp64 lanes)",
        "data": f"synthetic: residues uniform below each prime, seed 2048",
        "config": {"workload": f"{P} polynomial pairs x n={n} x primes {NTT_PRIMES[0]}, {NTT_PRIMES[1]}; "
                               f"c = iNTT(NTT(a) . NTT(b)) per prime", "polys": P, "n": n,
                   "primes": list(NTT_PRIMES)},
        "roofline": {"kernel": "k_polymul (fused forward NTTs, pointwise product, inverse NTT)", "bound": "hbm",
                     "achieved": round(bytes_ / k_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bytes_ / k_s / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes": bytes_},
        "modmul": {"per_launch": modmul, "butterflies": bfly, "achieved_G_per_s": round(modmul / k_s / 1e9, 1)},
    }
    if not args.no_cpu:
        from oracle.vo import Oracle, Ring, u64p

        o = Oracle()
        r = Ring.make(n, *NTT_PRIMES)
        S = min(P, args.ntt_sample)
        pick = np.random.default_rng(7).choice(P, size=S, replace=False)
        want = np.zeros((S, 2, n), np.uint64)
        asub = np.ascontiguousarray(a[pick])
        bsub = np.ascontiguousarray(b[pick])
        t0 = time.perf_counter()
        for i in range(S):
            o.poly_mul(r, want[i].ctypes.data_as(u64p), asub[i].ctypes.data_as(u64p), bsub[i].ctypes.data_as(u64p))
        t_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(S / t_cpu, 2), "unit": "RNS products/s", "cores": 1, "kind": "port",
                               "sample": f"{S} of the {P} products (C restatement of ring/ntt.rs, one thread, "
                                         f"{t_cpu:.2f} s)"}
        out["bit_exact_vs_cpu"] = bool(np.array_equal(want, c[pick]))
    print(json.dumps(out))
    for d in (da, db, dc):
        ctx.free(d)
    ctx.close()


def setup_many_measure(via_gpu, ctx, pp, clients=64):
    """Client::setup of `clients` clients through via_client_setup_many (host
    threads on lanes of the context), each its own stream
    b"via-gpu/setup-many/<c>"; client 0's keys compared word for word with a
    one-at-a-time via_client_setup of the same stream."""
    prgs = [via_gpu.Prg(b"via-gpu/setup-many/%d" % c) for c in range(clients)]
    via_gpu.client_setup_many(ctx, pp, prgs[:8])  # warm the lanes
    for g in prgs:
        g.close()
    prgs = [via_gpu.Prg(b"via-gpu/setup-many/%d" % c) for c in range(clients)]
    t0 = time.perf_counter()
    many = via_gpu.client_setup_many(ctx, pp, prgs)
    el = time.perf_counter() - t0
    for g in prgs:
        g.close()
    g = via_gpu.Prg(b"via-gpu/setup-many/0")
    one = via_gpu.client_setup(ctx, pp, g)
    g.close()
    same = all(np.array_equal(one[k], many[0][k]) for k in ("sk1", "sk2", "cascade_key", "conv_key", "rsk"))
    return {"clients": clients, "setups_per_s": round(clients / el, 2), "s": round(el, 3), "same_as_one": bool(same)}


def keygen_run(via_gpu, args):
    """Client::setup and Client::query of the library (PRG stream on the GPU,
    host parse, GPU products) at the paper stack; every key word checked
    against the C restatement on the first setup."""
    pp = via_gpu.paper_params(args.rows, args.cols)
    ctx = via_gpu.Context(0)
    g = via_gpu.Prg(b"bench/client-keys")
    for _ in range(args.warmup):
        keys = via_gpu.client_setup(ctx, pp, g)
    t_setup = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        keys = via_gpu.client_setup(ctx, pp, g)
        t_setup.append(time.perf_counter() - t)
    elapsed = time.perf_counter() - t0
    t_query = []
    for i in range(4 * args.steps):
        t = time.perf_counter()
        via_gpu.client_query(ctx, pp, keys["sk1"], (7919 * i) % (4 * args.rows * args.cols), g)
        t_query.append(time.perf_counter() - t)
    wc, wv, wr = via_gpu.client_key_words(pp)
    out = {
        "metric": "client setups/sec (Client::setup: S1, S2, cascade key, conversion key, ring-switch key), "
                  "bit-exact vs CPU",
        "value": round(args.steps / elapsed, 2), "unit": "setups/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "p50_ms": round(float(np.median(t_setup)) * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int (exact modular arithmetic, u64 residues)",
        "data": "seeded: keys drawn from Shake256Prg(b'bench/client-keys')",
        "config": {"workload": f"paper stack n1=2048 n2=512, L_CK=18 (base 18), L_RSK=8; keys of "
                               f"{8 * (wc + wv + wr) / 2**20:.2f} MiB per client; query of "
                               f"{pp.query_lwes} LWEs at I={args.rows} J={args.cols}",
                   "num_rows": args.rows, "num_cols": args.cols},
        "query_ms_p50": round(float(np.median(t_query)) * 1e3, 3),
        "one_at_a_time_setups_per_s": round(args.steps / elapsed, 2),
    }
    # many clients at once: the throughput a server sees admitting clients
    out["many"] = setup_many_measure(via_gpu, ctx, pp, clients=max(32, args.steps * 4))
    out["value"] = out["many"]["setups_per_s"]
    out["ms_per_step"] = round(1e3 / out["value"], 3)
    if not args.no_cpu:
        from oracle.vo import TERNARY, Oracle, paper_params, u64p

        o = Oracle()
        opp = paper_params(args.rows, args.cols)
        og = o.prg(b"bench/client-keys/cpu")
        bufs = [np.zeros(2 * pp.n1, np.uint64), np.zeros(pp.n2, np.uint64), np.zeros(wc, np.uint64),
                np.zeros(wv, np.uint64), np.zeros(wr, np.uint64)]
        t0 = time.perf_counter()
        o.client_setup(opp, TERNARY, TERNARY, TERNARY, og.h, *[b.ctypes.data_as(u64p) for b in bufs])
        t_cpu = time.perf_counter() - t0
        g2 = via_gpu.Prg(b"bench/client-keys/cpu")
        mine = via_gpu.client_setup(ctx, pp, g2)
        same = all(np.array_equal(mine[k], b) for k, b in zip(("sk1", "sk2", "cascade_key", "conv_key", "rsk"), bufs))
        out["cpu_baseline"] = {"value": round(1.0 / t_cpu, 3), "unit": "setups/s", "cores": 1, "kind": "port",
                               "sample": f"one Client::setup by the C restatement, one thread ({t_cpu:.2f} s)"}
        out["bit_exact_vs_cpu"] = bool(same)
        g2.close()
    print(json.dumps(out))
    g.close()
    ctx.close()


def tree_modes(ctx, answer, n=30):
    """The DMux and CMux step times (ms, HIP events) of `n` answers with the
    trees launched level by level (mode 0) and as one persistent cooperative
    launch per tree (mode 1); the context's mode is restored to the tuned
    choice afterwards."""
    out = {}
    for mode in (0, 1):
        ctx.set_tree_mode(mode)
        answer()
        ctx.synchronize()
        acc = np.zeros(7)
        for _ in range(n):
            answer()
            ctx.synchronize()
            acc += np.array(ctx.step_times())
        out["persistent" if mode else "per_level"] = {"dmux_ms": round(acc[1] / n * 1e3, 4),
                                                      "cmux_ms": round(acc[4] / n * 1e3, 4)}
    ctx.set_tree_mode(-1)
    return out


def batched_run(via_gpu, ctx, pp, db, Q, args, seed=31):
    """Q queries from Q distinct clients answered per pass (via_answer_many):
    client q's keys are Client::setup over b"via-gpu/<tag>/batch/setup/<q>",
    its query Client::query at an index drawn from b"via-gpu/<tag>/batch/query".
    Autotuned for Q, warmed, then `steps` timed passes."""
    tag = f"paper-{pp.num_rows}x{pp.num_cols}".encode()
    n_rec = (pp.n1 // pp.n_rec) * pp.num_rows * pp.num_cols
    g_q = via_gpu.Prg(b"via-gpu/%s/batch/query" % tag)
    keys, queries = [], []
    for q in range(Q):
        g = via_gpu.Prg(b"via-gpu/%s/batch/setup/%d" % (tag, q))
        k = via_gpu.client_setup(ctx, pp, g)
        keys.append(via_gpu.PreparedKeys(ctx, pp, k["cascade_key"], k["conv_key"], k["rsk"]))
        queries.append(via_gpu.client_query(ctx, pp, k["sk1"], g_q.uniform_below(ctx, n_rec), g))
        g.close()
    g_q.close()
    d_q = [ctx.alloc(q.nbytes) for q in queries]
    for d, q in zip(d_q, queries):
        ctx.to_device(d, q)
    d_m = [ctx.alloc(8 * pp.n2) for _ in range(Q)]
    d_b = [ctx.alloc(8 * pp.n2) for _ in range(Q)]
    L = via_gpu.lib()
    err = via_gpu.ErrDetail()
    vp = C.c_void_p
    Karr = (vp * Q)(*[k.h.value for k in keys])
    Qarr = (vp * Q)(*[d.value for d in d_q])
    Marr = (vp * Q)(*[d.value for d in d_m])
    Barr = (vp * Q)(*[d.value for d in d_b])

    def answer():
        st = L.via_answer_many_device(ctx.h, db.h, Q, Karr, Qarr, pp.query_lwes, Marr, Barr, C.byref(err))
        if st:
            raise via_gpu.ViaError(st, msg=L.via_status_str(st).decode())

    if not args.no_tune:
        ctx.autotune(db, keys, d_q, pp.query_lwes, d_m, d_b, reps=2)
    ctx.set_step_timing(False)  # nothing in the timed passes reads the step events
    for _ in range(max(1, args.warmup)):
        answer()
    ctx.synchronize()
    lat = []
    steps = max(3, args.steps // 2)
    t_start = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        answer()
        ctx.synchronize()
        lat.append(time.perf_counter() - t0)
    elapsed = time.perf_counter() - t_start
    lat_ms = np.array(lat) * 1e3
    ctx.set_step_timing(True)
    trees = tree_modes(ctx, answer, n=5)
    ctx.set_step_timing(False)
    for d in d_q + d_m + d_b:
        ctx.free(d)
    for k in keys:
        k.close()
    return {"queries_per_pass": Q, "answers_per_s": round(steps * Q / elapsed, 2), "passes": steps,
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 3), "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
            "tree_tuning": ctx.tree_tuning(Q), "tree_modes": trees}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=256)
    ap.add_argument("--cols", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=1, help="queries per batched pass (distinct clients)")
    ap.add_argument("--batched", type=int, default=32,
                    help="with one query per pass: also measure this many queries per pass (0: skip)")
    ap.add_argument("--serve", action="store_true", help="open-loop Poisson arrivals through the serving loop")
    ap.add_argument("--qps", type=float, default=0.0, help="serving arrival rate (0: sweep)")
    ap.add_argument("--rates", type=str, default="", help="serving: comma-separated arrival rates to sweep")
    ap.add_argument("--clients", type=int, default=64)
    ap.add_argument("--requests", type=int, default=600)
    ap.add_argument("--mix", action="store_true", help="serving: 30%% VIA-B requests against a second database")
    ap.add_argument("--vb", type=int, default=0, help="VIA-B batch size T (answer_batch), 0: VIA-C")
    ap.add_argument("--indices", type=int, default=64, help="VIA-B indices per step")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (heuristic digit groups)")
    ap.add_argument("--graphs", action="store_true", help="also time the answers replayed from a hipGraph")
    ap.add_argument("--ntt", action="store_true", help="NTT->pointwise->iNTT microbench")
    ap.add_argument("--polys", type=int, default=8192, help="NTT microbench: RNS polynomial pairs")
    ap.add_argument("--ntt-sample", type=int, default=8192, help="NTT microbench: products checked on the CPU")
    ap.add_argument("--keygen", action="store_true", help="client key generation and query encryption")
    ap.add_argument("--lat-samples", type=int, default=400,
                    help="one query per pass: answers timed one by one for the latency percentiles")
    ap.add_argument("--host-samples", type=int, default=100,
                    help="one query per pass: answers with host query and answer buffers (via_answer, PCIe included)")
    ap.add_argument("--vb-sweep", type=str, default="8,32,64",
                    help="default line: VIA-B batch sizes measured at --vb-rows x --vb-cols ('' to skip)")
    ap.add_argument("--vb-rows", type=int, default=32)
    ap.add_argument("--vb-cols", type=int, default=4096)
    ap.add_argument("--serve-qps", type=float, default=80.0,
                    help="default line: offered request rate of the mixed serving trace (0 to skip)")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-c-qps", type=float, default=5800.0,
                    help="default line: VIA-C-only arrival rate for the serving loop's saturation point (0: skip)")
    ap.add_argument("--serve-c-requests", type=int, default=3000)
    ap.add_argument("--keygen-clients", type=int, default=32,
                    help="default line: Client::setup of this many clients at once (0 to skip)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed region, one device copy of known size (256 MiB; one NTT operand "
                         "in --ntt mode): a plain copy kernel for counter calibration")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the answer path is single-GPU in this build "
              "(multi-GPU sharding is designed, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import __graft_entry__

    via_gpu = __graft_entry__.load_package()
    if args.vb:
        vb_run(via_gpu, args)
        return
    if args.ntt:
        ntt_run(via_gpu, args)
        return
    if args.keygen:
        keygen_run(via_gpu, args)
        return
    pp = via_gpu.paper_params(args.rows, args.cols)
    ctx = via_gpu.Context(0)
    tag = f"paper-{args.rows}x{args.cols}"
    ckeys, lwes, record_bytes, n_rec, index, prgs = seeded_inputs(via_gpu, ctx, pp, tag)
    cascade, conv, rsk = ckeys["cascade_key"], ckeys["conv_key"], ckeys["rsk"]
    t0 = time.perf_counter()
    db = via_gpu.PreparedDb(ctx, pp, record_bytes=record_bytes.tobytes(), n_records=n_rec)
    keys = via_gpu.PreparedKeys(ctx, pp, cascade, conv, rsk)
    t_setup = time.perf_counter() - t0

    if args.serve:
        if args.rates:
            rates = [float(x) for x in args.rates.split(",")]
        else:
            rates = [args.qps] if args.qps > 0 else [200.0, 400.0, 800.0, 1200.0, 1600.0, 2400.0]
        db_b = None
        if args.mix:
            pb = via_gpu.paper_b_params(args.rows, args.cols, T=8)
            nb_rec = (pb.n1 // pb.n_rec) * pb.num_rows * pb.num_cols
            g_b = via_gpu.Prg(b"via-gpu/serve/db-b")
            rb = g_b.fill(ctx, nb_rec * pb.n_rec // 2)
            g_b.close()
            db_b = via_gpu.PreparedDb(ctx, pb, record_bytes=rb.tobytes(), n_records=nb_rec)
            del rb
        tuned = None
        if not args.no_tune:
            tuned = tune_batches(via_gpu, ctx, pp, db, keys, [q for q in (1, 2, 4, 8, 16, 32) if q <= args.batch])
        loop, queries, bqueries, T = serve_setup(via_gpu, ctx, pp, db, args.clients, max_batch=args.batch, slots=2,
                                                 db_b=db_b)
        runs = []
        for r in rates:
            runs.append(serve_run(loop, queries, bqueries, T, r, args.requests, frac_b=0.3 if args.mix else 0.0))
            print(json.dumps({"serving": runs[-1]}), file=sys.stderr)
        loop.close()
        best = max(runs, key=lambda x: x["answers_per_s"])
        out = {
            "metric": "server answers/sec + p50/p99 answer latency, VIA-C n=2048, bit-exact vs CPU",
            "value": best["answers_per_s"], "unit": "answers/s", "n_gpus": 1, "steps": args.requests,
            "warmup": 0, "ms_per_step": round(1e3 / best["answers_per_s"], 3),
            "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "int (exact modular, residues < 2^38 in fp64 lanes)",
            "data": f"seeded: records fill_bytes of Shake256Prg(b'via-gpu/{tag}/db') (and b'via-gpu/serve/db-b'), "
                    f"client c's keys Client::setup of b'via-gpu/serve/setup/<c>' and its queries Client::query, "
                    f"indices uniform_below of b'via-gpu/serve/query'",
            "config": {"workload": (f"serving loop, Poisson arrivals from {args.clients} clients, "
                                    + ("70% VIA-C queries / 30% VIA-B requests of 4..64 records in T=8 batches, "
                                       if args.mix else "VIA-C queries, ")
                                    + f"VIA-C batches <= {args.batch}, 2 slots, I={pp.num_rows} J={pp.num_cols}")},
            "sweep": runs,
            "tuned_digit_groups": tuned,
        }
        print(json.dumps(out))
        ctx.close()
        return

    L = via_gpu.lib()
    Q = args.batch
    # clients 1 .. Q-1: their own Client::setup and a query at a seeded index
    client_keys, queries = [keys], [lwes]
    for q in range(1, Q):
        g = via_gpu.Prg(b"via-gpu/%s/setup/%d" % (tag.encode(), q))
        k = via_gpu.client_setup(ctx, pp, g)
        client_keys.append(via_gpu.PreparedKeys(ctx, pp, k["cascade_key"], k["conv_key"], k["rsk"]))
        queries.append(via_gpu.client_query(ctx, pp, k["sk1"], prgs[1].uniform_below(ctx, n_rec), g))
        g.close()
    d_q = [ctx.alloc(q.nbytes) for q in queries]
    for d, q in zip(d_q, queries):
        ctx.to_device(d, q)
    d_m = [ctx.alloc(8 * pp.n2) for _ in range(Q)]
    d_b = [ctx.alloc(8 * pp.n2) for _ in range(Q)]
    err = via_gpu.ErrDetail()
    vp = C.c_void_p
    Karr = (vp * Q)(*[k.h.value for k in client_keys])
    Qarr = (vp * Q)(*[d.value for d in d_q])
    Marr = (vp * Q)(*[d.value for d in d_m])
    Barr = (vp * Q)(*[d.value for d in d_b])

    def answer():
        st = L.via_answer_many_device(ctx.h, db.h, Q, Karr, Qarr, pp.query_lwes, Marr, Barr, C.byref(err))
        if st:
            raise via_gpu.ViaError(st, msg=L.via_status_str(st).decode())

    tuned = None
    if not args.no_tune:
        tuned = ctx.autotune(db, client_keys, d_q, pp.query_lwes, d_m, d_b, reps=2)
    # the timed answers record only the FirstDim events (the roofline's);
    # the step split is measured on answers of its own after them
    ctx.set_step_timing(False)
    for _ in range(args.warmup):
        answer()
    ctx.synchronize()
    lat = []
    fd_ms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        answer()
        ctx.synchronize()
        lat.append(time.perf_counter() - t0)
        fd_ms.append(ctx.first_dim_time() * 1e3)
    elapsed = time.perf_counter() - t_start
    ctx.set_step_timing(True)
    n_split = max(3, min(args.steps, 50))
    steps_acc = np.zeros(7)
    for _ in range(n_split):
        answer()
        ctx.synchronize()
        steps_acc += np.array(ctx.step_times())
    ctx.set_step_timing(False)
    if args.calib_copy:
        cb = 256 << 20
        c_src, c_dst = ctx.alloc(cb), ctx.alloc(cb)
        ctx.to_device(c_src, np.arange(cb // 8, dtype=np.uint64))  # a written source, as a real copy reads
        ctx.copy_device(c_dst, c_src, cb)
        ctx.synchronize()
        ctx.free(c_src)
        ctx.free(c_dst)
    gpu_mask = ctx.to_host(np.zeros(pp.n2, np.uint64), d_m[0])
    gpu_body = ctx.to_host(np.zeros(pp.n2, np.uint64), d_b[0])

    I, J, n = pp.num_rows, pp.num_cols, pp.n1
    fd_bytes = fd_plan_bytes(I, J, n, Q)
    fd_u64_bytes = I * J * n * 8 + Q * (2 * I * n * 8 + 2 * J * n * 8)
    fd_s = float(np.mean(fd_ms)) * 1e-3
    fd_gbs = fd_bytes / fd_s / 1e9
    lat_ms = np.array(lat) * 1e3
    value = args.steps * Q / elapsed

    out = {
        "metric": "server answers/sec + p50/p99 answer latency, VIA-C n=2048, bit-exact vs CPU",
        "value": round(value, 3),
        "unit": "answers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "batch": Q,
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int (exact modular, residues < 2^38 in fp64 lanes)",
        "data": f"seeded: records fill_bytes of Shake256Prg(b'via-gpu/{tag}/db'), keys Client::setup and "
                f"queries Client::query of b'via-gpu/{tag}/setup' (ternary), index {index} = uniform_below of "
                f"b'via-gpu/{tag}/query'",
        "recovered": bool(np.array_equal(via_gpu.client_recover(pp, ckeys["sk2"], gpu_mask, gpu_body),
                                         record_of(record_bytes, pp, index))),
        "config": {
            "workload": f"VIA-C paper stack, {Q} quer{'y' if Q == 1 else 'ies from distinct clients'} per pass, "
                        f"I={I} J={J}, {n_rec} records x 256 B, "
                        f"prepared DB {I * J * n * 5 / 2**30:.3f} GiB as 7-bit limb fragments "
                        f"({I * J * n * 8 / 2**30:.2f} GiB as u64 residues)",
            "num_rows": I, "num_cols": J, "n1": n, "n2": pp.n2, "query_lwes": pp.query_lwes,
        },
        "step_ms": {k: round(v / n_split * 1e3, 3) for k, v in zip(
            ["query_decomp", "dmux", "mod_switch", "first_dim", "cmux", "crot", "resp_comp"], steps_acc)},
        "step_ms_answers": n_split,
        "setup_s": round(t_setup, 3),
        "tuned_digit_groups": tuned,
        "step_work": step_work(STEP_MODMUL.get((I, J)), steps_acc / n_split, Q),
        "roofline": {
            "kernel": "k_fd_mfma (database limb stream, int8 MFMA multiply-accumulate)",
            "bound": "hbm",
            "achieved": round(fd_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(fd_gbs / HBM_PEAK_GBS, 4),
            "traffic": FD_TRAFFIC.get((I, J, Q, (ctx.tree_tuning(Q) or {}).get("fd_groups", 0))),
            "fd_groups": (ctx.tree_tuning(Q) or {}).get("fd_groups"),
            "algorithmic_bytes": fd_bytes,
            # the same launch priced at the reference's u64 database storage
            # (8 bytes per residue), which this layout does not move
            "u64_equivalent_bytes": fd_u64_bytes,
            "u64_equivalent_GBs": round(fd_u64_bytes / fd_s / 1e9, 1),
        },
    }
    out["tree_tuning"] = ctx.tree_tuning(Q)
    ctx.set_step_timing(True)
    out["tree_modes"] = tree_modes(ctx, answer)
    ctx.set_step_timing(False)
    if Q == 1 and args.lat_samples > 0:
        # the latency percentiles over many single answers (device buffers)
        slat = []
        for _ in range(args.lat_samples):
            t0 = time.perf_counter()
            answer()
            ctx.synchronize()
            slat.append(time.perf_counter() - t0)
        slat_ms = np.array(slat) * 1e3
        out["latency"] = {"samples": len(slat), "p50_ms": round(float(np.percentile(slat_ms, 50)), 3),
                          "p99_ms": round(float(np.percentile(slat_ms, 99)), 3),
                          "max_ms": round(float(slat_ms.max()), 3)}
        out["p50_ms"], out["p99_ms"] = out["latency"]["p50_ms"], out["latency"]["p99_ms"]
        out["p_samples"] = len(slat)
    if Q == 1 and args.host_samples > 0:
        # Server::answer's own contract: query and answer in host memory, the
        # ~1.2 MB of LWEs crossing PCIe inside each answer (via_answer)
        server = via_gpu.Server(ctx, pp, db, keys)
        server.answer(lwes)
        hlat = []
        t_h = time.perf_counter()
        for _ in range(args.host_samples):
            t0 = time.perf_counter()
            hm, hb = server.answer(lwes)
            hlat.append(time.perf_counter() - t0)
        h_el = time.perf_counter() - t_h
        hlat_ms = np.array(hlat) * 1e3
        out["host_buffers"] = {"answers_per_s": round(args.host_samples / h_el, 3), "samples": args.host_samples,
                               "p50_ms": round(float(np.percentile(hlat_ms, 50)), 3),
                               "p99_ms": round(float(np.percentile(hlat_ms, 99)), 3),
                               "same_answer": bool(np.array_equal(hm, gpu_mask) and np.array_equal(hb, gpu_body))}
    if args.graphs:
        # the same answers replayed from a captured hipGraph
        ctx.set_graphs(True)
        answer()  # eager run and capture
        ctx.synchronize()
        glat = []
        t_g = time.perf_counter()
        for _ in range(args.steps):
            t0 = time.perf_counter()
            answer()
            ctx.synchronize()
            glat.append(time.perf_counter() - t0)
        g_el = time.perf_counter() - t_g
        ctx.set_graphs(False)
        glat_ms = np.array(glat) * 1e3
        out["graph"] = {"answers_per_s": round(args.steps * Q / g_el, 3),
                        "p50_ms": round(float(np.percentile(glat_ms, 50)), 3),
                        "p99_ms": round(float(np.percentile(glat_ms, 99)), 3)}
    if args.batched > 1 and Q == 1:
        out["batched"] = batched_run(via_gpu, ctx, pp, db, args.batched, args)
        out["batched"]["vs_one_query"] = round(out["batched"]["answers_per_s"] / out["value"], 2)
    if Q == 1 and args.vb_sweep:
        out["via_b"] = [vb_measure(via_gpu, ctx, int(T), args.vb_rows, args.vb_cols, 64, 2, 1)
                        for T in args.vb_sweep.split(",")]
    if Q == 1 and args.serve_qps > 0:
        out["serving"] = serving_measure(via_gpu, ctx, pp, db, args)
        sat = out["serving"].get("via_c_saturation")
        if sat:
            sat["vs_one_query"] = round(sat["answers_per_s"] / out["value"], 2)
    if Q == 1 and args.keygen_clients > 0:
        out["client_setup"] = setup_many_measure(via_gpu, ctx, pp, clients=args.keygen_clients)
    if not args.no_cpu:
        am, ab, t_cpu, t_cpu_setup, cpu_steps = cpu_baseline(pp, cascade, conv, rsk, lwes, record_bytes, n_rec)
        from oracle.vo import Oracle

        cores = Oracle().host_threads()
        out["cpu_baseline"] = {
            "value": round(1.0 / t_cpu, 5),
            "unit": "answers/s",
            "cores": cores,
            "kind": "port",
            "sample": f"one answer of the same workload (C restatement, FirstDim over {cores} threads, "
                      f"{t_cpu:.2f} s; setup {t_cpu_setup:.2f} s)",
            "step_s": [round(x, 4) for x in cpu_steps],
        }
        out["bit_exact_vs_cpu"] = bool(np.array_equal(am, gpu_mask) and np.array_equal(ab, gpu_body))
    print(json.dumps(out))
    for d in d_q + d_m + d_b:
        ctx.free(d)
    ctx.close()


if __name__ == "__main__":
    main()
