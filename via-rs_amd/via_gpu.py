"""Host-side mirror of the via-rs server interface over the C ABI (include/via_gpu.h).

`Server.setup(config, records)` / `Server.answer(query)` follow
`via_server::Server::{setup, answer}` (crates/via-server/src/answer.rs:486-554);
errors are raised as `ViaError` with the reference's variants
(crates/via-protocol/src/error.rs:20-64). Every call goes to the HIP library
libvia_gpu.so built for gfx950; there is no other backend, and loading fails
loudly when the library is missing.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvia_gpu.so")

u64p = C.POINTER(C.c_uint64)

VIA_OK, VIA_E_DIM, VIA_E_QUERY_LEN, VIA_E_INDEX, VIA_E_NOMEM, VIA_E_INVALID, VIA_E_DEVICE, VIA_E_BATCH_LEN = range(8)
VIA_E_BUFFER_SHORT, VIA_E_FORMAT_TAG, VIA_E_PRG_SEED = 8, 9, 10


class Params(C.Structure):
    """via_params: PIRParams restricted to the server path (params.rs)."""

    _fields_ = [
        ("n1", C.c_uint32), ("n2", C.c_uint32), ("n_rec", C.c_uint32),
        ("q1_nl", C.c_uint32), ("q1", C.c_uint64 * 2),
        ("q2", C.c_uint64), ("q3", C.c_uint64), ("q4", C.c_uint64), ("p", C.c_uint64),
        ("b1", C.c_uint64), ("b2", C.c_uint64), ("b_rsk", C.c_uint64), ("ck_base", C.c_uint64),
        ("l_query", C.c_uint32), ("l_ck", C.c_uint32), ("l_rsk", C.c_uint32),
        ("num_rows", C.c_uint32), ("num_cols", C.c_uint32),
    ]

    @property
    def total_bits(self):
        lg = lambda x: (x - 1).bit_length() if x > 1 else 0
        return lg(self.num_rows) + lg(self.num_cols) + lg(self.n1 // self.n_rec)

    @property
    def query_lwes(self):
        return self.total_bits * self.l_query

    @property
    def lwe_words(self):
        return self.q1_nl * (self.n1 + 1)


def paper_params(num_rows, num_cols):
    """The paper VIA-C stack as the reference's paper e2e runs it
    (client_server_e2e_paper.rs:205-277): n1=2048, n2=512, q1 two-prime RNS,
    q2=17175674881, q3=8380417, q4=2^12, p=16, L_Q=2 (b1=55879, b2=81),
    L_CK=18 (base 18), L_RSK=8 (base 8)."""
    pp = Params()
    pp.n1, pp.n2, pp.n_rec, pp.q1_nl = 2048, 512, 512, 2
    pp.q1[0], pp.q1[1] = 137438822401, 274810798081
    pp.q2, pp.q3, pp.q4, pp.p = 17175674881, 8380417, 4096, 16
    pp.b1, pp.b2, pp.b_rsk, pp.ck_base = 55879, 81, 8, 18
    pp.l_query, pp.l_ck, pp.l_rsk = 2, 18, 8
    pp.num_rows, pp.num_cols = num_rows, num_cols
    return pp


def paper_b_params(num_rows, num_cols, T=8):
    """The paper VIA-B stack of batch_e2e_paper.rs:38-100: the paper VIA-C
    moduli with record degree n3 = n2 / T (n3 = 64 at T = 8)."""
    pp = paper_params(num_rows, num_cols)
    pp.n_rec = 512 // T
    return pp


class Dist(C.Structure):
    """via_dist: Distribution (sampling/distribution.rs:38-60), kind 0
    Ternary, kind 1 BoundedUniform { bound }."""

    _fields_ = [("kind", C.c_int32), ("bound", C.c_uint32)]


TERNARY = Dist(0, 0)


class ErrDetail(C.Structure):
    _fields_ = [("expected", C.c_uint64), ("got", C.c_uint64), ("msg", C.c_char * 128)]


class ServerStats(C.Structure):
    _fields_ = [("answered", C.c_uint64), ("batches", C.c_uint64), ("max_batch_seen", C.c_uint64),
                ("clients", C.c_uint64), ("key_bytes", C.c_uint64), ("pending", C.c_uint64),
                ("records", C.c_uint64), ("answered_b", C.c_uint64),
                ("p50_s", C.c_double), ("p99_s", C.c_double), ("p50_b_s", C.c_double), ("p99_b_s", C.c_double)]


class ViaError(Exception):
    """ViaError::{DimMismatch, QueryLengthMismatch, IndexOutOfRange} and the
    device-side failures of the C ABI."""

    def __init__(self, status, expected=None, got=None, msg=""):
        self.status = status
        self.expected = expected
        self.got = got
        super().__init__(f"via status {status}: {msg}")


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run python -c 'import __graft_entry__ as g; g.build()'")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(Params)
        vp = C.c_void_p
        sig = {
            "via_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
            "via_ctx_destroy": (None, [vp]),
            "via_ctx_step_times": (C.c_int, [vp, C.POINTER(C.c_double)]),
            "via_ctx_synchronize": (C.c_int, [vp]),
            "via_ctx_first_dim_time": (C.c_int, [vp, C.POINTER(C.c_double)]),
            "via_ctx_mark": (C.c_int, [vp, C.c_int]),
            "via_ctx_set_graphs": (C.c_int, [vp, C.c_int]),
            "via_ctx_set_step_timing": (C.c_int, [vp, C.c_int]),
            "via_debug_set_exact_paths": (C.c_int, [C.c_int]),
            "via_ctx_set_tree_mode": (C.c_int, [vp, C.c_int]),
            "via_ctx_tree_tuning": (C.c_int, [vp, C.c_uint32, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "via_debug_set_tree_spin": (C.c_int, [vp, C.c_long]),
            "via_ctx_pass_tuning": (C.c_int, [vp, C.c_uint32, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "via_answer_device_noalloc": (C.c_int, [vp, vp, vp, vp, C.c_size_t, vp, vp, C.POINTER(ErrDetail)]),
            "via_ctx_set_tuning": (C.c_int, [vp, C.c_uint32, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "via_ctx_autotune": (C.c_int, [vp, vp, C.c_uint32, C.POINTER(vp), C.POINTER(vp), C.c_size_t,
                                           C.POINTER(vp), C.POINTER(vp), C.c_int, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]),
            "via_ctx_marked_time": (C.c_int, [vp, C.POINTER(C.c_double)]),
            "via_db_create": (C.c_int, [vp, P, u64p, C.c_size_t, C.POINTER(vp)]),
            "via_db_create_bytes": (C.c_int, [vp, P, C.c_char_p, C.c_size_t, C.POINTER(vp)]),
            "via_db_destroy": (None, [vp]),
            "via_db_read_cells": (C.c_int, [vp, u64p, C.c_size_t]),
            "via_db_update": (C.c_int, [vp, u64p, C.c_size_t, u64p]),
            "via_client_recover": (C.c_int, [P, u64p, u64p, u64p, u64p]),
            "via_prg_create": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(vp)]),
            "via_prg_destroy": (None, [vp]),
            "via_prg_position": (C.c_uint64, [vp]),
            "via_prg_fill": (C.c_int, [vp, vp, C.c_void_p, C.c_size_t]),
            "via_prg_uniform_below": (C.c_int, [vp, vp, C.c_uint64, C.POINTER(C.c_uint64)]),
            "via_client_key_words": (C.c_int, [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                               C.POINTER(C.c_size_t)]),
            "via_client_setup": (C.c_int, [vp, P, Dist, Dist, Dist, vp, u64p, u64p, u64p, u64p, u64p]),
            "via_client_setup_many": (C.c_int, [vp, P, C.c_uint32, Dist, Dist, Dist, C.POINTER(vp), C.POINTER(vp),
                                                C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                                C.POINTER(C.c_int)]),
            "via_client_query": (C.c_int, [vp, P, u64p, C.c_size_t, Dist, vp, u64p]),
            "via_client_deinterleave": (C.c_int, [P, C.c_uint32, u64p, u64p]),
            "via_wire_uncompressed_size": (C.c_size_t, [C.c_size_t]),
            "via_wire_encode_uncompressed": (C.c_int, [u64p, C.c_size_t, C.c_char_p, C.c_size_t]),
            "via_wire_decode_uncompressed": (C.c_int, [C.c_char_p, C.c_size_t, u64p, C.c_size_t,
                                                       C.POINTER(C.c_size_t), C.POINTER(ErrDetail)]),
            "via_wire_decode_prg": (C.c_int, [C.c_char_p, C.c_size_t, C.c_char_p, u64p, C.c_size_t,
                                              C.POINTER(C.c_size_t), C.POINTER(ErrDetail)]),
            "via_prg_regenerate": (C.c_int, [vp, C.c_char_p, C.c_size_t, C.c_uint64, C.c_size_t, u64p]),
            "via_query_from_prg": (C.c_int, [vp, P, C.c_char_p, C.c_size_t, u64p, C.c_size_t, vp]),
            "via_first_dim": (C.c_int, [vp, vp, C.c_uint32, u64p, u64p]),
            "via_keys_create": (C.c_int, [vp, P, u64p, u64p, u64p, C.POINTER(vp)]),
            "via_keys_destroy": (None, [vp]),
            "via_answer": (C.c_int, [vp, vp, vp, u64p, C.c_size_t, u64p, u64p, C.POINTER(ErrDetail)]),
            "via_answer_device": (C.c_int, [vp, vp, vp, vp, C.c_size_t, vp, vp, C.POINTER(ErrDetail)]),
            "via_answer_many": (C.c_int, [vp, vp, C.c_uint32, C.POINTER(vp), C.POINTER(vp), C.c_size_t,
                                           C.POINTER(vp), C.POINTER(vp), C.POINTER(ErrDetail)]),
            "via_answer_many_device": (C.c_int, [vp, vp, C.c_uint32, C.POINTER(vp), C.POINTER(vp), C.c_size_t,
                                                  C.POINTER(vp), C.POINTER(vp), C.POINTER(ErrDetail)]),
            "via_answer_batch": (C.c_int, [vp, vp, vp, C.c_uint32, C.POINTER(vp), C.POINTER(C.c_size_t),
                                           C.c_size_t, u64p, u64p, C.POINTER(ErrDetail)]),
            "via_answer_batch_device": (C.c_int, [vp, vp, vp, C.c_uint32, vp, C.c_size_t, vp, vp,
                                                  C.POINTER(ErrDetail)]),
            "via_debug_stage": (C.c_int, [vp, C.c_int, u64p, C.c_size_t]),
            "via_poly_mul_device": (C.c_int, [vp, C.c_uint32, u64p, C.c_uint32, C.c_size_t, vp, vp, vp, C.c_size_t]),
            "via_ntt": (C.c_int, [vp, C.c_uint32, u64p, C.c_uint32, C.c_size_t, u64p, C.c_size_t, C.c_int]),
            "via_device_alloc": (C.c_int, [vp, C.c_size_t, C.POINTER(vp)]),
            "via_device_free": (C.c_int, [vp, vp]),
            "via_copy_to_device": (C.c_int, [vp, vp, vp, C.c_size_t]),
            "via_copy_to_host": (C.c_int, [vp, vp, vp, C.c_size_t]),
            "via_copy_device": (C.c_int, [vp, vp, vp, C.c_size_t]),
            "via_status_str": (C.c_char_p, [C.c_int]),
            "via_server_create": (C.c_int, [C.c_int, P, vp, vp, C.c_size_t, C.c_uint32, C.c_uint32, C.POINTER(vp)]),
            "via_server_submit_batch": (C.c_int, [vp, C.c_uint64, C.c_uint32, u64p, C.c_size_t,
                                                  C.POINTER(C.c_uint64)]),
            "via_server_destroy": (None, [vp]),
            "via_server_admit": (C.c_int, [vp, C.c_uint64, u64p, u64p, u64p]),
            "via_server_evict": (C.c_int, [vp, C.c_uint64]),
            "via_server_submit": (C.c_int, [vp, C.c_uint64, u64p, C.c_size_t, C.POINTER(C.c_uint64)]),
            "via_server_poll": (C.c_int, [vp, C.c_uint64, C.POINTER(C.c_int)]),
            "via_server_wait": (C.c_int, [vp, C.c_uint64, u64p, u64p, C.POINTER(C.c_double)]),
            "via_server_get_stats": (C.c_int, [vp, C.POINTER(ServerStats)]),
            "via_server_reset_stats": (C.c_int, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status, err=None):
    if status != VIA_OK:
        msg = lib().via_status_str(status).decode()
        if err is not None and status in (VIA_E_QUERY_LEN, VIA_E_BUFFER_SHORT, VIA_E_FORMAT_TAG):
            raise ViaError(status, err.expected, err.got, msg)
        raise ViaError(status, msg=msg)


# ---- wire formats (via-protocol wire/format.rs) ----------------------------------

def wire_encode(coeffs):
    """Uncompressed::serialize (format.rs:107-113): [0x00][count][coeffs]."""
    c = _u64(coeffs)
    n = lib().via_wire_uncompressed_size(c.size)
    buf = C.create_string_buffer(n)
    _check(lib().via_wire_encode_uncompressed(_p(c), c.size, buf, n))
    return buf.raw


def wire_decode(buf):
    """Uncompressed::deserialize (format.rs:115-131); ViaError with the
    reference's BufferTooShort {needed, got} / UnknownFormatTag(tag)."""
    b = bytes(buf)
    err, cnt = ErrDetail(), C.c_size_t()
    _check(lib().via_wire_decode_uncompressed(b, len(b), None, 0, C.byref(cnt), C.byref(err)), err)
    out = np.zeros(cnt.value, dtype=np.uint64)
    _check(lib().via_wire_decode_uncompressed(b, len(b), _p(out), out.size, C.byref(cnt), C.byref(err)), err)
    return out


def wire_decode_prg(buf):
    """PrgCompressed<32>::deserialize (format.rs:163-190): the seed and the
    body coefficients."""
    b = bytes(buf)
    err, cnt = ErrDetail(), C.c_size_t()
    seed = C.create_string_buffer(32)
    _check(lib().via_wire_decode_prg(b, len(b), seed, None, 0, C.byref(cnt), C.byref(err)), err)
    out = np.zeros(cnt.value, dtype=np.uint64)
    _check(lib().via_wire_decode_prg(b, len(b), seed, _p(out), out.size, C.byref(cnt), C.byref(err)), err)
    return seed.raw, out


class Prg:
    """Shake256Prg (sampling/prg.rs) over a seed of at most 64 bytes."""

    def __init__(self, seed):
        h = C.c_void_p()
        _check(lib().via_prg_create(seed, len(seed), C.byref(h)))
        self.h = h

    @property
    def position(self):
        return lib().via_prg_position(self.h)

    def fill(self, ctx, n):
        """fill_bytes: the next n stream bytes as uint8."""
        out = np.zeros(n, np.uint8)
        _check(lib().via_prg_fill(ctx.h, self.h, out.ctypes.data_as(C.c_void_p), n))
        return out

    def uniform_below(self, ctx, bound):
        v = C.c_uint64()
        _check(lib().via_prg_uniform_below(ctx.h, self.h, bound, C.byref(v)))
        return v.value

    def close(self):
        if self.h:
            lib().via_prg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def client_key_words(pp):
    """(cascade, conv, rsk) key sizes in words."""
    a, b, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
    _check(lib().via_client_key_words(C.byref(pp), C.byref(a), C.byref(b), C.byref(c)))
    return a.value, b.value, c.value


def client_setup(ctx, pp, prg, key1=TERNARY, key2=TERNARY, err=TERNARY):
    """Client::setup (client.rs:89-173): dict of sk1, sk2, cascade_key,
    conv_key, rsk drawn from prg."""
    wc, wv, wr = client_key_words(pp)
    out = {"sk1": np.zeros(2 * pp.n1, np.uint64), "sk2": np.zeros(pp.n2, np.uint64),
           "cascade_key": np.zeros(wc, np.uint64), "conv_key": np.zeros(wv, np.uint64),
           "rsk": np.zeros(wr, np.uint64)}
    _check(lib().via_client_setup(ctx.h, C.byref(pp), key1, key2, err, prg.h, _p(out["sk1"]), _p(out["sk2"]),
                                  _p(out["cascade_key"]), _p(out["conv_key"]), _p(out["rsk"])))
    return out


def client_setup_many(ctx, pp, prgs, key1=TERNARY, key2=TERNARY, err=TERNARY):
    """Client::setup for several clients at once (via_client_setup_many):
    one dict per PRG, each as client_setup gives for that stream."""
    wc, wv, wr = client_key_words(pp)
    n = len(prgs)
    outs = [{"sk1": np.zeros(2 * pp.n1, np.uint64), "sk2": np.zeros(pp.n2, np.uint64),
             "cascade_key": np.zeros(wc, np.uint64), "conv_key": np.zeros(wv, np.uint64),
             "rsk": np.zeros(wr, np.uint64)} for _ in range(n)]
    vp = C.c_void_p
    arr = lambda k: (vp * n)(*[o[k].ctypes.data for o in outs])
    G = (vp * n)(*[g.h.value if isinstance(g.h, C.c_void_p) else g.h for g in prgs])
    st = (C.c_int * n)()
    _check(lib().via_client_setup_many(ctx.h, C.byref(pp), n, key1, key2, err, G, arr("sk1"), arr("sk2"),
                                       arr("cascade_key"), arr("conv_key"), arr("rsk"), st))
    return outs


def client_query(ctx, pp, sk1, index, prg, err=TERNARY):
    """Client::query (client.rs:183-207): the query's raw LWEs."""
    out = np.zeros(pp.query_lwes * pp.lwe_words, np.uint64)
    _check(lib().via_client_query(ctx.h, C.byref(pp), _p(_u64(sk1)), index, err, prg.h, _p(out)))
    return out


def client_recover(pp, sk2, mask, body):
    """Client::recover (client.rs:222-239): the n2 plaintext coefficients."""
    out = np.zeros(pp.n2, dtype=np.uint64)
    _check(lib().via_client_recover(C.byref(pp), _p(_u64(sk2)), _p(_u64(mask)), _p(_u64(body)), _p(out)))
    return out


def client_deinterleave(pp, T, plain):
    """deinterleave_batch (batch.rs:31-52): T records of n_rec coefficients."""
    out = np.zeros((T, pp.n_rec), dtype=np.uint64)
    _check(lib().via_client_deinterleave(C.byref(pp), T, _p(_u64(plain)), _p(out)))
    return out


def answer_bytes(mask, body):
    """An answer on the wire: the Uncompressed mask, then the Uncompressed body."""
    return wire_encode(mask) + wire_encode(body)


def _u64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def _p(a):
    return a.ctypes.data_as(u64p)


class Context:
    """A device context (one HIP stream on one GPU)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().via_ctx_create(device, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()

    def adopt(self, child):
        """Databases and keys live on this context and are released before it."""
        self._children.add(child)

    def close(self):
        if self.h:
            for ch in list(self._children):
                ch.close()
            lib().via_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().via_ctx_synchronize(self.h))

    def step_times(self):
        t = (C.c_double * 7)()
        _check(lib().via_ctx_step_times(self.h, t))
        return list(t)

    def first_dim_time(self):
        t = C.c_double()
        _check(lib().via_ctx_first_dim_time(self.h, C.byref(t)))
        return t.value

    def prg_regenerate(self, seed, q, count):
        """PrgCompressed::regenerate_masks (format.rs:143-160) on the GPU."""
        out = np.zeros(count, dtype=np.uint64)
        _check(lib().via_prg_regenerate(self.h, bytes(seed), len(seed), q, count, _p(out)))
        return out

    def query_from_prg(self, pp, seed, bodies):
        """Device LWE array [n_lwe][limb][n1 + 1] of a query whose masks are
        regenerated from seed (limb-0-first RNS draws) and whose bodies
        [n_lwe][limb] are given; the caller frees it with free()."""
        b = _u64(bodies)
        n_lwe = b.size // 2
        d = self.alloc(8 * n_lwe * 2 * (pp.n1 + 1))
        _check(lib().via_query_from_prg(self.h, C.byref(pp), bytes(seed), len(seed), _p(b), n_lwe, d))
        return d

    def set_tuning(self, nq, shapes):
        """via_ctx_set_tuning: pin the twelve (groups, variant) launch shapes
        of batches of nq queries; shapes None erases them."""
        if shapes is None:
            _check(lib().via_ctx_set_tuning(self.h, nq, None, None))
            return
        g = (C.c_int * 12)(*[x[0] for x in shapes])
        v = (C.c_int * 12)(*[x[1] for x in shapes])
        _check(lib().via_ctx_set_tuning(self.h, nq, g, v))

    def autotune(self, db, keys, d_lwes, n_lwe, d_masks, d_bodies, reps=2):
        """via_ctx_autotune: per-step launch shapes for batches of len(keys)
        queries (device buffers); returns the twelve (groups, variant)
        choices (variant 0: 256 threads radix 8, 1: 1024 threads, 2: 256
        threads radix 16, 3: 512 threads radix 8)."""
        nq = len(keys)
        vp = C.c_void_p
        K = (vp * nq)(*[k.h.value for k in keys])
        Qp = (vp * nq)(*[d.value for d in d_lwes])
        M = (vp * nq)(*[d.value for d in d_masks])
        B = (vp * nq)(*[d.value for d in d_bodies])
        g = (C.c_int * 12)()
        w = (C.c_int * 12)()
        _check(lib().via_ctx_autotune(self.h, db.h, nq, K, Qp, n_lwe, M, B, reps, g, w))
        return list(zip(list(g), list(w)))

    def set_tree_mode(self, mode):
        """via_ctx_set_tree_mode: -1 tuned, 0 level by level, 1 persistent."""
        _check(lib().via_ctx_set_tree_mode(self.h, mode))

    def tree_tuning(self, nq):
        d, c = C.c_int(), C.c_int()
        if lib().via_ctx_tree_tuning(self.h, nq, C.byref(d), C.byref(c)) != VIA_OK:
            return None
        g, sp = C.c_int(), C.c_int()
        _check(lib().via_ctx_pass_tuning(self.h, nq, C.byref(g), C.byref(sp)))
        return {"dmux": d.value, "cmux": c.value, "fd_groups": g.value, "gate_split": sp.value}

    def set_step_timing(self, on):
        """Record the step events behind step_times (the default) or only the
        FirstDim pair: each event record idles the device for a few us."""
        _check(lib().via_ctx_set_step_timing(self.h, int(bool(on))))

    def set_graphs(self, on):
        """hipGraph mode: repeated identical answers replay a captured graph."""
        _check(lib().via_ctx_set_graphs(self.h, int(bool(on))))

    def mark(self, which):
        """Record the start (0) or stop (1) event of the stream timer."""
        _check(lib().via_ctx_mark(self.h, which))

    def marked_time(self):
        """Seconds between the last start and stop marks (waits for the stop)."""
        t = C.c_double()
        _check(lib().via_ctx_marked_time(self.h, C.byref(t)))
        return t.value

    def stage(self, k, words):
        out = np.zeros(words, dtype=np.uint64)
        _check(lib().via_debug_stage(self.h, k, _p(out), words))
        return out

    def ntt(self, n, moduli, polys, group=1, inverse=False):
        m = _u64(moduli)
        a = _u64(polys).copy()
        _check(lib().via_ntt(self.h, n, _p(m), len(m), group, _p(a), a.size // n, int(inverse)))
        return a

    # device buffers for callers that keep inputs resident in HBM
    def alloc(self, nbytes):
        d = C.c_void_p()
        _check(lib().via_device_alloc(self.h, nbytes, C.byref(d)))
        return d

    def free(self, d):
        _check(lib().via_device_free(self.h, d))

    def to_device(self, d, arr):
        arr = np.ascontiguousarray(arr)
        _check(lib().via_copy_to_device(self.h, d, arr.ctypes.data_as(C.c_void_p), arr.nbytes))

    def to_host(self, arr, d):
        _check(lib().via_copy_to_host(self.h, arr.ctypes.data_as(C.c_void_p), d, arr.nbytes))
        return arr

    def copy_device(self, dst, src, nbytes):
        _check(lib().via_copy_device(self.h, dst, src, nbytes))

    def poly_mul_device(self, n, moduli, group, d_a, d_b, d_c, count):
        m = _u64(moduli)
        _check(lib().via_poly_mul_device(self.h, n, _p(m), len(m), group, d_a, d_b, d_c, count))


class PreparedDb:
    """PreparedDb (prepared_db.rs): the HBM-resident eval-form database."""

    def __init__(self, ctx, pp, records=None, record_bytes=None, n_records=None):
        h = C.c_void_p()
        if record_bytes is not None:
            b = bytes(record_bytes)
            _check(lib().via_db_create_bytes(ctx.h, C.byref(pp), b, n_records, C.byref(h)))
        else:
            r = _u64(records)
            _check(lib().via_db_create(ctx.h, C.byref(pp), _p(r), r.size // pp.n_rec, C.byref(h)))
        self.h, self.ctx, self.pp = h, ctx, pp
        ctx.adopt(self)

    def first_dim(self, switched, nq=1):
        """first_dim (first_dim.rs:94-184) of nq queries: switched
        [q][I][part][n1] -> [q][J][part][n1], coefficient form at q2."""
        pp = self.pp
        sw = _u64(switched)
        if sw.size != nq * pp.num_rows * 2 * pp.n1:
            raise ValueError("switched holds %d words, expected %d" % (sw.size, nq * pp.num_rows * 2 * pp.n1))
        out = np.zeros(nq * pp.num_cols * 2 * pp.n1, dtype=np.uint64)
        _check(lib().via_first_dim(self.ctx.h, self.h, nq, _p(sw), _p(out)))
        return out

    def update(self, indices, records):
        """via_db_update: replace records in place ([n][n_rec] coefficients)."""
        idx = _u64(indices)
        recs = _u64(records)
        _check(lib().via_db_update(self.h, _p(idx), idx.size, _p(recs)))

    def cells(self):
        words = self.pp.num_rows * self.pp.num_cols * self.pp.n1
        out = np.zeros(words, dtype=np.uint64)
        _check(lib().via_db_read_cells(self.h, _p(out), words))
        return out

    def close(self):
        if self.h:
            lib().via_db_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PreparedKeys:
    """PreparedKeys (prepared_keys.rs:69-77): eval-form cascade, conversion
    and ring-switch keys, resident in HBM."""

    def __init__(self, ctx, pp, cascade_key, conv_key, rsk):
        h = C.c_void_p()
        ck, cv, rs = _u64(cascade_key), _u64(conv_key), _u64(rsk)
        _check(lib().via_keys_create(ctx.h, C.byref(pp), _p(ck), _p(cv), _p(rs), C.byref(h)))
        self.h, self.ctx = h, ctx
        ctx.adopt(self)

    def close(self):
        if self.h:
            lib().via_keys_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServerConfig:
    """ServerConfig (answer.rs:361-400): public parameters and keys."""

    def __init__(self, pp, cascade_key, conv_key, rsk):
        self.pp, self.cascade_key, self.conv_key, self.rsk = pp, cascade_key, conv_key, rsk


class Server:
    """via_server::Server on one MI355X."""

    def __init__(self, ctx, pp, db, keys):
        self.ctx, self.pp, self.db, self.keys = ctx, pp, db, keys

    @classmethod
    def setup(cls, config, records, ctx=None):
        """Server::setup (answer.rs:486-528): records [n][n_rec] in [0, p)."""
        ctx = ctx or Context(0)
        pp = config.pp
        db = PreparedDb(ctx, pp, records=records)
        keys = PreparedKeys(ctx, pp, config.cascade_key, config.conv_key, config.rsk)
        return cls(ctx, pp, db, keys)

    def answer(self, query):
        """Server::answer (answer.rs:534-554). `query` is the flat LWE array
        (n_lwe x q1_nl x (n1+1) words); returns (mask mod q3, body mod q4)."""
        q = _u64(query)
        n_lwe = q.size // self.pp.lwe_words
        am = np.zeros(self.pp.n2, dtype=np.uint64)
        ab = np.zeros(self.pp.n2, dtype=np.uint64)
        err = ErrDetail()
        st = lib().via_answer(self.ctx.h, self.db.h, self.keys.h, _p(q), n_lwe, _p(am), _p(ab), C.byref(err))
        _check(st, err)
        return am, ab

    def answer_batch(self, T, queries):
        """Server::answer_batch::<T> (answer.rs:561-602), VIA-B: `queries` is
        the BatchedQuery's list of T flat LWE arrays; returns one answer
        (mask mod q3, body mod q4) whose degree-n2 plaintext holds record t at
        slots t + (n2/n_rec) k (deinterleave_batch, via-client/src/batch.rs)."""
        qs = [_u64(q) for q in queries]
        nq = len(qs)
        vp = C.c_void_p
        Qp = (vp * max(nq, 1))(*[q.ctypes.data for q in qs])
        NL = (C.c_size_t * max(nq, 1))(*[q.size // self.pp.lwe_words for q in qs])
        am = np.zeros(self.pp.n2, dtype=np.uint64)
        ab = np.zeros(self.pp.n2, dtype=np.uint64)
        err = ErrDetail()
        st = lib().via_answer_batch(self.ctx.h, self.db.h, self.keys.h, T, Qp, NL, nq, _p(am), _p(ab),
                                    C.byref(err))
        _check(st, err)
        return am, ab


def answer_many(ctx, db, keys_list, queries, pp):
    """Answers several queries (each with its own client's PreparedKeys) in
    one batched pass over the pipeline (via_answer_many); each answer equals
    the one Server::answer gives for that query alone."""
    nq = len(queries)
    qs = [_u64(q) for q in queries]
    n_lwe = qs[0].size // pp.lwe_words
    masks = [np.zeros(pp.n2, dtype=np.uint64) for _ in range(nq)]
    bodies = [np.zeros(pp.n2, dtype=np.uint64) for _ in range(nq)]
    vp = C.c_void_p
    K = (vp * nq)(*[k.h.value for k in keys_list])
    Qp = (vp * nq)(*[q.ctypes.data for q in qs])
    M = (vp * nq)(*[m.ctypes.data for m in masks])
    B = (vp * nq)(*[b.ctypes.data for b in bodies])
    err = ErrDetail()
    _check(lib().via_answer_many(ctx.h, db.h, nq, K, Qp, n_lwe, M, B, C.byref(err)), err)
    return masks, bodies


class ServeLoop:
    """Continuous-batching request loop (via_server_*): clients admitted under
    an HBM key budget, queries submitted at any time, batches of up to
    max_batch queries launched on `slots` device contexts as they free up.
    With `db_b` (a database at a VIA-B record degree) the loop also takes
    VIA-B batches (submit_batch), each answered as answer_batch::<T>."""

    def __init__(self, pp, db, key_budget_bytes, max_batch=16, slots=2, device=0, db_b=None):
        h = C.c_void_p()
        _check(lib().via_server_create(device, C.byref(pp), db.h, db_b.h if db_b is not None else None,
                                       key_budget_bytes, max_batch, slots, C.byref(h)))
        self.h, self.pp, self.db, self.db_b = h, pp, db, db_b

    def close(self):
        if self.h:
            lib().via_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, client_id, cascade_key, conv_key, rsk):
        ck, cv, rs = _u64(cascade_key), _u64(conv_key), _u64(rsk)
        _check(lib().via_server_admit(self.h, client_id, _p(ck), _p(cv), _p(rs)))

    def evict(self, client_id):
        _check(lib().via_server_evict(self.h, client_id))

    def submit(self, client_id, query):
        q = _u64(query)
        t = C.c_uint64()
        _check(lib().via_server_submit(self.h, client_id, _p(q), q.size // self.pp.lwe_words, C.byref(t)))
        return t.value

    def submit_batch(self, client_id, T, queries):
        """One VIA-B request: the T queries (flat LWE arrays) of one client."""
        q = np.ascontiguousarray(np.concatenate([_u64(x) for x in queries]))
        t = C.c_uint64()
        n_lwe = q.size // (T * self.pp.lwe_words)
        _check(lib().via_server_submit_batch(self.h, client_id, T, _p(q), n_lwe, C.byref(t)))
        return t.value

    def done(self, ticket):
        d = C.c_int()
        _check(lib().via_server_poll(self.h, ticket, C.byref(d)))
        return bool(d.value)

    def wait(self, ticket):
        m = np.zeros(self.pp.n2, dtype=np.uint64)
        b = np.zeros(self.pp.n2, dtype=np.uint64)
        lat = C.c_double()
        _check(lib().via_server_wait(self.h, ticket, _p(m), _p(b), C.byref(lat)))
        return m, b, lat.value

    def stats(self):
        st = ServerStats()
        _check(lib().via_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def reset_stats(self):
        _check(lib().via_server_reset_stats(self.h))
