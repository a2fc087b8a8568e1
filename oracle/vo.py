"""ctypes binding of the CPU restatement (liboracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

u64p = C.POINTER(C.c_uint64)
i64p = C.POINTER(C.c_int64)


class Ring(C.Structure):
    _fields_ = [("n", C.c_uint32), ("nl", C.c_uint32), ("q", C.c_uint64 * 2)]

    @staticmethod
    def make(n, *qs):
        r = Ring()
        r.n = n
        r.nl = len(qs)
        r.q[0] = qs[0]
        r.q[1] = qs[1] if len(qs) > 1 else 0
        return r

    @property
    def big_q(self):
        return self.q[0] * (self.q[1] if self.nl == 2 else 1)

    @property
    def pw(self):
        return self.n * self.nl


class Dist(C.Structure):
    _fields_ = [("kind", C.c_int), ("bound", C.c_uint32)]


TERNARY = Dist(0, 0)


class Params(C.Structure):
    _fields_ = [
        ("n1", C.c_uint32), ("n2", C.c_uint32), ("n_rec", C.c_uint32),
        ("q1_nl", C.c_uint32), ("q1", C.c_uint64 * 2),
        ("q2", C.c_uint64), ("q3", C.c_uint64), ("q4", C.c_uint64), ("p", C.c_uint64),
        ("b1", C.c_uint64), ("b2", C.c_uint64), ("b_rsk", C.c_uint64), ("ck_base", C.c_uint64),
        ("l_query", C.c_uint32), ("l_ck", C.c_uint32), ("l_rsk", C.c_uint32),
        ("num_rows", C.c_uint32), ("num_cols", C.c_uint32),
    ]


def make_params(n1, n2, q1, q2, q3, q4, p, b1, b2, b_rsk, ck_base, l_query, l_ck, l_rsk,
                num_rows, num_cols, n_rec=None):
    pp = Params()
    pp.n1, pp.n2 = n1, n2
    pp.n_rec = n_rec if n_rec is not None else n2
    q1 = list(q1) if isinstance(q1, (list, tuple)) else [q1]
    pp.q1_nl = len(q1)
    pp.q1[0] = q1[0]
    pp.q1[1] = q1[1] if len(q1) > 1 else 0
    pp.q2, pp.q3, pp.q4, pp.p = q2, q3, q4, p
    pp.b1, pp.b2, pp.b_rsk, pp.ck_base = b1, b2, b_rsk, ck_base
    pp.l_query, pp.l_ck, pp.l_rsk = l_query, l_ck, l_rsk
    pp.num_rows, pp.num_cols = num_rows, num_cols
    return pp


# Paper VIA-C stack (client_server_e2e_paper.rs:205-277, presets.rs:232-258).
PAPER_Q1 = (137438822401, 274810798081)
PAPER_Q2 = 17175674881
PAPER_Q3 = 8380417


def paper_params(num_rows, num_cols):
    return make_params(2048, 512, PAPER_Q1, PAPER_Q2, PAPER_Q3, 4096, 16, 55879, 81, 8, 18,
                       2, 18, 8, num_rows, num_cols)


def toy_params(num_rows=2, num_cols=2):
    """VIA-C toy n8 stack of client_server_e2e.rs:30-54."""
    return make_params(8, 4, 1 << 36, 1 << 28, 1 << 20, 1 << 12, 16, 64, 64, 8, 64, 7, 7, 8,
                       num_rows, num_cols)


def paper_b_params(num_rows, num_cols, T=8):
    """Paper VIA-B stack (batch_e2e_paper.rs:38-100): the VIA-C paper moduli with
    record degree n3 = n2 / T (n3 = 64 at T = 8)."""
    return make_params(2048, 512, PAPER_Q1, PAPER_Q2, PAPER_Q3, 4096, 16, 55879, 81, 8, 18,
                       2, 18, 8, num_rows, num_cols, n_rec=512 // T)


def toy_b_params(num_rows=2, num_cols=2):
    """VIA-B toy n64 stack of batch_e2e_toy.rs:40-68 (n3 = 2, T = 8)."""
    return make_params(64, 16, 1 << 36, 1 << 28, 1 << 20, 1 << 12, 16, 64, 64, 8, 8, 7, 12, 8,
                       num_rows, num_cols, n_rec=2)


def ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.uint64:
        return a.ctypes.data_as(u64p)
    if a.dtype == np.int64:
        return a.ctypes.data_as(i64p)
    raise TypeError(a.dtype)


class Oracle:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
        L = C.CDLL(path)
        self.L = L
        R = C.POINTER(Ring)
        P = C.POINTER(Params)
        sig = {
            "vo_prg_new": (C.c_void_p, [C.c_char_p, C.c_size_t]),
            "vo_prg_free": (None, [C.c_void_p]),
            "vo_prg_fill": (None, [C.c_void_p, C.c_char_p, C.c_size_t]),
            "vo_prg_uniform_below": (C.c_uint64, [C.c_void_p, C.c_uint64]),
            "vo_sample": (None, [C.c_void_p, Dist, i64p, C.c_size_t]),
            "vo_ntt_friendly": (C.c_int, [R]),
            "vo_ntt_psi": (C.c_uint64, [C.c_uint64, C.c_uint32]),
            "vo_ntt_fwd": (None, [R, u64p]),
            "vo_ntt_inv": (None, [R, u64p]),
            "vo_poly_mul": (None, [R, u64p, u64p, u64p]),
            "vo_poly_random": (None, [R, C.c_void_p, u64p]),
            "vo_poly_from_centered": (None, [R, i64p, u64p]),
            "vo_mul_x_pow": (None, [R, u64p, C.c_size_t, u64p]),
            "vo_keygen": (None, [R, Dist, C.c_void_p, u64p]),
            "vo_encrypt": (None, [R, u64p, u64p, Dist, C.c_void_p, u64p]),
            "vo_encode": (None, [R, C.c_uint64, u64p, u64p]),
            "vo_decode": (None, [R, C.c_uint64, u64p, u64p]),
            "vo_gadget_vector": (None, [R, C.c_uint64, C.c_uint32, u64p]),
            "vo_encrypt_rlev": (None, [R, u64p, u64p, C.c_uint64, C.c_uint32, Dist, C.c_void_p, u64p]),
            "vo_encrypt_rgsw": (None, [R, u64p, u64p, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                       Dist, C.c_void_p, u64p]),
            "vo_gadget_decompose": (None, [R, u64p, C.c_uint64, C.c_uint32, i64p]),
            "vo_gadget_product": (None, [R, u64p, C.c_uint32, C.c_uint64, u64p, u64p]),
            "vo_external_product": (None, [R, u64p, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64,
                                           u64p, u64p]),
            "vo_key_switch": (None, [R, u64p, C.c_uint32, C.c_uint64, u64p, u64p]),
            "vo_rotate": (None, [R, u64p, C.c_size_t, u64p]),
            "vo_cmux": (None, [R, u64p, C.c_uint32, C.c_uint64, u64p, u64p, u64p]),
            "vo_dmux": (None, [R, u64p, C.c_uint32, C.c_uint64, u64p, u64p, u64p]),
            "vo_cmux_tree": (None, [R, u64p, C.c_uint32, C.c_uint32, C.c_uint64, u64p, u64p]),
            "vo_dmux_tree": (None, [R, u64p, C.c_uint32, C.c_uint32, C.c_uint64, u64p, u64p]),
            "vo_crot": (None, [R, C.c_int, u64p, C.c_uint32, C.c_uint32, C.c_uint64, u64p, u64p]),
            "vo_rlwe_to_rgsw": (None, [R, u64p, u64p, u64p, C.c_uint64, C.c_uint32, C.c_uint32, u64p]),
            "vo_mod_switch_sym": (None, [R, u64p, R, u64p]),
            "vo_mod_switch_asym": (None, [R, u64p, C.c_uint64, C.c_uint64, u64p, u64p]),
            "vo_mod_switch_rgsw": (None, [R, u64p, C.c_uint32, C.c_uint32, R, u64p]),
            "vo_rekey": (None, [R, u64p, R, u64p]),
            "vo_gen_rsk": (None, [R, u64p, R, u64p, C.c_uint64, C.c_uint32, Dist, C.c_void_p, u64p]),
            "vo_ring_switch": (None, [R, u64p, R, u64p, C.c_uint64, C.c_uint32, u64p]),
            "vo_decrypt_asym": (None, [R, u64p, u64p, u64p, C.c_uint64, C.c_uint64, u64p]),
            "vo_encrypt_lwe": (None, [R, u64p, C.c_uint64, C.c_uint64, Dist, C.c_void_p, u64p]),
            "vo_encrypt_lwe_raw": (None, [R, u64p, C.c_uint64, C.c_uint64, Dist, C.c_void_p, u64p]),
            "vo_cascade_key_words": (C.c_size_t, [R, C.c_uint32]),
            "vo_gen_cascade_key": (None, [R, u64p, C.c_uint64, C.c_uint32, Dist, C.c_void_p, u64p]),
            "vo_conv_step": (None, [R, C.c_uint32, u64p, u64p, u64p, C.c_uint64, C.c_uint32, u64p, u64p]),
            "vo_cascade": (None, [R, u64p, u64p, C.c_uint64, C.c_uint32, u64p]),
            "vo_words_cascade_key": (C.c_size_t, [P]),
            "vo_words_conv_key": (C.c_size_t, [P]),
            "vo_words_rsk": (C.c_size_t, [P]),
            "vo_words_lwe": (C.c_size_t, [P]),
            "vo_query_lwes": (C.c_uint32, [P]),
            "vo_client_setup": (C.c_int, [P, Dist, Dist, Dist, C.c_void_p, u64p, u64p, u64p, u64p, u64p]),
            "vo_client_query": (C.c_int, [P, u64p, C.c_size_t, Dist, C.c_void_p, u64p]),
            "vo_client_recover": (None, [P, u64p, u64p, u64p, u64p]),
            "vo_server_new": (C.c_int, [P, u64p, C.c_size_t, u64p, u64p, u64p, C.POINTER(C.c_void_p)]),
            "vo_server_free": (None, [C.c_void_p]),
            "vo_server_answer": (C.c_int, [C.c_void_p, u64p, C.c_size_t, u64p, u64p,
                                           C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
            "vo_server_step_times": (None, [C.c_void_p, C.POINTER(C.c_double)]),
            "vo_server_answer_stages": (C.c_int, [C.c_void_p, u64p, C.c_size_t, u64p, u64p, u64p, u64p,
                                                  u64p, u64p, u64p, u64p]),
            "vo_server_db": (u64p, [C.c_void_p]),
            "vo_setup_db": (None, [P, u64p, C.c_size_t, u64p]),
            "vo_first_dim": (None, [P, u64p, u64p, C.c_int, u64p]),
            "vo_round_mul_div": (None, [C.c_uint64] * 6 + [u64p, u64p]),
            "vo_server_answer_batch": (C.c_int, [C.c_void_p, C.c_uint32, u64p, C.c_size_t, u64p, u64p,
                                                 C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
            "vo_extr": (None, [R, u64p, C.c_uint32, u64p, u64p]),
            "vo_repack_key_words": (C.c_size_t, [C.c_uint32, C.c_uint32, C.c_uint32]),
            "vo_repack_key_from_cascade": (None, [R, u64p, C.c_uint64, C.c_uint32, C.c_uint32, u64p]),
            "vo_repack": (C.c_int, [R, C.c_uint32, C.c_uint32, u64p, u64p, C.c_uint64, C.c_uint32, u64p]),
            "vo_host_threads": (C.c_int, []),
            "vo_set_decomp_threads": (None, [C.c_int]),
            "vo_embed_at": (None, [C.c_uint32, C.c_uint32, C.c_uint32, u64p, C.c_size_t, u64p]),
            "vo_project_at": (None, [C.c_uint32, C.c_uint32, C.c_uint32, u64p, C.c_size_t, u64p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args

    def __getattr__(self, name):
        return getattr(self.L, "vo_" + name)

    # -- convenience --------------------------------------------------------
    def prg(self, seed):
        return Prg(self, seed)


class Prg:
    def __init__(self, o, seed):
        self.o = o
        self.h = o.L.vo_prg_new(seed, len(seed))

    def __del__(self):
        try:
            self.o.L.vo_prg_free(self.h)
        except Exception:
            pass

    def fill(self, n):
        buf = C.create_string_buffer(n)
        self.o.L.vo_prg_fill(self.h, buf, n)
        return buf.raw

    def uniform_below(self, b):
        return self.o.L.vo_prg_uniform_below(self.h, b)


def zeros(n):
    return np.zeros(n, dtype=np.uint64)
