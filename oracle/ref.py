"""ctypes binding of oracle/liblutheref.so (the C restatement in luthe_ref.c).

Test infrastructure only: see oracle/__init__.py.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

P = 0xFFFFFFFF00000001
CUTS_N2048 = (5, 6)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liblutheref.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblutheref.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        u64, u32, sz = C.c_uint64, C.c_uint32, C.c_size_t
        for name, res, args in [
            ("ref_gl_reduce128", u64, [u64, u64]),
            ("ref_gl_add", u64, [u64, u64]),
            ("ref_gl_sub", u64, [u64, u64]),
            ("ref_gl_mul", u64, [u64, u64]),
            ("ref_gl_pow", u64, [u64, u64]),
            ("ref_gl_shl", u64, [u64, u32]),
            ("ref_power_of_rou", u64, [u64, C.c_int64]),
            ("ref_ntt_sage", C.c_int, [u32, u32p, u32, u64p, u64p]),
            ("ref_intt_sage", C.c_int, [u32, u32p, u32, u64p, u64p]),
            ("ref_ntt", C.c_int, [u32, u32p, u32, u64p, u64p]),
            ("ref_intt", C.c_int, [u32, u32p, u32, u64p, u64p]),
            ("ref_ntt_phi", C.c_int, [u32, u32p, u32, C.c_int, u32, u64p]),
            ("ref_ntt_polymul_batch", C.c_int, [u32, u32p, u32, sz, u64p, u64p, u64p, C.c_int]),
            ("ref_negacyclic_mul", None, [u32, u64p, u64p, u64p]),
            ("ref_decompose", None, [u64, u32, u32, u32, i32p]),
            ("ref_decompose_batch", None, [u64p, sz, u32, u32, u32, i32p]),
            ("ref_modsw_p_to_2_64", u64, [u64]),
            ("ref_ks_body_switch", u32, [u64, u32]),
            ("ref_ks_to_2n", u32, [u32, u32, u32]),
            ("ref_monomial_mult", None, [u32, u64p, u32, u64p]),
            ("ref_sample_extract", None, [u32, u32, u64p, u32, u32, u64p]),
            ("ref_alu", C.c_int, [u32, sz, u64p, u64p, u32, u32, u32, u64p]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class RefParams(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "lwe_dim", "glwe_k", "poly_size", "pbs_base_log", "pbs_level",
        "ks_base_log", "ks_level", "ks_mod_log")]


def params(lwe_dim=724, glwe_k=1, poly_size=2048, pbs_base_log=23, pbs_level=1,
           ks_base_log=3, ks_level=5, ks_mod_log=21):
    return RefParams(lwe_dim, glwe_k, poly_size, pbs_base_log, pbs_level,
                     ks_base_log, ks_level, ks_mod_log)


def _cuts(cuts):
    return np.ascontiguousarray(cuts, dtype=np.uint32)


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def ntt(p, cuts=CUTS_N2048):
    p = _u64(p)
    out = np.empty_like(p)
    c = _cuts(cuts)
    assert lib().ref_ntt(len(p), c, len(c), p, out) == 0
    return out


def intt(x, cuts=CUTS_N2048):
    x = _u64(x)
    out = np.empty_like(x)
    c = _cuts(cuts)
    assert lib().ref_intt(len(x), c, len(c), x, out) == 0
    return out


def ntt_sage(x_rev, cuts):
    x = _u64(x_rev)
    out = np.empty_like(x)
    c = _cuts(cuts)
    assert lib().ref_ntt_sage(len(x), c, len(c), x, out) == 0
    return out


def intt_sage(x, cuts):
    x = _u64(x)
    out = np.empty_like(x)
    c = _cuts(cuts)
    assert lib().ref_intt_sage(len(x), c, len(c), x, out) == 0
    return out


def ntt_phi(N, cuts, bwd, idx):
    out = np.empty(N, np.uint64)
    c = _cuts(cuts)
    assert lib().ref_ntt_phi(N, c, len(c), int(bwd), idx, out) == 0
    return out


def ntt_polymul_batch(a, bn, cuts=CUTS_N2048, threads=1):
    a = _u64(a)
    bn = _u64(bn)
    nb, N = a.shape
    out = np.empty_like(a)
    c = _cuts(cuts)
    assert lib().ref_ntt_polymul_batch(N, c, len(c), nb, a, bn, out, threads) == 0
    return out


def negacyclic_mul(a, b):
    a = _u64(a)
    b = _u64(b)
    out = np.empty_like(a)
    lib().ref_negacyclic_mul(len(a), a, b, out)
    return out


def decompose(x, w, base_log, level):
    d = np.zeros(level, np.int32)
    lib().ref_decompose(int(x), w, base_log, level, d)
    return [int(v) for v in d]


def decompose_batch(x, w, base_log, level):
    x = _u64(x)
    out = np.zeros((x.size, level), np.int32)
    lib().ref_decompose_batch(x, x.size, w, base_log, level, out)
    return out


def modsw(e):
    return int(lib().ref_modsw_p_to_2_64(int(e)))


def monomial_mult(p, e):
    p = _u64(p)
    out = np.empty_like(p)
    lib().ref_monomial_mult(len(p), p, int(e), out)
    return out


def sample_extract(k, N, acc, b2n, log_lut_nb):
    acc = _u64(acc)
    out = np.empty(((1 << log_lut_nb), k * N + 1), np.uint64)
    lib().ref_sample_extract(k, N, acc, int(b2n), log_lut_nb, out)
    return out


def alu(opcode, a0, a1, mul_factor=0, msg_cst=0, useful_bits=5):
    """useful_bits = message + carry bits + 1 padding bit (5 for 2+2-bit blocks)."""
    a0 = _u64(a0)
    a1 = _u64(a1 if a1 is not None else np.zeros_like(a0))
    out = np.empty_like(a0)
    assert lib().ref_alu(opcode, len(a0), a0, a1, mul_factor, msg_cst, useful_bits, out) == 0
    return out


def _prm_ptr(prm):
    return C.byref(prm)


def keyswitch(prm, blwe, ksk):
    L = lib()
    f = L.ref_keyswitch
    f.restype = None
    n = prm.lwe_dim
    out = np.empty(n + 1, np.uint32)
    out2n = np.empty(n + 1, np.uint16)
    f(_prm_ptr(prm), _u64(blwe).ctypes.data_as(C.c_void_p),
      np.ascontiguousarray(ksk, np.uint32).ctypes.data_as(C.c_void_p),
      out.ctypes.data_as(C.c_void_p), out2n.ctypes.data_as(C.c_void_p))
    return out, out2n


def blind_rotate(prm, bsk_ntt, lut, lwe_2n, cuts=CUTS_N2048):
    L = lib()
    f = L.ref_blind_rotate
    f.restype = None
    c = _cuts(cuts)
    acc = np.empty((prm.glwe_k + 1) * prm.poly_size, np.uint64)
    f(_prm_ptr(prm), c.ctypes.data_as(C.c_void_p), C.c_uint32(len(c)),
      _u64(bsk_ntt).ctypes.data_as(C.c_void_p), _u64(lut).ctypes.data_as(C.c_void_p),
      np.ascontiguousarray(lwe_2n, np.uint16).ctypes.data_as(C.c_void_p),
      acc.ctypes.data_as(C.c_void_p))
    return acc


def pbs_batch(prm, bsk_ntt, ksk, luts, lut_index, blwe_in, threads=1, cuts=CUTS_N2048):
    L = lib()
    f = L.ref_pbs_batch
    f.restype = None
    c = _cuts(cuts)
    blwe_in = _u64(blwe_in)
    nct = blwe_in.shape[0]
    out = np.empty_like(blwe_in)
    f(_prm_ptr(prm), c.ctypes.data_as(C.c_void_p), C.c_uint32(len(c)),
      _u64(bsk_ntt).ctypes.data_as(C.c_void_p),
      np.ascontiguousarray(ksk, np.uint32).ctypes.data_as(C.c_void_p),
      _u64(luts).ctypes.data_as(C.c_void_p),
      np.ascontiguousarray(lut_index, np.uint32).ctypes.data_as(C.c_void_p),
      C.c_size_t(nct), blwe_in.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
      C.c_int(threads))
    return out


def bsk_to_ntt(prm, bsk_coef, threads=1, cuts=CUTS_N2048):
    L = lib()
    f = L.ref_bsk_to_ntt
    f.restype = None
    c = _cuts(cuts)
    bsk_coef = _u64(bsk_coef)
    out = np.empty_like(bsk_coef)
    f(_prm_ptr(prm), c.ctypes.data_as(C.c_void_p), C.c_uint32(len(c)),
      bsk_coef.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), C.c_int(threads))
    return out
