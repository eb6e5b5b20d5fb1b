"""The HIP kernels' per-thread code (ntt2048.h, pbs_core.h, gl64.h), executed by a
sequential CPU emulation of the 256-thread workgroup (tests/emu/kernel_emu.cpp),
against the oracle.  Checks index mappings, LDS layouts, twiddles and arithmetic of
the device code without a GPU."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import ref

HERE = os.path.dirname(os.path.abspath(__file__))
P = ref.P
rng = np.random.default_rng(777)


@pytest.fixture(scope="module")
def emu():
    import subprocess
    # make rebuilds the library when the kernel headers or the emulation changed
    subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "emu")])
    path = os.path.join(HERE, "emu", "libkernel_emu.so")
    E = C.CDLL(path)
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
    E.emu_ntt_fwd.argtypes = [u64p, u64p]
    E.emu_ntt_inv.argtypes = [u64p, u64p]
    for f in ("emu_gl_add", "emu_gl_sub", "emu_gl_mul", "emu_gl_reduce128", "emu_gl_mul_pow2", "emu_modswitch"):
        getattr(E, f).restype = C.c_uint64
    E.emu_gl_mul_pow2.argtypes = [C.c_uint64, C.c_uint32]
    E.emu_modswitch.argtypes = [C.c_uint64]
    E.emu_gl_reduce160.restype = C.c_uint64
    E.emu_gl_reduce160.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32]
    for f in ("emu_gl_add", "emu_gl_add_wrap", "emu_gl_sub", "emu_gl_mul", "emu_gl_reduce128"):
        getattr(E, f).argtypes = [C.c_uint64, C.c_uint64]
    E.emu_gl_add_wrap.restype = C.c_uint64
    E.emu_gl_canonical_fold.restype = C.c_uint64
    E.emu_gl_canonical_fold.argtypes = [C.c_uint64]
    return E


def test_device_arithmetic(emu):
    L = ref.lib()
    canon = [0, 1, 2, P - 1, P - 2, 2**32, 2**32 - 1, 2**63] + [int(v) for v in rng.integers(0, P, 3000, dtype=np.uint64)]
    anyv = [2**64 - 1, 2**64 - 2, P, P + 5] + canon
    for a, b in zip(canon[:1500], canon[1500:3000]):
        assert emu.emu_gl_add(a, b) == (a + b) % P
        assert emu.emu_gl_sub(a, b) == (a - b) % P
    # the wrap-only add: congruent, below 2^64, equal to the plain sum when it does not wrap,
    # and non-canonical exactly when the sum lies in [p, 2^64)
    edge = [(P - 1, P - 1), (P - 1, 1), (P - 1, 2**32), (2**63, 2**63), (2**63, 2**63 - 2**32),
            (P - 2**32, 2**32 - 1), (0, 0), (P - 1, 0)]
    for a, b in edge + list(zip(canon[:1500], canon[1500:3000])):
        r = emu.emu_gl_add_wrap(a, b)
        assert r % P == (a + b) % P and r < 2**64
        assert r == (a + b if a + b < 2**64 else a + b - P)
    assert emu.emu_gl_add_wrap(P - 1, 5) == P + 4  # in [p, 2^64): left as is
    for x in anyv + [P + 2**32 - 2, 2**64 - 2**32]:
        assert emu.emu_gl_canonical_fold(x) == x % P
    for a, b in zip(anyv[:1500], anyv[1500:3000]):
        assert emu.emu_gl_mul(a, b) == (a * b) % P
        assert emu.emu_gl_reduce128(a, b) == (a + (b << 64)) % P
    for i, (a, b) in enumerate(zip(anyv[:1500], anyv[1500:3000])):
        c = [0, 1, 2, 3, 7, 2**31, 2**32 - 2][i % 7]
        assert emu.emu_gl_reduce160(a, b, c) == (a + (b << 64) + (c << 128)) % P
    for x in anyv[:400]:
        for s in (0, 1, 3, 12, 24, 31, 32, 33, 48, 63, 64, 65, 72, 84, 93, 95, 96, 120, 144, 168, 191):
            assert emu.emu_gl_mul_pow2(x, s) == (x << s) % P
    limbs = [0, 1, 2**31 - 1, 2**31, 2**31 + 1, 2**31 + 2, 2**32 - 1]
    edges = [h * 2**32 + l for h in limbs + [2**31 + 2, 2**32 - 2, 2**32 - 1] for l in limbs]
    for e in canon + [e for e in edges if e < P]:
        assert emu.emu_modswitch(e) == ref.modsw(e)
    # compile-time shifts, every amount in [-192, 192): negated forms, the 2^96-free
    # reduction (amounts <= 32) and the swapped subtraction (amounts in [64, 96) + 96)
    out = np.zeros(384, np.uint64)
    emu.emu_gl_mul_pow2c_all.argtypes = [C.c_uint64, C.c_void_p]
    emu.emu_gl_mul_pow2c_n_all.argtypes = [C.c_uint64, C.c_void_p]
    for x in canon[:300]:
        exp = [(x * pow(2, s % 192, P)) % P for s in range(-192, 192)]
        emu.emu_gl_mul_pow2c_all(x, out.ctypes.data_as(C.c_void_p))
        assert [int(v) for v in out] == exp, x
        emu.emu_gl_mul_pow2c_n_all(x, out.ctypes.data_as(C.c_void_p))
        assert [int(v) for v in out] == exp, x
    # operands in [p, 2^64) (the DFTs' wrap-only sums): every amount with a nonzero shift part
    # (S mod 96 != 0) takes them; amounts that are 0 or a pure negation need canonical input
    for x in [P, P + 1, P + 5, 2**64 - 1, 2**64 - 2**32] + [int(v) for v in rng.integers(P, 2**64, 20, dtype=np.uint64)]:
        emu.emu_gl_mul_pow2c_n_all(x, out.ctypes.data_as(C.c_void_p))
        for i, s in enumerate(range(-192, 192)):
            if s % 96:
                assert int(out[i]) == (x * pow(2, s % 192, P)) % P, (x, s)


def test_kernel_ntt_matches_oracle(emu):
    for trial in range(4):
        a = rng.integers(0, P, 2048, dtype=np.uint64)
        if trial == 1:
            a[:] = P - 1
        if trial == 2:
            a[:] = 0
            a[3] = 1
        f = np.empty_like(a)
        emu.emu_ntt_fwd(a, f)
        assert np.array_equal(f, ref.ntt(a))
        back = np.empty_like(a)
        emu.emu_ntt_inv(f, back)
        assert np.array_equal(back, a)
        x = rng.integers(0, P, 2048, dtype=np.uint64)
        xi = np.empty_like(x)
        emu.emu_ntt_inv(x, xi)
        assert np.array_equal(xi, ref.intt(x))


@pytest.mark.parametrize("L,base_log", [(1, 23), (2, 15)])
def test_kernel_blind_rotation_matches_oracle(emu, L, base_log):
    n, K1, N = 5, 2, 2048
    bsk = rng.integers(0, P, (n, K1, L, K1, N), dtype=np.uint64)
    lut = rng.integers(0, 2**64, (K1, N), dtype=np.uint64)
    a2n = rng.integers(0, 4096, n + 1, dtype=np.uint16)
    a2n[:4] = [0, 2047, 2048, 4095]
    prm = ref.params(lwe_dim=n, pbs_base_log=base_log, pbs_level=L)
    exp = ref.blind_rotate(prm, ref.bsk_to_ntt(prm, bsk.reshape(-1)), lut.reshape(-1), a2n)
    bint = np.empty(n * 8 * K1 * L * K1 * 256, np.uint64)
    vp = C.c_void_p
    emu.emu_bsk_to_internal(C.c_int(n), C.c_int(L), np.ascontiguousarray(bsk.reshape(-1)).ctypes.data_as(vp),
                            bint.ctypes.data_as(vp))
    acc = np.empty(K1 * N, np.uint64)
    assert emu.emu_blind_rotate(C.c_int(n), C.c_int(L), bint.ctypes.data_as(vp),
                                np.ascontiguousarray(lut.reshape(-1)).ctypes.data_as(vp),
                                a2n.ctypes.data_as(vp), C.c_uint32(base_log), acc.ctypes.data_as(vp)) == 0
    assert np.array_equal(acc, exp)


def test_ntt_512_thread_passes(emu):
    """The 512-thread, four-values-per-thread NTT (ntt2048x4.h, the split blind rotation's
    transform) equals the 256-thread one (itself pinned to the oracle above) word for word,
    forward (from plain and from pretwisted coefficients) and inverse."""
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
    emu.emu_ntt4_fwd.argtypes = [u64p, u64p, C.c_int]
    emu.emu_ntt4_inv.argtypes = [u64p, u64p]
    for trial in range(4):
        a = rng.integers(0, P, 2048, dtype=np.uint64)
        if trial == 1:
            a[:] = P - 1
        if trial == 2:
            a[:] = 0
            a[7] = 1
        f, f4, f4p = np.empty_like(a), np.empty_like(a), np.empty_like(a)
        emu.emu_ntt_fwd(a, f)
        emu.emu_ntt4_fwd(a, f4, 0)
        emu.emu_ntt4_fwd(a, f4p, 1)
        assert np.array_equal(f4, f) and np.array_equal(f4p, f)
        b, b4 = np.empty_like(a), np.empty_like(a)
        emu.emu_ntt_inv(f, b)
        emu.emu_ntt4_inv(f, b4)
        assert np.array_equal(b4, b) and np.array_equal(b4, a)


@pytest.mark.parametrize("L,base_log", [(1, 23), (2, 15), (2, 20)])
def test_split_kernel_blind_rotation_matches_oracle(emu, L, base_log):
    """The 512-thread split kernel's CMUX (rotation and decomposition of the four-value layout,
    the six-pass transforms, the key read as thread t & 255 rows (t >> 8) + 2j) against the
    oracle's blind rotation; 2^20 x 2 takes the generic decomposition and unpretwisted lift."""
    n, K1, N = 4, 2, 2048
    bsk = rng.integers(0, P, (n, K1, L, K1, N), dtype=np.uint64)
    lut = rng.integers(0, 2**64, (K1, N), dtype=np.uint64)
    a2n = rng.integers(0, 4096, n + 1, dtype=np.uint16)
    a2n[:3] = [2047, 2048, 4095]
    prm = ref.params(lwe_dim=n, pbs_base_log=base_log, pbs_level=L)
    exp = ref.blind_rotate(prm, ref.bsk_to_ntt(prm, bsk.reshape(-1)), lut.reshape(-1), a2n)
    bint = np.empty(n * 8 * K1 * L * K1 * 256, np.uint64)
    vp = C.c_void_p
    emu.emu_bsk_to_internal(C.c_int(n), C.c_int(L), np.ascontiguousarray(bsk.reshape(-1)).ctypes.data_as(vp),
                            bint.ctypes.data_as(vp))
    acc = np.empty(K1 * N, np.uint64)
    assert emu.emu_blind_rotate_split(C.c_int(n), C.c_int(L), bint.ctypes.data_as(vp),
                                 np.ascontiguousarray(lut.reshape(-1)).ctypes.data_as(vp),
                                 a2n.ctypes.data_as(vp), C.c_uint32(base_log), acc.ctypes.data_as(vp)) == 0
    assert np.array_equal(acc, exp)


def test_pretwisted_digit_lift(emu):
    """The blind rotation lifts each digit already multiplied by the forward NTT's pre-twist
    2^(12 m1); every digit |d| <= 2^22 and every m1 against lift-then-shift (itself checked
    against exact integers above)."""
    emu.emu_pretwist_mismatches.restype = C.c_long
    assert emu.emu_pretwist_mismatches() == 0


def test_kernel_decomposition_and_extract(emu):
    vp = C.c_void_p
    emu.emu_decompose.argtypes = [C.c_uint64, C.c_uint32, C.c_int, vp]
    xs = [int(v) for v in rng.integers(0, 2**64, 3000, dtype=np.uint64)] + [0, 2**63, 2**64 - 1, 2**40, 2**41 - 1]
    for Lv, b in [(1, 23), (2, 15), (5, 3), (3, 7), (4, 2)]:
        d = np.zeros(Lv, np.int64)
        for x in xs:
            emu.emu_decompose(x, b, Lv, d.ctypes.data_as(vp))
            assert list(d) == ref.decompose(x, 64, b, Lv)
    # the high-word path (L*b <= 31) against the generic one: every top/round-bit pattern
    # for small bases, and the tie neighbourhoods (closest = 2^(b-1) with and without
    # rounding, all-ones top with the round bit) at the benchmark bases
    emu.emu_decompose_hi.argtypes = [C.c_uint64, C.c_uint32, C.c_int, vp]
    emu.emu_from_signed32.restype = C.c_uint64
    emu.emu_from_signed32.argtypes = [C.c_int32]
    for d in (0, 1, -1, 2**31 - 1, -2**31, 12345, -12345):
        assert emu.emu_from_signed32(d) == d % P
    def hi_cases(Lv, b):
        R = Lv * b
        ofs = 64 - R
        if R <= 9:
            tops = range(1 << R)
        else:
            t = [0, 1, (1 << R) - 1, (1 << (R - 1)), (1 << (R - 1)) - 1, (1 << (R - 1)) + 1]
            t += [(1 << (b - 1)) | (k << b) for k in range(4)] + [((1 << (b - 1)) - 1) | (k << b) for k in range(4)]
            tops = t + [int(v) for v in rng.integers(0, 1 << R, 200)]
        for top in tops:
            for rb in (0, 1):
                for low in (0, 1, (1 << (ofs - 1)) - 1):
                    yield (top << ofs) | (rb << (ofs - 1)) | low
    for Lv, bs in [(1, range(1, 32)), (2, range(1, 16)), (3, range(1, 11)), (4, range(1, 8)), (5, range(1, 7))]:
        for b in bs:
            d64 = np.zeros(Lv, np.int64)
            d32 = np.zeros(Lv, np.int32)
            for x in hi_cases(Lv, b):
                emu.emu_decompose(x, b, Lv, d64.ctypes.data_as(vp))
                assert emu.emu_decompose_hi(x, b, Lv, d32.ctypes.data_as(vp)) == 0
                assert list(d32) == list(d64), (Lv, b, hex(x))
    # exhaustive over every top-bit / round-bit pattern up to 22 bits of closest value
    # (the high-word path's closed-form carry chain against the generic chain)
    emu.emu_decompose_hi_mismatches.restype = C.c_long
    emu.emu_decompose_hi_mismatches.argtypes = [C.c_int, C.c_uint32]
    for Lv, bs in [(2, range(1, 12)), (3, range(1, 8)), (4, range(1, 6)), (5, range(1, 5))]:
        for b in bs:
            assert emu.emu_decompose_hi_mismatches(Lv, b) == 0, (Lv, b)
    # the same structured cases through the generic path against the oracle, at the
    # PBS and key-switch parameter sets
    for Lv, b in [(1, 23), (2, 15), (5, 3), (1, 5), (2, 4)]:
        d64 = np.zeros(Lv, np.int64)
        for x in hi_cases(Lv, b):
            emu.emu_decompose(x, b, Lv, d64.ctypes.data_as(vp))
            assert list(d64) == ref.decompose(x, 64, b, Lv), (Lv, b, hex(x))
    emu.emu_extract_coef.restype = C.c_uint64
    emu.emu_extract_coef.argtypes = [vp, C.c_int, C.c_int, C.c_uint32, C.c_uint32]
    acc = rng.integers(0, 2**64, 2 * 2048, dtype=np.uint64)
    for b in (0, 1, 2047, 2048, 3000, 4095):
        for log_nb in (0, 2, 3):
            exp = ref.sample_extract(1, 2048, acc, b, log_nb)
            for lut in range(1 << log_nb):
                offs = lut * (2048 >> log_nb)
                for m in (0, 1, 100, 2047):
                    assert emu.emu_extract_coef(acc.ctypes.data_as(vp), 0, m, b, offs) == int(exp[lut][m])
                assert emu.emu_extract_coef(acc.ctypes.data_as(vp), 1, 0, b, offs) == int(exp[lut][2048])
