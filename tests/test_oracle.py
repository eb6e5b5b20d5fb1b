"""Pins of the CPU oracle (oracle/luthe_ref.c) before it is trusted as the checker.

Every function is checked against something that does not restate its own formula:
the reference's committed phi ROMs and phi package (golden fixtures), the constants
of hw/module/hpu/simu/scripts/run.sh, the decomposition notebook KAT, an independent
formulation of the same algorithm (Carl's org_decompose vs the hardware form), closed
forms (polynomial evaluation, big-integer arithmetic) and brute force on small sizes.
"""
import os

import numpy as np
import pytest

from oracle import ref

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
P = ref.P
rng = np.random.default_rng(12345)


def _gl():
    return ref.lib()


# ---------------------------------------------------------------- field
def test_field_reduction_against_bigint():
    L = _gl()
    vals = [0, 1, P - 1, P, P + 1, 2**64 - 1, 2**32, 2**32 - 1, 2**63] + \
        [int(v) for v in rng.integers(0, 2**64, 2000, dtype=np.uint64)]
    for lo, hi in zip(vals, vals[::-1]):
        assert L.ref_gl_reduce128(lo, hi) == (lo + (hi << 64)) % P
    for a, b in zip(vals[:500], vals[500:1000]):
        assert L.ref_gl_mul(a, b) == (a * b) % P
        assert L.ref_gl_add(a, b) == (a + b) % P
        assert L.ref_gl_sub(a, b) == (a - b) % P
    for k in (0, 1, 3, 31, 32, 63, 64, 95, 96, 97, 150, 191):
        assert L.ref_gl_shl(12345678901234567, k) == (12345678901234567 << k) % P


def test_roots_of_unity_constants():
    """power_of_rou against the (omega_N, omega_2N) pairs hard-coded for the Goldilocks
    gtv flow in hw/module/hpu/simu/scripts/run.sh:429-458, and omega_64 = 8."""
    L = _gl()
    table = {32: (64, 8), 64: (8, 2198989700608), 128: (2198989700608, 14041890976876060974),
             256: (14041890976876060974, 14430643036723656017),
             512: (14430643036723656017, 4440654710286119610),
             1024: (4440654710286119610, 8816101479115663336),
             2048: (8816101479115663336, 10974926054405199669),
             4096: (10974926054405199669, 1206500561358145487)}
    for n, (wn, w2n) in table.items():
        assert L.ref_power_of_rou(n, 1) == wn, n
        assert L.ref_power_of_rou(2 * n, 1) == w2n, n
    assert L.ref_power_of_rou(64, 1) == 8


def test_phi_package_tables():
    """Powers of w_2n against ntt_core_gf64_phi_phi_pkg.sv (fixtures)."""
    L = _gl()
    for n in (2048, 1024, 64, 32, 4):
        f = np.load(os.path.join(GOLDEN, f"pkg_phi_fwd_N{n}.npy"))
        b = np.load(os.path.join(GOLDEN, f"pkg_phi_bwd_N{n}.npy"))
        w = np.load(os.path.join(GOLDEN, f"pkg_phi_bwdwdiv_N{n}.npy"))
        inv_n = pow(n, P - 2, P)
        idx = range(2 * n) if n <= 64 else rng.choice(2 * n, 200, replace=False)
        for p_ in idx:
            p_ = int(p_)
            assert int(f[p_]) == L.ref_power_of_rou(2 * n, p_)
            assert int(b[p_]) == L.ref_power_of_rou(2 * n, -p_)
            assert int(w[p_]) == L.ref_power_of_rou(2 * n, -p_) * inv_n % P


@pytest.mark.parametrize("cuts", [(5, 6), (4, 7)])
def test_phi_roms(cuts):
    """Negacyclic phi lists against the committed ROMs
    hw/memory_file/twiddle/NTT_CORE_ARCH_GF64/R2_PSI{8,16}/twd_phi_{fwd,bwd}_N2048_*.mem."""
    tag = f"{cuts[0]}_{cuts[1]}"
    fwd = np.load(os.path.join(GOLDEN, f"rom_phi_fwd_N2048_cut{tag}.npy"))
    bwd = np.load(os.path.join(GOLDEN, f"rom_phi_bwd_N2048_cut{tag}.npy"))
    assert np.array_equal(ref.ntt_phi(2048, cuts, 0, 0), fwd)
    assert np.array_equal(ref.ntt_phi(2048, cuts, 1, 0), bwd)


# ---------------------------------------------------------------- NTT
def _poly_eval(p, x):
    acc = 0
    for c in reversed([int(v) for v in p]):
        acc = (acc * x + c) % P
    return acc


def test_ntt_output_order_is_evaluation_at_odd_powers():
    """ntt() with cut [5,6]: index 64k + w holds P(zeta^(2(k + 32 w) + 1)), zeta the 4096th
    root -- a closed form that fixes the NTT-domain order used by the BSK and the NTT product."""
    L = _gl()
    zeta = L.ref_power_of_rou(4096, 1)
    p = rng.integers(0, P, 2048, dtype=np.uint64)
    out = ref.ntt(p)
    for idx in [0, 1, 7, 63, 64, 65, 1000, 2047] + list(rng.choice(2048, 8, replace=False)):
        k, w = divmod(int(idx), 64)
        assert int(out[idx]) == _poly_eval(p, pow(zeta, 2 * (k + 32 * w) + 1, P)), idx


@pytest.mark.parametrize("N,cuts", [(64, (3, 3)), (128, (5, 2)), (256, (4, 4)), (256, (2, 3, 3)), (64, (2, 2, 2))])
def test_ntt_convolution_small(N, cuts):
    """intt(ntt(a) . ntt(b)) == a*b mod (X^N + 1, p) by schoolbook (ntt_gf64.sage:515-578)."""
    a = rng.integers(0, P, N, dtype=np.uint64)
    b = rng.integers(0, P, N, dtype=np.uint64)
    c = rng.integers(0, P, N, dtype=np.uint64)
    A, B, Cn = ref.ntt(a, cuts), ref.ntt(b, cuts), ref.ntt(c, cuts)
    L = _gl()
    ab = np.array([L.ref_gl_mul(int(x), int(y)) for x, y in zip(A, B)], dtype=np.uint64)
    abc = np.array([L.ref_gl_mul(int(x), int(y)) for x, y in zip(ab, Cn)], dtype=np.uint64)
    assert np.array_equal(ref.intt(ab, cuts), ref.negacyclic_mul(a, b))
    assert np.array_equal(ref.intt(abc, cuts), ref.negacyclic_mul(ref.negacyclic_mul(a, b), c))


@pytest.mark.parametrize("N,cuts", [(2048, (5, 6)), (1024, (5, 5)), (2048, (4, 4, 3))])
def test_ntt_roundtrip(N, cuts):
    """NTT then INTT is the identity (tb_ntt_core_gf64_without_pp.sv:8-14,382-388)."""
    a = rng.integers(0, P, N, dtype=np.uint64)
    assert np.array_equal(ref.intt(ref.ntt(a, cuts), cuts), a)


def test_ntt_multiply_by_monomial():
    """At N = 2048: a * X^j computed through the NTT equals the negacyclic shift."""
    a = rng.integers(0, P, 2048, dtype=np.uint64)
    for j in (1, 5, 2047):
        xj = np.zeros(2048, np.uint64)
        xj[j] = 1
        A, XJ = ref.ntt(a), ref.ntt(xj)
        prod = np.array([_gl().ref_gl_mul(int(x), int(y)) for x, y in zip(A, XJ)], dtype=np.uint64)
        exp = np.empty(2048, dtype=object)
        for i in range(2048):
            src = i - j
            exp[i] = int(a[src]) if src >= 0 else (P - int(a[src + 2048])) % P
        assert np.array_equal(ref.intt(prod), exp.astype(np.uint64))


def test_ntt_rejects_single_and_oversized_cuts():
    with pytest.raises(AssertionError):
        ref.ntt(np.zeros(32, np.uint64), (5,))
    with pytest.raises(AssertionError):
        ref.ntt(np.zeros(2048, np.uint64), (4, 7))


# ---------------------------------------------------------------- decomposition
def org_decompose(val, base, level, q):
    """Carl's reference formulation (balanced_decomp_unfix_sign_round_bit.py:12-39):
    arithmetic closest value and carry rule, a different algorithm from the hardware
    bit-level form the oracle implements."""
    y = (val * 2 * base**level) // q
    bit = y % 2
    state = (y + 1) // 2
    if state > base**level / 2 or (state == base**level / 2 and bit == 1):
        state = state - base**level
    digits = []
    for _ in range(level):
        res = state % base
        state = (state - res) // base
        if (res > base // 2) or ((res == base // 2) and ((state % base) >= base // 2)):
            state += 1
            res -= base
        digits.append(res)
    return digits


def test_decomposition_notebook_kat():
    """sw/decomposer/Decomposition.ipynb cell 3: 2617242525, q = 2^32, L = 3, B = 2^7
    -> [-50, 0, -2] most significant level first."""
    assert ref.decompose(2617242525, 32, 7, 3) == [-2, 0, -50]


@pytest.mark.parametrize("base_log", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("level", [1, 2, 3])
def test_decomposition_exhaustive_vs_org(base_log, level):
    """The self-check of balanced_decomp_unfix_sign_round_bit.py:143-157: for every
    value below q = 2 * B^L, hardware form == org form (digits, hence reconstruction
    and weight)."""
    base = 1 << base_log
    q = 2 * base**level
    w = base_log * level + 1
    xs = np.arange(q, dtype=np.uint64)
    got = ref.decompose_batch(xs, w, base_log, level)
    for k in range(q):
        assert list(got[k]) == org_decompose(k, base, level, q), (base_log, level, k)


@pytest.mark.parametrize("base_log,level", [(4, 4), (5, 5), (7, 3), (3, 5)])
def test_decomposition_sampled_vs_org(base_log, level):
    base = 1 << base_log
    q = 2 * base**level
    w = base_log * level + 1
    xs = rng.integers(0, q, 4000, dtype=np.uint64)
    got = ref.decompose_batch(xs, w, base_log, level)
    for i, k in enumerate(xs):
        assert list(got[i]) == org_decompose(int(k), base, level, q)


@pytest.mark.parametrize("base_log,level", [(23, 1), (15, 2), (3, 5), (7, 3)])
def test_decomposition_64bit_vs_org(base_log, level):
    """Full 64-bit operands (q = 2^64): org form with the same rounding."""
    xs = np.concatenate([rng.integers(0, 2**64, 3000, dtype=np.uint64),
                         np.array([0, 1, 2**63, 2**63 - 1, 2**64 - 1, 1 << (64 - base_log * level - 1)],
                                  dtype=np.uint64)])
    got = ref.decompose_batch(xs, 64, base_log, level)
    base = 1 << base_log
    for i, k in enumerate(xs):
        d = list(got[i])
        assert d == org_decompose(int(k), base, level, 2**64)
        assert all(abs(v) <= base // 2 for v in d)


# ---------------------------------------------------------------- switches
def test_modswitch_against_bigint():
    """mod_switch_to_2powerN.sv:62-64,132: z = ((e * floor(2^160/p)) >> 96) + bit95."""
    C = (1 << 160) // P
    vals = [0, 1, 2, P - 1, P - 2, 2**63, 2**63 - 1, 2**32, 2**32 - 1] + \
        [int(v) for v in rng.integers(0, P, 20000, dtype=np.uint64)]
    for e in vals:
        t = e * C
        exp = ((t >> 96) + ((t >> 95) & 1)) % 2**64
        z = ref.modsw(e)
        assert z == exp
        # and it is the rounded e * 2^64 / p up to one unit
        assert abs(((e << 64) + P // 2) // P % 2**64 - z) in (0, 1, 2**64 - 1)


def test_ks_switches():
    """pep_ks_blram_format.sv:241 and pep_ks_out_process.sv:336-341, exhaustive on 21 bits."""
    L = _gl()
    v = np.arange(1 << 21, dtype=np.int64)
    exp = ((v >> 9) + ((v >> 8) & 1)) % 4096
    got = np.array([L.ref_ks_to_2n(int(x), 21, 12) for x in v[::97]])
    assert np.array_equal(got, exp[::97])
    for b in [0, 2**64 - 1, 2**42, 2**43 - 1] + [int(x) for x in rng.integers(0, 2**64, 2000, dtype=np.uint64)]:
        assert L.ref_ks_body_switch(b, 21) == ((b >> 43) + ((b >> 42) & 1)) % 2**21


# ---------------------------------------------------------------- rotation / extract / ALU
def monomial_mult_definition(poly, e):
    """monomial_mult.py:156-176: multiply by X, e times."""
    out = [int(v) for v in poly]
    for _ in range(e):
        out = out[-1:] + out[:-1]
        out[0] = (-out[0]) % 2**64
    return out


def test_monomial_mult_all_exponents():
    N = 16
    p = rng.integers(0, 2**64, N, dtype=np.uint64)
    for e in range(2 * N):
        assert [int(v) for v in ref.monomial_mult(p, e)] == monomial_mult_definition(p, e)


def sample_extract_tb(acc, k, N, b, lut_nb):
    """tb_pep_mmacc_sample_extract.sv:646-712 (rotation by b, extraction per LUT offset)."""
    M = 2**64
    outs = []
    v = [[0] * N for _ in range(k + 1)]
    for g in range(k + 1):
        for n in range(N):
            idx = n + b
            sign = N <= idx < 2 * N
            val = int(acc[g * N + idx % N])
            v[g][n] = (M - val) % M if sign else val
    for lut in range(lut_nb):
        offs = lut * N // lut_nb
        blwe = []
        for g in range(k):
            for n in range(N):
                ci = offs - n
                neg = ci < 0
                val = v[g][N + ci] if neg else v[g][ci]
                blwe.append((M - val) % M if neg else val)
        blwe.append(v[k][offs])
        outs.append(blwe)
    return outs


def test_sample_extract_all_rotations():
    N, k = 16, 1
    acc = rng.integers(0, 2**64, (k + 1) * N, dtype=np.uint64)
    for log_nb in range(4):
        for b in range(2 * N):
            got = ref.sample_extract(k, N, acc, b, log_nb)
            exp = sample_extract_tb(acc, k, N, b, 1 << log_nb)
            assert [[int(x) for x in row] for row in got] == exp


@pytest.mark.parametrize("useful", [5, 3, 7, 9])
def test_alu_closed_forms(useful):
    """pea_alu_core.sv:55-160: the scalar constant is in_msg_cst[USEFUL_BIT-1:0] placed at the
    top of the body only (pea_alu_core.sv:59, pe_alu.sv:469-475), USEFUL_BIT = message + carry
    + padding bits; MULS multiplies by the truncated 5-bit constant.  Widths other than 5 catch
    a constant hard-wired to the 2+2-bit blocks' 5 bits."""
    n = 2049
    a0 = rng.integers(0, 2**64, n, dtype=np.uint64)
    a1 = rng.integers(0, 2**64, n, dtype=np.uint64)
    M = 2**64
    x, y = [int(v) for v in a0], [int(v) for v in a1]
    cst = 0b11010110110  # 11-bit DOp field, wider than every useful width
    top = (cst % (1 << useful)) * (1 << (64 - useful))
    cext = [0] * (n - 1) + [top]
    cases = {0x01: [(p + q) % M for p, q in zip(x, y)],
             0x02: [(p - q) % M for p, q in zip(x, y)],
             0x05: [(p * 13 + q) % M for p, q in zip(x, y)],
             0x09: [(p + c) % M for p, c in zip(x, cext)],
             0x0A: [(p - c) % M for p, c in zip(x, cext)],
             0x0B: [(c - p) % M for p, c in zip(x, cext)],
             0x0C: [(p * (cst % 32)) % M for p in x]}
    for op, exp in cases.items():
        got = ref.alu(op, a0, a1, mul_factor=13, msg_cst=cst, useful_bits=useful)
        assert [int(v) for v in got] == exp, hex(op)


# ---------------------------------------------------------------- key switch / blind rotation
def test_keyswitch_against_integer_dot_product():
    """Closed form out_j = [j=n] b' - sum d_il * KSK[i][l][j] mod 2^21, then rounded to 12 bits."""
    prm = ref.params(lwe_dim=40)
    kN, L, n = 2048, 5, 40
    blwe = rng.integers(0, 2**64, kN + 1, dtype=np.uint64)
    ksk = rng.integers(0, 2**21, (kN, L, n + 1), dtype=np.uint32)
    out, out2n = ref.keyswitch(prm, blwe, ksk)
    d = ref.decompose_batch(blwe[:kN], 64, 3, 5).astype(np.int64)  # (kN, L)
    acc = np.einsum("il,ilj->j", d, ksk.astype(np.int64))
    b = int(blwe[kN])
    bp = ((b >> 43) + ((b >> 42) & 1)) % 2**21
    exp = (-acc) % 2**21
    exp[n] = (bp - acc[n]) % 2**21
    assert np.array_equal(out.astype(np.int64), exp)
    assert np.array_equal(out2n.astype(np.int64), ((exp >> 9) + ((exp >> 8) & 1)) % 4096)


def test_blind_rotate_against_schoolbook():
    """A.3 composed from independent pieces: rotation by the monomial definition,
    org_decompose digits, schoolbook negacyclic products mod p (no NTT), big-integer
    p -> 2^64 switch.  N = 64 (cuts [3,3]), k = 1, L = 2, base 2^10, n = 3."""
    N, k, L, blog, n = 64, 1, 2, 10, 3
    prm = ref.params(lwe_dim=n, poly_size=N, pbs_base_log=blog, pbs_level=L)
    cuts = (3, 3)
    bsk = rng.integers(0, P, (n, k + 1, L, k + 1, N), dtype=np.uint64)
    lut = rng.integers(0, 2**64, (k + 1) * N, dtype=np.uint64)
    a2n = np.array([5, 0, 100, 17], dtype=np.uint16)
    got = ref.blind_rotate(prm, ref.bsk_to_ntt(prm, bsk.reshape(-1), cuts=cuts), lut, a2n, cuts=cuts)
    Cst = (1 << 160) // P
    M = 2**64
    acc = [[int(v) for v in lut[g * N:(g + 1) * N]] for g in range(k + 1)]
    for i in range(n):
        a = int(a2n[i])
        ct10 = [[(r - s) % M for r, s in zip(monomial_mult_definition(acc[g], a), acc[g])] for g in range(k + 1)]
        dig = [[org_decompose(x, 1 << blog, L, M) for x in ct10[g]] for g in range(k + 1)]
        for j in range(k + 1):
            e = np.zeros(N, np.uint64)
            for g in range(k + 1):
                for l in range(L):
                    d = np.array([dig[g][m][l] % P for m in range(N)], dtype=np.uint64)
                    prod = ref.negacyclic_mul(d, bsk[i, g, l, j])
                    e = np.array([(int(x) + int(y)) % P for x, y in zip(e, prod)], dtype=np.uint64)
            for m in range(N):
                t = int(e[m]) * Cst
                acc[j][m] = (acc[j][m] + (t >> 96) + ((t >> 95) & 1)) % M
    assert [int(v) for v in got] == acc[0] + acc[1]
