"""GPU: the device builds of the Goldilocks primitives (inline-asm carry chains included)
and of the high-word decomposition, at edge operands and random ones, against
exact integer arithmetic and the oracle's modulus switch and decomposition."""
import ctypes as C
import itertools
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 2**64 - 2**32 + 1
EPS = 2**32 - 1
M64 = 2**64 - 1
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def prims():
    path = os.path.join(HERE, "gpu_prims", "libprims.so")
    if not os.path.exists(path):
        pytest.fail("tests/gpu_prims/libprims.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    lib.prims_run.restype = C.c_int
    lib.prims_run.argtypes = [C.c_void_p] * 3 + [C.c_int, C.c_void_p]
    return lib


def test_device_primitives_exact(prims, oracle):
    edges = [0, 1, 2, 3, EPS - 1, EPS, EPS + 1, 2**32, 2**31, 2**31 + 1, 2**63 - 1, 2**63, 2**63 + 1,
             2**63 + 2**31, P - 2, P - 1, P, P + 1, 2**64 - 2**32, 2**64 - 2, 2**64 - 1,
             0xFFFFFFFE_FFFFFFFF, 0x80000000_80000000, 0x7FFFFFFF_80000001]
    rng = np.random.default_rng(77)
    pairs = list(itertools.product(edges, edges))
    pairs += [(int(a), int(b)) for a, b in zip(rng.integers(0, 2**64, 20000, dtype=np.uint64),
                                                 rng.integers(0, 2**64, 20000, dtype=np.uint64))]
    zs = [0, 1, 2**31 - 1, 2**31, 2**32 - 2, 2**32 - 1, 12345] + [int(v) for v in rng.integers(0, 2**32, 64)]
    n = len(pairs)
    a = np.array([p[0] for p in pairs], np.uint64)
    b = np.array([p[1] for p in pairs], np.uint64)
    c = np.array([zs[i % len(zs)] for i in range(n)], np.uint32)
    out = np.zeros((n, 67), np.uint64)
    print(f"case: {n} operand pairs through the device primitives", flush=True)
    assert prims.prims_run(a.ctypes.data, b.ctypes.data, c.ctypes.data, n, out.ctypes.data) == 0
    for i, (x, y) in enumerate(pairs):
        z = int(c[i])
        xc, yc = x % P, y % P
        zz = 0 if z == 2**32 - 1 else z
        zs32 = z - 2**32 if z >= 2**31 else z
        exp = [xc, (xc + yc) % P, (xc - yc) % P, (x * y) % P, (x + (y << 64)) % P,
               (x + ((y & EPS) << 64)) % P, (x + (y << 64) + (zz << 128)) % P,
               x - yc if x >= yc else x - yc + 2**64 - EPS,  # sub_fold: exact 64-bit result
               xc * 2**24 % P, xc * 2**48 % P, xc * 2**72 % P, xc * pow(2, 120, P) % P, xc * pow(2, 168, P) % P,
               oracle.modsw(xc), zs32 % P, oracle.decompose(x, 64, 23, 1)[0] & M64]
        acc = (x + (y << 64) + ((z & 3) << 128)) + (y + (x << 64))
        exp += [acc & M64, (acc >> 64) & M64, acc >> 128]
        exp += [(xc - yc) % P, (yc - xc) % P, (xc - (~x & M64) % P) % P, ((x ^ y) % P)]
        nx, ny = ~x & M64, ~y & M64
        exp += [x * y % P, y * x % P, (x ^ y) * y % P, nx * x % P]
        exp += [(x + (y << 64) + (zz << 128)) % P, (y + (x << 64) + ((zz >> 1) << 128)) % P,
                (x + (x << 64)) % P, (y + (ny << 64) + (0xFFFFFFFE << 128)) % P]
        xs = x - 2**64 if x >= 2**63 else x  # x as a signed 64-bit value
        exp += [xs % P, (xs >> 17) % P]
        def sf(a, b):  # sub_fold
            return a - b if a >= b else a - b + 2**64 - EPS
        xy = x ^ y
        exp += [sf(x, yc), sf(y, xc), sf(xy, yc), sf(nx, xc), sf(y, xy % P)]
        for lo, hi, cin, pl, ph in ((x, y, z & 3, y, x), (y, x, (z >> 2) & 3, x, y), (xy, nx, 1, ny, x),
                                    (nx, ny, 0, nx, ny)):
            acc = (lo + (hi << 64) + (cin << 128)) + (pl + (ph << 64))
            exp += [acc & M64, (acc >> 64) & M64, acc >> 128]
        exp += [v * pow(2, s, P) % P for v, s in zip((xc, yc, xc, yc, xy % P, xc, yc, xc),
                                                    (24, 48, 72, 120, 168, 0, 96, 36))]
        et, eh = (x, y, xy, nx), (z, z >> 1, y & EPS, EPS)
        tails = [(t + h * EPS) % P for t, h in zip(et, eh)]
        exp += tails[:2] + tails[1:4] + tails
        got = [int(v) for v in out[i]]
        assert got == exp, (hex(x), hex(y), z, [k for k in range(67) if got[k] != exp[k]])
