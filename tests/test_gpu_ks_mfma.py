"""The matrix-core key switch (int8 MFMA over three balanced byte planes of the KSK) against
the VALU key switch bit for bit, and sampled rows against the oracle, over parameter sets
that exercise every digit width the int8 path admits (balanced digits up to +-64), the
widest KSK modulus (2^24), an output width that is not a multiple of the 128-column tile,
and batch sizes around the 64-row tile (a single row, partial tiles, several tiles) and on
each tier of the VALU kernel's small-batch shapes (<= 4, <= 16, more ciphertexts)."""
import os

import numpy as np
import pytest

import luthe_amd as la

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
VALU = 1 << 30  # ks_mfma_min above every batch: the VALU kernel


def _inputs(prm, n, seed):
    rng = np.random.default_rng(seed)
    blwe = rng.integers(0, 2**64, (n, prm.blen), dtype=np.uint64)
    edge = [0, 2**64 - 1, 2**63, 2**63 - 1, 2**64 - 2**40, 2**40]
    for r in range(min(n, len(edge))):
        blwe[r, : prm.blen // 2] = edge[r]
    if n > 2:  # values on the rounding boundary of the top digit bits of each level width
        R = prm.ks_base_log * prm.ks_level
        blwe[2, ::3] = np.uint64(1) << np.uint64(63 - R)
    return blwe


@pytest.mark.parametrize("over", [dict(), dict(lwe_dim=200, ks_base_log=4, ks_level=4),
                                  dict(lwe_dim=200, ks_base_log=2, ks_level=8, ks_mod_log=24),
                                  dict(lwe_dim=1024, ks_base_log=6, ks_level=4, ks_mod_log=24),
                                  dict(lwe_dim=100, ks_base_log=7, ks_level=3)])
def test_mfma_keyswitch_matches(oracle, over):
    print(f"case: matrix-core key switch, parameters {over}", flush=True)
    prm = la.default_params(**over)
    keys = la.Keys(prm, seed=61)
    ctx = la.Context(prm, device=0)
    ctx.upload_ksk(keys.ksk)
    kw = {k: v for k, v in over.items() if k in ("lwe_dim", "ks_base_log", "ks_level", "ks_mod_log")}
    rp = oracle.params(**kw)
    for n in (1, 3, 5, 16, 17, 64, 65, 200):
        blwe = _inputs(prm, n, 62 + n)
        ctx.set_tuning(ks_mfma_min=1)
        got = ctx.keyswitch_batch(blwe)
        ctx.set_tuning(ks_mfma_min=VALU)
        ref = ctx.keyswitch_batch(blwe)
        assert np.array_equal(got, ref), f"batch {n}: matrix-core and VALU key switches differ"
        for c in sorted({0, 1, 2, n // 2, n - 1} & set(range(n))):
            _, exp = oracle.keyswitch(rp, blwe[c], keys.ksk)
            assert np.array_equal(got[c], exp), f"batch {n}, ciphertext {c}"
    ctx.close()


def test_mfma_keyswitch_pbs_large_batch(oracle):
    """A 1,000-ciphertext PBS batch through the matrix-core key switch equals the same batch
    through the VALU key switch, and sampled rows equal the oracle's PBS."""
    print("case: PBS batch of 1000 with both key switches (n=724, PBS 2^23 x 1)", flush=True)
    prm = la.default_params()
    keys = la.Keys(prm, seed=71)
    ctx = la.Context(prm, device=0)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    lut = la.lut_encode(prm, np.arange(16, dtype=np.uint8))
    ctx.write_lut(0, lut)
    n = 1000
    cts = keys.encrypt_msg(np.random.default_rng(72).integers(0, 16, n), seed=73)
    gids = np.zeros(n, np.uint32)
    ctx.set_tuning(ks_mfma_min=1)
    got = ctx.pbs_batch(cts, gids)
    ctx.set_tuning(ks_mfma_min=VALU)
    ref = ctx.pbs_batch(cts, gids)
    assert np.array_equal(got, ref)
    rp = oracle.params()
    bsk_ntt = oracle.bsk_to_ntt(rp, keys.bsk.reshape(-1), threads=THREADS)
    pick = np.array([0, 63, 64, 999])
    exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, lut.reshape(1, -1), np.zeros(pick.size, np.uint32), cts[pick],
                           threads=THREADS)
    assert np.array_equal(got[pick], exp)
    ctx.close()
