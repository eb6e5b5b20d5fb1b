"""GPU: the 512-thread latency variant of the blind rotation (small cohorts) gives the same
bits as the throughput kernel and as the oracle, at PBS level 1 and 2, with Many-LUT."""
import os

import numpy as np
import pytest

import luthe_amd as la

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


@pytest.mark.parametrize("level,base_log", [(1, 23), (2, 15)])
def test_latency_variant_bit_exact(oracle, level, base_log):
    print(f"case: latency blind rotation, L={level}", flush=True)
    prm = la.default_params(pbs_level=level, pbs_base_log=base_log, lut_slots=64)
    keys = la.Keys(prm, seed=70 + level)
    ctx = la.Context(prm, device=0)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    ctx.set_tuning(br_split_max=0)  # small batches on the 512-thread kernel, not the split variant
    rng = np.random.default_rng(level)
    tables = rng.integers(0, 16, (8, 16), dtype=np.uint8)
    luts = np.stack([la.lut_encode(prm, t) for t in tables])
    ctx.write_lut(0, luts)
    ml = la.lut_encode_many(prm, rng.integers(0, 16, (2, 8), dtype=np.uint8), 1)
    ctx.write_lut(10, ml)
    lat_max = la.lib().luthe_get_br_latency_max(ctx.handle)
    assert lat_max > 0
    for n in (1, 5, lat_max, lat_max + 3):
        print(f"case: latency blind rotation, L={level}, batch {n}", flush=True)
        msgs = rng.integers(0, 16, n)
        cts = keys.encrypt_msg(msgs, seed=n)
        gids = rng.integers(0, 8, n).astype(np.uint32)
        assert la.lib().luthe_set_br_latency_max(ctx.handle, lat_max) == 0
        a = ctx.pbs_batch(cts, gids)
        assert la.lib().luthe_set_br_latency_max(ctx.handle, 0) == 0
        b = ctx.pbs_batch(cts, gids)
        assert np.array_equal(a, b), n
        if n <= 5:
            assert la.lib().luthe_set_br_latency_max(ctx.handle, lat_max) == 0
            m = ctx.pbs_batch(cts[:4] if n > 4 else cts, np.full(min(n, 4), 10, np.uint32), log_lut_nb=1)
            assert la.lib().luthe_set_br_latency_max(ctx.handle, 0) == 0
            m2 = ctx.pbs_batch(cts[:4] if n > 4 else cts, np.full(min(n, 4), 10, np.uint32), log_lut_nb=1)
            assert np.array_equal(m, m2)
    # one ciphertext against the oracle through the latency kernel
    rp = oracle.params(pbs_base_log=base_log, pbs_level=level)
    bsk_ntt = oracle.bsk_to_ntt(rp, keys.bsk.reshape(-1), threads=THREADS)
    cts = keys.encrypt_msg([3, 12], seed=99)
    assert la.lib().luthe_set_br_latency_max(ctx.handle, lat_max) == 0
    got = ctx.pbs_batch(cts, np.array([2, 6], np.uint32))
    exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, luts, np.array([2, 6], np.uint32), cts, threads=THREADS)
    assert np.array_equal(got, exp)
    ctx.close()
