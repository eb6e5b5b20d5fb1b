"""GPU edge cases, each against the oracle: an empty batch, the last LUT id and device LUT
ids beyond the LUT memory (clamped to its last slot), the last slot of a 65,536-slot
ciphertext memory through LD/ST, a larger LWE dimension, other key-switch parameters, and
a ragged batch of 4,097 ciphertexts (four waves of the persistent kernel plus one)."""
import ctypes as C
import os

import numpy as np
import pytest

import luthe_amd as la
from luthe_amd import dop

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _setup(oracle, seed, **over):
    prm = la.default_params(**over)
    keys = la.Keys(prm, seed=seed)
    ctx = la.Context(prm, device=0)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    kw = {k: v for k, v in over.items() if k in ("lwe_dim", "pbs_base_log", "pbs_level", "ks_base_log",
                                                  "ks_level", "ks_mod_log")}
    rp = oracle.params(**kw)
    bsk_ntt = oracle.bsk_to_ntt(rp, keys.bsk.reshape(-1), threads=THREADS)
    return prm, keys, ctx, rp, bsk_ntt


def test_empty_batch_and_lut_ids(oracle):
    print("case: empty batch, LUT id 4095, device LUT ids clamped", flush=True)
    prm, keys, ctx, rp, bsk_ntt = _setup(oracle, 401, lwe_dim=64, lut_slots=4096)
    assert la.lib().luthe_pbs_batch(ctx.handle, None, None, None, 0, 0, None) == 0
    assert la.lib().luthe_alu_batch(ctx.handle, 1, None, None, None, 0, 0, 0, None) == 0
    table = (np.arange(16) * 11 % 16).astype(np.uint8)
    lut = la.lut_encode(prm, table)
    ctx.write_lut(4095, lut)
    cts = keys.encrypt_msg([2, 9, 15], seed=402)
    got = ctx.pbs_batch(cts, np.array([4095, 4096, 0xFFFFFFFF], np.uint32))
    exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, lut.reshape(1, -1), np.zeros(3, np.uint32), cts, threads=3)
    assert np.array_equal(got, exp)
    with pytest.raises(la.LutheError):  # host-side ids are checked
        ctx.write_lut(4096, lut)
    ctx.close()


def test_last_slot_of_a_large_memory(oracle):
    print("case: LD/ST on slot 65535 of 65536", flush=True)
    prm, keys, ctx, rp, bsk_ntt = _setup(oracle, 403, lwe_dim=64, ct_slots=65536)
    ctx.write_lut(1, la.lut_encode(prm, np.arange(16, dtype=np.uint8)))
    ct = keys.encrypt_msg([6], seed=404)
    ctx.write_ct(65535, ct)
    t = ctx.submit_dops([dop.ld(3, 65535), dop.pbs(4, 3, 1, flush=True), dop.st(65534, 4), dop.st(0, 3)])
    ctx.wait(t)
    exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, la.lut_encode(prm, np.arange(16, dtype=np.uint8)).reshape(1, -1),
                           np.zeros(1, np.uint32), ct, threads=1)
    assert np.array_equal(ctx.read_ct(65534, 1), exp)
    assert np.array_equal(ctx.read_ct(0, 1), ct)
    assert np.array_equal(ctx.read_reg(4, 1), exp)  # architectural register after renaming
    with pytest.raises(la.LutheError):
        ctx.write_ct(65536, ct)
    ctx.close()


@pytest.mark.parametrize("over", [dict(lwe_dim=1024), dict(lwe_dim=200, ks_base_log=4, ks_level=4),
                                  dict(lwe_dim=200, ks_base_log=2, ks_level=8, ks_mod_log=24),
                                  dict(lwe_dim=200, pbs_base_log=20, pbs_level=2),
                                  dict(lwe_dim=200, pbs_base_log=28, pbs_level=1)])
def test_other_parameters(oracle, over):
    """Every blind-rotation kernel: the split variant (two workgroups per ciphertext) that
    small batches take by default, the 512-thread latency variant (split threshold 0) and the
    persistent throughput kernel (latency threshold 0 as well); the last two sets exceed the
    high-word decomposition's bounds (L * base_log > 31 or base_log > 23), so they run the
    generic 64-bit decomposition and unpretwisted lift."""
    print(f"case: parameters {over}", flush=True)
    prm, keys, ctx, rp, bsk_ntt = _setup(oracle, 405, **over)
    tables = np.random.default_rng(406).integers(0, 16, (4, 16), dtype=np.uint8)
    luts = np.stack([la.lut_encode(prm, t) for t in tables])
    ctx.write_lut(0, luts)
    cts = keys.encrypt_msg([0, 5, 10, 15], seed=407)
    gids = np.arange(4, dtype=np.uint32)
    exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, luts, gids, cts, threads=THREADS)
    got = ctx.pbs_batch(cts, gids)
    assert np.array_equal(got, exp), "split kernel"
    ctx.set_tuning(br_split_max=0)
    got = ctx.pbs_batch(cts, gids)
    assert np.array_equal(got, exp), "latency kernel"
    assert la.lib().luthe_set_br_latency_max(ctx.handle, 0) == 0
    got = ctx.pbs_batch(cts, gids)
    assert np.array_equal(got, exp), "throughput kernel"
    ctx.close()


def test_ragged_batch_4097(oracle):
    print("case: batch of 4097 (n=724, PBS 2^23 x 1), sampled against the oracle", flush=True)
    prm, keys, ctx, rp, bsk_ntt = _setup(oracle, 408)
    ctx.write_lut(0, la.lut_encode(prm, np.arange(16, dtype=np.uint8)))
    n = 4097
    msgs = np.random.default_rng(409).integers(0, 16, n)
    cts = keys.encrypt_msg(msgs, seed=410)
    got = ctx.pbs_batch(cts, np.zeros(n, np.uint32))
    pick = np.array([0, 1, 1023, 1024, 2047, 4095, 4096])
    exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, la.lut_encode(prm, np.arange(16, dtype=np.uint8)).reshape(1, -1),
                           np.zeros(pick.size, np.uint32), cts[pick], threads=THREADS)
    assert np.array_equal(got[pick], exp)
    assert (keys.decrypt_msg(got) == msgs).mean() > 0.9
    ctx.close()
