"""GPU: the split variant of the blind rotation (one 512-thread workgroup per GLWE polynomial,
the two exchanging NTT-domain digits through global memory once per CMUX) gives the bits of the
512-thread latency kernel and of the oracle, at PBS level 1 and 2, with Many-LUT, for pair counts
that do and do not fill a group of eight, over launches issued back to back on one exchange area
(no stale round tags), and over a batch list whose empty slots both workgroups of a pair skip."""
import os

import numpy as np
import pytest

import luthe_amd as la

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module", params=[(1, 23), (2, 15)], ids=["L1", "L2"])
def env(request, oracle):
    level, base_log = request.param
    prm = la.default_params(pbs_level=level, pbs_base_log=base_log, lut_slots=64)
    keys = la.Keys(prm, seed=90 + level)
    ctx = la.Context(prm, device=0)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    rng = np.random.default_rng(91 + level)
    tables = rng.integers(0, 16, (8, 16), dtype=np.uint8)
    luts = np.stack([la.lut_encode(prm, t) for t in tables])
    ctx.write_lut(0, luts)
    ctx.write_lut(10, la.lut_encode_many(prm, rng.integers(0, 16, (2, 8), dtype=np.uint8), 1))
    ctx.write_lut(11, la.lut_encode_many(prm, rng.integers(0, 16, (4, 4), dtype=np.uint8), 2))
    yield level, base_log, prm, keys, ctx, luts, rng
    ctx.close()


def _both(ctx, smax, cts, gids, log_lut_nb=0):
    """(split variant, 512-thread latency variant) outputs of one batch."""
    ctx.set_tuning(br_split_max=smax)
    a = ctx.pbs_batch(cts, gids, log_lut_nb=log_lut_nb)
    ctx.set_tuning(br_split_max=0)
    b = ctx.pbs_batch(cts, gids, log_lut_nb=log_lut_nb)
    ctx.set_tuning(br_split_max=smax)
    return a, b


def test_split_equals_latency_kernel(env):
    level, _, prm, keys, ctx, _, rng = env
    smax = ctx.tuning()["br_split_max"]
    assert smax >= 8
    for n in (1, 8, 9, 23, smax):
        print(f"case: split blind rotation, L={level}, batch {n}", flush=True)
        cts = keys.encrypt_msg(rng.integers(0, 16, n), seed=1000 + n)
        gids = rng.integers(0, 8, n).astype(np.uint32)
        a, b = _both(ctx, smax, cts, gids)
        assert np.array_equal(a, b), n
    for lg, gid in ((1, 10), (2, 11)):
        print(f"case: split blind rotation, L={level}, Many-LUT 2^{lg}", flush=True)
        cts = keys.encrypt_msg(rng.integers(0, 16 >> lg, 5), seed=1100 + lg)
        a, b = _both(ctx, smax, cts, np.full(5, gid, np.uint32), log_lut_nb=lg)
        assert np.array_equal(a, b), lg


def test_split_exchange_paths(env):
    """Both exchange paths (the shared XCC's L2, the default on MI355X, and agent-scope stores)
    give the latency kernel's bits; an invalid setting is refused."""
    level, _, prm, keys, ctx, _, rng = env
    lib, h = la.lib(), ctx.handle
    assert lib.luthe_get_br_split_local(h) == 1  # MI355X places workgroups b and b ^ 8 on one XCC
    smax = ctx.tuning()["br_split_max"]
    for n in (1, 9):
        cts = keys.encrypt_msg(rng.integers(0, 16, n), seed=1500 + n)
        gids = rng.integers(0, 8, n).astype(np.uint32)
        ref = None
        for local in (0, 1):
            print(f"case: split exchange path local={local}, L={level}, batch {n}", flush=True)
            assert lib.luthe_set_br_split_local(h, local) == 0
            a, b = _both(ctx, smax, cts, gids)
            assert np.array_equal(a, b), (n, local)
            ref = a if ref is None else ref
            assert np.array_equal(a, ref), (n, local)
    assert lib.luthe_set_br_split_local(h, 2) == -1  # LUTHE_ERR_ARG
    assert lib.luthe_get_br_split_local(h) == 1


def test_split_vs_oracle(env, oracle):
    level, base_log, prm, keys, ctx, luts, _ = env
    print(f"case: split blind rotation against the oracle, L={level}", flush=True)
    rp = oracle.params(pbs_base_log=base_log, pbs_level=level)
    bsk_ntt = oracle.bsk_to_ntt(rp, keys.bsk.reshape(-1), threads=THREADS)
    cts = keys.encrypt_msg([3, 12, 7], seed=99)
    gids = np.array([2, 6, 0], np.uint32)
    ctx.set_tuning(br_split_max=-1)
    got = ctx.pbs_batch(cts, gids)
    exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, luts, gids, cts, threads=THREADS)
    assert np.array_equal(got, exp)


def test_split_back_to_back_launches(env):
    """Launches queued without a host synchronisation reuse one exchange area: each starts from
    a zeroed area, so no word of an earlier launch passes as a word of the current one."""
    level, _, prm, keys, ctx, _, rng = env
    print(f"case: split launches back to back, L={level}", flush=True)
    ctx.set_tuning(br_split_max=-1)
    sizes = (5, 5, 17, 2, 17)
    cts = [keys.encrypt_msg(rng.integers(0, 16, n), seed=1200 + k) for k, n in enumerate(sizes)]
    gids = [rng.integers(0, 8, n).astype(np.uint32) for n in sizes]
    exp = [ctx.pbs_batch(c, g) for c, g in zip(cts, gids)]
    d_in = [la.DeviceBuffer.from_numpy(c) for c in cts]
    d_g = [la.DeviceBuffer.from_numpy(g) for g in gids]
    d_out = [la.DeviceBuffer((n, prm.blen), np.uint64) for n in sizes]
    for i in range(len(sizes)):
        ctx.pbs_batch_dev(d_in[i], d_out[i], d_g[i], sizes[i])
    ctx.sync()
    for i in range(len(sizes)):
        assert np.array_equal(d_out[i].numpy(), exp[i]), i


def test_split_batch_list_empty_slots(env):
    level, _, prm, keys, ctx, _, rng = env
    print(f"case: split variant over a batch list with empty slots, L={level}", flush=True)
    ctx.set_tuning(br_split_max=-1)
    spec = [(2, 0), (3, 1)]
    bufs, exp = [], []
    for k, (n, lg) in enumerate(spec):
        cts = keys.encrypt_msg(rng.integers(0, 16 >> lg, n), seed=1300 + k)
        gids = np.full(n, 10 if lg else 1, np.uint32)
        d_in, d_gid = la.DeviceBuffer.from_numpy(cts), la.DeviceBuffer.from_numpy(gids)
        d_out = la.DeviceBuffer((n << lg, prm.blen), np.uint64)
        d_out.upload(np.zeros((n << lg, prm.blen), np.uint64))
        exp.append(ctx.pbs_batch(cts, gids, log_lut_nb=lg))
        bufs.append((d_in, d_gid, d_out))
    descs = la.DeviceBuffer.from_numpy(la.batch_descs(
        [(b[0].ptr.value, b[2].ptr.value, b[1].ptr.value, n, lg) for b, (n, lg) in zip(bufs, spec)]))
    ctx.pbs_batch_list_dev(descs, len(spec), 13)  # slots 5..12 stay empty
    ctx.sync()
    for k in range(len(spec)):
        assert np.array_equal(bufs[k][2].numpy(), exp[k]), k


def test_split_wait_limit(env):
    """The split kernel's bounded wait: with no waiting allowed (br_split_wait_us = 0) a workgroup
    that finds a partner word missing gives up, its partner then does the same, the next
    synchronising call reports LUTHE_ERR_HIP once, and the context works as before afterwards."""
    level, _, prm, keys, ctx, _, rng = env
    print(f"case: split variant with no waiting allowed, L={level}", flush=True)
    cts = keys.encrypt_msg(rng.integers(0, 16, 3), seed=1400)
    gids = np.zeros(3, np.uint32)
    ctx.set_tuning(br_split_max=-1, br_split_wait_us=-1)
    exp = ctx.pbs_batch(cts, gids)
    ctx.set_tuning(br_split_wait_us=0)
    d_in, d_g = la.DeviceBuffer.from_numpy(cts), la.DeviceBuffer.from_numpy(gids)
    d_out = la.DeviceBuffer((3, prm.blen), np.uint64)
    ctx.pbs_batch_dev(d_in, d_out, d_g, 3)
    with pytest.raises(la.LutheError) as e:
        ctx.sync()
    assert e.value.code == -3  # LUTHE_ERR_HIP
    ctx.sync()  # reported once
    ctx.set_tuning(br_split_wait_us=-1)
    assert np.array_equal(ctx.pbs_batch(cts, gids), exp)
    with pytest.raises(la.LutheError):
        ctx.set_tuning(br_split_wait_us=-2)
