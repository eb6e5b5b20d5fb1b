"""GPU: luthe_pbs_batch_noalloc, the PBS batch entry point that allocates nothing (its scratch
is a fixed device array), called through the C ABI and compared bit for bit with
luthe_pbs_batch on the same inputs: two calls in a row, the second with a smaller batch, plus
the refusal of a batch beyond the fixed scratch."""
import ctypes as C

import numpy as np
import pytest

import luthe_amd as la

pytestmark = pytest.mark.gpu


def test_noalloc_matches_pbs_batch():
    print("case: luthe_pbs_batch_noalloc vs luthe_pbs_batch, n=724, batches 600 then 37", flush=True)
    prm = la.default_params(lut_slots=64)
    keys = la.Keys(prm, seed=9001)
    ctx = la.Context(prm, device=0)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    rng = np.random.default_rng(9002)
    tables = rng.integers(0, 16, (64, 16), dtype=np.uint8)
    ctx.write_lut(0, np.stack([la.lut_encode(prm, t) for t in tables]))
    L = la.lib()
    bytes0 = ctx.device_bytes()
    for n, log_nb in ((600, 0), (37, 1)):  # the throughput kernel, then the latency variant
        msgs = rng.integers(0, 16 >> log_nb, n)
        cts = keys.encrypt_msg(msgs, seed=9003 + n)
        gids = rng.integers(0, 64, n).astype(np.uint32)
        d_in, d_gid = la.DeviceBuffer.from_numpy(cts), la.DeviceBuffer.from_numpy(gids)
        out_a = la.DeviceBuffer((n << log_nb, prm.blen), np.uint64)
        out_b = la.DeviceBuffer((n << log_nb, prm.blen), np.uint64)
        # the no-allocation call first, on a context whose own scratch has not been grown yet
        rc = L.luthe_pbs_batch_noalloc(ctx.handle, d_in.ptr, out_b.ptr, d_gid.ptr, log_nb, n, None)
        assert rc == 0, la.lib().luthe_last_error(ctx.handle)
        ctx.sync()
        if n == 600:
            assert ctx.device_bytes() == bytes0, "the no-allocation call allocated context memory"
        rc = L.luthe_pbs_batch(ctx.handle, d_in.ptr, out_a.ptr, d_gid.ptr, log_nb, n, None)
        assert rc == 0
        ctx.sync()
        a, b = out_a.numpy(), out_b.numpy()
        assert np.array_equal(a, b), f"batch {n}: outputs differ"
        if log_nb == 0:
            assert (keys.decrypt_msg(b) == tables[gids, msgs]).mean() > 0.9
    # beyond the fixed scratch: refused before any work
    big = 8192  # LUTHE_NOALLOC_MAX_BATCH
    rc = L.luthe_pbs_batch_noalloc(ctx.handle, C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), 0, big + 1, None)
    assert rc == -1
    ctx.close()
