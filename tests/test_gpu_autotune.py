"""GPU: the autotuner times its candidates, writes and re-reads its JSON, and every setting
it can pick (blind-rotation width, 256- or 512-thread throughput kernel, split- and
latency-variant thresholds, NTT polynomials per workgroup) leaves the results bit-identical."""
import json
import os
import tempfile

import numpy as np
import pytest

import luthe_amd as la

pytestmark = pytest.mark.gpu


def test_autotune_and_settings_are_bit_exact():
    print("case: autotune (n=64, batch 1024)", flush=True)
    prm = la.default_params(lwe_dim=64)
    keys = la.Keys(prm, seed=61)
    ctx = la.Context(prm, device=0)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    ctx.write_lut(0, la.lut_encode(prm, np.arange(16, dtype=np.uint8)))
    rng = np.random.default_rng(62)
    cts = keys.encrypt_msg(rng.integers(0, 16, 700), seed=63)
    gids = np.zeros(700, np.uint32)
    ref = ctx.pbs_batch(cts, gids)
    a = rng.integers(0, la.P, (37, 2048), dtype=np.uint64)
    bn = rng.integers(0, la.P, (37, 2048), dtype=np.uint64)
    ref_ntt = ctx.ntt_polymul(a, bn)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tuning.json")
        t = ctx.autotune(1024, path)
        doc = json.load(open(path))
        assert doc["br_grid"] == t["br_grid"] and doc["ntt_polys_per_wg"] == t["ntt_polys_per_wg"]
        assert len(doc["measured"]["blind_rotation_width"]) >= 1
        assert t["cohort_max"] > 0 and t["ntt_polys_per_wg"] in (1, 2, 4)
        ctx.set_tuning(br_grid=0, br_latency_max=0, ntt_polys_per_wg=2, cohort_max=0)
        ctx.load_tuning(path)
        assert ctx.tuning() == {**t, "br_latency_max": t["br_latency_max"]}
    for grid, lat, np_, pair, split in [(256, 0, 1, 0, 0), (512, 700, 4, 0, 0), (1024, 0, 2, 0, 0),
                                        (0, -1, 2, -1, -1), (0, 0, 2, 1, 0), (300, 0, 2, 1, 0), (0, 0, 2, 0, 700)]:
        print(f"case: settings grid={grid} latency_max={lat} ntt_polys_per_wg={np_} pair={pair} split_max={split}",
              flush=True)
        ctx.set_tuning(br_grid=grid, br_latency_max=lat, ntt_polys_per_wg=np_, br_pair=pair, br_split_max=split)
        assert np.array_equal(ctx.pbs_batch(cts, gids), ref)
        assert np.array_equal(ctx.ntt_polymul(a, bn), ref_ntt)
    print("case: split variant over the first 100 ciphertexts", flush=True)
    ctx.set_tuning(br_split_max=100)
    assert np.array_equal(ctx.pbs_batch(cts[:100], gids[:100]), ref[:100])
    with pytest.raises(la.LutheError):
        ctx.set_tuning(ntt_polys_per_wg=3)
    with pytest.raises(la.LutheError):
        ctx.set_tuning(br_pair=2)
    with pytest.raises(la.LutheError):
        ctx.set_tuning(br_split_max=-2)
    ctx.close()
