"""GPU: keys, LUTs and ciphertexts given as HPU memory images run the same PBS, bit for bit,
as the engine's own layouts; the DOp path's trace log holds REFILL / ISSUE / RETIRE records
that pack into the HPU trace format."""
import os

import numpy as np
import pytest

import luthe_amd as la
from luthe_amd import dop

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def test_bsk_lut_ct_images_vs_engine_layouts(oracle):
    print("case: HPU images (BSK cut [3,4,4] chunk 32, 2-channel KSK image, LUT bodies, 2-channel ct image), "
          "n=64", flush=True)
    prm = la.default_params(lwe_dim=64)
    keys = la.Keys(prm, seed=91)
    rp = oracle.params(lwe_dim=64)
    bsk_ntt = oracle.bsk_to_ntt(rp, keys.bsk.reshape(-1), threads=THREADS).reshape(prm.bsk_shape())
    lay = la.hpu_layout(axi_bytes=64, pem_pc=2, regf_coef_nb=32, bsk_chunk=32, ntt_cut=(3, 4, 4))
    image = la.hpu_bsk_to_image(prm, lay, bsk_ntt)
    rng = np.random.default_rng(92)
    table = rng.integers(0, 16, 16, dtype=np.uint8)
    lut = la.lut_encode(prm, table)
    msgs = rng.integers(0, 16, 24)
    cts = keys.encrypt_msg(msgs, seed=93)
    ct_img = la.hpu_ct_to_image(prm, lay, cts)
    klay = la.hpu_ksk_layout(axi_bytes=64, ksk_pc=2, ksk_cut_nb=4, lbx=3, lby=64, lbz=3)
    ksk_img = la.hpu_ksk_to_image(prm, klay, keys.ksk)

    ref_ctx = la.Context(prm, device=0)
    ref_ctx.upload_bsk(keys.bsk, la.BSK_COEF)
    ref_ctx.upload_ksk(keys.ksk)
    ref_ctx.write_lut(5, lut)
    exp = ref_ctx.pbs_batch(cts, np.full(24, 5, np.uint32))
    ref_ctx.close()

    ctx = la.Context(prm, device=0)
    ctx.upload_bsk_hpu_image(lay, image)
    ctx.upload_ksk_hpu_image(klay, ksk_img)
    ctx.write_lut_hpu_image(5, lut[2048:])
    got = ctx.pbs_batch(la.hpu_ct_from_image(prm, lay, ct_img, 24), np.full(24, 5, np.uint32))
    ctx.close()
    assert np.array_equal(got, exp)
    assert np.array_equal(la.hpu_ct_from_image(prm, lay, la.hpu_ct_to_image(prm, lay, got), 24), got)
    assert (keys.decrypt_msg(got) == table[msgs]).mean() >= 0.9


def test_dop_trace_records(oracle):
    print("case: DOp-path trace (LD, MAC, PBS, ST)", flush=True)
    prm = la.default_params(lwe_dim=32)
    keys = la.Keys(prm, seed=94)
    ctx = la.Context(prm, device=0)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    ctx.write_lut(3, la.lut_encode(prm, np.arange(16, dtype=np.uint8)))
    ctx.write_ct(10, keys.encrypt_msg([1, 2], seed=95))
    ctx.trace_enable(1000)
    prog = [dop.ld(0, 10), dop.ld(1, 11), dop.mac(2, 0, 1, 2), dop.pbs(4, 2, 3), dop.pbs(5, 0, 3),
            dop.st(20, 4), dop.st(21, 5), dop.sync()]
    t = ctx.submit_dops(prog)
    ctx.wait(t)
    ev = ctx.trace_read()
    ctx.close()
    by = {c: [e for e in ev if e["cmd"] == c] for c in (la.TRC_REFILL, la.TRC_ISSUE, la.TRC_RETIRE)}
    words = [w & 0xFFFFFFFF for w in prog]
    assert [e["insn"] for e in by[la.TRC_REFILL]] == words
    assert sorted(e["insn"] for e in by[la.TRC_ISSUE]) == sorted(words[:-1])  # SYNC is never issued
    assert [e["insn"] for e in by[la.TRC_RETIRE]] == words
    pbs_ret = [e for e in by[la.TRC_RETIRE] if e["insn"] in (words[3], words[4])]
    assert [e["pe_reserved"] for e in pbs_ret] == [1, 1]  # the two PBS sit in different levels
    assert len({e["sync_id"] for e in ev}) == 1
    ts = [e["timestamp"] for e in ev]
    assert ts == sorted(ts)
    assert la.trace_unpack(prm.lwe_dim, la.trace_pack(prm.lwe_dim, ev)) == ev
