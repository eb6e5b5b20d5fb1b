"""HPU memory images and trace records (host only).

* bsk_order: the reference's own function's outputs (tests/golden/bsk_order_fixtures.npz,
  written by tests/golden/make_bsk_order_fixtures.py) for ten cuts; the cut [5,6] must equal
  this engine's NTT-domain order, whose closed form (position 64k+w holds the evaluation at
  zeta^(2(k+32w)+1)) is pinned in test_oracle.py.
* BSK image: the loop order of gen_stimuli.py:108-136 checked element by element against
  closed forms, round trips for several layouts.
* Ciphertext image: channel striping, body word and page-aligned slots (pem_load.sv,
  pem_common_param_pkg.sv), round trips.
* Trace records: field positions of the packed isc_trace_t, round trips.
"""
import os

import numpy as np
import pytest

import luthe_amd as la

HERE = os.path.dirname(os.path.abspath(__file__))
rng = np.random.default_rng(777)


def _fixtures():
    return np.load(os.path.join(HERE, "golden", "bsk_order_fixtures.npz"))


def test_bsk_order_matches_reference_function():
    fx = _fixtures()
    assert len(fx.files) == 10
    for name in fx.files:
        cut = [int(c) for c in name.split("_")[1:]]
        assert np.array_equal(la.hpu_bsk_order(cut), fx[name]), cut


def test_bsk_order_56_is_engine_ntt_order():
    order = la.hpu_bsk_order([5, 6])
    k, w = np.meshgrid(np.arange(32), np.arange(64), indexing="ij")
    assert np.array_equal(order[(64 * k + w).ravel()], (k + 32 * w).ravel())
    for cut in ([11], [3, 4, 4], [1] * 11):
        o = la.hpu_bsk_order(cut)
        assert np.array_equal(np.sort(o), np.arange(2048))
    assert np.array_equal(la.hpu_bsk_order([11]), np.arange(2048))


@pytest.mark.parametrize("L", [1, 2])
@pytest.mark.parametrize("cut,chunk", [([5, 6], 2048), ([5, 6], 64), ([3, 4, 4], 32), ([2, 3, 3, 3], 128)])
def test_bsk_image_layout_and_round_trip(L, cut, chunk):
    p = la.default_params(lwe_dim=3, pbs_level=L)
    lay = la.hpu_layout(bsk_chunk=chunk, ntt_cut=cut)
    bsk = rng.integers(0, la.P, p.bsk_shape(), dtype=np.uint64)   # [i][g][l][j][N], ntt() order cut [5,6]
    img = la.hpu_bsk_to_image(p, lay, bsk)
    assert img.size == la.lib().luthe_hpu_bsk_image_words(la.C.byref(p))
    assert np.array_equal(la.hpu_bsk_from_image(p, lay, img), bsk)
    # closed form: image[i][j][s][g][l][c] = bsk[i][g][l][j][pos56(order_hw[s*chunk + c])]
    K1, N = 2, 2048
    pos56 = np.argsort(la.hpu_bsk_order([5, 6]))
    hw = la.hpu_bsk_order(cut)
    v = img.reshape(p.lwe_dim, K1, N // chunk, K1, L, chunk)
    for i, j, g, l in [(0, 0, 0, 0), (2, 1, 1, L - 1), (1, 1, 0, 0), (1, 0, 1, L - 1)]:
        poly = v[i, j, :, g, l, :].reshape(N)
        assert np.array_equal(poly, bsk[i, g, l, j][pos56[hw]]), (i, j, g, l)


def test_bsk_image_rejects_bad_layouts():
    p = la.default_params(lwe_dim=2)
    bsk = np.zeros(p.bsk_shape(), np.uint64)
    for lay in (la.hpu_layout(ntt_cut=[5, 5]), la.hpu_layout(bsk_chunk=48), la.hpu_layout(bsk_chunk=0)):
        with pytest.raises(la.LutheError):
            la.hpu_bsk_to_image(p, lay, bsk)


@pytest.mark.parametrize("axi,pc,rc", [(64, 1, 32), (64, 2, 32), (32, 2, 64), (16, 1, 16), (64, 2, 128)])
def test_ct_image_layout_and_round_trip(axi, pc, rc):
    p = la.default_params(lwe_dim=4)
    lay = la.hpu_layout(axi_bytes=axi, pem_pc=pc, regf_coef_nb=rc)
    cts = rng.integers(0, 2**64, (5, p.blen), dtype=np.uint64)
    img = la.hpu_ct_to_image(p, lay, cts)
    sb = la.hpu_ct_slot_bytes(p, lay)
    words_per_pc = 2048 * 8 // axi // pc
    assert sb % 4096 == 0 and sb >= (words_per_pc + 1) * axi and sb - 4096 < (words_per_pc + 1) * axi
    assert np.array_equal(la.hpu_ct_from_image(p, lay, img, 5), cts)
    region = img.reshape(pc, 5, sb)
    for s in range(5):
        # body: first u64 of the extra word of channel 0
        assert region[0, s, words_per_pc * axi:words_per_pc * axi + 8].view(np.uint64)[0] == cts[s, -1]
        # mask: register-file word w, channel p holds coefficients w*rc + p*rc/pc + [0, rc/pc)
        for ch in range(pc):
            stream = region[ch, s, :words_per_pc * axi].view(np.uint64)
            idx = np.arange(2048).reshape(-1, rc)[:, ch * (rc // pc):(ch + 1) * (rc // pc)].ravel()
            assert np.array_equal(stream, cts[s, idx]), (s, ch)


def test_ct_image_rejects_bad_layouts():
    p = la.default_params(lwe_dim=4)
    for lay in (la.hpu_layout(pem_pc=3), la.hpu_layout(axi_bytes=48), la.hpu_layout(regf_coef_nb=33)):
        assert la.hpu_ct_slot_bytes(p, lay) == 0


def test_lut_image():
    p = la.default_params(lwe_dim=4)
    bodies = rng.integers(0, 2**64, (3, 2048), dtype=np.uint64)
    g = la.hpu_lut_from_image(p, bodies)
    assert g.shape == (3, p.lut_len)
    assert np.all(g[:, :2048] == 0) and np.array_equal(g[:, 2048:], bodies)
    lut = la.lut_encode(p, np.arange(16, dtype=np.uint8))
    assert np.array_equal(la.hpu_lut_from_image(p, lut[2048:])[0], lut)  # trivial LUTs round-trip


def test_trace_record_fields():
    assert la.lib().luthe_trace_record_bits(724) == 105
    assert la.lib().luthe_trace_record_bits(1024) == 105
    assert la.lib().luthe_trace_record_bits(1025) == 106

    def one(**f):
        e = {k: 0 for k, _ in la.TraceEvent._fields_}
        e.update(f)
        return la.trace_pack(724, [e])[0]

    def bit(w, b):
        return int((int(w[b // 32]) >> (b % 32)) & 1)

    # isc_trace_t, most significant field first: timestamp[104:73] insn[72:41] cmd[40:38]
    # sync_id[37:34] rd_lock[33:27] wr_lock[26:20] issue_lock[19:13] vld[12] rd_pdg[11]
    # pdg[10] pe_reserved[9:0]
    for field, value, pos in [("pe_reserved", 1, 0), ("pdg", 1, 10), ("rd_pdg", 1, 11), ("vld", 1, 12),
                              ("issue_lock", 1, 13), ("wr_lock", 1, 20), ("rd_lock", 1, 27), ("sync_id", 1, 34),
                              ("cmd", 1, 38), ("insn", 1, 41), ("timestamp", 1, 73), ("timestamp", 1 << 31, 104)]:
        w = one(**{field: value})
        assert w.shape == (4,)
        assert [b for b in range(128) if bit(w, b)] == [pos], field
    evs = []
    for _ in range(50):
        evs.append({"timestamp": int(rng.integers(0, 2**32)), "insn": int(rng.integers(0, 2**32)),
                    "cmd": int(rng.integers(0, 8)), "sync_id": int(rng.integers(0, 16)),
                    "rd_lock": int(rng.integers(0, 128)), "wr_lock": int(rng.integers(0, 128)),
                    "issue_lock": int(rng.integers(0, 128)), "vld": int(rng.integers(0, 2)),
                    "rd_pdg": int(rng.integers(0, 2)), "pdg": int(rng.integers(0, 2)),
                    "pe_reserved": int(rng.integers(0, 1024))})
    back = la.trace_unpack(724, la.trace_pack(724, evs))
    assert back == evs


# ---------------------------------------------------------------- KSK image
def _tb_ksk_if_image(prm, lay, ksk):
    """The ksk_if testbench's image writer restated (tb_ksk_if.sv:40-43,63-98 for the per-channel
    cut counts, :470-525 for the loop order, the run padding and the AXI packing), with the
    coefficient of (line, level, column) read from this engine's [kN][L][n+1] layout through
    the header's reverse-on-N line order.  The testbench pads runs with all-ones dummies; the
    engine writes zeros, and so does this restatement."""
    W, lbx, lby, lbz, npc, cut_nb = prm.ks_mod_log, lay.lbx, lay.lby, lay.lbz, lay.ksk_pc, lay.ksk_cut_nb
    N, kN, L, n1 = prm.poly_size, prm.glwe_k * prm.poly_size, prm.ks_level, prm.lwe_dim + 1
    cols, lines, lgs = -(-n1 // lbx), -(-kN // lby), -(-L // lbz)
    acs = 64 if lbz * W > 32 else 32
    cpw = lay.axi_bytes * 8 // acs
    fcoef = -(-lby // cut_nb)                                     # KSK_CUT_FCOEF_NB
    dist, cnt, per_pc = -(-cut_nb // npc), cut_nb, []             # get_cut_per_pc
    for _ in range(npc):
        per_pc.append(min(cnt, dist))
        cnt -= dist
    ofs = [sum(per_pc[:i]) for i in range(npc)]                    # get_cut_ofs
    nb = [-(-(lby * lines * lgs * lbz) // (cut_nb // per_pc[i])) for i in range(npc)]  # get_bcol_coef
    rnd = [-(-nb[i] // (cpw * lbz)) * cpw * lbz for i in range(npc)]                   # round
    q = [[] for _ in range(npc)]
    for c in range(cols):
        for x in range(lbx):
            for ln in range(lines):
                for lg in range(lgs):
                    for y in range(lby):
                        for z in range(lbz):
                            line, lev, j = ln * lby + y, lg * lbz + z, c * lbx + x
                            v = 0
                            if line < kN and lev < L and j < n1:
                                g, m = divmod(line, N)
                                v = int(ksk[g * N + (N - 1 - m), lev, j])
                            for pc in range(npc):
                                if ofs[pc] * fcoef <= y < (ofs[pc] + per_pc[pc]) * fcoef:
                                    q[pc].append(v)
                                    if len(q[pc]) % rnd[pc] == nb[pc]:
                                        q[pc] += [0] * (rnd[pc] - nb[pc])
    out = []
    for pc in range(npc):
        words = []
        for w0 in range(0, len(q[pc]), cpw * lbz):
            word = 0
            for i in range(cpw):
                for z in range(lbz):
                    word |= q[pc][w0 + i * lbz + z] << (i * acs + z * W)
            words.append(word.to_bytes(lay.axi_bytes, "little"))
        out.append(np.frombuffer(b"".join(words), np.uint8))
    return out


@pytest.mark.parametrize("lay_kw", [dict(axi_bytes=64, ksk_pc=1, ksk_cut_nb=1, lbx=3, lby=64, lbz=3),
                                    dict(axi_bytes=32, ksk_pc=2, ksk_cut_nb=4, lbx=2, lby=16, lbz=1),
                                    dict(axi_bytes=16, ksk_pc=4, ksk_cut_nb=8, lbx=5, lby=32, lbz=2)])
def test_ksk_image_matches_ksk_if_testbench_order(lay_kw):
    prm = la.default_params(lwe_dim=10)
    lay = la.hpu_ksk_layout(**lay_kw)
    ksk = rng.integers(0, 1 << prm.ks_mod_log, prm.ksk_shape(), dtype=np.uint32)
    got = la.hpu_ksk_to_image(prm, lay, ksk)
    exp = _tb_ksk_if_image(prm, lay, ksk)
    assert len(got) == lay.ksk_pc
    for g, e in zip(got, exp):
        assert g.size == e.size == la.hpu_ksk_image_bytes(prm, lay)
        assert np.array_equal(g, e)
    assert np.array_equal(la.hpu_ksk_from_image(prm, lay, got), ksk)


def test_ksk_image_round_trip_and_limits():
    prm = la.default_params()  # n = 724, KS 2^3 x 5, w = 21
    lay = la.hpu_ksk_layout(axi_bytes=64, ksk_pc=2, ksk_cut_nb=2, lbx=3, lby=64, lbz=3)
    ksk = rng.integers(0, 1 << 21, prm.ksk_shape(), dtype=np.uint32)
    imgs = la.hpu_ksk_to_image(prm, lay, ksk)
    assert np.array_equal(la.hpu_ksk_from_image(prm, lay, imgs), ksk)
    # every coefficient appears once: the images hold the key's bits and zero padding only
    assert sum(int(np.unpackbits(i).sum()) for i in imgs) == sum(bin(int(v)).count("1") for v in ksk.ravel())
    for bad in (dict(lbz=4), dict(ksk_cut_nb=3), dict(axi_bytes=24), dict(lby=63), dict(lbx=0)):
        kw = dict(axi_bytes=64, ksk_pc=2, ksk_cut_nb=2, lbx=3, lby=64, lbz=3)
        kw.update(bad)
        assert la.hpu_ksk_image_bytes(prm, la.hpu_ksk_layout(**kw)) == 0, bad
