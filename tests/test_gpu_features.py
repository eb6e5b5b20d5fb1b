"""GPU: PBS level-2 parameters (base 2^15 x 2), Many-LUT PBS, batched ALU,
HIP-graph capture of an uneven batch program, IOp submission and error paths --
all bit-exact against the oracle through the C ABI."""
import os

import numpy as np
import pytest

import luthe_amd as la
from luthe_amd import dop

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def prm_l2(oracle):
    """PBS level-2 parameters: n=724, k=1, N=2048, PBS base 2^15 level 2, KS 2^3 x 5."""
    prm = la.default_params(pbs_base_log=15, pbs_level=2, lut_slots=512)
    ctx = la.Context(prm, device=0)
    keys = la.Keys(prm, seed=303)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    rp = oracle.params(pbs_base_log=15, pbs_level=2)
    bsk_ntt = oracle.bsk_to_ntt(rp, keys.bsk.reshape(-1), threads=THREADS)
    yield prm, ctx, keys, rp, bsk_ntt
    ctx.close()


def test_pbs_level2_vs_oracle(prm_l2, oracle):
    print("case: PBS L=2 B=2^15, 12 ciphertexts, per-ciphertext LUTs", flush=True)
    prm, ctx, keys, rp, bsk_ntt = prm_l2
    rng = np.random.default_rng(31)
    tables = rng.integers(0, 16, (12, 16), dtype=np.uint8)
    luts = np.stack([la.lut_encode(prm, t) for t in tables])
    ctx.write_lut(0, luts)
    msgs = rng.integers(0, 16, 12)
    cts = keys.encrypt_msg(msgs, seed=5)
    gids = np.arange(12, dtype=np.uint32)
    got = ctx.pbs_batch(cts, gids)
    exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, luts, gids, cts, threads=THREADS)
    assert np.array_equal(got, exp)
    assert (keys.decrypt_msg(got) == tables[gids, msgs]).mean() >= 0.75


@pytest.mark.parametrize("log_nb", [1, 2, 3])
def test_many_lut_vs_oracle(prm_l2, oracle, log_nb):
    print(f"case: Many-LUT log_lut_nb={log_nb}", flush=True)
    prm, ctx, keys, rp, bsk_ntt = prm_l2
    rng = np.random.default_rng(40 + log_nb)
    nb, tlen = 1 << log_nb, 16 >> log_nb
    tabs = rng.integers(0, 16, (nb, tlen), dtype=np.uint8)
    lut = la.lut_encode_many(prm, tabs, log_nb)
    ctx.write_lut(100, lut)
    msgs = rng.integers(0, tlen, 4)
    cts = keys.encrypt_msg(msgs, seed=log_nb)
    got = ctx.pbs_batch(cts, np.full(4, 100, np.uint32), log_lut_nb=log_nb)
    assert got.shape == (4 * nb, prm.blen)
    for c in range(4):
        _, a2n = oracle.keyswitch(rp, cts[c], keys.ksk)
        acc = oracle.blind_rotate(rp, bsk_ntt, lut, a2n)
        exp = oracle.sample_extract(1, 2048, acc, int(a2n[-1]), log_nb)
        assert np.array_equal(got[c * nb:(c + 1) * nb], exp), c
    dec = keys.decrypt_msg(got).reshape(4, nb)
    assert (dec == tabs[:, msgs].T).mean() >= 0.75


def test_alu_batch_vs_oracle(prm_l2, oracle):
    print("case: ALU batch, all opcodes", flush=True)
    prm, ctx, _, _, _ = prm_l2
    rng = np.random.default_rng(50)
    n = 33
    a = rng.integers(0, 2**64, (n, prm.blen), dtype=np.uint64)
    b = rng.integers(0, 2**64, (n, prm.blen), dtype=np.uint64)
    da, db = la.DeviceBuffer.from_numpy(a), la.DeviceBuffer.from_numpy(b)
    dout = la.DeviceBuffer((n, prm.blen), np.uint64)
    for op in (dop.ADD, dop.SUB, dop.MAC, dop.ADDS, dop.SUBS, dop.SSUB, dop.MULS):
        ctx.alu_batch_dev(op, da, db, dout, n, mul_factor=19, msg_cst=0x15)
        ctx.sync()
        got = dout.numpy()
        for c in (0, 17, 32):
            exp = oracle.alu(op, a[c], b[c], mul_factor=19, msg_cst=0x15)
            assert np.array_equal(got[c], exp), hex(op)


def test_uneven_batches_graph_replay(prm_l2, oracle):
    """The uneven-batch workload at reduced size: batches of uneven sizes, per-ciphertext LUTs, an
    ADD and a MAC before each PBS, captured once as a HIP graph and replayed three times
    with eager work in between; every replay equals the eager result, which equals the
    oracle on sampled ciphertexts."""
    print("case: uneven batches captured as a HIP graph", flush=True)
    prm, ctx, keys, rp, bsk_ntt = prm_l2
    rng = np.random.default_rng(60)
    sizes = [1, 37, 5, 120, 64, 3]
    total = sum(sizes)
    tables = rng.integers(0, 16, (64, 16), dtype=np.uint8)
    luts = np.stack([la.lut_encode(prm, t) for t in tables])
    ctx.write_lut(200, luts)
    ma, mb = rng.integers(0, 4, total), rng.integers(0, 4, total)
    A, B = keys.encrypt_msg(ma, seed=61), keys.encrypt_msg(mb, seed=62)
    gids = (200 + rng.integers(0, 64, total)).astype(np.uint32)
    dA, dB = la.DeviceBuffer.from_numpy(A), la.DeviceBuffer.from_numpy(B)
    dT, dU = la.DeviceBuffer((total, prm.blen), np.uint64), la.DeviceBuffer((total, prm.blen), np.uint64)
    dG = la.DeviceBuffer.from_numpy(gids)
    dO = la.DeviceBuffer((total, prm.blen), np.uint64)
    ctx.reserve(max(sizes))
    blen8 = prm.blen * 8

    def program():
        off = 0
        for s in sizes:
            a_p, b_p = dA.ptr.value + off * blen8, dB.ptr.value + off * blen8
            t_p, u_p, o_p = dT.ptr.value + off * blen8, dU.ptr.value + off * blen8, dO.ptr.value + off * blen8
            g_p = dG.ptr.value + off * 4
            ctx.alu_batch_dev(dop.ADD, la.C.c_void_p(a_p), la.C.c_void_p(b_p), la.C.c_void_p(t_p), s)
            ctx.alu_batch_dev(dop.MAC, la.C.c_void_p(t_p), la.C.c_void_p(b_p), la.C.c_void_p(u_p), s, mul_factor=1)
            ctx.pbs_batch_dev(la.C.c_void_p(u_p), la.C.c_void_p(o_p), la.C.c_void_p(g_p), s)
            off += s

    program()
    ctx.sync()
    eager = dO.numpy()
    ctx.capture_begin()
    program()
    g = ctx.capture_end()
    for rep in range(3):
        dO.zero()
        g.launch()
        ctx.sync()
        assert np.array_equal(dO.numpy(), eager), f"replay {rep}"
        # eager work between replays
        ctx.pbs_batch(A[:3], gids[:3])
    g.destroy()
    for c in (0, 1, 40, total - 1):
        t = oracle.alu(dop.ADD, A[c], B[c])
        u = oracle.alu(dop.MAC, t, B[c], mul_factor=1)
        exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, luts, np.array([gids[c] - 200], np.uint32), u.reshape(1, -1),
                               threads=1)[0]
        assert np.array_equal(eager[c], exp), c


def test_iop_submission(prm_l2, oracle):
    """An IOp with a registered DOp template program behaves like the patched DOps."""
    print("case: IOp path", flush=True)
    prm, ctx, keys, rp, bsk_ntt = prm_l2
    rng = np.random.default_rng(70)
    lut = la.lut_encode(prm, (np.arange(16) * 7 % 16).astype(np.uint8))
    ctx.write_lut(300, lut)
    # template: dst block j = PBS(src0 block j + src1 block j) for 2 blocks
    tpl = []
    for j in range(2):
        tpl += [dop.ld(0, dop.tslot(0, j), dop.MEM_SRC), dop.ld(1, dop.tslot(1, j), dop.MEM_SRC),
                dop.add(2, 0, 1), dop.pbs(3, 2, 300, flush=(j == 1)), dop.st(dop.tslot(0, j), 3, dop.MEM_DST)]
    ctx.register_iop(0x20, 2, tpl)
    x, y = keys.encrypt_msg([1, 2], seed=71), keys.encrypt_msg([3, 0], seed=72)
    ctx.write_ct(500, x)
    ctx.write_ct(600, y)
    t = ctx.submit_iop(dop.iop_words(0x20, [(700, 2)], [(500, 2), (600, 2)]))
    ctx.wait(t)
    out = ctx.read_ct(700, 2)
    for j in range(2):
        s = oracle.alu(dop.ADD, x[j], y[j])
        exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, lut.reshape(1, -1), np.zeros(1, np.uint32), s.reshape(1, -1),
                               threads=1)[0]
        assert np.array_equal(out[j], exp)
    assert ctx.counters()["iops"] == 1


def test_error_paths(oracle):
    print("case: error paths", flush=True)
    prm = la.default_params(lwe_dim=8)
    ctx = la.Context(prm, device=0)
    cts = np.zeros((2, prm.blen), np.uint64)
    with pytest.raises(la.LutheError) as e:          # keys missing
        ctx.pbs_batch(cts, np.zeros(2, np.uint32))
    assert e.value.code == -7
    with pytest.raises(la.LutheError) as e:          # unknown ticket
        ctx.wait(12345)
    assert e.value.code == -7
    with pytest.raises(la.LutheError) as e:          # malformed stream: nothing runs
        ctx.submit_dops([dop.ld(0, 1), 0x3F << 26])
    assert e.value.code == -5
    with pytest.raises(la.LutheError) as e:          # slot out of range
        ctx.write_ct(4096, cts[:1])
    assert e.value.code == -6
    # a LUT id beyond the LUT memory is clamped, never read out of bounds
    keys = la.Keys(prm, seed=1)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    lut = la.lut_encode(prm, np.arange(16, dtype=np.uint8))
    ctx.write_lut(prm.lut_slots - 1, lut)
    ok = ctx.pbs_batch(keys.encrypt_msg([3]), np.array([prm.lut_slots - 1], np.uint32))
    clamped = ctx.pbs_batch(keys.encrypt_msg([3]), np.array([999999], np.uint32))
    assert np.array_equal(ok, clamped)
    ctx.close()


@pytest.mark.parametrize("msg,carry", [(1, 1), (3, 3)])
def test_alu_constant_other_block_widths(oracle, msg, carry):
    """ADDS/SUBS/SSUB place the constant in the top msg + carry + 1 bits of the body
    (pea_alu_core.sv:59, USEFUL_BIT of param_tfhe_pkg.sv:32): checked for blocks other than
    2+2 bits, on the batched ALU and on the DOp path, against the oracle."""
    useful = msg + carry + 1
    print(f"case: ALU constant, {msg}+{carry}-bit blocks (useful {useful} bits)", flush=True)
    prm = la.default_params(lwe_dim=16, msg_bits=msg, carry_bits=carry, ct_slots=64)
    ctx = la.Context(prm, device=0)
    try:
        rng = np.random.default_rng(60 + useful)
        n = 5
        a = rng.integers(0, 2**64, (n, prm.blen), dtype=np.uint64)
        b = rng.integers(0, 2**64, (n, prm.blen), dtype=np.uint64)
        da, db = la.DeviceBuffer.from_numpy(a), la.DeviceBuffer.from_numpy(b)
        dout = la.DeviceBuffer((n, prm.blen), np.uint64)
        cst = 0x5B3  # 11-bit field, wider than the useful width
        for op in (dop.ADDS, dop.SUBS, dop.SSUB, dop.MULS):
            ctx.alu_batch_dev(op, da, db, dout, n, msg_cst=cst)
            ctx.sync()
            got = dout.numpy()
            for c in range(n):
                exp = oracle.alu(op, a[c], None, msg_cst=cst, useful_bits=useful)
                assert np.array_equal(got[c], exp), (hex(op), c)
        # DOp path: LD, ADDS / SSUB on registers, ST
        ctx.write_ct(0, a[:2])
        words = [dop.ld(1, 0), dop.ld(2, 1), dop.adds(3, 1, cst), dop.ssub(4, 2, cst),
                 dop.st(10, 3), dop.st(11, 4)]
        ctx.wait(ctx.submit_dops(words))
        got = ctx.read_ct(10, 2)
        assert np.array_equal(got[0], oracle.alu(dop.ADDS, a[0], None, msg_cst=cst, useful_bits=useful))
        assert np.array_equal(got[1], oracle.alu(dop.SSUB, a[1], None, msg_cst=cst, useful_bits=useful))
    finally:
        ctx.close()


def test_pbs_batch_list(prm_l2, oracle):
    """luthe_pbs_batch_list: batches in separate allocations with their own LUT ids and Many-LUT
    counts in one launch equal luthe_pbs_batch on each batch (and the oracle on sampled rows),
    also replayed from a graph; a total below the descriptors' sum drops the trailing PBS
    (their outputs stay untouched) and a larger total runs nothing extra."""
    print("case: PBS batch list, 5 batches (1, 300, 7, 64, 129; Many-LUT 0/1/2), graph replay", flush=True)
    prm, ctx, keys, rp, bsk_ntt = prm_l2
    rng = np.random.default_rng(70)
    tables = rng.integers(0, 16, (40, 16), dtype=np.uint8)
    luts = np.stack([la.lut_encode(prm, t) for t in tables])
    ctx.write_lut(0, luts)
    spec = [(1, 0), (300, 0), (7, 2), (64, 1), (129, 0)]
    bufs, exp = [], []
    for k, (n, lg) in enumerate(spec):
        cts = keys.encrypt_msg(rng.integers(0, 16 >> lg, n), seed=700 + k)
        gids = rng.integers(0, 40, n).astype(np.uint32)
        d_in, d_gid = la.DeviceBuffer.from_numpy(cts), la.DeviceBuffer.from_numpy(gids)
        d_out = la.DeviceBuffer((n << lg, prm.blen), np.uint64)
        d_out.upload(np.full((n << lg, prm.blen), 0xA5A5A5A5A5A5A5A5, np.uint64))
        exp.append(ctx.pbs_batch(cts, gids, log_lut_nb=lg))
        bufs.append((cts, gids, d_in, d_gid, d_out))
    total = sum(n for n, _ in spec)
    descs = la.DeviceBuffer.from_numpy(la.batch_descs(
        [(b[2].ptr.value, b[4].ptr.value, b[3].ptr.value, n, lg) for b, (n, lg) in zip(bufs, spec)]))
    ctx.pbs_batch_list_dev(descs, len(spec), total)
    ctx.sync()
    for k in range(len(spec)):
        assert np.array_equal(bufs[k][4].numpy(), exp[k]), k
    cts, gids = bufs[1][0], bufs[1][1]
    rows = np.array([0, 151, 299])
    o = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, luts, gids[rows], cts[rows], threads=THREADS)
    assert np.array_equal(bufs[1][4].numpy()[rows], o)
    print("case: batch list captured in a graph, replayed twice", flush=True)
    for b in bufs:
        b[4].upload(np.zeros(b[4].shape, np.uint64))
    ctx.reserve(total + 64)
    ctx.capture_begin()
    ctx.pbs_batch_list_dev(descs, len(spec), total + 64)  # 64 slots no descriptor fills
    g = ctx.capture_end()
    for _ in range(2):
        g.launch()
        ctx.sync()
        for k in range(len(spec)):
            assert np.array_equal(bufs[k][4].numpy(), exp[k]), k
        ctx.pbs_batch(cts[:3], gids[:3])  # eager work between the replays
    g.destroy()
    print("case: batch list with a total below the sum", flush=True)
    last = bufs[-1][4]
    last.upload(np.zeros(last.shape, np.uint64))
    ctx.pbs_batch_list_dev(descs, len(spec), total - 29)
    ctx.sync()
    got = last.numpy()
    assert np.array_equal(got[:100], exp[-1][:100])
    assert not got[100:].any()
