"""GPU parity at the benchmark shapes, bit for bit against the oracle through the C ABI.

* batched PBS + KS: one 8192-ciphertext launch at n=724, k=1, N=2048, PBS 2^23 x 1,
  KS 2^3 x 5 (the persistent kernel's eight rounds of 1024 workgroups), 32 oracle rows spread
  over all rounds, and four of them recomputed alone;
* uneven batches: the 64-batch program of the benchmark (sizes uniform in [1, 4096] from its
  seed, ADD and MAC before a PBS with per-ciphertext LUTs, PBS 2^15 x 2) captured as one HIP
  graph, oracle rows from 8 batches, and the same PBS as one batch list equal to it;
* serving: 32-block ADD, CMP_GT and MUL through the server at n=724, decrypted and compared
  alone against in a mix, and an 8-block ADD compared node by node with the oracle.

The reference compares the HPU's output ciphertexts bit for bit with its software model's
(hw/module/hpu/simu/rtl/tb_hpu.sv:1430-1470); here the model is the oracle.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import luthe_amd as la

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def test_batched_pbs_8192(oracle):
    print("case: batched PBS, 8192 ciphertexts, n=724 PBS 2^23x1 KS 2^3x5", flush=True)
    prm = la.default_params()
    keys = la.Keys(prm, seed=1234)
    ctx = la.Context(prm, device=0)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    lut = la.lut_encode(prm, np.arange(16, dtype=np.uint8))
    ctx.write_lut(0, lut)
    n = 8192
    msgs = np.random.default_rng(99).integers(0, 16, n)
    cts = keys.encrypt_msg(msgs, seed=4321)
    assert la.lib().luthe_set_br_latency_max(ctx.handle, 0) == 0  # the persistent kernel
    grid = ctx.br_grid()
    got = ctx.pbs_batch(cts, np.zeros(n, np.uint32))
    rounds = (n + grid - 1) // grid
    print(f"case: {grid} workgroups, {rounds} persistent rounds; oracle on 32 rows", flush=True)
    assert rounds >= 2
    rng = np.random.default_rng(5)
    pick = []
    for r in range(rounds):  # 32 rows spread over every round, first and last workgroup included
        lo, hi = r * grid, min(n, (r + 1) * grid)
        pick += [lo, hi - 1] + [int(x) for x in rng.integers(lo, hi, max(2, 32 // rounds - 2))]
    pick = np.array(sorted(set(pick))[:max(32, 4 * rounds)])
    assert pick.size >= 32
    rp = oracle.params()
    bsk_ntt = oracle.bsk_to_ntt(rp, keys.bsk.reshape(-1), threads=THREADS)
    exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, lut.reshape(1, -1), np.zeros(pick.size, np.uint32), cts[pick],
                           threads=THREADS)
    assert np.array_equal(got[pick], exp)
    assert (keys.decrypt_msg(got) == msgs).mean() > 0.9
    print("case: four rows of the 8192 batch alone", flush=True)
    alone_rows = pick[[0, pick.size // 3, 2 * pick.size // 3, pick.size - 1]]
    for c in alone_rows:
        alone = ctx.pbs_batch(cts[c:c + 1], np.zeros(1, np.uint32))
        assert np.array_equal(alone[0], got[c]), int(c)
    ctx.close()


def test_uneven_batches_benchmark_shape(oracle):
    seed = 3
    prm = la.default_params(pbs_base_log=15, pbs_level=2, lut_slots=4096)
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(1, 4097, 64)]
    total = sum(sizes)
    print(f"case: uneven batches, 64 batches / {total} PBS, PBS 2^15x2, one graph", flush=True)
    keys = la.Keys(prm, seed=5678)
    ctx = la.Context(prm, device=0)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    tables = np.random.default_rng(seed + 1).integers(0, 16, (4096, 16), dtype=np.uint8)
    luts = np.stack([la.lut_encode(prm, t) for t in tables])
    ctx.write_lut(0, luts)
    ma, mb = rng.integers(0, 4, total), rng.integers(0, 4, total)
    A, B = keys.encrypt_msg(ma, seed=11), keys.encrypt_msg(mb, seed=12)
    dA, dB = la.DeviceBuffer.from_numpy(A), la.DeviceBuffer.from_numpy(B)
    dT = la.DeviceBuffer((total, prm.blen), np.uint64)
    dU = la.DeviceBuffer((total, prm.blen), np.uint64)
    dO = la.DeviceBuffer((total, prm.blen), np.uint64)
    gids = np.concatenate([np.arange(s, dtype=np.uint32) for s in sizes])
    dG = la.DeviceBuffer.from_numpy(gids)
    ctx.reserve(max(sizes))
    b8 = prm.blen * 8
    ctx.capture_begin()
    off = 0
    for s in sizes:
        a_p, b_p, t_p = dA.ptr.value + off * b8, dB.ptr.value + off * b8, dT.ptr.value + off * b8
        u_p, o_p, g_p = dU.ptr.value + off * b8, dO.ptr.value + off * b8, dG.ptr.value + off * 4
        ctx.alu_batch_dev(0x01, a_p, b_p, t_p, s)
        ctx.alu_batch_dev(0x05, t_p, b_p, u_p, s, mul_factor=2)
        ctx.pbs_batch_dev(u_p, o_p, g_p, s)
        off += s
    g = ctx.capture_end()
    g.launch()
    ctx.sync()
    out = dO.numpy()
    g.destroy()
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    print("case: the same 64 batches' PBS as one batch list", flush=True)
    dO2 = la.DeviceBuffer((total, prm.blen), np.uint64)
    descs = la.DeviceBuffer.from_numpy(la.batch_descs(
        [(dU.ptr.value + int(o) * b8, dO2.ptr.value + int(o) * b8, dG.ptr.value + int(o) * 4, s, 0)
         for o, s in zip(starts, sizes)]))
    ctx.reserve(total)
    ctx.pbs_batch_list_dev(descs, len(sizes), total)
    ctx.sync()
    assert np.array_equal(dO2.numpy(), out)
    del dO2
    pick_b = np.random.default_rng(6).choice(64, 8, replace=False)
    rows = []
    for bi in pick_b:  # first and last ciphertext of 8 batches
        rows += [int(starts[bi]), int(starts[bi] + sizes[bi] - 1)]
    rows = np.array(sorted(set(rows)))
    print(f"case: oracle on {rows.size} rows of batches {sorted(int(b) for b in pick_b)}", flush=True)
    rp = oracle.params(pbs_base_log=15, pbs_level=2)
    bsk_ntt = oracle.bsk_to_ntt(rp, keys.bsk.reshape(-1), threads=THREADS)
    u = np.stack([oracle.alu(0x05, oracle.alu(0x01, A[r], B[r]), B[r], mul_factor=2) for r in rows])
    exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, luts, gids[rows], u, threads=THREADS)
    assert np.array_equal(out[rows], exp)
    dec = keys.decrypt_msg(out)
    assert (dec == tables[gids, 2 * (ma + mb) + mb]).mean() > 0.9
    ctx.close()


def _program_luts(prm):
    luts = []
    for lut in range(la.IOP_LUTS):
        t, lognb = la.iop_lut(lut)
        if lognb:
            luts.append(la.lut_encode_many(prm, t.reshape(1 << lognb, 16 >> lognb), lognb))
        else:
            luts.append(la.lut_encode(prm, t))
    return np.stack(luts)


def _oracle_program_levels(oracle, rp, bsk_ntt, ksk, luts, op, a_cts, b_cts):
    """An IOp program on the oracle, its PBS nodes batched by dependency level (the oracle's
    threads over single-LUT PBS, a thread pool over Many-LUT ones)."""
    info, nodes, outs = la.iop_program(op, a_cts.shape[0])
    v = [None] * info.rows
    for i in range(info.nblocks):
        v[info.a_row0 + i] = a_cts[i]
        v[info.b_row0 + i] = b_cts[i]
    todo = list(nodes)

    def ml(n):
        _, a2n = oracle.keyswitch(rp, v[n.src0], ksk)
        acc = oracle.blind_rotate(rp, bsk_ntt, luts[n.lut], a2n)
        return n, oracle.sample_extract(1, rp.poly_size, acc, int(a2n[-1]), n.log_lut_nb)

    with ThreadPoolExecutor(THREADS) as pool:
        while todo:
            ready = [n for n in todo if v[n.src0] is not None
                     and (n.is_pbs or n.opcode not in (0x01, 0x02, 0x05) or v[n.src1] is not None)]
            assert ready
            ready_ids = {id(n) for n in ready}
            todo = [n for n in todo if id(n) not in ready_ids]
            for n in ready:
                if not n.is_pbs:
                    v[n.dst] = oracle.alu(n.opcode, v[n.src0], v[n.src1], n.mul_factor, 0)
            single = [n for n in ready if n.is_pbs and n.log_lut_nb == 0]
            if single:
                res = oracle.pbs_batch(rp, bsk_ntt, ksk, luts, np.array([n.lut for n in single], np.uint32),
                                       np.stack([v[n.src0] for n in single]), threads=THREADS)
                for n, r in zip(single, res):
                    v[n.dst] = r
            for n, res in pool.map(ml, [n for n in ready if n.is_pbs and n.log_lut_nb]):
                for j in range(1 << n.log_lut_nb):
                    v[n.dst + j] = res[j]
    return np.stack([v[r] for r in outs])


def _expected(op, a, b, nb):
    m = (1 << (2 * nb)) - 1
    if op == la.IOP_ADD:
        return la.to_blocks((a + b) & m, nb)
    if op == la.IOP_MUL:
        return la.to_blocks((a * b) & m, nb)
    return np.array([1 if a > b else 0], np.uint32)


def test_serving_32_blocks_and_oracle_program(oracle):
    """n = 724 with low key-switch noise (so that the 1,733-PBS MUL decrypts exactly)."""
    prm = la.default_params(lwe_noise_std=2.0**-25)
    ctx = la.Context(prm, device=0)
    keys = la.Keys(prm, seed=717)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    rng = np.random.default_rng(718)
    jobs = []
    for k, op in enumerate((la.IOP_ADD, la.IOP_CMP_GT, la.IOP_MUL)):
        a, b = (int(x) for x in rng.integers(0, 1 << 62, 2))
        a |= int(rng.integers(0, 4)) << 62
        b |= int(rng.integers(0, 4)) << 62
        jobs.append((op, 32, a, b, keys.encrypt_msg(la.to_blocks(a, 32), seed=40 + k),
                     keys.encrypt_msg(la.to_blocks(b, 32), seed=50 + k)))
    filler = [(la.IOP_ADD, 16), (la.IOP_MUL, 8), (la.IOP_CMP_GT, 16), (la.IOP_MUL, 16)]
    with la.Server(ctx) as srv:
        alone = []
        for op, nb, a, b, ca, cb in jobs:
            print(f"case: server, {la.IOP_NAMES[op]}32 alone", flush=True)
            alone.append(srv.wait(srv.submit(op, ca, cb), timeout_us=120_000_000)[0])
        print("case: server, the three 32-block IOps in a mix of 7", flush=True)
        ids = [srv.submit(op, ca, cb) for op, nb, a, b, ca, cb in jobs]
        for op, nb in filler:
            x = keys.encrypt_msg(la.to_blocks(int(rng.integers(0, 1 << 30)), nb), seed=60 + nb)
            ids.append(srv.submit(op, x, x))
        res = [srv.wait(i, timeout_us=120_000_000) for i in ids]
        for (op, nb, a, b, _, _), al, (mix, info) in zip(jobs, alone, res[:3]):
            assert np.array_equal(al, mix), la.IOP_NAMES[op]
            assert np.array_equal(keys.decrypt_msg(mix), _expected(op, a, b, nb)), la.IOP_NAMES[op]
            assert info["npbs"] == la.iop_program(op, nb)[0].npbs
        print("case: server ADD8 at n=724 vs the oracle, node by node", flush=True)
        a, b = (int(x) for x in rng.integers(0, 1 << 16, 2))
        ca = keys.encrypt_msg(la.to_blocks(a, 8), seed=70)
        cb = keys.encrypt_msg(la.to_blocks(b, 8), seed=71)
        got, _ = srv.wait(srv.submit(la.IOP_ADD, ca, cb), timeout_us=120_000_000)
    rp = oracle.params()  # the oracle computes on given keys and ciphertexts: noise plays no part
    bsk_ntt = oracle.bsk_to_ntt(rp, keys.bsk.reshape(-1), threads=THREADS)
    exp = _oracle_program_levels(oracle, rp, bsk_ntt, keys.ksk, _program_luts(prm), la.IOP_ADD, ca, cb)
    assert np.array_equal(got, exp)
    assert la.from_blocks(keys.decrypt_msg(got)) == (a + b) & 0xFFFF
    ctx.close()
