"""GPU: the serving engine (continuous batching of radix IOps).

* every IOp program run by the server equals the same program evaluated node by node on
  the CPU oracle (KS + blind rotation + sample extract + ALU), bit for bit;
* a request's result is the same alone and in a mix of concurrent requests (cohort
  composition does not change results);
* results decrypt to the integer semantics (low-noise keys);
* HBM admission (requests wait for arena space; oversize requests are refused), cohort
  fill threshold with timeout and explicit flush, and the error paths.
"""
import os
import time

import numpy as np
import pytest

import luthe_amd as la

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
OPS = [la.IOP_ADD, la.IOP_CMP_GT, la.IOP_MUL, la.IOP_ADD_RIPPLE]


def _program_luts(prm):
    luts = []
    for lut in range(la.IOP_LUTS):
        t, lognb = la.iop_lut(lut)
        if lognb:
            luts.append(la.lut_encode_many(prm, t.reshape(1 << lognb, 16 >> lognb), lognb))
        else:
            luts.append(la.lut_encode(prm, t))
    return np.stack(luts)


def _oracle_program(oracle, rp, bsk_ntt, ksk, luts, op, a_cts, b_cts):
    """Runs an IOp program node by node on the oracle."""
    info, nodes, outs = la.iop_program(op, a_cts.shape[0])
    v = [None] * info.rows
    for i in range(info.nblocks):
        v[info.a_row0 + i] = a_cts[i]
        v[info.b_row0 + i] = b_cts[i]
    for n in nodes:
        if n.is_pbs:
            if n.log_lut_nb == 0:
                v[n.dst] = oracle.pbs_batch(rp, bsk_ntt, ksk, luts, np.array([n.lut], np.uint32),
                                            v[n.src0][None, :], threads=1)[0]
            else:
                _, a2n = oracle.keyswitch(rp, v[n.src0], ksk)
                acc = oracle.blind_rotate(rp, bsk_ntt, luts[n.lut], a2n)
                res = oracle.sample_extract(1, rp.poly_size, acc, int(a2n[-1]), n.log_lut_nb)
                for j in range(1 << n.log_lut_nb):
                    v[n.dst + j] = res[j]
        else:
            v[n.dst] = oracle.alu(n.opcode, v[n.src0], v[n.src1], n.mul_factor, 0)
    return np.stack([v[r] for r in outs])


def _expected(op, a, b, nb):
    m = (1 << (2 * nb)) - 1
    if op in (la.IOP_ADD, la.IOP_ADD_RIPPLE):
        return la.to_blocks((a + b) & m, nb)
    if op == la.IOP_MUL:
        return la.to_blocks((a * b) & m, nb)
    return np.array([1 if a > b else 0], np.uint32)


@pytest.fixture(scope="module")
def small(oracle):
    """n = 128 LWE (fast oracle), low key-switch noise so results also decrypt exactly."""
    prm = la.default_params(lwe_dim=128, lwe_noise_std=2.0**-25)
    ctx = la.Context(prm, device=0)
    keys = la.Keys(prm, seed=515)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    rp = oracle.params(lwe_dim=128)
    bsk_ntt = oracle.bsk_to_ntt(rp, keys.bsk.reshape(-1), threads=THREADS)
    yield prm, ctx, keys, rp, bsk_ntt
    ctx.close()


@pytest.fixture(scope="module")
def full():
    """The PBS level-1 shapes (n = 724, N = 2048, PBS 2^23 x 1, KS 2^3 x 5) with low key-switch noise."""
    prm = la.default_params(lwe_noise_std=2.0**-25)
    ctx = la.Context(prm, device=0)
    keys = la.Keys(prm, seed=616)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    yield prm, ctx, keys
    ctx.close()


def test_server_programs_vs_oracle(small, oracle):
    prm, ctx, keys, rp, bsk_ntt = small
    luts = _program_luts(prm)
    rng = np.random.default_rng(7)
    with la.Server(ctx) as srv:
        assert srv.stats()["lut_base"] == prm.lut_slots - la.IOP_LUTS
        for op in OPS:
            for nb in (1, 2):
                print(f"case: server vs oracle, {la.IOP_NAMES[op]} on {nb} blocks", flush=True)
                a, b = (int(x) for x in rng.integers(0, 1 << (2 * nb), 2))
                ca = keys.encrypt_msg(la.to_blocks(a, nb), seed=int(rng.integers(1, 1 << 30)))
                cb = keys.encrypt_msg(la.to_blocks(b, nb), seed=int(rng.integers(1, 1 << 30)))
                got, info = srv.wait(srv.submit(op, ca, cb))
                lut_rows = luts  # luts indexed by logical LUT id, as the server's slots are
                exp = _oracle_program(oracle, rp, bsk_ntt, keys.ksk, lut_rows, op, ca, cb)
                assert np.array_equal(got, exp), (op, nb)
                assert np.array_equal(keys.decrypt_msg(got), _expected(op, a, b, nb)), (op, nb, a, b)
                assert info["status"] == 0 and info["npbs"] == la.iop_program(op, nb)[0].npbs


def test_server_alone_equals_mix_and_decrypts(full):
    prm, ctx, keys = full
    rng = np.random.default_rng(11)
    jobs = []
    for k in range(14):
        op = OPS[k % 4]
        nb = [4, 8, 16][k % 3]
        a, b = (int(x) for x in rng.integers(0, 1 << min(62, 2 * nb), 2))
        ca = keys.encrypt_msg(la.to_blocks(a, nb), seed=1000 + k)
        cb = keys.encrypt_msg(la.to_blocks(b, nb), seed=2000 + k)
        jobs.append((op, nb, a, b, ca, cb))
    with la.Server(ctx) as srv:
        print("case: request 5 alone", flush=True)
        op, nb, a, b, ca, cb = jobs[5]
        alone, _ = srv.wait(srv.submit(op, ca, cb))
        srv.stats(reset=True)
        print("case: 14 concurrent requests", flush=True)
        ids = [srv.submit(j[0], j[4], j[5]) for j in jobs]
        res = [srv.wait(i, timeout_us=120_000_000) for i in ids]
        st = srv.stats()
    assert np.array_equal(res[5][0], alone), "result changed when sharing cohorts"
    total_pbs = sum(la.iop_program(j[0], j[1])[0].npbs for j in jobs)
    assert st["pbs"] == total_pbs and st["completed"] == len(jobs)
    assert st["cohorts"] < total_pbs  # PBS of different requests were merged
    assert st["latency_p99_us"] >= st["latency_p50_us"] > 0
    for (op, nb, a, b, _, _), (out, info) in zip(jobs, res):
        assert np.array_equal(keys.decrypt_msg(out), _expected(op, a, b, nb)), (la.IOP_NAMES[op], nb)
        assert info["cohorts"] >= info["depth"]


def test_server_admission(full):
    prm, ctx, keys = full
    row_bytes = prm.blen * 8
    info8 = la.iop_program(la.IOP_MUL, 8)[0]
    rows8 = info8.rows + info8.nout
    arena_rows = rows8 + 40  # room for one 8-block MUL at a time
    rng = np.random.default_rng(12)
    with la.Server(ctx, arena_bytes=arena_rows * row_bytes) as srv:
        print("case: admission, 16-block MUL larger than the arena", flush=True)
        big = keys.encrypt_msg(np.zeros(16, np.uint32))
        with pytest.raises(la.LutheError) as e:
            srv.submit(la.IOP_MUL, big, big)
        assert e.value.code == -4
        print("case: admission, three 8-block MULs through a one-request arena", flush=True)
        vals = [(int(x), int(y)) for x, y in rng.integers(0, 1 << 16, (3, 2))]
        ids = [srv.submit(la.IOP_MUL, keys.encrypt_msg(la.to_blocks(a, 8), seed=3 + i),
                          keys.encrypt_msg(la.to_blocks(b, 8), seed=30 + i)) for i, (a, b) in enumerate(vals)]
        outs = [srv.wait(i, timeout_us=120_000_000) for i in ids]
        st = srv.stats()
    assert st["rejected"] == 1 and st["completed"] == 3
    assert st["arena_rows_peak"] <= arena_rows
    assert outs[2][1]["queue_us"] > 0
    for (a, b), (out, _) in zip(vals, outs):
        assert la.from_blocks(keys.decrypt_msg(out)) == (a * b) & 0xFFFF


def test_server_cohort_timeout_and_flush(full):
    prm, ctx, keys = full
    ca = keys.encrypt_msg(la.to_blocks(9, 2), seed=1)
    cb = keys.encrypt_msg(la.to_blocks(5, 2), seed=2)
    print("case: cohort fill threshold with a 20 ms timeout", flush=True)
    with la.Server(ctx, cohort_min=100000, cohort_timeout_us=20000) as srv:
        out, _ = srv.wait(srv.submit(la.IOP_ADD, ca, cb), timeout_us=60_000_000)
        st = srv.stats()
    assert st["cohorts_by_timeout"] == st["cohorts"] >= 2
    assert la.from_blocks(keys.decrypt_msg(out)) == 14
    print("case: cohort fill threshold without timeout, explicit flushes", flush=True)
    with la.Server(ctx, cohort_min=100000, cohort_timeout_us=0xFFFFFFFF) as srv:
        rid = srv.submit(la.IOP_ADD, ca, cb)
        with pytest.raises(la.LutheError) as e:
            srv.wait(rid, timeout_us=200_000)
        assert e.value.code == -8
        done = None
        for _ in range(40):
            srv.flush()
            time.sleep(0.05)
            done = srv.poll(rid)
            if done is not None:
                break
        assert done is not None
        st = srv.stats()
    assert np.array_equal(done[0], out)  # same keys and inputs: same ciphertext
    assert st["cohorts_by_timeout"] == 0 and st["cohorts"] >= 2


def test_server_errors(full):
    prm, ctx, keys = full
    ca = keys.encrypt_msg(np.zeros(2, np.uint32))
    with la.Server(ctx) as srv:
        print("case: server error paths", flush=True)
        with pytest.raises(la.LutheError) as e:
            srv.submit(7, ca, ca)
        assert e.value.code == -1
        with pytest.raises(la.LutheError) as e:
            la.Server(ctx)
        assert e.value.code == -7
        with pytest.raises(la.LutheError) as e:  # batch calls would share the cohorts' scratch
            ctx.pbs_batch(ca[:1], np.zeros(1, np.uint32))
        assert e.value.code == -10
        with pytest.raises(la.LutheError) as e:  # the server thread reads the keys and launch settings
            ctx.upload_ksk(keys.ksk)
        assert e.value.code == -10
        with pytest.raises(la.LutheError) as e:
            ctx.close()
        assert e.value.code == -7
        with pytest.raises(la.LutheError) as e:
            srv.wait(12345)
        assert e.value.code == -7
    # the context works normally once the server is gone
    out = ctx.pbs_batch(ca[:1], np.zeros(1, np.uint32))
    assert out.shape == (1, prm.blen)


def test_server_short_request_overtakes_long_ones(full):
    """Cohort policy SHORTEST (default): ready PBS with the fewest PBS levels left on their
    request's critical path go first, so an 8-block ADD submitted after four 32-block MULs
    (whose first level alone is ~4,200 PBS) completes before all of them; its bits equal the
    same request run alone, and the oldest-first policy gives the same bits."""
    prm, ctx, keys = full
    rng = np.random.default_rng(13)
    muls = []
    for k in range(4):
        a, b = (int(x) for x in rng.integers(0, 1 << 62, 2))
        muls.append((a, b, keys.encrypt_msg(la.to_blocks(a, 32), seed=500 + k),
                     keys.encrypt_msg(la.to_blocks(b, 32), seed=600 + k)))
    x, y = 12345, 54321
    cx, cy = keys.encrypt_msg(la.to_blocks(x, 8), seed=700), keys.encrypt_msg(la.to_blocks(y, 8), seed=701)
    results = {}
    for policy in (la.COHORT_SHORTEST, la.COHORT_OLDEST):
        with la.Server(ctx, policy=policy) as srv:
            print(f"case: policy {policy}: ADD8 alone", flush=True)
            alone, _ = srv.wait(srv.submit(la.IOP_ADD, cx, cy), timeout_us=60_000_000)
            print(f"case: policy {policy}: four MUL32 then ADD8", flush=True)
            ids = [srv.submit(la.IOP_MUL, ca, cb) for _, _, ca, cb in muls]
            add_id = srv.submit(la.IOP_ADD, cx, cy)
            add_out, add_info = srv.wait(add_id, timeout_us=120_000_000)
            mul_res = [srv.wait(i, timeout_us=120_000_000) for i in ids]
        assert np.array_equal(add_out, alone)
        assert la.from_blocks(keys.decrypt_msg(add_out)) == (x + y) & 0xFFFF
        for (a, b, _, _), (out, _) in zip(muls, mul_res):
            assert la.from_blocks(keys.decrypt_msg(out)) == (a * b) & ((1 << 64) - 1)
        results[policy] = (add_out, [m[0] for m in mul_res], add_info["latency_us"],
                           min(m[1]["latency_us"] for m in mul_res))
        print(f"    ADD8 {add_info['latency_us'] / 1e3:.1f} ms, first MUL32 "
              f"{results[policy][3] / 1e3:.1f} ms", flush=True)
    add_s, muls_s, lat_add, lat_mul = results[la.COHORT_SHORTEST]
    assert lat_add < lat_mul  # the short request finished first
    add_o, muls_o, _, _ = results[la.COHORT_OLDEST]
    assert np.array_equal(add_s, add_o)
    for p, q in zip(muls_s, muls_o):
        assert np.array_equal(p, q)


def test_dop_tickets_beside_a_server(full):
    """DOp/IOp tickets submitted while a server is attached run on the same stream, between its
    steps, with scratch of their own: a registered-IOp ticket and a raw DOp stream give the same
    ciphertexts as with no server, while radix requests keep flowing through the server."""
    prm, ctx, keys = full
    from luthe_amd import dop
    rng = np.random.default_rng(14)
    lut_base = prm.lut_slots - la.IOP_LUTS
    ctx.write_iop_luts(lut_base)
    ctx.register_iop_program(la.IOP_ADD, 8, lut_base)
    _, st = la.iop_program_dops(la.IOP_ADD, 8, prm.regfile_size, lut_base)
    a, b = (int(x) for x in rng.integers(0, 1 << 16, 2))
    ctx.write_ct(200, keys.encrypt_msg(la.to_blocks(a, 8), seed=900))
    ctx.write_ct(240, keys.encrypt_msg(la.to_blocks(b, 8), seed=901))
    iop = dop.iop_words(st["iop_opcode"], [(280, 8)], [(200, 8), (240, 8)], align=(st["align_bits"], st["align_bits"]))
    raw = [dop.ld(1, 200), dop.pbs(2, 1, lut_base), dop.st(300, 2)]
    print("case: IOp and DOp tickets without a server", flush=True)
    ctx.wait(ctx.submit_iop(iop))
    ctx.wait(ctx.submit_dops(raw))
    exp_iop, exp_raw = ctx.read_ct(280, 8), ctx.read_ct(300, 1)
    ctx.write_ct(280, np.zeros((8, prm.blen), np.uint64))
    ctx.write_ct(300, np.zeros((1, prm.blen), np.uint64))
    jobs = []
    for k in range(6):
        x, y = (int(v) for v in rng.integers(0, 1 << 32, 2))
        jobs.append((x, y, keys.encrypt_msg(la.to_blocks(x, 16), seed=910 + k),
                     keys.encrypt_msg(la.to_blocks(y, 16), seed=920 + k)))
    with la.Server(ctx) as srv:
        print("case: the same tickets while six MUL16 requests run through the server", flush=True)
        ids = [srv.submit(la.IOP_MUL, cx, cy) for _, _, cx, cy in jobs]
        t1 = ctx.submit_iop(iop)
        t2 = ctx.submit_dops(raw)
        ctx.wait(t1)
        ctx.wait(t2)
        res = [srv.wait(i, timeout_us=120_000_000) for i in ids]
    assert np.array_equal(ctx.read_ct(280, 8), exp_iop)
    assert np.array_equal(ctx.read_ct(300, 1), exp_raw)
    assert la.from_blocks(keys.decrypt_msg(exp_iop)) == (a + b) & 0xFFFF
    for (x, y, _, _), (out, _) in zip(jobs, res):
        assert la.from_blocks(keys.decrypt_msg(out)) == (x * y) & 0xFFFFFFFF
