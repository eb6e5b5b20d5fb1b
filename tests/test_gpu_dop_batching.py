"""GPU: continuous batching across DOp/IOp submissions.

Registered radix IOps (the firmware path: IOp words patched into DOp templates) are submitted
back to back; their PBS share batched launches across tickets, so the launches are far fewer
than the PBS levels of the IOps run one at a time, and every result is bit-identical to the
one-at-a-time run.  The launch rules: opportunistic by default, fill threshold with flush bit
and timeout, and host-driven launch on poll/wait (the BPIP start conditions of
pep_sequencer.sv:584-608)."""
import time

import numpy as np
import pytest

import luthe_amd as la
from luthe_amd import dop

pytestmark = pytest.mark.gpu


def _expected(op, a, b, nb):
    m = (1 << (2 * nb)) - 1
    if op == la.IOP_ADD:
        return la.to_blocks((a + b) & m, nb)
    if op == la.IOP_MUL:
        return la.to_blocks((a * b) & m, nb)
    return np.array([1 if a > b else 0], np.uint32)


@pytest.fixture(scope="module")
def env():
    prm = la.default_params(lwe_noise_std=2.0**-25)
    ctx = la.Context(prm, device=0)
    keys = la.Keys(prm, seed=919)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    lut_base = prm.lut_slots - la.IOP_LUTS
    ctx.write_iop_luts(lut_base)
    kinds = [(la.IOP_ADD, 8), (la.IOP_CMP_GT, 8), (la.IOP_MUL, 4), (la.IOP_ADD, 16)]
    stats = {}
    for op, nb in kinds:
        ctx.register_iop_program(op, nb, lut_base)
        stats[(op, nb)] = la.iop_program_dops(op, nb, prm.regfile_size, lut_base)[1]
    rng = np.random.default_rng(920)
    jobs = []
    for k in range(16):
        op, nb = kinds[k % len(kinds)]
        a, b = (int(x) for x in rng.integers(0, 1 << (2 * nb), 2))
        src_a, src_b, dst = 100 * k, 100 * k + 40, 100 * k + 80
        ctx.write_ct(src_a, keys.encrypt_msg(la.to_blocks(a, nb), seed=1000 + k))
        ctx.write_ct(src_b, keys.encrypt_msg(la.to_blocks(b, nb), seed=2000 + k))
        st = stats[(op, nb)]
        nout = 1 if op == la.IOP_CMP_GT else nb
        words = dop.iop_words(st["iop_opcode"], [(dst, nout)], [(src_a, nb), (src_b, nb)],
                              align=(st["align_bits"], st["align_bits"]))
        jobs.append((op, nb, a, b, dst, nout, words))
    yield prm, ctx, keys, jobs
    ctx.close()


def _clear_outputs(ctx, prm, jobs):
    for _, _, _, _, dst, nout, _ in jobs:
        ctx.write_ct(dst, np.zeros((nout, prm.blen), np.uint64))


def test_tickets_share_launches(env):
    prm, ctx, keys, jobs = env
    ctx.set_dop_batching(fill_pbs=0, timeout_us=0, max_groups=1)
    print("case: 16 registered IOps one at a time", flush=True)
    alone, levels = [], 0
    for op, nb, a, b, dst, nout, words in jobs:
        c0 = ctx.counters()
        ctx.wait(ctx.submit_iop(words))
        levels += ctx.counters()["pbs_batches"] - c0["pbs_batches"]
        alone.append(ctx.read_ct(dst, nout))
    total_pbs = sum(la.iop_program(op, nb)[0].npbs for op, nb, *_ in jobs)
    print(f"    {total_pbs} PBS in {levels} launches", flush=True)
    _clear_outputs(ctx, prm, jobs)
    print("case: the same 16 IOps back to back", flush=True)
    c0 = ctx.counters()
    tickets = [ctx.submit_iop(w) for *_, w in jobs]
    for t in tickets:
        ctx.wait(t)
    c1 = ctx.counters()
    launches = c1["pbs_batches"] - c0["pbs_batches"]
    print(f"    {c1['pbs'] - c0['pbs']} PBS in {launches} launches, {c1['groups'] - c0['groups']} groups",
          flush=True)
    assert c1["pbs"] - c0["pbs"] == total_pbs
    assert launches < levels
    for (op, nb, a, b, dst, nout, _), x in zip(jobs, alone):
        y = ctx.read_ct(dst, nout)
        assert np.array_equal(x, y), (la.IOP_NAMES[op], nb)
        assert np.array_equal(keys.decrypt_msg(y), _expected(op, a, b, nb)), (la.IOP_NAMES[op], nb)


def test_fill_threshold_flush_and_timeout(env):
    prm, ctx, keys, jobs = env
    ctx.write_lut(0, la.lut_encode(prm, (np.arange(16) * 5 % 16).astype(np.uint8)))
    msgs = np.arange(16) % 16
    ctx.write_ct(2000, keys.encrypt_msg(msgs, seed=76))
    streams = [[dop.ld(1, 2000 + k), dop.pbs(2, 1, 0), dop.pbs(3, 2, 0), dop.st(2100 + k, 3)] for k in range(16)]
    print("case: 16 two-PBS tickets one at a time", flush=True)
    ctx.set_dop_batching(fill_pbs=0, timeout_us=0, max_groups=1)
    c0 = ctx.counters()
    for w in streams:
        ctx.wait(ctx.submit_dops(w))
    assert ctx.counters()["pbs_batches"] - c0["pbs_batches"] == 32
    alone = ctx.read_ct(2100, 16)
    ctx.write_ct(2100, np.zeros((16, prm.blen), np.uint64))
    print("case: fill threshold: nothing launches until the host waits", flush=True)
    ctx.set_dop_batching(fill_pbs=1 << 30, timeout_us=0, max_groups=1)
    assert ctx.dop_batching() == {"fill_pbs": 1 << 30, "timeout_us": 0, "max_groups": 1}
    c0 = ctx.counters()
    tickets = [ctx.submit_dops(w) for w in streams]
    assert ctx.counters()["groups"] == c0["groups"]
    assert ctx.poll(tickets[-1]) in (0, 1)  # polling a ticket of the open group launches it
    ctx.wait(tickets[0])
    for t in tickets[1:]:
        ctx.wait(t)
    c1 = ctx.counters()
    assert c1["groups"] - c0["groups"] == 1
    assert c1["pbs_batches"] - c0["pbs_batches"] == 2  # 32 PBS, two dependency levels
    assert np.array_equal(ctx.read_ct(2100, 16), alone)
    assert np.array_equal(keys.decrypt_msg(alone), (msgs * 25) % 16)

    print("case: registered IOp templates carry the flush bit: the first launches at once", flush=True)
    c0 = ctx.counters()
    t = ctx.submit_iop(jobs[0][6])
    assert ctx.counters()["groups"] == c0["groups"] + 1
    ctx.wait(t)

    print("case: a PBS with the flush bit launches the open group", flush=True)
    ctx.write_lut(0, la.lut_encode(prm, np.arange(16, dtype=np.uint8)))
    ct = keys.encrypt_msg([5], seed=77)
    ctx.write_ct(3000, ct)
    c0 = ctx.counters()
    t1 = ctx.submit_dops([dop.ld(1, 3000), dop.pbs(2, 1, 0), dop.st(3001, 2)])
    assert ctx.counters()["groups"] == c0["groups"]  # below the threshold, no flush bit: waits
    t2 = ctx.submit_dops([dop.ld(3, 3000), dop.pbs(4, 3, 0, flush=True), dop.st(3002, 4)])
    assert ctx.counters()["groups"] == c0["groups"] + 1  # flushed: both tickets in one group
    ctx.wait(t1)
    ctx.wait(t2)
    assert ctx.counters()["pbs_batches"] == c0["pbs_batches"] + 1  # the two PBS shared one launch
    x, y = ctx.read_ct(3001, 1), ctx.read_ct(3002, 1)
    assert np.array_equal(x, y)

    print("case: the fill threshold's timeout", flush=True)
    ctx.set_dop_batching(fill_pbs=1 << 30, timeout_us=2000, max_groups=1)
    c0 = ctx.counters()
    t1 = ctx.submit_dops([dop.ld(1, 3000), dop.pbs(2, 1, 0), dop.st(3003, 2)])
    assert ctx.counters()["groups"] == c0["groups"]
    time.sleep(0.02)
    t2 = ctx.submit_dops([dop.ld(1, 3000), dop.pbs(2, 1, 0), dop.st(3004, 2)])  # the call that sees it
    assert ctx.counters()["groups"] == c0["groups"] + 1  # t1's group timed out; t2 opened a new one
    ctx.wait(t1)
    ctx.wait(t2)
    assert ctx.counters()["groups"] == c0["groups"] + 2
    assert np.array_equal(ctx.read_ct(3003, 1), x) and np.array_equal(ctx.read_ct(3004, 1), x)
    ctx.set_dop_batching()


def test_in_order_semantics_across_tickets(env):
    """A later ticket reads what an earlier one stored, and overwrites a slot an earlier one
    reads: merged into one group, the results are those of in-order execution."""
    prm, ctx, keys, jobs = env
    ctx.set_dop_batching(fill_pbs=1 << 30)
    ctx.write_lut(0, la.lut_encode(prm, (np.arange(16) * 3 % 16).astype(np.uint8)))
    ct = keys.encrypt_msg([3], seed=78)
    ctx.write_ct(3100, ct)
    t1 = ctx.submit_dops([dop.ld(1, 3100), dop.pbs(2, 1, 0), dop.st(3101, 2)])           # 3101 = f(3)
    t2 = ctx.submit_dops([dop.ld(1, 3101), dop.pbs(2, 1, 0), dop.st(3102, 2),             # 3102 = f(f(3))
                          dop.ld(5, 3100), dop.adds(5, 5, 1), dop.st(3100, 5)])            # 3100 += 1
    t3 = ctx.submit_dops([dop.ld(1, 3100), dop.st(3103, 1)])                              # the new 3100
    for t in (t1, t2, t3):
        ctx.wait(t)
    ctx.set_dop_batching()
    f = lambda m: (m * 3) % 16  # noqa: E731
    assert keys.decrypt_msg(ctx.read_ct(3101, 1))[0] == f(3)
    assert keys.decrypt_msg(ctx.read_ct(3102, 1))[0] == f(f(3))
    assert keys.decrypt_msg(ctx.read_ct(3103, 1))[0] == 4
