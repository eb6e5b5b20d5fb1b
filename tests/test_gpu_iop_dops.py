"""GPU: the radix IOp programs emitted as DOp templates, registered in the translation
table and run through luthe_submit_iop (the firmware path), give the same ciphertexts as
the serving engine for the same inputs and decrypt to the integer results."""
import time

import numpy as np
import pytest

import luthe_amd as la
from luthe_amd import dop

pytestmark = pytest.mark.gpu


def _expected(op, a, b, nb):
    m = (1 << (2 * nb)) - 1
    if op in (la.IOP_ADD, la.IOP_ADD_RIPPLE):
        return la.to_blocks((a + b) & m, nb)
    if op == la.IOP_MUL:
        return la.to_blocks((a * b) & m, nb)
    return np.array([1 if a > b else 0], np.uint32)


def test_registered_programs_match_server():
    prm = la.default_params(lwe_noise_std=2.0**-25)
    ctx = la.Context(prm, device=0)
    keys = la.Keys(prm, seed=818)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    lut_base = prm.lut_slots - la.IOP_LUTS  # the server's default LUT slots
    ctx.write_iop_luts(lut_base)
    rng = np.random.default_rng(19)
    jobs = []
    for op, nb in [(o, n) for o in (la.IOP_ADD, la.IOP_CMP_GT, la.IOP_MUL, la.IOP_ADD_RIPPLE) for n in (4, 8)] + [
            (la.IOP_MUL, 16)]:  # MUL on 16 blocks spills to 192 heap slots
        if True:
            a, b = (int(x) for x in rng.integers(0, 1 << (2 * nb), 2))
            jobs.append((op, nb, a, b, keys.encrypt_msg(la.to_blocks(a, nb), seed=10 * op + nb),
                         keys.encrypt_msg(la.to_blocks(b, nb), seed=100 + 10 * op + nb)))
    got_dop = []
    for k, (op, nb, a, b, ca, cb) in enumerate(jobs):
        print(f"case: DOp-template IOp {la.IOP_NAMES[op]} on {nb} blocks", flush=True)
        ctx.register_iop_program(op, nb, lut_base)
        _, st = la.iop_program_dops(op, nb, prm.regfile_size, lut_base)
        src_a, src_b, dst = 100 * k, 100 * k + 40, 100 * k + 80
        ctx.write_ct(src_a, ca)
        ctx.write_ct(src_b, cb)
        words = dop.iop_words(st["iop_opcode"], [(dst, 1 if op == la.IOP_CMP_GT else nb)],
                              [(src_a, nb), (src_b, nb)], align=(st["align_bits"], st["align_bits"]))
        t0 = time.perf_counter()
        ctx.wait(ctx.submit_iop(words))
        ms = (time.perf_counter() - t0) * 1e3
        out = ctx.read_ct(dst, 1 if op == la.IOP_CMP_GT else nb)
        print(f"    {st['pbs']} PBS, {st['dops']} DOps, {ms:.1f} ms", flush=True)
        assert np.array_equal(ctx.read_ct(src_a, nb), ca)  # sources untouched
        got_dop.append(out)
    with la.Server(ctx) as srv:
        print("case: the same IOps through the server", flush=True)
        ids = [srv.submit(op, ca, cb) for op, nb, a, b, ca, cb in jobs]
        got_srv = [srv.wait(i)[0] for i in ids]
    for (op, nb, a, b, _, _), x, y in zip(jobs, got_dop, got_srv):
        assert np.array_equal(x, y), (la.IOP_NAMES[op], nb)
        assert np.array_equal(keys.decrypt_msg(x), _expected(op, a, b, nb)), (la.IOP_NAMES[op], nb)
    ctx.close()
