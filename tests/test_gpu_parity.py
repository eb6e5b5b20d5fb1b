"""GPU parity: every HIP kernel path against the CPU oracle, bit for bit, through the C ABI.

Sizes are chosen so that the oracle finishes in seconds; full-size workloads are
checked through size-independent properties (round trips, batch independence).
"""
import os

import numpy as np
import pytest

import luthe_amd as la
from luthe_amd import dop

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _ctx(**over):
    prm = la.default_params(**over)
    return prm, la.Context(prm, device=0)


@pytest.fixture(scope="module")
def prm_l1(oracle):
    """PBS level-1 parameters: n=724, k=1, N=2048, PBS 2^23 x 1, KS 2^3 x 5, w_ks=21, seeded keys."""
    prm, ctx = _ctx()
    keys = la.Keys(prm, seed=11)
    ctx.upload_bsk(keys.bsk, la.BSK_COEF)
    ctx.upload_ksk(keys.ksk)
    rp = oracle.params()
    bsk_ntt = oracle.bsk_to_ntt(rp, keys.bsk.reshape(-1), threads=THREADS)
    yield prm, ctx, keys, rp, bsk_ntt
    ctx.close()


# ------------------------------------------------------------------ NTT core (4096-polynomial product kernel)
def test_ntt_forward_inverse_vs_oracle(oracle):
    print("case: NTT fwd/inv, 32 polys", flush=True)
    _, ctx = _ctx(lwe_dim=8)
    rng = np.random.default_rng(1)
    a = rng.integers(0, la.P, (32, 2048), dtype=np.uint64)
    a[0] = 0
    a[1] = la.P - 1
    a[2, :] = 0
    a[2, 0] = 1
    f = ctx.ntt_forward(a)
    for i in range(32):
        assert np.array_equal(f[i], oracle.ntt(a[i])), f"forward poly {i}"
    inv = ctx.ntt_inverse(f)
    assert np.array_equal(inv, a)
    x = rng.integers(0, la.P, (8, 2048), dtype=np.uint64)
    xi = ctx.ntt_inverse(x)
    for i in range(8):
        assert np.array_equal(xi[i], oracle.intt(x[i])), f"inverse poly {i}"
    ctx.close()


def test_ntt_polymul_vs_oracle(oracle):
    print("case: NTT polymul, 64 polys + odd count", flush=True)
    _, ctx = _ctx(lwe_dim=8)
    rng = np.random.default_rng(2)
    for nb in (64, 7):
        a = rng.integers(0, la.P, (nb, 2048), dtype=np.uint64)
        bn = rng.integers(0, la.P, (nb, 2048), dtype=np.uint64)
        got = ctx.ntt_polymul(a, bn)
        exp = oracle.ntt_polymul_batch(a, bn, threads=THREADS)
        assert np.array_equal(got, exp)
    ctx.close()


def test_ntt_polymul_full_4096(oracle):
    """The NTT product at full size: 4096 polynomials; 48 sampled rows against the oracle and
    the size-independent identity  polymul(a, NTT(1)) == a  on every row."""
    print("case: NTT polymul 4096 x 2048", flush=True)
    _, ctx = _ctx(lwe_dim=8)
    rng = np.random.default_rng(3)
    a = rng.integers(0, la.P, (4096, 2048), dtype=np.uint64)
    bn = rng.integers(0, la.P, (4096, 2048), dtype=np.uint64)
    got = ctx.ntt_polymul(a, bn)
    idx = rng.choice(4096, 48, replace=False)
    exp = oracle.ntt_polymul_batch(a[idx], bn[idx], threads=THREADS)
    assert np.array_equal(got[idx], exp)
    one = np.zeros(2048, np.uint64)
    one[0] = 1
    ones = np.tile(oracle.ntt(one), (4096, 1))
    assert np.array_equal(ctx.ntt_polymul(a, ones), a)
    ctx.close()


# ------------------------------------------------------------------ key switch
def test_keyswitch_vs_oracle(prm_l1, oracle):
    print("case: KS 40 ciphertexts, n=724 PBS 2^23 x 1 parameters", flush=True)
    prm, ctx, keys, rp, _ = prm_l1
    rng = np.random.default_rng(4)
    blwe = rng.integers(0, 2**64, (40, prm.blen), dtype=np.uint64)
    blwe[0] = 0
    blwe[1] = 2**64 - 1
    got = ctx.keyswitch_batch(blwe)
    for c in range(40):
        _, exp = oracle.keyswitch(rp, blwe[c], keys.ksk)
        assert np.array_equal(got[c], exp), f"ciphertext {c}"


# ------------------------------------------------------------------ blind rotation
def test_blind_rotate_vs_oracle(prm_l1, oracle):
    print("case: BR 6 ciphertexts, n=724 PBS 2^23 x 1 (incl. rotations 0, 2047, 2048, 4095)", flush=True)
    prm, ctx, keys, rp, bsk_ntt = prm_l1
    rng = np.random.default_rng(5)
    luts = rng.integers(0, 2**64, (2, prm.lut_len), dtype=np.uint64)
    ctx.write_lut(0, luts)
    a2n = rng.integers(0, 4096, (6, prm.lwe_dim + 1), dtype=np.uint16)
    a2n[0, :4] = [0, 2047, 2048, 4095]
    gids = np.array([0, 1, 0, 1, 1, 0], np.uint32)
    got = ctx.blind_rotate_batch(a2n, gids)
    for c in range(6):
        acc = oracle.blind_rotate(rp, bsk_ntt, luts[gids[c]], a2n[c])
        exp = oracle.sample_extract(1, 2048, acc, int(a2n[c, -1]), 0)[0]
        assert np.array_equal(got[c], exp), f"ciphertext {c}"


# ------------------------------------------------------------------ full PBS
def test_pbs_vs_oracle_and_decrypt(prm_l1, oracle):
    print("case: PBS 16 ciphertexts, n=724 PBS 2^23 x 1, per-ciphertext LUTs", flush=True)
    prm, ctx, keys, rp, bsk_ntt = prm_l1
    rng = np.random.default_rng(6)
    tables = rng.integers(0, 16, (4, 16), dtype=np.uint8)
    tables[0] = np.arange(16)
    luts = np.stack([la.lut_encode(prm, t) for t in tables])
    ctx.write_lut(0, luts)
    msgs = rng.integers(0, 16, 16)
    gids = rng.integers(0, 4, 16).astype(np.uint32)
    cts = keys.encrypt_msg(msgs, seed=13)
    got = ctx.pbs_batch(cts, gids)
    exp = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, luts, gids, cts, threads=THREADS)
    assert np.array_equal(got, exp)
    dec = keys.decrypt_msg(got)
    ok = (dec == tables[gids, msgs]).mean()
    assert ok >= 0.75, f"decryption success {ok}"


def test_pbs_batch_independence(prm_l1):
    """Same ciphertext, alone and inside a large batch: identical bits."""
    print("case: PBS alone vs in a batch of 1500", flush=True)
    prm, ctx, keys, _, _ = prm_l1
    lut = la.lut_encode(prm, np.arange(16, dtype=np.uint8))
    ctx.write_lut(0, lut)
    cts = keys.encrypt_msg(np.arange(1500) % 16, seed=21)
    alone = ctx.pbs_batch(cts[7:8], np.zeros(1, np.uint32))
    mix = ctx.pbs_batch(cts, np.zeros(1500, np.uint32))
    assert np.array_equal(mix[7], alone[0])
    dec = keys.decrypt_msg(mix)
    assert (dec == (np.arange(1500) % 16)).mean() > 0.8


# ------------------------------------------------------------------ DOp stream
def test_dop_stream_vs_oracle(prm_l1, oracle):
    print("case: DOp stream LD/ALU/PBS/ST", flush=True)
    prm, ctx, keys, rp, bsk_ntt = prm_l1
    lut = la.lut_encode(prm, (np.arange(16) * 5 % 16).astype(np.uint8))
    ctx.write_lut(3, lut)
    cts = keys.encrypt_msg([1, 2], seed=31)
    ctx.write_ct(10, cts)
    prog = [dop.ld(0, 10), dop.ld(1, 11), dop.mac(2, 0, 1, 2), dop.adds(3, 2, 1),
            dop.pbs(4, 3, 3, flush=True), dop.st(20, 4), dop.sub(5, 4, 0), dop.st(21, 5), dop.sync()]
    t = ctx.submit_dops(prog)
    ctx.wait(t)
    r2 = oracle.alu(dop.MAC, cts[0], cts[1], mul_factor=2)
    r3 = oracle.alu(dop.ADDS, r2, None, msg_cst=1)
    r4 = oracle.pbs_batch(rp, bsk_ntt, keys.ksk, lut.reshape(1, -1), np.zeros(1, np.uint32), r3.reshape(1, -1),
                          threads=1)[0]
    r5 = oracle.alu(dop.SUB, r4, cts[0])
    assert np.array_equal(ctx.read_ct(20)[0], r4)
    assert np.array_equal(ctx.read_ct(21)[0], r5)
