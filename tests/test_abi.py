"""C ABI without a GPU: library load and exports, parameter handling, client-side
helpers, and the host DOp/IOp front end (checked against the reference's own
patch_dop compiled from fw/ublaze/src/hpu_dop_fmt.h into oracle/_ref/)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import luthe_amd as la
from luthe_amd import dop
from oracle import ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
rng = np.random.default_rng(4242)


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "luthe.h")).read()
    declared = sorted(set(re.findall(r"\b(luthe_[a-z0-9_]+)\s*\(", hdr)))
    assert len(declared) > 30
    L = la.lib()
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing


def test_params_and_create_without_gpu():
    p = la.default_params()
    assert (p.lwe_dim, p.glwe_k, p.poly_size, p.pbs_base_log, p.pbs_level) == (724, 1, 2048, 23, 1)
    assert (p.ks_base_log, p.ks_level, p.ks_mod_log, p.regfile_size, p.ct_slots) == (3, 5, 21, 64, 4096)
    if la.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(la.LutheError) as e:
        la.Context(p)
    assert e.value.code == -9  # LUTHE_ERR_NODEV: no CPU fallback
    bad = la.default_params(poly_size=1024)
    h = C.c_void_p()
    assert la.lib().luthe_create(C.byref(bad), 0, C.byref(h)) == -2


def test_keygen_is_deterministic_and_decrypts():
    p = la.default_params(lwe_dim=32)
    k1 = la.Keys(p, seed=9)
    k2 = la.Keys(p, seed=9)
    k3 = la.Keys(p, seed=10)
    assert np.array_equal(k1.bsk, k2.bsk) and np.array_equal(k1.ksk, k2.ksk)
    assert not np.array_equal(k1.lwe_sk, k3.lwe_sk)
    assert np.all(k1.bsk < np.uint64(la.P))
    assert np.all(k1.ksk < (1 << 21))
    msgs = rng.integers(0, 32, 200)
    cts = k1.encrypt_msg(msgs, seed=3)
    assert np.array_equal(k1.decrypt_msg(cts), msgs)


def test_bsk_rows_decrypt_to_gadget_multiples():
    """GGSW row (g, l) of key bit s_i: phase = -s_i * S_g * gadget_l (mask rows) or
    s_i * gadget_l (body row), gadget_l = round(p / 2^(b (L - l))), plus small noise."""
    p = la.default_params(lwe_dim=4, pbs_level=2, pbs_base_log=15)
    keys = la.Keys(p, seed=77)
    P = la.P
    S = keys.glwe_sk.astype(object)
    for i in range(4):
        for g in range(2):
            for l in range(2):
                row = keys.bsk[i, g, l]
                A = [int(x) for x in row[0]]
                B = [int(x) for x in row[1]]
                # phase = B - A * S (negacyclic, mod p)
                AS = ref.negacyclic_mul(np.array(A, np.uint64), keys.glwe_sk.astype(np.uint64))
                phase = [(b - int(x)) % P for b, x in zip(B, AS)]
                gad = (P + (1 << (15 * (2 - l) - 1))) >> (15 * (2 - l))
                si = int(keys.lwe_sk[i])
                expect = [0] * 2048
                if g == 1:
                    expect[0] = si * gad % P
                else:
                    expect = [(-si * gad * int(S[m])) % P for m in range(2048)]
                noise = [((ph - e + P // 2) % P) - P // 2 for ph, e in zip(phase, expect)]
                assert max(abs(x) for x in noise) < 2**20, (i, g, l)


def test_lut_encoding_layout():
    p = la.default_params()
    table = np.arange(16, dtype=np.uint8)[::-1].copy()
    lut = la.lut_encode(p, table)
    d = la.delta(p)
    assert np.all(lut[:2048] == 0)
    body = lut[2048:]
    assert int(body[0]) == int(table[0]) * d
    assert int(body[64]) == int(table[1]) * d and int(body[63]) == int(table[0]) * d
    assert int(body[2047]) == (-int(table[0]) * d) % 2**64
    tabs = np.array([[1, 2, 3, 0, 1, 1, 2, 2], [3, 3, 1, 2, 0, 0, 1, 1]], np.uint8)
    many = la.lut_encode_many(p, tabs, 1)
    assert int(many[2048 + 0]) == 1 * d and int(many[2048 + 1024]) == 3 * d
    assert int(many[2048 + 1024 - 1]) == 3 * d      # top half box of segment 0 -> f1(0)
    assert int(many[2048 + 2047]) == (-1 * d) % 2**64  # negacyclic wrap -> -f0(0)
    assert int(many[2048 + 1024 + 128]) == 3 * d and int(many[2048 + 128 * 3]) == 0


# ---------------------------------------------------------------- DOp front end
def test_dop_decode_validation():
    ok = [dop.ld(0, 5), dop.pbs(2, 0, 7), dop.st(9, 2), dop.sync()]
    n, lv = la.dop_schedule(ok)
    assert n == 3 and list(lv) == [0, 1, 2, -1]
    bad_cases = {
        -5: [dop.ld(0, 5, mode=dop.MEM_SRC)],          # template in a concrete stream
        -6: [dop.ld(70, 5)],                           # register beyond 64
    }
    for code, prog in bad_cases.items():
        with pytest.raises(la.LutheError) as e:
            la.dop_schedule(prog)
        assert e.value.code == code
    with pytest.raises(la.LutheError):
        la.dop_schedule([0x3F << 26])                  # unknown opcode
    with pytest.raises(la.LutheError):
        la.dop_schedule([dop.pbs(3, 0, 1, log_lut_nb=2)])  # ML4 destination not aligned
    with pytest.raises(la.LutheError):
        la.dop_schedule([dop.pbs(2, 0, 300)])          # LUT id beyond 256 slots
    with pytest.raises(la.LutheError):
        la.dop_schedule([dop.ld(0, 5000)])             # slot beyond 4096
    with pytest.raises(la.LutheError):
        la.dop_schedule([dop.arith_msg(dop.ADDS, 0, 1, 3, mode=dop.IMM_VAR)])


def test_dop_schedule_respects_hazards():
    prog = [dop.ld(0, 1), dop.ld(1, 2),
            dop.pbs(2, 0, 0), dop.pbs(3, 1, 0),      # independent PBS: same level
            dop.add(4, 2, 3),                        # RAW on both
            dop.ld(0, 3),                            # WAR on r0 (read by the first PBS)
            dop.st(1, 4),                            # WAR on slot 1 (read by the first LD)
            dop.pbs(4, 4, 1),                        # RAW+WAW on r4 (in place)
            dop.pbs(8, 0, 2, log_lut_nb=2)]          # ML4 writes r8..r11, reads r0
    n, lv = la.dop_schedule(prog)
    lv = list(lv)
    assert lv[2] == lv[3] == 1
    assert lv[4] == 2
    assert lv[5] > lv[2]
    assert lv[6] > lv[0] and lv[6] > lv[4]
    assert lv[7] == max(lv[4], lv[6]) + 1           # WAR: the ST reads r4 first
    assert lv[8] > lv[5]
    # a serial chain has one level per DOp
    chain = [dop.ld(0, 0)] + [dop.pbs(0, 0, 0) for _ in range(5)] + [dop.st(0, 0)]
    n, lv = la.dop_schedule(chain)
    assert n == 7 and list(lv) == list(range(7))


def _ref_patch():
    path = os.path.join(ROOT, "oracle", "_ref", "libdopfmt_ref.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    R = C.CDLL(path)
    R.dopref_patch.restype = C.c_uint32
    R.dopref_patch.argtypes = [C.c_uint32, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    R.dopref_heap_start.restype = C.c_int
    return R


def test_iop_translation_matches_reference_patch_dop():
    R = _ref_patch()
    assert R.dopref_heap_start() == 4095
    for trial in range(200):
        ndst, nsrc, nimm = rng.integers(1, 4), rng.integers(1, 4), rng.integers(0, 3)
        nb = int(rng.integers(1, 17))
        dsts = [(int(rng.integers(0, 3000)), nb) for _ in range(ndst)]
        srcs = [(int(rng.integers(0, 3000)), nb) for _ in range(nsrc)]
        imms = [(int(rng.integers(0, 2**32)), nb) for _ in range(nimm)]
        words = dop.iop_words(0x10, dsts, srcs, imms)
        tpl = []
        for _ in range(12):
            kind = rng.integers(0, 5)
            if kind == 0:
                tpl.append(dop.ld(int(rng.integers(0, 64)), dop.tslot(int(rng.integers(0, nsrc)), int(rng.integers(0, nb))), dop.MEM_SRC))
            elif kind == 1:
                tpl.append(dop.st(dop.tslot(int(rng.integers(0, ndst)), int(rng.integers(0, nb))), int(rng.integers(0, 64)), dop.MEM_DST))
            elif kind == 2:
                tpl.append(dop.ld(int(rng.integers(0, 64)), int(rng.integers(0, 64)), dop.MEM_HEAP))
            elif kind == 3 and nimm:
                tpl.append(dop.arith_msg(dop.ADDS, int(rng.integers(0, 64)), int(rng.integers(0, 64)),
                                         (int(rng.integers(0, nimm)) << 8) | int(rng.integers(0, nb)), dop.IMM_VAR))
            else:
                tpl.append(dop.pbs(int(rng.integers(0, 64)), int(rng.integers(0, 64)), int(rng.integers(0, 256))))
        got = la.iop_translate(words, tpl, ct_slots=4096)
        dst_c = np.array([c for c, _ in dsts], np.uint16)
        src_c = np.array([c for c, _ in srcs], np.uint16)
        imm = np.zeros((max(nimm, 1), 6), np.uint16)
        for i, (val, n_) in enumerate(imms):
            for j in range(0, n_, 8):
                m = 0
                for e in range(8):
                    if j + e < n_:
                        m |= ((val >> (2 * (j + e))) & 3) << (2 * e)
                imm[i, j // 8] = m
        for w, g in zip(tpl, got):
            exp = R.dopref_patch(w, dst_c.ctypes.data_as(C.c_void_p), ndst, src_c.ctypes.data_as(C.c_void_p), nsrc,
                                 imm.ctypes.data_as(C.c_void_p), nimm)
            assert int(g) == exp, (hex(w), hex(int(g)), hex(exp))


def test_iop_translation_rejects_out_of_range_templates():
    words = dop.iop_words(0x10, [(100, 4)], [(200, 4)])
    with pytest.raises(la.LutheError):
        la.iop_translate(words, [dop.ld(0, dop.tslot(1, 0), dop.MEM_SRC)])   # operand 1 absent
    with pytest.raises(la.LutheError):
        la.iop_translate(words, [dop.ld(0, dop.tslot(0, 9), dop.MEM_SRC)])   # block beyond width
    with pytest.raises(la.LutheError):
        la.iop_translate(words[:-1], [dop.sync()])                           # truncated IOp


def _exec_in_order(words, regs, slots_init):
    """Reference semantics of a concrete DOp stream on plaintext integers (ALU mod 2^16,
    PBS = x -> (5x + gid) mod 2^16 per output j, + j)."""
    R = {r: 1000 + r for r in range(regs)}
    S = dict(slots_init)
    for w in words:
        op = w >> 26
        if op in (0x01, 0x02, 0x05):
            d, a, b, mf = w & 127, (w >> 7) & 127, (w >> 14) & 127, (w >> 21) & 31
            R[d] = (R[a] + R[b]) % 65536 if op == 1 else ((R[a] - R[b]) % 65536 if op == 2 else (R[a] * mf + R[b]) % 65536)
        elif op in (0x09, 0x0A):
            d, a, cst = w & 127, (w >> 7) & 127, (w >> 15) & 2047
            R[d] = (R[a] + cst) % 65536 if op == 0x09 else (R[a] - cst) % 65536
        elif op == 0x20:
            R[w & 127] = S[(w >> 10) & 0xFFFF]
        elif op == 0x21:
            S[(w >> 10) & 0xFFFF] = R[w & 127]
        elif (op & 0x34) == 0x30:
            d, s, gid, nb = w & 127, (w >> 7) & 127, (w >> 14) & 4095, 1 << (op & 3)
            x = R[s]
            for j in range(nb):
                R[d + j] = (5 * x + gid + j) % 65536
    return R, S


def _exec_plan(words, regs, nslots, slots_init):
    """Executes the renamed plan level by level the way the GPU does (copies, then ALU,
    then PBS of a level; then the copy-back)."""
    nlev, ent, cb, rows = la.dop_plan(words, regs, nslots, 256)
    mem = {r: 1000 + r for r in range(regs)}
    for s, v in slots_init.items():
        mem[regs + s] = v
    for L in range(nlev):
        idx = [i for i in range(len(words)) if ent[i, 0] == L]
        for kind in ("mem", "alu", "pbs"):
            upd, reads = {}, set()
            for i in idx:
                w, (_, d, a, b) = words[i], ent[i]
                op = w >> 26
                if kind == "mem" and op in (0x20, 0x21):
                    upd[d] = mem[a]
                    reads.add(int(a))
                elif kind == "alu" and op in (0x01, 0x02, 0x05, 0x09, 0x0A):
                    mf, cst = (w >> 21) & 31, (w >> 15) & 2047
                    upd[d] = {0x01: mem[a] + mem[b], 0x02: mem[a] - mem[b], 0x05: mem[a] * mf + mem[b],
                              0x09: mem[a] + cst, 0x0A: mem[a] - cst}[op] % 65536
                    reads.update((int(a), int(b)) if op in (0x01, 0x02, 0x05) else (int(a),))
                elif kind == "pbs" and (op & 0x34) == 0x30:
                    for j in range(1 << (op & 3)):
                        upd[d + j] = (5 * mem[a] + ((w >> 14) & 4095) + j) % 65536
                    reads.add(int(a))
            # one launch never reads a row another op of the same launch writes
            assert not (reads & set(upd)), (L, kind)
            mem.update(upd)
    for r, row in cb:
        mem[int(r)] = mem[int(row)]
    return nlev, {r: mem[r] for r in range(regs)}, {s: mem[regs + s] for s in range(nslots) if regs + s in mem}


def test_renamed_plan_equals_in_order_execution():
    """Random DOp streams over few registers and slots (dense hazards): the executor's
    renamed level plan gives the same registers and slots as in-order execution, and its
    level count never exceeds the hazard-ordered schedule's."""
    r = np.random.default_rng(31337)
    for trial in range(300):
        regs, nslots = 8, 6
        words = []
        for _ in range(int(r.integers(5, 60))):
            k = int(r.integers(0, 6))
            if k == 0:
                words.append(dop.arith(int(r.choice([1, 2, 5])), int(r.integers(0, regs)), int(r.integers(0, regs)),
                                       int(r.integers(0, regs)), int(r.integers(0, 32))))
            elif k == 1:
                words.append(dop.arith_msg(int(r.choice([9, 10])), int(r.integers(0, regs)), int(r.integers(0, regs)),
                                           int(r.integers(0, 2048))))
            elif k == 2:
                words.append(dop.ld(int(r.integers(0, regs)), int(r.integers(0, nslots))))
            elif k == 3:
                words.append(dop.st(int(r.integers(0, nslots)), int(r.integers(0, regs))))
            elif k == 4:
                lognb = int(r.integers(0, 3))
                nb = 1 << lognb
                words.append(dop.pbs(int(r.integers(0, regs // nb)) * nb, int(r.integers(0, regs)),
                                     int(r.integers(0, 256)), lognb, flush=bool(r.integers(0, 2))))
            else:
                words.append(dop.sync())
        words = [int(w) & 0xFFFFFFFF for w in words]
        init = {s: 7 * s + 3 for s in range(nslots)}
        R0, S0 = _exec_in_order(words, regs, init)
        nlev, R1, S1 = _exec_plan(words, regs, nslots, init)
        assert R1 == R0, trial
        assert S1 == S0, trial
        nlev_hazard, _ = la.dop_schedule(words, regfile_size=regs, ct_slots=nslots, lut_slots=256)
        assert nlev <= nlev_hazard, trial


def test_keyswitch_parameter_limits():
    """Digits are int8 and the key-switch MACs are 24-bit: parameters outside that are refused
    (the reference's sets use KS_B_W 2-3, KS_L 5-8, MOD_KSK_W 21)."""
    h = C.c_void_p()
    # b*L > w: the gadget weight 2^(w - b*(L-l)) of the top level would need a negative shift
    for over in (dict(ks_base_log=8, ks_level=4), dict(ks_mod_log=25), dict(ks_mod_log=32), dict(ks_base_log=0),
                 dict(ks_base_log=7, ks_level=8, ks_mod_log=21), dict(ks_base_log=3, ks_level=8, ks_mod_log=23)):
        p = la.default_params(**over)
        assert la.lib().luthe_create(C.byref(p), 0, C.byref(h)) == -2, over
