"""DOp / IOp word encoders (formats of fw/ublaze/src/hpu_dop_fmt.h:50-116 and
fw/ublaze/src/hpu_iop_fmt.h:19-76; opcodes of hpu_common_instruction_pkg.sv:42-68)."""

ADD, SUB, MAC, ADDS, SUBS, SSUB, MULS = 0x01, 0x02, 0x05, 0x09, 0x0A, 0x0B, 0x0C
SYNC, LD, ST, PBS = 0x10, 0x20, 0x21, 0x30
MEM_ADDR, MEM_HEAP, MEM_SRC, MEM_DST = 0, 1, 2, 3
IMM_CST, IMM_VAR = 0, 1
SYNC_WORD = 0x4000FFFF  # ucore.h:43 SYNC_DOP_WORD


def arith(op, dst, src0, src1, mul_factor=0):
    return (op << 26) | ((mul_factor & 31) << 21) | ((src1 & 127) << 14) | ((src0 & 127) << 7) | (dst & 127)


def arith_msg(op, dst, src0, cst, mode=IMM_CST):
    return (op << 26) | ((cst & 2047) << 15) | ((mode & 1) << 14) | ((src0 & 127) << 7) | (dst & 127)


def add(d, a, b): return arith(ADD, d, a, b)
def sub(d, a, b): return arith(SUB, d, a, b)
def mac(d, a, b, f): return arith(MAC, d, a, b, f)
def adds(d, a, c): return arith_msg(ADDS, d, a, c)
def subs(d, a, c): return arith_msg(SUBS, d, a, c)
def ssub(d, a, c): return arith_msg(SSUB, d, a, c)
def muls(d, a, c): return arith_msg(MULS, d, a, c)


def mem(op, rid, slot, mode=MEM_ADDR):
    return (op << 26) | ((slot & 0xFFFF) << 10) | ((mode & 3) << 8) | (rid & 127)


def ld(rid, slot, mode=MEM_ADDR): return mem(LD, rid, slot, mode)
def st(slot, rid, mode=MEM_ADDR): return mem(ST, rid, slot, mode)
def tslot(tid, bid): return ((tid & 0xFF) << 8) | (bid & 0xFF)  # TS[tid].bid / TD[tid].bid


def pbs(dst, src, gid, log_lut_nb=0, flush=False):
    op = PBS | (8 if flush else 0) | (log_lut_nb & 3)
    return (op << 26) | ((gid & 4095) << 14) | ((src & 127) << 7) | (dst & 127)


def sync(sid=0xFFFF): return (SYNC << 26) | (sid & 0x3FFFFFF)


# ---- IOp words (hpu_iop_fmt.h)
KIND_SRC, KIND_DST, KIND_IMM = 0, 1, 2


def iop_header(opcode, dst_align, src_align, has_imm=False, fw_mode=0):
    return (src_align & 0xFF) | ((dst_align & 0xFF) << 8) | ((opcode & 0xFF) << 16) | \
        ((1 if has_imm else 0) << 24) | ((fw_mode & 1) << 25)


def iop_operand(base_cid, nblocks, kind, is_last, vec=1):
    """hpu_iop_fmt.h:38-45: base_cid, block = blocks per integer, vec_size = integers - 1."""
    return (base_cid & 0xFFFF) | ((nblocks & 0xFF) << 16) | (((vec - 1) & 31) << 24) | \
        ((1 if is_last else 0) << 29) | ((kind & 3) << 30)


def iop_words(opcode, dsts, srcs, imms=(), nblocks=None, align=None):
    """dsts/srcs: lists of (base_cid, nblocks); imms: list of (value, nblocks).  The header's
    alignment fields are the operands' block counts unless `align` = (dst_align, src_align)
    is given (e.g. bit widths, as the programs of luthe_register_iop_program use)."""
    dal, sal = align if align is not None else (max(n for _, n in dsts), max(n for _, n in srcs))
    w = [iop_header(opcode, dal, sal, bool(imms))]
    for i, (cid, n) in enumerate(dsts):
        w.append(iop_operand(cid, n, KIND_DST, i == len(dsts) - 1))
    for i, (cid, n) in enumerate(srcs):
        w.append(iop_operand(cid, n, KIND_SRC, i == len(srcs) - 1))
    for i, (val, nb) in enumerate(imms):
        digits = [(val >> (2 * j)) & 3 for j in range(nb)]
        msgs = []
        for j in range(0, max(nb, 1), 8):
            m = 0
            for e, dgt in enumerate(digits[j:j + 8]):
                m |= dgt << (2 * e)
            msgs.append(m)
        hdr = (msgs[0] & 0xFFFF) | ((nb & 0xFFF) << 16) | ((1 if i == len(imms) - 1 else 0) << 28) | (KIND_IMM << 30)
        w.append(hdr)
        rest = msgs[1:]
        # ucore_fw.c:188-197 reads one more word while block > 8 * msg_pos
        pos = 1
        while nb > 8 * pos:
            lo = rest[pos - 1] if pos - 1 < len(rest) else 0
            hi = rest[pos] if pos < len(rest) else 0
            w.append(lo | (hi << 16))
            pos += 2
    del align
    return w
