// DOp / IOp words: decoding, validation, dependency levels and template patching.
// Host-only code.
//
// Word formats (LSB first) restate fw/ublaze/src/hpu_dop_fmt.h:50-116 and the packed
// structs of hw/module/hpu/module/hpu_common_instruction/rtl/hpu_common_instruction_pkg.sv:101-153:
//   arith     dst[6:0] src0[13:7] src1[20:14] mul_factor[25:21] opcode[31:26]
//   arith_msg dst[6:0] src0[13:7] msg_mode[14] msg_cst[25:15]    opcode[31:26]
//   mem       rid[6:0] pad[7]     mode[9:8]   slot[25:10]         opcode[31:26]
//   pbs       dst[6:0] src[13:7]  gid[25:14]                      opcode[31:26]
//   sync      sid[25:0]                                           opcode[31:26]
// Opcodes (hpu_common_instruction_pkg.sv:42-68): kind = opcode[5:4]
//   ADD 0x01 SUB 0x02 MAC 0x05 ADDS 0x09 SUBS 0x0A SSUB 0x0B MULS 0x0C SYNC 0x10
//   LD 0x20 ST 0x21 PBS 0x30|log_lut_nb (ML2/4/8), +0x08 for the flush variants.
// IOp words restate fw/ublaze/src/hpu_iop_fmt.h:19-76.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace dop {

enum Kind : uint32_t { KIND_ARITH = 0, KIND_SYNC = 1, KIND_MEM = 2, KIND_PBS = 3 };
enum Op : uint32_t {
  OP_ADD = 0x01, OP_SUB = 0x02, OP_MAC = 0x05, OP_ADDS = 0x09, OP_SUBS = 0x0A, OP_SSUB = 0x0B,
  OP_MULS = 0x0C, OP_SYNC = 0x10, OP_LD = 0x20, OP_ST = 0x21, OP_PBS = 0x30
};
enum MemMode : uint32_t { MEM_ADDR = 0, MEM_HEAP = 1, MEM_SRC = 2, MEM_DST = 3 };
constexpr uint32_t IMM_FLAG = 0x08;  // arith opcodes with an immediate (DOpArithFlag)

struct Dop {
  uint32_t word = 0, opcode = 0, kind = 0;
  uint32_t dst = 0, src0 = 0, src1 = 0;   // registers (arith, pbs: dst/src0)
  uint32_t mul_factor = 0, msg_cst = 0, msg_mode = 0;
  uint32_t rid = 0, mode = 0, slot = 0;   // mem
  uint32_t gid = 0, log_lut_nb = 0, flush = 0;
};

// Decode and validate a concrete DOp stream (templates rejected).  Returns 0 or a
// negative luthe error code; `why` names the first offending word.
int decode_stream(const uint32_t* words, size_t n, uint32_t regs, uint32_t slots, uint32_t luts,
                  std::vector<Dop>& out, std::string& why);
// Validate a template program (memory modes and IMM_VAR allowed).
int check_template(const uint32_t* words, size_t n, uint32_t regs, uint32_t luts, std::string& why);
// Dependency level per DOp (SYNC gets -1) on registers and slots (read-after-write,
// write-after-read, write-after-write).  Returns the number of levels.
int schedule_levels(const std::vector<Dop>& prog, uint32_t regs, std::vector<int>& level);

// Level schedule with register renaming: every register write goes to a fresh row
// rename_base + k, so only true data dependencies on registers remain (slots keep
// read-after-write, write-after-read and write-after-write).  Rows: register r = r,
// slot s = slot_base + s.  copy_back lists (architectural register, final row) pairs
// that restore the register file after the last level.  Results equal in-order
// execution.  Returns the number of levels (copy-back excluded).
struct Renamed {
  int level = -1;
  uint32_t dst = 0, src0 = 0, src1 = 0;  // rows
};
int schedule_renamed(const std::vector<Dop>& prog, uint32_t regs, uint32_t slot_base, uint32_t rename_base,
                     std::vector<Renamed>& out, std::vector<std::pair<uint32_t, uint32_t>>& copy_back,
                     uint32_t& rename_rows);

struct Operand {
  uint32_t cid_ofst = 0, len = 0;  // len: blocks the operand spans
};
struct IopBundle {
  uint32_t header = 0;
  uint8_t opcode = 0, src_align = 0, dst_align = 0, max_align = 0, has_imm = 0, fw_mode = 0;
  std::vector<Operand> dst, src;
  std::vector<std::vector<uint16_t>> imm;  // packed 2-bit digits, msg[0] = 8 lsb digits
  size_t words_used = 0;
};
// parse_iop of fw/ublaze/src/ucore_fw.c:135-205 with bounds checks.
int parse_iop(const uint32_t* words, size_t n, IopBundle& out, std::string& why);
// patch_dop of fw/ublaze/src/hpu_dop_fmt.h:148-229 over a whole program.
int patch_program(const std::vector<uint32_t>& tpl, const IopBundle& iop, uint32_t ct_slots,
                  std::vector<uint32_t>& out, std::string& why);
uint32_t patch_word(uint32_t w, const IopBundle& iop, uint32_t ct_slots, int& err, std::string& why);

// Encoders (for hosts and tests).
inline uint32_t enc_arith(uint32_t op, uint32_t dst, uint32_t s0, uint32_t s1, uint32_t mf) {
  return (op << 26) | ((mf & 31) << 21) | ((s1 & 127) << 14) | ((s0 & 127) << 7) | (dst & 127);
}
inline uint32_t enc_arith_msg(uint32_t op, uint32_t dst, uint32_t s0, uint32_t cst, uint32_t mode) {
  return (op << 26) | ((cst & 2047) << 15) | ((mode & 1) << 14) | ((s0 & 127) << 7) | (dst & 127);
}
inline uint32_t enc_mem(uint32_t op, uint32_t rid, uint32_t mode, uint32_t slot) {
  return (op << 26) | ((slot & 0xFFFF) << 10) | ((mode & 3) << 8) | (rid & 127);
}
inline uint32_t enc_pbs(uint32_t dst, uint32_t src, uint32_t gid, uint32_t log_lut_nb, uint32_t flush) {
  return ((OP_PBS | (flush ? 8u : 0u) | (log_lut_nb & 3)) << 26) | ((gid & 4095) << 14) | ((src & 127) << 7) | (dst & 127);
}
inline uint32_t enc_sync(uint32_t sid) { return (OP_SYNC << 26) | (sid & 0x3FFFFFF); }

}  // namespace dop
