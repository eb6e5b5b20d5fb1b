// Radix integer operations on 2-bit blocks expanded into block-level DAGs of ALU and
// PBS nodes (the role of the tfhe-rs "fw" translation programs, which are not in the
// reference checkout; the DOp vocabulary is the reference's: docs/dop.md:74-165).
//
// A block ciphertext carries a 4-bit value (2 message + 2 carry bits, plus padding);
// inputs and outputs are "clean" blocks (values < 4).  Every node writes fresh rows
// (single assignment), so dependencies are exactly producer -> consumer.
#pragma once
#include <stdint.h>

#include <vector>

namespace iop {

// OP_ADD is the carry-lookahead adder (Sklansky prefix over carry states, depth
// 2 + ceil(log2(n-1))); OP_ADD_RIPPLE the ripple-carry adder (n PBS, depth n).
// OP_MUL reduces partial-product columns to two rows, then adds them with the same
// carry-lookahead adder.
enum Op : uint32_t { OP_ADD = 0, OP_CMP_GT = 1, OP_MUL = 2, OP_ADD_RIPPLE = 3, OP_COUNT = 4 };

// Logical LUTs used by the programs (mapped to LUT-memory slots by the server).
enum Lut : uint32_t {
  LUT_MSG = 0,        // v & 3
  LUT_CARRY = 1,      // v >> 2
  LUT_MSG_CARRY = 2,  // Many-LUT (2): [v & 3, v >> 2] for v < 8
  LUT_CMP = 3,        // v = 4a + b -> 0 (a < b), 1 (a == b), 2 (a > b)
  LUT_CMP_MERGE = 4,  // v = 4 hi + lo (hi, lo in {0,1,2}) -> hi == 1 ? lo : hi
  LUT_IS_GT = 5,      // v in {0,1,2} -> v == 2
  LUT_MUL_LSB = 6,    // v = 4a + b -> (a b) & 3
  LUT_MUL_MSB = 7,    // v = 4a + b -> (a b) >> 2
  LUT_ADD_STATE = 8,  // s (<= 6) -> carry state: 2 generate (s >= 4), 1 propagate (s == 3), 0 kill
  LUT_ADD_SPLIT = 9,  // Many-LUT (2): [s & 3, carry state of s] for s < 8
  LUT_ADD_FINAL = 10, // v = 4 P + m (P = carry state of the blocks below, m < 4) -> (m + (P == 2)) & 3
  LUT_COUNT = 11
};

// Table of each logical LUT on 4-bit inputs; Many-LUT entries hold two 8-entry tables.
void lut_tables(uint32_t lut, std::vector<uint8_t>& table, uint32_t& log_lut_nb);

struct Node {
  uint8_t pbs = 0;          // 1 = PBS node, 0 = ALU node
  uint8_t opcode = 0;       // ALU DOp code
  uint8_t lut = 0;          // logical LUT (PBS)
  uint8_t log_lut_nb = 0;   // PBS Many-LUT factor
  uint32_t dst = 0, src0 = 0, src1 = 0;  // rows local to the request workspace
  uint32_t mul_factor = 0, msg_cst = 0;
  std::vector<uint32_t> succ;
  uint32_t npred = 0;
};

struct Program {
  uint32_t op = 0, nblocks = 0;
  uint32_t rows = 0;           // workspace rows
  uint32_t a_row0 = 0, b_row0 = 0;
  std::vector<uint32_t> out_rows;
  std::vector<Node> nodes;
  uint32_t npbs = 0, depth = 0;  // PBS count, PBS levels on the critical path
  std::vector<uint32_t> height;  // per node: PBS levels from it to the end (filled by the server)
};

// Builds the DAG of an IOp on `nblocks`-block operands (1..64).  Returns false if the
// op or size is not supported.
bool build(uint32_t op, uint32_t nblocks, Program& out);

// Plaintext evaluation of a program (PBS = table lookup, ALU = integer op mod 32): the
// host-side check of a program's logic.
void eval_plain(const Program& p, const uint32_t* a, const uint32_t* b, uint32_t* out);

// DOp template of a program for a register file of `regs` registers (>= 4): sources
// TS[0].i (a) and TS[1].i (b), results TD[0].i, spills to heap slots; PBS use LUT ids
// lut_base + logical LUT.  Nodes are issued in as-soon-as-possible level order, evicting
// the register whose next use is farthest; the last PBS of each level carries the flush
// bit.  Returns false if regs is too small.
struct EmitStats {
  uint32_t dops = 0, ld = 0, st = 0, alu = 0, pbs = 0, heap_slots = 0, levels = 0;
};
bool emit_dops(const Program& p, uint32_t regs, uint32_t lut_base, std::vector<uint32_t>& words, EmitStats& st);

// IOp opcode of the reference's IOp table (docs/iop.md:176-196) for each program.
uint8_t iop_opcode(uint32_t op);

}  // namespace iop
