// DOp / IOp front end (host only): see dop.h for the word formats.
#include "dop.h"

#include <unordered_map>

#include <stdio.h>

#include <algorithm>

#include "../../include/luthe.h"

namespace dop {

namespace {

std::string fmt(const char* f, size_t i, uint32_t w, const char* what) {
  char buf[160];
  snprintf(buf, sizeof buf, f, i, w, what);
  return buf;
}

bool valid_opcode(uint32_t op) {
  switch (op) {
    case OP_ADD: case OP_SUB: case OP_MAC: case OP_ADDS: case OP_SUBS: case OP_SSUB: case OP_MULS:
    case OP_SYNC: case OP_LD: case OP_ST:
      return true;
    default:
      return (op & 0x34) == 0x30;  // 0b11_f0ll : PBS / ML2 / ML4 / ML8 with optional flush
  }
}

Dop decode_word(uint32_t w) {
  Dop d;
  d.word = w;
  d.opcode = w >> 26;
  d.kind = d.opcode >> 4;
  switch (d.kind) {
    case KIND_ARITH:
      d.dst = w & 127;
      d.src0 = (w >> 7) & 127;
      if (d.opcode & IMM_FLAG) {
        d.msg_mode = (w >> 14) & 1;
        d.msg_cst = (w >> 15) & 2047;
      } else {
        d.src1 = (w >> 14) & 127;
        d.mul_factor = (w >> 21) & 31;
      }
      break;
    case KIND_MEM:
      d.rid = w & 127;
      d.mode = (w >> 8) & 3;
      d.slot = (w >> 10) & 0xFFFF;
      break;
    case KIND_PBS:
      d.dst = w & 127;
      d.src0 = (w >> 7) & 127;
      d.gid = (w >> 14) & 4095;
      d.log_lut_nb = d.opcode & 3;
      d.flush = (d.opcode >> 3) & 1;
      break;
    default:
      break;
  }
  return d;
}

int check_regs(const Dop& d, uint32_t regs, size_t i, std::string& why) {
  auto bad = [&](const char* what) {
    why = fmt("DOp %zu (0x%08x): %s", i, d.word, what);
    return LUTHE_ERR_RANGE;
  };
  switch (d.kind) {
    case KIND_ARITH:
      if (d.dst >= regs || d.src0 >= regs) return bad("register out of range");
      if (!(d.opcode & IMM_FLAG) && d.src1 >= regs) return bad("register out of range");
      break;
    case KIND_MEM:
      if (d.rid >= regs) return bad("register out of range");
      break;
    case KIND_PBS: {
      uint32_t nb = 1u << d.log_lut_nb;
      if (d.src0 >= regs || d.dst + nb > regs) return bad("register out of range");
      if (d.dst % nb) return bad("Many-LUT destination not aligned on the LUT count");
      break;
    }
    default:
      break;
  }
  return 0;
}

}  // namespace

int decode_stream(const uint32_t* words, size_t n, uint32_t regs, uint32_t slots, uint32_t luts,
                  std::vector<Dop>& out, std::string& why) {
  out.clear();
  out.reserve(n);
  for (size_t i = 0; i < n; i++) {
    Dop d = decode_word(words[i]);
    if (!valid_opcode(d.opcode)) {
      why = fmt("DOp %zu (0x%08x): %s", i, d.word, "unknown opcode");
      return LUTHE_ERR_DOP;
    }
    int rc = check_regs(d, regs, i, why);
    if (rc) return rc;
    if (d.kind == KIND_MEM) {
      if (d.mode != MEM_ADDR) {
        why = fmt("DOp %zu (0x%08x): %s", i, d.word, "templated memory operand in a concrete stream");
        return LUTHE_ERR_DOP;
      }
      if (d.slot >= slots) {
        why = fmt("DOp %zu (0x%08x): %s", i, d.word, "ciphertext slot out of range");
        return LUTHE_ERR_RANGE;
      }
    }
    if (d.kind == KIND_ARITH && (d.opcode & IMM_FLAG) && d.msg_mode != 0) {
      why = fmt("DOp %zu (0x%08x): %s", i, d.word, "templated immediate in a concrete stream");
      return LUTHE_ERR_DOP;
    }
    if (d.kind == KIND_PBS && d.gid >= luts) {
      why = fmt("DOp %zu (0x%08x): %s", i, d.word, "LUT id out of range");
      return LUTHE_ERR_RANGE;
    }
    out.push_back(d);
  }
  return 0;
}

int check_template(const uint32_t* words, size_t n, uint32_t regs, uint32_t luts, std::string& why) {
  for (size_t i = 0; i < n; i++) {
    Dop d = decode_word(words[i]);
    if (!valid_opcode(d.opcode)) {
      why = fmt("DOp %zu (0x%08x): %s", i, d.word, "unknown opcode");
      return LUTHE_ERR_DOP;
    }
    int rc = check_regs(d, regs, i, why);
    if (rc) return rc;
    if (d.kind == KIND_PBS && d.gid >= luts) {
      why = fmt("DOp %zu (0x%08x): %s", i, d.word, "LUT id out of range");
      return LUTHE_ERR_RANGE;
    }
  }
  return 0;
}

int schedule_renamed(const std::vector<Dop>& prog, uint32_t regs, uint32_t slot_base, uint32_t rename_base,
                     std::vector<Renamed>& out, std::vector<std::pair<uint32_t, uint32_t>>& copy_back,
                     uint32_t& rename_rows) {
  std::vector<uint32_t> cur(regs);       // row holding each register's current value
  std::vector<int> ready(regs, 0);       // level from which that row can be read
  for (uint32_t r = 0; r < regs; r++) cur[r] = r;
  std::unordered_map<uint32_t, std::pair<int, int>> slots;  // slot -> (last write, last read)
  out.assign(prog.size(), Renamed());
  copy_back.clear();
  uint32_t next = 0;
  int nlev = 0;
  auto fresh = [&](uint32_t count) {
    const uint32_t row = rename_base + next;
    next += count;
    return row;
  };
  for (size_t i = 0; i < prog.size(); i++) {
    const Dop& d = prog[i];
    Renamed& o = out[i];
    int lv = 0;
    switch (d.kind) {
      case KIND_ARITH: {
        lv = ready[d.src0];
        o.src0 = cur[d.src0];
        if (!(d.opcode & IMM_FLAG)) {
          lv = std::max(lv, ready[d.src1]);
          o.src1 = cur[d.src1];
        } else {
          o.src1 = o.src0;
        }
        o.dst = fresh(1);
        cur[d.dst] = o.dst;
        ready[d.dst] = lv + 1;
        break;
      }
      case KIND_PBS: {
        lv = ready[d.src0];
        o.src0 = cur[d.src0];
        const uint32_t nb = 1u << d.log_lut_nb;
        o.dst = fresh(nb);
        for (uint32_t j = 0; j < nb; j++) {
          cur[d.dst + j] = o.dst + j;
          ready[d.dst + j] = lv + 1;
        }
        break;
      }
      case KIND_MEM: {
        auto& st = slots.emplace(d.slot, std::make_pair(-1, -1)).first->second;
        if (d.opcode == OP_LD) {
          lv = st.first + 1;
          o.src0 = slot_base + d.slot;
          o.dst = fresh(1);
          cur[d.rid] = o.dst;
          ready[d.rid] = lv + 1;
          st.second = std::max(st.second, lv);
        } else {
          lv = std::max({ready[d.rid], st.first + 1, st.second + 1});
          o.src0 = cur[d.rid];
          o.dst = slot_base + d.slot;
          st.first = lv;
        }
        break;
      }
      default:
        continue;  // SYNC
    }
    o.level = lv;
    nlev = std::max(nlev, lv + 1);
  }
  for (uint32_t r = 0; r < regs; r++)
    if (cur[r] != r) copy_back.emplace_back(r, cur[r]);
  rename_rows = next;
  return nlev;
}

int schedule_levels(const std::vector<Dop>& prog, uint32_t regs, std::vector<int>& level) {
  // Resources: registers [0, regs), ciphertext slots from `regs` on (sparse map).
  std::vector<int> reg_w(regs, -1), reg_r(regs, -1);
  std::vector<std::pair<uint32_t, std::pair<int, int>>> slot_state;  // slot -> (last write, last read)
  auto slot_ref = [&](uint32_t s) -> std::pair<int, int>& {
    for (auto& e : slot_state)
      if (e.first == s) return e.second;
    slot_state.push_back({s, {-1, -1}});
    return slot_state.back().second;
  };
  level.assign(prog.size(), -1);
  int nlev = 0;
  std::vector<uint32_t> rr, rw;  // registers read / written
  for (size_t i = 0; i < prog.size(); i++) {
    const Dop& d = prog[i];
    rr.clear();
    rw.clear();
    int slot_read = -1, slot_write = -1;
    switch (d.kind) {
      case KIND_ARITH:
        rr.push_back(d.src0);
        if (!(d.opcode & IMM_FLAG)) rr.push_back(d.src1);
        rw.push_back(d.dst);
        break;
      case KIND_MEM:
        if (d.opcode == OP_LD) {
          slot_read = (int)d.slot;
          rw.push_back(d.rid);
        } else {
          rr.push_back(d.rid);
          slot_write = (int)d.slot;
        }
        break;
      case KIND_PBS:
        rr.push_back(d.src0);
        for (uint32_t j = 0; j < (1u << d.log_lut_nb); j++) rw.push_back(d.dst + j);
        break;
      default:
        continue;  // SYNC: completion is tracked per ticket
    }
    int lv = 0;
    for (uint32_t r : rr) lv = std::max(lv, reg_w[r] + 1);
    for (uint32_t r : rw) lv = std::max({lv, reg_w[r] + 1, reg_r[r] + 1});
    if (slot_read >= 0) lv = std::max(lv, slot_ref((uint32_t)slot_read).first + 1);
    if (slot_write >= 0) {
      auto& st = slot_ref((uint32_t)slot_write);
      lv = std::max({lv, st.first + 1, st.second + 1});
    }
    level[i] = lv;
    nlev = std::max(nlev, lv + 1);
    for (uint32_t r : rr) reg_r[r] = std::max(reg_r[r], lv);
    for (uint32_t r : rw) {
      reg_w[r] = lv;
      reg_r[r] = -1;
    }
    if (slot_read >= 0) {
      auto& st = slot_ref((uint32_t)slot_read);
      st.second = std::max(st.second, lv);
    }
    if (slot_write >= 0) {
      auto& st = slot_ref((uint32_t)slot_write);
      st.first = lv;
      st.second = -1;
    }
  }
  return nlev;
}

int parse_iop(const uint32_t* w, size_t n, IopBundle& out, std::string& why) {
  out = IopBundle();
  size_t pos = 0;
  auto next = [&](uint32_t& v) {
    if (pos >= n) return false;
    v = w[pos++];
    return true;
  };
  uint32_t h;
  if (!next(h)) return why = "empty IOp", LUTHE_ERR_DOP;
  out.header = h;
  out.src_align = h & 0xFF;
  out.dst_align = (h >> 8) & 0xFF;
  out.opcode = (h >> 16) & 0xFF;
  out.has_imm = (h >> 24) & 1;
  out.fw_mode = (h >> 25) & 1;
  out.max_align = std::max(out.dst_align, out.src_align);
  for (int list = 0; list < 2; list++) {
    std::vector<Operand>& ops = list == 0 ? out.dst : out.src;
    for (;;) {
      uint32_t o;
      if (!next(o)) return why = "truncated IOp operand list", LUTHE_ERR_DOP;
      uint32_t kind = o >> 30;
      if (kind != (list == 0 ? 1u : 0u)) return why = "IOp operand of the wrong kind", LUTHE_ERR_DOP;
      if (ops.size() >= 16) return why = "too many IOp operands", LUTHE_ERR_DOP;
      // hpu_iop_fmt.h:38-45: base_cid, block (blocks per integer), vec_size (integers - 1);
      // a template addresses blocks 0 .. (vec_size + 1) * block - 1 of the operand
      const uint32_t blocks = (o >> 16) & 0xFF, vec = ((o >> 24) & 31) + 1;
      ops.push_back(Operand{o & 0xFFFF, vec * (blocks ? blocks : 1u)});
      if ((o >> 29) & 1) break;  // is_last
    }
  }
  if (out.has_imm) {
    for (;;) {
      uint32_t ih;
      if (!next(ih)) return why = "truncated IOp immediate", LUTHE_ERR_DOP;
      if ((ih >> 30) != 2) return why = "IOp immediate of the wrong kind", LUTHE_ERR_DOP;
      uint32_t block = (ih >> 16) & 0xFFF;
      std::vector<uint16_t> msg;
      msg.push_back((uint16_t)(ih & 0xFFFF));
      // ucore_fw.c:188-197: one more word (two 16-bit slots) while block > 8 * msg_pos
      while (block > 8 * msg.size()) {
        uint32_t b;
        if (!next(b)) return why = "truncated IOp immediate payload", LUTHE_ERR_DOP;
        msg.push_back((uint16_t)(b & 0xFFFF));
        msg.push_back((uint16_t)(b >> 16));
      }
      out.imm.push_back(msg);
      if (out.imm.size() > 16) return why = "too many IOp immediates", LUTHE_ERR_DOP;
      if ((ih >> 28) & 1) break;
    }
  }
  out.words_used = pos;
  if (pos != n) return why = "trailing words after IOp", LUTHE_ERR_DOP;
  return 0;
}

uint32_t patch_word(uint32_t w, const IopBundle& iop, uint32_t ct_slots, int& err, std::string& why) {
  err = 0;
  uint32_t opcode = w >> 26, kind = opcode >> 4;
  if (kind == KIND_MEM) {
    uint32_t mode = (w >> 8) & 3, slot = (w >> 10) & 0xFFFF, out_slot = slot;
    switch (mode) {
      case MEM_ADDR:
        return w;
      case MEM_HEAP:  // hpu_dop_fmt.h:153-157: HEAP_START_SLOT - slot, heap grows down
        if (slot > ct_slots - 1) {
          err = LUTHE_ERR_RANGE;
          why = "heap offset beyond memory";
          return w;
        }
        out_slot = (ct_slots - 1) - slot;
        break;
      case MEM_SRC:
      case MEM_DST: {  // hpu_dop_fmt.h:159-174: operand[tid].cid_ofst + bid
        uint32_t tid = (slot >> 8) & 0xFF, bid = slot & 0xFF;
        const std::vector<Operand>& ops = mode == MEM_SRC ? iop.src : iop.dst;
        if (tid >= ops.size() || bid >= ops[tid].len) {
          err = LUTHE_ERR_RANGE;
          why = "templated operand block out of range";
          return w;
        }
        out_slot = ops[tid].cid_ofst + bid;
        break;
      }
    }
    if (out_slot >= ct_slots) {
      err = LUTHE_ERR_RANGE;
      why = "patched slot out of range";
      return w;
    }
    return (w & ~((0xFFFFu << 10) | (3u << 8))) | (out_slot << 10) | (MEM_ADDR << 8);
  }
  if (kind == KIND_ARITH && (opcode & IMM_FLAG)) {
    uint32_t mode = (w >> 14) & 1, cst = (w >> 15) & 2047;
    if (mode == 0) return w;
    // hpu_dop_fmt.h:180-199: digit bid of immediate tid, MSG_WIDTH = 2 bits
    uint32_t tid = (cst >> 8) & 0xFF, bid = cst & 0xFF;
    uint32_t slot_id = (bid * 2) / 16, offset = (bid * 2) % 16;
    if (tid >= iop.imm.size() || slot_id >= iop.imm[tid].size()) {
      err = LUTHE_ERR_RANGE;
      why = "templated immediate digit out of range";
      return w;
    }
    uint32_t v = ((uint32_t)iop.imm[tid][slot_id] >> offset) & 3;
    return (w & ~((2047u << 15) | (1u << 14))) | (v << 15);
  }
  return w;
}

int patch_program(const std::vector<uint32_t>& tpl, const IopBundle& iop, uint32_t ct_slots,
                  std::vector<uint32_t>& out, std::string& why) {
  out.clear();
  out.reserve(tpl.size());
  for (size_t i = 0; i < tpl.size(); i++) {
    int err = 0;
    uint32_t p = patch_word(tpl[i], iop, ct_slots, err, why);
    if (err) {
      char buf[64];
      snprintf(buf, sizeof buf, " (template word %zu)", i);
      why += buf;
      return err;
    }
    out.push_back(p);
  }
  return 0;
}

}  // namespace dop
