// Radix IOp programs over 2-bit blocks (see iop_programs.h).
//
//  ADD    : s_i = a_i + b_i; a PBS maps each s_i (i < n-1) to its carry state
//           (kill / propagate / generate); a Sklansky parallel prefix merges states with
//           (hi, lo) -> hi == propagate ? lo : hi; block i >= 1 is then
//           PBS(4 P_(i-1) + s_i) = (s_i + carry_in) & 3.  About n + (n/2) log2 n PBS,
//           depth 2 + ceil(log2(n-1)).
//  ADD_RIPPLE : ripple t_i = s_i + c_(i-1), one Many-LUT PBS per block yields
//           (t_i & 3, t_i >> 2) (docs/dop.md "ManyCarryMsg"); n PBS, depth n.
//  CMP_GT : c_i = CMP(4 a_i + b_i) in {lt, eq, gt}; a pairwise tree merges
//           (hi, lo) -> hi == eq ? lo : hi; a last PBS maps gt -> 1.  2n PBS, depth
//           2 + ceil(log2 n).
//  MUL    : partial products (4 a_i + b_j -> lsb, msb) for i + j < n, then rounds of
//           column reduction: groups of terms whose value bound stays <= 15 are summed
//           with ALU adds and split into msg (same column) and carry (next column); a
//           group bounded by 7 uses one Many-LUT PBS, one bounded by 3 needs none.
#include "iop_programs.h"

#include <algorithm>
#include <deque>

#include "dop.h"

namespace iop {

void lut_tables(uint32_t lut, std::vector<uint8_t>& t, uint32_t& log_lut_nb) {
  t.assign(16, 0);
  log_lut_nb = 0;
  for (uint32_t v = 0; v < 16; v++) {
    uint32_t a = v >> 2, b = v & 3;
    switch (lut) {
      case LUT_MSG: t[v] = (uint8_t)(v & 3); break;
      case LUT_CARRY: t[v] = (uint8_t)(v >> 2); break;
      case LUT_CMP: t[v] = (uint8_t)(a < b ? 0 : (a == b ? 1 : 2)); break;
      case LUT_CMP_MERGE: t[v] = (uint8_t)((a == 1) ? b : a); break;
      case LUT_IS_GT: t[v] = (uint8_t)(v == 2 ? 1 : 0); break;
      case LUT_MUL_LSB: t[v] = (uint8_t)((a * b) & 3); break;
      case LUT_MUL_MSB: t[v] = (uint8_t)((a * b) >> 2); break;
      case LUT_ADD_STATE: t[v] = (uint8_t)(v >= 4 ? 2 : (v == 3 ? 1 : 0)); break;
      case LUT_ADD_FINAL: t[v] = (uint8_t)((b + (a == 2 ? 1 : 0)) & 3); break;
      default: break;
    }
  }
  if (lut == LUT_MSG_CARRY || lut == LUT_ADD_SPLIT) {
    log_lut_nb = 1;
    for (uint32_t v = 0; v < 8; v++) {
      t[v] = (uint8_t)(v & 3);
      t[8 + v] = (uint8_t)(lut == LUT_MSG_CARRY ? v >> 2 : (v >= 4 ? 2 : (v == 3 ? 1 : 0)));
    }
  }
}

namespace {

constexpr uint8_t ALU_ADD = 0x01, ALU_MAC = 0x05;

struct Builder {
  Program& p;
  std::vector<int> writer;     // row -> producing node (-1: program input)
  std::vector<uint32_t> bound;  // row -> largest value the row can hold
  std::vector<uint32_t> level;  // row -> PBS levels needed to produce it

  explicit Builder(Program& prog) : p(prog) {}

  uint32_t new_rows(uint32_t n, uint32_t bnd, uint32_t lev) {
    uint32_t r = p.rows;
    for (uint32_t i = 0; i < n; i++) {
      writer.push_back(-1);
      bound.push_back(bnd);
      level.push_back(lev);
    }
    p.rows += n;
    return r;
  }

  void link(Node& n, uint32_t idx, uint32_t src) {
    int w = writer[src];
    if (w < 0) return;
    auto& succ = p.nodes[(size_t)w].succ;
    if (std::find(succ.begin(), succ.end(), idx) != succ.end()) return;  // operand used twice
    succ.push_back(idx);
    n.npred++;
  }

  uint32_t alu(uint8_t op, uint32_t a, uint32_t b, uint32_t mf = 0) {
    uint32_t bnd = (op == ALU_MAC) ? bound[a] * mf + bound[b] : bound[a] + bound[b];
    uint32_t r = new_rows(1, bnd, std::max(level[a], level[b]));
    Node n;
    n.opcode = op;
    n.dst = r;
    n.src0 = a;
    n.src1 = b;
    n.mul_factor = mf;
    uint32_t idx = (uint32_t)p.nodes.size();
    p.nodes.push_back(n);
    link(p.nodes.back(), idx, a);
    link(p.nodes.back(), idx, b);
    writer[r] = (int)idx;
    return r;
  }

  // PBS of `src` through logical LUT `lut`; returns the first output row.
  uint32_t pbs(uint32_t lut, uint32_t src, uint32_t out_bound) {
    std::vector<uint8_t> t;
    uint32_t lognb;
    lut_tables(lut, t, lognb);
    uint32_t nout = 1u << lognb;
    uint32_t r = new_rows(nout, out_bound, level[src] + 1);
    Node n;
    n.pbs = 1;
    n.lut = (uint8_t)lut;
    n.log_lut_nb = (uint8_t)lognb;
    n.dst = r;
    n.src0 = src;
    uint32_t idx = (uint32_t)p.nodes.size();
    p.nodes.push_back(n);
    link(p.nodes.back(), idx, src);
    for (uint32_t i = 0; i < nout; i++) writer[r + i] = (int)idx;
    p.npbs++;
    p.depth = std::max(p.depth, level[src] + 1);
    return r;
  }
};

void build_add_ripple(Builder& B, uint32_t n) {
  uint32_t carry = 0;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t s = B.alu(ALU_ADD, B.p.a_row0 + i, B.p.b_row0 + i);
    uint32_t t = (i == 0) ? s : B.alu(ALU_ADD, s, carry);
    if (i + 1 == n) {
      B.p.out_rows.push_back(B.pbs(LUT_MSG, t, 3));
    } else {
      uint32_t r = B.pbs(LUT_MSG_CARRY, t, 3);  // rows r (msg), r + 1 (carry)
      B.bound[r + 1] = 1;
      B.p.out_rows.push_back(r);
      carry = r + 1;
    }
  }
}

// Carry-propagating addition of per-block sums s[i] (value bound <= 6 each): the
// result blocks are (sum_i s_i 4^i) mod 4^n.  Block i's low part m_i = s_i & 3 and its
// carry state (2 generate, 1 propagate, 0 kill) come from one Many-LUT PBS, or from the
// state PBS alone when s_i <= 3 already; a Sklansky parallel prefix merges states; block
// i >= 1 is PBS(4 P_(i-1) + m_i).
void prefix_add(Builder& B, const std::vector<uint32_t>& s) {
  const uint32_t n = (uint32_t)s.size();
  std::vector<uint32_t> m(n), P(n > 1 ? n - 1 : 0);
  for (uint32_t i = 0; i < n; i++) {
    const bool need_state = i + 1 < n, small = B.bound[s[i]] <= 3;
    if (small) {
      m[i] = s[i];
      if (need_state) P[i] = B.pbs(LUT_ADD_STATE, s[i], 2);
    } else if (need_state) {
      uint32_t r = B.pbs(LUT_ADD_SPLIT, s[i], 3);  // rows r (s & 3), r + 1 (state)
      B.bound[r + 1] = 2;
      m[i] = r;
      P[i] = r + 1;
    } else {
      m[i] = B.pbs(LUT_MSG, s[i], 3);
    }
  }
  B.p.out_rows.push_back(m[0]);
  const uint32_t np = (uint32_t)P.size();
  for (uint32_t d = 0; (1u << d) < np; d++) {
    std::vector<uint32_t> next = P;
    for (uint32_t i = 0; i < np; i++) {
      if (!((i >> d) & 1)) continue;
      uint32_t j = ((i >> d) << d) - 1;  // last element of the lower half of i's group
      uint32_t u = B.alu(ALU_MAC, P[i], P[j], 4);
      next[i] = B.pbs(LUT_CMP_MERGE, u, 2);
    }
    P.swap(next);
  }
  for (uint32_t i = 1; i < n; i++) {
    uint32_t v = B.alu(ALU_MAC, P[i - 1], m[i], 4);
    B.p.out_rows.push_back(B.pbs(LUT_ADD_FINAL, v, 3));
  }
}

void build_add_prefix(Builder& B, uint32_t n) {
  std::vector<uint32_t> s(n);
  for (uint32_t i = 0; i < n; i++) s[i] = B.alu(ALU_ADD, B.p.a_row0 + i, B.p.b_row0 + i);
  prefix_add(B, s);
}

void build_cmp_gt(Builder& B, uint32_t n) {
  std::vector<uint32_t> c;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t t = B.alu(ALU_MAC, B.p.a_row0 + i, B.p.b_row0 + i, 4);
    c.push_back(B.pbs(LUT_CMP, t, 2));
  }
  while (c.size() > 1) {
    std::vector<uint32_t> next;
    for (size_t j = 0; j + 1 < c.size(); j += 2) {
      uint32_t u = B.alu(ALU_MAC, c[j + 1], c[j], 4);  // 4 * hi + lo
      next.push_back(B.pbs(LUT_CMP_MERGE, u, 2));
    }
    if (c.size() & 1) next.push_back(c.back());
    c.swap(next);
  }
  B.p.out_rows.push_back(B.pbs(LUT_IS_GT, c[0], 1));
}

void build_mul(Builder& B, uint32_t n) {
  std::vector<std::vector<uint32_t>> col(n);
  for (uint32_t i = 0; i < n; i++)
    for (uint32_t j = 0; i + j < n; j++) {
      uint32_t t = B.alu(ALU_MAC, B.p.a_row0 + i, B.p.b_row0 + j, 4);
      col[i + j].push_back(B.pbs(LUT_MUL_LSB, t, 3));
      if (i + j + 1 < n) col[i + j + 1].push_back(B.pbs(LUT_MUL_MSB, t, 2));
    }
  // Reduce columns holding more than two terms (all terms bounded by 3) ...
  for (;;) {
    bool busy = false;
    for (auto& c : col) busy |= c.size() > 2;
    if (!busy) break;
    std::vector<std::vector<uint32_t>> next(n);
    for (uint32_t k = 0; k < n; k++) {
      auto& terms = col[k];
      if (terms.size() <= 2) {
        for (uint32_t r : terms) next[k].push_back(r);
        continue;
      }
      size_t i = 0;
      while (i < terms.size()) {
        // greedy group with value bound <= 15
        uint32_t acc = terms[i], bnd = B.bound[terms[i]];
        size_t j = i + 1;
        while (j < terms.size() && bnd + B.bound[terms[j]] <= 15) {
          acc = B.alu(ALU_ADD, acc, terms[j]);
          bnd = B.bound[acc];
          j++;
        }
        if (bnd <= 3) {
          next[k].push_back(acc);  // a single term, or a sum that cannot carry
        } else if (k + 1 == n) {
          next[k].push_back(B.pbs(LUT_MSG, acc, 3));  // top column: carry dropped (mod 4^n)
        } else if (bnd <= 7) {
          uint32_t r = B.pbs(LUT_MSG_CARRY, acc, 3);
          B.bound[r + 1] = bnd >> 2;
          next[k].push_back(r);
          next[k + 1].push_back(r + 1);
        } else {
          next[k].push_back(B.pbs(LUT_MSG, acc, 3));
          next[k + 1].push_back(B.pbs(LUT_CARRY, acc, bnd >> 2));
        }
        i = j;
      }
    }
    col.swap(next);
  }
  // ... then one carry-propagating addition of the two remaining rows.
  std::vector<uint32_t> s(n);
  for (uint32_t k = 0; k < n; k++)
    s[k] = col[k].size() == 2 ? B.alu(ALU_ADD, col[k][0], col[k][1]) : col[k][0];
  prefix_add(B, s);
}

}  // namespace

bool build(uint32_t op, uint32_t nblocks, Program& out) {
  if (nblocks < 1 || nblocks > 64 || op >= OP_COUNT) return false;
  out = Program();
  out.op = op;
  out.nblocks = nblocks;
  Builder B(out);
  out.a_row0 = B.new_rows(nblocks, 3, 0);
  out.b_row0 = B.new_rows(nblocks, 3, 0);
  switch (op) {
    case OP_ADD: build_add_prefix(B, nblocks); break;
    case OP_ADD_RIPPLE: build_add_ripple(B, nblocks); break;
    case OP_CMP_GT: build_cmp_gt(B, nblocks); break;
    case OP_MUL: build_mul(B, nblocks); break;
    default: return false;
  }
  return true;
}

uint8_t iop_opcode(uint32_t op) {
  switch (op) {
    case OP_ADD: return 0xE0;
    case OP_CMP_GT: return 0xC0;
    case OP_MUL: return 0xE4;
    default: return 0x60;  // user range (0x00..0x7f): ripple-carry ADD
  }
}

bool emit_dops(const Program& p, uint32_t regs, uint32_t lut_base, std::vector<uint32_t>& words, EmitStats& st) {
  words.clear();
  st = EmitStats();
  if (regs < 4 || regs > 128) return false;
  const size_t nn = p.nodes.size();
  // ASAP level of every node (inputs at level 0)
  std::vector<int> writer(p.rows, -1);
  for (size_t i = 0; i < nn; i++) {
    const Node& n = p.nodes[i];
    for (uint32_t k = 0; k < (n.pbs ? (1u << n.log_lut_nb) : 1u); k++) writer[n.dst + k] = (int)i;
  }
  std::vector<uint32_t> lvl(nn, 0);
  for (size_t i = 0; i < nn; i++) {
    const Node& n = p.nodes[i];
    uint32_t l = 0;
    int w0 = writer[n.src0];
    if (w0 >= 0) l = std::max(l, lvl[(size_t)w0] + 1);
    if (!n.pbs) {
      int w1 = writer[n.src1];
      if (w1 >= 0) l = std::max(l, lvl[(size_t)w1] + 1);
    }
    lvl[i] = l;
  }
  std::vector<uint32_t> order(nn);
  for (size_t i = 0; i < nn; i++) order[i] = (uint32_t)i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return lvl[x] < lvl[y]; });
  // uses of every row, as positions in the issue order
  std::vector<std::vector<uint32_t>> uses(p.rows);
  for (uint32_t pos = 0; pos < nn; pos++) {
    const Node& n = p.nodes[order[pos]];
    uses[n.src0].push_back(pos);
    if (!n.pbs && n.src1 != n.src0) uses[n.src1].push_back(pos);
  }
  std::vector<int> out_block(p.rows, -1);
  for (size_t k = 0; k < p.out_rows.size(); k++) out_block[p.out_rows[k]] = (int)k;
  std::vector<size_t> next_use(p.rows, 0);  // index into uses[row]
  auto next_pos = [&](uint32_t row) -> uint32_t {
    return next_use[row] < uses[row].size() ? uses[row][next_use[row]] : 0xFFFFFFFFu;
  };
  std::vector<int> reg_of(p.rows, -1), heap_of(p.rows, -1), content(regs, -1);
  std::vector<uint32_t> pinned(regs, 0);
  std::deque<uint32_t> free_regs;
  for (uint32_t r = 0; r < regs; r++) free_regs.push_back(r);
  std::vector<uint32_t> free_heap;
  uint32_t heap_top = 0;
  auto mem_slot = [&](uint32_t row, uint32_t& mode, uint32_t& slot) -> bool {
    if (row >= p.a_row0 && row < p.a_row0 + p.nblocks) {
      mode = dop::MEM_SRC, slot = ((0u) << 8) | (row - p.a_row0);
      return true;
    }
    if (row >= p.b_row0 && row < p.b_row0 + p.nblocks) {
      mode = dop::MEM_SRC, slot = ((1u) << 8) | (row - p.b_row0);
      return true;
    }
    if (heap_of[row] >= 0) {
      mode = dop::MEM_HEAP, slot = (uint32_t)heap_of[row];
      return true;
    }
    return false;
  };
  auto release = [&](uint32_t r) {
    content[r] = -1;
    free_regs.push_back(r);
  };
  // frees a register holding a value: spill it first if it is still needed and not in memory
  auto evict = [&](uint32_t r) {
    const int row = content[r];
    if (row >= 0) {
      uint32_t mode, slot;
      if (next_pos((uint32_t)row) != 0xFFFFFFFFu && !mem_slot((uint32_t)row, mode, slot)) {
        uint32_t h;
        if (!free_heap.empty()) {
          h = free_heap.back();
          free_heap.pop_back();
        } else {
          h = heap_top++;
        }
        heap_of[row] = (int)h;
        words.push_back(dop::enc_mem(dop::OP_ST, r, dop::MEM_HEAP, h));
        st.st++;
      }
      reg_of[row] = -1;
    }
    content[r] = -1;
  };
  // victim: unpinned register whose value is next used farthest away
  auto victim = [&](uint32_t align, uint32_t count) -> int {
    int best = -1;
    uint64_t best_key = 0;
    for (uint32_t r0 = 0; r0 + count <= regs; r0 += align) {
      uint64_t key = ~0ull;
      bool ok = true;
      for (uint32_t k = 0; k < count && ok; k++) {
        if (pinned[r0 + k]) ok = false;
        else if (content[r0 + k] >= 0) key = std::min<uint64_t>(key, next_pos((uint32_t)content[r0 + k]));
      }
      if (ok && (best < 0 || key > best_key)) best = (int)r0, best_key = key;
    }
    return best;
  };
  // `count` consecutive registers aligned on `align`, evicting as needed
  auto grab = [&](uint32_t align, uint32_t count) -> int {
    if (count == 1) {
      for (size_t i = 0; i < free_regs.size(); i++) {
        uint32_t r = free_regs[i];
        if (!pinned[r] && content[r] < 0) {
          free_regs.erase(free_regs.begin() + (ptrdiff_t)i);
          return (int)r;
        }
      }
    } else {
      for (uint32_t r0 = 0; r0 + count <= regs; r0 += align) {
        bool ok = true;
        for (uint32_t k = 0; k < count; k++) ok &= !pinned[r0 + k] && content[r0 + k] < 0;
        if (ok) {
          for (uint32_t k = 0; k < count; k++)
            free_regs.erase(std::find(free_regs.begin(), free_regs.end(), r0 + k));
          return (int)r0;
        }
      }
    }
    int r0 = victim(align, count);
    if (r0 < 0) return -1;
    for (uint32_t k = 0; k < count; k++) {
      evict((uint32_t)r0 + k);
      auto it = std::find(free_regs.begin(), free_regs.end(), (uint32_t)r0 + k);
      if (it != free_regs.end()) free_regs.erase(it);
    }
    return r0;
  };
  auto load = [&](uint32_t row) -> int {
    if (reg_of[row] >= 0) return reg_of[row];
    uint32_t mode, slot;
    if (!mem_slot(row, mode, slot)) return -1;  // never produced
    int r = grab(1, 1);
    if (r < 0) return -1;
    words.push_back(dop::enc_mem(dop::OP_LD, (uint32_t)r, mode, slot));
    st.ld++;
    reg_of[row] = r;
    content[r] = (int)row;
    return r;
  };
  auto drop_if_dead = [&](uint32_t row) {
    if (next_pos(row) != 0xFFFFFFFFu) return;
    if (reg_of[row] >= 0) release((uint32_t)reg_of[row]), reg_of[row] = -1;
    if (heap_of[row] >= 0) free_heap.push_back((uint32_t)heap_of[row]), heap_of[row] = -1;
  };
  auto store_output = [&](uint32_t row, uint32_t r) {
    if (out_block[row] < 0) return;
    words.push_back(dop::enc_mem(dop::OP_ST, r, dop::MEM_DST, ((0u) << 8) | (uint32_t)out_block[row]));
    st.st++;
    out_block[row] = -1;
  };
  size_t last_pbs_word = (size_t)-1;
  uint32_t cur_level = 0;
  for (uint32_t pos = 0; pos < nn; pos++) {
    const Node& n = p.nodes[order[pos]];
    if (lvl[order[pos]] != cur_level) {
      if (last_pbs_word != (size_t)-1) words[last_pbs_word] |= 8u << 26;  // flush the level's last PBS
      last_pbs_word = (size_t)-1;
      cur_level = lvl[order[pos]];
    }
    // sources into registers, pinned for this node
    int rs0 = load(n.src0);
    if (rs0 < 0) return false;
    pinned[(uint32_t)rs0] = 1;
    int rs1 = rs0;
    if (!n.pbs) {
      rs1 = load(n.src1);
      if (rs1 < 0) return false;
      pinned[(uint32_t)rs1] = 1;
    }
    // consume this use of the sources
    next_use[n.src0]++;
    if (!n.pbs && n.src1 != n.src0) next_use[n.src1]++;
    pinned[(uint32_t)rs0] = 0;
    pinned[(uint32_t)rs1] = 0;
    const uint32_t cnt = n.pbs ? (1u << n.log_lut_nb) : 1u;
    // a dying source register may take the result (no Many-LUT alignment issue for 1 reg)
    for (uint32_t row : {n.src0, n.src1}) {
      if (n.pbs && row != n.src0) continue;
      if (next_pos(row) == 0xFFFFFFFFu && reg_of[row] >= 0 && !(cnt > 1)) {
        uint32_t r = (uint32_t)reg_of[row];
        reg_of[row] = -1;
        content[r] = -1;
        if (heap_of[row] >= 0) free_heap.push_back((uint32_t)heap_of[row]), heap_of[row] = -1;
        free_regs.push_back(r);
      }
    }
    pinned[(uint32_t)rs0] = reg_of[n.src0] >= 0 ? 1 : 0;
    if (!n.pbs && reg_of[n.src1] >= 0) pinned[(uint32_t)rs1] = 1;
    int rd = grab(cnt, cnt);
    if (rd < 0) return false;
    pinned[(uint32_t)rs0] = 0;
    pinned[(uint32_t)rs1] = 0;
    if (n.pbs) {
      words.push_back(dop::enc_pbs((uint32_t)rd, (uint32_t)rs0, lut_base + n.lut, n.log_lut_nb, 0));
      last_pbs_word = words.size() - 1;
      st.pbs++;
    } else {
      words.push_back(dop::enc_arith(n.opcode, (uint32_t)rd, (uint32_t)rs0, (uint32_t)rs1, n.mul_factor));
      st.alu++;
    }
    for (uint32_t k = 0; k < cnt; k++) {
      const uint32_t row = n.dst + k;
      reg_of[row] = rd + (int)k;
      content[(uint32_t)rd + k] = (int)row;
      store_output(row, (uint32_t)rd + k);
      drop_if_dead(row);
    }
    drop_if_dead(n.src0);
    if (!n.pbs) drop_if_dead(n.src1);
  }
  if (last_pbs_word != (size_t)-1) words[last_pbs_word] |= 8u << 26;
  // outputs that no node produced (inputs passed through)
  for (size_t k = 0; k < p.out_rows.size(); k++) {
    const uint32_t row = p.out_rows[k];
    if (out_block[row] < 0) continue;
    int r = load(row);
    if (r < 0) return false;
    store_output(row, (uint32_t)r);
  }
  st.dops = (uint32_t)words.size();
  st.heap_slots = heap_top;
  st.levels = nn ? cur_level + 1 : 0;
  return true;
}

void eval_plain(const Program& p, const uint32_t* a, const uint32_t* b, uint32_t* out) {
  std::vector<uint32_t> v(p.rows, 0);
  for (uint32_t i = 0; i < p.nblocks; i++) {
    v[p.a_row0 + i] = a[i] & 3;
    v[p.b_row0 + i] = b[i] & 3;
  }
  for (const Node& n : p.nodes) {
    if (n.pbs) {
      std::vector<uint8_t> t;
      uint32_t lognb;
      lut_tables(n.lut, t, lognb);
      uint32_t tlen = 16u >> lognb;
      uint32_t x = v[n.src0] % 32;
      for (uint32_t lu = 0; lu < (1u << lognb); lu++) v[n.dst + lu] = x < tlen ? t[lu * tlen + x] : 0xFFFF;
    } else {
      uint32_t x = v[n.src0], y = v[n.src1];
      v[n.dst] = (n.opcode == 0x05 ? x * n.mul_factor + y : x + y) % 32;
    }
  }
  for (size_t i = 0; i < p.out_rows.size(); i++) out[i] = v[p.out_rows[i]];
}

}  // namespace iop
