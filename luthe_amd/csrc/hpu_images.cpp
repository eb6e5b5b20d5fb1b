// HPU memory images and trace records (host only).
//
// * BSK image: the order in which the HPU's bsk_if streams GGSWs from HBM, as written by
//   the pe_pbs testbench generator (hw/module/hpu/simu/scripts/gen_stimuli.py:108-136):
//   for each CMUX i, output column j, stage iteration s, input row g, gadget level l
//   (least significant first), a chunk of PSI*R NTT-domain coefficients.  Within a
//   polynomial the NTT-domain order is the HW NTT's: natural order permuted by
//   bsk_order(cut) (sw/ntt/ntt_gf64_bsk_order.py:9-35).
// * LUT image: body polynomials only, gid * N * 8 bytes apart (pep_load_glwe.sv:9,83,204).
// * Ciphertext image: per pseudo-channel p, slot s at s * CT_MEM_BYTES (page aligned,
//   pem_common_param_pkg.sv:24-36); register-file word w of REGF_COEF_NB coefficients
//   sends coefficients [w*RC + p*RC/PC, w*RC + (p+1)*RC/PC) to channel p; channel 0 holds
//   one extra AXI word with the body (pem_load.sv:90-156).
// * KSK image: per pseudo-channel pc, the order in which ksk_if reads the key from HBM, as the
//   ksk_if testbench writes it (hw/module/ksk/module/ksk_if/simu/rtl/tb_ksk_if.sv:470-525):
//   for block column c, column x < LBX, block line l, level group lg, line y of this channel's
//   cuts, level z < LBZ, one MOD_KSK_W-bit coefficient; each (c, x) run is padded to whole AXI
//   words; an AXI word holds AXI4_DATA_W / KSK_ACS_W slots of LBZ coefficients
//   (ksk_if_common_param_pkg.sv:20-21).  Output column j = c*LBX + x (j = n is the body),
//   level = lg*LBZ + z (the decomposition's digit index, least significant first), and line
//   l*LBY + y = g*N + (N - 1 - m) holds mask coefficient m of polynomial g: "reverse order in
//   the BLWE_K dimension, on N basis" (ksk_manager.sv:11).  Columns, lines and levels beyond
//   n + 1, k*N and KS_L, and the padding, are zero.
// * Trace record: isc_trace_t (isc_common_param_pkg.sv:35-48), packed MSB-first,
//   written as 32-bit AXI words, least significant word first (trace_manager.sv:245-264).
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/luthe.h"

namespace {

constexpr uint32_t PAGE_BYTES = 4096;  // axi_if_common_param_pkg.sv:14

int log2_exact(uint32_t v) {
  int l = 0;
  while ((1u << l) < v) l++;
  return (1u << l) == v ? l : -1;
}

// order[pos] = natural NTT-domain index stored at position pos.
void bsk_order_rec(const uint32_t* cut, uint32_t ncut, std::vector<uint32_t>& out) {
  uint32_t s = 0;
  for (uint32_t i = 0; i < ncut; i++) s += cut[i];
  const uint32_t n = 1u << s;
  out.resize(n);
  if (ncut == 1) {
    for (uint32_t i = 0; i < n; i++) out[i] = i;
    return;
  }
  const uint32_t r0 = 1u << cut[0];          // stride of the first stage
  const uint32_t r1 = 1u << (s - cut[0]);     // size of the remaining sub-transform
  std::vector<uint32_t> inner;
  bsk_order_rec(cut + 1, ncut - 1, inner);
  // blocks j = 0..r0-1, each holding the sub-transform order scaled by r0, offset j
  for (uint32_t j = 0; j < n / r1; j++)
    for (uint32_t t = 0; t < r1; t++) out[(size_t)j * r1 + t] = r0 * inner[t] + j;
}

bool valid_cut(const luthe_params* p, const luthe_hpu_layout* lay) {
  if (!lay || lay->ntt_ncut < 1 || lay->ntt_ncut > 8) return false;
  uint32_t s = 0;
  for (uint32_t i = 0; i < lay->ntt_ncut; i++) {
    if (lay->ntt_cut[i] < 1) return false;
    s += lay->ntt_cut[i];
  }
  return (1u << s) == p->poly_size;
}

// position in this engine's NTT order (reference ntt() with cut [5,6]) of each natural index
std::vector<uint32_t> engine_pos_of_natural(uint32_t N) {
  const uint32_t cut56[2] = {5, 6};
  std::vector<uint32_t> ord;
  bsk_order_rec(cut56, 2, ord);
  std::vector<uint32_t> pos(N);
  for (uint32_t q = 0; q < N; q++) pos[ord[q]] = q;
  return pos;
}

int bsk_image_map(const luthe_params* p, const luthe_hpu_layout* lay, std::vector<size_t>& img_to_engine) {
  if (!p || p->poly_size != 2048 || !valid_cut(p, lay)) return LUTHE_ERR_ARG;
  const uint32_t N = p->poly_size, K1 = p->glwe_k + 1, L = p->pbs_level, chunk = lay->bsk_chunk;
  if (chunk == 0 || N % chunk) return LUTHE_ERR_ARG;
  std::vector<uint32_t> hw;
  bsk_order_rec(lay->ntt_cut, lay->ntt_ncut, hw);
  std::vector<uint32_t> pos = engine_pos_of_natural(N);
  const size_t per_i = (size_t)K1 * L * K1 * N;
  img_to_engine.resize(per_i);
  size_t q = 0;
  for (uint32_t j = 0; j < K1; j++)
    for (uint32_t s = 0; s < N / chunk; s++)
      for (uint32_t g = 0; g < K1; g++)
        for (uint32_t l = 0; l < L; l++)
          for (uint32_t c = 0; c < chunk; c++) {
            const uint32_t nat = hw[s * chunk + c];
            img_to_engine[q++] = (((size_t)g * L + l) * K1 + j) * N + pos[nat];
          }
  return LUTHE_OK;
}

struct CtGeom {
  uint32_t words_per_blwe, words_per_pc, slot_bytes, coef_per_word, rc_per_pc;
};

int ct_geom(const luthe_params* p, const luthe_hpu_layout* lay, CtGeom& g) {
  if (!p || !lay) return LUTHE_ERR_ARG;
  const uint32_t axi = lay->axi_bytes, pc = lay->pem_pc, rc = lay->regf_coef_nb;
  if (log2_exact(axi) < 3 || axi > 64 || (pc != 1 && pc != 2) || rc == 0 || rc % pc) return LUTHE_ERR_ARG;
  const uint32_t kN = p->glwe_k * p->poly_size;
  if (kN % rc) return LUTHE_ERR_ARG;
  g.coef_per_word = axi / 8;
  g.words_per_blwe = (kN * 8 + axi - 1) / axi;
  if (g.words_per_blwe % pc) return LUTHE_ERR_ARG;  // PEM_PC must divide the word count
  g.words_per_pc = g.words_per_blwe / pc;
  g.slot_bytes = (((g.words_per_pc + 1) * axi + PAGE_BYTES - 1) / PAGE_BYTES) * PAGE_BYTES;
  g.rc_per_pc = rc / pc;
  return LUTHE_OK;
}

// channel and word-stream index of mask coefficient e
void ct_place(const CtGeom& g, const luthe_hpu_layout* lay, uint32_t e, uint32_t& pc, uint32_t& idx) {
  const uint32_t rc = lay->regf_coef_nb;
  const uint32_t w = e / rc, in = e % rc;
  pc = in / g.rc_per_pc;
  idx = w * g.rc_per_pc + in % g.rc_per_pc;
}

constexpr int SYNC_ID_W = 4, LOCK_W = 7;  // POOL_SLOT_NB 64, MIN_IOP_SIZE 4

int lwe_k_w(uint32_t lwe_dim) {
  int w = 0;
  while ((1u << w) < lwe_dim) w++;
  return w == 0 ? 1 : w;
}

struct Bits {  // little-endian bit vector of 32-bit words
  uint32_t* w;
  void put(int at, int width, uint32_t v) {
    for (int b = 0; b < width; b++)
      if ((v >> b) & 1u) w[(at + b) >> 5] |= 1u << ((at + b) & 31);
  }
  uint32_t get(int at, int width) const {
    uint32_t v = 0;
    for (int b = 0; b < width; b++) v |= ((w[(at + b) >> 5] >> ((at + b) & 31)) & 1u) << b;
    return v;
  }
};

// KSK image geometry (tb_ksk_if.sv:40-43,63-98 with the cut counts dividing evenly).
struct KskGeom {
  uint32_t W = 0, acs_w = 0, cpw = 0;         // MOD_KSK_W, KSK_ACS_W, slots per AXI word
  uint32_t cols = 0, lines = 0, lgs = 0;      // KS_BLOCK_COL_NB, KS_BLOCK_LINE_NB, KS_LG_NB
  uint32_t y_per_pc = 0;                      // lines y of one channel (LBY / KSK_PC)
  size_t run = 0, run_padded = 0;             // coefficients per (c, x) run of one channel
  size_t pc_bytes = 0;
};

int ksk_geom(const luthe_params* p, const luthe_hpu_ksk_layout* lay, KskGeom& g) {
  if (!p || !lay) return LUTHE_ERR_ARG;
  if (lay->lbx < 1 || lay->lby < 1 || lay->lbz < 1 || lay->ksk_pc < 1 || lay->ksk_cut_nb < 1) return LUTHE_ERR_ARG;
  if (lay->ksk_cut_nb % lay->ksk_pc || lay->lby % lay->ksk_cut_nb) return LUTHE_ERR_ARG;  // even cuts
  if (log2_exact(lay->axi_bytes) < 3 || lay->axi_bytes > 128) return LUTHE_ERR_ARG;
  g.W = p->ks_mod_log;
  if ((uint64_t)lay->lbz * g.W > 64) return LUTHE_ERR_ARG;
  g.acs_w = lay->lbz * g.W > 32 ? 64 : 32;
  if (lay->axi_bytes * 8 < g.acs_w) return LUTHE_ERR_ARG;
  g.cpw = lay->axi_bytes * 8 / g.acs_w;
  g.cols = (p->lwe_dim + 1 + lay->lbx - 1) / lay->lbx;
  g.lines = (p->glwe_k * p->poly_size + lay->lby - 1) / lay->lby;
  g.lgs = (p->ks_level + lay->lbz - 1) / lay->lbz;
  g.y_per_pc = lay->lby / lay->ksk_pc;
  g.run = (size_t)g.y_per_pc * g.lines * g.lgs * lay->lbz;
  const size_t per_word = (size_t)g.cpw * lay->lbz;
  g.run_padded = (g.run + per_word - 1) / per_word * per_word;
  g.pc_bytes = (size_t)g.cols * lay->lbx * (g.run_padded / per_word) * lay->axi_bytes;
  return LUTHE_OK;
}

// Visits every coefficient position of channel pc's image in stream order: f(byte offset of
// the coefficient's slot, bit offset inside the slot, KSK index or SIZE_MAX for zero fill).
template <class F>
void ksk_walk(const luthe_params* p, const luthe_hpu_ksk_layout* lay, const KskGeom& g, uint32_t pc, F f) {
  const uint32_t N = p->poly_size, kN = p->glwe_k * N, n1 = p->lwe_dim + 1, L = p->ks_level;
  const size_t per_word = (size_t)g.cpw * lay->lbz;
  size_t word = 0;
  for (uint32_t c = 0; c < g.cols; c++)
    for (uint32_t x = 0; x < lay->lbx; x++) {
      size_t e = 0;  // coefficient index inside this run
      auto put = [&](size_t idx) {
        const size_t w = word + e / per_word, in = e % per_word;
        f(w * lay->axi_bytes + (in / lay->lbz) * (g.acs_w / 8), (uint32_t)(in % lay->lbz) * g.W, idx);
        e++;
      };
      const uint32_t j = c * lay->lbx + x;
      for (uint32_t l = 0; l < g.lines; l++)
        for (uint32_t lg = 0; lg < g.lgs; lg++)
          for (uint32_t yy = 0; yy < g.y_per_pc; yy++)
            for (uint32_t z = 0; z < lay->lbz; z++) {
              const uint32_t line = l * lay->lby + pc * g.y_per_pc + yy, lev = lg * lay->lbz + z;
              size_t idx = SIZE_MAX;
              if (j < n1 && line < kN && lev < L) {
                const uint32_t gpoly = line / N, m = N - 1 - line % N;  // reversed on N basis
                idx = ((size_t)(gpoly * N + m) * L + lev) * n1 + j;
              }
              put(idx);
            }
      while (e < g.run_padded) put(SIZE_MAX);
      word += g.run_padded / per_word;
    }
}

}  // namespace

extern "C" {

size_t luthe_hpu_ksk_image_bytes(const luthe_params* p, const luthe_hpu_ksk_layout* lay) {
  KskGeom g;
  return ksk_geom(p, lay, g) ? 0 : g.pc_bytes;
}

int luthe_hpu_ksk_to_image(const luthe_params* p, const luthe_hpu_ksk_layout* lay, const uint32_t* ksk,
                           uint8_t* const* images) {
  KskGeom g;
  int rc = ksk_geom(p, lay, g);
  if (rc) return rc;
  if (!ksk || !images) return LUTHE_ERR_ARG;
  const uint64_t wmask = (1ull << g.W) - 1;
  for (uint32_t pc = 0; pc < lay->ksk_pc; pc++) {
    if (!images[pc]) return LUTHE_ERR_ARG;
    memset(images[pc], 0, g.pc_bytes);
    uint8_t* img = images[pc];
    ksk_walk(p, lay, g, pc, [&](size_t byte, uint32_t bit, size_t idx) {
      if (idx == SIZE_MAX) return;
      uint64_t slot = 0;
      memcpy(&slot, img + byte, g.acs_w / 8);
      slot |= ((uint64_t)ksk[idx] & wmask) << bit;
      memcpy(img + byte, &slot, g.acs_w / 8);  // little-endian AXI words
    });
  }
  return LUTHE_OK;
}

int luthe_hpu_ksk_from_image(const luthe_params* p, const luthe_hpu_ksk_layout* lay, const uint8_t* const* images,
                             uint32_t* ksk) {
  KskGeom g;
  int rc = ksk_geom(p, lay, g);
  if (rc) return rc;
  if (!ksk || !images) return LUTHE_ERR_ARG;
  const uint64_t wmask = (1ull << g.W) - 1;
  for (uint32_t pc = 0; pc < lay->ksk_pc; pc++) {
    if (!images[pc]) return LUTHE_ERR_ARG;
    const uint8_t* img = images[pc];
    ksk_walk(p, lay, g, pc, [&](size_t byte, uint32_t bit, size_t idx) {
      if (idx == SIZE_MAX) return;
      uint64_t slot = 0;
      memcpy(&slot, img + byte, g.acs_w / 8);
      ksk[idx] = (uint32_t)((slot >> bit) & wmask);
    });
  }
  return LUTHE_OK;
}

int luthe_upload_ksk_hpu_image(luthe_ctx* ctx, const luthe_hpu_ksk_layout* lay, const uint8_t* const* images) {
  if (!ctx) return LUTHE_ERR_ARG;
  luthe_params p;
  luthe_get_params(ctx, &p);
  std::vector<uint32_t> ksk((size_t)p.glwe_k * p.poly_size * p.ks_level * (p.lwe_dim + 1));
  int rc = luthe_hpu_ksk_from_image(&p, lay, images, ksk.data());
  if (rc) return rc;
  return luthe_upload_ksk(ctx, ksk.data());
}


int luthe_hpu_bsk_order(const uint32_t* cut, uint32_t ncut, uint32_t* order) {
  if (!cut || !order || ncut < 1 || ncut > 16) return LUTHE_ERR_ARG;
  uint32_t s = 0;
  for (uint32_t i = 0; i < ncut; i++) {
    if (cut[i] < 1) return LUTHE_ERR_ARG;
    s += cut[i];
  }
  if (s > 16) return LUTHE_ERR_ARG;
  std::vector<uint32_t> o;
  bsk_order_rec(cut, ncut, o);
  memcpy(order, o.data(), o.size() * sizeof(uint32_t));
  return LUTHE_OK;
}

size_t luthe_hpu_bsk_image_words(const luthe_params* p) {
  if (!p) return 0;
  const size_t K1 = p->glwe_k + 1;
  return (size_t)p->lwe_dim * K1 * p->pbs_level * K1 * p->poly_size;
}

int luthe_hpu_bsk_from_image(const luthe_params* p, const luthe_hpu_layout* lay, const uint64_t* image,
                             uint64_t* bsk_ntt) {
  if (!image || !bsk_ntt) return LUTHE_ERR_ARG;
  std::vector<size_t> map;
  int rc = bsk_image_map(p, lay, map);
  if (rc) return rc;
  const size_t per_i = map.size();
  for (uint32_t i = 0; i < p->lwe_dim; i++) {
    const uint64_t* src = image + (size_t)i * per_i;
    uint64_t* dst = bsk_ntt + (size_t)i * per_i;
    for (size_t q = 0; q < per_i; q++) dst[map[q]] = src[q];
  }
  return LUTHE_OK;
}

int luthe_hpu_bsk_to_image(const luthe_params* p, const luthe_hpu_layout* lay, const uint64_t* bsk_ntt,
                           uint64_t* image) {
  if (!image || !bsk_ntt) return LUTHE_ERR_ARG;
  std::vector<size_t> map;
  int rc = bsk_image_map(p, lay, map);
  if (rc) return rc;
  const size_t per_i = map.size();
  for (uint32_t i = 0; i < p->lwe_dim; i++) {
    const uint64_t* src = bsk_ntt + (size_t)i * per_i;
    uint64_t* dst = image + (size_t)i * per_i;
    for (size_t q = 0; q < per_i; q++) dst[q] = src[map[q]];
  }
  return LUTHE_OK;
}

int luthe_upload_bsk_hpu_image(luthe_ctx* ctx, const luthe_hpu_layout* lay, const uint64_t* image) {
  if (!ctx || !lay || !image) return LUTHE_ERR_ARG;
  luthe_params p;
  int rc = luthe_get_params(ctx, &p);
  if (rc) return rc;
  std::vector<uint64_t> bsk(luthe_hpu_bsk_image_words(&p));
  rc = luthe_hpu_bsk_from_image(&p, lay, image, bsk.data());
  if (rc) return rc;
  return luthe_upload_bsk(ctx, bsk.data(), LUTHE_BSK_NTT);
}

size_t luthe_hpu_ct_slot_bytes(const luthe_params* p, const luthe_hpu_layout* lay) {
  CtGeom g;
  return ct_geom(p, lay, g) ? 0 : g.slot_bytes;
}

int luthe_hpu_ct_to_image(const luthe_params* p, const luthe_hpu_layout* lay, const uint64_t* blwe, uint32_t count,
                          uint8_t* image) {
  CtGeom g;
  if (!blwe || !image || ct_geom(p, lay, g)) return LUTHE_ERR_ARG;
  const uint32_t kN = p->glwe_k * p->poly_size;
  const size_t region = (size_t)count * g.slot_bytes;
  memset(image, 0, region * lay->pem_pc);
  for (uint32_t s = 0; s < count; s++) {
    const uint64_t* ct = blwe + (size_t)s * (kN + 1);
    for (uint32_t e = 0; e < kN; e++) {
      uint32_t pc, idx;
      ct_place(g, lay, e, pc, idx);
      memcpy(image + pc * region + (size_t)s * g.slot_bytes + (size_t)idx * 8, &ct[e], 8);
    }
    // body: first coefficient of the extra AXI word of channel 0
    memcpy(image + (size_t)s * g.slot_bytes + (size_t)g.words_per_pc * lay->axi_bytes, &ct[kN], 8);
  }
  return LUTHE_OK;
}

int luthe_hpu_ct_from_image(const luthe_params* p, const luthe_hpu_layout* lay, const uint8_t* image, uint32_t count,
                            uint64_t* blwe) {
  CtGeom g;
  if (!blwe || !image || ct_geom(p, lay, g)) return LUTHE_ERR_ARG;
  const uint32_t kN = p->glwe_k * p->poly_size;
  const size_t region = (size_t)count * g.slot_bytes;
  for (uint32_t s = 0; s < count; s++) {
    uint64_t* ct = blwe + (size_t)s * (kN + 1);
    for (uint32_t e = 0; e < kN; e++) {
      uint32_t pc, idx;
      ct_place(g, lay, e, pc, idx);
      memcpy(&ct[e], image + pc * region + (size_t)s * g.slot_bytes + (size_t)idx * 8, 8);
    }
    memcpy(&ct[kN], image + (size_t)s * g.slot_bytes + (size_t)g.words_per_pc * lay->axi_bytes, 8);
  }
  return LUTHE_OK;
}

int luthe_hpu_lut_from_image(const luthe_params* p, const uint64_t* bodies, uint32_t count, uint64_t* glwe) {
  if (!p || !bodies || !glwe) return LUTHE_ERR_ARG;
  const size_t N = p->poly_size, kN = (size_t)p->glwe_k * N;
  for (uint32_t c = 0; c < count; c++) {
    uint64_t* g = glwe + (size_t)c * (kN + N);
    memset(g, 0, kN * sizeof(uint64_t));  // trivial GLWE: the mask is not stored
    memcpy(g + kN, bodies + (size_t)c * N, N * sizeof(uint64_t));
  }
  return LUTHE_OK;
}

int luthe_write_lut_hpu_image(luthe_ctx* ctx, uint32_t gid, const uint64_t* bodies, uint32_t count) {
  if (!ctx || !bodies) return LUTHE_ERR_ARG;
  luthe_params p;
  int rc = luthe_get_params(ctx, &p);
  if (rc) return rc;
  std::vector<uint64_t> g((size_t)count * (p.glwe_k + 1) * p.poly_size);
  rc = luthe_hpu_lut_from_image(&p, bodies, count, g.data());
  if (rc) return rc;
  return luthe_write_lut(ctx, gid, g.data(), count);
}

int luthe_trace_record_bits(uint32_t lwe_dim) {
  if (lwe_dim < 1) return LUTHE_ERR_ARG;
  return 32 + 32 + 3 + SYNC_ID_W + 3 * LOCK_W + 3 + lwe_k_w(lwe_dim);
}

int luthe_trace_pack(uint32_t lwe_dim, const luthe_trace_event* ev, size_t n, uint32_t* words) {
  if (lwe_dim < 1 || (n && (!ev || !words))) return LUTHE_ERR_ARG;
  const int kw = lwe_k_w(lwe_dim), bits = luthe_trace_record_bits(lwe_dim), nw = (bits + 31) / 32;
  for (size_t i = 0; i < n; i++) {
    Bits b{words + i * nw};
    memset(b.w, 0, nw * sizeof(uint32_t));
    const luthe_trace_event& e = ev[i];
    int at = 0;  // from the least significant field (last declared) upwards
    b.put(at, kw, e.pe_reserved), at += kw;
    b.put(at, 1, e.pdg), at += 1;
    b.put(at, 1, e.rd_pdg), at += 1;
    b.put(at, 1, e.vld), at += 1;
    b.put(at, LOCK_W, e.issue_lock), at += LOCK_W;
    b.put(at, LOCK_W, e.wr_lock), at += LOCK_W;
    b.put(at, LOCK_W, e.rd_lock), at += LOCK_W;
    b.put(at, SYNC_ID_W, e.sync_id), at += SYNC_ID_W;
    b.put(at, 3, e.cmd), at += 3;
    b.put(at, 32, e.insn), at += 32;
    b.put(at, 32, e.timestamp);
  }
  return nw;
}

int luthe_trace_unpack(uint32_t lwe_dim, const uint32_t* words, size_t n, luthe_trace_event* ev) {
  if (lwe_dim < 1 || (n && (!ev || !words))) return LUTHE_ERR_ARG;
  const int kw = lwe_k_w(lwe_dim), bits = luthe_trace_record_bits(lwe_dim), nw = (bits + 31) / 32;
  for (size_t i = 0; i < n; i++) {
    Bits b{const_cast<uint32_t*>(words + i * nw)};
    luthe_trace_event& e = ev[i];
    memset(&e, 0, sizeof e);
    int at = 0;
    e.pe_reserved = b.get(at, kw), at += kw;
    e.pdg = (uint8_t)b.get(at, 1), at += 1;
    e.rd_pdg = (uint8_t)b.get(at, 1), at += 1;
    e.vld = (uint8_t)b.get(at, 1), at += 1;
    e.issue_lock = (uint8_t)b.get(at, LOCK_W), at += LOCK_W;
    e.wr_lock = (uint8_t)b.get(at, LOCK_W), at += LOCK_W;
    e.rd_lock = (uint8_t)b.get(at, LOCK_W), at += LOCK_W;
    e.sync_id = (uint8_t)b.get(at, SYNC_ID_W), at += SYNC_ID_W;
    e.cmd = (uint8_t)b.get(at, 3), at += 3;
    e.insn = b.get(at, 32), at += 32;
    e.timestamp = b.get(at, 32);
  }
  return nw;
}

}  // extern "C"
