// C ABI of the MI355X PBS engine: context, device memories, key upload, batched hot
// path and the DOp/IOp front end.
//
// DOp execution model.  A submitted stream is validated whole, then every register
// write is renamed to a fresh row after the ciphertext memory and the stream is
// scheduled by dependency level: a DOp's level is one more than the latest level of the
// DOps producing its operands (and, for LD/ST, of earlier conflicting accesses to the
// same slot).  DOps of one level are independent, so a level runs as at most three
// launches: all its LD/ST (one copy kernel), all its ALU DOps (one ALU kernel) and all
// its PBS (one key-switch + one blind-rotation launch); a last copy launch writes the
// registers' final values back.  Each PBS is a pure function of its source row and LUT
// and every ciphertext runs its CMUX chain in the fixed order i = 0..n-1, so the results
// equal in-order execution and do not depend on which other DOps share a batch.  This
// replaces the HPU instruction scheduler's issue logic (hw/module/instruction_scheduler),
// which orders register reuse with locks, and the BPIP batch former
// (pep_sequencer.sv:584-608) with a host-side renamed level schedule.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/luthe.h"
#include "ctx.h"
#include "dop.h"
#include "kernels.h"
#include "ntt2048.h"

using lk::AluOp;
using lk::CopyOp;
using lk::PbsJob;

namespace lx {

int fail(luthe_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    std::lock_guard<std::mutex> g(c->err_mu);  // an attached server's thread reports here too
    c->err = buf;
  }
  return code;
}

int dev_alloc(luthe_ctx* c, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(c, LUTHE_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  c->dev_bytes += bytes;
  return LUTHE_OK;
}

int check_params(const luthe_params* p, std::string& why) {
  if (p->glwe_k != 1) return why = "glwe_k must be 1", -1;
  if (p->poly_size != 2048) return why = "poly_size must be 2048", -1;
  if (p->lwe_dim < 1 || p->lwe_dim > 4095) return why = "lwe_dim out of range [1, 4095]", -1;
  if (p->pbs_level < 1 || p->pbs_level > 2) return why = "pbs_level must be 1 or 2", -1;
  if (p->pbs_base_log < 1 || p->pbs_base_log * p->pbs_level >= 64) return why = "pbs_base_log*pbs_level must be < 64", -1;
  if (p->ks_level < 1 || p->ks_level > 8) return why = "ks_level must be in [1, 8]", -1;
  // key-switch digits are held as int8 (|d| <= 2^(b-1)) and multiplied with 24-bit MACs
  if (p->ks_base_log < 1 || p->ks_base_log > 7) return why = "ks_base_log must be in [1, 7]", -1;
  if (p->ks_base_log * p->ks_level >= 64) return why = "ks_base_log*ks_level must be < 64", -1;
  if (p->ks_mod_log < 13 || p->ks_mod_log > 24) return why = "ks_mod_log must be in [13, 24]", -1;
  // the gadget weights 2^(w - b*(L-l)) of the KSK live mod 2^w
  if (p->ks_base_log * p->ks_level > p->ks_mod_log) return why = "ks_base_log*ks_level must be <= ks_mod_log", -1;
  if (p->msg_bits + p->carry_bits < 1 || p->msg_bits + p->carry_bits > 8) return why = "msg_bits+carry_bits in [1, 8]", -1;
  if (p->regfile_size < 1 || p->regfile_size > 128) return why = "regfile_size in [1, 128]", -1;
  if (p->ct_slots > 65536) return why = "ct_slots <= 65536", -1;
  if (p->lut_slots < 1 || p->lut_slots > 4096) return why = "lut_slots in [1, 4096]", -1;
  return 0;
}

void retire_scratch(luthe_ctx* c, void* p, size_t bytes) {
  if (!p) return;
  if (!c->graphs.empty()) {
    c->retired.emplace_back(p, bytes);  // a captured graph may still read or write it
    return;
  }
  (void)hipFree(p);
  c->dev_bytes -= bytes;
}

// Scratch grows by allocating the new buffer first: on failure the context keeps the old
// buffer and capacity.
int ensure_lwe2n(luthe_ctx* c, PbsScratch& sc, size_t n) {
  if (n <= sc.lwe2n_cap) return LUTHE_OK;
  const size_t cap = std::max(n, sc.lwe2n_cap * 2);
  uint16_t* nb = nullptr;
  int rc = dev_alloc(c, (void**)&nb, cap * c->lwe_stride * sizeof(uint16_t));
  if (rc) return rc;
  if (sc.lwe2n) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    retire_scratch(c, sc.lwe2n, sc.lwe2n_cap * c->lwe_stride * sizeof(uint16_t));
  }
  sc.lwe2n = nb;
  sc.lwe2n_cap = cap;
  return LUTHE_OK;
}

// Rows after the ciphertext memory that hold renamed register values of a DOp stream;
// growing keeps the register file and ciphertext memory.
int ensure_rename_rows(luthe_ctx* c, size_t n) {
  if (n <= c->rename_cap) return LUTHE_OK;
  const size_t keep = (size_t)c->prm.regfile_size + c->prm.ct_slots;
  const size_t cap = std::max(n, c->rename_cap * 2);
  uint64_t* nm = nullptr;
  int rc = dev_alloc(c, (void**)&nm, (keep + cap) * c->blen * 8);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(nm, c->d_mem, keep * c->blen * 8, hipMemcpyDeviceToDevice));
  HIPCHK(c, hipFree(c->d_mem));
  c->dev_bytes -= (keep + c->rename_cap) * c->blen * 8;
  c->d_mem = nm;
  c->rename_cap = cap;
  return LUTHE_OK;
}

int ensure_jobs(luthe_ctx* c, size_t n) {
  if (n <= c->jobs_cap) return LUTHE_OK;
  const size_t cap = std::max(n, c->jobs_cap * 2);
  PbsJob* nb = nullptr;
  int rc = dev_alloc(c, (void**)&nb, cap * sizeof(PbsJob));
  if (rc) return rc;
  if (c->d_jobs) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    retire_scratch(c, c->d_jobs, c->jobs_cap * sizeof(PbsJob));
  }
  c->d_jobs = nb;
  c->jobs_cap = cap;
  return LUTHE_OK;
}

// Words of the split key switch's partial sums for a batch of n ciphertexts.
static size_t ks_part_words(const luthe_ctx* c, size_t n) {
  const int kN = (int)(c->prm.glwe_k * c->prm.poly_size), nl = (int)c->prm.lwe_dim;
  const int s = lk::keyswitch_split((int)std::min<size_t>(n, 1u << 20), kN, nl);
  return s > 1 ? (size_t)s * n * (nl + 1) : 0;
}

int ensure_ks_part(luthe_ctx* c, PbsScratch& sc, size_t n) {
  if (n <= sc.ks_part_for) return LUTHE_OK;
  // the largest split scratch of any batch size up to n: the split depends on the number of
  // 32-ciphertext tiles only, so the maximum sits at the last size of some tile count
  // (and the finer splits of batches of at most 4 and at most 16 ciphertexts)
  size_t need = std::max(ks_part_words(c, std::min<size_t>(n, 4)), ks_part_words(c, std::min<size_t>(n, 16)));
  for (size_t tiles = 1; tiles <= (n + 31) / 32; tiles++) {
    const size_t w = ks_part_words(c, std::min(n, tiles * 32));
    if (w == 0) break;  // more tiles: no split
    need = std::max(need, w);
  }
  if (need <= sc.ks_part_cap) {
    sc.ks_part_for = n;
    return LUTHE_OK;
  }
  uint32_t* nb = nullptr;
  int rc = dev_alloc(c, (void**)&nb, need * sizeof(uint32_t));
  if (rc) return rc;
  if (sc.ks_part) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    retire_scratch(c, sc.ks_part, sc.ks_part_cap * sizeof(uint32_t));
  }
  sc.ks_part = nb;
  sc.ks_part_cap = need;
  sc.ks_part_for = n;
  return LUTHE_OK;
}

int split_max(luthe_ctx* c) {
  if (c->br_split_max < 0) {
    (void)hipSetDevice(c->device);
    c->br_split_max = lk::blind_rotate_split_slots();
  }
  return c->br_split_max;
}

int ensure_xch(luthe_ctx* c, PbsScratch& sc, size_t n) {
  n = std::min(n, (size_t)std::max(split_max(c), 0));
  if (n <= sc.xch_cap) return LUTHE_OK;
  const size_t cap = std::min(std::max(n, sc.xch_cap * 2), (size_t)split_max(c));
  const size_t bytes = lk::blind_rotate_split_words((int)c->prm.pbs_level) * 8;
  uint64_t* nb = nullptr;
  int rc = dev_alloc(c, (void**)&nb, cap * bytes);
  if (rc) return rc;
  if (sc.xch) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    retire_scratch(c, sc.xch, sc.xch_cap * bytes);
  }
  sc.xch = nb;
  sc.xch_cap = cap;
  return LUTHE_OK;
}

std::string last_error(luthe_ctx* c) {
  std::lock_guard<std::mutex> g(c->err_mu);
  return c->err;
}

int check_device_status(luthe_ctx* c) {
  if (!c->h_status || __atomic_exchange_n(c->h_status, 0u, __ATOMIC_ACQ_REL) == 0) return LUTHE_OK;
  return fail(c, LUTHE_ERR_HIP,
              "split blind rotation: a workgroup gave up waiting for its partner (not resident?); "
              "the outputs of that launch are undefined");
}

// Staging of per-submission descriptors: host pinned buffer -> device buffer.
int stage_upload(luthe_ctx* c, const std::vector<uint8_t>& bytes) {
  size_t n = bytes.size();
  if (n == 0) return LUTHE_OK;
  // the pinned buffer may still be read by the previous submission's copy
  HIPCHK(c, hipEventSynchronize(c->stage_copied));
  if (n > c->h_stage_cap) {
    const size_t cap = std::max(n, c->h_stage_cap * 2);
    uint8_t* nh = nullptr;
    HIPCHK(c, hipHostMalloc((void**)&nh, cap, hipHostMallocDefault));
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    c->h_stage = nh;
    c->h_stage_cap = cap;
  }
  if (n > c->stage_cap) {
    const size_t cap = std::max(n, c->stage_cap * 2);
    uint8_t* nd = nullptr;
    int rc = dev_alloc(c, (void**)&nd, cap);
    if (rc) return rc;
    if (c->d_stage) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      (void)hipFree(c->d_stage);
      c->dev_bytes -= c->stage_cap;
    }
    c->d_stage = nd;
    c->stage_cap = cap;
  }
  memcpy(c->h_stage, bytes.data(), n);
  HIPCHK(c, hipMemcpyAsync(c->d_stage, c->h_stage, n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipEventRecord(c->stage_copied, c->stream));
  return LUTHE_OK;
}

hipStream_t pick_stream(luthe_ctx* c, void* s) { return s ? (hipStream_t)s : c->stream; }

uint32_t useful_bits(const luthe_ctx* c) { return c->prm.msg_bits + c->prm.carry_bits + 1; }

// Throughput kernel: the 512-thread two-per-CU pair variant by default at PBS level 2 (its
// four waves per SIMD against the 256-thread kernel's two there), the 256-thread kernel at
// level 1; luthe_set_tuning / the autotuner override.
bool use_pair(const luthe_ctx* c) { return c->br_pair < 0 ? c->prm.pbs_level == 2 : c->br_pair == 1; }

int latency_max(luthe_ctx* c) {
  if (c->br_latency_max < 0) {
    (void)hipSetDevice(c->device);
    c->br_latency_max = lk::blind_rotate_latency_slots((int)c->prm.pbs_level);
  }
  return c->br_latency_max;
}

lk::KsArgs ks_args(const luthe_ctx* c, int lwe_stride) {
  const luthe_params& p = c->prm;
  return lk::KsArgs{(int)(p.glwe_k * p.poly_size), (int)p.lwe_dim, lwe_stride, p.ks_base_log, p.ks_mod_log, 12};
}

// The matrix-core key switch for batches of at least this many ciphertexts (when the KSK
// planes exist); smaller ones take the VALU kernel, split over the inputs.  Default: one
// ciphertext per CU (measured on MI355X, 256 CUs: 0.13 against 0.24 ms at 256 ciphertexts, a
// tie at 128, the VALU kernel ahead below).
int ks_mfma_min(luthe_ctx* c) {
  if (c->ks_mfma_min < 0) {
    int cus = 0;
    (void)hipSetDevice(c->device);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus < 1)
      cus = 256;
    c->ks_mfma_min = cus;
  }
  return c->ks_mfma_min;
}

int launch_ks(luthe_ctx* c, const PbsJob* d_jobs, int n, uint16_t* lwe2n, int lwe_stride, uint32_t* ks_part,
              hipStream_t s) {
  const luthe_params& p = c->prm;
  const lk::KsArgs ka = ks_args(c, lwe_stride);
  if (c->d_kskp && n >= ks_mfma_min(c)) {
    HIPCHK(c, lk::launch_keyswitch_mfma(d_jobs, n, c->d_kskp, lwe2n, ka, (int)p.ks_level, s));
    return LUTHE_OK;
  }
  const int nsplit = ks_part ? lk::keyswitch_split(n, ka.kN, ka.n_lwe) : 1;
  HIPCHK(c, lk::launch_keyswitch(d_jobs, n, c->d_ksk, lwe2n, ka, (int)p.ks_level, s, ks_part, nsplit));
  return LUTHE_OK;
}

// KS + BR over jobs already on the device, with the mod-switched LWEs in `lwe2n`
// (n rows of lwe_stride values).  Allocates nothing.
int run_pbs_on(luthe_ctx* c, const PbsJob* d_jobs, int n, hipStream_t s,
               uint16_t* lwe2n, uint32_t* ks_part, uint64_t* xch) {
  if (!c->bsk_ok || !c->ksk_ok) return fail(c, LUTHE_ERR_STATE, "keys not uploaded");
  if (n == 0) return LUTHE_OK;
  const luthe_params& p = c->prm;
  int rc = launch_ks(c, d_jobs, n, lwe2n, c->lwe_stride, ks_part, s);
  if (rc) return rc;
  lk::BrArgs ba{(int)p.lwe_dim, c->lwe_stride, p.pbs_base_log, c->br_split_wait_us};
  HIPCHK(c, lk::launch_blind_rotate(c->d_bsk, c->d_luts, lwe2n, d_jobs, n, c->d_tab, ba,
                                    (int)p.pbs_level, c->br_grid_override, s, latency_max(c), use_pair(c),
                                    split_max(c), xch, c->d_status, c->split_local));
  {
    std::lock_guard<std::mutex> g(c->cnt_mu);  // an attached server's thread counts here too
    c->cnt.pbs += (uint64_t)n;
    c->cnt.pbs_batches += 1;
    c->cnt.pbs_max_batch = std::max<uint64_t>(c->cnt.pbs_max_batch, (uint64_t)n);
  }
  return LUTHE_OK;
}

// KS + BR over jobs already on the device, on the context's (grown as needed) scratch.
int run_pbs(luthe_ctx* c, PbsScratch& sc, const PbsJob* d_jobs, int n, hipStream_t s) {
  if (!c->bsk_ok || !c->ksk_ok) return fail(c, LUTHE_ERR_STATE, "keys not uploaded");
  if (n == 0) return LUTHE_OK;
  int rc = ensure_lwe2n(c, sc, (size_t)n);
  if (!rc) rc = ensure_ks_part(c, sc, (size_t)n);
  if (!rc) rc = ensure_xch(c, sc, (size_t)n);
  if (rc) return rc;
  return run_pbs_on(c, d_jobs, n, s, sc.lwe2n, sc.ks_part, (size_t)n <= sc.xch_cap ? sc.xch : nullptr);
}

void trace_push(luthe_ctx* c, const luthe_trace_event& e) {
  if (!c->trace_cap) return;
  if (c->trace.size() >= c->trace_cap) c->trace.pop_front();
  c->trace.push_back(e);
}

uint32_t trace_now(luthe_ctx* c) {
  return (uint32_t)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() -
                                                                        c->trace_t0).count();
}

// RETIRE events of a completed submission.
void trace_retire(luthe_ctx* c, uint64_t ticket) {
  auto it = c->trace_retire.find(ticket);
  if (it == c->trace_retire.end()) return;
  const uint32_t now = trace_now(c);
  for (luthe_trace_event e : it->second) {
    e.timestamp = now;
    trace_push(c, e);
  }
  c->trace_retire.erase(it);
}

}  // namespace lx

using namespace lx;

// ===================================================================================
extern "C" {

void luthe_params_default(luthe_params* p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  p->lwe_dim = 724;
  p->glwe_k = 1;
  p->poly_size = 2048;
  p->pbs_base_log = 23;
  p->pbs_level = 1;
  p->ks_base_log = 3;
  p->ks_level = 5;
  p->ks_mod_log = 21;
  p->msg_bits = 2;
  p->carry_bits = 2;
  p->lwe_noise_std = 1.0 / 32768.0;            // 2^-15
  p->glwe_noise_std = 1.0 / 2251799813685248.0;  // 2^-51
  p->regfile_size = 64;
  p->ct_slots = 4096;
  p->lut_slots = 256;
}

int luthe_create(const luthe_params* p, int device, luthe_ctx** out) {
  if (!p || !out) return LUTHE_ERR_ARG;
  *out = nullptr;
  std::string why;
  if (check_params(p, why)) return LUTHE_ERR_PARAMS;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return LUTHE_ERR_NODEV;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return LUTHE_ERR_NODEV;
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return LUTHE_ERR_NODEV;
  luthe_ctx* c = new luthe_ctx();
  c->prm = *p;
  c->device = device;
  c->blen = (size_t)p->glwe_k * p->poly_size + 1;
  c->lut_len = (size_t)(p->glwe_k + 1) * p->poly_size;
  c->lwe_stride = (int)((p->lwe_dim + 1 + 7) & ~7u);
  int rc = LUTHE_OK;
  auto bail = [&](int code) {
    luthe_destroy(c);
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(LUTHE_ERR_HIP);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(LUTHE_ERR_HIP);
  if ((rc = noalloc_init(c, device))) return bail(rc);
  if (hipHostMalloc((void**)&c->h_status, 64, hipHostMallocMapped) != hipSuccess) return bail(LUTHE_ERR_HIP);
  *c->h_status = 0;
  if (hipHostGetDevicePointer((void**)&c->d_status, c->h_status, 0) != hipSuccess) return bail(LUTHE_ERR_HIP);
  if (hipEventCreateWithFlags(&c->stage_copied, hipEventDisableTiming) != hipSuccess) return bail(LUTHE_ERR_HIP);
  size_t rows = (size_t)p->regfile_size + p->ct_slots;
  if ((rc = dev_alloc(c, (void**)&c->d_mem, rows * c->blen * 8))) return bail(rc);
  if ((rc = dev_alloc(c, (void**)&c->d_luts, (size_t)p->lut_slots * c->lut_len * 8))) return bail(rc);
  if (hipMemset(c->d_mem, 0, rows * c->blen * 8) != hipSuccess) return bail(LUTHE_ERR_HIP);
  if (hipMemset(c->d_luts, 0, (size_t)p->lut_slots * c->lut_len * 8) != hipSuccess) return bail(LUTHE_ERR_HIP);
  std::vector<uint64_t> tab(ntt::TABLE_WORDS);
  ntt::make_twist_tables(tab.data());
  if ((rc = dev_alloc(c, (void**)&c->d_tab, tab.size() * 8))) return bail(rc);
  if (hipMemcpy(c->d_tab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice) != hipSuccess) return bail(LUTHE_ERR_HIP);
  c->split_local = c->split_local_ok = lk::split_pairs_share_xcc(c->stream);
  *out = c;
  return LUTHE_OK;
}

int luthe_destroy(luthe_ctx* c) {
  if (!c) return LUTHE_ERR_ARG;
  if (c->server) return fail(c, LUTHE_ERR_STATE, "destroy the attached server first");
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->groups)  // open (unlaunched) tickets are dropped
    if (kv.second.done) (void)hipEventDestroy(kv.second.done);
  for (luthe_graph* g : c->graphs) g->ctx = nullptr;  // their scratch is freed below
  for (auto& r : c->retired) (void)hipFree(r.first);
  void* bufs[] = {c->d_mem,     c->d_luts,      c->d_bsk,   c->d_ksk, c->d_kskp, c->d_tab,     c->bs.lwe2n,
                  c->bs.ks_part, c->ds.lwe2n, c->ds.ks_part, c->d_jobs, c->d_stage, c->bs.xch, c->ds.xch};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  if (c->h_status) (void)hipHostFree(c->h_status);
  if (c->stage_copied) (void)hipEventDestroy(c->stage_copied);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return LUTHE_OK;
}

const char* luthe_last_error(const luthe_ctx* c) { return c ? c->err.c_str() : "null context"; }

int luthe_get_params(const luthe_ctx* c, luthe_params* out) {
  if (!c || !out) return LUTHE_ERR_ARG;
  *out = c->prm;
  return LUTHE_OK;
}

int luthe_get_counters(const luthe_ctx* c, luthe_counters* out) {
  if (!c || !out) return LUTHE_ERR_ARG;
  std::lock_guard<std::mutex> g(c->cnt_mu);
  *out = c->cnt;
  return LUTHE_OK;
}

uint64_t luthe_device_bytes(const luthe_ctx* c) { return c ? c->dev_bytes : 0; }

void* luthe_stream(luthe_ctx* c) { return c ? (void*)c->stream : nullptr; }

int luthe_set_br_grid(luthe_ctx* c, int workgroups) {
  if (!c || workgroups < 0) return LUTHE_ERR_ARG;
  if (c->server) return fail(c, LUTHE_ERR_BUSY, "keys and launch settings are fixed while a server is attached");
  c->br_grid_override = workgroups;
  return LUTHE_OK;
}

int luthe_set_br_latency_max(luthe_ctx* c, int max_cohort) {
  if (!c || max_cohort < 0) return LUTHE_ERR_ARG;
  if (c->server) return fail(c, LUTHE_ERR_BUSY, "keys and launch settings are fixed while a server is attached");
  c->br_latency_max = max_cohort;
  return LUTHE_OK;
}

int luthe_set_br_split_local(luthe_ctx* c, int local) {
  if (!c || (local != 0 && local != 1)) return LUTHE_ERR_ARG;
  if (c->server) return fail(c, LUTHE_ERR_BUSY, "keys and launch settings are fixed while a server is attached");
  if (local && !c->split_local_ok)
    return fail(c, LUTHE_ERR_STATE, "the split blind rotation's workgroup pairs do not share an XCC on this device");
  c->split_local = local != 0;
  return LUTHE_OK;
}

int luthe_get_br_split_local(luthe_ctx* c) {
  if (!c) return LUTHE_ERR_ARG;
  return c->split_local ? 1 : 0;
}

int luthe_get_br_latency_max(luthe_ctx* c) {
  if (!c) return LUTHE_ERR_ARG;
  return latency_max(c);
}

int luthe_get_br_grid(luthe_ctx* c) {
  if (!c) return LUTHE_ERR_ARG;
  if (c->br_grid_override) return c->br_grid_override;
  (void)hipSetDevice(c->device);
  return lk::blind_rotate_occupancy((int)c->prm.pbs_level, use_pair(c));
}

int luthe_microbench_modmul(luthe_ctx* c, double* gmodmul_per_s) {
  if (!c || !gmodmul_per_s) return LUTHE_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  int cus = 0;
  HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  const int blocks = cus * 8, iters = 4096;
  uint64_t* d_out = nullptr;
  int rc = dev_alloc(c, (void**)&d_out, 64);
  if (rc) return rc;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  HIPCHK(c, lk::launch_modmul_peak(d_out, blocks, 64, c->stream));  // warm up
  HIPCHK(c, hipEventRecord(e0, c->stream));
  HIPCHK(c, lk::launch_modmul_peak(d_out, blocks, iters, c->stream));
  HIPCHK(c, hipEventRecord(e1, c->stream));
  HIPCHK(c, hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(d_out);
  c->dev_bytes -= 64;
  *gmodmul_per_s = (double)blocks * 256.0 * 8.0 * iters / (ms * 1e-3) / 1e9;
  return LUTHE_OK;
}

int luthe_sync(luthe_ctx* c) {
  if (!c) return LUTHE_ERR_ARG;
  int rc = dop_progress(c, true);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_device_status(c);
}

// ---------------------------------------------------------------- keys and memories
int luthe_upload_bsk(luthe_ctx* c, const uint64_t* bsk, int layout) {
  if (!c || !bsk) return LUTHE_ERR_ARG;
  if (c->server) return fail(c, LUTHE_ERR_BUSY, "keys and launch settings are fixed while a server is attached");
  if (layout != LUTHE_BSK_COEF && layout != LUTHE_BSK_NTT) return fail(c, LUTHE_ERR_ARG, "bad BSK layout %d", layout);
  const luthe_params& p = c->prm;
  const int K1 = (int)p.glwe_k + 1, rows = K1 * (int)p.pbs_level;
  size_t elems = (size_t)p.lwe_dim * rows * K1 * p.poly_size;
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->d_bsk) {
    int rc = dev_alloc(c, (void**)&c->d_bsk, elems * 8);
    if (rc) return rc;
  }
  uint64_t* tmp = nullptr;
  int rc = dev_alloc(c, (void**)&tmp, elems * 8);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(tmp, bsk, elems * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = lk::launch_bsk_prepare(tmp, c->d_bsk, (int)p.lwe_dim, rows, K1, layout == LUTHE_BSK_NTT, c->d_tab, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  c->dev_bytes -= elems * 8;
  if (e != hipSuccess) return fail(c, LUTHE_ERR_HIP, "BSK upload: %s", hipGetErrorString(e));
  c->bsk_ok = true;
  return LUTHE_OK;
}

int luthe_upload_ksk(luthe_ctx* c, const uint32_t* ksk) {
  if (!c || !ksk) return LUTHE_ERR_ARG;
  if (c->server) return fail(c, LUTHE_ERR_BUSY, "keys and launch settings are fixed while a server is attached");
  const luthe_params& p = c->prm;
  size_t elems = (size_t)p.glwe_k * p.poly_size * p.ks_level * (p.lwe_dim + 1);
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->d_ksk) {
    int rc = dev_alloc(c, (void**)&c->d_ksk, elems * 4);
    if (rc) return rc;
  }
  HIPCHK(c, hipMemcpy(c->d_ksk, ksk, elems * 4, hipMemcpyHostToDevice));
  // the byte planes of the matrix-core key switch, when the parameters allow it
  const lk::KsArgs ka = ks_args(c, (int)p.lwe_dim + 1);
  if (lk::keyswitch_mfma_ok(ka, (int)p.ks_level)) {
    if (!c->d_kskp) {
      int rc = dev_alloc(c, (void**)&c->d_kskp, lk::ksk_planes_bytes(ka.kN, ka.n_lwe, (int)p.ks_level));
      if (rc) return rc;
    }
    HIPCHK(c, lk::launch_ksk_planes(c->d_ksk, c->d_kskp, ka, (int)p.ks_level, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  c->ksk_ok = true;
  return LUTHE_OK;
}

static int rows_io(luthe_ctx* c, uint64_t* dbase, uint32_t first, uint32_t count, uint32_t limit, const uint64_t* hin,
                   uint64_t* hout, size_t len, const char* what) {
  if (!c || (!hin && !hout)) return LUTHE_ERR_ARG;
  if ((uint64_t)first + count > limit) return fail(c, LUTHE_ERR_RANGE, "%s %u+%u out of range %u", what, first, count, limit);
  HIPCHK(c, hipSetDevice(c->device));
  int rc = dop_progress(c, true);  // submitted DOps come first, in submission order
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if ((rc = check_device_status(c))) return rc;
  if (hin)
    HIPCHK(c, hipMemcpy(dbase + (size_t)first * len, hin, (size_t)count * len * 8, hipMemcpyHostToDevice));
  else
    HIPCHK(c, hipMemcpy(hout, dbase + (size_t)first * len, (size_t)count * len * 8, hipMemcpyDeviceToHost));
  return LUTHE_OK;
}

int luthe_write_lut(luthe_ctx* c, uint32_t gid, const uint64_t* glwe, uint32_t count) {
  if (!c) return LUTHE_ERR_ARG;
  return rows_io(c, c->d_luts, gid, count, c->prm.lut_slots, glwe, nullptr, c->lut_len, "lut");
}
int luthe_write_ct(luthe_ctx* c, uint32_t cid, const uint64_t* blwe, uint32_t count) {
  if (!c) return LUTHE_ERR_ARG;
  return rows_io(c, c->d_mem + (size_t)c->prm.regfile_size * c->blen, cid, count, c->prm.ct_slots, blwe, nullptr,
                 c->blen, "ct slot");
}
int luthe_read_ct(luthe_ctx* c, uint32_t cid, uint64_t* blwe, uint32_t count) {
  if (!c) return LUTHE_ERR_ARG;
  return rows_io(c, c->d_mem + (size_t)c->prm.regfile_size * c->blen, cid, count, c->prm.ct_slots, nullptr, blwe,
                 c->blen, "ct slot");
}
int luthe_write_reg(luthe_ctx* c, uint32_t rid, const uint64_t* blwe, uint32_t count) {
  if (!c) return LUTHE_ERR_ARG;
  return rows_io(c, c->d_mem, rid, count, c->prm.regfile_size, blwe, nullptr, c->blen, "register");
}
int luthe_read_reg(luthe_ctx* c, uint32_t rid, uint64_t* blwe, uint32_t count) {
  if (!c) return LUTHE_ERR_ARG;
  return rows_io(c, c->d_mem, rid, count, c->prm.regfile_size, nullptr, blwe, c->blen, "register");
}

// ---------------------------------------------------------------- batched hot path
int luthe_keyswitch_batch(luthe_ctx* c, const uint64_t* d_blwe, uint16_t* d_lwe2n, size_t n, void* stream) {
  if (!c || (!d_blwe && n) || (!d_lwe2n && n)) return LUTHE_ERR_ARG;
  if (c->server) return fail(c, LUTHE_ERR_BUSY, "batch scratch is in use by an attached server");
  if (!c->ksk_ok) return fail(c, LUTHE_ERR_STATE, "KSK not uploaded");
  if (n == 0) return LUTHE_OK;
  hipStream_t s = pick_stream(c, stream);
  int rc = ensure_jobs(c, n);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, lk::launch_make_jobs(c->d_jobs, d_blwe, nullptr, nullptr, (int)n, 0, c->prm.lut_slots, (int)c->blen, s));
  rc = ensure_ks_part(c, c->bs, n);
  if (rc) return rc;
  return launch_ks(c, c->d_jobs, (int)n, d_lwe2n, (int)c->prm.lwe_dim + 1, c->bs.ks_part, s);
}

int luthe_blind_rotate_batch(luthe_ctx* c, const uint16_t* d_lwe2n, uint64_t* d_out, const uint32_t* d_gid,
                             uint32_t log_lut_nb, size_t n, void* stream) {
  if (!c || (n && (!d_lwe2n || !d_out || !d_gid))) return LUTHE_ERR_ARG;
  if (c->server) return fail(c, LUTHE_ERR_BUSY, "batch scratch is in use by an attached server");
  if (!c->bsk_ok) return fail(c, LUTHE_ERR_STATE, "BSK not uploaded");
  if (log_lut_nb > 3) return fail(c, LUTHE_ERR_ARG, "log_lut_nb must be <= 3");
  if (n == 0) return LUTHE_OK;
  hipStream_t s = pick_stream(c, stream);
  int rc = ensure_jobs(c, n);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, lk::launch_make_jobs(c->d_jobs, nullptr, d_out, d_gid, (int)n, log_lut_nb, c->prm.lut_slots, (int)c->blen,
                                 s));
  if ((rc = ensure_xch(c, c->bs, n))) return rc;
  const luthe_params& p = c->prm;
  lk::BrArgs ba{(int)p.lwe_dim, (int)p.lwe_dim + 1, p.pbs_base_log, c->br_split_wait_us};
  HIPCHK(c, lk::launch_blind_rotate(c->d_bsk, c->d_luts, d_lwe2n, c->d_jobs, (int)n, c->d_tab, ba,
                                    (int)p.pbs_level, c->br_grid_override, s, latency_max(c), use_pair(c),
                                    split_max(c), n <= c->bs.xch_cap ? c->bs.xch : nullptr, c->d_status,
                                    c->split_local));
  return LUTHE_OK;
}

int luthe_pbs_batch(luthe_ctx* c, const uint64_t* d_in, uint64_t* d_out, const uint32_t* d_gid, uint32_t log_lut_nb,
                    size_t n, void* stream) {
  if (!c || (n && (!d_in || !d_out || !d_gid))) return LUTHE_ERR_ARG;
  if (c->server) return fail(c, LUTHE_ERR_BUSY, "batch scratch is in use by an attached server");
  if (log_lut_nb > 3) return fail(c, LUTHE_ERR_ARG, "log_lut_nb must be <= 3");
  if (n == 0) return LUTHE_OK;
  if (n > (size_t)1 << 28) return fail(c, LUTHE_ERR_ARG, "batch too large");
  hipStream_t s = pick_stream(c, stream);
  HIPCHK(c, hipSetDevice(c->device));
  int rc = ensure_jobs(c, n);
  if (rc) return rc;
  HIPCHK(c, lk::launch_make_jobs(c->d_jobs, d_in, d_out, d_gid, (int)n, log_lut_nb, c->prm.lut_slots, (int)c->blen, s));
  return run_pbs(c, c->bs, c->d_jobs, (int)n, s);
}

int luthe_pbs_batch_list(luthe_ctx* c, const luthe_pbs_batch_desc* d_descs, uint32_t count, size_t total,
                         void* stream) {
  static_assert(sizeof(luthe_pbs_batch_desc) == sizeof(lk::PbsBatchDesc), "descriptor layouts agree");
  if (!c || (count && !d_descs)) return LUTHE_ERR_ARG;
  if (c->server) return fail(c, LUTHE_ERR_BUSY, "batch scratch is in use by an attached server");
  if (count == 0 || total == 0) return LUTHE_OK;
  if (total > (size_t)1 << 28 || count > (1u << 20)) return fail(c, LUTHE_ERR_ARG, "batch list too large");
  if (!c->bsk_ok || !c->ksk_ok) return fail(c, LUTHE_ERR_STATE, "keys not uploaded");
  hipStream_t s = pick_stream(c, stream);
  HIPCHK(c, hipSetDevice(c->device));
  int rc = ensure_jobs(c, total);
  if (!rc) rc = ensure_lwe2n(c, c->bs, total);
  if (rc) return rc;
  HIPCHK(c, lk::launch_make_jobs_list(c->d_jobs, (const lk::PbsBatchDesc*)d_descs, (int)count, (int)total,
                                      c->prm.lut_slots, (int)c->blen, s));
  return run_pbs(c, c->bs, c->d_jobs, (int)total, s);
}

int luthe_alu_batch(luthe_ctx* c, uint32_t opcode, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_out,
                    uint32_t mul_factor, uint32_t msg_cst, size_t n, void* stream) {
  if (!c || (n && (!d_a || !d_out))) return LUTHE_ERR_ARG;
  bool two = opcode == dop::OP_ADD || opcode == dop::OP_SUB || opcode == dop::OP_MAC;
  bool one = opcode == dop::OP_ADDS || opcode == dop::OP_SUBS || opcode == dop::OP_SSUB || opcode == dop::OP_MULS;
  if (!two && !one) return fail(c, LUTHE_ERR_DOP, "opcode 0x%02x is not an ALU DOp", opcode);
  if (two && n && !d_b) return LUTHE_ERR_ARG;
  if (n > (size_t)1 << 28) return fail(c, LUTHE_ERR_ARG, "batch too large");
  HIPCHK(c, lk::launch_alu_batch(opcode, d_a, d_b, d_out, mul_factor, msg_cst, (int)n, (int)c->blen,
                                 useful_bits(c), pick_stream(c, stream)));
  c->cnt.alu += n;
  return LUTHE_OK;
}

int luthe_reserve(luthe_ctx* c, size_t max_batch) {
  if (!c) return LUTHE_ERR_ARG;
  if (c->server) return fail(c, LUTHE_ERR_BUSY, "batch scratch is in use by an attached server");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = ensure_jobs(c, max_batch);
  if (!rc) rc = ensure_ks_part(c, c->bs, max_batch);
  if (!rc) rc = ensure_xch(c, c->bs, max_batch);
  if (rc) return rc;
  return ensure_lwe2n(c, c->bs, max_batch);
}

int luthe_capture_begin(luthe_ctx* c) {
  if (!c) return LUTHE_ERR_ARG;
  if (c->server) return fail(c, LUTHE_ERR_BUSY, "context stream is owned by an attached server");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = dop_progress(c, true);  // submitted DOps are launched, not captured
  if (rc) return rc;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  return LUTHE_OK;
}

int luthe_capture_end(luthe_ctx* c, luthe_graph** out) {
  if (!c || !out) return LUTHE_ERR_ARG;
  *out = nullptr;
  luthe_graph* g = new luthe_graph();
  hipError_t e = hipStreamEndCapture(c->stream, &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    luthe_graph_destroy(g);
    return fail(c, LUTHE_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
  }
  g->ctx = c;
  c->graphs.push_back(g);
  *out = g;
  return LUTHE_OK;
}

int luthe_graph_launch(luthe_ctx* c, luthe_graph* g) {
  if (!c || !g || !g->exec) return LUTHE_ERR_ARG;
  if (g->ctx != c) return fail(c, LUTHE_ERR_ARG, "graph was captured on another context");
  HIPCHK(c, hipGraphLaunch(g->exec, c->stream));
  return LUTHE_OK;
}

int luthe_graph_destroy(luthe_graph* g) {
  if (!g) return LUTHE_ERR_ARG;
  luthe_ctx* c = g->ctx;
  if (c) {
    // a replay may still be running on the context stream
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->graphs.erase(std::remove(c->graphs.begin(), c->graphs.end(), g), c->graphs.end());
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  if (c && c->graphs.empty()) {
    for (auto& r : c->retired) {
      (void)hipFree(r.first);
      c->dev_bytes -= r.second;
    }
    c->retired.clear();
  }
  return LUTHE_OK;
}

int luthe_ntt_polymul_batch(luthe_ctx* c, const uint64_t* d_a, const uint64_t* d_bn, uint64_t* d_out, size_t npoly,
                            void* stream) {
  if (!c || (npoly && (!d_a || !d_bn || !d_out))) return LUTHE_ERR_ARG;
  HIPCHK(c, lk::launch_ntt(2, d_a, d_bn, d_out, npoly, c->d_tab, pick_stream(c, stream), c->ntt_polys_per_wg));
  return LUTHE_OK;
}
int luthe_ntt_forward_batch(luthe_ctx* c, const uint64_t* d_in, uint64_t* d_out, size_t npoly, void* stream) {
  if (!c || (npoly && (!d_in || !d_out))) return LUTHE_ERR_ARG;
  HIPCHK(c, lk::launch_ntt(0, d_in, nullptr, d_out, npoly, c->d_tab, pick_stream(c, stream), c->ntt_polys_per_wg));
  return LUTHE_OK;
}
int luthe_ntt_inverse_batch(luthe_ctx* c, const uint64_t* d_in, uint64_t* d_out, size_t npoly, void* stream) {
  if (!c || (npoly && (!d_in || !d_out))) return LUTHE_ERR_ARG;
  HIPCHK(c, lk::launch_ntt(1, d_in, nullptr, d_out, npoly, c->d_tab, pick_stream(c, stream), c->ntt_polys_per_wg));
  return LUTHE_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- DOp stream
// Tickets are not launched one by one: submissions join an open group, and a group is
// scheduled as one DOp program (the tickets' streams concatenated in submission order), so
// the ready PBS of every ticket in it share one key-switch + blind-rotation launch per
// dependency level.  Concatenation keeps in-order semantics across tickets (register
// renaming and slot hazards span the whole group), and every PBS is a pure function of its
// input and LUT, so each ticket's results equal running the tickets one at a time.
// A group is sealed (launched) when a group slot is free (fewer than max_groups sealed
// groups still running) and the fill rule holds -- always when fill_pbs = 0, else once it
// holds fill_pbs PBS, a PBS carries the flush bit, or its oldest ticket waited timeout_us --
// and at once when the host needs a result (poll or wait on one of its tickets, luthe_flush,
// memory reads and writes, luthe_sync, capture, destroy).  The rules are checked whenever
// the host calls into the context.  This is the BPIP batch former (pep_sequencer.sv:584-608:
// full batch, timeout, flush) with opportunistic start (bpip_use_opportunism,
// hpu_config.toml.j2:9-11) applied to whole dependency-level schedules.

// Schedules and launches one group; returns its id (0 for an empty program).
static int execute_group(luthe_ctx* c, const std::vector<dop::Dop>& prog, const std::vector<uint64_t>& dop_ticket,
                         uint64_t* group_id) {
  const luthe_params& p = c->prm;
  const uint32_t R = p.regfile_size;
  // Register renaming: each register write lands in a fresh row after the ciphertext
  // memory, so only true data dependencies order the levels; a final level copies the
  // registers' last values back (results equal in-order execution).
  std::vector<dop::Renamed> rn;
  std::vector<std::pair<uint32_t, uint32_t>> copy_back;
  uint32_t rename_rows = 0;
  const int nlev_ops = dop::schedule_renamed(prog, R, R, R + p.ct_slots, rn, copy_back, rename_rows);
  const int nlev = nlev_ops + (copy_back.empty() ? 0 : 1);
  std::vector<int> level(prog.size());
  for (size_t i = 0; i < prog.size(); i++) level[i] = rn[i].level;
  int rc = ensure_rename_rows(c, rename_rows);
  if (rc) return rc;
  // descriptors per level, packed into one staging blob
  struct LevelPlan {
    size_t copy_off = 0, alu_off = 0, pbs_off = 0;
    int ncopy = 0, nalu = 0, npbs = 0;
  };
  std::vector<LevelPlan> plan(nlev);
  std::vector<std::vector<CopyOp>> copies(nlev);
  std::vector<std::vector<AluOp>> alus(nlev);
  std::vector<std::vector<PbsJob>> pbss(nlev);
  for (size_t i = 0; i < prog.size(); i++) {
    const dop::Dop& d = prog[i];
    const dop::Renamed& o = rn[i];
    int L = o.level;
    if (L < 0) continue;  // SYNC
    switch (d.kind) {
      case dop::KIND_MEM:
        copies[L].push_back(CopyOp{o.dst, o.src0});
        break;
      case dop::KIND_ARITH:
        alus[L].push_back(AluOp{d.opcode, o.dst, o.src0, o.src1, d.mul_factor, d.msg_cst});
        break;
      case dop::KIND_PBS:
        pbss[L].push_back(PbsJob{(uint64_t)(uintptr_t)(c->d_mem + (size_t)o.src0 * c->blen),
                                 (uint64_t)(uintptr_t)(c->d_mem + (size_t)o.dst * c->blen), d.gid, d.log_lut_nb});
        break;
      default:
        break;
    }
  }
  for (const auto& cb : copy_back) copies[nlev - 1].push_back(CopyOp{cb.first, cb.second});
  std::vector<uint8_t> blob;
  auto append = [&](const void* src, size_t bytes) {
    size_t off = (blob.size() + 15) & ~(size_t)15;
    blob.resize(off + bytes);
    if (bytes) memcpy(blob.data() + off, src, bytes);
    return off;
  };
  size_t maxp = 0;
  for (int L = 0; L < nlev; L++) {
    plan[L].ncopy = (int)copies[L].size();
    plan[L].nalu = (int)alus[L].size();
    plan[L].npbs = (int)pbss[L].size();
    plan[L].copy_off = append(copies[L].data(), copies[L].size() * sizeof(CopyOp));
    plan[L].alu_off = append(alus[L].data(), alus[L].size() * sizeof(AluOp));
    plan[L].pbs_off = append(pbss[L].data(), pbss[L].size() * sizeof(PbsJob));
    maxp = std::max(maxp, pbss[L].size());
  }
  if (maxp && (!c->bsk_ok || !c->ksk_ok)) return fail(c, LUTHE_ERR_STATE, "keys not uploaded");
  HIPCHK(c, hipSetDevice(c->device));
  rc = stage_upload(c, blob);
  if (rc) return rc;
  const int blen = (int)c->blen;
  // trace: ISSUE per level as it is launched; RETIRE events (recorded at submission) are
  // stamped when their ticket completes
  std::vector<std::vector<uint32_t>> by_level;
  if (c->trace_cap) {
    by_level.resize(nlev);
    std::map<uint64_t, size_t> nth;  // position of the DOp in its own submission
    for (size_t i = 0; i < prog.size(); i++) {
      const size_t k = nth[dop_ticket[i]]++;
      if (level[i] < 0) continue;
      by_level[level[i]].push_back((uint32_t)i);
      if (prog[i].kind == dop::KIND_PBS) {  // RETIRE of a PBS: the size of the launch it ran in
        auto it = c->trace_retire.find(dop_ticket[i]);
        if (it != c->trace_retire.end() && k < it->second.size())
          it->second[k].pe_reserved = (uint32_t)pbss[level[i]].size();
      }
    }
  }
  for (int L = 0; L < nlev; L++) {
    const LevelPlan& lp = plan[L];
    if (lp.ncopy)
      HIPCHK(c, lk::launch_copy(c->d_mem, (const CopyOp*)(c->d_stage + lp.copy_off), lp.ncopy, blen, c->stream));
    if (lp.nalu)
      HIPCHK(c, lk::launch_alu(c->d_mem, (const AluOp*)(c->d_stage + lp.alu_off), lp.nalu, blen, useful_bits(c),
                               c->stream));
    if (lp.npbs) {
      rc = run_pbs(c, c->ds, (const PbsJob*)(c->d_stage + lp.pbs_off), lp.npbs, c->stream);
      if (rc) return rc;
    }
    c->cnt.alu += (uint64_t)lp.nalu;
    c->cnt.mem += (uint64_t)lp.ncopy;
    if (c->trace_cap) {
      const uint32_t now = trace_now(c);
      for (uint32_t i : by_level[L]) {
        luthe_trace_event e{};
        e.timestamp = now;
        e.insn = prog[i].word;
        e.cmd = LUTHE_TRC_ISSUE;
        e.sync_id = (uint8_t)(dop_ticket[i] & 15);
        e.vld = 1;
        e.pdg = 1;
        trace_push(c, e);
      }
    }
  }
  c->cnt.levels += (uint64_t)nlev;
  c->cnt.groups += 1;
  DopGroup g;
  HIPCHK(c, hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
  HIPCHK(c, hipEventRecord(g.done, c->stream));
  const uint64_t gid = c->next_group++;
  c->groups[gid] = g;
  c->groups_running.push_back(gid);
  *group_id = gid;
  return LUTHE_OK;
}

namespace lx {

// Launches the open group now.  On failure its tickets complete with the error.
int dop_seal(luthe_ctx* c) {
  if (c->open_ids.empty()) return LUTHE_OK;
  std::vector<dop::Dop> prog;
  std::vector<uint64_t> ids, dop_ticket;
  prog.swap(c->open_prog);
  ids.swap(c->open_ids);
  dop_ticket.swap(c->open_dop_ticket);
  c->open_pbs = 0;
  c->open_flush = false;
  uint64_t gid = 0;
  const int rc = execute_group(c, prog, dop_ticket, &gid);
  for (uint64_t id : ids) {
    Ticket& t = c->tickets[id];
    if (rc) {
      t.status = rc;
    } else {
      t.group = gid;
      c->groups[gid].tickets++;
    }
  }
  return rc;
}

// Checks the sealing rules; force seals the open group whatever they say.
int dop_progress(luthe_ctx* c, bool force) {
  while (!c->groups_running.empty()) {  // sealed groups that have finished free their slot
    auto it = c->groups.find(c->groups_running.front());
    if (it != c->groups.end() && hipEventQuery(it->second.done) == hipErrorNotReady) break;
    c->groups_running.pop_front();
  }
  if (c->open_ids.empty()) return LUTHE_OK;
  if (!force) {
    const luthe_dop_batching& b = c->batching;
    if (c->groups_running.size() >= b.max_groups) return LUTHE_OK;
    if (b.fill_pbs) {
      const double waited = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() -
                                                                      c->open_since).count();
      const bool due = c->open_pbs >= b.fill_pbs || c->open_flush || (b.timeout_us && waited >= b.timeout_us);
      if (!due) return LUTHE_OK;
    }
  }
  return dop_seal(c);
}

}  // namespace lx

// Appends a validated stream to the open group as a new ticket.
static int enqueue_ticket(luthe_ctx* c, std::vector<dop::Dop>& prog, uint64_t* ticket) {
  int rc = dop_progress(c, false);  // frees finished group slots; may seal the open group
  if (rc) return rc;
  const uint64_t id = c->next_ticket++;
  if (c->open_ids.empty()) c->open_since = std::chrono::steady_clock::now();
  if (c->trace_cap) {  // REFILL: the DOps enter the pool
    const uint32_t now = trace_now(c);
    std::vector<luthe_trace_event>& ret = c->trace_retire[id];
    for (const dop::Dop& d : prog) {
      luthe_trace_event e{};
      e.timestamp = now;
      e.insn = d.word;
      e.cmd = LUTHE_TRC_REFILL;
      e.sync_id = (uint8_t)(id & 15);
      e.vld = 1;
      trace_push(c, e);
      e.cmd = LUTHE_TRC_RETIRE;
      ret.push_back(e);
    }
  }
  for (const dop::Dop& d : prog) {
    c->open_dop_ticket.push_back(id);
    if (d.kind == dop::KIND_PBS) {
      c->open_pbs++;
      if (d.flush) c->open_flush = true;
    }
  }
  c->open_prog.insert(c->open_prog.end(), prog.begin(), prog.end());
  c->open_ids.push_back(id);
  c->tickets[id] = Ticket{};
  c->cnt.dops += prog.size();
  c->cnt.tickets += 1;
  if (ticket) *ticket = id;
  if (c->open_prog.size() >= ((size_t)1 << 20)) return dop_seal(c);  // bounds one schedule's size
  return dop_progress(c, false);
}

// Releases a completed ticket (and its group's event once no ticket refers to it).
static void release_ticket(luthe_ctx* c, std::map<uint64_t, Ticket>::iterator it) {
  const uint64_t gid = it->second.group;
  const uint64_t id = it->first;
  c->tickets.erase(it);
  trace_retire(c, id);
  auto g = c->groups.find(gid);
  if (g != c->groups.end() && --g->second.tickets == 0) {
    (void)hipEventDestroy(g->second.done);
    c->groups_running.erase(std::remove(c->groups_running.begin(), c->groups_running.end(), gid),
                            c->groups_running.end());
    c->groups.erase(g);
  }
}

// The ticket's group, sealing the open group first if the ticket is in it.  Returns the
// ticket's error status if its group failed to launch.
static int ticket_group(luthe_ctx* c, uint64_t ticket, std::map<uint64_t, Ticket>::iterator& it) {
  it = c->tickets.find(ticket);
  if (it == c->tickets.end()) return fail(c, LUTHE_ERR_STATE, "unknown ticket %llu", (unsigned long long)ticket);
  if (it->second.group == 0 && it->second.status == 0) {
    (void)dop_seal(c);
    it = c->tickets.find(ticket);
  }
  const int st = it->second.status;
  if (st) {
    c->tickets.erase(it);
    c->trace_retire.erase(ticket);
    return st;
  }
  return LUTHE_OK;
}

extern "C" {

int luthe_submit_dops(luthe_ctx* c, const uint32_t* words, size_t n, uint64_t* ticket) {
  if (!c || (!words && n)) return LUTHE_ERR_ARG;
  std::vector<dop::Dop> prog;
  std::string why;
  int rc = dop::decode_stream(words, n, c->prm.regfile_size, c->prm.ct_slots, c->prm.lut_slots, prog, why);
  if (rc) return fail(c, rc, "%s", why.c_str());
  return enqueue_ticket(c, prog, ticket);
}

int luthe_poll(luthe_ctx* c, uint64_t ticket) {
  if (!c) return LUTHE_ERR_ARG;
  std::map<uint64_t, Ticket>::iterator it;
  int rc = ticket_group(c, ticket, it);
  if (rc) return rc;
  (void)dop_progress(c, false);
  hipError_t e = hipEventQuery(c->groups[it->second.group].done);
  if (e == hipErrorNotReady) return 0;
  if (e != hipSuccess) return fail(c, LUTHE_ERR_HIP, "ticket: %s", hipGetErrorString(e));
  release_ticket(c, it);
  if ((rc = check_device_status(c))) return rc;
  return 1;
}

int luthe_wait(luthe_ctx* c, uint64_t ticket, int64_t timeout_us) {
  if (!c) return LUTHE_ERR_ARG;
  std::map<uint64_t, Ticket>::iterator it;
  int rc = ticket_group(c, ticket, it);
  if (rc) return rc;
  hipEvent_t done = c->groups[it->second.group].done;
  if (timeout_us < 0) {
    HIPCHK(c, hipEventSynchronize(done));
  } else {
    auto t0 = std::chrono::steady_clock::now();
    for (;;) {
      hipError_t e = hipEventQuery(done);
      if (e == hipSuccess) break;
      if (e != hipErrorNotReady) return fail(c, LUTHE_ERR_HIP, "ticket: %s", hipGetErrorString(e));
      auto us = std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
      if (us >= timeout_us) return fail(c, LUTHE_ERR_TIMEOUT, "ticket %llu not done after %lld us",
                                       (unsigned long long)ticket, (long long)timeout_us);
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
  release_ticket(c, it);
  (void)dop_progress(c, false);  // a group slot is free: the open group may start
  return check_device_status(c);
}

int luthe_flush(luthe_ctx* c) {
  if (!c) return LUTHE_ERR_ARG;
  return dop_progress(c, true);
}

int luthe_set_dop_batching(luthe_ctx* c, const luthe_dop_batching* b) {
  if (!c || !b) return LUTHE_ERR_ARG;
  if (b->max_groups < 1 || b->max_groups > 64) return fail(c, LUTHE_ERR_ARG, "max_groups must be in [1, 64]");
  int rc = dop_progress(c, true);  // the open group was formed under the old rules
  if (rc) return rc;
  c->batching = *b;
  return LUTHE_OK;
}

int luthe_get_dop_batching(const luthe_ctx* c, luthe_dop_batching* out) {
  if (!c || !out) return LUTHE_ERR_ARG;
  *out = c->batching;
  return LUTHE_OK;
}

int luthe_trace_enable(luthe_ctx* c, size_t capacity) {
  if (!c) return LUTHE_ERR_ARG;
  c->trace_cap = capacity;
  c->trace.clear();
  c->trace_retire.clear();
  c->trace_t0 = std::chrono::steady_clock::now();
  return LUTHE_OK;
}

int luthe_trace_read(luthe_ctx* c, luthe_trace_event* out, size_t cap, size_t* n) {
  if (!c || !n || (cap && !out)) return LUTHE_ERR_ARG;
  size_t k = std::min(cap, c->trace.size());
  for (size_t i = 0; i < k; i++) out[i] = c->trace[i];
  c->trace.erase(c->trace.begin(), c->trace.begin() + (ptrdiff_t)k);
  *n = k;
  return LUTHE_OK;
}

int luthe_dop_schedule(const uint32_t* words, size_t n, uint32_t regs, uint32_t slots, uint32_t luts,
                       int32_t* levels) {
  if ((!words && n) || (!levels && n)) return LUTHE_ERR_ARG;
  if (regs < 1 || regs > 128 || slots > 65536 || luts < 1 || luts > 4096) return LUTHE_ERR_ARG;
  std::vector<dop::Dop> prog;
  std::string why;
  int rc = dop::decode_stream(words, n, regs, slots, luts, prog, why);
  if (rc) return rc;
  std::vector<int> lv;
  int nlev = dop::schedule_levels(prog, regs, lv);
  for (size_t i = 0; i < n; i++) levels[i] = lv[i];
  return nlev;
}

int luthe_dop_plan(const uint32_t* words, size_t n, uint32_t regs, uint32_t slots, uint32_t luts,
                   luthe_dop_plan_entry* entries, uint32_t* copy_back, size_t copy_back_cap, size_t* n_copy_back,
                   uint32_t* rename_rows) {
  if ((!words && n) || (!entries && n) || !n_copy_back || !rename_rows) return LUTHE_ERR_ARG;
  if (regs < 1 || regs > 128 || slots > 65536 || luts < 1 || luts > 4096) return LUTHE_ERR_ARG;
  std::vector<dop::Dop> prog;
  std::string why;
  int rc = dop::decode_stream(words, n, regs, slots, luts, prog, why);
  if (rc) return rc;
  std::vector<dop::Renamed> rn;
  std::vector<std::pair<uint32_t, uint32_t>> cb;
  uint32_t rows = 0;
  int nlev = dop::schedule_renamed(prog, regs, regs, regs + slots, rn, cb, rows);
  for (size_t i = 0; i < n; i++) entries[i] = luthe_dop_plan_entry{rn[i].level, rn[i].dst, rn[i].src0, rn[i].src1};
  *n_copy_back = cb.size();
  *rename_rows = rows;
  if (copy_back) {
    if (copy_back_cap < cb.size()) return LUTHE_ERR_ARG;
    for (size_t i = 0; i < cb.size(); i++) copy_back[2 * i] = cb[i].first, copy_back[2 * i + 1] = cb[i].second;
  }
  return nlev;
}

int luthe_iop_translate(const uint32_t* iop_words, size_t n_iop, const uint32_t* tpl, size_t n_tpl,
                        uint32_t ct_slots, uint32_t* out) {
  if (!iop_words || n_iop == 0 || (!tpl && n_tpl) || (!out && n_tpl) || ct_slots == 0 || ct_slots > 65536)
    return LUTHE_ERR_ARG;
  dop::IopBundle iop;
  std::string why;
  int rc = dop::parse_iop(iop_words, n_iop, iop, why);
  if (rc) return rc;
  std::vector<uint32_t> patched;
  rc = dop::patch_program(std::vector<uint32_t>(tpl, tpl + n_tpl), iop, ct_slots, patched, why);
  if (rc) return rc;
  for (size_t i = 0; i < n_tpl; i++) out[i] = patched[i];
  return LUTHE_OK;
}

// ---------------------------------------------------------------- IOp path
int luthe_register_iop(luthe_ctx* c, uint8_t opcode, uint8_t max_align, const uint32_t* dops, size_t n) {
  if (!c || (!dops && n)) return LUTHE_ERR_ARG;
  std::string why;
  int rc = dop::check_template(dops, n, c->prm.regfile_size, c->prm.lut_slots, why);
  if (rc) return fail(c, rc, "IOp 0x%02x template: %s", opcode, why.c_str());
  c->iop_table[((uint32_t)max_align << 8) | opcode] = std::vector<uint32_t>(dops, dops + n);
  return LUTHE_OK;
}

int luthe_submit_iop(luthe_ctx* c, const uint32_t* words, size_t n, uint64_t* ticket) {
  if (!c || !words || n == 0) return LUTHE_ERR_ARG;
  dop::IopBundle iop;
  std::string why;
  int rc = dop::parse_iop(words, n, iop, why);
  if (rc) return fail(c, rc, "%s", why.c_str());
  auto it = c->iop_table.find(((uint32_t)iop.max_align << 8) | iop.opcode);
  if (it == c->iop_table.end())
    return fail(c, LUTHE_ERR_DOP, "no DOp program registered for IOp 0x%02x align %u", iop.opcode, iop.max_align);
  std::vector<uint32_t> patched;
  rc = dop::patch_program(it->second, iop, c->prm.ct_slots, patched, why);
  if (rc) return fail(c, rc, "%s", why.c_str());
  std::vector<dop::Dop> prog;
  rc = dop::decode_stream(patched.data(), patched.size(), c->prm.regfile_size, c->prm.ct_slots, c->prm.lut_slots, prog,
                          why);
  if (rc) return fail(c, rc, "%s", why.c_str());
  c->cnt.iops += 1;
  return enqueue_ticket(c, prog, ticket);
}

}  // extern "C"
