// Context internals shared by the C ABI translation units (luthe_api.cpp, serve.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/luthe.h"
#include "dop.h"
#include "kernels.h"

// A submission: in the open group (group 0) until sealed, then in a launched group.
struct Ticket {
  uint64_t group = 0;
  int status = 0;  // error of a group that failed to launch
};

// Tickets scheduled and launched together; `done` is recorded after its last level.
struct DopGroup {
  hipEvent_t done = nullptr;
  uint32_t tickets = 0;  // tickets not yet released
};

struct luthe_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  luthe_ctx* ctx = nullptr;  // cleared when the context is destroyed first
};

// Per-launch scratch of the PBS path: the key switch's mod-switched LWEs and the partial sums
// of its split form.  A context keeps two sets: one for batch calls and an attached server's
// cohorts, one for DOp/IOp ticket groups, so that groups launched by the submitting thread
// never share scratch with a server thread's cohorts on the same stream.
struct PbsScratch {
  uint16_t* lwe2n = nullptr;
  size_t lwe2n_cap = 0;    // ciphertexts
  uint32_t* ks_part = nullptr;
  size_t ks_part_cap = 0;  // words
  size_t ks_part_for = 0;  // batch sizes up to this are covered
  uint64_t* xch = nullptr;  // the split blind rotation's exchange area
  size_t xch_cap = 0;       // ciphertexts
};

struct luthe_ctx {
  luthe_params prm{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  size_t blen = 0;           // k*N + 1
  size_t lut_len = 0;        // (k+1)*N
  // device memories
  uint64_t* d_mem = nullptr;  // [regfile_size + ct_slots + rename_cap] BLWE rows
  size_t rename_cap = 0;      // rows for renamed registers of DOp streams
  uint64_t* d_luts = nullptr;
  uint64_t* d_bsk = nullptr;
  uint32_t* d_ksk = nullptr;
  int8_t* d_kskp = nullptr;   // the KSK as balanced byte planes for the matrix-core key switch
  uint64_t* d_tab = nullptr;  // NTT twiddle tables (ntt2048.h layout)
  bool bsk_ok = false, ksk_ok = false;
  // scratch (grow-only)
  PbsScratch bs;  // batch calls and the server's cohorts
  PbsScratch ds;  // DOp/IOp ticket groups
  int lwe_stride = 0;
  lk::PbsJob* d_jobs = nullptr;
  size_t jobs_cap = 0;
  uint8_t* d_stage = nullptr;  // per-submission descriptors
  size_t stage_cap = 0;
  uint8_t* h_stage = nullptr;  // pinned mirror
  size_t h_stage_cap = 0;
  hipEvent_t stage_copied = nullptr;
  // Captured graphs hold raw pointers to the scratch above: while any exists, scratch that
  // grows is retired here instead of freed, and freed when the last graph is destroyed.
  std::vector<luthe_graph*> graphs;
  std::vector<std::pair<void*, size_t>> retired;  // (pointer, bytes)
  uint64_t dev_bytes = 0;
  luthe_counters cnt{};
  int br_split_max = -1;        // split blind rotation up to this many ciphertexts (-1: CUs / 2)
  int br_split_wait_us = -1;    // its longest wait for a partner workgroup (-1: 2 s)
  int ks_mfma_min = -1;         // smallest batch for the matrix-core key switch (-1: default)
  uint32_t* h_status = nullptr;  // pinned: set by a split pair whose exchange timed out
  uint32_t* d_status = nullptr;  // its device address
  bool split_local_ok = false;   // split pairs share an XCC (probed at creation)
  bool split_local = false;      // their exchange goes through that XCC's L2 (needs _ok)
  mutable std::mutex cnt_mu;  // counters and the error message are also written by a server thread
  std::mutex err_mu;
  uint64_t next_ticket = 1;
  std::map<uint64_t, Ticket> tickets;
  // DOp/IOp submissions: the open group (streams concatenated in submission order) and the
  // launched groups; groups_running lists launched groups not yet seen complete, oldest first.
  std::vector<dop::Dop> open_prog;
  std::vector<uint64_t> open_dop_ticket;  // ticket of each open DOp
  std::vector<uint64_t> open_ids;         // tickets of the open group
  uint32_t open_pbs = 0;
  bool open_flush = false;                // a PBS of the open group carries the flush bit
  std::chrono::steady_clock::time_point open_since;
  std::map<uint64_t, DopGroup> groups;
  std::deque<uint64_t> groups_running;
  uint64_t next_group = 1;
  luthe_dop_batching batching{0, 0, 1};
  std::map<uint32_t, std::vector<uint32_t>> iop_table;  // key = max_align << 8 | opcode
  int br_grid_override = 0;
  int br_pair = -1;          // throughput kernel: -1 default, 0 256-thread, 1 512-thread pair
  int br_latency_max = -1;  // cohorts up to this size use the latency kernel; -1 = its resident capacity
  int ntt_polys_per_wg = 2;  // polynomials per NTT workgroup (1, 2, 4)
  int cohort_max = 0;        // serving cohort size chosen by the autotuner (0 = blind-rotation width)
  luthe_server* server = nullptr;  // attached serving engine (owns the stream while set)
  // DOp-path trace (isc_trace_t events), a ring of trace_cap events
  size_t trace_cap = 0;
  std::deque<luthe_trace_event> trace;
  std::chrono::steady_clock::time_point trace_t0;
  std::map<uint64_t, std::vector<luthe_trace_event>> trace_retire;  // ticket -> RETIRE events
};

namespace lx {

// Records the message on the context and returns `code`.
int fail(luthe_ctx* c, int code, const char* fmt, ...);
int dev_alloc(luthe_ctx* c, void** p, size_t bytes);
// Frees a scratch buffer that is being replaced, or retires it while captured graphs exist.
void retire_scratch(luthe_ctx* c, void* p, size_t bytes);
int ensure_lwe2n(luthe_ctx* c, PbsScratch& sc, size_t n);
int ensure_jobs(luthe_ctx* c, size_t n);
// Scratch for the split key switch of batches of up to n ciphertexts.
int ensure_ks_part(luthe_ctx* c, PbsScratch& sc, size_t n);
// Exchange area of the split blind rotation for cohorts of up to n ciphertexts (capped at the
// split threshold; nothing for larger n).
int ensure_xch(luthe_ctx* c, PbsScratch& sc, size_t n);
int split_max(luthe_ctx* c);
// After a synchronisation: LUTHE_ERR_HIP (once) if a split blind rotation's pair exchange
// timed out since the last check.
int check_device_status(luthe_ctx* c);
// A copy of the context's last error message (taken under its lock).
std::string last_error(luthe_ctx* c);
int ensure_rename_rows(luthe_ctx* c, size_t n);
// Key switch + blind rotation + sample extract of n jobs whose descriptors are on the
// device (their BLWE addresses inside).
int run_pbs(luthe_ctx* c, PbsScratch& sc, const lk::PbsJob* d_jobs, int n, hipStream_t s);
// The same with the mod-switched LWEs in caller-provided device memory; allocates nothing.
// ks_part: the split key switch's scratch (NULL: no split).
int run_pbs_on(luthe_ctx* c, const lk::PbsJob* d_jobs, int n, hipStream_t s, uint16_t* lwe2n, uint32_t* ks_part,
               uint64_t* xch = nullptr);
// Looks up the no-allocation entry point's device scratch on `device` (current).
int noalloc_init(luthe_ctx* c, int device);
// DOp groups: seal (launch) the open group now; check the sealing rules (force: seal).
int dop_seal(luthe_ctx* c);
int dop_progress(luthe_ctx* c, bool force);
// Width of the ALU scalar constant in the body: msg + carry + 1 padding bit.
uint32_t useful_bits(const luthe_ctx* c);
// Whether batches above the latency threshold run the 512-thread pair kernel.
bool use_pair(const luthe_ctx* c);
// Largest cohort that runs the 512-thread latency kernel.
int latency_max(luthe_ctx* c);
// Key-switch arguments of the context's parameters with outputs lwe_stride apart.
lk::KsArgs ks_args(const luthe_ctx* c, int lwe_stride);
// Smallest batch that takes the matrix-core key switch (when the KSK planes exist).
int ks_mfma_min(luthe_ctx* c);
// Key switch of n jobs into lwe2n: the matrix-core kernel, or the VALU kernel split over the
// inputs with ks_part (NULL: no split).  Allocates nothing.
int launch_ks(luthe_ctx* c, const lk::PbsJob* d_jobs, int n, uint16_t* lwe2n, int lwe_stride, uint32_t* ks_part,
              hipStream_t s);

}  // namespace lx

#define HIPCHK(c, expr)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return lx::fail((c), LUTHE_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
