// Serving engine: radix IOps on 2-bit blocks with continuous batching of PBS cohorts.
//
// Replaces, for an MI355X, the path from the tfhe-rs HPU backend's IOp work queue to
// its ack (fw/ublaze/src/ucore_fw.c:54-232), the instruction scheduler's out-of-order
// issue (hw/module/instruction_scheduler/rtl/instruction_scheduler.sv) and the BPIP
// batch former with its fill threshold and timeout (pep_sequencer.sv:584-620,
// hpu_regif_core_cfg_1in3.toml:240-260).
//
// Model.  A request is one IOp (op, nblocks, two operands).  Its program (a DAG of ALU
// and PBS nodes over single-assignment rows, iop_programs.h) runs in a private
// workspace of BLWE rows carved from an HBM arena sized at creation; a request is
// admitted only when its workspace and host staging fit (HBM admission).  A server
// thread issues "steps" on the context stream.  One step is:
//   inputs of newly admitted requests (host -> HBM)
//   ALU nodes that became ready, in dependency rounds (one k_alu launch per round)
//   one PBS cohort: up to cohort_max ready PBS nodes of all in-flight requests (one
//     key-switch + one blind-rotation launch), chosen by the cohort policy: oldest first, or
//     (default) the nodes with the fewest PBS levels left on their request's critical path
//     first, so that short requests are not held behind the fan-out of long ones, with
//     nodes that have waited age_limit_us taken first whatever their height
//   ALU nodes that the cohort made ready, in rounds
//   outputs of requests whose nodes are all issued (gather + HBM -> host)
//   an event.
// Successors are marked ready when their producer is issued, not when it finishes:
// everything is ordered on one stream, so step k + 1 may read rows step k writes.  At
// most pipeline_depth steps are in flight; a request arriving meanwhile joins the next
// step (continuous batching).  Every PBS is a pure function of its input row and LUT
// and the blind rotation's CMUX order is fixed, so a request's results do not depend on
// which other requests share its cohorts.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

#include "../../include/luthe.h"
#include "ctx.h"
#include "iop_programs.h"
#include "kernels.h"

using lk::AluOp;
using lk::CopyOp;
using lk::PbsJob;
using Clock = std::chrono::steady_clock;

namespace {

double us_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::micro>(b - a).count();
}

// First-fit allocator over [0, cap) in whole units (rows).
class RangeAlloc {
 public:
  void reset(uint64_t cap) {
    free_.clear();
    cap_ = cap;
    used_ = 0;
    peak_ = 0;
    if (cap) free_[0] = cap;
  }
  bool alloc(uint64_t n, uint64_t& at) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      at = it->first;
      uint64_t rest = it->second - n;
      free_.erase(it);
      if (rest) free_[at + n] = rest;
      used_ += n;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(uint64_t at, uint64_t n) {
    auto it = free_.emplace(at, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
    used_ -= n;
  }
  uint64_t cap() const { return cap_; }
  uint64_t used() const { return used_; }
  uint64_t peak() const { return peak_; }
  void reset_peak() { peak_ = used_; }

 private:
  std::map<uint64_t, uint64_t> free_;
  uint64_t cap_ = 0, used_ = 0, peak_ = 0;
};

struct Request {
  uint64_t id = 0;
  uint32_t op = 0, nblocks = 0;
  const iop::Program* prog = nullptr;
  std::vector<uint64_t> input;  // 2 * nblocks BLWE rows (a then b)
  uint64_t* out_host = nullptr;
  uint32_t nout = 0;
  // set by the server thread
  uint64_t row0 = 0, rows = 0;  // arena workspace: program rows, then nout packed output rows
  uint64_t pin0 = 0, pin_rows = 0;
  std::vector<uint32_t> waiting;  // per node: predecessors not yet issued
  uint32_t left = 0;              // nodes not yet issued
  uint32_t cohorts = 0;
  uint64_t last_cohort = ~0ull;
  Clock::time_point t_submit, t_admit, t_done;
  // guarded by the server mutex
  bool done = false;
  int status = LUTHE_OK;
};

struct ReadyPbs {
  Request* r;
  uint32_t node;
  Clock::time_point since;
};

// Ready PBS nodes, indexed by arrival (seq) and by cohort priority (height, seq).
struct ReadyQueue {
  std::map<uint64_t, ReadyPbs> by_age;
  std::set<std::pair<uint32_t, uint64_t>> by_prio;
  std::map<uint64_t, uint32_t> prio_of;  // seq -> priority key
  uint64_t next_seq = 0;

  bool empty() const { return by_age.empty(); }
  size_t size() const { return by_age.size(); }
  Clock::time_point oldest() const { return by_age.begin()->second.since; }
  void push(const ReadyPbs& rp, uint32_t prio) {
    const uint64_t q = next_seq++;
    by_age.emplace(q, rp);
    by_prio.emplace(prio, q);
    prio_of.emplace(q, prio);
  }
  ReadyPbs take(uint64_t q) {
    auto it = by_age.find(q);
    ReadyPbs rp = it->second;
    by_age.erase(it);
    auto p = prio_of.find(q);
    by_prio.erase(std::make_pair(p->second, q));
    prio_of.erase(p);
    return rp;
  }
  // Up to n nodes: those waited at least age_us first (oldest first), then by priority.
  std::vector<ReadyPbs> pop(size_t n, Clock::time_point now, uint32_t age_us) {
    std::vector<ReadyPbs> out;
    out.reserve(n);
    while (age_us && out.size() < n && !by_age.empty() && us_between(oldest(), now) >= (double)age_us)
      out.push_back(take(by_age.begin()->first));
    while (out.size() < n && !by_prio.empty()) out.push_back(take(by_prio.begin()->second));
    return out;
  }
};

struct StepSlot {  // descriptor staging of one in-flight step
  uint8_t* host = nullptr;
  uint8_t* dev = nullptr;
  size_t cap = 0;
};

struct Step {
  hipEvent_t done = nullptr;
  int slot = 0;
  std::vector<Request*> finished;
};

}  // namespace

struct luthe_server {
  luthe_ctx* c = nullptr;
  luthe_server_config cfg{};
  std::string err;
  size_t blen = 0, row_bytes = 0;
  uint64_t* d_arena = nullptr;
  uint64_t* h_pinned = nullptr;
  RangeAlloc arena, pinned;
  std::map<uint64_t, iop::Program> programs;  // key = op << 8 | nblocks

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;                         // submitted, not admitted
  std::map<uint64_t, std::unique_ptr<Request>> reqs;    // all not yet collected
  uint64_t next_id = 1;
  bool stop = false, flush = false;
  int fatal = LUTHE_OK;
  luthe_server_stats st{};
  std::vector<double> lat_us;

  // server thread only
  std::thread th;
  ReadyQueue ready_pbs;
  std::deque<Step> inflight;
  std::vector<StepSlot> slots;
  std::vector<hipEvent_t> free_events;
  uint64_t nsteps = 0, ncohorts = 0;
  int admitted_live = 0;
};

namespace {

int sfail(luthe_server* s, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  std::lock_guard<std::mutex> g(s->mu);
  s->err = buf;
  return code;
}

#define SHIP(s, expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return sfail((s), LUTHE_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

const iop::Program* program_for(luthe_server* s, uint32_t op, uint32_t nb) {
  uint64_t key = ((uint64_t)op << 8) | nb;
  auto it = s->programs.find(key);
  if (it != s->programs.end()) return &it->second;
  iop::Program p;
  if (!iop::build(op, nb, p)) return nullptr;
  // PBS levels from each node to the end of the program (nodes are in topological order)
  p.height.assign(p.nodes.size(), 0);
  for (size_t i = p.nodes.size(); i-- > 0;) {
    uint32_t h = 0;
    for (uint32_t succ : p.nodes[i].succ) h = std::max(h, p.height[succ]);
    p.height[i] = h + p.nodes[i].pbs;
  }
  return &s->programs.emplace(key, std::move(p)).first->second;
}

// Ready lists of the server thread: ALU nodes are collected per request and issued in
// rounds; PBS nodes queue for cohorts.
struct Issue {
  luthe_server* s;
  std::vector<std::pair<Request*, uint32_t>> alu;  // current round
  Clock::time_point now;
  std::vector<Request*> finished;  // requests whose last node was issued

  void mark_ready(Request* r, uint32_t node) {
    const iop::Node& n = r->prog->nodes[node];
    if (n.pbs)
      s->ready_pbs.push(ReadyPbs{r, node, now}, s->cfg.policy == LUTHE_COHORT_OLDEST ? 0u : r->prog->height[node]);
    else
      alu.emplace_back(r, node);
  }
  // Node issued: its successors lose one predecessor.
  void issued(Request* r, uint32_t node) {
    if (--r->left == 0) finished.push_back(r);
    for (uint32_t succ : r->prog->nodes[node].succ)
      if (--r->waiting[succ] == 0) mark_ready(r, succ);
  }
  // Drains the ALU ready list into rounds of AluOp descriptors.
  void alu_rounds(std::vector<std::vector<AluOp>>& rounds) {
    while (!alu.empty()) {
      std::vector<std::pair<Request*, uint32_t>> cur;
      cur.swap(alu);
      std::vector<AluOp> ops;
      ops.reserve(cur.size());
      for (auto& rn : cur) {
        const iop::Node& n = rn.first->prog->nodes[rn.second];
        const uint32_t b = (uint32_t)rn.first->row0;
        ops.push_back(AluOp{n.opcode, b + n.dst, b + n.src0, b + n.src1, n.mul_factor, n.msg_cst});
      }
      for (auto& rn : cur) issued(rn.first, rn.second);
      rounds.push_back(std::move(ops));
    }
  }
};

int slot_reserve(luthe_server* s, StepSlot& sl, size_t bytes) {
  if (bytes <= sl.cap) return LUTHE_OK;
  size_t cap = std::max(bytes, sl.cap * 2);
  if (sl.dev) SHIP(s, hipFree(sl.dev));
  if (sl.host) SHIP(s, hipHostFree(sl.host));
  sl.dev = nullptr;
  sl.host = nullptr;
  sl.cap = 0;
  SHIP(s, hipMalloc((void**)&sl.dev, cap));
  SHIP(s, hipHostMalloc((void**)&sl.host, cap, hipHostMallocDefault));
  sl.cap = cap;
  return LUTHE_OK;
}

bool cohort_due(luthe_server* s, Clock::time_point now, bool force) {
  if (s->ready_pbs.empty()) return false;
  if (force || s->ready_pbs.size() >= s->cfg.cohort_min) return true;
  return us_between(s->ready_pbs.oldest(), now) >= (double)s->cfg.cohort_timeout_us;
}

// Builds and issues one step.  Returns LUTHE_OK, 1 if there was nothing to issue, or an
// error code.
int issue_step(luthe_server* s, std::vector<Request*>& admitted) {
  luthe_ctx* c = s->c;
  const Clock::time_point now = Clock::now();
  bool force;  // explicit flush, or draining at destroy: ignore the fill threshold
  {
    std::lock_guard<std::mutex> g(s->mu);
    force = s->flush || s->stop;
  }
  const bool cohort = cohort_due(s, now, force);
  if (admitted.empty() && !cohort) return 1;

  Issue is{s, {}, now};
  std::vector<std::vector<AluOp>> pre, post;
  for (Request* r : admitted)
    for (uint32_t i = 0; i < r->prog->nodes.size(); i++)
      if (r->waiting[i] == 0) is.mark_ready(r, i);
  is.alu_rounds(pre);
  // cohort (nodes made ready by the pre rounds may join it)
  std::vector<PbsJob> jobs;
  bool by_timeout = false;
  if (cohort_due(s, now, force)) {
    by_timeout = !force && s->ready_pbs.size() < s->cfg.cohort_min;
    const size_t n = std::min<size_t>(s->ready_pbs.size(), s->cfg.cohort_max);
    jobs.reserve(n);
    std::vector<ReadyPbs> take = s->ready_pbs.pop(n, now, s->cfg.policy == LUTHE_COHORT_OLDEST ? 0 : s->cfg.age_limit_us);
    for (const ReadyPbs& rp : take) {
      const iop::Node& nd = rp.r->prog->nodes[rp.node];
      const uint32_t b = (uint32_t)rp.r->row0;
      jobs.push_back(PbsJob{(uint64_t)(uintptr_t)(s->d_arena + (size_t)(b + nd.src0) * s->blen),
                            (uint64_t)(uintptr_t)(s->d_arena + (size_t)(b + nd.dst) * s->blen), s->cfg.lut_base + nd.lut,
                            nd.log_lut_nb});
      if (rp.r->last_cohort != s->ncohorts) {
        rp.r->last_cohort = s->ncohorts;
        rp.r->cohorts++;
      }
    }
    for (const ReadyPbs& rp : take) is.issued(rp.r, rp.node);
    is.alu_rounds(post);
  }
  // requests whose nodes are all issued: gather their outputs
  std::vector<Request*> finished;
  finished.swap(is.finished);
  std::vector<CopyOp> gathers;
  for (Request* r : finished) {
    const uint32_t b = (uint32_t)r->row0, packed = b + r->prog->rows;
    for (uint32_t i = 0; i < r->nout; i++) gathers.push_back(CopyOp{packed + i, b + r->prog->out_rows[i]});
  }
  // descriptor blob
  std::vector<std::pair<size_t, size_t>> pre_off, post_off;  // (offset, count)
  size_t off = 0;
  auto place = [&](size_t bytes) {
    size_t o = (off + 255) & ~(size_t)255;
    off = o + bytes;
    return o;
  };
  for (auto& r : pre) pre_off.emplace_back(place(r.size() * sizeof(AluOp)), r.size());
  const size_t jobs_off = place(jobs.size() * sizeof(PbsJob));
  for (auto& r : post) post_off.emplace_back(place(r.size() * sizeof(AluOp)), r.size());
  const size_t gat_off = place(gathers.size() * sizeof(CopyOp));
  const int slot = (int)(s->nsteps % s->slots.size());
  StepSlot& sl = s->slots[(size_t)slot];
  int rc = slot_reserve(s, sl, off);
  if (rc) return rc;
  for (size_t i = 0; i < pre.size(); i++) memcpy(sl.host + pre_off[i].first, pre[i].data(), pre[i].size() * sizeof(AluOp));
  if (!jobs.empty()) memcpy(sl.host + jobs_off, jobs.data(), jobs.size() * sizeof(PbsJob));
  for (size_t i = 0; i < post.size(); i++)
    memcpy(sl.host + post_off[i].first, post[i].data(), post[i].size() * sizeof(AluOp));
  if (!gathers.empty()) memcpy(sl.host + gat_off, gathers.data(), gathers.size() * sizeof(CopyOp));

  hipStream_t st = c->stream;
  const int blen = (int)s->blen;
  for (Request* r : admitted)
    SHIP(s, hipMemcpyAsync(s->d_arena + r->row0 * s->blen, s->h_pinned + r->pin0 * s->blen,
                           (size_t)2 * r->nblocks * s->row_bytes, hipMemcpyHostToDevice, st));
  if (off) SHIP(s, hipMemcpyAsync(sl.dev, sl.host, off, hipMemcpyHostToDevice, st));
  uint64_t nalu = 0;
  for (size_t i = 0; i < pre.size(); i++) {
    SHIP(s, lk::launch_alu(s->d_arena, (const AluOp*)(sl.dev + pre_off[i].first), (int)pre_off[i].second, blen,
                             lx::useful_bits(c), st));
    nalu += pre_off[i].second;
  }
  if (!jobs.empty()) {
    rc = lx::run_pbs(c, c->bs, (const PbsJob*)(sl.dev + jobs_off), (int)jobs.size(), st);
    if (rc) return sfail(s, rc, "cohort: %s", lx::last_error(c).c_str());
  }
  for (size_t i = 0; i < post.size(); i++) {
    SHIP(s, lk::launch_alu(s->d_arena, (const AluOp*)(sl.dev + post_off[i].first), (int)post_off[i].second, blen,
                             lx::useful_bits(c), st));
    nalu += post_off[i].second;
  }
  if (!gathers.empty())
    SHIP(s, lk::launch_copy(s->d_arena, (const CopyOp*)(sl.dev + gat_off), (int)gathers.size(), blen, st));
  for (Request* r : finished)
    SHIP(s, hipMemcpyAsync(s->h_pinned + (r->pin0 + 2ull * r->nblocks) * s->blen,
                           s->d_arena + (r->row0 + r->prog->rows) * s->blen, (size_t)r->nout * s->row_bytes,
                           hipMemcpyDeviceToHost, st));
  Step step;
  if (!s->free_events.empty()) {
    step.done = s->free_events.back();
    s->free_events.pop_back();
  } else {
    SHIP(s, hipEventCreateWithFlags(&step.done, hipEventDisableTiming));
  }
  SHIP(s, hipEventRecord(step.done, st));
  step.slot = slot;
  step.finished = std::move(finished);
  s->inflight.push_back(std::move(step));
  s->nsteps++;
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (!jobs.empty()) {
      if (force) s->flush = false;
      s->ncohorts++;
      s->st.cohorts++;
      s->st.pbs += jobs.size();
      s->st.cohorts_by_timeout += by_timeout ? 1 : 0;
      size_t bucket = std::min<size_t>(7, jobs.size() * 8 / std::max<uint32_t>(1, s->cfg.cohort_max));
      if (jobs.size() == s->cfg.cohort_max) bucket = 7;
      s->st.fill_hist[bucket]++;
    }
    s->st.alu += nalu;
    s->st.alu_launches += pre.size() + post.size();
    s->st.steps++;
  }
  return LUTHE_OK;
}

// Completion of the oldest in-flight step, if its event has passed.
int retire(luthe_server* s, bool block) {
  if (s->inflight.empty()) return 1;
  Step& step = s->inflight.front();
  if (block) {
    SHIP(s, hipEventSynchronize(step.done));
  } else {
    hipError_t e = hipEventQuery(step.done);
    if (e == hipErrorNotReady) return 1;
    if (e != hipSuccess) return sfail(s, LUTHE_ERR_HIP, "step event: %s", hipGetErrorString(e));
  }
  if (int rc = lx::check_device_status(s->c)) return sfail(s, rc, "step: %s", lx::last_error(s->c).c_str());
  const Clock::time_point now = Clock::now();
  for (Request* r : step.finished)
    memcpy(r->out_host, s->h_pinned + (r->pin0 + 2ull * r->nblocks) * s->blen, (size_t)r->nout * s->row_bytes);
  {
    std::lock_guard<std::mutex> g(s->mu);
    for (Request* r : step.finished) {
      s->arena.release(r->row0, r->rows);
      s->pinned.release(r->pin0, r->pin_rows);
      r->pin_rows = 0;
      r->t_done = now;
      r->done = true;
      s->admitted_live--;
      s->st.completed++;
      s->lat_us.push_back(us_between(r->t_submit, r->t_done));
    }
  }
  s->free_events.push_back(step.done);
  s->inflight.pop_front();
  s->cv_done.notify_all();
  return LUTHE_OK;
}

// Moves pending requests into the arena while they fit (FIFO).
void admit(luthe_server* s, std::vector<Request*>& out) {
  std::lock_guard<std::mutex> g(s->mu);
  while (!s->pending.empty()) {
    Request* r = s->pending.front();
    if (s->cfg.max_inflight && (uint32_t)s->admitted_live >= s->cfg.max_inflight) break;
    uint64_t at = 0, pat = 0;
    const uint64_t rows = (uint64_t)r->prog->rows + r->nout;
    const uint64_t prows = 2ull * r->nblocks + r->nout;
    if (!s->arena.alloc(rows, at)) break;
    if (!s->pinned.alloc(prows, pat)) {
      s->arena.release(at, rows);
      break;
    }
    r->row0 = at;
    r->rows = rows;
    r->pin0 = pat;
    r->pin_rows = prows;
    r->t_admit = Clock::now();
    s->pending.pop_front();
    s->admitted_live++;
    s->st.admitted++;
    out.push_back(r);
  }
  s->st.arena_rows_peak = s->arena.peak();
}

void prepare(luthe_server* s, Request* r) {
  memcpy(s->h_pinned + r->pin0 * s->blen, r->input.data(), r->input.size() * sizeof(uint64_t));
  std::vector<uint64_t>().swap(r->input);
  r->waiting.resize(r->prog->nodes.size());
  for (size_t i = 0; i < r->prog->nodes.size(); i++) r->waiting[i] = r->prog->nodes[i].npred;
  r->left = (uint32_t)r->prog->nodes.size();
}

void fail_all(luthe_server* s, int code) {
  std::lock_guard<std::mutex> g(s->mu);
  s->fatal = code;
  for (auto& kv : s->reqs) {
    Request* r = kv.second.get();
    if (!r->done) {
      r->done = true;
      r->status = code;
      r->t_done = Clock::now();
    }
  }
  s->pending.clear();
  s->cv_done.notify_all();
}

void server_main(luthe_server* s) {
  if (hipSetDevice(s->c->device) != hipSuccess) {
    fail_all(s, LUTHE_ERR_HIP);
    return;
  }
  for (;;) {
    int rc;
    while ((rc = retire(s, false)) == LUTHE_OK) {
    }
    if (rc < 0) return fail_all(s, rc);
    bool issued = false;
    if (s->inflight.size() < s->cfg.pipeline_depth) {
      std::vector<Request*> admitted;
      admit(s, admitted);
      for (Request* r : admitted) prepare(s, r);
      rc = issue_step(s, admitted);
      if (rc < 0) return fail_all(s, rc);
      issued = rc == LUTHE_OK;
    }
    if (issued) continue;
    std::unique_lock<std::mutex> lk(s->mu);
    const bool idle = s->inflight.empty() && s->ready_pbs.empty() && s->pending.empty();
    if (s->stop && idle) return;
    if (!s->inflight.empty() || !s->ready_pbs.empty()) {
      // a step is running, or ready PBS wait for their cohort threshold / timeout
      s->cv_work.wait_for(lk, std::chrono::microseconds(100));
    } else {
      s->flush = false;  // nothing to flush
      s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
    }
  }
}

}  // namespace

extern "C" {

void luthe_server_config_default(luthe_server_config* cfg) {
  if (!cfg) return;
  memset(cfg, 0, sizeof *cfg);
  cfg->arena_bytes = 4ull << 30;
  cfg->pinned_bytes = 256ull << 20;
  cfg->cohort_max = 0;  // blind-rotation grid width
  cfg->cohort_min = 1;
  cfg->cohort_timeout_us = 0;
  cfg->pipeline_depth = 2;
  cfg->lut_base = 0xFFFFFFFFu;  // top LUTHE_IOP_LUTS slots
  cfg->max_inflight = 0;
  cfg->queue_max = 65536;
  cfg->policy = LUTHE_COHORT_SHORTEST;
  cfg->age_limit_us = 500000;
}

int luthe_server_create(luthe_ctx* c, const luthe_server_config* cfg_in, luthe_server** out) {
  if (!c || !out) return LUTHE_ERR_ARG;
  *out = nullptr;
  if (c->server) return lx::fail(c, LUTHE_ERR_STATE, "a server is already attached to this context");
  {
    int rc0 = lx::dop_progress(c, true);  // submitted DOps run before the server owns the stream
    if (rc0) return rc0;
  }
  if (!c->bsk_ok || !c->ksk_ok) return lx::fail(c, LUTHE_ERR_STATE, "keys not uploaded");
  const luthe_params& p = c->prm;
  if (p.msg_bits != 2 || p.carry_bits != 2)
    return lx::fail(c, LUTHE_ERR_PARAMS, "radix programs need 2 message + 2 carry bits");
  luthe_server_config cfg;
  if (cfg_in)
    cfg = *cfg_in;
  else
    luthe_server_config_default(&cfg);
  if (cfg.lut_base == 0xFFFFFFFFu) cfg.lut_base = p.lut_slots >= LUTHE_IOP_LUTS ? p.lut_slots - LUTHE_IOP_LUTS : 0;
  if ((uint64_t)cfg.lut_base + LUTHE_IOP_LUTS > p.lut_slots)
    return lx::fail(c, LUTHE_ERR_RANGE, "LUT slots %u..%u beyond the LUT memory (%u)", cfg.lut_base,
                    cfg.lut_base + LUTHE_IOP_LUTS - 1, p.lut_slots);
  if (cfg.pipeline_depth == 0) cfg.pipeline_depth = 2;
  if (cfg.pipeline_depth > 4) return lx::fail(c, LUTHE_ERR_ARG, "pipeline_depth must be <= 4");
  if (cfg.cohort_min == 0) cfg.cohort_min = 1;
  if (cfg.queue_max == 0) cfg.queue_max = 65536;
  HIPCHK(c, hipSetDevice(c->device));
  (void)lx::latency_max(c);  // resolved here, read only by the server thread
  (void)lx::ks_mfma_min(c);  // likewise
  if (cfg.cohort_max == 0) cfg.cohort_max = c->cohort_max > 0 ? (uint32_t)c->cohort_max : (uint32_t)luthe_get_br_grid(c);
  if (cfg.cohort_max > (1u << 20)) return lx::fail(c, LUTHE_ERR_ARG, "cohort_max must be <= 2^20");
  if (cfg.policy > LUTHE_COHORT_SHORTEST) return lx::fail(c, LUTHE_ERR_ARG, "unknown cohort policy %u", cfg.policy);

  auto s = std::make_unique<luthe_server>();
  s->c = c;
  s->cfg = cfg;
  s->blen = c->blen;
  s->row_bytes = c->blen * sizeof(uint64_t);
  const uint64_t arena_rows = cfg.arena_bytes / s->row_bytes;
  const uint64_t pin_rows = cfg.pinned_bytes / s->row_bytes;
  if (arena_rows < 8 || arena_rows > 0xFFFFFFFFull)
    return lx::fail(c, LUTHE_ERR_ARG, "arena of %llu rows out of range", (unsigned long long)arena_rows);
  if (pin_rows < 4) return lx::fail(c, LUTHE_ERR_ARG, "pinned staging too small");
  // LUTs of the programs
  std::vector<uint64_t> glwe(c->lut_len);
  for (uint32_t l = 0; l < iop::LUT_COUNT; l++) {
    std::vector<uint8_t> t;
    uint32_t lognb;
    iop::lut_tables(l, t, lognb);
    int rc = lognb ? luthe_lut_encode_many(&p, t.data(), 16u >> lognb, lognb, glwe.data())
                   : luthe_lut_encode(&p, t.data(), 16, glwe.data());
    if (rc) return lx::fail(c, rc, "LUT encoding");
    rc = luthe_write_lut(c, cfg.lut_base + l, glwe.data(), 1);
    if (rc) return rc;
  }
  int rc = lx::ensure_jobs(c, 1);
  // the cohorts' scratch is sized once here: the server thread never grows context buffers
  if (!rc) rc = lx::ensure_lwe2n(c, c->bs, cfg.cohort_max);
  if (!rc) rc = lx::ensure_ks_part(c, c->bs, cfg.cohort_max);
  if (!rc) rc = lx::ensure_xch(c, c->bs, cfg.cohort_max);
  if (rc) return rc;
  rc = lx::dev_alloc(c, (void**)&s->d_arena, arena_rows * s->row_bytes);
  if (rc) return rc;
  if (hipHostMalloc((void**)&s->h_pinned, pin_rows * s->row_bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(s->d_arena);
    c->dev_bytes -= arena_rows * s->row_bytes;
    return lx::fail(c, LUTHE_ERR_NOMEM, "pinned staging of %llu bytes", (unsigned long long)(pin_rows * s->row_bytes));
  }
  s->arena.reset(arena_rows);
  s->pinned.reset(pin_rows);
  s->slots.resize(cfg.pipeline_depth + 1);
  s->st.arena_rows = arena_rows;
  s->st.cohort_max = cfg.cohort_max;
  s->st.lut_base = cfg.lut_base;
  c->server = s.get();
  luthe_server* raw = s.release();
  raw->th = std::thread(server_main, raw);
  *out = raw;
  return LUTHE_OK;
}

int luthe_server_destroy(luthe_server* s) {
  if (!s) return LUTHE_ERR_ARG;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->flush = true;
  }
  s->cv_work.notify_all();
  if (s->th.joinable()) s->th.join();
  luthe_ctx* c = s->c;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (Step& st : s->inflight) s->free_events.push_back(st.done);
  for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
  for (StepSlot& sl : s->slots) {
    if (sl.dev) (void)hipFree(sl.dev);
    if (sl.host) (void)hipHostFree(sl.host);
  }
  if (s->d_arena) {
    (void)hipFree(s->d_arena);
    c->dev_bytes -= s->arena.cap() * s->row_bytes;
  }
  if (s->h_pinned) (void)hipHostFree(s->h_pinned);
  c->server = nullptr;
  delete s;
  return LUTHE_OK;
}

const char* luthe_server_last_error(luthe_server* s) {
  if (!s) return "null server";
  std::lock_guard<std::mutex> g(s->mu);
  return s->err.c_str();
}

int luthe_server_submit(luthe_server* s, uint32_t op, uint32_t nblocks, const uint64_t* a, const uint64_t* b,
                        uint64_t* out, uint64_t* id) {
  if (!s || !a || !b || !out || !id) return LUTHE_ERR_ARG;
  if (op >= LUTHE_IOP_COUNT || nblocks < 1 || nblocks > 64)
    return sfail(s, LUTHE_ERR_ARG, "IOp %u on %u blocks not supported", op, nblocks);
  auto r = std::make_unique<Request>();
  r->t_submit = Clock::now();
  r->op = op;
  r->nblocks = nblocks;
  r->out_host = out;
  r->input.resize((size_t)2 * nblocks * s->blen);
  memcpy(r->input.data(), a, (size_t)nblocks * s->row_bytes);
  memcpy(r->input.data() + (size_t)nblocks * s->blen, b, (size_t)nblocks * s->row_bytes);
  std::lock_guard<std::mutex> g(s->mu);
  if (s->fatal) {
    s->err = "server stopped after a device error";
    return s->fatal;
  }
  if (s->stop) return LUTHE_ERR_STATE;
  r->prog = program_for(s, op, nblocks);
  if (!r->prog) return LUTHE_ERR_ARG;
  r->nout = (uint32_t)r->prog->out_rows.size();
  if ((uint64_t)r->prog->rows + r->nout > s->arena.cap() || 2ull * nblocks + r->nout > s->pinned.cap()) {
    s->st.rejected++;
    s->err = "request larger than the HBM arena or host staging";
    return LUTHE_ERR_NOMEM;
  }
  if (s->pending.size() >= s->cfg.queue_max) {
    s->st.rejected++;
    s->err = "admission queue full";
    return LUTHE_ERR_BUSY;
  }
  r->id = s->next_id++;
  *id = r->id;
  s->pending.push_back(r.get());
  s->reqs.emplace(r->id, std::move(r));
  s->st.submitted++;
  s->cv_work.notify_one();
  return LUTHE_OK;
}

static void fill_info(const Request* r, luthe_request_info* info) {
  if (!info) return;
  memset(info, 0, sizeof *info);
  info->op = r->op;
  info->nblocks = r->nblocks;
  info->npbs = r->prog ? r->prog->npbs : 0;
  info->depth = r->prog ? r->prog->depth : 0;
  info->cohorts = r->cohorts;
  info->status = r->status;
  info->queue_us = r->t_admit.time_since_epoch().count() ? us_between(r->t_submit, r->t_admit) : 0.0;
  info->latency_us = us_between(r->t_submit, r->t_done);
}

int luthe_server_poll(luthe_server* s, uint64_t id, luthe_request_info* info) {
  if (!s) return LUTHE_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return LUTHE_ERR_STATE;
  if (!it->second->done) return 0;
  fill_info(it->second.get(), info);
  int st = it->second->status;
  s->reqs.erase(it);
  return st ? st : 1;
}

int luthe_server_wait(luthe_server* s, uint64_t id, int64_t timeout_us, luthe_request_info* info) {
  if (!s) return LUTHE_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return LUTHE_ERR_STATE;
  Request* r = it->second.get();
  if (timeout_us < 0) {
    s->cv_done.wait(lk, [&] { return r->done; });
  } else if (!s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), [&] { return r->done; })) {
    return LUTHE_ERR_TIMEOUT;
  }
  fill_info(r, info);
  int st = r->status;
  s->reqs.erase(id);
  return st;
}

int luthe_server_flush(luthe_server* s) {
  if (!s) return LUTHE_ERR_ARG;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->flush = true;
  }
  s->cv_work.notify_one();
  return LUTHE_OK;
}

int luthe_server_get_stats(luthe_server* s, luthe_server_stats* out, int reset) {
  if (!s || !out) return LUTHE_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  s->st.pending = (uint32_t)s->pending.size();
  s->st.inflight = (uint32_t)s->admitted_live;
  s->st.arena_rows_used = s->arena.used();
  s->st.arena_rows_peak = s->arena.peak();
  std::vector<double> v = s->lat_us;
  std::sort(v.begin(), v.end());
  auto pct = [&](double q) {
    if (v.empty()) return 0.0;
    size_t i = (size_t)(q * (double)(v.size() - 1) + 0.5);
    return v[std::min(i, v.size() - 1)];
  };
  s->st.latency_p50_us = pct(0.50);
  s->st.latency_p99_us = pct(0.99);
  s->st.latency_max_us = v.empty() ? 0.0 : v.back();
  *out = s->st;
  if (reset) {
    const luthe_server_stats keep = s->st;
    memset(&s->st, 0, sizeof s->st);
    s->st.arena_rows = keep.arena_rows;
    s->st.cohort_max = keep.cohort_max;
    s->st.lut_base = keep.lut_base;
    s->lat_us.clear();
    s->arena.reset_peak();
  }
  return LUTHE_OK;
}

// ---------------------------------------------------------------- host-only program access
int luthe_iop_program(uint32_t op, uint32_t nblocks, luthe_iop_info* info, luthe_iop_node* nodes, size_t nodes_cap,
                      uint32_t* out_rows, size_t out_cap) {
  if (op >= LUTHE_IOP_COUNT || nblocks < 1 || nblocks > 64) return LUTHE_ERR_ARG;
  iop::Program p;
  if (!iop::build(op, nblocks, p)) return LUTHE_ERR_ARG;
  if (info) {
    memset(info, 0, sizeof *info);
    info->nblocks = nblocks;
    info->rows = p.rows;
    info->nodes = (uint32_t)p.nodes.size();
    info->npbs = p.npbs;
    info->nalu = (uint32_t)p.nodes.size() - p.npbs;
    info->depth = p.depth;
    info->nout = (uint32_t)p.out_rows.size();
    info->a_row0 = p.a_row0;
    info->b_row0 = p.b_row0;
  }
  if (nodes) {
    if (nodes_cap < p.nodes.size()) return LUTHE_ERR_ARG;
    for (size_t i = 0; i < p.nodes.size(); i++) {
      const iop::Node& n = p.nodes[i];
      nodes[i] = luthe_iop_node{n.pbs, n.opcode, n.lut, n.log_lut_nb, n.dst, n.src0, n.src1, n.mul_factor};
    }
  }
  if (out_rows) {
    if (out_cap < p.out_rows.size()) return LUTHE_ERR_ARG;
    for (size_t i = 0; i < p.out_rows.size(); i++) out_rows[i] = p.out_rows[i];
  }
  return LUTHE_OK;
}

int luthe_iop_program_dops(uint32_t op, uint32_t nblocks, uint32_t regs, uint32_t lut_base, uint32_t* words,
                           size_t cap, size_t* n, luthe_iop_dop_stats* stats) {
  if (op >= LUTHE_IOP_COUNT || nblocks < 1 || nblocks > 64 || !n) return LUTHE_ERR_ARG;
  if (lut_base + LUTHE_IOP_LUTS > 4096) return LUTHE_ERR_RANGE;
  iop::Program p;
  if (!iop::build(op, nblocks, p)) return LUTHE_ERR_ARG;
  std::vector<uint32_t> w;
  iop::EmitStats es;
  if (!iop::emit_dops(p, regs, lut_base, w, es)) return LUTHE_ERR_ARG;
  *n = w.size();
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->dops = es.dops, stats->ld = es.ld, stats->st = es.st, stats->alu = es.alu, stats->pbs = es.pbs;
    stats->heap_slots = es.heap_slots, stats->levels = es.levels;
    stats->iop_opcode = iop::iop_opcode(op);
    stats->align_bits = (op == LUTHE_IOP_CMP_GT) ? 2 * nblocks : 2 * nblocks;
  }
  if (words) {
    if (cap < w.size()) return LUTHE_ERR_ARG;
    memcpy(words, w.data(), w.size() * sizeof(uint32_t));
  }
  return LUTHE_OK;
}

int luthe_write_iop_luts(luthe_ctx* c, uint32_t lut_base) {
  if (!c) return LUTHE_ERR_ARG;
  const luthe_params& p = c->prm;
  if (p.msg_bits != 2 || p.carry_bits != 2)
    return lx::fail(c, LUTHE_ERR_PARAMS, "radix programs need 2 message + 2 carry bits");
  if ((uint64_t)lut_base + LUTHE_IOP_LUTS > p.lut_slots) return lx::fail(c, LUTHE_ERR_RANGE, "LUT slots out of range");
  std::vector<uint64_t> glwe(c->lut_len);
  for (uint32_t l = 0; l < iop::LUT_COUNT; l++) {
    std::vector<uint8_t> t;
    uint32_t lognb;
    iop::lut_tables(l, t, lognb);
    int rc = lognb ? luthe_lut_encode_many(&p, t.data(), 16u >> lognb, lognb, glwe.data())
                   : luthe_lut_encode(&p, t.data(), 16, glwe.data());
    if (rc) return lx::fail(c, rc, "LUT encoding");
    rc = luthe_write_lut(c, lut_base + l, glwe.data(), 1);
    if (rc) return rc;
  }
  return LUTHE_OK;
}

int luthe_register_iop_program(luthe_ctx* c, uint32_t op, uint32_t nblocks, uint32_t lut_base) {
  if (!c) return LUTHE_ERR_ARG;
  size_t n = 0;
  luthe_iop_dop_stats st;
  int rc = luthe_iop_program_dops(op, nblocks, c->prm.regfile_size, lut_base, nullptr, 0, &n, &st);
  if (rc) return lx::fail(c, rc, "IOp %u on %u blocks: no program for %u registers", op, nblocks, c->prm.regfile_size);
  std::vector<uint32_t> w(n);
  rc = luthe_iop_program_dops(op, nblocks, c->prm.regfile_size, lut_base, w.data(), w.size(), &n, &st);
  if (rc) return rc;
  if (st.heap_slots > c->prm.ct_slots / 2)
    return lx::fail(c, LUTHE_ERR_RANGE, "program needs %u heap slots", st.heap_slots);
  return luthe_register_iop(c, st.iop_opcode, (uint8_t)st.align_bits, w.data(), w.size());
}

int luthe_iop_lut(uint32_t lut, uint8_t* table16, uint32_t* log_lut_nb) {
  if (lut >= LUTHE_IOP_LUTS || !table16 || !log_lut_nb) return LUTHE_ERR_ARG;
  std::vector<uint8_t> t;
  iop::lut_tables(lut, t, *log_lut_nb);
  memcpy(table16, t.data(), 16);
  return LUTHE_OK;
}

int luthe_iop_eval_plain(uint32_t op, uint32_t nblocks, const uint32_t* a, const uint32_t* b, uint32_t* out) {
  if (op >= LUTHE_IOP_COUNT || nblocks < 1 || nblocks > 64 || !a || !b || !out) return LUTHE_ERR_ARG;
  iop::Program p;
  if (!iop::build(op, nblocks, p)) return LUTHE_ERR_ARG;
  iop::eval_plain(p, a, b, out);
  return LUTHE_OK;
}

}  // extern "C"
