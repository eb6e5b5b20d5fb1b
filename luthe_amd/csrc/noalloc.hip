// Batched PBS with no allocation: luthe_pbs_batch_noalloc.
//
// luthe_pbs_batch keeps two scratch buffers on the context (job descriptors and the key
// switch's mod-switched LWEs) and grows them with hipMalloc when a batch is larger than any
// before.  Callers that may not allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid it) use this variant instead: the same kernels, with both buffers
// carved from one file-scope device array of fixed size.  Each GPU holds its own copy of the
// array (its code object's), at its own address, for the life of the process; the host never
// uses the array's name as an address (in host code it is a stand-in, not device memory) but
// the pointer hipGetSymbolAddress returns for the current device, looked up once per device
// when a context is created there.
//
// Every call on one GPU shares that GPU's array: calls that overlap in time on one GPU (other
// streams, other contexts, other threads) race and must be serialised by the caller.  Nothing
// in the array needs to start at zero: each call's kernels write every descriptor and every
// LWE coefficient they later read.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ctx.h"
#include "kernels.h"

namespace {

constexpr size_t JOBS_BYTES = (size_t)LUTHE_NOALLOC_MAX_BATCH * sizeof(lk::PbsJob);  // 192 KiB
constexpr size_t LWE_BYTES = (size_t)LUTHE_NOALLOC_MAX_LWE_WORDS * sizeof(uint16_t);  // 16 MiB
constexpr size_t SCRATCH_BYTES = JOBS_BYTES + LWE_BYTES;
static_assert(JOBS_BYTES % 256 == 0, "the LWE region starts 256-byte aligned");

static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};  // device address of `scratch` on each device, once looked up

}  // namespace

namespace lx {

// Called by luthe_create with `device` current.  Looks the array's address up once per device.
int noalloc_init(luthe_ctx* c, int device) {
  if (device < 0 || device >= MAX_DEVICES) return fail(c, LUTHE_ERR_NODEV, "device %d beyond %d", device, MAX_DEVICES);
  std::lock_guard<std::mutex> g(g_mu);
  if (g_ptr[device]) return LUTHE_OK;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(c, LUTHE_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
  g_ptr[device] = ptr;
  return LUTHE_OK;
}

}  // namespace lx

extern "C" int luthe_pbs_batch_noalloc(luthe_ctx* c, const uint64_t* d_in, uint64_t* d_out, const uint32_t* d_gid,
                                       uint32_t log_lut_nb, size_t n, void* stream) {
  if (!c || (n && (!d_in || !d_out || !d_gid))) return LUTHE_ERR_ARG;
  if (log_lut_nb > 3) return lx::fail(c, LUTHE_ERR_ARG, "log_lut_nb must be <= 3");
  if (!c->bsk_ok || !c->ksk_ok) return lx::fail(c, LUTHE_ERR_STATE, "keys not uploaded");
  if (n == 0) return LUTHE_OK;
  const size_t stride = (size_t)c->lwe_stride;
  if (n > LUTHE_NOALLOC_MAX_BATCH || n * stride > LUTHE_NOALLOC_MAX_LWE_WORDS)
    return lx::fail(c, LUTHE_ERR_ARG,
                    "batch of %zu exceeds the fixed scratch (at most %d ciphertexts and %d mod-switched LWE "
                    "words; %zu words per ciphertext)",
                    n, LUTHE_NOALLOC_MAX_BATCH, LUTHE_NOALLOC_MAX_LWE_WORDS, stride);
  void* base;
  {
    std::lock_guard<std::mutex> g(g_mu);
    base = (c->device >= 0 && c->device < MAX_DEVICES) ? g_ptr[c->device] : nullptr;
  }
  if (!base) return lx::fail(c, LUTHE_ERR_STATE, "no-allocation scratch not set up on device %d", c->device);
  lk::PbsJob* jobs = (lk::PbsJob*)base;
  uint16_t* lwe2n = (uint16_t*)((uint8_t*)base + JOBS_BYTES);
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, lk::launch_make_jobs(jobs, d_in, d_out, d_gid, (int)n, log_lut_nb, c->prm.lut_slots, (int)c->blen, s));
  return lx::run_pbs_on(c, jobs, (int)n, s, lwe2n, nullptr);  // the unsplit key switch: no partial-sum scratch
}
