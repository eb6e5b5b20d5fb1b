// HIP kernels of the MI355X PBS engine (gfx950, wave64).
//
//  k_ntt_*        batched negacyclic NTT / INTT / NTT-multiply-INTT, N = 2048
//  k_bsk_prepare  bootstrapping key -> NTT domain, blind-rotation layout
//  (k_keyswitch, the batched LWE key switch, is in keyswitch.hip, its matrix-core form in
//  ks_mfma.hip; the 512-thread latency and pair blind rotations in latency.hip; the split
//  blind rotation for the smallest cohorts in split.hip)
//  k_blind_rotate persistent blind rotation: one ciphertext per 256-thread workgroup,
//                 accumulator in registers, NTT work buffer in LDS, BSK streamed from
//                 HBM/L2 in a coalesced per-thread layout; sample extract fused
//  k_alu, k_copy  DOp ALU and LD/ST on BLWE rows
//
// The reduction tails of the four-lane reductions and shift twists go up to four lanes per
// asm block here (GL_GROUPED_TAILS): at four waves per SIMD the fewer wait states between
// blocks outweigh the less flexible schedule.  The latency kernels, with fewer waves to hide
// a longer dependency chain, keep one block per lane.
#define GL_GROUPED_TAILS 1
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"
#include "passes.h"
#include "pbs_core.h"

using namespace ntt;

namespace lk {

constexpr int MAX_GRID_Y = 65535;  // launches split their per-op grid dimension at this

// ------------------------------------------------------------------------------------
// NTT kernels: NP polynomials per 256-thread workgroup.
// ------------------------------------------------------------------------------------
// mode 0: forward (out NTT domain), 1: inverse (in NTT domain), 2: polymul
template <int NP, int MODE>
__global__ __launch_bounds__(256) void k_ntt(const uint64_t* __restrict__ a, const uint64_t* __restrict__ bn,
                                             uint64_t* __restrict__ out, size_t npoly,
                                             const uint64_t* __restrict__ tab) {
  __shared__ uint64_t lds[NP * N];
  const int t = threadIdx.x;
  const int r = t & 63, m0 = t >> 6, k = t >> 3, w1 = t & 7;
  const size_t base = (size_t)blockIdx.x * NP;
  uint64_t v[NP][8];
  if (MODE == 1) {
#pragma unroll
    for (int p = 0; p < NP; p++) {
      size_t poly = base + p < npoly ? base + p : npoly - 1;
#pragma unroll
      for (int w0 = 0; w0 < 8; w0++) v[p][w0] = a[poly * N + 64 * k + w1 + 8 * w0];
    }
  } else {
#pragma unroll
    for (int p = 0; p < NP; p++) {
      size_t poly = base + p < npoly ? base + p : npoly - 1;
#pragma unroll
      for (int m1 = 0; m1 < 8; m1++) v[p][m1] = a[poly * N + 64 * (m0 + 4 * m1) + r];
    }
    fwd_passes<NP>(v, lds, tab, t);
  }
  if (MODE == 0) {
#pragma unroll
    for (int p = 0; p < NP; p++) {
      if (base + p >= npoly) break;
#pragma unroll
      for (int w0 = 0; w0 < 8; w0++) out[(base + p) * N + 64 * k + w1 + 8 * w0] = gl::canonical(v[p][w0]);
    }
    return;
  }
  if (MODE == 2) {
#pragma unroll
    for (int p = 0; p < NP; p++) {
      size_t poly = base + p < npoly ? base + p : npoly - 1;
#pragma unroll
      for (int w0 = 0; w0 < 8; w0++) v[p][w0] = gl::mul(v[p][w0], bn[poly * N + 64 * k + w1 + 8 * w0]);
    }
  }
  inv_passes<NP>(v, lds, tab, t);
#pragma unroll
  for (int p = 0; p < NP; p++) {
    if (base + p >= npoly) break;
#pragma unroll
    for (int m1 = 0; m1 < 8; m1++) out[(base + p) * N + 64 * (m0 + 4 * m1) + r] = v[p][m1];
  }
}

// ------------------------------------------------------------------------------------
// Bootstrapping key preparation: one workgroup per polynomial (i, gl, j).
// Internal layout [i][w0][K1*L][K1][T]: thread t of the blind rotation reads its own
// 8 values per (gl, j) with fully coalesced loads.
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bsk_prepare(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst,
                                                     int rows, int K1, int is_ntt, const uint64_t* __restrict__ tab) {
  __shared__ uint64_t lds[N];
  const int t = threadIdx.x;
  const int r = t & 63, m0 = t >> 6, k = t >> 3, w1 = t & 7;
  const size_t poly = blockIdx.x;  // = (i * rows + gl) * K1 + j
  const size_t j = poly % K1, gl_ = (poly / K1) % rows, i = poly / ((size_t)K1 * rows);
  uint64_t v[1][8];
  if (is_ntt) {
#pragma unroll
    for (int w0 = 0; w0 < 8; w0++) v[0][w0] = src[poly * N + 64 * k + w1 + 8 * w0];
  } else {
#pragma unroll
    for (int m1 = 0; m1 < 8; m1++) v[0][m1] = src[poly * N + 64 * (m0 + 4 * m1) + r];
    fwd_passes<1>(v, lds, tab, t);
  }
#pragma unroll
  for (int w0 = 0; w0 < 8; w0++)
    dst[(((i * 8 + w0) * rows + gl_) * K1 + j) * T + t] = gl::canonical(v[0][w0]);
}

// ------------------------------------------------------------------------------------
// Blind rotation + sample extract.  Persistent: workgroup b handles ciphertexts
// b, b + gridDim.x, ...  The CMUX order i = 0..n-1 is fixed per ciphertext, so the
// result does not depend on batch composition.
// ------------------------------------------------------------------------------------
// Resident workgroups per CU requested from the register allocator: 32 KB of LDS per
// workgroup at L = 1 allows 4 (4 waves per SIMD at <= 128 VGPRs); L = 2 needs 64 KB.
// HI: L * base_log <= 31 and base_log <= 23: the decomposition reads only the high word of
// each coefficient and its digit lift applies the forward NTT's pre-twist.
template <int K1, int L, bool HI>
__global__ __launch_bounds__(256, (L == 1 ? 4 : 2)) void k_blind_rotate(const uint64_t* __restrict__ bsk,
                                                      const uint64_t* __restrict__ luts,
                                                      const uint16_t* __restrict__ lwe2n,
                                                      const PbsJob* __restrict__ jobs, int nct,
                                                      const uint64_t* __restrict__ tab, BrArgs a) {
  constexpr int NPF = K1 * L;
  __shared__ uint64_t lds[NPF * N];
  const int t = threadIdx.x;
  const int r = t & 63, m0 = t >> 6;
  const size_t blen = (size_t)(K1 - 1) * N + 1;
  const size_t bsk_stride = (size_t)8 * NPF * K1 * T;
  // The key through a buffer resource (its size in bytes, below 2^31, checked at launch):
  // each key word is read at the lane offset 8 t plus a scalar offset, so the MAC's loads
  // take no vector address arithmetic.
  const __amdgpu_buffer_rsrc_t bsk_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)bsk, (short)0, (int)((size_t)a.n_lwe * bsk_stride * 8), 0x00020000);
  for (int c = blockIdx.x; c < nct; c += gridDim.x) {
    const PbsJob job = jobs[c];
    if (job.out_addr == 0) continue;  // an empty slot of a batch list (uniform per workgroup)
    const uint16_t* a2n = lwe2n + (size_t)c * a.lwe_stride;
    uint64_t acc[K1][8];
    const uint64_t* lut = luts + (size_t)job.gid * K1 * N;
#pragma unroll
    for (int g = 0; g < K1; g++)
#pragma unroll
      for (int m1 = 0; m1 < 8; m1++) acc[g][m1] = lut[g * N + 64 * (m0 + 4 * m1) + r];
    uint64_t v[NPF][8];
    for (int i = 0; i < a.n_lwe; i++) {
      const uint32_t rot = a2n[i];
      // Opaque copy of the thread id: keeps the compiler from hoisting every derived
      // shift amount, twiddle and LDS address out of the CMUX loop (which would hold
      // hundreds of registers live); they are recomputed per pass instead.
      int tt = t;
      asm volatile("" : "+v"(tt));
      const int rr = tt & 63, mm0 = tt >> 6;
      __syncthreads();
#pragma unroll
      for (int g = 0; g < K1; g++)
#pragma unroll
        for (int m1 = 0; m1 < 8; m1++) lds[g * N + 64 * (mm0 + 4 * m1) + rr] = acc[g][m1];
      __syncthreads();
#pragma unroll
      for (int g = 0; g < K1; g++) pbs::rotate_decompose<L, HI>(lds + g * N, acc[g], tt, rot, a.base_log, &v[g * L]);
      __syncthreads();
      fwd_passes<NPF, HI>(v, lds, tab, tt);
      const int slice_off = (int)((size_t)i * bsk_stride * 8);
      const int lane_off = tt * 8;
      pbs::mac_with<K1, L>(v, [&](int e) {
        const auto w = __builtin_amdgcn_raw_buffer_load_b64(bsk_rsrc, lane_off, slice_off + e * T * 8, 0);
        return ((uint64_t)w[1] << 32) | w[0];
      });
      inv_passes<K1>(v, lds, tab, tt);
#pragma unroll
      for (int g = 0; g < K1; g++)
#pragma unroll
        for (int m1 = 0; m1 < 8; m1++) acc[g][m1] += gl::modswitch_p_to_q(v[g][m1]);
    }
    // sample extract
    __syncthreads();
#pragma unroll
    for (int g = 0; g < K1; g++)
#pragma unroll
      for (int m1 = 0; m1 < 8; m1++) lds[g * N + 64 * (m0 + 4 * m1) + r] = acc[g][m1];
    __syncthreads();
    const uint32_t b = a2n[a.n_lwe];
    const int lut_nb = 1 << job.log_lut_nb;
    for (int lu = 0; lu < lut_nb; lu++) {
      const uint32_t offs = (uint32_t)(lu * (N >> job.log_lut_nb));
      uint64_t* o = reinterpret_cast<uint64_t*>(job.out_addr) + (size_t)lu * blen;
#pragma unroll
      for (int g = 0; g < K1 - 1; g++)
#pragma unroll
        for (int e = 0; e < N / T; e++) {
          int m = t + T * e;
          o[g * N + m] = pbs::extract_coef(lds, g, m, b, offs);
        }
      if (t == 0) o[(K1 - 1) * N] = pbs::extract_coef(lds, K1 - 1, 0, b, offs);
    }
  }
}

// ------------------------------------------------------------------------------------
// ALU DOps (pea_alu_core.sv:55-160) and LD/ST row copies.
// ------------------------------------------------------------------------------------
// The scalar constant of ADDS/SUBS/SSUB occupies the top `useful` bits of the body only:
// msg_cst[useful-1:0] << (64 - useful), useful = msg + carry + padding bits
// (pea_alu_core.sv:59, param_tfhe_pkg.sv:32); 1 <= useful <= 11 (checked by the host).
__device__ __forceinline__ uint64_t alu_body_cst(uint32_t msg_cst, uint32_t useful) {
  return ((uint64_t)msg_cst & ((1ull << useful) - 1)) << (64 - useful);
}

__global__ __launch_bounds__(256) void k_alu(uint64_t* __restrict__ base, const AluOp* __restrict__ ops, int blen,
                                             uint32_t useful) {
  const AluOp op = ops[blockIdx.y];
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= blen) return;
  const uint64_t x = base[(size_t)op.src0_row * blen + e];
  const uint64_t y = (op.opcode == 0x01 || op.opcode == 0x02 || op.opcode == 0x05)
                         ? base[(size_t)op.src1_row * blen + e] : 0;
  const uint64_t cst = (e == blen - 1) ? alu_body_cst(op.msg_cst, useful) : 0;
  uint64_t z;
  switch (op.opcode) {
    case 0x01: z = x + y; break;
    case 0x02: z = x - y; break;
    case 0x05: z = x * (uint64_t)(op.mul_factor & 31u) + y; break;
    case 0x09: z = x + cst; break;
    case 0x0A: z = x - cst; break;
    case 0x0B: z = cst - x; break;
    default: z = x * (uint64_t)(op.msg_cst & 31u); break;  // 0x0C MULS
  }
  // Reads of all ops of a launch happen before writes only when rows do not alias
  // across ops; the scheduler guarantees that (one dependency level per launch).
  base[(size_t)op.dst_row * blen + e] = z;
}

__global__ __launch_bounds__(256) void k_alu_batch(uint32_t opcode, const uint64_t* __restrict__ a,
                                                   const uint64_t* __restrict__ b, uint64_t* __restrict__ out,
                                                   uint32_t mul_factor, uint32_t msg_cst, int blen, uint32_t useful) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= blen) return;
  const size_t off = (size_t)blockIdx.y * blen + e;
  const uint64_t x = a[off];
  const uint64_t y = (opcode == 0x01 || opcode == 0x02 || opcode == 0x05) ? b[off] : 0;
  const uint64_t cst = (e == blen - 1) ? alu_body_cst(msg_cst, useful) : 0;
  uint64_t z;
  switch (opcode) {
    case 0x01: z = x + y; break;
    case 0x02: z = x - y; break;
    case 0x05: z = x * (uint64_t)(mul_factor & 31u) + y; break;
    case 0x09: z = x + cst; break;
    case 0x0A: z = x - cst; break;
    case 0x0B: z = cst - x; break;
    default: z = x * (uint64_t)(msg_cst & 31u); break;
  }
  out[off] = z;
}

__global__ __launch_bounds__(256) void k_copy(uint64_t* __restrict__ base, const CopyOp* __restrict__ ops, int blen) {
  const CopyOp op = ops[blockIdx.y];
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < blen) base[(size_t)op.dst_row * blen + e] = base[(size_t)op.src_row * blen + e];
}

// Jobs of a list of batches (luthe_pbs_batch_list): block b writes batch b's jobs at the sum
// of the earlier batches' sizes; jobs at or beyond `total` are dropped, and slots of [0, total)
// that no batch fills are marked empty (address 0), which the key switch and the blind
// rotation skip, so a total that disagrees with the descriptors cannot send them to unset
// addresses.  Many-LUT counts are clamped to 8 outputs, LUT ids to the LUT memory.
__global__ __launch_bounds__(256) void k_make_jobs_list(PbsJob* __restrict__ jobs,
                                                        const PbsBatchDesc* __restrict__ descs, int count, int total,
                                                        uint32_t lut_slots, int blen) {
  __shared__ uint64_t part[256];
  const int b = blockIdx.x, t = threadIdx.x;
  if (b == count) {  // the extra block: mark [sum of all sizes, total) empty
    uint64_t s = 0;
    for (int i = t; i < count; i += 256) s += descs[i].n;
    part[t] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) part[t] += part[t + w];
      __syncthreads();
    }
    for (uint64_t c = part[0] + t; c < (uint64_t)total; c += 256) jobs[c] = PbsJob{0, 0, 0, 0};
    return;
  }
  uint64_t s = 0;
  for (int i = t; i < b; i += 256) s += descs[i].n;
  part[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) part[t] += part[t + w];
    __syncthreads();
  }
  const uint64_t off = part[0];
  const PbsBatchDesc d = descs[b];
  const uint32_t lg = d.log_lut_nb < 3 ? d.log_lut_nb : 3;
  for (uint32_t c = t; c < d.n; c += 256) {
    if (off + c >= (uint64_t)total) break;
    uint32_t g = d.gid ? d.gid[c] : 0u;
    g = g < lut_slots ? g : lut_slots - 1;
    jobs[off + c] = PbsJob{(uint64_t)(uintptr_t)(d.in + (size_t)c * blen),
                           (uint64_t)(uintptr_t)(d.out + ((size_t)c << lg) * blen), g, lg};
  }
}

// LUT ids come from caller device memory: ids beyond the LUT memory are clamped to its
// last slot so that a bad id can never read outside the allocation.
__global__ void k_make_jobs(PbsJob* jobs, const uint64_t* in, uint64_t* out, const uint32_t* gid, int n,
                            uint32_t log_lut_nb, uint32_t lut_slots, int blen) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) {
    uint32_t g = gid ? gid[c] : 0u;
    g = g < lut_slots ? g : lut_slots - 1;
    jobs[c] = PbsJob{(uint64_t)(uintptr_t)(in + (size_t)c * blen),
                     (uint64_t)(uintptr_t)(out + ((size_t)c << log_lut_nb) * blen), g, log_lut_nb};
  }
}

// ------------------------------------------------------------------------------------
// Throughput of a minimal Goldilocks modular multiply (64x64 -> 128-bit product with
// four 32-bit multiply-adds, then the 2^64 = 2^32 - 1, 2^96 = -1 fold with no final
// canonicalisation) at full occupancy, 8 independent chains per thread.  Denominator of
// the blind rotation's VALU roofline; deliberately independent of the kernels' own
// primitives so the denominator does not move when they change.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t modmul_min(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  gl::mul_full(a, b, lo, hi);
  uint64_t hh = hi >> 32, hl = hi & gl::EPS;
  uint64_t t0;
  if (__builtin_sub_overflow(lo, hh, &t0)) t0 -= gl::EPS;
  uint64_t t1 = (hl << 32) - hl, r;
  if (__builtin_add_overflow(t0, t1, &r)) r += gl::EPS;
  return r;
}
__global__ __launch_bounds__(256) void k_modmul_peak(uint64_t* out, uint64_t seed, int iters) {
  uint64_t x[8];
  const uint64_t y = seed | 0x8000000000000001ull;
#pragma unroll
  for (int j = 0; j < 8; j++) x[j] = seed * (uint64_t)(threadIdx.x + 977 * j + 1) + blockIdx.x;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int j = 0; j < 8; j++) x[j] = modmul_min(x[j], y);
  }
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < 8; j++) s ^= x[j];
  if (s == 0x123456789ull) out[0] = s;  // keeps the chains live; practically never stores
}

// ------------------------------------------------------------------------------------
// Host launchers
// ------------------------------------------------------------------------------------
template <int NP>
static void ntt_launch(int mode, const uint64_t* a, const uint64_t* bn, uint64_t* out, size_t npoly,
                       const uint64_t* tab, hipStream_t s) {
  dim3 grid((unsigned)((npoly + NP - 1) / NP));
  if (mode == 0)
    k_ntt<NP, 0><<<grid, 256, 0, s>>>(a, bn, out, npoly, tab);
  else if (mode == 1)
    k_ntt<NP, 1><<<grid, 256, 0, s>>>(a, bn, out, npoly, tab);
  else
    k_ntt<NP, 2><<<grid, 256, 0, s>>>(a, bn, out, npoly, tab);
}

hipError_t launch_ntt(int mode, const uint64_t* a, const uint64_t* bn, uint64_t* out, size_t npoly,
                      const uint64_t* tab, hipStream_t s, int polys_per_wg) {
  if (npoly == 0) return hipSuccess;
  switch (polys_per_wg) {
    case 1: ntt_launch<1>(mode, a, bn, out, npoly, tab, s); break;
    case 4: ntt_launch<4>(mode, a, bn, out, npoly, tab, s); break;
    default: ntt_launch<2>(mode, a, bn, out, npoly, tab, s); break;
  }
  return hipGetLastError();
}

hipError_t launch_bsk_prepare(const uint64_t* src, uint64_t* dst, int n_lwe, int rows, int K1, int is_ntt,
                              const uint64_t* tab, hipStream_t s) {
  k_bsk_prepare<<<dim3((unsigned)((size_t)n_lwe * rows * K1)), 256, 0, s>>>(src, dst, rows, K1, is_ntt, tab);
  return hipGetLastError();
}

// Occupancy figures are computed once per process (function-local statics: thread-safe
// initialisation, as an attached server's thread may launch at the same time as the host).
template <int K1, int L>
static int br_max_grid() {
  static const int cached = [] {
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_blind_rotate<K1, L, true>, 256, 0);
    return (cus > 0 ? cus : 256) * (per > 0 ? per : 1);
  }();
  return cached;
}

template <int L, bool HI>
static void br_launch(const uint64_t* bsk, const uint64_t* luts, const uint16_t* lwe2n, const PbsJob* jobs, int nct,
                      const uint64_t* tab, const BrArgs& a, int grid_override, hipStream_t s,
                      bool latency, bool pair, uint64_t* xch, uint32_t* status, bool split_local) {
  if (xch) {
    launch_blind_rotate_split(L, HI, split_local, bsk, luts, lwe2n, jobs, nct, tab, a, s, xch, status);
  } else if (latency || pair) {
    launch_blind_rotate_lat(L, HI, !latency, latency ? 0 : grid_override, bsk, luts, lwe2n, jobs, nct, tab, a, s);
  } else {
    int g = grid_override > 0 ? grid_override : br_max_grid<2, L>();
    g = g < nct ? g : nct;
    k_blind_rotate<2, L, HI><<<dim3(g), 256, 0, s>>>(bsk, luts, lwe2n, jobs, nct, tab, a);
  }
}

hipError_t launch_blind_rotate(const uint64_t* bsk, const uint64_t* luts, const uint16_t* lwe2n, const PbsJob* jobs,
                               int nct, const uint64_t* tab, const BrArgs& a,
                               int pbs_level, int grid_override, hipStream_t s, int latency_max, bool pair,
                               int split_max, uint64_t* xch, uint32_t* status, bool split_local) {
  if (nct == 0) return hipSuccess;
  if (pbs_level != 1 && pbs_level != 2) return hipErrorInvalidValue;
  // the batched kernel addresses the key through a buffer resource of at most 2^31 - 1 bytes
  if ((size_t)a.n_lwe * 8 * (2 * pbs_level) * 2 * T * 8 > 0x7FFFFFFFull) return hipErrorInvalidValue;
  const bool latency = nct <= latency_max;
  if (!(xch && status && nct <= split_max)) xch = nullptr;
  const bool hi = (uint32_t)pbs_level * a.base_log <= 31 && a.base_log <= 23;
  if (pbs_level == 1) {
    if (hi)
      br_launch<1, true>(bsk, luts, lwe2n, jobs, nct, tab, a, grid_override, s, latency, pair, xch, status, split_local);
    else
      br_launch<1, false>(bsk, luts, lwe2n, jobs, nct, tab, a, grid_override, s, latency, pair, xch, status, split_local);
  } else {
    if (hi)
      br_launch<2, true>(bsk, luts, lwe2n, jobs, nct, tab, a, grid_override, s, latency, pair, xch, status, split_local);
    else
      br_launch<2, false>(bsk, luts, lwe2n, jobs, nct, tab, a, grid_override, s, latency, pair, xch, status, split_local);
  }
  return hipGetLastError();
}

int blind_rotate_occupancy(int pbs_level, bool pair) {
  if (pair) return blind_rotate_lat_max_grid(pbs_level, true);
  return pbs_level == 1 ? br_max_grid<2, 1>() : br_max_grid<2, 2>();
}

int blind_rotate_latency_slots(int pbs_level) { return blind_rotate_lat_max_grid(pbs_level, false); }

hipError_t launch_alu(uint64_t* base, const AluOp* ops, int nops, int blen, uint32_t useful_bits, hipStream_t s) {
  if (useful_bits < 1 || useful_bits > 11) return hipErrorInvalidValue;
  for (int o = 0; o < nops; o += MAX_GRID_Y) {
    const int n = (nops - o) < MAX_GRID_Y ? (nops - o) : MAX_GRID_Y;
    k_alu<<<dim3((unsigned)((blen + 255) / 256), (unsigned)n), 256, 0, s>>>(base, ops + o, blen, useful_bits);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_alu_batch(uint32_t opcode, const uint64_t* a, const uint64_t* b, uint64_t* out, uint32_t mul_factor,
                            uint32_t msg_cst, int n, int blen, uint32_t useful_bits, hipStream_t s) {
  if (useful_bits < 1 || useful_bits > 11) return hipErrorInvalidValue;
  for (int o = 0; o < n; o += MAX_GRID_Y) {
    const int m = (n - o) < MAX_GRID_Y ? (n - o) : MAX_GRID_Y;
    const size_t off = (size_t)o * blen;
    k_alu_batch<<<dim3((unsigned)((blen + 255) / 256), (unsigned)m), 256, 0, s>>>(
        opcode, a + off, b ? b + off : nullptr, out + off, mul_factor, msg_cst, blen, useful_bits);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_copy(uint64_t* base, const CopyOp* ops, int nops, int blen, hipStream_t s) {
  for (int o = 0; o < nops; o += MAX_GRID_Y) {
    const int n = (nops - o) < MAX_GRID_Y ? (nops - o) : MAX_GRID_Y;
    k_copy<<<dim3((unsigned)((blen + 255) / 256), (unsigned)n), 256, 0, s>>>(base, ops + o, blen);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_modmul_peak(uint64_t* out, int blocks, int iters, hipStream_t s) {
  k_modmul_peak<<<dim3(blocks), 256, 0, s>>>(out, 0x9E3779B97F4A7C15ull, iters);
  return hipGetLastError();
}

hipError_t launch_make_jobs_list(PbsJob* jobs, const PbsBatchDesc* descs, int count, int total, uint32_t lut_slots,
                                 int blen, hipStream_t s) {
  if (count <= 0 || total <= 0) return hipSuccess;
  k_make_jobs_list<<<dim3((unsigned)count + 1), 256, 0, s>>>(jobs, descs, count, total, lut_slots, blen);
  return hipGetLastError();
}

hipError_t launch_make_jobs(PbsJob* jobs, const uint64_t* in, uint64_t* out, const uint32_t* gid, int n,
                            uint32_t log_lut_nb, uint32_t lut_slots, int blen, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_make_jobs<<<dim3((unsigned)((n + 255) / 256)), 256, 0, s>>>(jobs, in, out, gid, n, log_lut_nb, lut_slots, blen);
  return hipGetLastError();
}

}  // namespace lk
