// The split blind rotation (one ciphertext per pair of 512-thread workgroups on two CUs of one
// XCC) and its placement probe, in a translation unit of their own so that they can be
// scheduled with the compiler's iterative-minreg strategy (faster for this kernel at PBS level
// 2, equal at level 1; the throughput kernels in kernels.hip are faster under the default).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"
#include "ntt2048x4.h"
#include "pbs_core.h"

using namespace ntt;

namespace lk {

// ------------------------------------------------------------------------------------
// Split variant of the blind rotation for the smallest cohorts: ciphertext c runs on a pair
// of 512-thread workgroups, one per GLWE polynomial, so that each CU carries half of every
// CMUX with two waves per SIMD: the rotation and decomposition of its polynomial, the forward
// NTTs of its L digit polynomials (ntt2048x4.h, four values per thread), output column h of
// the external product and its inverse NTT.  The product also needs the partner's digits:
// thread t of one workgroup sends its NTT-domain digits to thread t of the other through
// global memory, as 64-bit words of 32 data bits under a 32-bit tag (CMUX index + 1), polled
// with relaxed agent-scope atomic loads (which miss the CU's L1 and read the L2).  A word whose
// tag matches carries its own data, so neither fences nor flags are needed.  LOCAL: the stores
// are relaxed atomics of workgroup scope, written through the L1 into the L2 of the XCC that
// both workgroups run on, where the partner's loads find them (on MI355X a partner's word
// arrives about 0.4 us per CMUX sooner than with agent-scope stores, which go past the L2);
// otherwise they are agent-scope stores.  LOCAL needs both workgroups of every pair on one
// XCC, which the host checks once per context (split_pairs_share_xcc); the pair also checks
// it at the start of each launch: each publishes its XCC id (hardware register XCC_ID) in a
// handshake word with agent-scope atomics and reads its partner's.  The two parities of the exchange area
// alternate per CMUX; a workgroup overwrites parity i & 1 at CMUX i + 2 only after receiving
// its partner's CMUX i + 1 words, which the partner sends after it has read CMUX i.  The area
// must hold no tag of this launch when the launch starts (the launcher zeroes it first).
// Workgroups bid and bid ^ 8 form a pair, so that round-robin dispatch over the 8 XCDs puts
// both on one XCD's L2.  Every wait ends: a partner silent for SPLIT_WAIT_TICKS of the
// 100 MHz real-time clock makes the workgroup set *status and leave, and its partner then
// does the same (outputs undefined; the host reports the status as an error).
// BrArgs::split_wait_us overrides the limit (0: give up at the first poll that finds a word
// missing, which exercises this path in tests).  Arithmetic and CMUX order are those of
// k_blind_rotate, so results are bit-identical to it.
// ------------------------------------------------------------------------------------
constexpr uint64_t SPLIT_WAIT_TICKS = 200000000ull;  // 2 s

template <int L>
constexpr size_t split_xch_words() {  // [parity 2][side 2][L][j 4][half 2][512 threads]
  return (size_t)2 * 2 * L * 8 * ntt4::T4;
}
template <int L>
constexpr size_t split_words() {  // the exchange words, then 8 words: the handshake [side]
  return split_xch_words<L>() + 8;
}

// This thread's 8 L tagged words of its digits v to snd (word k at snd[k * T4]); LOCAL: the
// partner runs on this workgroup's XCC (the words stay in the shared L2).
template <int L, bool LOCAL>
__device__ __forceinline__ void split_send(uint64_t* snd, uint64_t tag, const uint64_t (*v)[4]) {
#pragma unroll
  for (int l = 0; l < L; l++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const uint64_t w = tag | (uint32_t)(v[l][j] >> (32 * q));
        uint64_t* p = snd + ((l * 4 + j) * 2 + q) * ntt4::T4;
        if (LOCAL)
          __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else
          __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
}

__global__ __launch_bounds__(256) void k_zero_words(uint64_t* __restrict__ p, size_t n) {
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) p[e] = 0;
}

// 512-thread passes (ntt2048x4.h) over NP polynomials, with the two LDS buffers p and q used
// alternately (one barrier per pass): A1a -> q, A1b -> p, A2 -> q, C1 -> p, C2 -> q, C3 reads q.
// Entry: no thread still reads q, and p only by reads that precede the first barrier here.
template <int NP, bool PRE>
__device__ __forceinline__ void fwd_passes_x4(uint64_t (*v)[4], uint64_t* p, uint64_t* q, const ntt4::Tw4& tw, int t) {
  const ntt4::Map m(t);
  const int W = wave_uniform(t >> 6);
#pragma unroll
  for (int pp = 0; pp < NP; pp++) ntt4::fwd_a1a_compute<PRE>(v[pp], W >> 2);
#pragma unroll
  for (int pp = 0; pp < NP; pp++) ntt4::fwd_a1a_store(v[pp], q + pp * N, m);
  __syncthreads();
#pragma unroll
  for (int pp = 0; pp < NP; pp++) {
    ntt4::fwd_a1b_load(v[pp], q + pp * N, m);
    ntt4::fwd_a1b_compute(v[pp], W);
  }
#pragma unroll
  for (int pp = 0; pp < NP; pp++) ntt4::fwd_a1b_store(v[pp], p + pp * N, m);
  __syncthreads();
#pragma unroll
  for (int pp = 0; pp < NP; pp++) {
    ntt4::fwd_a2_load(v[pp], p + pp * N, m);
    ntt4::fwd_a2_compute(v[pp], tw.tw);
  }
#pragma unroll
  for (int pp = 0; pp < NP; pp++) ntt4::fwd_a2_store(v[pp], q + pp * N, m);
  __syncthreads();
#pragma unroll
  for (int pp = 0; pp < NP; pp++) {
    ntt4::fwd_c1_load(v[pp], q + pp * N, m);
    ntt4::fwd_c1_compute(v[pp], tw.c1);
  }
#pragma unroll
  for (int pp = 0; pp < NP; pp++) ntt4::fwd_c1_store(v[pp], p + pp * N, m);
  __syncthreads();
#pragma unroll
  for (int pp = 0; pp < NP; pp++) {
    ntt4::fwd_c2_load(v[pp], p + pp * N, m);
    ntt4::fwd_c2_compute(v[pp], W & 3);
  }
#pragma unroll
  for (int pp = 0; pp < NP; pp++) ntt4::fwd_c2_store(v[pp], q + pp * N, m);
  __syncthreads();
#pragma unroll
  for (int pp = 0; pp < NP; pp++) {
    ntt4::fwd_c3_load(v[pp], q + pp * N, m);
    ntt4::fwd_c3_compute(v[pp]);
  }
}
// The inverse of one polynomial: IC3 -> p, IC2 -> q, IC1 -> p, IA2 -> q, IA1b -> p, IA1a reads
// p.  Entry: no thread still reads p, and q only by reads that precede the first barrier here.
__device__ __forceinline__ void inv_passes_x4(uint64_t* v, uint64_t* p, uint64_t* q, const ntt4::Tw4& tw, int t) {
  const ntt4::Map m(t);
  const int W = wave_uniform(t >> 6);
  ntt4::inv_c3_compute(v);
  ntt4::inv_c3_store(v, p, m);
  __syncthreads();
  ntt4::inv_c2_load(v, p, m);
  ntt4::inv_c2_compute(v, W & 3);
  ntt4::inv_c2_store(v, q, m);
  __syncthreads();
  ntt4::inv_c1_load(v, q, m);
  ntt4::inv_c1_compute(v, tw.ic1, tw.itw);
  ntt4::inv_c1_store(v, p, m);
  __syncthreads();
  ntt4::inv_a2_load(v, p, m);
  ntt4::inv_a2_compute(v, W);
  ntt4::inv_a2_store(v, q, m);
  __syncthreads();
  ntt4::inv_a1b_load(v, q, m);
  ntt4::inv_a1b_compute(v);
  ntt4::inv_a1b_store(v, p, m);
  __syncthreads();
  ntt4::inv_a1a_load(v, p, m);
  ntt4::inv_a1a_compute(v, W >> 2);
}

// The split kernel's workgroup: 512 threads on one GLWE polynomial.  Thread t holds
// NTT-domain indices 64k + w1 + 8 (wd + 2j) (ntt2048x4.h's C3 map), so it reads the key words
// of thread t & 255 of the blind-rotation key layout, rows w0 = wd + 2j, and exchanges 8 L
// words with thread t of its partner.
template <int L, bool HI, bool LOCAL>
__global__ __launch_bounds__(512, 1) void k_blind_rotate_split(const uint64_t* __restrict__ bsk,
                                                               const uint64_t* __restrict__ luts,
                                                               const uint16_t* __restrict__ lwe2n,
                                                               const PbsJob* __restrict__ jobs, int nct,
                                                               const uint64_t* __restrict__ tab, BrArgs a,
                                                               uint64_t* __restrict__ xch, uint32_t* status) {
  constexpr int K1 = 2, NPF = K1 * L, T4 = ntt4::T4;
  __shared__ uint64_t lds[2 * L * N];
  __shared__ int abort_;
  const int bid = blockIdx.x;
  const int c = (bid >> 4) * 8 + (bid & 7);
  const int h = (bid >> 3) & 1;
  if (c >= nct) return;  // both workgroups of a pair leave together
  const PbsJob job = jobs[c];
  if (job.out_addr == 0) return;  // an empty slot of a batch list
  const int t = threadIdx.x;
  const int r = t & 63, m0 = (t >> 6) & 3, g = t >> 8;
  const size_t bsk_stride = (size_t)8 * NPF * K1 * T;
  const __amdgpu_buffer_rsrc_t bsk_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)bsk, (short)0, (int)((size_t)a.n_lwe * bsk_stride * 8), 0x00020000);
  const uint16_t* a2n = lwe2n + (size_t)c * a.lwe_stride;
  const uint64_t* lut = luts + (size_t)job.gid * K1 * N;
  uint64_t* xc = xch + (size_t)c * split_words<L>();
  ntt4::Tw4 tw;
  tw.load(tab, t);
  const uint64_t wait_ticks = a.split_wait_us < 0 ? SPLIT_WAIT_TICKS : (uint64_t)a.split_wait_us * 100;
  if (t == 0) abort_ = 0;
  if (LOCAL && t == 0) {
    // handshake: bit 63 marks a published word, bits 0..3 the XCC id; a partner on another
    // XCC would not see this workgroup's words, so the pair gives up as for a silent partner
    uint32_t xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    xcc &= 15u;
    uint64_t* hs = xc + split_xch_words<L>();
    __hip_atomic_store(hs + h, (1ull << 63) | xcc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t pw, deadline = 0;
    int ab = 0;
    for (;;) {
      pw = __hip_atomic_load(hs + (1 - h), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (pw >> 63) break;
      const uint64_t now = __builtin_amdgcn_s_memrealtime();
      if (deadline == 0) deadline = now + wait_ticks;
      if (now >= deadline) {
        __hip_atomic_store(status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        ab = 1;
        break;
      }
    }
    if (!ab && (uint32_t)(pw & 15u) != xcc) {
      __hip_atomic_store(status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      ab = 1;
    }
    abort_ = ab;
  }
  uint64_t acc[4];
#pragma unroll
  for (int i = 0; i < 4; i++) acc[i] = lut[h * N + 64 * (m0 + 4 * (2 * g + (i & 1) + 4 * (i >> 1))) + r];
  uint32_t rot_next = a2n[0];
  // Buffers per CMUX (x: accumulator, y: the other), one barrier per pass: acc -> x |
  // rotation reads x | forward A1a -> y, A1b -> x, A2 -> y, C1 -> x, C2 -> y, C3 reads y |
  // inverse IC3 -> x, IC2 -> y, IC1 -> x, IA2 -> y, IA1b -> x, IA1a reads x; the next CMUX
  // swaps x and y.  Every store's buffer was last read before the preceding barrier.
  uint64_t* bx = lds;
  uint64_t* by = lds + L * N;
  __syncthreads();  // abort_ initialised
  for (int i = 0; i < a.n_lwe; i++) {
    const uint32_t rot = rot_next;
    rot_next = a2n[i + 1];  // i + 1 <= n_lwe: the body slot, always present
    if (abort_) break;  // set before the last barrier of the previous CMUX: uniform here
    // key words of output column h for this thread's rows w0 = wd + 2j: own rows first
    const int slice_off = (int)((size_t)i * bsk_stride * 8);
    const int lane_off = (t & 255) * 8;
    uint64_t kw[4][2][L];
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int side = 0; side < 2; side++)
#pragma unroll
        for (int l = 0; l < L; l++) {
          const int gg = side ? 1 - h : h;
          // g is uniform per wave: its readfirstlane copy keeps the scalar offset in SGPRs
          // (a per-lane value there compiles to a loop over the distinct lane values)
          const int e = ((wave_uniform(g) + 2 * j) * NPF + gg * L + l) * K1 + h;
          const auto w = __builtin_amdgcn_raw_buffer_load_b64(bsk_rsrc, lane_off, slice_off + e * T * 8, 0);
          kw[j][side][l] = ((uint64_t)w[1] << 32) | w[0];
        }
#pragma unroll
    for (int k = 0; k < 4; k++) bx[64 * (m0 + 4 * (2 * g + (k & 1) + 4 * (k >> 1))) + r] = acc[k];
    __syncthreads();
    uint64_t v[L][4];
    pbs::rotate_decompose4<L, HI>(bx, acc, t, rot, a.base_log, v);
    fwd_passes_x4<L, HI>(v, bx, by, tw, t);
    // send this polynomial's digit rows h * L + l
    const uint64_t tag = (uint64_t)(i + 1) << 32;
    uint64_t* snd = xc + (size_t)((i & 1) * 2 + h) * L * 8 * T4 + t;
    split_send<L, LOCAL>(snd, tag, v);
    uint64_t lo[4], hi[4];
    uint32_t cc[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      cc[j] = 0;
      gl::mul_full(v[0][j], kw[j][0][0], lo[j], hi[j]);
    }
#pragma unroll
    for (int l = 1; l < L; l++) {
      uint64_t plo[4], phi[4];
#pragma unroll
      for (int j = 0; j < 4; j++) gl::mul_full(v[l][j], kw[j][0][l], plo[j], phi[j]);
      gl::acc128x4(lo, hi, cc, plo, phi);
    }
    const uint64_t* rcv = xc + (size_t)((i & 1) * 2 + (1 - h)) * L * 8 * T4 + t;
    uint64_t got[L][4][2];
    uint64_t deadline = 0;
    for (;;) {
      bool ready = true;
#pragma unroll
      for (int l = 0; l < L; l++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
          for (int q = 0; q < 2; q++) {
            got[l][j][q] = __hip_atomic_load(rcv + ((l * 4 + j) * 2 + q) * T4, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
            ready &= (got[l][j][q] >> 32) == (uint64_t)(i + 1);
          }
      if (ready) break;
      const uint64_t now = __builtin_amdgcn_s_memrealtime();
      if (deadline == 0) deadline = now + wait_ticks;
      if (now >= deadline) {
        __hip_atomic_store(status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        abort_ = 1;  // read by every thread after the next barrier
        break;
      }
    }
#pragma unroll
    for (int l = 0; l < L; l++) {
      uint64_t plo[4], phi[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint64_t d = (got[l][j][1] << 32) | (uint32_t)got[l][j][0];
        gl::mul_full(d, kw[j][1][l], plo[j], phi[j]);
      }
      gl::acc128x4(lo, hi, cc, plo, phi);
    }
    uint64_t z[4];
    gl::reduce160x4(z, lo, hi, cc);
    inv_passes_x4(z, bx, by, tw, t);
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += gl::modswitch_p_to_q(z[k]);
    uint64_t* sw = bx;
    bx = by;
    by = sw;
  }
  __syncthreads();
  if (abort_) return;
  // sample extract: the mask from polynomial 0, the body from polynomial 1
#pragma unroll
  for (int k = 0; k < 4; k++) lds[64 * (m0 + 4 * (2 * g + (k & 1) + 4 * (k >> 1))) + r] = acc[k];
  __syncthreads();
  const uint32_t b = a2n[a.n_lwe];
  const size_t blen = (size_t)(K1 - 1) * N + 1;
  const int lut_nb = 1 << job.log_lut_nb;
  for (int lu = 0; lu < lut_nb; lu++) {
    const uint32_t offs = (uint32_t)(lu * (N >> job.log_lut_nb));
    uint64_t* o = reinterpret_cast<uint64_t*>(job.out_addr) + (size_t)lu * blen;
    if (h == 0) {
#pragma unroll
      for (int e = 0; e < N / T4; e++) {
        const int mm = t + T4 * e;
        o[mm] = pbs::extract_coef(lds, 0, mm, b, offs);
      }
    } else if (t == 0) {
      o[N] = pbs::extract_coef(lds, 0, 0, b, offs);
    }
  }
}


// Whether workgroups b and b ^ 8 of a split-kernel grid (512 threads, up to 256 workgroups)
// run on one XCC: a probe kernel records each workgroup's XCC id.  false on any error.
__global__ __launch_bounds__(512) void k_xcc_ids(uint32_t* __restrict__ out) {
  uint32_t x;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
  if (threadIdx.x == 0) out[blockIdx.x] = x & 15u;
}
bool split_pairs_share_xcc(hipStream_t s) {
  constexpr int G = 256;
  uint32_t* d = nullptr;
  uint32_t h[G];
  if (hipMalloc(&d, G * sizeof(uint32_t)) != hipSuccess) return false;
  bool ok = false;
  k_xcc_ids<<<dim3(G), 512, 0, s>>>(d);
  if (hipGetLastError() == hipSuccess && hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, s) == hipSuccess &&
      hipStreamSynchronize(s) == hipSuccess) {
    ok = true;
    for (int b = 0; b < G; b++) ok = ok && h[b] == h[b ^ 8];
  }
  (void)hipFree(d);
  return ok;
}

size_t blind_rotate_split_words(int pbs_level) { return pbs_level == 1 ? split_words<1>() : split_words<2>(); }

int blind_rotate_split_slots() {
  static const int cached = [] {
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return (cus > 0 ? cus : 256) / 2;
  }();
  return cached;
}


void launch_blind_rotate_split(int pbs_level, bool hi, bool local, const uint64_t* bsk, const uint64_t* luts,
                               const uint16_t* lwe2n, const PbsJob* jobs, int nct, const uint64_t* tab,
                               const BrArgs& a, hipStream_t s, uint64_t* xch, uint32_t* status) {
  const size_t words = (size_t)nct * blind_rotate_split_words(pbs_level);
  k_zero_words<<<dim3((unsigned)std::min<size_t>((words + 255) / 256, 1024)), 256, 0, s>>>(xch, words);
  const dim3 grid((unsigned)(16 * ((nct + 7) / 8)));
#define SPLIT_CASE(L, H, LOC) \
  if (pbs_level == L && hi == H && local == LOC) \
    k_blind_rotate_split<L, H, LOC><<<grid, 512, 0, s>>>(bsk, luts, lwe2n, jobs, nct, tab, a, xch, status);
  SPLIT_CASE(1, true, true) SPLIT_CASE(1, true, false) SPLIT_CASE(1, false, true) SPLIT_CASE(1, false, false)
  SPLIT_CASE(2, true, true) SPLIT_CASE(2, true, false) SPLIT_CASE(2, false, true) SPLIT_CASE(2, false, false)
#undef SPLIT_CASE
}

}  // namespace lk
