// The four NTT passes of one 256-thread group (ntt2048.h) over NP polynomials held in
// registers, with the LDS exchanges and barriers between them: shared by the 256-thread
// kernels (kernels.hip) and the 512-thread latency and pair kernels (latency.hip), each half
// of which runs them on its own polynomials.
#pragma once
#include <hip/hip_runtime.h>

#include "ntt2048.h"

namespace lk {

using namespace ntt;

// tab: twiddle tables (ntt2048.h TABLE_WORDS layout).
// PRE: the inputs are already pretwisted by psi^(4 m1) (blind-rotation digits, HI path).
template <int NP, bool PRE = false>
__device__ __forceinline__ void fwd_passes(uint64_t (*v)[8], uint64_t* lds, const uint64_t* tab, int t) {
  const int r = t & 63, m0 = t >> 6, k = t >> 3, j = t & 7;
  const int m0u = wave_uniform(m0);
  const uint64_t* tw = tab + TW_OFF;
  const uint64_t* c64 = tab + C64_OFF;
#pragma unroll
  for (int p = 0; p < NP; p++) fwd_a1_compute(v[p], m0u, PRE);
#pragma unroll
  for (int p = 0; p < NP; p++) fwd_a1_store(v[p], lds + p * N, r, m0);
  __syncthreads();
#pragma unroll
  for (int p = 0; p < NP; p++) {
    fwd_a2_load(v[p], lds + p * N, r, m0);
    fwd_a2_compute(v[p], r, m0u, tw);
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < NP; p++) fwd_a2_store(v[p], lds + p * N, r, m0);
  __syncthreads();
  fwd_c1_load<NP>(v, lds, t);
  fwd_c1_compute<NP>(v, t, c64);
  __syncthreads();
  fwd_c1_store<NP>(v, lds, t);
  __syncthreads();
#pragma unroll
  for (int p = 0; p < NP; p++) {
    fwd_c2_load(v[p], lds + p * N, k, j);
    fwd_c2_compute(v[p]);
  }
}

// Entry: v in NTT domain (thread (k, w1) layout), LDS free for writing only after a
// barrier that orders the previous readers.  Exit: v[p][m1] canonical coefficients.
template <int NP>
__device__ __forceinline__ void inv_passes(uint64_t (*v)[8], uint64_t* lds, const uint64_t* tab, int t) {
  // Opaque copy of the thread id: the index values below are derived here rather than
  // kept live across the caller's code (which would spill at 128 VGPRs).
  asm volatile("" : "+v"(t));
  const int r = t & 63, m0 = t >> 6, k = t >> 3, j = t & 7;
  const int m0u = wave_uniform(m0);
  const uint64_t* itw = tab + ITW_OFF;
  const uint64_t* ic64 = tab + IC64_OFF;
#pragma unroll
  for (int p = 0; p < NP; p++) inv_c2_compute(v[p]);
  __syncthreads();
#pragma unroll
  for (int p = 0; p < NP; p++) inv_c2_store(v[p], lds + p * N, k, j);
  __syncthreads();
  inv_c1_load<NP>(v, lds, t);
  inv_c1_compute<NP>(v, t, itw, ic64);
  __syncthreads();
  inv_c1_store<NP>(v, lds, t);
  __syncthreads();
#pragma unroll
  for (int p = 0; p < NP; p++) {
    inv_a2_load(v[p], lds + p * N, r, m0);
    inv_a2_compute(v[p], m0u);
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < NP; p++) inv_a2_store(v[p], lds + p * N, r, m0);
  __syncthreads();
#pragma unroll
  for (int p = 0; p < NP; p++) {
    inv_a1_load(v[p], lds + p * N, r, m0);
    inv_a1_compute(v[p]);
  }
}

}  // namespace lk
