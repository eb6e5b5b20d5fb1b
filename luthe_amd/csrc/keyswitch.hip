// Batched LWE key switch (gfx950, wave64): an integer GEMM mod 2^32 on VALU with 24-bit
// MACs.  Kept in a translation unit of its own so that it can be scheduled with the
// compiler's max-ILP strategy (10 % faster on this kernel; the blind rotation is faster
// under the default strategy).
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "pbs_core.h"

namespace lk {

namespace {
constexpr int KS_MAX_GRID_Y = 65535;  // grid.y limit: launches split their ciphertexts at this
}  // namespace

// ------------------------------------------------------------------------------------
// Key switch.  Workgroup = CT ciphertexts x 256 output coefficients; K dimension =
// (k*N mask coefficients) x KSL levels, streamed in chunks of CHUNK coefficients whose
// digits are decomposed once into LDS and broadcast to all threads.  CT = 32, CHUNK = 64; for
// batches of at most 16 ciphertexts CT = 16, CHUNK = 32, and of at most 4 CT = 4, CHUNK = 16
// (no MACs for empty ciphertext slots, and the split below can go two or four times finer).
// PART (small batches): workgroup z of gridDim.z takes only the mask coefficients
// [z * kN / gridDim.z, (z + 1) * kN / gridDim.z) and writes its sums to part[z][c][j];
// k_ks_finish adds the gridDim.z partial sums (mod 2^32, in a fixed order) and finishes.
// The sums are exact mod 2^32, so the split changes no result.
// ------------------------------------------------------------------------------------
constexpr int KS_CT = 32;
constexpr int KS_CHUNK = 64;
// (CT, CHUNK, largest split) of a batch of nct ciphertexts
struct KsShape {
  int ct, chunk, smax;
};
inline KsShape ks_shape(int nct) {
  if (nct <= 4) return KsShape{4, 16, 128};
  if (nct <= 16) return KsShape{16, 32, 64};
  return KsShape{KS_CT, KS_CHUNK, 32};
}

template <int KSL, bool PART, int CT = KS_CT, int CHUNK = KS_CHUNK>
__global__ __launch_bounds__(256) void k_keyswitch(const PbsJob* __restrict__ jobs,
                                                   int nct, const uint32_t* __restrict__ ksk,
                                                   uint16_t* __restrict__ lwe2n, uint32_t* __restrict__ part,
                                                   KsArgs a) {
  constexpr int KS_CT = CT, KS_CHUNK = CHUNK;
  __shared__ int32_t dig[KS_CHUNK][KSL][KS_CT / 4];  // 4 signed 8-bit digits per word
  const int t = threadIdx.x;
  const int c0 = blockIdx.y * KS_CT;
  const int j = blockIdx.x * 256 + t;
  const int n1 = a.n_lwe + 1;
  uint32_t acc[KS_CT];  // sums mod 2^32 (the result is taken mod 2^w, w <= 24)
#pragma unroll
  for (int c = 0; c < KS_CT; c++) acc[c] = 0;
  int8_t* dig8 = reinterpret_cast<int8_t*>(&dig[0][0][0]);
  const int span = PART ? a.kN / (int)gridDim.z : a.kN;  // a multiple of KS_CHUNK (host-checked)
  const int i_begin = PART ? (int)blockIdx.z * span : 0, i_end = i_begin + span;
  for (int i0 = i_begin; i0 < i_end; i0 += KS_CHUNK) {
    __syncthreads();
    // CT x CHUNK coefficients (32 x 64: 8 per thread); a row's coefficients are contiguous.
#pragma unroll
    for (int e = 0; e < (KS_CT * KS_CHUNK + 255) / 256; e++) {
      const int idx = e * 256 + t;
      if (KS_CT * KS_CHUNK % 256 == 0 || idx < KS_CT * KS_CHUNK) {
        const int c = idx / KS_CHUNK, ii = idx % KS_CHUNK;
        uint64_t x = 0;
        if (c0 + c < nct && i0 + ii < a.kN) {
          const uint64_t in = jobs[c0 + c].in_addr;  // 0: an empty slot of a batch list
          if (in) x = reinterpret_cast<const uint64_t*>(in)[i0 + ii];
        }
        int64_t d[KSL];
        pbs::decompose<KSL>(x, a.base_log, d);
  #pragma unroll
        for (int l = 0; l < KSL; l++) dig8[(ii * KSL + l) * KS_CT + c] = (int8_t)d[l];
      }
    }
    __syncthreads();
    if (j < n1) {
      const int iimax = (a.kN - i0) < KS_CHUNK ? (a.kN - i0) : KS_CHUNK;
      for (int ii = 0; ii < iimax; ii++) {
#pragma unroll
        for (int l = 0; l < KSL; l++) {
          int32_t kv = (int32_t)ksk[((size_t)(i0 + ii) * KSL + l) * n1 + j];
#pragma unroll
          for (int c4 = 0; c4 < KS_CT / 4; c4++) {
            int32_t w = dig[ii][l][c4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
              int32_t dv = (int32_t)(int8_t)(w >> (8 * e));
              acc[c4 * 4 + e] += (uint32_t)__mul24(dv, kv);
            }
          }
        }
      }
    }
  }
  if (PART) {
    if (j < n1) {
#pragma unroll
      for (int c = 0; c < KS_CT; c++)
        if (c0 + c < nct) part[((size_t)blockIdx.z * nct + c0 + c) * n1 + j] = acc[c];
    }
    return;
  }
  if (j < n1) {
#pragma unroll
    for (int c = 0; c < KS_CT; c++) {
      if (c0 + c >= nct) break;
      uint32_t body = 0;
      if (j == a.n_lwe) {
        const uint64_t in = jobs[c0 + c].in_addr;
        if (in) body = pbs::ks_body(reinterpret_cast<const uint64_t*>(in)[a.kN], a.mod_log);
      }
      lwe2n[(size_t)(c0 + c) * a.lwe_stride + j] = (uint16_t)pbs::ks_finish((uint32_t)acc[c], body, a.mod_log, a.lc);
    }
  }
}

// Sum of the split key switch's partial sums for output j of ciphertext c, then the body and
// the switch to 2N as in k_keyswitch.
__global__ __launch_bounds__(256) void k_ks_finish(const PbsJob* __restrict__ jobs, int nct, int nsplit,
                                                   const uint32_t* __restrict__ part, uint16_t* __restrict__ lwe2n,
                                                   KsArgs a) {
  const int n1 = a.n_lwe + 1;
  const int j = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
  if (j >= n1 || c >= nct) return;
  uint32_t acc = 0;
#pragma unroll 8
  for (int z = 0; z < nsplit; z++) acc += part[((size_t)z * nct + c) * n1 + j];
  uint32_t body = 0;
  if (j == a.n_lwe) {
    const uint64_t in = jobs[c].in_addr;
    if (in) body = pbs::ks_body(reinterpret_cast<const uint64_t*>(in)[a.kN], a.mod_log);
  }
  lwe2n[(size_t)c * a.lwe_stride + j] = (uint16_t)pbs::ks_finish(acc, body, a.mod_log, a.lc);
}

int keyswitch_split(int nct, int kN, int n_lwe) {
  // A workgroup's time is its pass over all k*N inputs, whatever the batch: small batches
  // split that pass until the launch has up to 512 workgroups (2 per CU; measured: splitting an
  // 8192-ciphertext batch in two is slower, 3.4 -> 4.3 ms); each keeps >= one chunk.
  const KsShape sh = ks_shape(nct);
  const long wgs = (long)((n_lwe + 1 + 255) / 256) * ((nct + sh.ct - 1) / sh.ct);
  int s = 1;
  while (s < sh.smax && wgs * s * 2 <= 512 && (kN % (s * 2 * sh.chunk)) == 0) s *= 2;
  return s;
}

hipError_t launch_keyswitch(const PbsJob* jobs, int nct, const uint32_t* ksk, uint16_t* lwe2n,
                            const KsArgs& a, int ks_level, hipStream_t s, uint32_t* part, int nsplit) {
  if (ks_level < 1 || ks_level > 8) return hipErrorInvalidValue;
  if (nsplit > 1) {
    const KsShape sh = ks_shape(nct);
    if (!part || nct > KS_MAX_GRID_Y || a.kN % (nsplit * sh.chunk)) return hipErrorInvalidValue;
    dim3 grid((unsigned)((a.n_lwe + 1 + 255) / 256), (unsigned)((nct + sh.ct - 1) / sh.ct), (unsigned)nsplit);
    if (sh.ct == 4) {
      switch (ks_level) {
#define KS_SCASE(L) \
  case L: k_keyswitch<L, true, 4, 16><<<grid, 256, 0, s>>>(jobs, nct, ksk, nullptr, part, a); break;
        KS_SCASE(1) KS_SCASE(2) KS_SCASE(3) KS_SCASE(4) KS_SCASE(5) KS_SCASE(6) KS_SCASE(7) KS_SCASE(8)
#undef KS_SCASE
      }
    } else if (sh.ct == 16) {
      switch (ks_level) {
#define KS_MCASE(L) \
  case L: k_keyswitch<L, true, 16, 32><<<grid, 256, 0, s>>>(jobs, nct, ksk, nullptr, part, a); break;
        KS_MCASE(1) KS_MCASE(2) KS_MCASE(3) KS_MCASE(4) KS_MCASE(5) KS_MCASE(6) KS_MCASE(7) KS_MCASE(8)
#undef KS_MCASE
      }
    } else {
      switch (ks_level) {
#define KS_PCASE(L) \
  case L: k_keyswitch<L, true><<<grid, 256, 0, s>>>(jobs, nct, ksk, nullptr, part, a); break;
        KS_PCASE(1) KS_PCASE(2) KS_PCASE(3) KS_PCASE(4) KS_PCASE(5) KS_PCASE(6) KS_PCASE(7) KS_PCASE(8)
#undef KS_PCASE
      }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_ks_finish<<<dim3((unsigned)((a.n_lwe + 1 + 255) / 256), (unsigned)nct), 256, 0, s>>>(jobs, nct, nsplit, part,
                                                                                          lwe2n, a);
    return hipGetLastError();
  }
  // grid.y is limited to 65535: chunks of at most 65535 * KS_CT ciphertexts
  for (int c0 = 0; c0 < nct; c0 += KS_MAX_GRID_Y * KS_CT) {
    const int nc = (nct - c0) < KS_MAX_GRID_Y * KS_CT ? (nct - c0) : KS_MAX_GRID_Y * KS_CT;
    dim3 grid((unsigned)((a.n_lwe + 1 + 255) / 256), (unsigned)((nc + KS_CT - 1) / KS_CT));
    const PbsJob* jb = jobs + c0;
    uint16_t* out = lwe2n + (size_t)c0 * a.lwe_stride;
    switch (ks_level) {
#define KS_CASE(L) \
  case L: k_keyswitch<L, false><<<grid, 256, 0, s>>>(jb, nc, ksk, out, nullptr, a); break;
      KS_CASE(1) KS_CASE(2) KS_CASE(3) KS_CASE(4) KS_CASE(5) KS_CASE(6) KS_CASE(7) KS_CASE(8)
#undef KS_CASE
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace lk
