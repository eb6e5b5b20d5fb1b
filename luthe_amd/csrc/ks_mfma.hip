// LWE key switch on the int8 matrix cores (gfx950 v_mfma_i32_32x32x32_i8).
//
// The key switch is an exact integer GEMM: out_j = [j = n] b' - sum_{i,l} d_{i,l} KSK[i][l][j]
// (mod 2^w) with balanced digits |d| <= 2^(b-1) <= 64 and w <= 24
// (pep_ks_mult_node.sv:93-105, pep_ks_out_process.sv:276-282).  Each KSK word v (mod 2^w) is
// held as three balanced bytes v == s0 + 2^8 s1 + 2^16 s2 (mod 2^24), s_p in [-128, 127], so
// the sum is P0 + 2^8 P1 + 2^16 P2 with P_p = sum d * s_p: three int8 GEMMs with int32
// accumulation, exact (|P_p| <= K * 64 * 128 < 2^31 for K = kN * L <= 2^18) and combined mod
// 2^32 before the mod-2^w output stage.  No floating point anywhere.
//
// Operand layouts (lane l of a wave; checked on the device): A fragment = row l % 32,
// k = 16 (l / 32) + [0, 16) of a 32-deep step; B fragment = column l % 32, same k; output
// register r of lane l = C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
//
// K order: step ks = ch * L + l covers level l of mask coefficients 32 ch + [0, 32).  The key
// planes are stored per (plane, column block, step) as 64 lanes x 16 bytes in fragment order,
// so every B fragment is one coalesced 1 KB load.  The digits are decomposed in the kernel
// (high-word closed form of pbs_core.h) into an LDS tile per 32-coefficient chunk.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "pbs_core.h"

namespace lk {

namespace {
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
// Global-address-space view of an input row: loads through a generic pointer made from an
// integer address compile to flat loads, which also count against the LDS wait counter, so
// every LDS wait would drain the prefetched inputs too.
typedef const uint32_t __attribute__((address_space(1)))* gptr_u32;

constexpr int KM_ROWS = 64;    // ciphertexts per workgroup: two 32-row blocks
constexpr int KM_COLS = 128;   // outputs per workgroup: four 32-column blocks
constexpr int KM_GRID_Y = 65535;

__host__ __device__ inline int ceil_div(int a, int b) { return (a + b - 1) / b; }
}  // namespace

__host__ __device__ int ksk_planes_cols(int n_lwe) { return ceil_div(n_lwe + 1, KM_COLS) * KM_COLS; }

size_t ksk_planes_bytes(int kN, int n_lwe, int ks_level) {
  return (size_t)3 * (ksk_planes_cols(n_lwe) / 32) * (size_t)(kN / 32) * ks_level * 1024;
}

bool keyswitch_mfma_ok(const KsArgs& a, int ks_level) {
  return a.kN % 32 == 0 && ks_level >= 1 && ks_level <= 8 && (int)a.base_log * ks_level <= 31 &&
         a.mod_log <= 24 && (size_t)a.kN * ks_level <= (1u << 18);
}

// One thread per (plane, column block, step, lane): the lane's 16 key bytes of one plane.
__global__ __launch_bounds__(256) void k_ksk_planes(const uint32_t* __restrict__ ksk, int8_t* __restrict__ out,
                                                    int kN, int n1, int ksl, int jb_count, uint32_t w) {
  const int ks_count = (kN / 32) * ksl;
  const size_t total = (size_t)3 * jb_count * ks_count * 64;
  const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= total) return;
  const int q = (int)(g & 63);
  size_t rest = g >> 6;
  const int ks = (int)(rest % ks_count);
  rest /= ks_count;
  const int jb = (int)(rest % jb_count);
  const int p = (int)(rest / jb_count);
  const int j = jb * 32 + (q & 31);
  const int ch = ks / ksl, l = ks % ksl;
  const int i0 = ch * 32 + 16 * (q >> 5);
  const uint32_t wmask = (1u << w) - 1;
  int8_t b[16];
#pragma unroll
  for (int e = 0; e < 16; e++) {
    const int64_t v = j < n1 ? (int64_t)(ksk[((size_t)(i0 + e) * ksl + l) * n1 + j] & wmask) : 0;
    const int64_t s0 = (int8_t)(v & 0xFF);
    const int64_t v1 = (v - s0) >> 8;
    const int64_t s1 = (int8_t)(v1 & 0xFF);
    const int64_t v2 = (v1 - s1) >> 8;
    const int64_t s2 = (int8_t)(v2 & 0xFF);
    b[e] = (int8_t)(p == 0 ? s0 : p == 1 ? s1 : s2);
  }
  int8_t* o = out + g * 16;
#pragma unroll
  for (int e = 0; e < 16; e++) o[e] = b[e];
}

hipError_t launch_ksk_planes(const uint32_t* ksk, int8_t* planes, const KsArgs& a, int ks_level, hipStream_t s) {
  const int jb_count = ksk_planes_cols(a.n_lwe) / 32;
  const size_t total = (size_t)3 * jb_count * (a.kN / 32) * ks_level * 64;
  k_ksk_planes<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(ksk, planes, a.kN, a.n_lwe + 1, ks_level, jb_count,
                                                             a.mod_log);
  return hipGetLastError();
}

// Barrier for the LDS tile only: LDS-scope fences, so that the key-fragment and input loads
// issued for the next chunk stay in flight across it (a full __syncthreads() fence would wait
// for them).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The digits of one thread's 8 coefficients (high words xh; zeros unless live) packed as bytes
// into its 8-byte slot of each level's fragment tile.
template <int KSL>
__device__ __forceinline__ void ks_decompose(const uint32_t* xh, bool live, uint32_t b, uint8_t* dst) {
  uint32_t pk[KSL][2];
#pragma unroll
  for (int l = 0; l < KSL; l++) pk[l][0] = pk[l][1] = 0;
#pragma unroll
  for (int e = 0; e < 8; e++) {
    int32_t d[KSL];
    pbs::decompose_hi<KSL>(live ? xh[e] : 0u, b, d);  // zero: zero digits
#pragma unroll
    for (int l = 0; l < KSL; l++) pk[l][e >> 2] |= ((uint32_t)d[l] & 0xFFu) << (8 * (e & 3));
  }
#pragma unroll
  for (int l = 0; l < KSL; l++) *reinterpret_cast<uint2*>(dst + l * 2048) = make_uint2(pk[l][0], pk[l][1]);
}

// Workgroup: 64 ciphertexts x 128 outputs, 4 waves, one workgroup per CU (the wave keeps
// 512 registers: the accumulators, two sets of key fragments and the next input words); wave w
// owns column block w of the workgroup for both row blocks and all three planes (6
// accumulators), so every key fragment is loaded by exactly one wave.  The digit tile is double
// buffered in LDS: chunk ch's MFMAs read tile ch & 1 while each thread decomposes its 8
// coefficients of chunk ch + 1 into the other tile, one barrier per chunk.  Loads run one
// chunk ahead (key fragments) and two ahead (input words).
template <int KSL>
__global__ __launch_bounds__(256, 1) void k_keyswitch_mfma(const PbsJob* __restrict__ jobs, int nct,
                                                           const int8_t* __restrict__ planes,
                                                           uint16_t* __restrict__ lwe2n, KsArgs a) {
  __shared__ v4i atile[2][KSL][2][64];  // [buffer][level][row block][lane]
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int c0 = blockIdx.y * KM_ROWS;
  const int jb_count = ksk_planes_cols(a.n_lwe) / 32, ks_count = (a.kN / 32) * KSL;
  const int jb = blockIdx.x * (KM_COLS / 32) + w;
  const int nch = a.kN / 32;
  v16i acc[2][3];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[rb][p][r] = 0;
  // decomposition role: row drow, coefficients 8 dq + [0, 8) of each chunk; they are bytes
  // 8 (dq & 1) + [0, 8) of fragment lane (drow & 31) + 32 (dq >> 1) of row block drow >> 5
  const int drow = t >> 2, dq = t & 3;
  const uint64_t in = c0 + drow < nct ? jobs[c0 + drow].in_addr : 0;  // 0: no input (zero digits)
  // high words of the coefficients (the decomposition reads only those); rows without input
  // read the key planes instead (at least 384 kN bytes, of which 8 kN are read) and use zeros,
  // so that every load is unconditional and the loads can stay in flight across iterations
  const gptr_u32 xin = (gptr_u32)(in ? in : (uint64_t)(uintptr_t)planes) + dq * 16 + 1;
  uint8_t* tile8 = reinterpret_cast<uint8_t*>(&atile[0][0][0][0]);
  const int toff = ((drow >> 5) * 64 + (drow & 31) + 32 * (dq >> 1)) * 16 + 8 * (dq & 1);
  const uint32_t b = a.base_log;
  // this wave's fragments: v4i index (p * jb_count + jb) * ks_count * 64 + ks * 64 + lane
  const v4i* bsrc = reinterpret_cast<const v4i*>(planes) + (size_t)jb * ks_count * 64 + lane;
  const size_t pstride = (size_t)jb_count * ks_count * 64;  // v4i per plane
  const bool live = in != 0;
  // two register sets for the fragments and the input words, used in turn (no copies between
  // iterations, so a load's wait falls where its data is first used)
  v4i bf0[KSL][3], bf1[KSL][3];
  uint32_t x0[8], x1[8];
  auto load_b = [&](v4i(&bf)[KSL][3], int ch) {
#pragma unroll
    for (int l = 0; l < KSL; l++)
#pragma unroll
      for (int p = 0; p < 3; p++) bf[l][p] = bsrc[p * pstride + (size_t)(ch * KSL + l) * 64];
  };
  auto load_x = [&](uint32_t(&x)[8], int ch) {
#pragma unroll
    for (int e = 0; e < 8; e++) x[e] = xin[2 * (ch * 32 + e)];
  };
  // one chunk: MFMAs of chunk ch from tile ch & 1 with fragments bf, then chunk ch + 1's digits
  // (words x) into the other tile, whose last readers were chunk ch - 1's MFMAs, before the
  // previous barrier; loads of the fragments of ch + 1 into bn and the words of ch + 2 into x
  // (past the last chunk they reload it)
  auto step = [&](int ch, v4i(&bf)[KSL][3], v4i(&bn)[KSL][3], uint32_t(&x)[8], uint32_t(&xnext)[8]) {
    load_b(bn, ch + 1 < nch ? ch + 1 : ch);
    load_x(xnext, ch + 2 < nch ? ch + 2 : nch - 1);
    const v4i(*tile)[2][64] = atile[ch & 1];
#pragma unroll
    for (int l = 0; l < KSL; l++) {
      const v4i a0 = tile[l][0][lane], a1 = tile[l][1][lane];
#pragma unroll
      for (int p = 0; p < 3; p++) {
        acc[0][p] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, bf[l][p], acc[0][p], 0, 0, 0);
        acc[1][p] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, bf[l][p], acc[1][p], 0, 0, 0);
      }
    }
    if (ch + 1 < nch) ks_decompose<KSL>(x, live, b, tile8 + ((ch + 1) & 1) * (KSL * 2048) + toff);
    lds_barrier();
  };
  load_b(bf0, 0);
  load_x(x1, 0);
  ks_decompose<KSL>(x1, live, b, tile8 + toff);  // chunk 0 into buffer 0
  load_x(x0, nch > 1 ? 1 : 0);
  lds_barrier();
  for (int ch = 0; ch < nch; ch += 2) {
    step(ch, bf0, bf1, x0, x1);
    if (ch + 1 < nch) step(ch + 1, bf1, bf0, x1, x0);
  }
  // combine the planes mod 2^32, then the output stage of k_keyswitch
  const int n1 = a.n_lwe + 1;
  const int j = jb * 32 + (lane & 31);
  if (j < n1) {
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int c = c0 + rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (c >= nct) continue;
        const uint32_t s = (uint32_t)acc[rb][0][r] + ((uint32_t)acc[rb][1][r] << 8) + ((uint32_t)acc[rb][2][r] << 16);
        uint32_t body = 0;
        if (j == a.n_lwe) {
          const uint64_t cin = jobs[c].in_addr;
          if (cin) body = pbs::ks_body(reinterpret_cast<const uint64_t*>(cin)[a.kN], a.mod_log);
        }
        lwe2n[(size_t)c * a.lwe_stride + j] = (uint16_t)pbs::ks_finish(s, body, a.mod_log, a.lc);
      }
  }
}

hipError_t launch_keyswitch_mfma(const PbsJob* jobs, int nct, const int8_t* planes, uint16_t* lwe2n, const KsArgs& a,
                                 int ks_level, hipStream_t s) {
  if (!keyswitch_mfma_ok(a, ks_level)) return hipErrorInvalidValue;
  const unsigned gx = (unsigned)(ksk_planes_cols(a.n_lwe) / KM_COLS);
  for (int c0 = 0; c0 < nct; c0 += KM_GRID_Y * KM_ROWS) {
    const int nc = (nct - c0) < KM_GRID_Y * KM_ROWS ? (nct - c0) : KM_GRID_Y * KM_ROWS;
    const dim3 grid(gx, (unsigned)ceil_div(nc, KM_ROWS));
    const PbsJob* jb = jobs + c0;
    uint16_t* out = lwe2n + (size_t)c0 * a.lwe_stride;
    switch (ks_level) {
#define KM_CASE(L) \
  case L: k_keyswitch_mfma<L><<<grid, 256, 0, s>>>(jb, nc, planes, out, a); break;
      KM_CASE(1) KM_CASE(2) KM_CASE(3) KM_CASE(4) KM_CASE(5) KM_CASE(6) KM_CASE(7) KM_CASE(8)
#undef KM_CASE
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace lk
