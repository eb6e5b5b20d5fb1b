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

// Workgroup: 64 ciphertexts x 128 outputs, 4 waves; wave w owns row block w & 1 and column
// blocks 2 (w >> 1) + {0, 1} of the workgroup, all three planes (6 accumulators).  Per chunk
// of 32 mask coefficients: each thread decomposes 8 coefficients of one row into the LDS
// tile, then every wave runs L steps of 6 MFMAs.
template <int KSL>
__global__ __launch_bounds__(256) void k_keyswitch_mfma(const PbsJob* __restrict__ jobs, int nct,
                                                        const int8_t* __restrict__ planes,
                                                        uint16_t* __restrict__ lwe2n, KsArgs a) {
  __shared__ v4i atile[KSL][2][64];  // [level][row block][lane]
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int c0 = blockIdx.y * KM_ROWS;
  const int jb_count = ksk_planes_cols(a.n_lwe) / 32, ks_count = (a.kN / 32) * KSL;
  const int rb = w & 1, jb0 = blockIdx.x * (KM_COLS / 32) + 2 * (w >> 1);
  v16i acc[2][3];
#pragma unroll
  for (int cb = 0; cb < 2; cb++)
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[cb][p][r] = 0;
  // decomposition role: row drow, coefficients 8 dq + [0, 8) of each chunk; they are bytes
  // 8 (dq & 1) + [0, 8) of fragment lane (drow & 31) + 32 (dq >> 1) of row block drow >> 5
  const int drow = t >> 2, dq = t & 3;
  const uint64_t in = c0 + drow < nct ? jobs[c0 + drow].in_addr : 0;  // 0: no input (zero digits)
  const uint64_t* xin = reinterpret_cast<const uint64_t*>(in);
  uint8_t* tile8 = reinterpret_cast<uint8_t*>(&atile[0][0][0]);
  const int toff = ((drow >> 5) * 64 + (drow & 31) + 32 * (dq >> 1)) * 16 + 8 * (dq & 1);
  const uint32_t b = a.base_log;
  const size_t pstride = (size_t)jb_count * ks_count * 1024;  // bytes per plane
  for (int ch = 0; ch < a.kN / 32; ch++) {
    uint32_t pk[KSL][2];
    if (in) {
#pragma unroll
      for (int l = 0; l < KSL; l++) pk[l][0] = pk[l][1] = 0;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const uint32_t xh = (uint32_t)(xin[ch * 32 + dq * 8 + e] >> 32);
        int32_t d[KSL];
        pbs::decompose_hi<KSL>(xh, b, d);
#pragma unroll
        for (int l = 0; l < KSL; l++) pk[l][e >> 2] |= ((uint32_t)d[l] & 0xFFu) << (8 * (e & 3));
      }
    }
    __syncthreads();  // the previous chunk's fragments are read
    if (in || ch == 0) {  // rows without input hold zero digits from the first chunk on
#pragma unroll
      for (int l = 0; l < KSL; l++) {
        uint2 v;
        v.x = in ? pk[l][0] : 0u;
        v.y = in ? pk[l][1] : 0u;
        *reinterpret_cast<uint2*>(tile8 + l * 2048 + toff) = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < KSL; l++) {
      const int ks = ch * KSL + l;
      const v4i af = atile[l][rb][lane];
#pragma unroll
      for (int cb = 0; cb < 2; cb++)
#pragma unroll
        for (int p = 0; p < 3; p++) {
          const v4i bf = *reinterpret_cast<const v4i*>(planes + p * pstride +
                                                       (((size_t)(jb0 + cb) * ks_count + ks) * 64 + lane) * 16);
          acc[cb][p] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, bf, acc[cb][p], 0, 0, 0);
        }
    }
  }
  // combine the planes mod 2^32, then the output stage of k_keyswitch
  const int n1 = a.n_lwe + 1;
#pragma unroll
  for (int cb = 0; cb < 2; cb++) {
    const int j = (jb0 + cb) * 32 + (lane & 31);
    if (j >= n1) continue;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int c = c0 + rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      if (c >= nct) continue;
      const uint32_t s = (uint32_t)acc[cb][0][r] + ((uint32_t)acc[cb][1][r] << 8) + ((uint32_t)acc[cb][2][r] << 16);
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
