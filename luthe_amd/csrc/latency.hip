// Latency and pair blind rotations (one ciphertext per 512-thread workgroup), in their own
// translation unit: they keep one asm block per reduction tail (GL_GROUPED_TAILS unset), which
// suits their two waves per SIMD (latency variant) and barrier-coupled halves (pair variant)
// better than the grouped tails of the 256-thread throughput kernel in kernels.hip.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "passes.h"
#include "pbs_core.h"

namespace lk {

// ------------------------------------------------------------------------------------
// Latency variant of the blind rotation for small cohorts: one ciphertext per 512-thread
// workgroup.  Half h (256 threads, same per-thread mapping as above) owns GLWE
// polynomial h: its rotation and decomposition, the forward NTTs of its L digit
// polynomials, output column j = h of the external product and its inverse NTT.  The
// halves exchange NTT-domain digits through LDS before the product.  Arithmetic is the
// same canonical Goldilocks arithmetic, so results equal k_blind_rotate's bit for bit.
// The key words are read in groups of four outputs through a buffer resource and the
// products summed as 128-bit values with one reduction per output, as in k_blind_rotate.
// TP (the throughput pair variant): two workgroups per CU (at most 128 VGPRs, so four
// waves per SIMD); without TP one workgroup per CU (the latency variant for small cohorts).
// ------------------------------------------------------------------------------------
template <int K1, int L, bool HI, bool TP>
__global__ __launch_bounds__(512, TP ? 2 : 1) void k_blind_rotate_lat(const uint64_t* __restrict__ bsk,
                                                             const uint64_t* __restrict__ luts,
                                                             const uint16_t* __restrict__ lwe2n,
                                                             const PbsJob* __restrict__ jobs, int nct,
                                                             const uint64_t* __restrict__ tab, BrArgs a) {
  static_assert(K1 == 2, "one half per GLWE polynomial");
  constexpr int NPF = K1 * L;
  __shared__ uint64_t lds[NPF * N];  // acc (K1 polys) / digit NTTs (K1*L) / exchange (K1*L*8*T)
  const int t2 = threadIdx.x;
  const int h = t2 >> 8, t = t2 & 255;
  const int r = t & 63, m0 = t >> 6;
  const size_t blen = (size_t)(K1 - 1) * N + 1;
  const size_t bsk_stride = (size_t)8 * NPF * K1 * T;
  const __amdgpu_buffer_rsrc_t bsk_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)bsk, (short)0, (int)((size_t)a.n_lwe * bsk_stride * 8), 0x00020000);
  for (int c = blockIdx.x; c < nct; c += gridDim.x) {
    const PbsJob job = jobs[c];
    if (job.out_addr == 0) continue;  // an empty slot of a batch list (uniform per workgroup)
    const uint16_t* a2n = lwe2n + (size_t)c * a.lwe_stride;
    const uint64_t* lut = luts + (size_t)job.gid * K1 * N;
    uint64_t acc[8];
#pragma unroll
    for (int m1 = 0; m1 < 8; m1++) acc[m1] = lut[h * N + 64 * (m0 + 4 * m1) + r];
    uint32_t rot_next = a2n[0];
    for (int i = 0; i < a.n_lwe; i++) {
      const uint32_t rot = rot_next;
      rot_next = a2n[i + 1];  // i + 1 <= n_lwe: the body slot, always present
      int tt = t;
      asm volatile("" : "+v"(tt));
      const int rr = tt & 63, mm0 = tt >> 6;
      __syncthreads();
#pragma unroll
      for (int m1 = 0; m1 < 8; m1++) lds[h * N + 64 * (mm0 + 4 * m1) + rr] = acc[m1];
      __syncthreads();
      uint64_t v[L][8];
      pbs::rotate_decompose<L, HI>(lds + h * N, acc, tt, rot, a.base_log, v);
      __syncthreads();
      fwd_passes<L, HI>(v, lds + h * L * N, tab, tt);
      // exchange the digit NTTs: row gl = g * L + l at [(gl * 8 + w0) * T + t]
      __syncthreads();
#pragma unroll
      for (int l = 0; l < L; l++)
#pragma unroll
        for (int w0 = 0; w0 < 8; w0++) lds[((h * L + l) * 8 + w0) * T + tt] = v[l][w0];
      __syncthreads();
      uint64_t z[1][8];
      // outputs w0 in groups of four: the group's key words first, then the 128-bit sums
      const int hu = wave_uniform(h);
      const int slice_off = (int)((size_t)i * bsk_stride * 8);
      const int lane_off = tt * 8;
#pragma unroll
      for (int q = 0; q < 8; q += 4) {
        uint64_t kw[4][NPF];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
          for (int gl_ = 0; gl_ < NPF; gl_++) {
            const int e = ((q + u) * NPF + gl_) * K1 + hu;
            const auto w = __builtin_amdgcn_raw_buffer_load_b64(bsk_rsrc, lane_off, slice_off + e * T * 8, 0);
            kw[u][gl_] = ((uint64_t)w[1] << 32) | w[0];
          }
        uint64_t lo[4], hi[4], o[4];
        uint32_t cc[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          cc[u] = 0;
          gl::mul_full(lds[(q + u) * T + tt], kw[u][0], lo[u], hi[u]);
        }
#pragma unroll
        for (int gl_ = 1; gl_ < NPF; gl_++) {
          uint64_t plo[4], phi[4];
#pragma unroll
          for (int u = 0; u < 4; u++) gl::mul_full(lds[(gl_ * 8 + q + u) * T + tt], kw[u][gl_], plo[u], phi[u]);
          gl::acc128x4(lo, hi, cc, plo, phi);
        }
        gl::reduce160x4(o, lo, hi, cc);
#pragma unroll
        for (int u = 0; u < 4; u++) z[0][q + u] = o[u];
      }
      inv_passes<1>(z, lds + h * N, tab, tt);
#pragma unroll
      for (int m1 = 0; m1 < 8; m1++) acc[m1] += gl::modswitch_p_to_q(z[0][m1]);
    }
    // sample extract
    __syncthreads();
#pragma unroll
    for (int m1 = 0; m1 < 8; m1++) lds[h * N + 64 * (m0 + 4 * m1) + r] = acc[m1];
    __syncthreads();
    const uint32_t b = a2n[a.n_lwe];
    const int lut_nb = 1 << job.log_lut_nb;
    for (int lu = 0; lu < lut_nb; lu++) {
      const uint32_t offs = (uint32_t)(lu * (N >> job.log_lut_nb));
      uint64_t* o = reinterpret_cast<uint64_t*>(job.out_addr) + (size_t)lu * blen;
#pragma unroll
      for (int g = 0; g < K1 - 1; g++)
#pragma unroll
        for (int e = 0; e < N / (2 * T); e++) {
          int m = t2 + 2 * T * e;
          o[g * N + m] = pbs::extract_coef(lds, g, m, b, offs);
        }
      if (t2 == 0) o[(K1 - 1) * N] = pbs::extract_coef(lds, K1 - 1, 0, b, offs);
    }
  }
}

// Occupancy figures are computed once per process (function-local statics: thread-safe
// initialisation, as an attached server's thread may launch at the same time as the host).
template <int K1, int L, bool TP>
static int br_lat_max_grid() {
  static const int cached = [] {
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_blind_rotate_lat<K1, L, true, TP>, 512, 0);
    return (cus > 0 ? cus : 256) * (per > 0 ? per : 1);
  }();
  return cached;
}

int blind_rotate_lat_max_grid(int pbs_level, bool pair) {
  if (pbs_level == 1) return pair ? br_lat_max_grid<2, 1, true>() : br_lat_max_grid<2, 1, false>();
  return pair ? br_lat_max_grid<2, 2, true>() : br_lat_max_grid<2, 2, false>();
}

template <int L, bool HI, bool TP>
static void lat_launch(int grid_override, const uint64_t* bsk, const uint64_t* luts, const uint16_t* lwe2n,
                       const PbsJob* jobs, int nct, const uint64_t* tab, const BrArgs& a, hipStream_t s) {
  int g = grid_override > 0 ? grid_override : br_lat_max_grid<2, L, TP>();
  g = g < nct ? g : nct;
  k_blind_rotate_lat<2, L, HI, TP><<<dim3(g), 512, 0, s>>>(bsk, luts, lwe2n, jobs, nct, tab, a);
}

void launch_blind_rotate_lat(int pbs_level, bool hi, bool pair, int grid_override, const uint64_t* bsk,
                             const uint64_t* luts, const uint16_t* lwe2n, const PbsJob* jobs, int nct,
                             const uint64_t* tab, const BrArgs& a, hipStream_t s) {
#define LAT_CASE(L_, HI_, TP_)                                                  \
  if (pbs_level == L_ && hi == HI_ && pair == TP_)                              \
    return lat_launch<L_, HI_, TP_>(grid_override, bsk, luts, lwe2n, jobs, nct, tab, a, s);
  LAT_CASE(1, true, true)
  LAT_CASE(1, true, false)
  LAT_CASE(1, false, true)
  LAT_CASE(1, false, false)
  LAT_CASE(2, true, true)
  LAT_CASE(2, true, false)
  LAT_CASE(2, false, true)
  LAT_CASE(2, false, false)
#undef LAT_CASE
}

}  // namespace lk
