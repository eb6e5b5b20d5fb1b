// Test library: runs the device builds of the Goldilocks primitives (gl64.h) and the
// high-word decomposition on caller-given operands, so that the device code paths
// (inline-asm carry chains included) are checked against exact integer arithmetic at edge
// values.  Test infrastructure only.  The four-lane reductions and batched shifts below run
// with grouped reduction tails (several lanes per asm block), as in the throughput kernels;
// the single-lane forms (o[4] - o[6]) cover the per-lane blocks.
#define GL_GROUPED_TAILS 1
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../luthe_amd/csrc/pbs_core.h"

namespace {
constexpr int NOUT = 67;  // results per operand triple, in the order listed in prims_run

__global__ void k_prims(const uint64_t* a, const uint64_t* b, const uint32_t* c, int n, uint64_t* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t x = a[i], y = b[i];
  const uint32_t z = c[i];
  uint64_t* o = out + (size_t)i * NOUT;
  const uint64_t xc = gl::canonical(x), yc = gl::canonical(y);
  o[0] = gl::canonical(x);
  o[1] = gl::add(xc, yc);
  o[2] = gl::sub(xc, yc);
  o[3] = gl::mul(x, y);
  o[4] = gl::reduce128(x, y);
  o[5] = gl::reduce96(x, y & gl::EPS);
  o[6] = gl::reduce160(x, y, z == 0xFFFFFFFFu ? 0 : z);
  o[7] = gl::sub_fold(x, yc);
  o[8] = gl::mul_pow2c<24>(xc);
  o[9] = gl::mul_pow2c<48>(xc);
  o[10] = gl::mul_pow2c<72>(xc);
  o[11] = gl::mul_pow2c<120>(xc);
  o[12] = gl::mul_pow2c<168>(xc);
  o[13] = gl::modswitch_p_to_q(xc);
  o[14] = gl::from_signed32((int32_t)z);
  int32_t d[1];
  pbs::decompose_hi<1>((uint32_t)(x >> 32), 23, d);
  o[15] = (uint64_t)(int64_t)d[0];
  uint64_t lo = x, hi = y;
  uint32_t cc = z & 3;
  gl::acc128(lo, hi, cc, y, x);  // (x + y 2^64 + cc 2^128) + (y + x 2^64)
  o[16] = lo;
  o[17] = hi;
  o[18] = cc;
  // four interleaved subtractions with operands crossed so that every lane differs
  gl::sub4(o[19], o[20], o[21], o[22], xc, yc, xc, gl::canonical(x ^ y), yc, xc, gl::canonical(~x), 0);
  // four-lane products and 160-bit reductions
  const uint64_t ma[4] = {x, y, x ^ y, ~x}, mb[4] = {y, x, y, x};
  gl::mul4(o + 23, ma, mb);
  const uint32_t zc = z == 0xFFFFFFFFu ? 0 : z;
  const uint64_t rl[4] = {x, y, x, y}, rh[4] = {y, x, x, ~y};
  const uint32_t rc[4] = {zc, zc >> 1, 0, 0xFFFFFFFEu};
  gl::reduce160x4(o + 27, rl, rh, rc);
  // signed lift into Z_p, full range and a shifted (small-magnitude) value
  o[31] = gl::lift64((int64_t)x);
  o[32] = gl::lift64((int64_t)x >> 17);
  // interleaved sub_folds (exact 64-bit results, operands any a and canonical b)
  gl::sub2(o[33], o[34], x, y, yc, xc);
  gl::sub3(o[35], o[36], o[37], x ^ y, ~x, y, yc, xc, gl::canonical(x ^ y));
  // four interleaved 128-bit accumulations
  uint64_t al[4] = {x, y, x ^ y, ~x}, ah[4] = {y, x, ~x, ~y};
  uint32_t ac[4] = {z & 3, (z >> 2) & 3, 1, 0};
  const uint64_t pl[4] = {y, x, ~y, ~x}, ph[4] = {x, y, x, ~y};
  gl::acc128x4(al, ah, ac, pl, ph);
  for (int i = 0; i < 4; i++) o[38 + 3 * i] = al[i], o[39 + 3 * i] = ah[i], o[40 + 3 * i] = ac[i];
  // batched compile-time shifts (every kind: e = 0, <= 32, < 64, < 96, each negated or not)
  uint64_t tv[8] = {xc, yc, xc, yc, gl::canonical(x ^ y), xc, yc, xc};
  gl::mul_pow2c_n<24, 48, 72, 120, 168, 0, 96, 36>(tv);
  for (int i = 0; i < 8; i++) o[50 + i] = tv[i];
  // interleaved reduction tails t0 + hl * EPS over two, three and four lanes
  const uint64_t et[4] = {x, y, x ^ y, ~x};
  const uint32_t eh[4] = {z, z >> 1, (uint32_t)y, 0xFFFFFFFFu};
  gl::add_eps_prodn<2>(o + 58, et, eh);
  gl::add_eps_prodn<3>(o + 60, et + 1, eh + 1);
  gl::add_eps_prodn<4>(o + 63, et, eh);
}
}  // namespace

// a, b, c: host arrays of n operands; out: host array of n * NOUT results.  Returns 0 or the
// HIP error code.
extern "C" int prims_run(const uint64_t* a, const uint64_t* b, const uint32_t* c, int n, uint64_t* out) {
  uint64_t *da = nullptr, *db = nullptr, *dout = nullptr;
  uint32_t* dc = nullptr;
  hipError_t e = hipSuccess;
  const size_t n8 = (size_t)n * 8;
  if ((e = hipMalloc(&da, n8)) != hipSuccess || (e = hipMalloc(&db, n8)) != hipSuccess ||
      (e = hipMalloc(&dc, (size_t)n * 4)) != hipSuccess || (e = hipMalloc(&dout, n8 * NOUT)) != hipSuccess)
    goto done;
  if ((e = hipMemcpy(da, a, n8, hipMemcpyHostToDevice)) != hipSuccess) goto done;
  if ((e = hipMemcpy(db, b, n8, hipMemcpyHostToDevice)) != hipSuccess) goto done;
  if ((e = hipMemcpy(dc, c, (size_t)n * 4, hipMemcpyHostToDevice)) != hipSuccess) goto done;
  k_prims<<<dim3((unsigned)((n + 255) / 256)), 256>>>(da, db, dc, n, dout);
  if ((e = hipGetLastError()) != hipSuccess) goto done;
  e = hipMemcpy(out, dout, n8 * NOUT, hipMemcpyDeviceToHost);
done:
  (void)hipFree(da);
  (void)hipFree(db);
  (void)hipFree(dc);
  (void)hipFree(dout);
  return (int)e;
}
