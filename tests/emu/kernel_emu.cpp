// Sequential CPU execution of the HIP kernels' per-thread code (test harness only).
// Each pass runs its load/compute for all 256 "threads", then all stores, which is
// what the barriers in the kernels guarantee; the arithmetic is the device code itself.
#include <stdint.h>
#include <string.h>
#include <vector>

#include "../../luthe_amd/csrc/ntt2048.h"

using namespace ntt;

namespace {
struct Tables {
  uint64_t tw[2048], itw[2048];
  Tables() { make_twist_tables(tw, itw); }
};
const Tables& tables() {
  static Tables t;
  return t;
}
}  // namespace

extern "C" {

// Forward NTT of one polynomial: out in NTT-domain order (index 64k + w).
void emu_ntt_fwd(const uint64_t* in, uint64_t* out) {
  const Tables& tb = tables();
  std::vector<uint64_t> lds(N);
  uint64_t v[T][8];
  for (int t = 0; t < T; t++) {
    int r = t & 63, m0 = t >> 6;
    for (int m1 = 0; m1 < 8; m1++) v[t][m1] = in[64 * (m0 + 4 * m1) + r];
    fwd_a1_compute(v[t], m0);
  }
  for (int t = 0; t < T; t++) fwd_a1_store(v[t], lds.data(), t & 63, t >> 6);
  for (int t = 0; t < T; t++) {
    fwd_a2_load(v[t], lds.data(), t & 63, t >> 6);
    fwd_a2_compute(v[t], t & 63, t >> 6, tb.tw);
  }
  for (int t = 0; t < T; t++) fwd_a2_store(v[t], lds.data(), t & 63, t >> 6);
  for (int t = 0; t < T; t++) {
    fwd_c1_load(v[t], lds.data(), t >> 3, t & 7);
    fwd_c1_compute(v[t], t & 7);
  }
  for (int t = 0; t < T; t++) fwd_c1_store(v[t], lds.data(), t >> 3, t & 7);
  for (int t = 0; t < T; t++) {
    fwd_c2_load(v[t], lds.data(), t >> 3, t & 7);
    fwd_c2_compute(v[t]);
    int k = t >> 3, w1 = t & 7;
    for (int w0 = 0; w0 < 8; w0++) out[64 * k + w1 + 8 * w0] = gl::canonical(v[t][w0]);
  }
}

// Inverse NTT of one polynomial: in NTT-domain order, out canonical coefficients.
void emu_ntt_inv(const uint64_t* in, uint64_t* out) {
  const Tables& tb = tables();
  std::vector<uint64_t> lds(N);
  uint64_t v[T][8];
  for (int t = 0; t < T; t++) {
    int k = t >> 3, w1 = t & 7;
    for (int w0 = 0; w0 < 8; w0++) v[t][w0] = in[64 * k + w1 + 8 * w0];
    inv_c2_compute(v[t], w1);
  }
  for (int t = 0; t < T; t++) inv_c2_store(v[t], lds.data(), t >> 3, t & 7);
  for (int t = 0; t < T; t++) {
    inv_c1_load(v[t], lds.data(), t >> 3, t & 7);
    inv_c1_compute(v[t], t >> 3, t & 7, tb.itw);
  }
  for (int t = 0; t < T; t++) inv_c1_store(v[t], lds.data(), t >> 3, t & 7);
  for (int t = 0; t < T; t++) {
    inv_a2_load(v[t], lds.data(), t & 63, t >> 6);
    inv_a2_compute(v[t], t >> 6);
  }
  for (int t = 0; t < T; t++) inv_a2_store(v[t], lds.data(), t & 63, t >> 6);
  for (int t = 0; t < T; t++) {
    int r = t & 63, m0 = t >> 6;
    inv_a1_load(v[t], lds.data(), r, m0);
    inv_a1_compute(v[t], m0);
    for (int m1 = 0; m1 < 8; m1++) out[64 * (m0 + 4 * m1) + r] = v[t][m1];
  }
}

uint64_t emu_gl_mul(uint64_t a, uint64_t b) { return gl::mul(a, b); }
uint64_t emu_gl_add(uint64_t a, uint64_t b) { return gl::add(a, b); }
uint64_t emu_gl_sub(uint64_t a, uint64_t b) { return gl::sub(a, b); }
uint64_t emu_gl_mul_pow2_rt(uint64_t a, uint32_t s) { return gl::mul_pow2_rt(a, s); }
uint64_t emu_gl_mul_pow2(uint64_t a, uint32_t s) { return gl::mul_pow2(a, s); }
uint64_t emu_gl_reduce128(uint64_t lo, uint64_t hi) { return gl::reduce128(lo, hi); }
uint64_t emu_modswitch(uint64_t e) { return gl::modswitch_p_to_q(e); }
}
