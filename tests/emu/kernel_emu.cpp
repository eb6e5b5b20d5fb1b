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
  uint64_t tab[TABLE_WORDS];
  const uint64_t* tw = tab + TW_OFF;
  const uint64_t* itw = tab + ITW_OFF;
  const uint64_t* c64 = tab + C64_OFF;
  const uint64_t* ic64 = tab + IC64_OFF;
  Tables() { make_twist_tables(tab); }
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
    fwd_c1_load<1>(reinterpret_cast<uint64_t(*)[8]>(v[t]), lds.data(), t);
    fwd_c1_compute<1>(reinterpret_cast<uint64_t(*)[8]>(v[t]), t, tb.c64);
  }
  for (int t = 0; t < T; t++) fwd_c1_store<1>(reinterpret_cast<uint64_t(*)[8]>(v[t]), lds.data(), t);
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
    inv_c2_compute(v[t]);
  }
  for (int t = 0; t < T; t++) inv_c2_store(v[t], lds.data(), t >> 3, t & 7);
  for (int t = 0; t < T; t++) {
    inv_c1_load<1>(reinterpret_cast<uint64_t(*)[8]>(v[t]), lds.data(), t);
    inv_c1_compute<1>(reinterpret_cast<uint64_t(*)[8]>(v[t]), t, tb.itw, tb.ic64);
  }
  for (int t = 0; t < T; t++) inv_c1_store<1>(reinterpret_cast<uint64_t(*)[8]>(v[t]), lds.data(), t);
  for (int t = 0; t < T; t++) {
    inv_a2_load(v[t], lds.data(), t & 63, t >> 6);
    inv_a2_compute(v[t], t >> 6);
  }
  for (int t = 0; t < T; t++) inv_a2_store(v[t], lds.data(), t & 63, t >> 6);
  for (int t = 0; t < T; t++) {
    int r = t & 63, m0 = t >> 6;
    inv_a1_load(v[t], lds.data(), r, m0);
    inv_a1_compute(v[t]);
    for (int m1 = 0; m1 < 8; m1++) out[64 * (m0 + 4 * m1) + r] = v[t][m1];
  }
}

uint64_t emu_gl_mul(uint64_t a, uint64_t b) { return gl::mul(a, b); }
uint64_t emu_gl_add(uint64_t a, uint64_t b) { return gl::add(a, b); }
uint64_t emu_gl_add_wrap(uint64_t a, uint64_t b) { return gl::add_wrap(a, b); }
uint64_t emu_gl_canonical_fold(uint64_t x) { return gl::canonical_fold(x); }
uint64_t emu_gl_sub(uint64_t a, uint64_t b) { return gl::sub(a, b); }
uint64_t emu_gl_mul_pow2(uint64_t a, uint32_t s) { return gl::mul_pow2(a, s); }
// every compile-time shift: out[i] = x * 2^(i - 192) for i < 384 (negative amounts included)
void emu_gl_mul_pow2c_all(uint64_t x, uint64_t* out);
uint64_t emu_gl_reduce128(uint64_t lo, uint64_t hi) { return gl::reduce128(lo, hi); }
uint64_t emu_gl_reduce160(uint64_t lo, uint64_t hi, uint32_t c) { return gl::reduce160(lo, hi, c); }
uint64_t emu_modswitch(uint64_t e) { return gl::modswitch_p_to_q(e); }
}

#include "../../luthe_amd/csrc/pbs_core.h"

namespace {
// Internal BSK layout of the blind-rotation kernel: [i][w0][K1*L][K1][T].
void bsk_to_internal(int n, int K1, int L, const uint64_t* bsk_coef, uint64_t* out) {
  std::vector<uint64_t> z(N);
  for (int i = 0; i < n; i++)
    for (int gl_ = 0; gl_ < K1 * L; gl_++)
      for (int j = 0; j < K1; j++) {
        const uint64_t* src = bsk_coef + (((size_t)i * K1 * L + gl_) * K1 + j) * N;
        emu_ntt_fwd(src, z.data());
        for (int t = 0; t < T; t++) {
          int k = t >> 3, w1 = t & 7;
          for (int w0 = 0; w0 < 8; w0++)
            out[(((size_t)i * 8 + w0) * (K1 * L) + gl_) * K1 * T + (size_t)j * T + t] = z[64 * k + w1 + 8 * w0];
        }
      }
}

template <int K1, int L>
void blind_rotate(int n, const uint64_t* bsk_int, const uint64_t* lut, const uint16_t* a2n, uint32_t base_log,
                  uint64_t* acc_out) {
  const bool hi = L * base_log <= 31 && base_log <= 23;  // the kernels' HI variant
  constexpr int NPF = K1 * L;
  const Tables& tb = tables();
  std::vector<uint64_t> lds(NPF * N);
  static uint64_t acc[T][K1][8];
  static uint64_t v[T][NPF][8];
  for (int t = 0; t < T; t++)
    for (int g = 0; g < K1; g++)
      for (int m1 = 0; m1 < 8; m1++) acc[t][g][m1] = lut[g * N + pbs::coef_index(t, m1)];
  for (int i = 0; i < n; i++) {
    uint32_t a = a2n[i];
    for (int t = 0; t < T; t++)
      for (int g = 0; g < K1; g++)
        for (int m1 = 0; m1 < 8; m1++) lds[g * N + pbs::coef_index(t, m1)] = acc[t][g][m1];
    for (int t = 0; t < T; t++)
      for (int g = 0; g < K1; g++) {
        if (hi)
          pbs::rotate_decompose<L, true>(lds.data() + g * N, acc[t][g], t, a, base_log, &v[t][g * L]);
        else
          pbs::rotate_decompose<L, false>(lds.data() + g * N, acc[t][g], t, a, base_log, &v[t][g * L]);
      }
    for (int t = 0; t < T; t++)
      for (int p = 0; p < NPF; p++) fwd_a1_compute(v[t][p], t >> 6, hi);
    for (int t = 0; t < T; t++)
      for (int p = 0; p < NPF; p++) fwd_a1_store(v[t][p], lds.data() + p * N, t & 63, t >> 6);
    for (int t = 0; t < T; t++)
      for (int p = 0; p < NPF; p++) {
        fwd_a2_load(v[t][p], lds.data() + p * N, t & 63, t >> 6);
        fwd_a2_compute(v[t][p], t & 63, t >> 6, tb.tw);
      }
    for (int t = 0; t < T; t++)
      for (int p = 0; p < NPF; p++) fwd_a2_store(v[t][p], lds.data() + p * N, t & 63, t >> 6);
    for (int t = 0; t < T; t++) {
      fwd_c1_load<NPF>(v[t], lds.data(), t);
      fwd_c1_compute<NPF>(v[t], t, tb.c64);
    }
    for (int t = 0; t < T; t++) fwd_c1_store<NPF>(v[t], lds.data(), t);
    const uint64_t* bsk_i = bsk_int + (size_t)i * 8 * NPF * K1 * T;
    for (int t = 0; t < T; t++) {
      for (int p = 0; p < NPF; p++) {
        fwd_c2_load(v[t][p], lds.data() + p * N, t >> 3, t & 7);
        fwd_c2_compute(v[t][p]);
      }
      pbs::mac<K1, L>(v[t], bsk_i, t);
      for (int j = 0; j < K1; j++) inv_c2_compute(v[t][j]);
    }
    for (int t = 0; t < T; t++)
      for (int j = 0; j < K1; j++) inv_c2_store(v[t][j], lds.data() + j * N, t >> 3, t & 7);
    for (int t = 0; t < T; t++) {
      inv_c1_load<K1>(v[t], lds.data(), t);
      inv_c1_compute<K1>(v[t], t, tb.itw, tb.ic64);
    }
    for (int t = 0; t < T; t++) inv_c1_store<K1>(v[t], lds.data(), t);
    for (int t = 0; t < T; t++)
      for (int j = 0; j < K1; j++) {
        inv_a2_load(v[t][j], lds.data() + j * N, t & 63, t >> 6);
        inv_a2_compute(v[t][j], t >> 6);
      }
    for (int t = 0; t < T; t++)
      for (int j = 0; j < K1; j++) inv_a2_store(v[t][j], lds.data() + j * N, t & 63, t >> 6);
    for (int t = 0; t < T; t++)
      for (int j = 0; j < K1; j++) {
        inv_a1_load(v[t][j], lds.data() + j * N, t & 63, t >> 6);
        inv_a1_compute(v[t][j]);
        for (int m1 = 0; m1 < 8; m1++) acc[t][j][m1] += gl::modswitch_p_to_q(v[t][j][m1]);
      }
  }
  for (int t = 0; t < T; t++)
    for (int g = 0; g < K1; g++)
      for (int m1 = 0; m1 < 8; m1++) acc_out[g * N + pbs::coef_index(t, m1)] = acc[t][g][m1];
}
}  // namespace

extern "C" {
void emu_bsk_to_internal(int n, int L, const uint64_t* bsk_coef, uint64_t* out) {
  bsk_to_internal(n, 2, L, bsk_coef, out);
}
int emu_blind_rotate(int n, int L, const uint64_t* bsk_int, const uint64_t* lut, const uint16_t* a2n,
                     uint32_t base_log, uint64_t* acc_out) {
  if (L == 1)
    blind_rotate<2, 1>(n, bsk_int, lut, a2n, base_log, acc_out);
  else if (L == 2)
    blind_rotate<2, 2>(n, bsk_int, lut, a2n, base_log, acc_out);
  else
    return -1;
  return 0;
}
void emu_decompose(uint64_t x, uint32_t b, int L, int64_t* d) {
  switch (L) {
    case 1: pbs::decompose<1>(x, b, d); break;
    case 2: pbs::decompose<2>(x, b, d); break;
    case 3: pbs::decompose<3>(x, b, d); break;
    case 4: pbs::decompose<4>(x, b, d); break;
    case 5: pbs::decompose<5>(x, b, d); break;
    default: break;
  }
}
// decompose_hi for L * b <= 31 (digits as int32); returns -1 when not applicable
int emu_decompose_hi(uint64_t x, uint32_t b, int L, int32_t* d) {
  if ((uint32_t)L * b > 31) return -1;
  const uint32_t xh = (uint32_t)(x >> 32);
  switch (L) {
    case 1: pbs::decompose_hi<1>(xh, b, d); break;
    case 2: pbs::decompose_hi<2>(xh, b, d); break;
    case 3: pbs::decompose_hi<3>(xh, b, d); break;
    case 4: pbs::decompose_hi<4>(xh, b, d); break;
    case 5: pbs::decompose_hi<5>(xh, b, d); break;
    default: return -1;
  }
  return 0;
}
uint64_t emu_from_signed32(int32_t d) { return gl::from_signed32(d); }
// decompose_hi<L> against decompose<L> for every top-bit / round-bit pattern of an R = L*b
// <= 22 bit closest value (bits below the round bit zero, one, or all ones); returns the
// number of mismatching inputs, or -1 when (L, b) is out of range.
long emu_decompose_hi_mismatches(int L, uint32_t b) {
  const uint32_t R = (uint32_t)L * b;
  if (L < 1 || L > 5 || b < 1 || R > 22) return -1;
  const int ofs = 64 - (int)R;
  long bad = 0;
  int64_t d64[5];
  int32_t d32[5];
  for (uint64_t top = 0; top < (1ull << R); top++)
    for (uint64_t rb = 0; rb < 2; rb++)
      for (int lowk = 0; lowk < 3; lowk++) {
        const uint64_t low = lowk == 0 ? 0 : lowk == 1 ? 1 : (1ull << (ofs - 1)) - 1;
        const uint64_t x = (top << ofs) | (rb << (ofs - 1)) | low;
        emu_decompose(x, b, L, d64);
        emu_decompose_hi(x, b, L, d32);
        for (int i = 0; i < L; i++) bad += d64[i] != (int64_t)d32[i];
      }
  return bad;
}
// pretwist_digit<m1>(d) against mul_pow2c<12 m1>(from_signed32(d)) for every |d| <= 2^22 and
// every m1; returns the number of mismatches.
long emu_pretwist_mismatches();
uint64_t emu_extract_coef(const uint64_t* acc, int g, int m, uint32_t b, uint32_t offs) {
  return pbs::extract_coef(acc, g, m, b, offs);
}
}

namespace {
template <int S>
void pow2c_all(uint64_t x, uint64_t* out) {
  if constexpr (S < 192) {
    out[S + 192] = gl::mul_pow2c<S>(x);
    pow2c_all<S + 1>(x, out);
  }
}
// the batched form the DFT networks use, one lane per amount
template <int S>
void pow2c_n_all(uint64_t x, uint64_t* out) {
  if constexpr (S < 192) {
    out[S + 192] = x;
    gl::mul_pow2c_n<S>(out + S + 192);
    pow2c_n_all<S + 1>(x, out);
  }
}
}  // namespace

extern "C" void emu_gl_mul_pow2c_all(uint64_t x, uint64_t* out) { pow2c_all<-192>(x, out); }
extern "C" void emu_gl_mul_pow2c_n_all(uint64_t x, uint64_t* out) { pow2c_n_all<-192>(x, out); }

namespace {
template <int M1>
long pretwist_check() {
  long bad = 0;
  for (int32_t d = -(1 << 22); d <= (1 << 22); d++)
    bad += gl::pretwist_digit<M1>(d) != gl::mul_pow2c<12 * M1>(gl::from_signed32(d));
  return bad;
}
}  // namespace

extern "C" long emu_pretwist_mismatches() {
  return pretwist_check<0>() + pretwist_check<1>() + pretwist_check<2>() + pretwist_check<3>() +
         pretwist_check<4>() + pretwist_check<5>() + pretwist_check<6>() + pretwist_check<7>();
}

// ---------------------------------------------------------------- the 512-thread NTT
#include "../../luthe_amd/csrc/ntt2048x4.h"

namespace {
struct Tw4All {
  std::vector<ntt4::Tw4> tw;
  Tw4All() : tw(ntt4::T4) {
    for (int t = 0; t < ntt4::T4; t++) tw[t].load(tables().tab, t);
  }
};
const Tw4All& tw4() {
  static Tw4All t;
  return t;
}
}  // namespace

extern "C" {

// Forward NTT of one polynomial through the 512-thread passes: in coefficient order, out in
// NTT-domain order (index 64k + w), as emu_ntt_fwd.  pretwisted: the input already carries
// psi^(4 m1) (the digit lift's form).
void emu_ntt4_fwd(const uint64_t* in, uint64_t* out, int pretwisted) {
  const Tw4All& tw = tw4();
  std::vector<uint64_t> lds(ntt::N);
  static uint64_t v[ntt4::T4][4];
  for (int t = 0; t < ntt4::T4; t++) {
    const ntt4::Map m(t);
    const int r = m.r(), m0 = m.m0(), g = m.hi2();
    const int m1s[4] = {2 * g, 2 * g + 1, 2 * g + 4, 2 * g + 5};
    for (int i = 0; i < 4; i++) {
      uint64_t x = in[64 * (m0 + 4 * m1s[i]) + r];
      v[t][i] = pretwisted ? gl::mul_pow2(x, 12 * m1s[i]) : x;
    }
    if (pretwisted)
      ntt4::fwd_a1a_compute<true>(v[t], g);
    else
      ntt4::fwd_a1a_compute<false>(v[t], g);
  }
  for (int t = 0; t < ntt4::T4; t++) ntt4::fwd_a1a_store(v[t], lds.data(), ntt4::Map(t));
  for (int t = 0; t < ntt4::T4; t++) {
    ntt4::fwd_a1b_load(v[t], lds.data(), ntt4::Map(t));
    ntt4::fwd_a1b_compute(v[t], ntt4::Map(t).W);
  }
  for (int t = 0; t < ntt4::T4; t++) ntt4::fwd_a1b_store(v[t], lds.data(), ntt4::Map(t));
  for (int t = 0; t < ntt4::T4; t++) {
    ntt4::fwd_a2_load(v[t], lds.data(), ntt4::Map(t));
    ntt4::fwd_a2_compute(v[t], tw.tw[t].tw);
  }
  for (int t = 0; t < ntt4::T4; t++) ntt4::fwd_a2_store(v[t], lds.data(), ntt4::Map(t));
  for (int t = 0; t < ntt4::T4; t++) {
    ntt4::fwd_c1_load(v[t], lds.data(), ntt4::Map(t));
    ntt4::fwd_c1_compute(v[t], tw.tw[t].c1);
  }
  for (int t = 0; t < ntt4::T4; t++) ntt4::fwd_c1_store(v[t], lds.data(), ntt4::Map(t));
  for (int t = 0; t < ntt4::T4; t++) {
    ntt4::fwd_c2_load(v[t], lds.data(), ntt4::Map(t));
    ntt4::fwd_c2_compute(v[t], ntt4::Map(t).ra());
  }
  for (int t = 0; t < ntt4::T4; t++) ntt4::fwd_c2_store(v[t], lds.data(), ntt4::Map(t));
  for (int t = 0; t < ntt4::T4; t++) {
    const ntt4::Map m(t);
    ntt4::fwd_c3_load(v[t], lds.data(), m);
    ntt4::fwd_c3_compute(v[t]);
    for (int j = 0; j < 4; j++) out[64 * m.k3() + m.w1() + 8 * (m.wd() + 2 * j)] = gl::canonical(v[t][j]);
  }
}

// Inverse: NTT-domain order in, canonical coefficients out, as emu_ntt_inv.
void emu_ntt4_inv(const uint64_t* in, uint64_t* out) {
  const Tw4All& tw = tw4();
  std::vector<uint64_t> lds(ntt::N);
  static uint64_t v[ntt4::T4][4];
  for (int t = 0; t < ntt4::T4; t++) {
    const ntt4::Map m(t);
    for (int j = 0; j < 4; j++) v[t][j] = in[64 * m.k3() + m.w1() + 8 * (m.wd() + 2 * j)];
    ntt4::inv_c3_compute(v[t]);
  }
  for (int t = 0; t < ntt4::T4; t++) ntt4::inv_c3_store(v[t], lds.data(), ntt4::Map(t));
  for (int t = 0; t < ntt4::T4; t++) {
    ntt4::inv_c2_load(v[t], lds.data(), ntt4::Map(t));
    ntt4::inv_c2_compute(v[t], ntt4::Map(t).ra());
  }
  for (int t = 0; t < ntt4::T4; t++) ntt4::inv_c2_store(v[t], lds.data(), ntt4::Map(t));
  for (int t = 0; t < ntt4::T4; t++) {
    ntt4::inv_c1_load(v[t], lds.data(), ntt4::Map(t));
    ntt4::inv_c1_compute(v[t], tw.tw[t].ic1, tw.tw[t].itw);
  }
  for (int t = 0; t < ntt4::T4; t++) ntt4::inv_c1_store(v[t], lds.data(), ntt4::Map(t));
  for (int t = 0; t < ntt4::T4; t++) {
    ntt4::inv_a2_load(v[t], lds.data(), ntt4::Map(t));
    ntt4::inv_a2_compute(v[t], ntt4::Map(t).W);
  }
  for (int t = 0; t < ntt4::T4; t++) ntt4::inv_a2_store(v[t], lds.data(), ntt4::Map(t));
  for (int t = 0; t < ntt4::T4; t++) {
    ntt4::inv_a1b_load(v[t], lds.data(), ntt4::Map(t));
    ntt4::inv_a1b_compute(v[t]);
  }
  for (int t = 0; t < ntt4::T4; t++) ntt4::inv_a1b_store(v[t], lds.data(), ntt4::Map(t));
  for (int t = 0; t < ntt4::T4; t++) {
    const ntt4::Map m(t);
    ntt4::inv_a1a_load(v[t], lds.data(), m);
    ntt4::inv_a1a_compute(v[t], m.hi2());
    const int r = m.r(), m0 = m.m0(), g = m.hi2();
    const int m1s[4] = {2 * g, 2 * g + 1, 2 * g + 4, 2 * g + 5};
    for (int i = 0; i < 4; i++) out[64 * (m0 + 4 * m1s[i]) + r] = v[t][i];
  }
}
}

// ---------------------------------------------------------------- the 512-thread blind rotation
// The split kernel's per-thread CMUX (k_blind_rotate_split): both GLWE polynomials of one
// ciphertext, each on 512 "threads" with four values, the key read in the 256-thread internal
// layout (thread t & 255, rows w0 = (t >> 8) + 2j), the partner's digits taken directly.
namespace {
template <int L>
void blind_rotate_split(int n, const uint64_t* bsk_int, const uint64_t* lut, const uint16_t* a2n, uint32_t base_log,
                   uint64_t* acc_out) {
  constexpr int K1 = 2, NPF = K1 * L, T4 = ntt4::T4;
  const bool hi = L * base_log <= 31 && base_log <= 23;
  const Tw4All& tw = tw4();
  std::vector<uint64_t> lds(ntt::N);
  static uint64_t acc[K1][T4][4], v[NPF][T4][4], z[K1][T4][4];
  auto coef = [](int t, int i) {
    const int g = t >> 8;
    return 64 * (((t >> 6) & 3) + 4 * (2 * g + (i & 1) + 4 * (i >> 1))) + (t & 63);
  };
  for (int p = 0; p < K1; p++)
    for (int t = 0; t < T4; t++)
      for (int i = 0; i < 4; i++) acc[p][t][i] = lut[p * ntt::N + coef(t, i)];
  for (int it = 0; it < n; it++) {
    const uint32_t a = a2n[it];
    for (int p = 0; p < K1; p++) {
      for (int t = 0; t < T4; t++)
        for (int i = 0; i < 4; i++) lds[coef(t, i)] = acc[p][t][i];
      for (int t = 0; t < T4; t++) {
        uint64_t d[L][4];
        if (hi)
          pbs::rotate_decompose4<L, true>(lds.data(), acc[p][t], t, a, base_log, d);
        else
          pbs::rotate_decompose4<L, false>(lds.data(), acc[p][t], t, a, base_log, d);
        for (int l = 0; l < L; l++)
          for (int i = 0; i < 4; i++) v[p * L + l][t][i] = d[l][i];
      }
    }
    // forward: one digit polynomial at a time through the six passes
    for (int q = 0; q < NPF; q++) {
      for (int t = 0; t < T4; t++) {
        if (hi)
          ntt4::fwd_a1a_compute<true>(v[q][t], t >> 8);
        else
          ntt4::fwd_a1a_compute<false>(v[q][t], t >> 8);
      }
      for (int t = 0; t < T4; t++) ntt4::fwd_a1a_store(v[q][t], lds.data(), ntt4::Map(t));
      for (int t = 0; t < T4; t++) {
        ntt4::fwd_a1b_load(v[q][t], lds.data(), ntt4::Map(t));
        ntt4::fwd_a1b_compute(v[q][t], t >> 6);
      }
      for (int t = 0; t < T4; t++) ntt4::fwd_a1b_store(v[q][t], lds.data(), ntt4::Map(t));
      for (int t = 0; t < T4; t++) {
        ntt4::fwd_a2_load(v[q][t], lds.data(), ntt4::Map(t));
        ntt4::fwd_a2_compute(v[q][t], tw.tw[t].tw);
      }
      for (int t = 0; t < T4; t++) ntt4::fwd_a2_store(v[q][t], lds.data(), ntt4::Map(t));
      for (int t = 0; t < T4; t++) {
        ntt4::fwd_c1_load(v[q][t], lds.data(), ntt4::Map(t));
        ntt4::fwd_c1_compute(v[q][t], tw.tw[t].c1);
      }
      for (int t = 0; t < T4; t++) ntt4::fwd_c1_store(v[q][t], lds.data(), ntt4::Map(t));
      for (int t = 0; t < T4; t++) {
        ntt4::fwd_c2_load(v[q][t], lds.data(), ntt4::Map(t));
        ntt4::fwd_c2_compute(v[q][t], (t >> 6) & 3);
      }
      for (int t = 0; t < T4; t++) ntt4::fwd_c2_store(v[q][t], lds.data(), ntt4::Map(t));
      for (int t = 0; t < T4; t++) {
        ntt4::fwd_c3_load(v[q][t], lds.data(), ntt4::Map(t));
        ntt4::fwd_c3_compute(v[q][t]);
      }
    }
    // external product: output column j of thread t, rows w0 = (t >> 8) + 2 i of key thread t & 255
    const uint64_t* bsk_i = bsk_int + (size_t)it * 8 * NPF * K1 * ntt::T;
    for (int j = 0; j < K1; j++)
      for (int t = 0; t < T4; t++)
        for (int i = 0; i < 4; i++) {
          uint64_t lo = 0, hi_ = 0;
          uint32_t c = 0;
          for (int q = 0; q < NPF; q++) {
            const int e = (((t >> 8) + 2 * i) * NPF + q) * K1 + j;
            uint64_t plo, phi;
            gl::mul_full(v[q][t][i], bsk_i[(size_t)e * ntt::T + (t & 255)], plo, phi);
            if (q == 0)
              lo = plo, hi_ = phi;
            else
              gl::acc128(lo, hi_, c, plo, phi);
          }
          z[j][t][i] = gl::reduce160(lo, hi_, c);
        }
    for (int j = 0; j < K1; j++) {
      for (int t = 0; t < T4; t++) {
        ntt4::inv_c3_compute(z[j][t]);
        ntt4::inv_c3_store(z[j][t], lds.data(), ntt4::Map(t));
      }
      for (int t = 0; t < T4; t++) {
        ntt4::inv_c2_load(z[j][t], lds.data(), ntt4::Map(t));
        ntt4::inv_c2_compute(z[j][t], (t >> 6) & 3);
      }
      for (int t = 0; t < T4; t++) ntt4::inv_c2_store(z[j][t], lds.data(), ntt4::Map(t));
      for (int t = 0; t < T4; t++) {
        ntt4::inv_c1_load(z[j][t], lds.data(), ntt4::Map(t));
        ntt4::inv_c1_compute(z[j][t], tw.tw[t].ic1, tw.tw[t].itw);
      }
      for (int t = 0; t < T4; t++) ntt4::inv_c1_store(z[j][t], lds.data(), ntt4::Map(t));
      for (int t = 0; t < T4; t++) {
        ntt4::inv_a2_load(z[j][t], lds.data(), ntt4::Map(t));
        ntt4::inv_a2_compute(z[j][t], t >> 6);
      }
      for (int t = 0; t < T4; t++) ntt4::inv_a2_store(z[j][t], lds.data(), ntt4::Map(t));
      for (int t = 0; t < T4; t++) {
        ntt4::inv_a1b_load(z[j][t], lds.data(), ntt4::Map(t));
        ntt4::inv_a1b_compute(z[j][t]);
      }
      for (int t = 0; t < T4; t++) ntt4::inv_a1b_store(z[j][t], lds.data(), ntt4::Map(t));
      for (int t = 0; t < T4; t++) {
        ntt4::inv_a1a_load(z[j][t], lds.data(), ntt4::Map(t));
        ntt4::inv_a1a_compute(z[j][t], t >> 8);
        for (int i = 0; i < 4; i++) acc[j][t][i] += gl::modswitch_p_to_q(z[j][t][i]);
      }
    }
  }
  for (int p = 0; p < K1; p++)
    for (int t = 0; t < T4; t++)
      for (int i = 0; i < 4; i++) acc_out[p * ntt::N + coef(t, i)] = acc[p][t][i];
}
}  // namespace

extern "C" int emu_blind_rotate_split(int n, int L, const uint64_t* bsk_int, const uint64_t* lut, const uint16_t* a2n,
                                 uint32_t base_log, uint64_t* acc_out) {
  if (L == 1)
    blind_rotate_split<1>(n, bsk_int, lut, a2n, base_log, acc_out);
  else if (L == 2)
    blind_rotate_split<2>(n, bsk_int, lut, a2n, base_log, acc_out);
  else
    return -1;
  return 0;
}
