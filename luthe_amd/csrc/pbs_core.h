// Per-thread building blocks of the blind rotation, key switch and sample extract,
// shared by the HIP kernels and by the sequential CPU emulation used in tests.
//
// Thread mapping follows ntt2048.h: in the coefficient domain thread (r, m0) owns
// coefficients n = 64 (m0 + 4 m1) + r of every polynomial (m1 < 8); in the NTT domain
// thread t = 8k + w1 owns indices 64k + w1 + 8 w0 (w0 < 8).
#pragma once
#include "gl64.h"
#include "ntt2048.h"

namespace pbs {

constexpr int N = ntt::N;
constexpr int T = ntt::T;

GL_HD int coef_index(int t, int m1) { return 64 * ((t >> 6) + 4 * m1) + (t & 63); }

// Balanced signed decomposition of a w-bit word into L digits of b bits, LSB level
// first: closest representable value (top L*b bits + round bit), sign of the closest
// representable, then the carry chain of decomp_parallel.sv:56-150 (same algorithm as
// decomp_balseq_core.sv:10-173 and balanced_decomp_unfix_sign_round_bit.py:42-105).
// With v = u + c (raw digit plus incoming carry) and H = 2^(b-1), the chain's
// propagate = ov | gt | (ge & msb(next)) with gt = [v > H], ge = [v >= H], ov = [v = 2^b]
// is [v + msb(next) > H]; the top level's "next" msb is the sign.
// Requires L*b < 64.  Digits are returned as signed ints.
template <int L>
GL_HD void decompose(uint64_t x, uint32_t b, int64_t* d) {
  const uint32_t R = (uint32_t)L * b;
  const uint32_t ofs = 64 - R;
  const uint64_t bmask = (1ull << b) - 1;
  const uint64_t H = 1ull << (b - 1);
  uint64_t rb = (x >> (ofs - 1)) & 1;
  uint64_t top = x >> ofs;                                 // R bits
  uint64_t closest = (top + rb) & ((1ull << R) - 1);
  uint64_t msb = x >> 63;
  uint64_t below_mask = (1ull << (R - 1)) - 1;
  uint64_t below = top & below_mask;
  uint64_t sign = (msb & (uint64_t)((below != 0) | rb)) | ((msb ^ 1) & (uint64_t)(below == below_mask) & rb);
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    uint64_t v = ((closest >> (i * b)) & bmask) + c;
    uint64_t nmsb = (i + 1 < L) ? ((closest >> ((i + 1) * b + b - 1)) & 1) : sign;
    uint64_t prop = (v + nmsb) > H;
    d[i] = (int64_t)v - (int64_t)(prop << b);
    c = prop;
  }
}

// The same digits when L*b <= 31, from the high word xh = x >> 32 alone: the round bit
// (bit 63 - L*b) and every digit bit lie in it, and adding the round bit at bit 31 - L*b
// of xh gives closest in its top L*b bits.  For L = 1 the digit is closest read as a
// signed L*b-bit number, except closest = 2^(b-1) reached without rounding (top = 2^(b-1),
// rb = 0), whose sign is 0 and whose digit stays +2^(b-1).
//
// For L >= 2 the carry chain is one integer addition.  With y_i = chunk_i + c_i + nmsb_i +
// (H - 1), the chain's propagate [v_i + nmsb_i > H] is [y_i >= 2^b] and y_i < 2^(b+1), so
// the carries c_i are those of Y = closest + sum_i (H - 1 + nmsb_i) 2^(b i), and digit i is
// chunk i of Y minus (H - 1 + nmsb_i).  The sign (the top level's next msb) is the msb of x,
// flipped exactly when closest = 2^(R-1): for msb = 1 the rule's (below != 0) | rb is false
// only for top + rb = 2^(R-1); for msb = 0 its (below == all ones) & rb is true only there.
// Y < 2^(R+1) <= 2^32.  No branch, no compare chain.
template <int L>
GL_HD void decompose_hi(uint32_t xh, uint32_t b, int32_t* d) {
  const uint32_t R = (uint32_t)L * b;
  const uint32_t rounded = xh + (1u << (31 - R));
  if constexpr (L == 1) {
    const int32_t sx = (int32_t)rounded >> (32 - R);
    const bool tie = (xh >> (31 - R)) == (1u << R);
    d[0] = tie ? (int32_t)(1u << (R - 1)) : sx;
  } else {
    const uint32_t closest = rounded >> (32 - R);
    const uint32_t sign = (xh >> 31) ^ (uint32_t)(closest == (1u << (R - 1)));
    const uint32_t bmask = (1u << b) - 1, h1 = (1u << (b - 1)) - 1;
    uint32_t off[L];
    uint32_t y = closest;
#pragma unroll
    for (int i = 0; i < L; i++) {
      const uint32_t nmsb = (i + 1 < L) ? ((closest >> ((i + 2) * b - 1)) & 1) : sign;
      off[i] = h1 + nmsb;
      y += off[i] << (i * b);
    }
#pragma unroll
    for (int i = 0; i < L; i++) d[i] = (int32_t)((y >> (i * b)) & bmask) - (int32_t)off[i];
  }
}

// ct10 = X^a * ACC - ACC for the thread's coefficients of one polynomial, decomposed
// and lifted into Z_p: out[l][m1].  acc_lds is the polynomial in plain order.
// Rotation semantics of pep_mmacc_splitc_feed_read.sv:484-488 and feed_final.sv:203-330:
// coefficient n reads ACC[(n - a) mod 2N], negated when that index lands in [N, 2N).
// HI: L * base_log <= 31 and base_log <= 23 (decompose_hi on the high word; digits
// |d| <= 2^22, so the lift also applies the forward NTT's pre-twist psi^(4 m1) and the
// outputs are pretwisted, see fwd_a1_compute).
template <int L, bool HI, int M1 = 0>
GL_HD void rotate_decompose_from(const uint64_t* acc_lds, const uint64_t* acc_reg, int t, uint32_t a,
                                 uint32_t base_log, uint64_t (*out)[8]);
template <int L, bool HI>
GL_HD void rotate_decompose(const uint64_t* acc_lds, const uint64_t* acc_reg, int t, uint32_t a,
                            uint32_t base_log, uint64_t (*out)[8]) {
  rotate_decompose_from<L, HI, 0>(acc_lds, acc_reg, t, a, base_log, out);
}
// One coefficient per recursion step so that m1 is a compile-time constant (the pre-twist).
template <int L, bool HI, int M1>
GL_HD void rotate_decompose_from(const uint64_t* acc_lds, const uint64_t* acc_reg, int t, uint32_t a,
                                 uint32_t base_log, uint64_t (*out)[8]) {
  if constexpr (M1 < 8) {
    constexpr int m1 = M1;
    int n = coef_index(t, m1);
    uint32_t idx = ((uint32_t)n + 2 * N - a) & (2 * N - 1);
    uint64_t val = acc_lds[idx & (N - 1)];
    val = (idx >= (uint32_t)N) ? (0 - val) : val;
    uint64_t ct10 = val - acc_reg[m1];
    if constexpr (HI) {
      int32_t d[L];
      decompose_hi<L>((uint32_t)(ct10 >> 32), base_log, d);
#pragma unroll
      for (int l = 0; l < L; l++) out[l][m1] = gl::pretwist_digit<M1>(d[l]);
    } else {
      int64_t d[L];
      decompose<L>(ct10, base_log, d);
#pragma unroll
      for (int l = 0; l < L; l++) out[l][m1] = gl::from_signed(d[l]);
    }
    rotate_decompose_from<L, HI, M1 + 1>(acc_lds, acc_reg, t, a, base_log, out);
  }
}

// rotate_decompose for the 512-thread layout (ntt2048x4.h A1a map: thread (r, m0, g) owns
// coefficients 64 (m0 + 4 m1) + r for m1 = 2g, 2g+1, 2g+4, 2g+5, in that register order);
// g is uniform per wave, so the pre-twist exponents stay compile-time.
template <int L, bool HI, int G, int I = 0>
GL_HD void rotate_decompose4_from(const uint64_t* acc_lds, const uint64_t* acc_reg, int r, int m0, uint32_t a,
                                  uint32_t base_log, uint64_t (*out)[4]) {
  if constexpr (I < 4) {
    constexpr int m1 = 2 * G + (I & 1) + 4 * (I >> 1);
    const int n = 64 * (m0 + 4 * m1) + r;
    const uint32_t idx = ((uint32_t)n + 2 * N - a) & (2 * N - 1);
    uint64_t val = acc_lds[idx & (N - 1)];
    val = (idx >= (uint32_t)N) ? (0 - val) : val;
    const uint64_t ct10 = val - acc_reg[I];
    if constexpr (HI) {
      int32_t d[L];
      decompose_hi<L>((uint32_t)(ct10 >> 32), base_log, d);
#pragma unroll
      for (int l = 0; l < L; l++) out[l][I] = gl::pretwist_digit<m1>(d[l]);
    } else {
      int64_t d[L];
      decompose<L>(ct10, base_log, d);
#pragma unroll
      for (int l = 0; l < L; l++) out[l][I] = gl::from_signed(d[l]);
    }
    rotate_decompose4_from<L, HI, G, I + 1>(acc_lds, acc_reg, r, m0, a, base_log, out);
  }
}
template <int L, bool HI>
GL_HD void rotate_decompose4(const uint64_t* acc_lds, const uint64_t* acc_reg, int t, uint32_t a, uint32_t base_log,
                             uint64_t (*out)[4]) {
  const int r = t & 63, m0 = (t >> 6) & 3;
  if (ntt::wave_uniform(t >> 8) == 0)
    rotate_decompose4_from<L, HI, 0>(acc_lds, acc_reg, r, m0, a, base_log, out);
  else
    rotate_decompose4_from<L, HI, 1>(acc_lds, acc_reg, r, m0, a, base_log, out);
}

// Pointwise external product for one thread: z[gl][w0] are the NTT-domain digits of
// row gl = g*L + l; the key slice of this CMUX is in the internal layout [w0][K1*L][K1][T],
// and key(e) returns this thread's word e*T + t of it.  Result written over z[j][w0] for
// j < K1.
template <int K1, int L, class KeyWord>
GL_HD void mac_with(uint64_t (*z)[8], KeyWord key) {
  // Outputs in groups of four, (w0, j) with j fastest; a group covers whole w0 columns, so
  // every input z[.][w0] of the group is read before its outputs are written over it.
  static_assert(4 % K1 == 0, "mac groups four outputs over whole w0 columns");
#pragma unroll
  for (int q = 0; q < 8 * K1; q += 4) {
    uint64_t lo[4], hi[4], o[4];
    uint32_t c[4];
    // the group's key words first, so that their loads are in flight together
    uint64_t kw[4][K1 * L];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int gl_ = 0; gl_ < K1 * L; gl_++) kw[u][gl_] = key((((q + u) / K1) * (K1 * L) + gl_) * K1 + (q + u) % K1);
    // per output the sum of K1*L full 128-bit products, carries past 2^128 counted in c, one
    // reduction; the four outputs' accumulations interleaved (acc128x4)
#pragma unroll
    for (int u = 0; u < 4; u++) {
      c[u] = 0;
      gl::mul_full(z[0][(q + u) / K1], kw[u][0], lo[u], hi[u]);
    }
#pragma unroll
    for (int gl_ = 1; gl_ < K1 * L; gl_++) {
      uint64_t plo[4], phi[4];
#pragma unroll
      for (int u = 0; u < 4; u++) gl::mul_full(z[gl_][(q + u) / K1], kw[u][gl_], plo[u], phi[u]);
      gl::acc128x4(lo, hi, c, plo, phi);
    }
    gl::reduce160x4(o, lo, hi, c);
#pragma unroll
    for (int u = 0; u < 4; u++) z[(q + u) % K1][(q + u) / K1] = o[u];
  }
}

// mac_with over a plain pointer to the key slice of this CMUX (bsk_i).
template <int K1, int L>
GL_HD void mac(uint64_t (*z)[8], const uint64_t* bsk_i, int t) {
  const uint64_t* bt = bsk_i + t;
  mac_with<K1, L>(z, [bt](int e) { return bt[(size_t)e * T]; });
}

// Sample extract of one output coefficient with Many-LUT
// (tb_pep_mmacc_sample_extract.sv:646-690): v = X^(-b) ACC; mask[g][m] = v_g[offs - m]
// (negated across the wrap), body = v_k[offs].  acc_lds holds the K1 polynomials.
GL_HD uint64_t extract_coef(const uint64_t* acc_lds, int g, int m, uint32_t b, uint32_t offs) {
  int ci = (int)offs - m;
  uint32_t x = (ci >= 0) ? (uint32_t)ci : (uint32_t)(N + ci);
  uint64_t neg1 = (ci < 0);
  uint32_t idx = (x + b) & (2 * N - 1);
  uint64_t val = acc_lds[g * N + (idx & (N - 1))];
  uint64_t neg2 = (idx >= (uint32_t)N);
  return (neg1 ^ neg2) ? (0 - val) : val;
}

// Key-switch output stage (pep_ks_out_process.sv:276-282,336-341): out = [j==n] b' - acc
// mod 2^w, then rounded to LWE_COEF_W bits.
GL_HD uint32_t ks_finish(uint32_t acc, uint32_t body, uint32_t w, uint32_t lc) {
  uint32_t wmask = (w >= 32) ? 0xFFFFFFFFu : ((1u << w) - 1);
  uint32_t v = (body - acc) & wmask;
  uint32_t top = v >> (w - lc);
  uint32_t rb = (v >> (w - lc - 1)) & 1;
  return (top + rb) & ((1u << lc) - 1);
}

// Body switch 2^64 -> 2^w (pep_ks_blram_format.sv:236-246).
GL_HD uint32_t ks_body(uint64_t b, uint32_t w) {
  uint64_t top = b >> (64 - w);
  uint64_t rb = (b >> (64 - w - 1)) & 1;
  return (uint32_t)((top + rb) & ((1ull << w) - 1));
}

}  // namespace pbs
