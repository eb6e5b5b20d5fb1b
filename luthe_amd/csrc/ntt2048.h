// Negacyclic NTT of size N = 2048 over Goldilocks, as per-thread pass functions for a
// 256-thread workgroup (4 waves of 64).  The transform is a 32 x 64 four-step split:
//
//   p[64m + r], m < 32, r < 64
//   A  (per column r): 32-point negacyclic NTT over m, root psi = 8 = 2^3
//        A1: pre-twist psi^m, DFT-8 over m1 (m = m0 + 4 m1), twist Omega^(k1 m0)
//        A2: DFT-4 over m0, k = k1 + 8 k0
//   B  twist by zeta^((2k+1) r), zeta the primitive 4096-th root of unity (table)
//   C  (per row k): 64-point cyclic DFT over r, root omega = 8
//        C1: DFT-8 over r1 (r = r0 + 8 r1), twist omega^(w1 r0)
//        C2: DFT-8 over r0, w = w1 + 8 w0
//   NTT-domain index 64k + w holds P(zeta^(2(k + 32w) + 1)).
//
// That output order is exactly the order of the reference model's ntt() with cut
// [5,6] (sw/ntt/ntt_gf64.sage:367-445: negacyclic radix-32 column, network, ngc phi,
// cyclic radix-64 column), i.e. the HPU's n5c6 cut, so NTT-domain operands given in
// the reference order need no permutation.  All twiddles except B are powers of two
// (omega_64 = 8, sw/ntt/ntt_gf64.sage:35-38): those of the A passes are shifts by
// compile-time amounts (specialised per wave, whose m0 or k1 they depend on), and so are
// C's 8^(w1 r0) whenever a workgroup holds two or more polynomials (C1 map below; with one
// polynomial they come from a 64-entry table).  The inverse runs the forward DFT networks
// with reversed outputs (inverse roots), the inverse twiddles, and folds 1/N into the B
// twist table.
//
// Thread mappings (t = thread id in 0..255, wave w = t >> 6):
//   A map   : r = t & 63, m0 = t >> 6   owns coefficients n = 64 (m0 + 4 m1) + r, m1 < 8
//   A2 map  : r = t & 63, q = t >> 6    owns k1 in {2q, 2q+1}, all k0 < 4
//   C1 map  : k = t & 31, h = (t >> 5) & 1; r0 = 2w + h (one polynomial) or, with 2 or 4
//             polynomials per workgroup, tasks (polynomial h + 2i, r0 = 2w + j), j < 2:
//             r0 is then uniform per wave and task, so C1's twiddles 8^(w1 r0) (and the
//             inverse's 8^(-w1 r0), applied at the start of inverse C1) are compile-time
//             shifts; with one polynomial they come from a 64-entry table
//   C2 map  : k = t >> 3, w1 = t & 7
//   NTT domain: thread (k, w1) owns indices 64 k + w1 + 8 w0, w0 < 8.
//
// LDS layouts per polynomial (2048 u64), XOR-swizzled to be bank-conflict free for
// 64-bit accesses with 32 or 64 banks under every map above:
//   Y1(k1, m0, r) = (4 k1 + m0) * 64 + r
//   T (k, r)      = 64 k + (r ^ k)
//   U (k, w1, r0) = 64 k + (x ^ 2 (x >> 4) ^ k), x = 8 w1 + r0
// (checked for lane groups of 16 over 32 banks and of 32 over 64 banks).  k < 32 leaves
// bit 5 of r alone, so T(k, r + 32) = T(k, r) + 32: the C1 accesses of r1 and r1 + 4 are
// one base and an offset.
//
// Field values stay canonical (gl64.h).  Each pass is split into load / compute / store so that a kernel can put a barrier
// between all loads and all stores of the same buffer (in-place LDS reuse), and a
// sequential CPU emulation of the 256 threads executes the identical code.
#pragma once
#include "gl64.h"

namespace ntt {

constexpr int N = 2048;
constexpr int T = 256;

GL_HD int idx_y1(int k1, int m0, int r) { return (k1 * 4 + m0) * 64 + r; }
GL_HD int idx_t(int k, int r) { return k * 64 + (r ^ k); }
GL_HD int idx_u(int k, int w1, int r0) {
  const int x = 8 * w1 + r0;
  return k * 64 + (x ^ (2 * (x >> 4)) ^ k);
}

// DFT-8 in natural order with root 2^S1 (S1 = 24).  Each radix-2 stage forms its four sums
// first, then its four differences in one gl::sub4 (the operands are dead after it).
// WRAP: bit i set makes the last stage's sum s_i (output x[0], x[2], x[1], x[3] for i = 0..3) a
// gl::add_wrap (below 2^64 but possibly >= p), for outputs that only feed products or shift
// twists by a nonzero amount (both take any 64-bit operand).  Bit 4 (with bits 0..3: every
// output may be >= p) also wraps the earlier sums that reach the last stage only as its first
// operands (stage 1's x[0], stage 2's x[0] and x[4]): add_wrap and sub_fold take any 64-bit
// first operand with a canonical second one.
template <int S1, int WRAP = 0>
GL_HD void dft8(uint64_t* x) {
  static_assert(!(WRAP & 16) || (WRAP & 15) == 15, "bit 4 needs every output non-canonical");
  constexpr int DEEP = WRAP & 16;
  uint64_t s0, s1, s2, s3, d0, d1, d2, d3;
  // stage 1: pairs (j, j+4), twiddle root^j
  s0 = gl::add_w<DEEP>(x[0], x[4]); s1 = gl::add(x[1], x[5]); s2 = gl::add(x[2], x[6]); s3 = gl::add(x[3], x[7]);
  gl::sub4(d0, d1, d2, d3, x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7]);
  x[0] = s0; x[1] = s1; x[2] = s2; x[3] = s3;
  x[4] = d0;
  {
    uint64_t w[3] = {d1, d2, d3};
    gl::mul_pow2c_n<S1, 2 * S1, 3 * S1>(w);
    x[5] = w[0];
    x[6] = w[1];
    x[7] = w[2];
  }
  // stage 2: pairs (h+j, h+j+2), twiddle root^(2j)
  s0 = gl::add_w<DEEP>(x[0], x[2]); s1 = gl::add(x[1], x[3]);
  s2 = gl::add_w<DEEP>(x[4], x[6]); s3 = gl::add(x[5], x[7]);
  gl::sub4(d0, d1, d2, d3, x[0], x[1], x[4], x[5], x[2], x[3], x[6], x[7]);
  x[0] = s0; x[1] = s1; x[4] = s2; x[5] = s3;
  x[2] = d0;
  x[6] = d2;
  {
    uint64_t w[2] = {d1, d3};
    gl::mul_pow2c_n<2 * S1, 2 * S1>(w);
    x[3] = w[0];
    x[7] = w[1];
  }
  // stage 3: pairs (2j, 2j+1); the outputs land bit-reversed (1 <-> 4, 3 <-> 6), which
  // the assignment below undoes
  s0 = gl::add_w<WRAP & 1>(x[0], x[1]);
  s1 = gl::add_w<WRAP & 2>(x[2], x[3]);
  s2 = gl::add_w<WRAP & 4>(x[4], x[5]);
  s3 = gl::add_w<WRAP & 8>(x[6], x[7]);
  gl::sub4(d0, d1, d2, d3, x[0], x[2], x[4], x[6], x[1], x[3], x[5], x[7]);
  x[0] = s0; x[4] = d0; x[2] = s1; x[6] = d1; x[1] = s2; x[5] = d2; x[3] = s3; x[7] = d3;
}

// Two DFT-4s in natural order with root 2^S2 (S2 = 48), over x[0..3] and x[4..7]; each of
// the two radix-2 stages takes its four differences in one gl::sub4.  WRAP as in dft8, over
// the sums x[0], x[1], x[4], x[5] (bits 0..3); bit 4 also wraps the first-stage sums a0 and c0
// (first operands of the second stage only).
template <int S2, int WRAP = 0>
GL_HD void dft4x2(uint64_t* x) {
  static_assert(!(WRAP & 16) || (WRAP & 15) == 15, "bit 4 needs every output non-canonical");
  uint64_t a0 = gl::add_w<WRAP & 16>(x[0], x[2]), b0 = gl::add(x[1], x[3]);
  uint64_t c0 = gl::add_w<WRAP & 16>(x[4], x[6]), e0 = gl::add(x[5], x[7]);
  uint64_t a1, b1, c1, e1;
  gl::sub4(a1, b1, c1, e1, x[0], x[1], x[4], x[5], x[2], x[3], x[6], x[7]);
  {
    uint64_t w[2] = {b1, e1};
    gl::mul_pow2c_n<S2, S2>(w);
    b1 = w[0];
    e1 = w[1];
  }
  x[0] = gl::add_w<WRAP & 1>(a0, b0);
  x[1] = gl::add_w<WRAP & 2>(a1, b1);
  x[4] = gl::add_w<WRAP & 4>(c0, e0);
  x[5] = gl::add_w<WRAP & 8>(c1, e1);
  gl::sub4(x[2], x[3], x[6], x[7], a0, a1, c0, c1, b0, b1, e0, e1);
}

// Inverse DFT-8 / DFT-4 (roots 2^-24, 2^-48): sum_j x[j] w^(-jk) is output (n - k) mod n
// of the forward transform, so the inverse is the forward network (no negated twiddles)
// with outputs k and n - k exchanged -- a register renaming.
GL_HD void swap_(uint64_t& a, uint64_t& b) {
  uint64_t t = a;
  a = b;
  b = t;
}
// WRAP as in dft8 (bit i: the sum that ends in output x[0], x[6], x[7], x[5] for i = 0..3)
template <int WRAP = 0>
GL_HD void idft8(uint64_t* x) {
  dft8<24, WRAP>(x);
  swap_(x[1], x[7]);
  swap_(x[2], x[6]);
  swap_(x[3], x[5]);
}
// WRAP as in dft4x2 (bit i: the sum that ends in output x[0], x[3], x[4], x[7] for i = 0..3)
template <int WRAP = 0>
GL_HD void idft4x2(uint64_t* x) {
  dft4x2<48, WRAP>(x);
  swap_(x[1], x[3]);
  swap_(x[5], x[7]);
}

// Twists by compile-time powers of two (every shift amount a constant, so each wave-uniform
// case compiles to straight-line code; the lanes' subtractions are interleaved).
template <int M0>
GL_HD void a1_post_twist(uint64_t* v) {  // v[k1] *= 2^(3 M0 (1 + 2 k1))
  gl::mul_pow2c_n<3 * M0, 9 * M0, 15 * M0, 21 * M0, 27 * M0, 33 * M0, 39 * M0, 45 * M0>(v);
}
// v[h * 4 + m0] *= -2^(-3 m0 (2 k1 + 1)) = 2^(96 - 3 m0 (2 k1 + 1)), k1 = 2 Q + h
template <int Q>
GL_HD void ia2_twist(uint64_t* v) {
  constexpr int F = 3 * (4 * Q + 1), G = 3 * (4 * Q + 3);
  gl::mul_pow2c_n<96, 96 - F, 96 - 2 * F, 96 - 3 * F, 96, 96 - G, 96 - 2 * G, 96 - 3 * G>(v);
}
GL_HD void ia1_twist(uint64_t* v) {  // v[m1] *= -2^(-12 m1) = 2^(96 - 12 m1)
  gl::mul_pow2c_n<96, 84, 72, 60, 48, 36, 24, 12>(v);
}

// Wave-uniform copy of a value that is uniform within each 64-lane wave (the A-pass
// index m0 = t >> 6): lets shifts by it compile to scalar-controlled code.
GL_HD int wave_uniform(int x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_readfirstlane(x);
#else
  return x;
#endif
}

// ---------------------------------------------------------------- forward passes
// A1: in regs v[m1] = p[64 (m0 + 4 m1) + r].  psi^(m0 + 4 m1) Omega^(k1 m0) is split into
// a compile-time pre-twist psi^(4 m1) = 2^(12 m1) and a post-twist
// psi^m0 Omega^(k1 m0) = 2^(3 m0 (1 + 2 k1)) that is uniform per wave.
// pretwisted: v already holds p[n] psi^(4 m1) (the blind rotation's digit lift applies it).
// The DFT's last sums (outputs 0..3) skip canonicalisation: for m0 > 0 every output is twisted
// by a nonzero shift, which takes any 64-bit operand; for m0 = 0 (no twist) the four sums are
// made canonical afterwards, in the wave with the least twist work (m0 is uniform per wave).
GL_HD void fwd_a1_compute(uint64_t* v, int m0u, bool pretwisted = false) {
  if (!pretwisted) gl::mul_pow2c_n<12, 24, 36, 48, 60, 72, 84>(v + 1);
  dft8<24, 0xF>(v);
  switch (m0u) {
    case 1: a1_post_twist<1>(v); break;
    case 2: a1_post_twist<2>(v); break;
    case 3: a1_post_twist<3>(v); break;
    default:
#pragma unroll
      for (int i = 0; i < 4; i++) v[i] = gl::canonical_fold(v[i]);
      break;
  }
}
GL_HD void fwd_a1_store(const uint64_t* v, uint64_t* lds, int r, int m0) {
#pragma unroll
  for (int k1 = 0; k1 < 8; k1++) lds[idx_y1(k1, m0, r)] = v[k1];
}
// A2
GL_HD void fwd_a2_load(uint64_t* v, const uint64_t* lds, int r, int q) {
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int m0 = 0; m0 < 4; m0++) v[h * 4 + m0] = lds[idx_y1(2 * q + h, m0, r)];
}
GL_HD void fwd_a2_compute(uint64_t* v, int r, int q, const uint64_t* tw) {
  dft4x2<48, 0x1F>(v);  // the outputs feed the B twist products
#pragma unroll
  for (int h = 0; h < 2; h++) {
    uint64_t w[4];
#pragma unroll
    for (int k0 = 0; k0 < 4; k0++) w[k0] = tw[(2 * q + h + 8 * k0) * 64 + r];
    gl::mul4(v + h * 4, v + h * 4, w);
  }
}
GL_HD void fwd_a2_store(const uint64_t* v, uint64_t* lds, int r, int q) {
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int k0 = 0; k0 < 4; k0++) lds[idx_t(2 * q + h + 8 * k0, r)] = v[h * 4 + k0];
}
// C1 tasks (C1 map above): polynomial and r0 of task j.  Wave w takes the r0 pair
// R0_PAIR[w] rather than {2w, 2w+1}: the twiddle cost grows with r0 (r0 = 0 is free), and
// the pairing evens the per-wave VALU load of the C and A twists (wave 0 has no A1 twist).
constexpr int R0_PAIR[4][2] = {{3, 4}, {1, 6}, {2, 5}, {0, 7}};
GL_HD int r0_of(int w, int j) { return (int)((0x70526143u >> (4 * (2 * w + j))) & 15); }  // R0_PAIR[w][j]
template <int NP>
GL_HD int c1_poly(int j, int h) { return NP == 1 ? 0 : h + 2 * (j >> 1); }
template <int NP>
GL_HD int c1_r0(int j, int w, int h) { return NP == 1 ? r0_of(w, h) : r0_of(w, j & 1); }

// v[w1] *= 8^(S w1 R0) = 2^(3 S w1 R0), w1 = 1..7 (S = 1 forward, -1 inverse).
template <int S, int R0>
GL_HD void c_twist(uint64_t* v) {
  constexpr int F = 3 * S * R0;
  gl::mul_pow2c_n<F, 2 * F, 3 * F, 4 * F, 5 * F, 6 * F, 7 * F>(v + 1);
}
template <int S, int NP, int W>
GL_HD void c_twist_wave(uint64_t (*v)[8]) {
#pragma unroll
  for (int j = 0; j < NP; j++) {
    if (j & 1)
      c_twist<S, R0_PAIR[W][1]>(v[j]);
    else
      c_twist<S, R0_PAIR[W][0]>(v[j]);
  }
}
// Wave-specialised twist of every task (NP >= 2), or the per-lane table tab64 (NP = 1).
template <int S, int NP>
GL_HD void c_twist_all(uint64_t (*v)[8], int t, const uint64_t* tab64) {
  if constexpr (NP == 1) {
    const int r0 = c1_r0<1>(0, t >> 6, (t >> 5) & 1);
#pragma unroll
    for (int w1 = 1; w1 < 8; w1++) v[0][w1] = gl::mul(v[0][w1], tab64[8 * w1 + r0]);
  } else {
    switch (wave_uniform(t >> 6)) {
      case 0: c_twist_wave<S, NP, 0>(v); break;
      case 1: c_twist_wave<S, NP, 1>(v); break;
      case 2: c_twist_wave<S, NP, 2>(v); break;
      default: c_twist_wave<S, NP, 3>(v); break;
    }
  }
}

// C1: DFT-8 over r1, then omega^(w1 r0).  lds: the NP polynomials, 2048 words apart.
template <int NP>
GL_HD void fwd_c1_load(uint64_t (*v)[8], const uint64_t* lds, int t) {
  const int k = t & 31, h = (t >> 5) & 1, w = t >> 6;
#pragma unroll
  for (int j = 0; j < NP; j++) {
    const uint64_t* pl = lds + c1_poly<NP>(j, h) * N;
    const int r0 = c1_r0<NP>(j, w, h);
#pragma unroll
    for (int r1 = 0; r1 < 4; r1++) {
      const uint64_t* q = pl + idx_t(k, r0 + 8 * r1);
      v[j][r1] = q[0];
      v[j][r1 + 4] = q[32];
    }
  }
}
// The sums ending in outputs 1..7 (s1..s3) feed the twist: with NP = 1 a product (any 64-bit
// operand); with NP >= 2 shifts that are nonzero for the odd tasks (r0 = R0_PAIR[w][1] > 0), so
// those skip canonicalisation.  Output 0 is not twisted and stays canonical.
template <int NP>
GL_HD void fwd_c1_compute(uint64_t (*v)[8], int t, const uint64_t* c64) {
#pragma unroll
  for (int j = 0; j < NP; j++) {
    if (NP == 1 || (j & 1))
      dft8<24, 0xE>(v[j]);
    else
      dft8<24>(v[j]);
  }
  c_twist_all<1, NP>(v, t, c64);
}
template <int NP>
GL_HD void fwd_c1_store(const uint64_t (*v)[8], uint64_t* lds, int t) {
  const int k = t & 31, h = (t >> 5) & 1, w = t >> 6;
#pragma unroll
  for (int j = 0; j < NP; j++) {
    uint64_t* pl = lds + c1_poly<NP>(j, h) * N;
    const int r0 = c1_r0<NP>(j, w, h);
#pragma unroll
    for (int w1 = 0; w1 < 8; w1++) pl[idx_u(k, w1, r0)] = v[j][w1];
  }
}
// C2
GL_HD void fwd_c2_load(uint64_t* v, const uint64_t* lds, int k, int w1) {
#pragma unroll
  for (int r0 = 0; r0 < 8; r0++) v[r0] = lds[idx_u(k, w1, r0)];
}
// v[w0] = Z[64k + w1 + 8 w0], below 2^64 but possibly >= p: the values only feed products (the
// external product's MAC, the NTT product's pointwise multiply); stored results are made
// canonical by their kernels
GL_HD void fwd_c2_compute(uint64_t* v) { dft8<24, 0x1F>(v); }

// ---------------------------------------------------------------- inverse passes
// IC2: inverse DFT-8 over w0 (C2 map); the twiddle omega^(-w1 r0) is applied by IC1.
GL_HD void inv_c2_compute(uint64_t* v) { idft8(v); }
GL_HD void inv_c2_store(const uint64_t* v, uint64_t* lds, int k, int w1) {
#pragma unroll
  for (int r0 = 0; r0 < 8; r0++) lds[idx_u(k, w1, r0)] = v[r0];
}
// IC1 (C1 map): omega^(-w1 r0), inverse DFT-8 over w1, then the B untwist
// itw[(r0 + 8 r1) * 32 + k] (1/N folded in; r-major so that the lanes' k are contiguous).
template <int NP>
GL_HD void inv_c1_load(uint64_t (*v)[8], const uint64_t* lds, int t) {
  const int k = t & 31, h = (t >> 5) & 1, w = t >> 6;
#pragma unroll
  for (int j = 0; j < NP; j++) {
    const uint64_t* pl = lds + c1_poly<NP>(j, h) * N;
    const int r0 = c1_r0<NP>(j, w, h);
#pragma unroll
    for (int w1 = 0; w1 < 8; w1++) v[j][w1] = pl[idx_u(k, w1, r0)];
  }
}
template <int NP>
GL_HD void inv_c1_compute(uint64_t (*v)[8], int t, const uint64_t* itw, const uint64_t* ic64) {
  c_twist_all<-1, NP>(v, t, ic64);
  const int k = t & 31, h = (t >> 5) & 1, w = t >> 6;
#pragma unroll
  for (int j = 0; j < NP; j++) {
    idft8<0x1F>(v[j]);  // the outputs feed the B untwist products
    const int r0 = c1_r0<NP>(j, w, h);
    uint64_t w[8];
#pragma unroll
    for (int r1 = 0; r1 < 8; r1++) w[r1] = itw[(r0 + 8 * r1) * 32 + k];
    gl::mul4(v[j], v[j], w);
    gl::mul4(v[j] + 4, v[j] + 4, w + 4);
  }
}
template <int NP>
GL_HD void inv_c1_store(const uint64_t (*v)[8], uint64_t* lds, int t) {
  const int k = t & 31, h = (t >> 5) & 1, w = t >> 6;
#pragma unroll
  for (int j = 0; j < NP; j++) {
    uint64_t* pl = lds + c1_poly<NP>(j, h) * N;
    const int r0 = c1_r0<NP>(j, w, h);
#pragma unroll
    for (int r1 = 0; r1 < 4; r1++) {
      uint64_t* q = pl + idx_t(k, r0 + 8 * r1);
      q[0] = v[j][r1];
      q[32] = v[j][r1 + 4];
    }
  }
}
GL_HD void inv_a2_load(uint64_t* v, const uint64_t* lds, int r, int q) {
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int k0 = 0; k0 < 4; k0++) v[h * 4 + k0] = lds[idx_t(2 * q + h + 8 * k0, r)];
}
// IA2: Omega^(-k1 m0) psi^(-m0) = 2^(-3 m0 (2 k1 + 1)), uniform per wave (k1 = 2q + h).
// The psi^(-m0) part of the final untwist psi^-(m0 + 4 m1) is constant over IA1's
// DFT-8 and moves here; both stages carry a factor -1 (the cheaper shift forms).
// Outputs 0 and 4 take the twist -1 (a negation, which needs canonical input); the sums that
// end in outputs 3 and 7 meet nonzero shifts and skip canonicalisation.
GL_HD void inv_a2_compute(uint64_t* v, int qu) {
  idft4x2<0xA>(v);
  switch (qu) {
    case 0: ia2_twist<0>(v); break;
    case 1: ia2_twist<1>(v); break;
    case 2: ia2_twist<2>(v); break;
    default: ia2_twist<3>(v); break;
  }
}
GL_HD void inv_a2_store(const uint64_t* v, uint64_t* lds, int r, int q) {
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int m0 = 0; m0 < 4; m0++) lds[idx_y1(2 * q + h, m0, r)] = v[h * 4 + m0];
}
GL_HD void inv_a1_load(uint64_t* v, const uint64_t* lds, int r, int m0) {
#pragma unroll
  for (int k1 = 0; k1 < 8; k1++) v[k1] = lds[idx_y1(k1, m0, r)];
}
// Output v[m1] = p[64 (m0 + 4 m1) + r] (canonical): the remaining untwist psi^(-4 m1)
// times the -1 that cancels IA2's.
// Output 0's twist is -1 (needs canonical input); the sums ending in outputs 5, 6, 7 meet
// nonzero shifts and skip canonicalisation.
GL_HD void inv_a1_compute(uint64_t* v) {
  idft8<0xE>(v);
  ia1_twist(v);
}

// Twiddle tables (host side), one buffer of TABLE_WORDS u64:
//   [0, 2048)    tw [k*64 + r] = zeta^((2k+1) r)
//   [2048, 4096) itw[r*32 + k] = zeta^(-(2k+1) r) / 2048
//   [4096, 4160) c64 [8 w1 + r0] = 8^(w1 r0)
//   [4160, 4224) ic64[8 w1 + r0] = 8^(-w1 r0)
// zeta = OMG_2_32^(2^20) (sw/ntt/ntt_gf64.sage:44,69-75 power_of_rou(4096, 1)).
constexpr int TW_OFF = 0, ITW_OFF = 2048, C64_OFF = 4096, IC64_OFF = 4160, TABLE_WORDS = 4224;

inline void make_twist_tables(uint64_t* tab) {
  const uint64_t OMG_2_32 = 16334397945464290598ull;
  uint64_t zeta = gl::pow(OMG_2_32, 1ull << 20);
  uint64_t izeta = gl::inv(zeta);
  uint64_t inv_n = gl::inv(2048);
  for (int k = 0; k < 32; k++)
    for (int r = 0; r < 64; r++) {
      uint64_t e = (uint64_t)(2 * k + 1) * (uint64_t)r;
      tab[TW_OFF + k * 64 + r] = gl::pow(zeta, e);
      tab[ITW_OFF + r * 32 + k] = gl::mul(gl::pow(izeta, e), inv_n);
    }
  uint64_t i8 = gl::inv(8);
  for (int w1 = 0; w1 < 8; w1++)
    for (int r0 = 0; r0 < 8; r0++) {
      tab[C64_OFF + 8 * w1 + r0] = gl::pow(8, (uint64_t)(w1 * r0));
      tab[IC64_OFF + 8 * w1 + r0] = gl::pow(i8, (uint64_t)(w1 * r0));
    }
}

}  // namespace ntt
