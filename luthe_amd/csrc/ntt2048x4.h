// Negacyclic NTT of size N = 2048 over Goldilocks for a 512-thread workgroup, four values
// per thread (8 waves of 64): the same transform and NTT-domain order as ntt2048.h, with each
// DFT-8 of that four-step split into passes of at most four values per thread, so that one
// polynomial keeps two waves per SIMD busy.
//
//   p[64m + r], m = m0 + 4 m1 (m0 < 4, m1 < 8), r < 64
//   A1a  per (r, m0): pre-twist psi^(4 m1) = 2^(12 m1); radix-2 over the top bit of m1
//        (m1 = j0 + 4 j1): u_0[j0] = x[j0] + x[j0+4], u_1[j0] = (x[j0] - x[j0+4]) 2^(24 j0)
//   A1b  per (r, m0, s): DFT-4 over j0 (root 2^48) gives y[k1 = 2q + s] (the DFT-8 over m1 of
//        ntt2048.h's A1, split decimation-in-frequency); post-twist 2^(3 m0 (1 + 2 k1))
//   A2   per (r, k1): DFT-4 over m0, k = k1 + 8 k0; B twist zeta^((2k+1) r)
//   C    per row k, the 64-point DFT over r (root 8) as three radix-4 stages,
//        r = ra + 4 rb + 16 rc, w = wq + 4 wb + 16 wa:
//   C1   per (k, r' = ra + 4 rb): DFT-4 over rc (root 2^48) -> wq; twist 8^(r' wq)
//   C2   per (k, wq, ra): DFT-4 over rb -> wb; twist 2^(12 ra wb)
//   C3   per (k, wq, wb): DFT-4 over ra -> wa
//   NTT-domain index 64k + w, as in ntt2048.h.
//
// Thread maps (t < 512, wave W = t >> 6, lane = t & 63, h = lane >> 5):
//   A1a: r = lane, m0 = W & 3, g = W >> 2; registers m1 = 2g, 2g+1, 2g+4, 2g+5 (the
//        coefficient layout of the accumulator)
//   A1b: r = lane, m0 = W & 3, s = W >> 2
//   A2 : r = lane, k1 = W
//   C1 : k = lane & 31, r' = 2W + h
//   C2 : k = lane & 31, ra = W & 3, wq = 2 (W >> 2) + h
//   C3 : wd = t >> 8, k = (t >> 3) & 31, w1 = t & 7 (wq = w1 & 3, wb = (w1 >> 2) + 2 wd):
//        register j holds index 64k + w1 + 8 (wd + 2j), so thread t reads the key words of
//        thread t & 255 of ntt2048.h's C2 map, rows w0 = wd + 2j.
// Every twist is a compile-time shift in a wave-uniform index except C1's (r' differs
// between half-waves) and B: those come from a twiddle source (tables or registers).
//
// LDS layouts (2048 u64 per polynomial), bank-conflict free for 64-bit accesses of 32-lane
// groups: A-side index x * 64 + r (lanes run over r); C-side 64 k + (x ^ k) with x < 64
// (lanes run over k, ntt2048.h's T swizzle):
//   LA (m0, s, j0, r)   = (8 m0 + 4 s + j0) * 64 + r
//   Y1 (k1, m0, r)      = ntt2048.h's Y1
//   T  (k, r)           = ntt2048.h's T
//   U2 (k, wq, r')      = 64 k + ((16 wq + r') ^ k)
//   U3 (k, wb, wq, ra)  = 64 k + ((16 wb + 4 wq + ra) ^ k)
//
// The inverse runs the stages backwards with inverse DFT-4s (the forward network with
// outputs 1 and 3 exchanged), inverse twists, the B untwist with 1/N folded in (the
// ntt2048.h table), and psi^-m as there: 2^(96 - 3 m0 (2 k1 + 1)) after IA2 and
// 2^(96 - 12 m1) after IA1a (the two factors -1 cancel).
#pragma once
#include "ntt2048.h"

// member functions (GL_HD is static on the host)
#if defined(__HIPCC__)
#define GL_HDM __host__ __device__ __forceinline__
#else
#define GL_HDM inline
#endif

namespace ntt4 {

using ntt::N;
constexpr int T4 = 512;

GL_HD int idx_la(int m0, int s, int j0, int r) { return (8 * m0 + 4 * s + j0) * 64 + r; }
GL_HD int idx_u2(int k, int wq, int rp) { return 64 * k + ((16 * wq + rp) ^ k); }
GL_HD int idx_u3(int k, int wb, int wq, int ra) { return 64 * k + ((16 * wb + 4 * wq + ra) ^ k); }

// DFT-4 in natural order with root 2^S (S = 48 forward).  The subtractions go in pairs
// (gl::sub2): x0 - x2 with x1 - x3, then the reduction of (x1 - x3) 2^S (S in (32, 64): a
// 128-bit value lo + 2^64 hl + 2^96 hh, reduced as in gl::reduce128) with a0 - b0.
// WRAP: bit i set makes the final sum x[i] (i = 0, 1) a gl::add_wrap (below 2^64, possibly
// >= p), for outputs that only feed products or shift twists by a nonzero amount; bit 2 (with
// bits 0, 1: every output may be >= p) also wraps a0, which reaches the second stage only as a
// first operand (add_wrap and sub_fold take any 64-bit first operand with a canonical second).
template <int S, int WRAP = 0>
GL_HD void dft4(uint64_t* x) {
  static_assert(S > 32 && S < 64, "the paired reduction assumes a shift in (32, 64)");
  static_assert(!(WRAP & 4) || (WRAP & 3) == 3, "bit 2 needs every output non-canonical");
  const uint64_t a0 = gl::add_w<WRAP & 4>(x[0], x[2]), b0 = gl::add(x[1], x[3]);
  uint64_t a1, d, t;
  gl::sub2(a1, d, x[0], x[1], x[2], x[3]);
  gl::sub2(t, x[2], d << S, a0, d >> (96 - S), b0);
  const uint64_t b1 = gl::add_eps_prod(t, (uint32_t)(d >> (64 - S)));
  x[0] = gl::add_w<WRAP & 1>(a0, b0);
  x[1] = gl::add_w<WRAP & 2>(a1, b1);
  x[3] = gl::sub(a1, b1);
}
// WRAP as in dft4 (bit 0: output x[0], bit 1: output x[3])
template <int WRAP = 0>
GL_HD void idft4(uint64_t* x) {
  dft4<48, WRAP>(x);
  ntt::swap_(x[1], x[3]);
}

// Per-thread indices of the maps above.
struct Map {
  int t, W, lane, h;
  GL_HDM explicit Map(int t_) : t(t_), W(t_ >> 6), lane(t_ & 63), h((t_ >> 5) & 1) {}
  GL_HDM int r() const { return lane; }
  GL_HDM int m0() const { return W & 3; }
  GL_HDM int hi2() const { return W >> 2; }  // g (A1a) or s (A1b)
  GL_HDM int kc() const { return lane & 31; }          // k of C1 / C2
  GL_HDM int rp() const { return 2 * W + h; }          // r' of C1
  GL_HDM int ra() const { return W & 3; }              // ra of C2
  GL_HDM int wq2() const { return 2 * (W >> 2) + h; }  // wq of C2
  GL_HDM int k3() const { return (t >> 3) & 31; }
  GL_HDM int w1() const { return t & 7; }
  GL_HDM int wd() const { return t >> 8; }
  GL_HDM int wq3() const { return t & 3; }
  GL_HDM int wb3() const { return ((t >> 2) & 1) + 2 * (t >> 8); }
};

// ---------------------------------------------------------------- forward passes
// A1a: v[i] = p[64 (m0 + 4 m1_i) + r], m1 = (2G, 2G+1, 2G+4, 2G+5); PRE: already pretwisted.
// Out: v[2 s + e] = u_s[j0 = 2G + e].
template <int G, bool PRE>
GL_HD void a1a(uint64_t* v) {
  if constexpr (!PRE) gl::mul_pow2c_n<12 * (2 * G), 12 * (2 * G + 1), 12 * (2 * G + 4), 12 * (2 * G + 5)>(v);
  const uint64_t u00 = gl::add(v[0], v[2]), u01 = gl::add(v[1], v[3]);
  uint64_t d[2];
  gl::sub2(d[0], d[1], v[0], v[1], v[2], v[3]);
  gl::mul_pow2c_n<24 * (2 * G), 24 * (2 * G + 1)>(d);
  v[0] = u00;
  v[1] = u01;
  v[2] = d[0];
  v[3] = d[1];
}
template <bool PRE>
GL_HD void fwd_a1a_compute(uint64_t* v, int gu) {
  if (gu == 0)
    a1a<0, PRE>(v);
  else
    a1a<1, PRE>(v);
}
GL_HD void fwd_a1a_store(const uint64_t* v, uint64_t* lds, const Map& m) {
  const int g = m.hi2(), m0 = m.m0(), r = m.r();
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int e = 0; e < 2; e++) lds[idx_la(m0, s, 2 * g + e, r)] = v[2 * s + e];
}
// A1b: DFT-4 over j0, post-twist 2^(3 M0 (1 + 2 k1)), k1 = 2q + S.
GL_HD void fwd_a1b_load(uint64_t* v, const uint64_t* lds, const Map& m) {
  const int s = m.hi2(), m0 = m.m0(), r = m.r();
#pragma unroll
  for (int j0 = 0; j0 < 4; j0++) v[j0] = lds[idx_la(m0, s, j0, r)];
}
template <int M0, int S>
GL_HD void a1b(uint64_t* v) {
  dft4<48, M0 == 0 ? 0 : 3>(v);  // M0 > 0: every output meets a nonzero shift
  gl::mul_pow2c_n<3 * M0 * (1 + 2 * S), 3 * M0 * (1 + 2 * (2 + S)), 3 * M0 * (1 + 2 * (4 + S)),
                  3 * M0 * (1 + 2 * (6 + S))>(v);
}
GL_HD void fwd_a1b_compute(uint64_t* v, int Wu) {
  switch (Wu) {
    case 0: a1b<0, 0>(v); break;
    case 1: a1b<1, 0>(v); break;
    case 2: a1b<2, 0>(v); break;
    case 3: a1b<3, 0>(v); break;
    case 4: a1b<0, 1>(v); break;
    case 5: a1b<1, 1>(v); break;
    case 6: a1b<2, 1>(v); break;
    default: a1b<3, 1>(v); break;
  }
}
GL_HD void fwd_a1b_store(const uint64_t* v, uint64_t* lds, const Map& m) {
  const int s = m.hi2(), m0 = m.m0(), r = m.r();
#pragma unroll
  for (int q = 0; q < 4; q++) lds[ntt::idx_y1(2 * q + s, m0, r)] = v[q];
}
// A2: DFT-4 over m0, then B: v[k0] *= tw(k1 + 8 k0, r) (tw4 = the thread's four factors).
GL_HD void fwd_a2_load(uint64_t* v, const uint64_t* lds, const Map& m) {
#pragma unroll
  for (int m0 = 0; m0 < 4; m0++) v[m0] = lds[ntt::idx_y1(m.W, m0, m.r())];
}
GL_HD void fwd_a2_compute(uint64_t* v, const uint64_t* tw4) {
  dft4<48, 7>(v);  // the outputs feed the B twist products
  gl::mul4(v, v, tw4);
}
GL_HD void fwd_a2_store(const uint64_t* v, uint64_t* lds, const Map& m) {
#pragma unroll
  for (int k0 = 0; k0 < 4; k0++) lds[ntt::idx_t(m.W + 8 * k0, m.r())] = v[k0];
}
// C1: DFT-4 over rc, twist 8^(r' wq) (c1w[wq], c1w[0] = 1).
GL_HD void fwd_c1_load(uint64_t* v, const uint64_t* lds, const Map& m) {
  const int k = m.kc(), rp = m.rp();
#pragma unroll
  for (int rc = 0; rc < 4; rc++) v[rc] = lds[ntt::idx_t(k, rp + 16 * rc)];
}
GL_HD void fwd_c1_compute(uint64_t* v, const uint64_t* c1w) {
  dft4<48, 7>(v);  // the outputs feed the twist products
  gl::mul4(v, v, c1w);
}
GL_HD void fwd_c1_store(const uint64_t* v, uint64_t* lds, const Map& m) {
  const int k = m.kc(), rp = m.rp();
#pragma unroll
  for (int wq = 0; wq < 4; wq++) lds[idx_u2(k, wq, rp)] = v[wq];
}
// C2: DFT-4 over rb, twist 2^(12 RA wb).
GL_HD void fwd_c2_load(uint64_t* v, const uint64_t* lds, const Map& m) {
  const int k = m.kc(), wq = m.wq2(), ra = m.ra();
#pragma unroll
  for (int rb = 0; rb < 4; rb++) v[rb] = lds[idx_u2(k, wq, ra + 4 * rb)];
}
template <int RA>
GL_HD void c2(uint64_t* v) {
  dft4<48, RA == 0 ? 0 : 2>(v);  // output 1 meets a nonzero shift when RA > 0
  gl::mul_pow2c_n<12 * RA, 24 * RA, 36 * RA>(v + 1);
}
GL_HD void fwd_c2_compute(uint64_t* v, int rau) {
  switch (rau) {
    case 0: c2<0>(v); break;
    case 1: c2<1>(v); break;
    case 2: c2<2>(v); break;
    default: c2<3>(v); break;
  }
}
GL_HD void fwd_c2_store(const uint64_t* v, uint64_t* lds, const Map& m) {
  const int k = m.kc(), wq = m.wq2(), ra = m.ra();
#pragma unroll
  for (int wb = 0; wb < 4; wb++) lds[idx_u3(k, wb, wq, ra)] = v[wb];
}
// C3: DFT-4 over ra; v[j] = Z[64k + w1 + 8 (wd + 2j)].
GL_HD void fwd_c3_load(uint64_t* v, const uint64_t* lds, const Map& m) {
  const int k = m.k3(), wq = m.wq3(), wb = m.wb3();
#pragma unroll
  for (int ra = 0; ra < 4; ra++) v[ra] = lds[idx_u3(k, wb, wq, ra)];
}
// outputs below 2^64 but possibly >= p: they only feed the MAC's products
GL_HD void fwd_c3_compute(uint64_t* v) { dft4<48, 7>(v); }

// ---------------------------------------------------------------- inverse passes
GL_HD void inv_c3_compute(uint64_t* v) { idft4(v); }  // wa -> ra
GL_HD void inv_c3_store(const uint64_t* v, uint64_t* lds, const Map& m) {
  const int k = m.k3(), wq = m.wq3(), wb = m.wb3();
#pragma unroll
  for (int ra = 0; ra < 4; ra++) lds[idx_u3(k, wb, wq, ra)] = v[ra];
}
GL_HD void inv_c2_load(uint64_t* v, const uint64_t* lds, const Map& m) {
  const int k = m.kc(), wq = m.wq2(), ra = m.ra();
#pragma unroll
  for (int wb = 0; wb < 4; wb++) v[wb] = lds[idx_u3(k, wb, wq, ra)];
}
template <int RA>
GL_HD void ic2(uint64_t* v) {  // 2^(-12 RA wb), then wb -> rb
  gl::mul_pow2c_n<-12 * RA, -24 * RA, -36 * RA>(v + 1);
  idft4<7>(v);  // the outputs feed inverse C1's products
}
GL_HD void inv_c2_compute(uint64_t* v, int rau) {
  switch (rau) {
    case 0: ic2<0>(v); break;
    case 1: ic2<1>(v); break;
    case 2: ic2<2>(v); break;
    default: ic2<3>(v); break;
  }
}
GL_HD void inv_c2_store(const uint64_t* v, uint64_t* lds, const Map& m) {
  const int k = m.kc(), wq = m.wq2(), ra = m.ra();
#pragma unroll
  for (int rb = 0; rb < 4; rb++) lds[idx_u2(k, wq, ra + 4 * rb)] = v[rb];
}
GL_HD void inv_c1_load(uint64_t* v, const uint64_t* lds, const Map& m) {
  const int k = m.kc(), rp = m.rp();
#pragma unroll
  for (int wq = 0; wq < 4; wq++) v[wq] = lds[idx_u2(k, wq, rp)];
}
// 8^(-r' wq) (ic1w), wq -> rc, then the B untwist itw4[rc] = itw((r' + 16 rc), k).
GL_HD void inv_c1_compute(uint64_t* v, const uint64_t* ic1w, const uint64_t* itw4) {
  gl::mul4(v, v, ic1w);
  idft4<7>(v);  // the outputs feed the B untwist products
  gl::mul4(v, v, itw4);
}
GL_HD void inv_c1_store(const uint64_t* v, uint64_t* lds, const Map& m) {
  const int k = m.kc(), rp = m.rp();
#pragma unroll
  for (int rc = 0; rc < 4; rc++) lds[ntt::idx_t(k, rp + 16 * rc)] = v[rc];
}
GL_HD void inv_a2_load(uint64_t* v, const uint64_t* lds, const Map& m) {
#pragma unroll
  for (int k0 = 0; k0 < 4; k0++) v[k0] = lds[ntt::idx_t(m.W + 8 * k0, m.r())];
}
template <int K1>
GL_HD void ia2(uint64_t* v) {  // k0 -> m0, then 2^(96 - 3 m0 (2 K1 + 1))
  idft4<2>(v);  // output 3 meets a nonzero shift; output 0 a negation (canonical input)
  gl::mul_pow2c_n<96, 96 - 3 * (2 * K1 + 1), 96 - 6 * (2 * K1 + 1), 96 - 9 * (2 * K1 + 1)>(v);
}
GL_HD void inv_a2_compute(uint64_t* v, int Wu) {
  switch (Wu) {
    case 0: ia2<0>(v); break;
    case 1: ia2<1>(v); break;
    case 2: ia2<2>(v); break;
    case 3: ia2<3>(v); break;
    case 4: ia2<4>(v); break;
    case 5: ia2<5>(v); break;
    case 6: ia2<6>(v); break;
    default: ia2<7>(v); break;
  }
}
GL_HD void inv_a2_store(const uint64_t* v, uint64_t* lds, const Map& m) {
#pragma unroll
  for (int m0 = 0; m0 < 4; m0++) lds[ntt::idx_y1(m.W, m0, m.r())] = v[m0];
}
GL_HD void inv_a1b_load(uint64_t* v, const uint64_t* lds, const Map& m) {
  const int s = m.hi2(), m0 = m.m0(), r = m.r();
#pragma unroll
  for (int q = 0; q < 4; q++) v[q] = lds[ntt::idx_y1(2 * q + s, m0, r)];
}
GL_HD void inv_a1b_compute(uint64_t* v) { idft4(v); }  // q -> j0
GL_HD void inv_a1b_store(const uint64_t* v, uint64_t* lds, const Map& m) {
  const int s = m.hi2(), m0 = m.m0(), r = m.r();
#pragma unroll
  for (int j0 = 0; j0 < 4; j0++) lds[idx_la(m0, s, j0, r)] = v[j0];
}
// IA1a: v[2 s + e] = u_s[2g + e] -> x[m1] for m1 = (2g, 2g+1, 2g+4, 2g+5), untwisted.
GL_HD void inv_a1a_load(uint64_t* v, const uint64_t* lds, const Map& m) {
  const int g = m.hi2(), m0 = m.m0(), r = m.r();
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int e = 0; e < 2; e++) v[2 * s + e] = lds[idx_la(m0, s, 2 * g + e, r)];
}
template <int G>
GL_HD void ia1a(uint64_t* v) {
  uint64_t b[2] = {v[2], v[3]};
  gl::mul_pow2c_n<-24 * (2 * G), -24 * (2 * G + 1)>(b);
  // both sums meet nonzero shifts below, except x0 for G = 0 (a negation: canonical input)
  const uint64_t x0 = gl::add_w<G>(v[0], b[0]), x1 = gl::add_wrap(v[1], b[1]);
  gl::sub2(v[2], v[3], v[0], v[1], b[0], b[1]);
  v[0] = x0;
  v[1] = x1;
  gl::mul_pow2c_n<96 - 12 * (2 * G), 96 - 12 * (2 * G + 1), 96 - 12 * (2 * G + 4), 96 - 12 * (2 * G + 5)>(v);
}
GL_HD void inv_a1a_compute(uint64_t* v, int gu) {
  if (gu == 0)
    ia1a<0>(v);
  else
    ia1a<1>(v);
}

// The thread's twiddles: B (forward and inverse) and C1's 8^(+-r' wq), from the ntt2048.h
// tables (tw, itw) and computed shifts.
struct Tw4 {
  uint64_t tw[4], c1[4], ic1[4], itw[4];
  GL_HDM void load(const uint64_t* tab, int t) {
    const Map m(t);
    const int rp = m.rp(), k = m.kc();
#pragma unroll
    for (int k0 = 0; k0 < 4; k0++) tw[k0] = tab[ntt::TW_OFF + (m.W + 8 * k0) * 64 + m.r()];
#pragma unroll
    for (int wq = 0; wq < 4; wq++) {
      const uint32_t e = (uint32_t)(3 * rp * wq) % 192;
      c1[wq] = gl::mul_pow2(1, e);
      ic1[wq] = gl::mul_pow2(1, (192 - e) % 192);
    }
#pragma unroll
    for (int rc = 0; rc < 4; rc++) itw[rc] = tab[ntt::ITW_OFF + (rp + 16 * rc) * 32 + k];
  }
};

}  // namespace ntt4
