// Goldilocks field arithmetic (p = 2^64 - 2^32 + 1) for the MI355X PBS engine.
//
// Values are kept in a "weak" form: any u64 congruent to the field element, so the
// hot loops skip canonicalisation; canonical() is applied where bits leave the field
// (INTT output before the p -> 2^64 switch).  Reductions use 2^64 = 2^32 - 1 and
// 2^96 = -1 (mod p), the identities of the reference's Goldilocks reducer
// (hw/common_lib/mod_arithmetic/.../mod_reduct_64bgoldilocks_karatsuba.sv:8-22,112-114).
// Powers of two are multiplications by shifts: omega_64 = 8 = 2^3
// (sw/ntt/ntt_gf64.sage:35-38), so every twiddle inside a radix-64 block is 2^(3j).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GL_HD __host__ __device__ __forceinline__
#else
#define GL_HD static inline
#endif

namespace gl {

constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p

GL_HD uint64_t canonical(uint64_t x) { return x >= P ? x - P : x; }

// a + b (mod p), inputs and output weak.
GL_HD uint64_t add(uint64_t a, uint64_t b) {
  uint64_t s = a + b;
  uint64_t c = (uint64_t)(s < a);            // wrapped by 2^64 == EPS (mod p)
  uint64_t t = s + ((0 - c) & EPS);
  uint64_t c2 = (uint64_t)(t < s);           // at most one more wrap, then t < EPS
  return t + ((0 - c2) & EPS);
}

// a - b (mod p), inputs and output weak.
GL_HD uint64_t sub(uint64_t a, uint64_t b) {
  uint64_t d = a - b;
  uint64_t w = (uint64_t)(a < b);            // borrowed 2^64 == EPS
  uint64_t m = (0 - w) & EPS;
  uint64_t t = d - m;
  uint64_t w2 = (uint64_t)(d < m);
  return t - ((0 - w2) & EPS);
}

GL_HD uint64_t neg(uint64_t a) { return sub(0, a); }

// (lo + hi * 2^64) mod p, weak output.
GL_HD uint64_t reduce128(uint64_t lo, uint64_t hi) {
  uint64_t hh = hi >> 32, hl = hi & EPS;
  // hi * 2^64 = hh * 2^96 + hl * 2^64 == -hh + hl * (2^32 - 1)
  uint64_t t0 = lo - hh;
  t0 -= (0 - (uint64_t)(lo < hh)) & EPS;     // borrow: cannot wrap twice
  uint64_t t1 = (hl << 32) - hl;             // hl * 2^64 == hl * (2^32 - 1)
  uint64_t r = t0 + t1;
  r += (0 - (uint64_t)(r < t1)) & EPS;       // carry: cannot wrap twice
  return r;
}

GL_HD void mul_full(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  uint64_t p00 = (uint64_t)a0 * b0;
  uint64_t mid = (uint64_t)a0 * b1 + (p00 >> 32);
  uint64_t mid2 = (uint64_t)a1 * b0 + (uint32_t)mid;
  lo = (mid2 << 32) | (uint32_t)p00;
  hi = (uint64_t)a1 * b1 + (mid >> 32) + (mid2 >> 32);
#else
  unsigned __int128 m = (unsigned __int128)a * b;
  lo = (uint64_t)m;
  hi = (uint64_t)(m >> 64);
#endif
}

GL_HD uint64_t mul(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  mul_full(a, b, lo, hi);
  return reduce128(lo, hi);
}

// x * 2^s for 0 <= s < 96 using V = x << s = v0 + v1*2^64 + v2*2^96 with
// V == v0 + v1*(2^32-1) - v2 (mod p).
GL_HD uint64_t shl96(uint64_t x, uint32_t s) {
  uint64_t v0, v1, v2;
  if (s == 0) return x;
  if (s < 64) {
    v0 = x << s;
    uint64_t hi = x >> (64 - s);
    v1 = hi & EPS;
    v2 = hi >> 32;
  } else {
    v0 = 0;
    v1 = (x << (s - 64)) & EPS;
    v2 = x >> (96 - s);
  }
  return add(sub(v0, v2), (v1 << 32) - v1);
}

// x * 2^s for any 0 <= s < 192 (2^96 == -1), s known per call site.
GL_HD uint64_t mul_pow2(uint64_t x, uint32_t s) {
  s %= 192;
  if (s >= 96) return neg(shl96(x, s - 96));
  return shl96(x, s);
}

// Compile-time shift amount (the DFT-8/DFT-4 twiddles).
template <int S>
GL_HD uint64_t mul_pow2c(uint64_t x) {
  constexpr int s = ((S % 192) + 192) % 192;
  if constexpr (s == 0) {
    return x;
  } else if constexpr (s >= 96) {
    return neg(shl96(x, (uint32_t)(s - 96)));
  } else {
    return shl96(x, (uint32_t)s);
  }
}

// Branch-free runtime x * 2^s for 0 <= s < 192 where s varies across lanes.
GL_HD uint64_t mul_pow2_rt(uint64_t x, uint32_t s) {
  uint32_t negf = s >= 96;
  uint32_t s2 = negf ? s - 96 : s;           // 0..95
  uint32_t lo_sh = s2 & 63;                  // valid when s2 < 64
  uint64_t v0 = (s2 < 64) ? (x << lo_sh) : 0;
  uint64_t hi_a = (s2 == 0) ? 0 : (x >> ((64 - s2) & 63));           // s2 in 1..63
  uint64_t v1a = hi_a & EPS, v2a = hi_a >> 32;
  uint64_t v1b = (x << ((s2 - 64) & 63)) & EPS;                       // s2 in 64..95
  uint64_t v2b = x >> ((96 - s2) & 63);
  uint64_t v1 = (s2 < 64) ? v1a : v1b;
  uint64_t v2 = (s2 < 64) ? v2a : v2b;
  uint64_t r = add(sub(v0, v2), (v1 << 32) - v1);
  return negf ? neg(r) : r;
}

GL_HD uint64_t pow(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = mul(r, a);
    a = mul(a, a);
    e >>= 1;
  }
  return canonical(r);
}

GL_HD uint64_t inv(uint64_t a) { return pow(a, P - 2); }

// Digit lift: signed integer -> field element.
GL_HD uint64_t from_signed(int64_t d) { return d < 0 ? P - (uint64_t)(-d) : (uint64_t)d; }

// p -> 2^64 modulus switch of the reference (mod_switch_to_2powerN.sv:60-64,132-135,
// PRECISION_W = 96 per pe_pbs.sv:38): z = bits[159:96] of e*C + bit 95, with
// C = floor(2^160 / p) = 2^96 + 2^64 - 2, so e*C = (e*2^32 + e)*2^64 - 2e needs no
// multiplier.  e must be canonical (< p).
GL_HD uint64_t modswitch_p_to_q(uint64_t e) {
  // W = e*2^32 + e as (w_hi:w_lo), 97 bits
  uint64_t w_lo = (e << 32) + e;
  uint64_t w_hi = (e >> 32) + (uint64_t)(w_lo < e);
  // U = W - ceil(2e / 2^64) ... exact: V = W*2^64 - 2e ; U = V >> 64
  uint64_t two_e_hi = e >> 63, two_e_lo = e << 1;
  uint64_t t = two_e_hi + (uint64_t)(two_e_lo != 0);   // 0..2
  uint64_t u_lo = w_lo - t;
  uint64_t u_hi = w_hi - (uint64_t)(w_lo < t);
  // z = U[95:32] + U[31]
  uint64_t z = (u_lo >> 32) | (u_hi << 32);
  return z + ((u_lo >> 31) & 1);
}

}  // namespace gl
