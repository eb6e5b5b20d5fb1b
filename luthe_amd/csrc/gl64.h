// Goldilocks field arithmetic (p = 2^64 - 2^32 + 1) for the MI355X PBS engine.
//
// Field elements are kept canonical ([0, p)) so that every addition and subtraction
// needs exactly one conditional correction; products and shifts accept any 64-bit
// input and return canonical values.  Reductions use 2^64 = 2^32 - 1 and
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

// Reduction tails (add_eps_prod) of the four-lane reductions and of the batched shift twists:
// up to four lanes per asm block when a translation unit defines GL_GROUPED_TAILS to 1 before
// including this header, one block per lane otherwise.  The choice only changes the device
// instruction schedule; results are identical.
#ifndef GL_GROUPED_TAILS
#define GL_GROUPED_TAILS 0
#endif

namespace gl {

#if defined(__HIP_DEVICE_COMPILE__) && GL_GROUPED_TAILS
constexpr bool GROUPED_TAILS = true;
#else
constexpr bool GROUPED_TAILS = false;
#endif

constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p

GL_HD uint64_t canonical(uint64_t x) { return x >= P ? x - P : x; }

// canonical(x) as x + (EPS or 0): x - p == x + EPS (mod 2^64); the select picks a 32-bit addend
GL_HD uint64_t canonical_fold(uint64_t x) { return x + (uint64_t)(x > P - 1 ? 0xFFFFFFFFu : 0u); }

// a + b (mod p), canonical, for any 64-bit a, b with a + b < 2p: canonical a and b, or the
// reductions' any-64-bit a plus b <= p - 2^32.  When the 64-bit sum wrapped (it is then
// a + b - 2^64 < p - EPS) or is >= p, a + b - p is the 64-bit sum plus EPS (mod 2^64),
// and is < p.  The select picks the 32-bit addend (EPS or 0), not one of two 64-bit sums:
// one 32-bit select per add instead of two.
GL_HD uint64_t add(uint64_t a, uint64_t b) {
  uint64_t s;
  const bool wrapped = __builtin_add_overflow(a, b, &s);
  const uint32_t x = (wrapped | (s > P - 1)) ? 0xFFFFFFFFu : 0u;
  return s + (uint64_t)x;
}

// a + b (mod p) for a + b < 2p (canonical a and b), folded into 64 bits but not necessarily
// canonical: the wrap of the 64-bit sum is folded (it is then a + b - 2^64 < p - EPS, so adding
// EPS stays below 2^64), the result can lie in [p, 2^64).  For sums that only feed a product
// (mul, mul_full: any 64-bit operand): one compare and one select fewer than add().
GL_HD uint64_t add_wrap(uint64_t a, uint64_t b) {
  uint64_t s;
  const bool wrapped = __builtin_add_overflow(a, b, &s);
  return s + (uint64_t)(wrapped ? 0xFFFFFFFFu : 0u);
}

// add_wrap when W is nonzero, add otherwise (per-output choice in the DFT networks)
template <int W>
GL_HD uint64_t add_w(uint64_t a, uint64_t b) {
  if constexpr (W != 0) return add_wrap(a, b);
  else return add(a, b);
}

// t0 + hl * EPS (mod p), canonical, for any 64-bit t0 and hl < 2^32 (the tail of every
// reduction: hl * EPS <= p - 2^32, so the sum is < 2p and add()'s rule applies).  On the
// device the product and the sum are one v_mad_u64_u32, whose carry-out is the wrap flag;
// the addend EPS or 0 is then added by a second v_mad_u64_u32 (x * 1 + s).  The select
// mask is the SALU OR of the two VALU-written flags into a third register pair that no VALU
// writes, so the select may read it at once (only VALU-written SGPRs need wait states before
// a VALU reads them).
GL_HD uint64_t add_eps_prod(uint64_t t0, uint32_t hl) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t s, r, c, m, n, d;
  uint32_t x;
  asm("v_mad_u64_u32 %0, %3, %7, -1, %8\n\t"
      "v_cmp_lt_u64_e64 %4, %9, %0\n\t"
      "s_or_b64 %5, %4, %3\n\t"
      "v_cndmask_b32_e64 %2, 0, -1, %5\n\t"
      "v_mad_u64_u32 %1, %6, %2, 1, %0"
      : "=&v"(s), "=v"(r), "=&v"(x), "=&s"(c), "=&s"(m), "=&s"(n), "=&s"(d)
      : "v"(hl), "v"(t0), "s"(P - 1)
      : "scc");
  return r;
#else
  return add(t0, ((uint64_t)hl << 32) - hl);
#endif
}

// r_i = add_eps_prod(t0_i, hl_i) for N <= 4 independent lanes in one asm block, the lanes'
// steps interleaved.  One block instead of N: the compiler puts a wait state after an asm block
// whose results the next instruction touches, so N separate blocks in a row cost up to N of
// them.  Each select mask is the SALU OR into its own register pair, which no VALU writes.  The
// sums are formed in place over t0_i and the selects over hl_i (tied operands): no registers
// beyond the inputs.
template <int N>
GL_HD void add_eps_prodn(uint64_t* r, const uint64_t* t0, const uint32_t* hl) {
  static_assert(N >= 1 && N <= 4, "one to four lanes");
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (N == 1) {
    r[0] = add_eps_prod(t0[0], hl[0]);
  } else if constexpr (N == 2) {
    uint64_t r0 = t0[0], r1 = t0[1];
    uint32_t x0 = hl[0], x1 = hl[1];
    uint64_t c0, m0, n0, c1, m1, n1;
    asm("v_mad_u64_u32 %0, %4, %2, -1, %0\n\t"
        "v_mad_u64_u32 %1, %5, %3, -1, %1\n\t"
        "v_cmp_lt_u64_e64 %6, %10, %0\n\t"
        "v_cmp_lt_u64_e64 %7, %10, %1\n\t"
        "s_or_b64 %8, %6, %4\n\t"
        "s_or_b64 %9, %7, %5\n\t"
        "v_cndmask_b32_e64 %2, 0, -1, %8\n\t"
        "v_cndmask_b32_e64 %3, 0, -1, %9\n\t"
        "v_mad_u64_u32 %0, %4, %2, 1, %0\n\t"
        "v_mad_u64_u32 %1, %5, %3, 1, %1"
        : "+v"(r0), "+v"(r1), "+v"(x0), "+v"(x1),
          "=&s"(c0), "=&s"(c1), "=&s"(m0), "=&s"(m1), "=&s"(n0), "=&s"(n1)
        : "s"(P - 1)
        : "scc");
    r[0] = r0;
    r[1] = r1;
  } else if constexpr (N == 3) {
    uint64_t r0 = t0[0], r1 = t0[1], r2 = t0[2];
    uint32_t x0 = hl[0], x1 = hl[1], x2 = hl[2];
    uint64_t c0, m0, n0, c1, m1, n1, c2, m2, n2;
    asm("v_mad_u64_u32 %0, %6, %3, -1, %0\n\t"
        "v_mad_u64_u32 %1, %7, %4, -1, %1\n\t"
        "v_mad_u64_u32 %2, %8, %5, -1, %2\n\t"
        "v_cmp_lt_u64_e64 %9, %15, %0\n\t"
        "v_cmp_lt_u64_e64 %10, %15, %1\n\t"
        "v_cmp_lt_u64_e64 %11, %15, %2\n\t"
        "s_or_b64 %12, %9, %6\n\t"
        "s_or_b64 %13, %10, %7\n\t"
        "s_or_b64 %14, %11, %8\n\t"
        "v_cndmask_b32_e64 %3, 0, -1, %12\n\t"
        "v_cndmask_b32_e64 %4, 0, -1, %13\n\t"
        "v_cndmask_b32_e64 %5, 0, -1, %14\n\t"
        "v_mad_u64_u32 %0, %6, %3, 1, %0\n\t"
        "v_mad_u64_u32 %1, %7, %4, 1, %1\n\t"
        "v_mad_u64_u32 %2, %8, %5, 1, %2"
        : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(x0), "+v"(x1), "+v"(x2),
          "=&s"(c0), "=&s"(c1), "=&s"(c2), "=&s"(m0), "=&s"(m1), "=&s"(m2),
          "=&s"(n0), "=&s"(n1), "=&s"(n2)
        : "s"(P - 1)
        : "scc");
    r[0] = r0;
    r[1] = r1;
    r[2] = r2;
  } else {
    uint64_t r0 = t0[0], r1 = t0[1], r2 = t0[2], r3 = t0[3];
    uint32_t x0 = hl[0], x1 = hl[1], x2 = hl[2], x3 = hl[3];
    uint64_t c0, m0, n0, c1, m1, n1, c2, m2, n2, c3, m3, n3;
    asm("v_mad_u64_u32 %0, %8, %4, -1, %0\n\t"
        "v_mad_u64_u32 %1, %9, %5, -1, %1\n\t"
        "v_mad_u64_u32 %2, %10, %6, -1, %2\n\t"
        "v_mad_u64_u32 %3, %11, %7, -1, %3\n\t"
        "v_cmp_lt_u64_e64 %12, %20, %0\n\t"
        "v_cmp_lt_u64_e64 %13, %20, %1\n\t"
        "v_cmp_lt_u64_e64 %14, %20, %2\n\t"
        "v_cmp_lt_u64_e64 %15, %20, %3\n\t"
        "s_or_b64 %16, %12, %8\n\t"
        "s_or_b64 %17, %13, %9\n\t"
        "s_or_b64 %18, %14, %10\n\t"
        "s_or_b64 %19, %15, %11\n\t"
        "v_cndmask_b32_e64 %4, 0, -1, %16\n\t"
        "v_cndmask_b32_e64 %5, 0, -1, %17\n\t"
        "v_cndmask_b32_e64 %6, 0, -1, %18\n\t"
        "v_cndmask_b32_e64 %7, 0, -1, %19\n\t"
        "v_mad_u64_u32 %0, %8, %4, 1, %0\n\t"
        "v_mad_u64_u32 %1, %9, %5, 1, %1\n\t"
        "v_mad_u64_u32 %2, %10, %6, 1, %2\n\t"
        "v_mad_u64_u32 %3, %11, %7, 1, %3"
        : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3),
          "=&s"(c0), "=&s"(c1), "=&s"(c2), "=&s"(c3), "=&s"(m0), "=&s"(m1), "=&s"(m2), "=&s"(m3),
          "=&s"(n0), "=&s"(n1), "=&s"(n2), "=&s"(n3)
        : "s"(P - 1)
        : "scc");
    r[0] = r0;
    r[1] = r1;
    r[2] = r2;
    r[3] = r3;
  }
#else
  for (int i = 0; i < N; i++) r[i] = add_eps_prod(t0[i], hl[i]);
#endif
}

// add_eps_prodn over any number of lanes, four per block (r may alias t0)
template <int N>
GL_HD void add_eps_prod_all(uint64_t* r, const uint64_t* t0, const uint32_t* hl) {
  if constexpr (N > 4) {
    add_eps_prodn<4>(r, t0, hl);
    add_eps_prod_all<N - 4>(r + 4, t0 + 4, hl + 4);
  } else {
    add_eps_prodn<N>(r, t0, hl);
  }
}

// a - b with the borrow folded back as 2^64 == EPS: (a - b) mod 2^64, minus EPS when the
// subtraction borrowed.  For canonical a, b this is a - b (mod p), canonical (a - b + p
// when a < b); the reductions below use it on operands where the fold cannot wrap.
//
// On the device: d = a - b with borrow w (carry chain into VCC), then d - w*EPS as
// lo + w, hi - w + carry(lo + w) -- five VALU operations and no select.  Two wait states
// separate every carry written to an SGPR from its reader.
GL_HD uint64_t sub_fold(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t lo, hi;
  uint64_t c;
  asm("v_sub_co_u32 %0, vcc, %3, %5\n\t"
      "s_nop 1\n\t"
      "v_subb_co_u32 %1, vcc, %4, %6, vcc\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %0, %2, %0, 0, vcc\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %1, %2, %1, 0, %2\n\t"
      "v_subb_co_u32 %1, vcc, %1, 0, vcc"
      : "=&v"(lo), "=&v"(hi), "=&s"(c)
      : "v"((uint32_t)a), "v"((uint32_t)(a >> 32)), "v"((uint32_t)b), "v"((uint32_t)(b >> 32))
      : "vcc");
  return ((uint64_t)hi << 32) | lo;
#else
  uint64_t d;
  const bool borrow = __builtin_sub_overflow(a, b, &d);
  return borrow ? d - EPS : d;
#endif
}

GL_HD uint64_t sub(uint64_t a, uint64_t b) { return sub_fold(a, b); }

// Four independent r_i = sub_fold(a_i, b_i) (a_i - b_i (mod p) for canonical operands): the
// sub_fold sequence with the four carry chains interleaved, each chain on its own SGPR pairs.
// Every carry is read three instructions after it is written, which covers the two wait
// states without s_nop (sub_fold needs three of them).
GL_HD void sub4(uint64_t& r0, uint64_t& r1, uint64_t& r2, uint64_t& r3, uint64_t a0, uint64_t a1,
                uint64_t a2, uint64_t a3, uint64_t b0, uint64_t b1, uint64_t b2, uint64_t b3) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t l0, h0, l1, h1, l2, h2, l3, h3;
  uint64_t w0, w1, w2, w3, c0, c1, c2, c3;
  asm("v_sub_co_u32 %0, %8, %16, %24\n\t"
      "v_sub_co_u32 %2, %9, %18, %26\n\t"
      "v_sub_co_u32 %4, %10, %20, %28\n\t"
      "v_sub_co_u32 %6, %11, %22, %30\n\t"
      "v_subb_co_u32 %1, %8, %17, %25, %8\n\t"
      "v_subb_co_u32 %3, %9, %19, %27, %9\n\t"
      "v_subb_co_u32 %5, %10, %21, %29, %10\n\t"
      "v_subb_co_u32 %7, %11, %23, %31, %11\n\t"
      "v_addc_co_u32 %0, %12, %0, 0, %8\n\t"
      "v_addc_co_u32 %2, %13, %2, 0, %9\n\t"
      "v_addc_co_u32 %4, %14, %4, 0, %10\n\t"
      "v_addc_co_u32 %6, %15, %6, 0, %11\n\t"
      "v_addc_co_u32 %1, %12, %1, 0, %12\n\t"
      "v_addc_co_u32 %3, %13, %3, 0, %13\n\t"
      "v_addc_co_u32 %5, %14, %5, 0, %14\n\t"
      "v_addc_co_u32 %7, %15, %7, 0, %15\n\t"
      "v_subb_co_u32 %1, %8, %1, 0, %8\n\t"
      "v_subb_co_u32 %3, %9, %3, 0, %9\n\t"
      "v_subb_co_u32 %5, %10, %5, 0, %10\n\t"
      "v_subb_co_u32 %7, %11, %7, 0, %11"
      : "=&v"(l0), "=&v"(h0), "=&v"(l1), "=&v"(h1), "=&v"(l2), "=&v"(h2), "=&v"(l3), "=&v"(h3),
        "=&s"(w0), "=&s"(w1), "=&s"(w2), "=&s"(w3), "=&s"(c0), "=&s"(c1), "=&s"(c2), "=&s"(c3)
      : "v"((uint32_t)a0), "v"((uint32_t)(a0 >> 32)), "v"((uint32_t)a1), "v"((uint32_t)(a1 >> 32)),
        "v"((uint32_t)a2), "v"((uint32_t)(a2 >> 32)), "v"((uint32_t)a3), "v"((uint32_t)(a3 >> 32)),
        "v"((uint32_t)b0), "v"((uint32_t)(b0 >> 32)), "v"((uint32_t)b1), "v"((uint32_t)(b1 >> 32)),
        "v"((uint32_t)b2), "v"((uint32_t)(b2 >> 32)), "v"((uint32_t)b3), "v"((uint32_t)(b3 >> 32)));
  r0 = ((uint64_t)h0 << 32) | l0;
  r1 = ((uint64_t)h1 << 32) | l1;
  r2 = ((uint64_t)h2 << 32) | l2;
  r3 = ((uint64_t)h3 << 32) | l3;
#else
  r0 = sub_fold(a0, b0);
  r1 = sub_fold(a1, b1);
  r2 = sub_fold(a2, b2);
  r3 = sub_fold(a3, b3);
#endif
}

// Two and three independent sub_folds with their carry chains interleaved.  With two chains
// each carry is read two instructions after it is written when an s_nop 0 sits between the
// steps (three s_nop 0 instead of six s_nop 1 for two separate sub_folds); with three chains
// the other two chains' steps cover the wait states.
GL_HD void sub2(uint64_t& r0, uint64_t& r1, uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t l0, h0, l1, h1;
  uint64_t w0, w1, c0, c1;
  asm("v_sub_co_u32 %0, %4, %8, %12\n\t"
      "v_sub_co_u32 %2, %5, %10, %14\n\t"
      "s_nop 0\n\t"
      "v_subb_co_u32 %1, %4, %9, %13, %4\n\t"
      "v_subb_co_u32 %3, %5, %11, %15, %5\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32 %0, %6, %0, 0, %4\n\t"
      "v_addc_co_u32 %2, %7, %2, 0, %5\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32 %1, %6, %1, 0, %6\n\t"
      "v_addc_co_u32 %3, %7, %3, 0, %7\n\t"
      "v_subb_co_u32 %1, %4, %1, 0, %4\n\t"
      "v_subb_co_u32 %3, %5, %3, 0, %5"
      : "=&v"(l0), "=&v"(h0), "=&v"(l1), "=&v"(h1), "=&s"(w0), "=&s"(w1), "=&s"(c0), "=&s"(c1)
      : "v"((uint32_t)a0), "v"((uint32_t)(a0 >> 32)), "v"((uint32_t)a1), "v"((uint32_t)(a1 >> 32)),
        "v"((uint32_t)b0), "v"((uint32_t)(b0 >> 32)), "v"((uint32_t)b1), "v"((uint32_t)(b1 >> 32)));
  r0 = ((uint64_t)h0 << 32) | l0;
  r1 = ((uint64_t)h1 << 32) | l1;
#else
  r0 = sub_fold(a0, b0);
  r1 = sub_fold(a1, b1);
#endif
}

GL_HD void sub3(uint64_t& r0, uint64_t& r1, uint64_t& r2, uint64_t a0, uint64_t a1, uint64_t a2, uint64_t b0,
                uint64_t b1, uint64_t b2) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t l0, h0, l1, h1, l2, h2;
  uint64_t w0, w1, w2, c0, c1, c2;
  asm("v_sub_co_u32 %0, %6, %12, %18\n\t"
      "v_sub_co_u32 %2, %7, %14, %20\n\t"
      "v_sub_co_u32 %4, %8, %16, %22\n\t"
      "v_subb_co_u32 %1, %6, %13, %19, %6\n\t"
      "v_subb_co_u32 %3, %7, %15, %21, %7\n\t"
      "v_subb_co_u32 %5, %8, %17, %23, %8\n\t"
      "v_addc_co_u32 %0, %9, %0, 0, %6\n\t"
      "v_addc_co_u32 %2, %10, %2, 0, %7\n\t"
      "v_addc_co_u32 %4, %11, %4, 0, %8\n\t"
      "v_addc_co_u32 %1, %9, %1, 0, %9\n\t"
      "v_addc_co_u32 %3, %10, %3, 0, %10\n\t"
      "v_addc_co_u32 %5, %11, %5, 0, %11\n\t"
      "v_subb_co_u32 %1, %6, %1, 0, %6\n\t"
      "v_subb_co_u32 %3, %7, %3, 0, %7\n\t"
      "v_subb_co_u32 %5, %8, %5, 0, %8"
      : "=&v"(l0), "=&v"(h0), "=&v"(l1), "=&v"(h1), "=&v"(l2), "=&v"(h2), "=&s"(w0), "=&s"(w1), "=&s"(w2),
        "=&s"(c0), "=&s"(c1), "=&s"(c2)
      : "v"((uint32_t)a0), "v"((uint32_t)(a0 >> 32)), "v"((uint32_t)a1), "v"((uint32_t)(a1 >> 32)),
        "v"((uint32_t)a2), "v"((uint32_t)(a2 >> 32)), "v"((uint32_t)b0), "v"((uint32_t)(b0 >> 32)),
        "v"((uint32_t)b1), "v"((uint32_t)(b1 >> 32)), "v"((uint32_t)b2), "v"((uint32_t)(b2 >> 32)));
  r0 = ((uint64_t)h0 << 32) | l0;
  r1 = ((uint64_t)h1 << 32) | l1;
  r2 = ((uint64_t)h2 << 32) | l2;
#else
  r0 = sub_fold(a0, b0);
  r1 = sub_fold(a1, b1);
  r2 = sub_fold(a2, b2);
#endif
}

// r[i] = sub_fold(a[i], b[i]) for i < NS, in interleaved groups of four, three and two (a
// lone sub_fold only for NS = 1).
template <int NS>
GL_HD void subn(uint64_t* r, const uint64_t* a, const uint64_t* b) {
  if constexpr (NS == 1) {
    r[0] = sub_fold(a[0], b[0]);
  } else if constexpr (NS == 2) {
    sub2(r[0], r[1], a[0], a[1], b[0], b[1]);
  } else if constexpr (NS == 3) {
    sub3(r[0], r[1], r[2], a[0], a[1], a[2], b[0], b[1], b[2]);
  } else if constexpr (NS == 5 || NS == 6) {
    sub3(r[0], r[1], r[2], a[0], a[1], a[2], b[0], b[1], b[2]);
    subn<NS - 3>(r + 3, a + 3, b + 3);
  } else if constexpr (NS >= 4) {
    sub4(r[0], r[1], r[2], r[3], a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]);
    subn<NS - 4>(r + 4, a + 4, b + 4);
  }
}

GL_HD uint64_t neg(uint64_t a) { return sub(0, a); }

// (lo + hi * 2^64) mod p, canonical.  hi * 2^64 = hh * 2^96 + hl * 2^64 == -hh + hl * EPS.
// The final sum t0 + hl * EPS has t0 < 2^64 and hl * EPS <= p - 2^32, so one add() both
// folds its carry and canonicalises.
GL_HD uint64_t reduce128(uint64_t lo, uint64_t hi) {
  uint64_t hh = hi >> 32;
  uint64_t t0 = sub_fold(lo, hh);  // borrowed 2^64 == EPS; hh < 2^32: cannot wrap again
  return add_eps_prod(t0, (uint32_t)hi);  // + hl * EPS, hl * EPS <= p - 2^32
}

// (lo + hl * 2^64) mod p for hl < 2^32, canonical: hl * 2^64 == hl * EPS (the 2^96 part is
// zero, so no subtraction).
GL_HD uint64_t reduce96(uint64_t lo, uint64_t hl) { return add_eps_prod(lo, (uint32_t)hl); }

// (lo + hi * 2^64 + c * 2^128) mod p for c < 2^32 - 1, canonical: 2^128 == -2^32, so c
// joins the 2^96 part of hi as the high word of the subtrahend.
GL_HD uint64_t reduce160(uint64_t lo, uint64_t hi, uint32_t c) {
  uint64_t hh = (hi >> 32) | ((uint64_t)c << 32), hl = hi & EPS;
  uint64_t t0 = sub_fold(lo, hh);  // hh < 2^64 - 2^32: cannot wrap again
  return add_eps_prod(t0, (uint32_t)hl);
}

// (lo, hi, c) += (plo, phi): a 128-bit add whose carry past 2^128 is counted in c.  On the
// device one carry chain through the four words into c (five VALU operations, no compares).
GL_HD void acc128(uint64_t& lo, uint64_t& hi, uint32_t& c, uint64_t plo, uint64_t phi) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t l0 = (uint32_t)lo, l1 = (uint32_t)(lo >> 32), h0 = (uint32_t)hi, h1 = (uint32_t)(hi >> 32);
  asm("v_add_co_u32 %0, vcc, %0, %5\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %1, vcc, %1, %6, vcc\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %2, vcc, %2, %7, vcc\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %3, vcc, %3, %8, vcc\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %4, vcc, %4, 0, vcc"
      // early-clobber: an input may hold the same value as a word written before it is read
      : "+&v"(l0), "+&v"(l1), "+&v"(h0), "+&v"(h1), "+&v"(c)
      : "v"((uint32_t)plo), "v"((uint32_t)(plo >> 32)), "v"((uint32_t)phi), "v"((uint32_t)(phi >> 32))
      : "vcc");
  lo = ((uint64_t)l1 << 32) | l0;
  hi = ((uint64_t)h1 << 32) | h0;
#else
  uint64_t nlo, t1, nhi;
  const bool k1 = __builtin_add_overflow(lo, plo, &nlo);
  const bool o1 = __builtin_add_overflow(hi, phi, &t1);
  const bool o2 = __builtin_add_overflow(t1, (uint64_t)k1, &nhi);
  lo = nlo, hi = nhi;
  c += (uint32_t)o1 + (uint32_t)o2;
#endif
}

// acc128 over four lanes with the four carry chains interleaved: every carry is read three
// instructions after it is written, so no s_nop is needed (acc128 needs four of them).
GL_HD void acc128x4(uint64_t* lo, uint64_t* hi, uint32_t* c, const uint64_t* plo, const uint64_t* phi) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t w[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
    w[i][0] = (uint32_t)lo[i], w[i][1] = (uint32_t)(lo[i] >> 32), w[i][2] = (uint32_t)hi[i],
    w[i][3] = (uint32_t)(hi[i] >> 32);
  uint64_t k0, k1, k2, k3;
  asm("v_add_co_u32 %0, %20, %0, %24\n\t"
      "v_add_co_u32 %4, %21, %4, %28\n\t"
      "v_add_co_u32 %8, %22, %8, %32\n\t"
      "v_add_co_u32 %12, %23, %12, %36\n\t"
      "v_addc_co_u32 %1, %20, %1, %25, %20\n\t"
      "v_addc_co_u32 %5, %21, %5, %29, %21\n\t"
      "v_addc_co_u32 %9, %22, %9, %33, %22\n\t"
      "v_addc_co_u32 %13, %23, %13, %37, %23\n\t"
      "v_addc_co_u32 %2, %20, %2, %26, %20\n\t"
      "v_addc_co_u32 %6, %21, %6, %30, %21\n\t"
      "v_addc_co_u32 %10, %22, %10, %34, %22\n\t"
      "v_addc_co_u32 %14, %23, %14, %38, %23\n\t"
      "v_addc_co_u32 %3, %20, %3, %27, %20\n\t"
      "v_addc_co_u32 %7, %21, %7, %31, %21\n\t"
      "v_addc_co_u32 %11, %22, %11, %35, %22\n\t"
      "v_addc_co_u32 %15, %23, %15, %39, %23\n\t"
      "v_addc_co_u32 %16, %20, %16, 0, %20\n\t"
      "v_addc_co_u32 %17, %21, %17, 0, %21\n\t"
      "v_addc_co_u32 %18, %22, %18, 0, %22\n\t"
      "v_addc_co_u32 %19, %23, %19, 0, %23"
      // early-clobber: an input may hold the same value as a word written before it is read
      : "+&v"(w[0][0]), "+&v"(w[0][1]), "+&v"(w[0][2]), "+&v"(w[0][3]), "+&v"(w[1][0]), "+&v"(w[1][1]),
        "+&v"(w[1][2]), "+&v"(w[1][3]), "+&v"(w[2][0]), "+&v"(w[2][1]), "+&v"(w[2][2]), "+&v"(w[2][3]),
        "+&v"(w[3][0]), "+&v"(w[3][1]), "+&v"(w[3][2]), "+&v"(w[3][3]), "+&v"(c[0]), "+&v"(c[1]), "+&v"(c[2]),
        "+&v"(c[3]), "=&s"(k0), "=&s"(k1), "=&s"(k2), "=&s"(k3)
      : "v"((uint32_t)plo[0]), "v"((uint32_t)(plo[0] >> 32)), "v"((uint32_t)phi[0]), "v"((uint32_t)(phi[0] >> 32)),
        "v"((uint32_t)plo[1]), "v"((uint32_t)(plo[1] >> 32)), "v"((uint32_t)phi[1]), "v"((uint32_t)(phi[1] >> 32)),
        "v"((uint32_t)plo[2]), "v"((uint32_t)(plo[2] >> 32)), "v"((uint32_t)phi[2]), "v"((uint32_t)(phi[2] >> 32)),
        "v"((uint32_t)plo[3]), "v"((uint32_t)(plo[3] >> 32)), "v"((uint32_t)phi[3]), "v"((uint32_t)(phi[3] >> 32)));
#pragma unroll
  for (int i = 0; i < 4; i++) {
    lo[i] = ((uint64_t)w[i][1] << 32) | w[i][0];
    hi[i] = ((uint64_t)w[i][3] << 32) | w[i][2];
  }
#else
#pragma unroll
  for (int i = 0; i < 4; i++) acc128(lo[i], hi[i], c[i], plo[i], phi[i]);
#endif
}

// reduce128 over four lanes, the four borrow folds in one sub4.
GL_HD void reduce128x4(uint64_t* r, const uint64_t* lo, const uint64_t* hi) {
  uint64_t t[4];
  sub4(t[0], t[1], t[2], t[3], lo[0], lo[1], lo[2], lo[3], hi[0] >> 32, hi[1] >> 32, hi[2] >> 32,
       hi[3] >> 32);
  const uint32_t hl[4] = {(uint32_t)hi[0], (uint32_t)hi[1], (uint32_t)hi[2], (uint32_t)hi[3]};
  if constexpr (GROUPED_TAILS) {
    add_eps_prodn<4>(r, t, hl);
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++) r[i] = add_eps_prod(t[i], hl[i]);
  }
}

// reduce160 over four lanes (c_i < 2^32 - 1), the four borrow folds in one sub4.
GL_HD void reduce160x4(uint64_t* r, const uint64_t* lo, const uint64_t* hi, const uint32_t* c) {
  uint64_t t[4], hh[4];
#pragma unroll
  for (int i = 0; i < 4; i++) hh[i] = (hi[i] >> 32) | ((uint64_t)c[i] << 32);
  sub4(t[0], t[1], t[2], t[3], lo[0], lo[1], lo[2], lo[3], hh[0], hh[1], hh[2], hh[3]);
  const uint32_t hl[4] = {(uint32_t)hi[0], (uint32_t)hi[1], (uint32_t)hi[2], (uint32_t)hi[3]};
  if constexpr (GROUPED_TAILS) {
    add_eps_prodn<4>(r, t, hl);
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++) r[i] = add_eps_prod(t[i], hl[i]);
  }
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

// a * b mod p for any 64-bit a, b; canonical result.
GL_HD uint64_t mul(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  mul_full(a, b, lo, hi);
  return reduce128(lo, hi);
}

// r_i = a_i * b_i mod p for four lanes (r may alias a or b).
GL_HD void mul4(uint64_t* r, const uint64_t* a, const uint64_t* b) {
  uint64_t lo[4], hi[4];
#pragma unroll
  for (int i = 0; i < 4; i++) mul_full(a[i], b[i], lo[i], hi[i]);
  reduce128x4(r, lo, hi);
}

// x * 2^s mod p for 0 <= s < 96 and any 64-bit x; canonical result.
//  s < 64 : x << s is the 128-bit (x << s, x >> (64 - s)), reduced as above.
//  s >= 64: x << s = v1 * 2^64 + v2 * 2^96 with v1 = (x << (s - 64)) mod 2^32 and
//           v2 = x >> (96 - s), == v1 * EPS - v2 (mod p); v1 * EPS < p.
//  s <= 32 needs no 2^96 part (x >> (64 - s) < 2^32).  x >> (96 - s) < 2^63 < p.
GL_HD uint64_t shl96(uint64_t x, uint32_t s) {
  if (s == 0) return canonical(x);
  if (s <= 32) return reduce96(x << s, x >> (64 - s));
  if (s < 64) return reduce128(x << s, x >> (64 - s));
  uint64_t v1 = (x << (s - 64)) & EPS;
  uint64_t v2 = x >> (96 - s);
  return sub((v1 << 32) - v1, v2);
}

// x * 2^s for 0 <= s < 192 (2^96 == -1); s uniform or per lane.
GL_HD uint64_t mul_pow2(uint64_t x, uint32_t s) {
  if (s >= 96) return neg(shl96(x, s - 96));
  return shl96(x, s);
}

// Compile-time shift amount (the DFT and four-step twiddles), x canonical.  S is taken
// mod 192; S mod 192 >= 96 is the negated shift by S - 96.  For shifts in [64, 96) the
// result is v1 * EPS - v2, so the negated form is the swapped subtraction at no cost.
template <int S>
GL_HD uint64_t mul_pow2c(uint64_t x) {
  constexpr int s = ((S % 192) + 192) % 192;
  constexpr bool negate = s >= 96;
  constexpr int e = negate ? s - 96 : s;
  if constexpr (e == 0) {
    return negate ? neg(x) : x;
  } else if constexpr (e >= 64) {
    const uint64_t v1 = (x << (e - 64)) & EPS;
    const uint64_t v2 = x >> (96 - e);
    const uint64_t v1e = (v1 << 32) - v1;
    return negate ? sub(v2, v1e) : sub(v1e, v2);
  } else {
    const uint64_t r = (e <= 32) ? reduce96(x << e, x >> (64 - e)) : reduce128(x << e, x >> (64 - e));
    return negate ? neg(r) : r;
  }
}

// mul_pow2c over several lanes, v[i] *= 2^S_i, with the lanes' sub_folds issued together
// (subn): first the reductions of shifts in (32, 64) and the subtractions of shifts in
// [64, 96), then the negations of shifts in [96, 192) whose remainder is below 64.  Same
// results as mul_pow2c lane by lane.
struct Pw2 {
  int e;     // shift mod 96
  bool neg;  // S mod 192 >= 96
  int kind;  // 0: e = 0, 1: e <= 32, 2: e < 64, 3: e < 96
};
GL_HD constexpr Pw2 pw2_of(int S) {
  const int s = ((S % 192) + 192) % 192;
  const bool ng = s >= 96;
  const int e = ng ? s - 96 : s;
  return Pw2{e, ng, e == 0 ? 0 : e <= 32 ? 1 : e < 64 ? 2 : 3};
}
template <int... S>
GL_HD void mul_pow2c_n(uint64_t* v) {
  constexpr int K = sizeof...(S);
  constexpr Pw2 I[K] = {pw2_of(S)...};
  constexpr int N1 = ((pw2_of(S).kind >= 2 ? 1 : 0) + ... + 0);
  constexpr int N2 = ((pw2_of(S).neg && pw2_of(S).kind <= 2 ? 1 : 0) + ... + 0);
  uint64_t a[N1 > 0 ? N1 : 1], b[N1 > 0 ? N1 : 1], t[N1 > 0 ? N1 : 1];
  uint32_t hl[K];
  int j = 0;
#pragma unroll
  for (int i = 0; i < K; i++) {
    const uint64_t x = v[i];
    const int e = I[i].e;
    hl[i] = 0;
    if (I[i].kind == 1) {
      if constexpr (!GROUPED_TAILS) v[i] = reduce96(x << e, x >> (64 - e));
    } else if (I[i].kind == 2) {
      a[j] = x << e;
      b[j] = x >> (96 - e);  // the 2^96 part of the 128-bit x << e
      hl[i] = (uint32_t)(x >> (64 - e));
      j++;
    } else if (I[i].kind == 3) {
      const uint64_t v1 = (x << (e - 64)) & EPS, v2 = x >> (96 - e), v1e = (v1 << 32) - v1;
      a[j] = I[i].neg ? v2 : v1e;
      b[j] = I[i].neg ? v1e : v2;
      j++;
    }
  }
  if constexpr (N1 > 0) subn<N1>(t, a, b);
  j = 0;
  if constexpr (GROUPED_TAILS) {
    // the reduction tails of kind-1 and kind-2 lanes together, up to four per asm block
    constexpr int NE = ((pw2_of(S).kind == 1 || pw2_of(S).kind == 2 ? 1 : 0) + ... + 0);
    uint64_t et[NE > 0 ? NE : 1];
    uint32_t eh[NE > 0 ? NE : 1];
    int q = 0;
#pragma unroll
    for (int i = 0; i < K; i++) {
      if (I[i].kind == 1) et[q] = v[i] << I[i].e, eh[q++] = (uint32_t)(v[i] >> (64 - I[i].e));
      else if (I[i].kind == 2) et[q] = t[j++], eh[q++] = hl[i];
      else if (I[i].kind == 3) v[i] = t[j++];
    }
    if constexpr (NE > 0) add_eps_prod_all<NE>(et, et, eh);
    q = 0;
#pragma unroll
    for (int i = 0; i < K; i++)
      if (I[i].kind == 1 || I[i].kind == 2) v[i] = et[q++];
  } else {
#pragma unroll
    for (int i = 0; i < K; i++) {
      if (I[i].kind == 2) v[i] = add_eps_prod(t[j++], hl[i]);
      else if (I[i].kind == 3) v[i] = t[j++];
    }
  }
  if constexpr (N2 > 0) {
    uint64_t z[N2], nb[N2], r[N2];
    j = 0;
#pragma unroll
    for (int i = 0; i < K; i++)
      if (I[i].neg && I[i].kind <= 2) z[j] = 0, nb[j++] = v[i];
    subn<N2>(r, z, nb);
    j = 0;
#pragma unroll
    for (int i = 0; i < K; i++)
      if (I[i].neg && I[i].kind <= 2) v[i] = r[j++];
  }
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
GL_HD uint64_t from_signed32(int32_t d) {
  const uint64_t s = (uint64_t)(int64_t)d;  // d + 2^64 when negative; + p - 2^64 = - EPS
  return d < 0 ? s - EPS : s;
}

// v mod p for a signed 64-bit v, canonical: (uint64)v = 2^64 + v when negative, and
// 2^64 + v - EPS = p + v.  -EPS == 1 - 2^32 (mod 2^64) is the word pair (1, 2^32 - 1), i.e.
// (b, m) with m the sign word of v and b = m & 1: one 64-bit add, no compare or select.
GL_HD uint64_t lift64(int64_t v) {
  const uint32_t m = (uint32_t)(v >> 63);
  return (uint64_t)v + (((uint64_t)m << 32) | (m & 1u));
}

// lift64(d * f) for a signed 32-bit d and 0 < f < 2^31: the sign word of d * f is that of d,
// so on the device the product, its sign extension and the lift's addend are one
// v_mad_i64_i32 (d * f + (m & 1, m)), with no 64-bit shift of a sign-extended d.
GL_HD uint64_t lift_mul(int32_t d, int32_t f) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint32_t m = (uint32_t)(d >> 31);
  const uint64_t addend = ((uint64_t)m << 32) | (m & 1u);
  uint64_t r, c;
  asm("v_mad_i64_i32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(c) : "v"(d), "s"(f), "v"(addend));
  return r;
#else
  return lift64((int64_t)d * f);
#endif
}

// d * 2^(12 M1) mod p for a digit |d| <= 2^22, canonical: the forward NTT's pre-twist
// psi^(4 m1) (psi = 2^3) fused into the digit lift.  Products that fit a signed 64-bit word
// are lifted directly (2^12, 2^24, 2^36 = (d 2^8) 2^28); 2^72 is (d 2^8) * 2^64 ==
// (d 2^8) * EPS, also exact in 64 bits; 2^48, 2^60 and 2^84 are an exact product times a
// constant shift.
template <int M1>
GL_HD uint64_t pretwist_digit(int32_t d) {
  const int64_t D = d;
  if constexpr (M1 == 0) return lift_mul(d, 1);
  else if constexpr (M1 == 1) return lift_mul(d, 1 << 12);
  else if constexpr (M1 == 2) return lift_mul(d, 1 << 24);
  else if constexpr (M1 == 3) return lift_mul(d * (1 << 8), 1 << 28);
  else if constexpr (M1 == 4) return mul_pow2c<24>(lift_mul(d, 1 << 24));
  else if constexpr (M1 == 5) return mul_pow2c<24>(lift_mul(d * (1 << 8), 1 << 28));
  else if constexpr (M1 == 6) {
    const int64_t w = D * (1ll << 8);  // |w| <= 2^30: w * EPS fits
    return lift64(w * (1ll << 32) - w);
  } else {
    return mul_pow2c<72>(lift_mul(d, 1 << 12));
  }
}

// p -> 2^64 modulus switch of the reference (mod_switch_to_2powerN.sv:60-64,132-135,
// PRECISION_W = 96 per pe_pbs.sv:38): z = bits[159:96] of e*C + bit 95 (mod 2^64), with
// C = floor(2^160 / p) = 2^96 + 2^64 - 2.  For e = eh*2^32 + el, e*C + 2^95 over 2^96 is
// e + eh + (el*2^64 + 2^95 - 2e) / 2^96 with the last term in [0, 2), so
//   z = e + eh + [el*2^64 - 2e >= 2^95] = e + eh + [el > 2^31 + [e > 2^63]].
// e must be canonical (< p); then eh + bit < 2^32 (eh = 2^32 - 1 only for e = p - 1).
GL_HD uint64_t modswitch_p_to_q(uint64_t e) {
  const uint32_t el = (uint32_t)e, eh = (uint32_t)(e >> 32);
  const uint32_t thr = 0x80000000u + (uint32_t)(e > 0x8000000000000000ull);
  return e + (uint64_t)(eh + (uint32_t)(el > thr));
}

}  // namespace gl
