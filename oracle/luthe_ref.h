/*
 * luthe_ref.h -- CPU oracle for the TFHE PBS path (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain-C restatement of the reference's bit-exact algorithm for the
 * programmable bootstrap (KS -> modulus switch -> blind rotation -> sample
 * extract) and its building blocks.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker: the product
 * (libluthe.so) never links, loads or calls it.
 *
 * Reference files restated (paths relative to the reference checkout):
 *   Goldilocks arithmetic ..... hw/common_lib/mod_arithmetic/.../mod_reduct_64bgoldilocks_karatsuba.sv:8-22,112-114
 *   NTT / INTT ................ sw/ntt/ntt_gf64.sage:31-510, sw/ntt/ntt_gf64_network.py:20-53
 *   decomposition ............. sw/decomposer/balanced_decomp_unfix_sign_round_bit.py:42-105,
 *                               hw/module/decomposer/module/decomp_parallel/rtl/decomp_parallel.sv:56-150
 *   key switch ................ hw/module/pep_key_switch/.../pep_ks_blram_format.sv:159-173,236-246,
 *                               .../pep_ks_mult_node.sv:93-105, .../pep_ks_out_process.sv:276-282,336-341
 *   rotation / CMUX input ..... sw/monomial_mult/monomial_mult.py:156-176,
 *                               hw/module/pep_mono_mult_acc/.../pep_mmacc_splitc_feed_read.sv:484-488,
 *                               .../pep_mmacc_splitc_feed_final.sv:203-330
 *   p -> 2^64 switch .......... hw/common_lib/mod_arithmetic/module/mod_switch/module/mod_switch_to_2powerN/
 *                               rtl/mod_switch_to_2powerN.sv:56-147 with PRECISION_W = 96 (pe_pbs.sv:38)
 *   sample extract ............ hw/module/pep_mono_mult_acc/simu/tb_pep_mmacc_sample_extract/rtl/
 *                               tb_pep_mmacc_sample_extract.sv:624-712
 *   ALU DOps .................. hw/module/pe_alu/rtl/pea_alu_core.sv:55-160, pe_alu.sv:320-326,465-475
 *
 * Parity pinning status of each function is listed in oracle/README.md.
 */
#ifndef LUTHE_REF_H
#define LUTHE_REF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REF_P 0xFFFFFFFF00000001ULL /* Goldilocks prime 2^64 - 2^32 + 1 */

typedef struct {
  uint32_t lwe_dim;      /* n: small LWE dimension (LWE_K) */
  uint32_t glwe_k;       /* k: GLWE dimension */
  uint32_t poly_size;    /* N */
  uint32_t pbs_base_log; /* PBS_B_W */
  uint32_t pbs_level;    /* PBS_L */
  uint32_t ks_base_log;  /* KS_B_W */
  uint32_t ks_level;     /* KS_L */
  uint32_t ks_mod_log;   /* MOD_KSK_W */
} ref_params;

/* ---- Goldilocks field ---- */
uint64_t ref_gl_reduce128(uint64_t lo, uint64_t hi); /* canonical (lo + hi*2^64) mod p */
uint64_t ref_gl_add(uint64_t a, uint64_t b);
uint64_t ref_gl_sub(uint64_t a, uint64_t b);
uint64_t ref_gl_mul(uint64_t a, uint64_t b);
uint64_t ref_gl_pow(uint64_t a, uint64_t e);
uint64_t ref_gl_inv(uint64_t a);
uint64_t ref_gl_shl(uint64_t a, uint32_t k); /* a * 2^k mod p */
/* power_of_rou(root, power): OMG_2_32^(power * 2^32 / root), power may be negative */
uint64_t ref_power_of_rou(uint64_t root, int64_t power);

/* ---- NTT following ntt_gf64.sage (cut list: first cut negacyclic) ----
 * ref_ntt_sage: input p in bit-reversed order (as the sage model), output the
 * last radix column of ntt().  ref_intt_sage: input NTT-domain, output the last
 * column of intt() (bit-reversed coefficient order). */
int ref_ntt_sage(uint32_t N, const uint32_t* cuts, uint32_t ncut, const uint64_t* in_rev, uint64_t* out);
int ref_intt_sage(uint32_t N, const uint32_t* cuts, uint32_t ncut, const uint64_t* in, uint64_t* out_rev);
/* Natural-order wrappers: ntt(p) = ntt_sage(rev(p)), intt(P) = rev(intt_sage(P)). */
int ref_ntt(uint32_t N, const uint32_t* cuts, uint32_t ncut, const uint64_t* p, uint64_t* out);
int ref_intt(uint32_t N, const uint32_t* cuts, uint32_t ncut, const uint64_t* in, uint64_t* p);
/* phi tables of ntt_phi(cut_l, bwd) for list entry idx (N values). */
int ref_ntt_phi(uint32_t N, const uint32_t* cuts, uint32_t ncut, int bwd, uint32_t idx, uint64_t* out);
/* Batched config-1 core: out[b] = intt(ntt(a[b]) (.) bn[b]) for b < nb, cut [5,6] at N=2048. */
int ref_ntt_polymul_batch(uint32_t N, const uint32_t* cuts, uint32_t ncut, size_t nb,
                          const uint64_t* a, const uint64_t* bn, uint64_t* out, int threads);
/* Schoolbook negacyclic product mod p (for checking the NTT on small N). */
void ref_negacyclic_mul(uint32_t N, const uint64_t* a, const uint64_t* b, uint64_t* out);

/* ---- decomposition (Appendix: closest representable + balanced signed digits) ----
 * digits[0] is the LSB level, digits[L-1] the MSB level. */
void ref_decompose(uint64_t x, uint32_t w, uint32_t base_log, uint32_t level, int32_t* digits);
void ref_decompose_batch(const uint64_t* x, size_t n, uint32_t w, uint32_t base_log, uint32_t level, int32_t* digits);

/* ---- modulus switches ---- */
uint64_t ref_modsw_p_to_2_64(uint64_t e);                  /* mod_switch_to_2powerN, PRECISION_W=96 */
uint32_t ref_ks_body_switch(uint64_t b, uint32_t ks_mod_log); /* (b >> (64-w)) + bit(64-w-1) mod 2^w */
uint32_t ref_ks_to_2n(uint32_t v, uint32_t ks_mod_log, uint32_t log2_2n);

/* ---- monomial rotation and CMUX input ---- */
/* out = X^e * in  (negacyclic, mod 2^64), e in [0, 2N) */
void ref_monomial_mult(uint32_t N, const uint64_t* in, uint32_t e, uint64_t* out);

/* ---- key switch ----
 * blwe: k*N mask + body (u64).  ksk: [k*N][ks_level][n+1] u32 (values mod 2^ks_mod_log).
 * out_lwe: [n+1] values mod 2^ks_mod_log; out_2n: [n+1] values in [0, 2N). */
void ref_keyswitch(const ref_params* prm, const uint64_t* blwe, const uint32_t* ksk,
                   uint32_t* out_lwe, uint16_t* out_2n);

/* ---- blind rotation + sample extract ----
 * bsk_ntt: [n][(k+1)][L][(k+1)][N] u64, NTT domain (ntt() order, cut list cuts).
 * lut: (k+1)*N u64 (GLWE, mask polys then body poly).
 * lwe_2n: [n+1] (mask then body), values in [0,2N).
 * acc_out: (k+1)*N u64 (optional, may be NULL). */
void ref_blind_rotate(const ref_params* prm, const uint32_t* cuts, uint32_t ncut,
                      const uint64_t* bsk_ntt, const uint64_t* lut, const uint16_t* lwe_2n,
                      uint64_t* acc_out);
/* Sample extract with Many-LUT: writes lut_nb = 2^log_lut_nb BLWEs of k*N+1 u64. */
void ref_sample_extract(uint32_t k, uint32_t N, const uint64_t* acc, uint32_t b2n,
                        uint32_t log_lut_nb, uint64_t* out);
/* Full PBS DOp: KS -> BR -> SE.  out: lut_nb BLWEs. */
void ref_pbs(const ref_params* prm, const uint32_t* cuts, uint32_t ncut,
             const uint64_t* bsk_ntt, const uint32_t* ksk, const uint64_t* lut,
             const uint64_t* blwe_in, uint32_t log_lut_nb, uint64_t* out);
void ref_pbs_batch(const ref_params* prm, const uint32_t* cuts, uint32_t ncut,
                   const uint64_t* bsk_ntt, const uint32_t* ksk, const uint64_t* luts,
                   const uint32_t* lut_index, size_t nct, const uint64_t* blwe_in,
                   uint64_t* out, int threads);
/* Convert a coefficient-domain BSK [n][(k+1)L][(k+1)][N] mod p into ntt() order. */
void ref_bsk_to_ntt(const ref_params* prm, const uint32_t* cuts, uint32_t ncut,
                    const uint64_t* bsk_coef, uint64_t* bsk_ntt, int threads);

/* ---- ALU DOps on BLWE (k*N+1 u64), opcode = 6-bit DOp code; useful_bits = msg + carry + 1 ---- */
int ref_alu(uint32_t opcode, size_t blwe_len, const uint64_t* a0, const uint64_t* a1,
            uint32_t mul_factor, uint32_t msg_cst, uint32_t useful_bits, uint64_t* out);

#ifdef __cplusplus
}
#endif
#endif
