/*
 * luthe_ref.c -- CPU oracle for the TFHE PBS path (TEST INFRASTRUCTURE ONLY).
 *
 * Plain C restatement of the reference's bit-exact algorithms.  Each function
 * cites the reference file:line it follows.  This file is the checker for the
 * HIP path; it is never linked into, loaded by, or called from libluthe.so.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 */
#include "luthe_ref.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ==================================================================================== */
/* Goldilocks field                                                                      */
/* ==================================================================================== */

/* x = lo + hi*2^64 reduced with 2^64 = 2^32-1 and 2^96 = -1 (mod p), the identities
 * the Goldilocks reduction uses (mod_reduct_64bgoldilocks_karatsuba.sv:8-22,112-114:
 * msb = a[63:32]+a[95:64], lsb = a[31:0]-a[95:64]-a[..:96]).  Canonical result. */
uint64_t ref_gl_reduce128(uint64_t lo, uint64_t hi) {
  uint64_t hl = hi & 0xFFFFFFFFULL, hh = hi >> 32;
  __int128 t = (__int128)lo + (__int128)hl * 0xFFFFFFFFLL - (__int128)hh;
  const __int128 P = (__int128)REF_P;
  while (t < 0) t += P;
  while (t >= P) t -= P;
  return (uint64_t)t;
}

uint64_t ref_gl_add(uint64_t a, uint64_t b) {
  u128 s = (u128)(a % REF_P) + (b % REF_P);
  if (s >= REF_P) s -= REF_P;
  return (uint64_t)s;
}

uint64_t ref_gl_sub(uint64_t a, uint64_t b) {
  a %= REF_P;
  b %= REF_P;
  return a >= b ? a - b : (uint64_t)((u128)a + REF_P - b);
}

uint64_t ref_gl_mul(uint64_t a, uint64_t b) {
  u128 m = (u128)a * b;
  return ref_gl_reduce128((uint64_t)m, (uint64_t)(m >> 64));
}

uint64_t ref_gl_pow(uint64_t a, uint64_t e) {
  uint64_t r = 1, x = a % REF_P;
  while (e) {
    if (e & 1) r = ref_gl_mul(r, x);
    x = ref_gl_mul(x, x);
    e >>= 1;
  }
  return r;
}

uint64_t ref_gl_inv(uint64_t a) { return ref_gl_pow(a, REF_P - 2); }

static uint64_t g_pow2[193];
static pthread_once_t g_pow2_once = PTHREAD_ONCE_INIT;
static void init_pow2(void) {
  g_pow2[0] = 1;
  for (int i = 1; i <= 192; i++) g_pow2[i] = ref_gl_add(g_pow2[i - 1], g_pow2[i - 1]);
}

/* x << k in GF(p) (sage: IntegerMod << k multiplies by 2^k). */
uint64_t ref_gl_shl(uint64_t a, uint32_t k) {
  pthread_once(&g_pow2_once, init_pow2);
  return ref_gl_mul(a, g_pow2[k % 192]);
}

/* ntt_gf64.sage:44 OMG_2_32 and :69-75 power_of_rou(root, power) = OMG_2_32^(power*2^32/root) */
#define OMG_2_32 16334397945464290598ULL
uint64_t ref_power_of_rou(uint64_t root, int64_t power) {
  int64_t e = power * (int64_t)((1ULL << 32) / root);
  if (e >= 0) return ref_gl_pow(OMG_2_32, (uint64_t)e);
  return ref_gl_pow(ref_gl_inv(OMG_2_32), (uint64_t)(-e));
}

/* ==================================================================================== */
/* NTT model: ntt_gf64.sage:79-510, ntt_gf64_network.py:20-53                            */
/* ==================================================================================== */

static uint32_t ilog2u(uint64_t x) {
  uint32_t r = 0;
  while ((1ULL << r) < x) r++;
  return r;
}

/* ntt_gf64.sage:79-99 reverse_order(v, S) */
static uint32_t rev_bits(uint32_t v, uint32_t S) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < S; i++) r = (r << 1) | ((v >> i) & 1);
  return r;
}

#define MAX_CUTS 12
typedef struct {
  uint32_t N, S, ncut, cuts[MAX_CUTS];
  /* twiddles per processing column: [col][stage][R] shift amount and sign */
  uint16_t* fwd_pow2[MAX_CUTS];
  int8_t* fwd_sign[MAX_CUTS];
  uint16_t* bwd_pow2[MAX_CUTS];
  int8_t* bwd_sign[MAX_CUTS];
  /* phi per list entry: [N] general value, shift, sign */
  uint64_t* fwd_phi[MAX_CUTS];
  uint16_t* fwd_phi_pow2[MAX_CUTS];
  int8_t* fwd_phi_sign[MAX_CUTS];
  uint64_t* bwd_phi[MAX_CUTS];
  uint16_t* bwd_phi_pow2[MAX_CUTS];
  int8_t* bwd_phi_sign[MAX_CUTS];
} ntt_plan;

/* ntt_gf64.sage:236-270 ntt_twd + :104-127 convert_power_2.  Fills [stage][R]. */
static void build_twd(uint32_t R, int bwd, int ngc, uint16_t* pow2, int8_t* sign) {
  uint32_t logR = ilog2u(R);
  for (uint32_t s = 0; s < logR; s++) {
    uint32_t gp = 1u << (s + 1);
    /* bwd: the stage order is reversed, so stage s of the output uses the list of s' */
    uint32_t so = bwd ? (logR - 1 - s) : s;
    for (uint32_t i = 0; i < R; i++) {
      uint32_t v;
      if ((i % gp) < gp / 2) {
        v = 0;
      } else if (!ngc) {
        v = (i % (gp / 2)) * 64 / gp;
      } else {
        v = (1 + 2 * (i % (gp / 2))) * 64 / (2 * gp);
      }
      if (bwd) v = 64 - v;
      uint32_t e = v * 3;
      int8_t sg = 1;
      if (e >= 96) {
        sg = -1;
        e -= 96;
      }
      pow2[so * R + i] = (uint16_t)e;
      sign[so * R + i] = sg;
    }
  }
}

/* ntt_gf64.sage:132-186 ntt_phi_core -> values for index l*2^cut_end + a (R entries). */
static void build_phi_core(uint32_t R, uint32_t cut_start, uint32_t cut_end, int bwd, int ngc,
                           uint64_t* phi, uint16_t* pow2, int8_t* sign) {
  uint32_t pos = 0;
  for (uint32_t l = 0; l < (1u << cut_start); l++) {
    for (uint32_t a = 0; a < (1u << cut_end); a++) {
      uint64_t v;
      int64_t e;
      int8_t s = 1;
      if (bwd) {
        uint64_t ll = rev_bits(l, cut_start), aa = a;
        v = ref_power_of_rou(R, -(int64_t)(aa * ll));
        e = (int64_t)(((uint64_t)R - aa * ll) * 64 / R);
        s = ((e / 32) % 2) ? -1 : 1;
        e = 3 * (e % 32);
        if (ngc) v = ref_gl_mul(ref_gl_mul(v, ref_power_of_rou(2 * (uint64_t)R, -(int64_t)ll)), ref_gl_inv(R));
      } else {
        uint64_t ll = l, aa = rev_bits(a, cut_end);
        v = ref_power_of_rou(R, (int64_t)(aa * ll));
        e = (int64_t)(3 * (aa * ll * 64 / R));
        if (ngc) v = ref_gl_mul(v, ref_power_of_rou(2 * (uint64_t)R, (int64_t)aa));
      }
      phi[pos] = v;
      pow2[pos] = (uint16_t)e;
      sign[pos] = s;
      pos++;
    }
  }
}

static ntt_plan* plan_build(uint32_t N, const uint32_t* cuts, uint32_t ncut) {
  ntt_plan* P = (ntt_plan*)calloc(1, sizeof(ntt_plan));
  P->N = N;
  P->S = ilog2u(N);
  P->ncut = ncut;
  memcpy(P->cuts, cuts, ncut * sizeof(uint32_t));
  /* twiddles: fwd in cut order, bwd built in cut order then the list reversed
   * (ntt_gf64.sage:954-972) */
  for (uint32_t i = 0; i < ncut; i++) {
    uint32_t R = 1u << cuts[i], logR = cuts[i];
    P->fwd_pow2[i] = (uint16_t*)malloc(sizeof(uint16_t) * R * (logR ? logR : 1));
    P->fwd_sign[i] = (int8_t*)malloc(R * (logR ? logR : 1));
    build_twd(R, 0, i == 0, P->fwd_pow2[i], P->fwd_sign[i]);
    uint32_t bi = ncut - 1 - i;
    P->bwd_pow2[bi] = (uint16_t*)malloc(sizeof(uint16_t) * R * (logR ? logR : 1));
    P->bwd_sign[bi] = (int8_t*)malloc(R * (logR ? logR : 1));
    build_twd(R, 1, i == 0, P->bwd_pow2[bi], P->bwd_sign[bi]);
  }
  /* phis: ntt_gf64.sage:188-230 ntt_phi */
  for (int bwd = 0; bwd < 2; bwd++) {
    for (uint32_t i = 0; i + 1 < ncut; i++) {
      uint32_t cut_acc = 0;
      for (uint32_t c = i + 1; c < ncut; c++) cut_acc += cuts[c];
      uint32_t STG = cut_acc + cuts[i];
      uint32_t R = 1u << STG;
      uint64_t* ph = (uint64_t*)malloc(sizeof(uint64_t) * N);
      uint16_t* pw = (uint16_t*)malloc(sizeof(uint16_t) * N);
      int8_t* sg = (int8_t*)malloc(N);
      if (bwd)
        build_phi_core(R, cut_acc, cuts[i], 1, i == 0, ph, pw, sg);
      else
        build_phi_core(R, cuts[i], cut_acc, 0, i == 0, ph, pw, sg);
      for (uint32_t rep = 1; rep < N / R; rep++) { /* (N//R) * p_l : list repetition */
        memcpy(ph + rep * R, ph, R * sizeof(uint64_t));
        memcpy(pw + rep * R, pw, R * sizeof(uint16_t));
        memcpy(sg + rep * R, sg, R);
      }
      uint32_t idx = bwd ? (ncut - 2 - i) : i; /* bwd list is reversed */
      if (bwd) {
        P->bwd_phi[idx] = ph;
        P->bwd_phi_pow2[idx] = pw;
        P->bwd_phi_sign[idx] = sg;
      } else {
        P->fwd_phi[idx] = ph;
        P->fwd_phi_pow2[idx] = pw;
        P->fwd_phi_sign[idx] = sg;
      }
    }
  }
  return P;
}

static pthread_mutex_t g_plan_mu = PTHREAD_MUTEX_INITIALIZER;
static ntt_plan* g_plans[16];
static int g_nplans = 0;

static ntt_plan* plan_get(uint32_t N, const uint32_t* cuts, uint32_t ncut) {
  pthread_mutex_lock(&g_plan_mu);
  for (int i = 0; i < g_nplans; i++) {
    ntt_plan* p = g_plans[i];
    if (p->N == N && p->ncut == ncut && memcmp(p->cuts, cuts, ncut * sizeof(uint32_t)) == 0) {
      pthread_mutex_unlock(&g_plan_mu);
      return p;
    }
  }
  ntt_plan* p = plan_build(N, cuts, ncut);
  if (g_nplans < 16) g_plans[g_nplans++] = p;
  pthread_mutex_unlock(&g_plan_mu);
  return p;
}

/* Radix columns are only exact with omega_64-power twiddles: ngc cut <= 5, cyclic cut <= 6
 * (ntt_gf64.sage:46-51).  The phi tables alone are defined for any cut
 * (gen_gf64_rom_twd.sh:18 generates cut [4,7] ROMs), hence phi_only. */
static int check_cuts_ex(uint32_t N, const uint32_t* cuts, uint32_t ncut, int phi_only) {
  if (ncut == 0 || ncut > MAX_CUTS) return -1;
  uint32_t s = 0;
  for (uint32_t i = 0; i < ncut; i++) {
    if (cuts[i] == 0 || cuts[i] > 16) return -1;
    if (!phi_only && i == 0 && cuts[i] > 5) return -1; /* NGC_MAX_R = 32 */
    if (!phi_only && i > 0 && cuts[i] > 6) return -1;  /* CYC_MAX_R = 64 */
    s += cuts[i];
  }
  if ((1u << s) != N) return -1;
  return 0;
}
/* The inverse carries 1/N in the ngc phi (ntt_gf64.sage:168-169), so a transform needs >= 2 cuts. */
static int check_cuts(uint32_t N, const uint32_t* cuts, uint32_t ncut) {
  return ncut < 2 ? -1 : check_cuts_ex(N, cuts, ncut, 0);
}

/* ntt_gf64_network.py:20-53 ref_network (the hardware network() is checked equal to it
 * by that script's own self-check, ntt_gf64_network.py:306-355). */
static void ref_network(const ntt_plan* P, const uint64_t* data, int bwd, uint32_t rdx_col_id,
                        uint64_t* out) {
  uint32_t N = P->N, ncut = P->ncut, s_l = 0, rdx_1;
  if (bwd) {
    uint32_t cid = ncut - 1 - rdx_col_id;
    for (uint32_t i = 0; i < cid + 2; i++) s_l += P->cuts[ncut - 1 - i];
    rdx_1 = 1u << P->cuts[rdx_col_id - 1];
  } else {
    for (uint32_t i = rdx_col_id; i < ncut; i++) s_l += P->cuts[i];
    rdx_1 = (1u << s_l) >> P->cuts[rdx_col_id];
  }
  uint32_t n_l = 1u << s_l, pos = 0;
  for (uint32_t b = 0; b < N / n_l; b++)
    for (uint32_t r1 = 0; r1 < n_l / rdx_1; r1++)
      for (uint32_t i = 0; i < rdx_1; i++) out[pos++] = data[b * n_l + i * (n_l / rdx_1) + r1];
}

/* ntt_gf64.sage:275-314 ntt_radix (Cooley-Tukey, Rev/Nat) on one block */
static void ntt_radix(uint32_t R, uint64_t* x, const uint16_t* pow2, const int8_t* sign) {
  uint32_t logR = ilog2u(R);
  for (uint32_t s = 0; s < logR; s++) {
    uint32_t gp = 1u << (s + 1);
    for (uint32_t r = 0; r < R / gp; r++)
      for (uint32_t i = 0; i < gp / 2; i++) {
        uint32_t i1 = r * gp + i, i2 = i1 + gp / 2;
        uint64_t a = x[i1];
        uint64_t b = ref_gl_shl(x[i2], pow2[s * R + i2]);
        if (sign[s * R + i2] < 0) b = ref_gl_sub(0, b);
        x[i1] = ref_gl_add(a, b);
        x[i2] = ref_gl_sub(a, b);
      }
  }
}

/* ntt_gf64.sage:319-362 intt_radix (Gentleman-Sande, Nat/Rev) on one block */
static void intt_radix(uint32_t R, uint64_t* x, const uint16_t* pow2, const int8_t* sign) {
  uint32_t logR = ilog2u(R);
  for (uint32_t s = 0; s < logR; s++) {
    uint32_t gp = 1u << (logR - s);
    for (uint32_t r = 0; r < R / gp; r++)
      for (uint32_t i = 0; i < gp / 2; i++) {
        uint32_t i1 = r * gp + i, i2 = i1 + gp / 2;
        uint64_t a = ref_gl_add(x[i1], x[i2]);
        uint64_t b = ref_gl_sub(x[i1], x[i2]);
        b = ref_gl_shl(b, pow2[s * R + i2]);
        if (sign[s * R + i2] < 0) b = ref_gl_sub(0, b);
        x[i1] = a;
        x[i2] = b;
      }
  }
}

/* ntt_gf64.sage:367-445 ntt() : returns the last radix column */
static void ntt_core(const ntt_plan* P, const uint64_t* in_rev, uint64_t* out) {
  uint32_t N = P->N, ncut = P->ncut;
  uint64_t* cur = (uint64_t*)malloc(sizeof(uint64_t) * N);
  uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * N);
  for (uint32_t x = 0; x < N; x++) cur[x] = in_rev[x] % REF_P;
  for (uint32_t i = 0; i < ncut; i++) {
    uint32_t R = 1u << P->cuts[i];
    if (i > 0) {
      uint32_t ss = 0;
      for (uint32_t c = i; c < ncut; c++) ss += P->cuts[c];
      uint32_t blk_R = 1u << (ss + P->cuts[i - 1]);
      ref_network(P, cur, 0, i - 1, tmp);
      if (i == 1 || blk_R > 64) {
        for (uint32_t x = 0; x < N; x++) cur[x] = ref_gl_mul(tmp[x], P->fwd_phi[i - 1][x]);
      } else {
        for (uint32_t x = 0; x < N; x++) {
          uint64_t v = ref_gl_shl(tmp[x], P->fwd_phi_pow2[i - 1][x]);
          cur[x] = P->fwd_phi_sign[i - 1][x] < 0 ? ref_gl_sub(0, v) : v;
        }
      }
    }
    for (uint32_t j = 0; j < N / R; j++) ntt_radix(R, cur + j * R, P->fwd_pow2[i], P->fwd_sign[i]);
  }
  memcpy(out, cur, sizeof(uint64_t) * N);
  free(cur);
  free(tmp);
}

/* ntt_gf64.sage:450-510 intt() : returns the last radix column (bit-reversed order) */
static void intt_core(const ntt_plan* P, const uint64_t* in, uint64_t* out_rev) {
  uint32_t N = P->N, ncut = P->ncut;
  uint64_t* cur = (uint64_t*)malloc(sizeof(uint64_t) * N);
  uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * N);
  for (uint32_t x = 0; x < N; x++) cur[x] = in[x] % REF_P;
  for (uint32_t i = 0; i < ncut; i++) {
    uint32_t c = P->cuts[ncut - 1 - i]; /* rev_cut_l[i] */
    uint32_t R = 1u << c;
    int ngc = (i == ncut - 1);
    uint32_t rdx_col_id = ncut - 1 - i;
    if (i > 0) {
      uint32_t s = 0;
      for (uint32_t q = 0; q <= i; q++) s += P->cuts[ncut - 1 - q];
      uint32_t blk_R = 1u << s;
      ref_network(P, cur, 1, rdx_col_id + 1, tmp);
      if (ngc || blk_R > 64) {
        for (uint32_t x = 0; x < N; x++) cur[x] = ref_gl_mul(tmp[x], P->bwd_phi[i - 1][x]);
      } else {
        for (uint32_t x = 0; x < N; x++) {
          uint64_t v = ref_gl_shl(tmp[x], P->bwd_phi_pow2[i - 1][x]);
          cur[x] = P->bwd_phi_sign[i - 1][x] < 0 ? ref_gl_sub(0, v) : v;
        }
      }
    }
    for (uint32_t j = 0; j < N / R; j++) intt_radix(R, cur + j * R, P->bwd_pow2[i], P->bwd_sign[i]);
  }
  memcpy(out_rev, cur, sizeof(uint64_t) * N);
  free(cur);
  free(tmp);
}

int ref_ntt_sage(uint32_t N, const uint32_t* cuts, uint32_t ncut, const uint64_t* in_rev, uint64_t* out) {
  if (check_cuts(N, cuts, ncut)) return -1;
  ntt_core(plan_get(N, cuts, ncut), in_rev, out);
  return 0;
}

int ref_intt_sage(uint32_t N, const uint32_t* cuts, uint32_t ncut, const uint64_t* in, uint64_t* out_rev) {
  if (check_cuts(N, cuts, ncut)) return -1;
  intt_core(plan_get(N, cuts, ncut), in, out_rev);
  return 0;
}

int ref_ntt(uint32_t N, const uint32_t* cuts, uint32_t ncut, const uint64_t* p, uint64_t* out) {
  if (check_cuts(N, cuts, ncut)) return -1;
  uint32_t S = ilog2u(N);
  uint64_t* r = (uint64_t*)malloc(sizeof(uint64_t) * N);
  for (uint32_t i = 0; i < N; i++) r[i] = p[rev_bits(i, S)]; /* sage.test_convolution:539 */
  ntt_core(plan_get(N, cuts, ncut), r, out);
  free(r);
  return 0;
}

int ref_intt(uint32_t N, const uint32_t* cuts, uint32_t ncut, const uint64_t* in, uint64_t* p) {
  if (check_cuts(N, cuts, ncut)) return -1;
  uint32_t S = ilog2u(N);
  uint64_t* r = (uint64_t*)malloc(sizeof(uint64_t) * N);
  intt_core(plan_get(N, cuts, ncut), in, r);
  for (uint32_t i = 0; i < N; i++) p[i] = r[rev_bits(i, S)]; /* sage.test_convolution:565 */
  free(r);
  return 0;
}

int ref_ntt_phi(uint32_t N, const uint32_t* cuts, uint32_t ncut, int bwd, uint32_t idx, uint64_t* out) {
  if (check_cuts_ex(N, cuts, ncut, 1) || idx + 1 >= ncut) return -1;
  ntt_plan* P = plan_get(N, cuts, ncut);
  memcpy(out, bwd ? P->bwd_phi[idx] : P->fwd_phi[idx], sizeof(uint64_t) * N);
  return 0;
}

int ref_ntt_polymul_batch(uint32_t N, const uint32_t* cuts, uint32_t ncut, size_t nb,
                          const uint64_t* a, const uint64_t* bn, uint64_t* out, int threads) {
  if (check_cuts(N, cuts, ncut)) return -1;
  ntt_plan* P = plan_get(N, cuts, ncut);
  (void)P;
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 4)
  for (size_t b = 0; b < nb; b++) {
    uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t) * N);
    ref_ntt(N, cuts, ncut, a + b * N, t);
    for (uint32_t x = 0; x < N; x++) t[x] = ref_gl_mul(t[x], bn[b * N + x]);
    ref_intt(N, cuts, ncut, t, out + b * N);
    free(t);
  }
  return 0;
}

void ref_negacyclic_mul(uint32_t N, const uint64_t* a, const uint64_t* b, uint64_t* out) {
  for (uint32_t i = 0; i < N; i++) out[i] = 0;
  for (uint32_t i = 0; i < N; i++)
    for (uint32_t j = 0; j < N; j++) {
      uint64_t m = ref_gl_mul(a[i], b[j]);
      uint32_t d = i + j;
      if (d >= N)
        out[d - N] = ref_gl_sub(out[d - N], m);
      else
        out[d] = ref_gl_add(out[d], m);
    }
}

/* ==================================================================================== */
/* Decomposition: decomp_parallel.sv:56-150 (hw form of                                   */
/* balanced_decomp_unfix_sign_round_bit.py:42-105)                                        */
/* ==================================================================================== */
void ref_decompose(uint64_t x, uint32_t w, uint32_t b, uint32_t L, int32_t* digits) {
  uint32_t R = L * b;
  uint64_t wmask = (w >= 64) ? ~0ULL : ((1ULL << w) - 1);
  uint64_t rmask = (R >= 64) ? ~0ULL : ((1ULL << R) - 1);
  uint64_t bmask = (1ULL << b) - 1;
  uint64_t closest;
  int sign;
  x &= wmask;
  if (R >= w) { /* gen_entire (decomp_parallel.sv:76-79) */
    closest = x;
    sign = 0;
  } else {
    uint32_t ofs = w - R;
    uint64_t rb = (x >> (ofs - 1)) & 1;
    closest = ((x >> ofs) + rb) & rmask;
    uint64_t msb = (x >> (w - 1)) & 1;
    uint64_t below_mask = (R - 1 >= 64) ? ~0ULL : ((1ULL << (R - 1)) - 1);
    uint64_t below = (x >> ofs) & below_mask;
    /* decomp_parallel.sv:83-84 */
    sign = (msb && (below != 0 || rb)) || (!msb && below == below_mask && rb);
  }
  uint64_t u[65];
  for (uint32_t i = 0; i < L; i++) u[i] = (closest >> (i * b)) & bmask;
  u[L] = sign ? bmask : 0;
  uint64_t half_low_mask = (b >= 2) ? ((1ULL << (b - 1)) - 1) : 0;
  uint64_t c = 0;
  for (uint32_t i = 0; i < L; i++) { /* decomp_parallel.sv:104-148 */
    uint64_t t = u[i] + c;
    uint64_t half = (u[i] >> (b - 1)) & 1;
    uint64_t low = u[i] & half_low_mask;
    int gt = half && (low != 0 || c);
    int ge = half || (low == half_low_mask && c);
    int ov = (u[i] == bmask) && c;
    int prop = ov || gt || (ge && ((u[i + 1] >> (b - 1)) & 1));
    digits[i] = prop ? (int32_t)((int64_t)t - (int64_t)(1ULL << b)) : (int32_t)t;
    c = prop ? 1 : 0;
  }
}

void ref_decompose_batch(const uint64_t* x, size_t n, uint32_t w, uint32_t b, uint32_t L, int32_t* digits) {
  for (size_t i = 0; i < n; i++) ref_decompose(x[i], w, b, L, digits + i * L);
}

/* ==================================================================================== */
/* Modulus switches                                                                       */
/* ==================================================================================== */

/* mod_switch_to_2powerN.sv:60-64,132-135 with MOD_Q_W=64, MOD_P_W=64, PRECISION_W=96
 * (pe_pbs.sv:38 MODSW_2_PRECISION_W = MOD_NTT_W + 32):
 *   CST = floor(2^160 / p) (97 bits); z = bits[159:96] of a*CST, +1 if bit 95 is set. */
static uint64_t g_modsw_lo, g_modsw_hi;
static pthread_once_t g_modsw_once = PTHREAD_ONCE_INIT;
static void init_modsw(void) {
  /* CST = floor(2^160 / p) by long division over base-2^32 digits (most significant first) */
  const uint32_t num[6] = {1, 0, 0, 0, 0, 0};
  uint32_t q[6];
  u128 rem = 0;
  for (int i = 0; i < 6; i++) {
    rem = (rem << 32) | num[i];
    q[i] = (uint32_t)(rem / REF_P);
    rem = rem % REF_P;
  }
  /* q[5] holds bits 0..31, q[2] bits 96..127; CST < 2^97 so q[0] = q[1] = 0 */
  g_modsw_lo = ((uint64_t)q[4] << 32) | q[5];
  g_modsw_hi = ((uint64_t)q[2] << 32) | q[3];
}

uint64_t ref_modsw_p_to_2_64(uint64_t e) {
  pthread_once(&g_modsw_once, init_modsw);
  u128 p1 = (u128)e * g_modsw_lo;
  u128 m = (p1 >> 64) + (u128)e * g_modsw_hi; /* bits >= 64 of e*CST (< 2^98) */
  uint64_t z = (uint64_t)(m >> 32);           /* bits [159:96] */
  z += (uint64_t)((m >> 31) & 1);             /* bit 95 : rounding */
  return z;
}

/* pep_ks_blram_format.sv:236-246: b' = b[63 -: w] + b[63-w] */
uint32_t ref_ks_body_switch(uint64_t b, uint32_t w) {
  uint64_t top = b >> (64 - w);
  uint64_t rb = (b >> (64 - w - 1)) & 1;
  return (uint32_t)((top + rb) & ((1ULL << w) - 1));
}

/* pep_ks_out_process.sv:336-341: lwe[w-1 -: LWE_COEF_W] + lwe[w-1-LWE_COEF_W] */
uint32_t ref_ks_to_2n(uint32_t v, uint32_t w, uint32_t lc) {
  uint32_t top = v >> (w - lc);
  uint32_t rb = (v >> (w - lc - 1)) & 1;
  return (top + rb) & ((1u << lc) - 1);
}

/* ==================================================================================== */
/* Monomial multiplication: monomial_mult.py:156-176                                      */
/* ==================================================================================== */
void ref_monomial_mult(uint32_t N, const uint64_t* in, uint32_t e, uint64_t* out) {
  /* Equivalent closed form of repeated X-multiplication:  out[i] = +/- in[(i - e) mod N]
   * with the sign flipped once per wrap.  Checked against the repeated-shift definition
   * in tests (brute force over all 2N exponents). */
  e %= (2 * N);
  for (uint32_t i = 0; i < N; i++) {
    int64_t src = (int64_t)i - (int64_t)e;
    int neg = 0;
    while (src < 0) {
      src += N;
      neg ^= 1;
    }
    out[i] = neg ? (uint64_t)(0 - in[src]) : in[src];
  }
}

/* ==================================================================================== */
/* Key switch                                                                             */
/* ==================================================================================== */
void ref_keyswitch(const ref_params* prm, const uint64_t* blwe, const uint32_t* ksk,
                   uint32_t* out_lwe, uint16_t* out_2n) {
  uint32_t n = prm->lwe_dim, kN = prm->glwe_k * prm->poly_size, L = prm->ks_level;
  uint32_t w = prm->ks_mod_log;
  uint32_t lc = ilog2u(2 * prm->poly_size);
  uint32_t wmask = (w >= 32) ? 0xFFFFFFFFu : ((1u << w) - 1);
  uint32_t* acc = (uint32_t*)calloc(n + 1, sizeof(uint32_t));
  int32_t d[64];
  for (uint32_t i = 0; i < kN; i++) {
    /* pep_ks_blram_format.sv:159-173: decomp_parallel OP_W=64, L=KS_L, B_W=KS_B_W */
    ref_decompose(blwe[i], 64, prm->ks_base_log, L, d);
    for (uint32_t l = 0; l < L; l++) {
      if (d[l] == 0) continue;
      const uint32_t* row = ksk + ((size_t)i * L + l) * (n + 1);
      uint32_t dl = (uint32_t)d[l];
      /* pep_ks_mult_node.sv:93-105: digit x KSK mod 2^w, accumulated */
      for (uint32_t j = 0; j <= n; j++) acc[j] += dl * row[j];
    }
  }
  uint32_t bp = ref_ks_body_switch(blwe[kN], w);
  for (uint32_t j = 0; j <= n; j++) {
    /* pep_ks_out_process.sv:276-282: (0,..,0,b) - sum */
    uint32_t v = (((j == n) ? bp : 0u) - acc[j]) & wmask;
    if (out_lwe) out_lwe[j] = v;
    if (out_2n) out_2n[j] = (uint16_t)ref_ks_to_2n(v, w, lc);
  }
  free(acc);
}

/* ==================================================================================== */
/* Blind rotation                                                                         */
/* ==================================================================================== */
void ref_blind_rotate(const ref_params* prm, const uint32_t* cuts, uint32_t ncut,
                      const uint64_t* bsk_ntt, const uint64_t* lut, const uint16_t* lwe_2n,
                      uint64_t* acc_out) {
  uint32_t n = prm->lwe_dim, k = prm->glwe_k, N = prm->poly_size, L = prm->pbs_level;
  uint32_t K1 = k + 1;
  size_t poly = N;
  uint64_t* acc = (uint64_t*)malloc(sizeof(uint64_t) * K1 * poly);
  uint64_t* ct1 = (uint64_t*)malloc(sizeof(uint64_t) * poly);
  uint64_t* dig = (uint64_t*)malloc(sizeof(uint64_t) * K1 * L * poly);
  uint64_t* dn = (uint64_t*)malloc(sizeof(uint64_t) * K1 * L * poly);
  uint64_t* sum = (uint64_t*)malloc(sizeof(uint64_t) * poly);
  uint64_t* e = (uint64_t*)malloc(sizeof(uint64_t) * poly);
  int32_t d[64];
  memcpy(acc, lut, sizeof(uint64_t) * K1 * poly);
  for (uint32_t i = 0; i < n; i++) {
    uint32_t a = lwe_2n[i];
    /* feed: ct1 = X^a * ACC ; ct10 = ct1 - ACC (feed_read.sv:484-488, feed_final.sv:320-329) */
    for (uint32_t g = 0; g < K1; g++) {
      ref_monomial_mult(N, acc + g * poly, a, ct1);
      for (uint32_t m = 0; m < N; m++) {
        uint64_t ct10 = ct1[m] - acc[g * poly + m];
        /* decomp_balseq_core.sv:10-173: LSB level first, lifted to Z_p */
        ref_decompose(ct10, 64, prm->pbs_base_log, L, d);
        for (uint32_t l = 0; l < L; l++)
          dig[(g * L + l) * poly + m] = d[l] < 0 ? REF_P - (uint64_t)(-(int64_t)d[l]) : (uint64_t)d[l];
      }
    }
    for (uint32_t r = 0; r < K1 * L; r++) ref_ntt(N, cuts, ncut, dig + r * poly, dn + r * poly);
    const uint64_t* bski = bsk_ntt + (size_t)i * K1 * L * K1 * poly;
    for (uint32_t j = 0; j < K1; j++) {
      /* ntt_core_gf64_pp_core.sv: sum over (g,l) of NTT(D) x BSK, mod p */
      for (uint32_t x = 0; x < N; x++) {
        uint64_t s = 0;
        for (uint32_t g = 0; g < K1; g++)
          for (uint32_t l = 0; l < L; l++)
            s = ref_gl_add(s, ref_gl_mul(dn[(g * L + l) * poly + x], bski[((g * L + l) * K1 + j) * poly + x]));
        sum[x] = s;
      }
      ref_intt(N, cuts, ncut, sum, e); /* canonical in [0,p) */
      for (uint32_t m = 0; m < N; m++) acc[j * poly + m] += ref_modsw_p_to_2_64(e[m]); /* ct0 + pp_mod_q */
    }
  }
  if (acc_out) memcpy(acc_out, acc, sizeof(uint64_t) * K1 * poly);
  free(acc);
  free(ct1);
  free(dig);
  free(dn);
  free(sum);
  free(e);
}

/* tb_pep_mmacc_sample_extract.sv:646-690 */
void ref_sample_extract(uint32_t k, uint32_t N, const uint64_t* acc, uint32_t b2n, uint32_t log_lut_nb,
                        uint64_t* out) {
  uint32_t lut_nb = 1u << log_lut_nb;
  uint64_t* v = (uint64_t*)malloc(sizeof(uint64_t) * (k + 1) * N);
  for (uint32_t g = 0; g <= k; g++)
    for (uint32_t m = 0; m < N; m++) {
      uint32_t idx = m + b2n;
      int sg = (idx >= N && idx < 2 * N);
      uint64_t val = acc[g * N + (idx % N)];
      v[g * N + m] = sg ? (uint64_t)(0 - val) : val;
    }
  for (uint32_t lut = 0; lut < lut_nb; lut++) {
    uint32_t offs = lut * N / lut_nb;
    uint64_t* o = out + (size_t)lut * (k * N + 1);
    for (uint32_t g = 0; g < k; g++)
      for (uint32_t m = 0; m < N; m++) {
        int32_t ci = (int32_t)offs - (int32_t)m;
        o[g * N + m] = (ci < 0) ? (uint64_t)(0 - v[g * N + N + ci]) : v[g * N + ci];
      }
    o[k * N] = v[k * N + offs];
  }
  free(v);
}

void ref_pbs(const ref_params* prm, const uint32_t* cuts, uint32_t ncut, const uint64_t* bsk_ntt,
             const uint32_t* ksk, const uint64_t* lut, const uint64_t* blwe_in, uint32_t log_lut_nb,
             uint64_t* out) {
  uint32_t n = prm->lwe_dim, k = prm->glwe_k, N = prm->poly_size;
  uint16_t* a2n = (uint16_t*)malloc(sizeof(uint16_t) * (n + 1));
  uint64_t* acc = (uint64_t*)malloc(sizeof(uint64_t) * (k + 1) * N);
  ref_keyswitch(prm, blwe_in, ksk, NULL, a2n);
  ref_blind_rotate(prm, cuts, ncut, bsk_ntt, lut, a2n, acc);
  ref_sample_extract(k, N, acc, a2n[n], log_lut_nb, out);
  free(a2n);
  free(acc);
}

void ref_pbs_batch(const ref_params* prm, const uint32_t* cuts, uint32_t ncut, const uint64_t* bsk_ntt,
                   const uint32_t* ksk, const uint64_t* luts, const uint32_t* lut_index, size_t nct,
                   const uint64_t* blwe_in, uint64_t* out, int threads) {
  size_t blen = (size_t)prm->glwe_k * prm->poly_size + 1;
  size_t llen = (size_t)(prm->glwe_k + 1) * prm->poly_size;
  plan_get(prm->poly_size, cuts, ncut);
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 1)
  for (size_t c = 0; c < nct; c++)
    ref_pbs(prm, cuts, ncut, bsk_ntt, ksk, luts + (size_t)lut_index[c] * llen, blwe_in + c * blen, 0,
            out + c * blen);
}

void ref_bsk_to_ntt(const ref_params* prm, const uint32_t* cuts, uint32_t ncut, const uint64_t* bsk_coef,
                    uint64_t* bsk_ntt, int threads) {
  size_t np = (size_t)prm->lwe_dim * (prm->glwe_k + 1) * prm->pbs_level * (prm->glwe_k + 1);
  uint32_t N = prm->poly_size;
  plan_get(N, cuts, ncut);
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(static)
  for (size_t i = 0; i < np; i++) ref_ntt(N, cuts, ncut, bsk_coef + i * N, bsk_ntt + i * N);
}

/* ==================================================================================== */
/* ALU: pea_alu_core.sv:55-160.  The scalar constant sits in the top USEFUL_BIT bits of the  */
/* body: in_msg_cst[USEFUL_BIT-1:0] << (64 - USEFUL_BIT) (pea_alu_core.sv:59), USEFUL_BIT =   */
/* PAYLOAD_BIT + PADDING_BIT (param_tfhe_pkg.sv:32), i.e. msg + carry + 1 padding bit.      */
/* ==================================================================================== */
int ref_alu(uint32_t opcode, size_t len, const uint64_t* a0, const uint64_t* a1, uint32_t mul_factor,
            uint32_t msg_cst, uint32_t useful_bits, uint64_t* out) {
  if (useful_bits < 1 || useful_bits > 11) return -1;
  const uint64_t cst_body = ((uint64_t)msg_cst & ((1ull << useful_bits) - 1)) << (64 - useful_bits);
  for (size_t i = 0; i < len; i++) {
    int is_body = (i == len - 1);
    uint64_t cst = is_body ? cst_body : 0; /* body only: pe_alu.sv:469-475 */
    uint64_t x = a0[i];
    switch (opcode) {
      case 0x01: out[i] = x + a1[i]; break;                          /* ADD  */
      case 0x02: out[i] = x - a1[i]; break;                          /* SUB  */
      case 0x05: out[i] = x * (uint64_t)(mul_factor & 31u) + a1[i]; break; /* MAC  */
      case 0x09: out[i] = x + cst; break;                            /* ADDS */
      case 0x0A: out[i] = x - cst; break;                            /* SUBS */
      case 0x0B: out[i] = cst - x; break;                            /* SSUB */
      case 0x0C: out[i] = x * (uint64_t)(msg_cst & 31u); break;      /* MULS */
      default: return -1;
    }
  }
  return 0;
}
