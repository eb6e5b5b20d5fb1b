// Client-side helpers: seeded key generation, BLWE encryption/decryption and LUT
// encoding.  In the reference these live in tfhe-rs (outside the checkout); the
// conventions here are this engine's own and are documented in include/luthe.h:
//
//  * RNG: counter-based SplitMix64 streams, one per (seed, domain, index), so key
//    material is identical whatever the thread count.
//  * Noise: rounded Gaussian (Box-Muller on 53-bit uniforms), std given relative to
//    the modulus of the ciphertext being produced.
//  * GGSW rows (g, l) over Z_p: GLWE_p(0) + s_i * round(p / 2^(b (L - l))) added to
//    component g, coefficient 0 (l = 0 is the least significant gadget level,
//    matching the LSB-first digits of the PBS decomposition).
//  * KSK rows (i, l) over Z_(2^w): LWE(s_i * 2^(w - b (L - l))).
//  * LWE convention b = <a, s> + m + e.
#include <math.h>
#include <string.h>

#include <thread>
#include <vector>

#include "../../include/luthe.h"
#include "gl64.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t key, ctr = 0;
  Rng(uint64_t seed, uint64_t domain, uint64_t index) : key(mix64(seed ^ mix64(domain * 0x100000001B3ull + mix64(index)))) {}
  uint64_t next() { return mix64(key ^ mix64(++ctr)); }
  uint64_t uniform_p() {
    for (;;) {
      uint64_t v = next();
      if (v < gl::P) return v;
    }
  }
  double unit() {  // (0, 1]
    return ((double)(next() >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  }
  double gauss() {
    double u1 = unit(), u2 = unit();
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
};

enum : uint64_t { D_LWE_SK = 1, D_GLWE_SK = 2, D_BSK = 3, D_KSK = 4, D_ENC = 5 };

// Host negacyclic NTT mod p (radix-2, psi-twisted), used only to form A*S in keygen.
struct HostNtt {
  int n, logn;
  std::vector<uint64_t> psi_pow, ipsi_pow, w;
  uint64_t inv_n;
  explicit HostNtt(int n_) : n(n_) {
    logn = 0;
    while ((1 << logn) < n) logn++;
    const uint64_t OMG_2_32 = 16334397945464290598ull;
    uint64_t psi = gl::pow(OMG_2_32, (1ull << 32) / (2ull * n));
    uint64_t ipsi = gl::inv(psi);
    psi_pow.resize(n);
    ipsi_pow.resize(n);
    uint64_t a = 1, b = 1;
    for (int i = 0; i < n; i++) {
      psi_pow[i] = a;
      ipsi_pow[i] = b;
      a = gl::canonical(gl::mul(a, psi));
      b = gl::canonical(gl::mul(b, ipsi));
    }
    inv_n = gl::inv((uint64_t)n);
  }
  void dft(std::vector<uint64_t>& x, bool inverse) const {
    // iterative Cooley-Tukey with bit reversal, root = psi^2 (or its inverse)
    for (int i = 1, j = 0; i < n; i++) {
      int bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) std::swap(x[i], x[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
      // omega_len = psi^(2n/len)
      uint64_t wl = inverse ? ipsi_pow[(2 * n / len) % n] : psi_pow[(2 * n / len) % n];
      if (2 * n / len == n) wl = gl::P - 1;  // psi^n = -1
      for (int i = 0; i < n; i += len) {
        uint64_t wc = 1;
        for (int j = 0; j < len / 2; j++) {
          uint64_t u = x[i + j], v = gl::mul(x[i + j + len / 2], wc);
          x[i + j] = gl::canonical(gl::add(u, v));
          x[i + j + len / 2] = gl::canonical(gl::sub(u, v));
          wc = gl::canonical(gl::mul(wc, wl));
        }
      }
    }
  }
  void forward(std::vector<uint64_t>& x) const {
    for (int i = 0; i < n; i++) x[i] = gl::mul(x[i], psi_pow[i]);
    dft(x, false);
  }
  void inverse(std::vector<uint64_t>& x) const {
    dft(x, true);
    for (int i = 0; i < n; i++) x[i] = gl::canonical(gl::mul(gl::mul(x[i], ipsi_pow[i]), inv_n));
  }
};

int64_t round_noise(Rng& r, double std_abs) { return (int64_t)llround(r.gauss() * std_abs); }

}  // namespace

extern "C" {

int luthe_keygen_secret(const luthe_params* p, uint64_t seed, uint64_t* lwe_sk, uint64_t* glwe_sk) {
  if (!p || !lwe_sk || !glwe_sk) return LUTHE_ERR_ARG;
  Rng a(seed, D_LWE_SK, 0), b(seed, D_GLWE_SK, 0);
  for (uint32_t i = 0; i < p->lwe_dim; i++) lwe_sk[i] = a.next() & 1;
  for (uint32_t i = 0; i < p->glwe_k * p->poly_size; i++) glwe_sk[i] = b.next() & 1;
  return LUTHE_OK;
}

int luthe_keygen_bsk(const luthe_params* p, uint64_t seed, const uint64_t* lwe_sk, const uint64_t* glwe_sk,
                     uint64_t* bsk, int threads) {
  if (!p || !lwe_sk || !glwe_sk || !bsk) return LUTHE_ERR_ARG;
  const int N = (int)p->poly_size, k = (int)p->glwe_k, K1 = k + 1, L = (int)p->pbs_level;
  const uint32_t b = p->pbs_base_log;
  HostNtt ntt(N);
  std::vector<std::vector<uint64_t>> S(k, std::vector<uint64_t>(N));
  for (int g = 0; g < k; g++) {
    for (int m = 0; m < N; m++) S[g][m] = glwe_sk[g * N + m];
    ntt.forward(S[g]);
  }
  const double std_abs = p->glwe_noise_std * (double)gl::P;
  const size_t nrows = (size_t)p->lwe_dim * K1 * L;
  auto work = [&](size_t r0, size_t r1) {
    std::vector<uint64_t> acc(N), t(N);
    for (size_t row = r0; row < r1; row++) {
      size_t i = row / (K1 * L), g = (row / L) % K1, l = row % L;
      Rng rng(seed, D_BSK, row);
      uint64_t* out = bsk + row * K1 * N;  // [j][N]
      std::fill(acc.begin(), acc.end(), 0);
      for (int j = 0; j < k; j++) {
        for (int m = 0; m < N; m++) t[m] = out[j * N + m] = rng.uniform_p();
        ntt.forward(t);
        for (int m = 0; m < N; m++) acc[m] = gl::add(acc[m], gl::mul(t[m], S[j][m]));
      }
      ntt.inverse(acc);
      for (int m = 0; m < N; m++)
        out[(size_t)k * N + m] = gl::canonical(gl::add(acc[m], gl::from_signed(round_noise(rng, std_abs))));
      if (lwe_sk[i] & 1) {
        unsigned e = b * (unsigned)(L - (int)l);
        unsigned __int128 num = (unsigned __int128)gl::P + ((unsigned __int128)1 << (e - 1));
        uint64_t gad = (uint64_t)(num >> e);  // round(p / 2^e)
        out[g * N] = gl::canonical(gl::add(out[g * N], gad));
      }
    }
  };
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = (int)std::min<size_t>((size_t)nt, nrows);
  std::vector<std::thread> th;
  for (int w = 0; w < nt; w++) th.emplace_back(work, nrows * w / nt, nrows * (w + 1) / nt);
  for (auto& x : th) x.join();
  return LUTHE_OK;
}

int luthe_keygen_ksk(const luthe_params* p, uint64_t seed, const uint64_t* lwe_sk, const uint64_t* glwe_sk,
                     uint32_t* ksk) {
  if (!p || !lwe_sk || !glwe_sk || !ksk) return LUTHE_ERR_ARG;
  const uint32_t n = p->lwe_dim, L = p->ks_level, b = p->ks_base_log, w = p->ks_mod_log;
  const uint32_t kN = p->glwe_k * p->poly_size;
  const uint64_t mask = (w >= 32) ? 0xFFFFFFFFull : ((1ull << w) - 1);
  const double std_abs = p->lwe_noise_std * (double)(1ull << w);
  for (uint32_t i = 0; i < kN; i++)
    for (uint32_t l = 0; l < L; l++) {
      Rng rng(seed, D_KSK, (uint64_t)i * L + l);
      uint32_t* row = ksk + ((size_t)i * L + l) * (n + 1);
      uint64_t body = 0;
      for (uint32_t j = 0; j < n; j++) {
        uint64_t a = rng.next() & mask;
        row[j] = (uint32_t)a;
        body += a * (lwe_sk[j] & 1);
      }
      int64_t e = round_noise(rng, std_abs);
      uint64_t msg = (glwe_sk[i] & 1) ? (1ull << (w - b * (L - l))) : 0;
      row[n] = (uint32_t)((body + msg + (uint64_t)e) & mask);
    }
  return LUTHE_OK;
}

int luthe_encrypt_blwe(const luthe_params* p, uint64_t seed, const uint64_t* glwe_sk, const uint64_t* pt, size_t n,
                       uint64_t* blwe) {
  if (!p || !glwe_sk || (n && (!pt || !blwe))) return LUTHE_ERR_ARG;
  const size_t kN = (size_t)p->glwe_k * p->poly_size;
  const double std_abs = p->glwe_noise_std * 18446744073709551616.0;
  for (size_t c = 0; c < n; c++) {
    Rng rng(seed, D_ENC, c);
    uint64_t* o = blwe + c * (kN + 1);
    uint64_t body = 0;
    for (size_t i = 0; i < kN; i++) {
      o[i] = rng.next();
      body += o[i] * (glwe_sk[i] & 1);
    }
    o[kN] = body + pt[c] + (uint64_t)round_noise(rng, std_abs);
  }
  return LUTHE_OK;
}

int luthe_decrypt_blwe(const luthe_params* p, const uint64_t* glwe_sk, const uint64_t* blwe, size_t n,
                       uint64_t* phase) {
  if (!p || !glwe_sk || (n && (!blwe || !phase))) return LUTHE_ERR_ARG;
  const size_t kN = (size_t)p->glwe_k * p->poly_size;
  for (size_t c = 0; c < n; c++) {
    const uint64_t* o = blwe + c * (kN + 1);
    uint64_t s = 0;
    for (size_t i = 0; i < kN; i++) s += o[i] * (glwe_sk[i] & 1);
    phase[c] = o[kN] - s;
  }
  return LUTHE_OK;
}

uint64_t luthe_delta(const luthe_params* p) {
  if (!p) return 0;
  return 1ull << (64 - (p->msg_bits + p->carry_bits + 1));
}

uint32_t luthe_decode(const luthe_params* p, uint64_t phase) {
  if (!p) return 0;
  unsigned sh = 64 - (p->msg_bits + p->carry_bits + 1);
  uint64_t mod = 1ull << (p->msg_bits + p->carry_bits + 1);
  return (uint32_t)(((phase + (1ull << (sh - 1))) >> sh) & (mod - 1));
}

int luthe_lut_encode_many(const luthe_params* p, const uint8_t* tables, uint32_t table_len, uint32_t log_lut_nb,
                          uint64_t* glwe) {
  if (!p || !tables || !glwe) return LUTHE_ERR_ARG;
  const uint32_t bits = p->msg_bits + p->carry_bits;
  if (log_lut_nb > 3 || log_lut_nb > bits) return LUTHE_ERR_ARG;
  if (table_len != (1u << (bits - log_lut_nb))) return LUTHE_ERR_ARG;
  const uint32_t N = p->poly_size, k = p->glwe_k;
  const uint32_t box = N >> bits, seg = N >> log_lut_nb, nlut = 1u << log_lut_nb;
  const uint64_t delta = luthe_delta(p);
  memset(glwe, 0, sizeof(uint64_t) * (k + 1) * N);
  uint64_t* body = glwe + (size_t)k * N;
  for (uint32_t j = 0; j < N; j++) {
    uint32_t s = j / seg, rem = j % seg;
    uint32_t m = (rem + box / 2) / box;
    uint64_t v;
    if (m < table_len) {
      v = (uint64_t)tables[s * table_len + m] * delta;
    } else if (s + 1 < nlut) {  // top half box of segment s: input 0 of function s+1
      v = (uint64_t)tables[(s + 1) * table_len] * delta;
    } else {  // negacyclic wrap of function 0 at input 0
      v = 0 - (uint64_t)tables[0] * delta;
    }
    body[j] = v;
  }
  return LUTHE_OK;
}

int luthe_lut_encode(const luthe_params* p, const uint8_t* table, uint32_t table_len, uint64_t* glwe) {
  return luthe_lut_encode_many(p, table, table_len, 0, glwe);
}

}  // extern "C"
