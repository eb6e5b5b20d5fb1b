/* A C host driving libluthe.so the way firmware drives the HPU: keys and LUTs uploaded
 * once, ciphertexts written to memory slots, a DOp stream submitted and acknowledged,
 * then one radix IOp through the serving engine.  Exit status 0 when every decrypted
 * result is right.  Build: gcc -O2 -Iinclude dop_host.c -Lluthe_amd -lluthe. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "luthe.h"

#define CHECK(x)                                                                          \
  do {                                                                                    \
    int rc_ = (x);                                                                        \
    if (rc_) {                                                                            \
      fprintf(stderr, "%s failed: %d (%s)\n", #x, rc_, ctx ? luthe_last_error(ctx) : "");  \
      return 1;                                                                           \
    }                                                                                     \
  } while (0)

static uint32_t dop_arith(uint32_t op, uint32_t d, uint32_t a, uint32_t b) {
  return (op << 26) | (b << 14) | (a << 7) | d;
}
static uint32_t dop_mem(uint32_t op, uint32_t rid, uint32_t slot) { return (op << 26) | (slot << 10) | rid; }
static uint32_t dop_pbs(uint32_t d, uint32_t s, uint32_t gid) { return (0x38u << 26) | (gid << 14) | (s << 7) | d; }

int main(void) {
  luthe_ctx* ctx = NULL;
  luthe_params p;
  luthe_params_default(&p);
  p.lwe_noise_std = 1.0 / 33554432.0; /* 2^-25: results decrypt exactly */
  const size_t kN = (size_t)p.glwe_k * p.poly_size, blen = kN + 1;
  uint64_t* lwe_sk = calloc(p.lwe_dim, 8);
  uint64_t* glwe_sk = calloc(kN, 8);
  uint64_t* bsk = calloc((size_t)p.lwe_dim * 4 * p.pbs_level * p.poly_size, 8);
  uint32_t* ksk = calloc(kN * p.ks_level * (p.lwe_dim + 1), 4);
  if (!lwe_sk || !glwe_sk || !bsk || !ksk) return 2;
  CHECK(luthe_keygen_secret(&p, 2024, lwe_sk, glwe_sk));
  CHECK(luthe_keygen_bsk(&p, 2024, lwe_sk, glwe_sk, bsk, 0));
  CHECK(luthe_keygen_ksk(&p, 2024, lwe_sk, glwe_sk, ksk));
  int rc = luthe_create(&p, 0, &ctx);
  if (rc) {
    fprintf(stderr, "luthe_create: %d (no gfx950 device?)\n", rc);
    return 1;
  }
  CHECK(luthe_upload_bsk(ctx, bsk, LUTHE_BSK_COEF));
  CHECK(luthe_upload_ksk(ctx, ksk));
  /* LUT 7: x -> 15 - x */
  uint8_t table[16];
  for (int i = 0; i < 16; i++) table[i] = (uint8_t)(15 - i);
  uint64_t* lut = calloc(2 * p.poly_size, 8);
  CHECK(luthe_lut_encode(&p, table, 16, lut));
  CHECK(luthe_write_lut(ctx, 7, lut, 1));
  /* ciphertexts of 2 and 3 in slots 10, 11 */
  const uint64_t delta = luthe_delta(&p);
  uint64_t pt[2] = {2 * delta, 3 * delta};
  uint64_t* cts = calloc(3 * blen, 8);
  CHECK(luthe_encrypt_blwe(&p, 77, glwe_sk, pt, 2, cts));
  CHECK(luthe_write_ct(ctx, 10, cts, 2));
  /* R0 <- [10]; R1 <- [11]; R2 = R0 + R1; R3 = PBS_F(R2, LUT 7); [12] <- R3; SYNC */
  uint32_t prog[] = {dop_mem(0x20, 0, 10), dop_mem(0x20, 1, 11), dop_arith(0x01, 2, 0, 1), dop_pbs(3, 2, 7),
                     dop_mem(0x21, 3, 12), 0x10u << 26};
  uint64_t ticket;
  CHECK(luthe_submit_dops(ctx, prog, sizeof prog / 4, &ticket));
  CHECK(luthe_wait(ctx, ticket, -1));
  CHECK(luthe_read_ct(ctx, 12, cts + 2 * blen, 1));
  uint64_t phase;
  CHECK(luthe_decrypt_blwe(&p, glwe_sk, cts + 2 * blen, 1, &phase));
  const uint32_t got = luthe_decode(&p, phase);
  printf("DOp stream: PBS(2 + 3) = %u (expected %u)\n", got, table[5]);
  /* one 8-block radix ADD through the serving engine */
  luthe_server* srv = NULL;
  CHECK(luthe_server_create(ctx, NULL, &srv));
  const uint32_t nb = 8;
  const uint64_t a = 40000, b = 23456;
  uint64_t ptab[16];
  for (uint32_t i = 0; i < nb; i++) {
    ptab[i] = ((a >> (2 * i)) & 3) * delta;
    ptab[nb + i] = ((b >> (2 * i)) & 3) * delta;
  }
  uint64_t* ab = calloc(2 * nb * blen, 8);
  uint64_t* out = calloc(nb * blen, 8);
  CHECK(luthe_encrypt_blwe(&p, 78, glwe_sk, ptab, 2 * nb, ab));
  uint64_t id;
  CHECK(luthe_server_submit(srv, LUTHE_IOP_ADD, nb, ab, ab + nb * blen, out, &id));
  luthe_request_info info;
  CHECK(luthe_server_wait(srv, id, -1, &info));
  CHECK(luthe_server_destroy(srv));
  uint64_t sum = 0;
  for (uint32_t i = 0; i < nb; i++) {
    CHECK(luthe_decrypt_blwe(&p, glwe_sk, out + i * blen, 1, &phase));
    sum |= (uint64_t)luthe_decode(&p, phase) << (2 * i);
  }
  printf("server: ADD %llu + %llu = %llu (expected %llu), %u PBS, %.1f ms\n", (unsigned long long)a,
         (unsigned long long)b, (unsigned long long)sum, (unsigned long long)((a + b) & 0xFFFF), info.npbs,
         info.latency_us / 1e3);
  CHECK(luthe_destroy(ctx));
  return (got == table[5] && sum == ((a + b) & 0xFFFF)) ? 0 : 1;
}
