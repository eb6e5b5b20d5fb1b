// Autotuner: times candidate launch shapes on the device with the context's keys and
// picks (1) the persistent blind-rotation width, i.e. ciphertexts resident per CU,
// (2) the largest cohorts that run faster on the split variant (two workgroups per
// ciphertext) and on the 512-thread latency variant, (3) the polynomials per NTT workgroup,
// (4) the smallest batch for the matrix-core key switch and (5) the serving cohort size (one
// wave of the chosen width).  Results are applied to the context and persisted as JSON; luthe_load_tuning
// re-applies a saved file.  Inputs are synthetic: timing does not depend on values.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/luthe.h"
#include "ctx.h"
#include "kernels.h"

namespace {

struct Timer {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t s;
  explicit Timer(hipStream_t st) : s(st) {
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
  }
  ~Timer() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  void start() { (void)hipEventRecord(e0, s); }
  float stop() {
    float ms = 0;
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    return ms;
  }
};

// one warm-up and one timed PBS batch of n ciphertexts with the given kernel settings
int time_pbs(luthe_ctx* c, const uint64_t* d_in, uint64_t* d_out, int n, int grid, int lat_max, float& ms,
             int split_max = 0) {
  c->br_grid_override = grid;
  c->br_latency_max = lat_max;
  c->br_split_max = split_max;
  HIPCHK(c, lk::launch_make_jobs(c->d_jobs, d_in, d_out, nullptr, n, 0, c->prm.lut_slots, (int)c->blen, c->stream));
  int rc = lx::run_pbs(c, c->bs, c->d_jobs, n, c->stream);
  if (rc) return rc;
  Timer t(c->stream);
  t.start();
  rc = lx::run_pbs(c, c->bs, c->d_jobs, n, c->stream);
  ms = t.stop();
  return rc;
}

bool json_int(const std::string& s, const char* key, long& v) {
  std::string k = std::string("\"") + key + "\"";
  size_t p = s.find(k);
  if (p == std::string::npos) return false;
  p = s.find(':', p + k.size());
  if (p == std::string::npos) return false;
  char* end = nullptr;
  v = strtol(s.c_str() + p + 1, &end, 10);
  return end != s.c_str() + p + 1;
}

int apply(luthe_ctx* c, const luthe_tuning& t) {
  if (t.br_grid < 0 || t.br_latency_max < -1 || t.cohort_max < 0 || t.br_pair < -1 || t.br_pair > 1 ||
      t.br_split_max < -1 || t.br_split_wait_us < -1 || t.ks_mfma_min < -1)
    return LUTHE_ERR_ARG;
  if (t.ntt_polys_per_wg != 1 && t.ntt_polys_per_wg != 2 && t.ntt_polys_per_wg != 4) return LUTHE_ERR_ARG;
  c->br_grid_override = t.br_grid;
  c->br_latency_max = t.br_latency_max;
  c->ntt_polys_per_wg = t.ntt_polys_per_wg;
  c->cohort_max = t.cohort_max;
  c->br_pair = t.br_pair;
  c->br_split_max = t.br_split_max;
  c->br_split_wait_us = t.br_split_wait_us;
  c->ks_mfma_min = t.ks_mfma_min;
  return LUTHE_OK;
}

}  // namespace

extern "C" {

int luthe_get_tuning(luthe_ctx* c, luthe_tuning* out) {
  if (!c || !out) return LUTHE_ERR_ARG;
  out->br_grid = c->br_grid_override;
  out->br_latency_max = lx::latency_max(c);
  out->ntt_polys_per_wg = c->ntt_polys_per_wg;
  out->cohort_max = c->cohort_max;
  out->br_pair = c->br_pair;
  out->br_split_max = lx::split_max(c);
  out->br_split_wait_us = c->br_split_wait_us;
  out->ks_mfma_min = lx::ks_mfma_min(c);
  return LUTHE_OK;
}

int luthe_set_tuning(luthe_ctx* c, const luthe_tuning* t) {
  if (!c || !t) return LUTHE_ERR_ARG;
  if (c->server) return lx::fail(c, LUTHE_ERR_BUSY, "keys and launch settings are fixed while a server is attached");
  int rc = apply(c, *t);
  if (rc) return lx::fail(c, rc, "invalid tuning");
  return LUTHE_OK;
}

int luthe_load_tuning(luthe_ctx* c, const char* path) {
  if (!c || !path) return LUTHE_ERR_ARG;
  if (c->server) return lx::fail(c, LUTHE_ERR_BUSY, "keys and launch settings are fixed while a server is attached");
  FILE* f = fopen(path, "rb");
  if (!f) return lx::fail(c, LUTHE_ERR_ARG, "cannot open %s", path);
  std::string s;
  char buf[4096];
  size_t k;
  while ((k = fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, k);
  fclose(f);
  long pl = 0;
  if (json_int(s, "pbs_level", pl) && pl != (long)c->prm.pbs_level)
    return lx::fail(c, LUTHE_ERR_PARAMS, "%s was tuned for PBS level %ld", path, pl);
  luthe_tuning t{};
  long v;
  if (!json_int(s, "br_grid", v)) return lx::fail(c, LUTHE_ERR_ARG, "%s: no br_grid", path);
  t.br_grid = (int32_t)v;
  t.br_latency_max = json_int(s, "br_latency_max", v) ? (int32_t)v : -1;
  t.ntt_polys_per_wg = json_int(s, "ntt_polys_per_wg", v) ? (int32_t)v : 2;
  t.cohort_max = json_int(s, "cohort_max", v) ? (int32_t)v : 0;
  t.br_pair = json_int(s, "br_pair", v) ? (int32_t)v : -1;
  t.br_split_max = json_int(s, "br_split_max", v) ? (int32_t)v : -1;
  t.br_split_wait_us = c->br_split_wait_us;  // a robustness limit, not a tuned choice
  t.ks_mfma_min = json_int(s, "ks_mfma_min", v) ? (int32_t)v : -1;
  int rc = apply(c, t);
  if (rc) return lx::fail(c, rc, "%s: invalid tuning", path);
  return LUTHE_OK;
}

int luthe_autotune(luthe_ctx* c, uint32_t batch, const char* json_path, luthe_tuning* out) {
  if (!c || batch < 16 || batch > (1u << 20)) return LUTHE_ERR_ARG;
  if (c->server) return lx::fail(c, LUTHE_ERR_BUSY, "context stream is owned by an attached server");
  if (!c->bsk_ok || !c->ksk_ok) return lx::fail(c, LUTHE_ERR_STATE, "keys not uploaded");
  HIPCHK(c, hipSetDevice(c->device));
  int cus = 0;
  HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  const int L = (int)c->prm.pbs_level;
  const int occ = lk::blind_rotate_occupancy(L, false), per_cu = occ / (cus > 0 ? cus : 1);
  const int occ_pair = lk::blind_rotate_occupancy(L, true);
  const int lat_slots = lk::blind_rotate_latency_slots(L);
  const int split_slots = lk::blind_rotate_split_slots();
  int rc = lx::ensure_jobs(c, batch);
  if (!rc) rc = lx::ensure_lwe2n(c, c->bs, batch);
  if (rc) return rc;
  // synthetic inputs: any u64 values (the kernels' cost does not depend on them)
  uint64_t *d_in = nullptr, *d_out = nullptr, *d_a = nullptr, *d_b = nullptr, *d_o = nullptr;
  const size_t rows = batch;
  const size_t npoly = 4096;
  auto cleanup = [&]() {
    for (void* p : {(void*)d_in, (void*)d_out, (void*)d_a, (void*)d_b, (void*)d_o})
      if (p) (void)hipFree(p);
  };
  if (hipMalloc((void**)&d_in, rows * c->blen * 8) != hipSuccess ||
      hipMalloc((void**)&d_out, rows * c->blen * 8) != hipSuccess ||
      hipMalloc((void**)&d_a, npoly * 2048 * 8) != hipSuccess || hipMalloc((void**)&d_b, npoly * 2048 * 8) != hipSuccess ||
      hipMalloc((void**)&d_o, npoly * 2048 * 8) != hipSuccess) {
    cleanup();
    return lx::fail(c, LUTHE_ERR_NOMEM, "autotune buffers");
  }
  (void)hipMemsetAsync(d_in, 0x5A, rows * c->blen * 8, c->stream);
  (void)hipMemsetAsync(d_a, 0x31, npoly * 2048 * 8, c->stream);
  (void)hipMemsetAsync(d_b, 0x17, npoly * 2048 * 8, c->stream);
  const luthe_tuning saved{c->br_grid_override, c->br_latency_max, c->ntt_polys_per_wg, c->cohort_max, c->br_pair,
                           c->br_split_max, c->br_split_wait_us, c->ks_mfma_min};
  c->br_pair = 0;  // (1) sweeps the 256-thread kernel
  c->br_split_wait_us = -1;  // the split kernel's default wait limit while timing
  std::string meas;
  char line[256];
  // (1) resident ciphertexts per CU
  int best_grid = occ;
  float best_rate = 0;
  for (int k = 1; k <= per_cu; k++) {
    float ms = 0;
    rc = time_pbs(c, d_in, d_out, (int)batch, k * cus, 0, ms);
    if (rc) break;
    const float rate = batch / (ms * 1e-3f);
    snprintf(line, sizeof line, "%s{\"per_cu\": %d, \"grid\": %d, \"ms\": %.3f, \"pbs_per_s\": %.1f}",
             meas.empty() ? "" : ", ", k, k * cus, ms, rate);
    meas += line;
    if (rate > best_rate * 1.005f) best_rate = rate, best_grid = k * cus;
  }
  // (1b) the 512-thread pair kernel at its full width
  int best_pair = 0;
  if (!rc) {
    float ms = 0;
    c->br_pair = 1;
    rc = time_pbs(c, d_in, d_out, (int)batch, 0, 0, ms);
    const float rate = batch / (ms * 1e-3f);
    snprintf(line, sizeof line, ", {\"pair\": 1, \"grid\": %d, \"ms\": %.3f, \"pbs_per_s\": %.1f}", occ_pair, ms,
             rate);
    meas += line;
    if (!rc && rate > best_rate * 1.005f) best_rate = rate, best_pair = 1;
    c->br_pair = best_pair;
  }
  const int width = best_pair ? occ_pair : best_grid;
  // (2) split and latency variants: the largest cohorts where each beats the kernels of
  // larger cohorts (split: both others; latency: the throughput kernel)
  std::string lat;
  int lat_max = 0, split_best = 0;
  const int grid_thr = best_pair ? 0 : best_grid;
  for (int n : {1, cus / 8, cus / 4, cus / 2, cus, (cus * 3) / 2, 2 * cus, 3 * cus}) {
    if (rc || n < 1 || n > (int)batch || (n > lat_slots && n > split_slots)) continue;
    float t_split = 0, t_lat = 0, t_thr = 0;
    if (n <= split_slots) rc = time_pbs(c, d_in, d_out, n, grid_thr, 0, t_split, n);
    if (!rc && n <= lat_slots) rc = time_pbs(c, d_in, d_out, n, grid_thr, n, t_lat);
    if (!rc) rc = time_pbs(c, d_in, d_out, n, grid_thr, 0, t_thr);
    if (rc) break;
    snprintf(line, sizeof line, "%s{\"n\": %d, \"split_ms\": %.3f, \"latency_ms\": %.3f, \"throughput_ms\": %.3f}",
             lat.empty() ? "" : ", ", n, t_split, t_lat, t_thr);
    lat += line;
    if (n <= lat_slots && t_lat < t_thr) lat_max = n;
    if (n <= split_slots && t_split < t_thr && (n > lat_slots || t_split < t_lat)) split_best = n;
  }
  // (3) polynomials per NTT workgroup (forward, pointwise, inverse over 4096 polynomials)
  std::string ntt;
  int best_np = 2;
  float best_ntt = 1e30f;
  for (int np : {1, 2, 4}) {
    if (rc) break;
    hipError_t e = lk::launch_ntt(2, d_a, d_b, d_o, npoly, c->d_tab, c->stream, np);
    Timer t(c->stream);
    t.start();
    for (int r = 0; r < 3 && e == hipSuccess; r++) e = lk::launch_ntt(2, d_a, d_b, d_o, npoly, c->d_tab, c->stream, np);
    const float ms = t.stop() / 3;
    if (e != hipSuccess) {
      rc = lx::fail(c, LUTHE_ERR_HIP, "NTT tuning: %s", hipGetErrorString(e));
      break;
    }
    snprintf(line, sizeof line, "%s{\"polys_per_wg\": %d, \"ms\": %.4f}", ntt.empty() ? "" : ", ", np, ms);
    ntt += line;
    if (ms < best_ntt * 0.99f) best_ntt = ms, best_np = np;
  }
  // (4) key switch: the smallest tested batch from which the matrix-core kernel is faster
  // than the VALU kernel at that size and every larger tested size
  std::string ksm;
  int ks_min = 1 << 30;
  if (!rc && c->d_kskp) rc = lx::ensure_ks_part(c, c->bs, batch);
  if (!rc && c->d_kskp) {
    HIPCHK(c, lk::launch_make_jobs(c->d_jobs, d_in, d_out, nullptr, (int)batch, 0, c->prm.lut_slots, (int)c->blen,
                                   c->stream));
    const int sizes[] = {(int)batch, 3 * cus, cus, cus / 2, cus / 4, 32, 8, 1};
    bool wins = true;
    for (int n : sizes) {
      if (rc || n < 1 || n > (int)batch) continue;
      float ms[2] = {0, 0};
      for (int m = 0; m < 2 && !rc; m++) {
        c->ks_mfma_min = m ? 1 : (1 << 30);
        rc = lx::launch_ks(c, c->d_jobs, n, c->bs.lwe2n, c->lwe_stride, c->bs.ks_part, c->stream);
        Timer t(c->stream);
        t.start();
        for (int r = 0; r < 3 && !rc; r++)
          rc = lx::launch_ks(c, c->d_jobs, n, c->bs.lwe2n, c->lwe_stride, c->bs.ks_part, c->stream);
        ms[m] = t.stop() / 3;
      }
      if (rc) break;
      snprintf(line, sizeof line, "%s{\"n\": %d, \"valu_ms\": %.4f, \"mfma_ms\": %.4f}", ksm.empty() ? "" : ", ", n,
               ms[0], ms[1]);
      ksm += line;
      wins = wins && ms[1] < ms[0];
      if (wins) ks_min = n;
    }
    c->ks_mfma_min = saved.ks_mfma_min;
  }
  (void)hipStreamSynchronize(c->stream);
  cleanup();
  if (!rc) rc = lx::check_device_status(c);
  if (rc) {
    (void)apply(c, saved);
    return rc;
  }
  luthe_tuning t{best_pair || best_grid == occ ? 0 : best_grid, lat_max, best_np, width, best_pair, split_best,
                 saved.br_split_wait_us, c->d_kskp ? ks_min : saved.ks_mfma_min};
  (void)apply(c, t);
  if (out) *out = t;
  if (json_path) {
    FILE* f = fopen(json_path, "wb");
    if (!f) return lx::fail(c, LUTHE_ERR_ARG, "cannot write %s", json_path);
    hipDeviceProp_t prop;
    (void)hipGetDeviceProperties(&prop, c->device);
    fprintf(f,
            "{\n  \"device\": \"%s\", \"cus\": %d, \"lwe_dim\": %u, \"pbs_level\": %d, \"batch\": %u,\n"
            "  \"br_grid\": %d, \"br_pair\": %d, \"br_split_max\": %d, \"br_latency_max\": %d, \"ntt_polys_per_wg\": %d,\n"
            "  \"cohort_max\": %d, \"ks_mfma_min\": %d,\n"
            "  \"measured\": {\n    \"blind_rotation_width\": [%s],\n    \"latency_variant\": [%s],\n"
            "    \"ntt_polymul_4096\": [%s],\n    \"key_switch\": [%s]\n  }\n}\n",
            prop.gcnArchName, cus, c->prm.lwe_dim, L, batch, t.br_grid, t.br_pair, t.br_split_max, t.br_latency_max,
            t.ntt_polys_per_wg,
            t.cohort_max, t.ks_mfma_min, meas.c_str(), lat.c_str(), ntt.c_str(), ksm.c_str());
    fclose(f);
  }
  return LUTHE_OK;
}

}  // extern "C"
