// Kernel-side descriptors and host launchers shared by kernels.hip and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lk {

// One PBS of a batch: device address of its input BLWE (key switch input) and of its first
// output BLWE (lut_nb consecutive BLWEs for Many-LUT), LUT id, log2 of the number of LUTs.
// Absolute addresses let one launch span batches in different allocations.
struct PbsJob {
  uint64_t in_addr;
  uint64_t out_addr;
  uint32_t gid;
  uint32_t log_lut_nb;
};

// One batch of a batch list, in device memory (the layout of luthe_pbs_batch_desc).
struct PbsBatchDesc {
  const uint64_t* in;
  uint64_t* out;
  const uint32_t* gid;
  uint32_t n;
  uint32_t log_lut_nb;
};

struct AluOp {
  uint32_t opcode;
  uint32_t dst_row;
  uint32_t src0_row;
  uint32_t src1_row;
  uint32_t mul_factor;
  uint32_t msg_cst;
};

struct CopyOp {
  uint32_t dst_row;
  uint32_t src_row;
};

struct KsArgs {
  int kN;          // mask length of the input BLWE
  int n_lwe;       // output LWE dimension
  int lwe_stride;  // u16 elements between output rows (>= n_lwe + 1)
  uint32_t base_log, mod_log, lc;  // KS_B_W, MOD_KSK_W, log2(2N)
};

struct BrArgs {
  int n_lwe;
  int lwe_stride;
  uint32_t base_log;      // PBS_B_W
  int split_wait_us = -1;  // split kernel: longest wait for the partner (-1: 2 s; 0: none)
};

// polys_per_wg: polynomials per 256-thread workgroup (1, 2 or 4; tuned).
hipError_t launch_ntt(int mode, const uint64_t* a, const uint64_t* bn, uint64_t* out, size_t npoly,
                      const uint64_t* tab, hipStream_t s, int polys_per_wg = 2);
hipError_t launch_bsk_prepare(const uint64_t* src, uint64_t* dst, int n_lwe, int rows, int K1, int is_ntt,
                              const uint64_t* tab, hipStream_t s);
// nsplit > 1 (small batches): the input coefficients split over nsplit workgroups whose partial
// sums go to part (nsplit * nct * (n_lwe + 1) words), then added by a finishing kernel.
hipError_t launch_keyswitch(const PbsJob* jobs, int nct, const uint32_t* ksk, uint16_t* lwe2n,
                            const KsArgs& a, int ks_level, hipStream_t s, uint32_t* part = nullptr, int nsplit = 1);
// The split launch_keyswitch uses for a batch of nct ciphertexts (1 = none).
int keyswitch_split(int nct, int kN, int n_lwe);
// Key switch on the int8 matrix cores (ks_mfma.hip): the KSK as three balanced byte planes in
// fragment order (ksk_planes_bytes, built by launch_ksk_planes from the u32 KSK), digits
// decomposed in the kernel.  Allocates nothing; results equal launch_keyswitch's.
// keyswitch_mfma_ok: the parameters it covers (kN % 32 == 0, base_log * level <= 31).
bool keyswitch_mfma_ok(const KsArgs& a, int ks_level);
size_t ksk_planes_bytes(int kN, int n_lwe, int ks_level);
hipError_t launch_ksk_planes(const uint32_t* ksk, int8_t* planes, const KsArgs& a, int ks_level, hipStream_t s);
hipError_t launch_keyswitch_mfma(const PbsJob* jobs, int nct, const int8_t* planes, uint16_t* lwe2n, const KsArgs& a,
                                 int ks_level, hipStream_t s);
// Cohorts of at most latency_max ciphertexts run the 512-thread latency variant; larger ones
// the 256-thread persistent kernel, or with `pair` the 512-thread two-per-CU variant.
hipError_t launch_blind_rotate(const uint64_t* bsk, const uint64_t* luts, const uint16_t* lwe2n, const PbsJob* jobs,
                               int nct, const uint64_t* tab, const BrArgs& a,
                               int pbs_level, int grid_override, hipStream_t s, int latency_max, bool pair,
                               int split_max = 0, uint64_t* xch = nullptr, uint32_t* status = nullptr,
                               bool split_local = false);
// Split blind rotation (two workgroups per ciphertext) for cohorts of at most split_max
// ciphertexts when xch (blind_rotate_split_words(pbs_level) u64 per ciphertext) and status
// (a word the host can read; set to 1 if a pair's exchange timed out) are given.  split_local:
// the pair's words go through the shared L2 of its XCC (split_pairs_share_xcc() must hold).
size_t blind_rotate_split_words(int pbs_level);
// Whether the device places workgroups b and b ^ 8 of a split grid on one XCC (a probe kernel
// on stream s, synchronous; false on any error).
bool split_pairs_share_xcc(hipStream_t s);
// The split variant's launch (split.hip): clears the pairs' exchange area (xch, nct x
// blind_rotate_split_words(pbs_level) words), then the pairs; hi and local select the
// template instance as in launch_blind_rotate.
void launch_blind_rotate_split(int pbs_level, bool hi, bool local, const uint64_t* bsk, const uint64_t* luts,
                               const uint16_t* lwe2n, const PbsJob* jobs, int nct, const uint64_t* tab,
                               const BrArgs& a, hipStream_t s, uint64_t* xch, uint32_t* status);
// Ciphertexts the split variant runs with one workgroup per CU (CUs / 2).
int blind_rotate_split_slots();
// The 512-thread blind rotation (latency.hip): one ciphertext per workgroup, the halves owning
// one GLWE polynomial each; pair: two workgroups per CU (the level-2 throughput variant), else
// one (the latency variant for small cohorts).  grid_override > 0 fixes the grid.
void launch_blind_rotate_lat(int pbs_level, bool hi, bool pair, int grid_override, const uint64_t* bsk,
                             const uint64_t* luts, const uint16_t* lwe2n, const PbsJob* jobs, int nct,
                             const uint64_t* tab, const BrArgs& a, hipStream_t s);
// Resident workgroups of the 512-thread kernel (CUs x workgroups per CU).
int blind_rotate_lat_max_grid(int pbs_level, bool pair);
// Resident workgroups of the throughput kernel (CUs x workgroups per CU).
int blind_rotate_occupancy(int pbs_level, bool pair);
// Ciphertexts the latency variant holds resident at once (its workgroups per CU x CUs).
int blind_rotate_latency_slots(int pbs_level);
hipError_t launch_modmul_peak(uint64_t* out, int blocks, int iters, hipStream_t s);
// useful_bits = msg + carry + 1 padding bit: the ADDS/SUBS/SSUB constant's width in the body.
hipError_t launch_alu(uint64_t* base, const AluOp* ops, int nops, int blen, uint32_t useful_bits, hipStream_t s);
hipError_t launch_copy(uint64_t* base, const CopyOp* ops, int nops, int blen, hipStream_t s);
hipError_t launch_alu_batch(uint32_t opcode, const uint64_t* a, const uint64_t* b, uint64_t* out, uint32_t mul_factor,
                            uint32_t msg_cst, int n, int blen, uint32_t useful_bits, hipStream_t s);
// Jobs of `count` batches described in device memory, `total` job slots.
hipError_t launch_make_jobs_list(PbsJob* jobs, const PbsBatchDesc* descs, int count, int total, uint32_t lut_slots,
                                 int blen, hipStream_t s);
// Jobs of one contiguous batch: input c at in + c * blen, outputs at out + (c << log_lut_nb) * blen.
hipError_t launch_make_jobs(PbsJob* jobs, const uint64_t* in, uint64_t* out, const uint32_t* gid, int n,
                            uint32_t log_lut_nb, uint32_t lut_slots, int blen,
                            hipStream_t s);

}  // namespace lk
