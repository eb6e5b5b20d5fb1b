/*
 * luthe.h -- C ABI of the MI355X TFHE bootstrapping engine (libluthe.so).
 *
 * The engine takes the place of the HPU's PBS/ALU/memory processing elements behind
 * its digit-operation (DOp) surface: a C host hands it 32-bit DOp words (or IOp
 * words plus registered DOp templates) instead of streaming them to the FPGA.
 * Every entry point below names the reference interface it replaces.  Reference
 * paths are relative to the reference checkout.
 *
 * Conventions
 *  - Ciphertexts are "BLWE": k*N mask coefficients (polynomial g at [g*N, (g+1)*N),
 *    natural coefficient order) followed by the body, all u64 mod 2^64.
 *  - LUTs are GLWE: (k+1)*N u64, mask polynomials then the body polynomial.
 *  - Status: every int-returning call returns LUTHE_OK (0) or a negative LUTHE_ERR_*;
 *    luthe_last_error() gives the message of the last failure on that context.
 *  - Threading: a context is externally synchronised (one submitting thread at a
 *    time).  Device memory belongs to the context and is freed by luthe_destroy().
 *    "d_" pointers are device pointers on the context's device; other pointers are
 *    host memory owned by the caller.
 *  - The library requires a gfx950 GPU; there is no CPU execution path.
 */
#ifndef LUTHE_H
#define LUTHE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LUTHE_ABI_VERSION 1

enum {
  LUTHE_OK = 0,
  LUTHE_ERR_ARG = -1,         /* bad argument (null pointer, size) */
  LUTHE_ERR_PARAMS = -2,      /* parameter set not supported */
  LUTHE_ERR_HIP = -3,         /* HIP runtime error */
  LUTHE_ERR_NOMEM = -4,       /* device memory / HBM admission refused */
  LUTHE_ERR_DOP = -5,         /* malformed DOp or IOp word */
  LUTHE_ERR_RANGE = -6,       /* register, slot or LUT id out of range */
  LUTHE_ERR_STATE = -7,       /* keys not uploaded, unknown ticket, ... */
  LUTHE_ERR_TIMEOUT = -8,     /* luthe_wait timed out */
  LUTHE_ERR_NODEV = -9,       /* no usable gfx950 device */
  LUTHE_ERR_BUSY = -10        /* admission queue full, or the context is owned by a server */
};

/* TFHE parameter set.  Replaces the compile-time parameter packages
 * hw/module/param/module/<name>/rtl/param_tfhe_definition_pkg.sv (N, GLWE_K, PBS_L,
 * PBS_B_W, LWE_K, KS_L, KS_B_W, MOD_KSK_W, PAYLOAD/PADDING) and the memory sizes of
 * fw/ublaze/src/hpu_dop_fmt.h:43 (CT_MEM_SLOT), hpu_config.toml.j2:22 (lut_mem) and
 * docs/hpu.md:31-32 (64 registers). */
typedef struct luthe_params {
  uint32_t lwe_dim;        /* n (LWE_K) */
  uint32_t glwe_k;         /* k (GLWE_K), must be 1 */
  uint32_t poly_size;      /* N, must be 2048 */
  uint32_t pbs_base_log;   /* PBS_B_W */
  uint32_t pbs_level;      /* PBS_L in {1, 2} */
  uint32_t ks_base_log;    /* KS_B_W in [1, 7] (the reference's parameter sets use 2 and 3) */
  uint32_t ks_level;       /* KS_L in [1, 8] */
  uint32_t ks_mod_log;     /* MOD_KSK_W in [13, 24] (the reference's sets use 21) */
  uint32_t msg_bits;       /* message bits of a block (2) */
  uint32_t carry_bits;     /* carry bits of a block (2) */
  double lwe_noise_std;    /* keygen/encryption only: std relative to the modulus */
  double glwe_noise_std;
  uint32_t regfile_size;   /* BLWE registers, <= 128 (RID_W = 7) */
  uint32_t ct_slots;       /* ciphertext memory slots, <= 65536 (CID_W = 16) */
  uint32_t lut_slots;      /* LUT GLWE slots, <= 4096 (GID_W = 12) */
} luthe_params;

typedef struct luthe_ctx luthe_ctx;

/* Counters.  Replaces the HPU info/counter registers (pep_common_param_pkg.sv:277-330,
 * hpu_regif_counter_manager.sv): PBS issued, batches, fill, DOps by kind. */
typedef struct luthe_counters {
  uint64_t pbs;            /* PBS (KS + blind rotation + extract) executed */
  uint64_t pbs_batches;    /* blind-rotation launches */
  uint64_t pbs_max_batch;  /* largest batch launched */
  uint64_t alu;            /* ALU DOps executed */
  uint64_t mem;            /* LD/ST DOps executed */
  uint64_t dops;           /* DOps accepted */
  uint64_t iops;           /* IOps accepted */
  uint64_t tickets;        /* submissions */
  uint64_t levels;         /* dependency levels scheduled */
  uint64_t groups;         /* DOp groups launched (submissions scheduled together) */
} luthe_counters;

/* ---------------------------------------------------------------- context */

/* Defaults: n=724, k=1, N=2048, PBS 2^23 x 1, KS 2^3 x 5, w_ks=21,
 * 2+2 bit blocks, noise 2^-15 / 2^-51, 64 registers, 4096 ct slots, 256 LUT slots. */
void luthe_params_default(luthe_params* p);

/* Create a context on HIP device `device`.  Allocates the register file, ciphertext
 * and LUT memories.  Replaces the card bring-up and info-register read-back
 * (docs/hpu.md:42-48, hpu_regif_core_cfg_1in3.toml:43-200). */
int luthe_create(const luthe_params* p, int device, luthe_ctx** out);
int luthe_destroy(luthe_ctx* ctx);
const char* luthe_last_error(const luthe_ctx* ctx);
int luthe_get_params(const luthe_ctx* ctx, luthe_params* out);
int luthe_get_counters(const luthe_ctx* ctx, luthe_counters* out);
/* HBM bytes held by the context (keys, memories, scratch). */
uint64_t luthe_device_bytes(const luthe_ctx* ctx);

/* ---------------------------------------------------------------- keys and memories */

#define LUTHE_BSK_COEF 0 /* [n][k+1][L][k+1][N] coefficient domain, values < p */
#define LUTHE_BSK_NTT 1  /* same shape, NTT domain in the order of the reference model's
                            ntt() with cut [5,6] (sw/ntt/ntt_gf64.sage:367-445) */

/* Upload the bootstrapping key (GGSW_i of LWE key bit i, mod p = 2^64-2^32+1).
 * Row (g, l): g-th GLWE component gadget level l (l = 0 is the least significant
 * level, gadget round(p / 2^(b*(L-l)))), columns j = GLWE components.  Converted on
 * the device into the blind-rotation layout.  Replaces the BSK image written to HBM
 * and streamed by hw/module/bsk/rtl/bsk_manager.sv (header: reverse order (R,N)). */
int luthe_upload_bsk(luthe_ctx* ctx, const uint64_t* bsk, int layout);

/* Upload the key-switching key [k*N][ks_level][n+1] (u32, values mod 2^ks_mod_log;
 * level 0 = least significant digit, weight 2^(w - b*(L-l))).  Replaces the KSK
 * image streamed by hw/module/ksk/rtl/ksk_manager.sv. */
int luthe_upload_ksk(luthe_ctx* ctx, const uint32_t* ksk);

/* LUT memory (GLWE indexed by gid, hpu_dop_fmt.h:94-99 / GID_W=12). */
int luthe_write_lut(luthe_ctx* ctx, uint32_t gid, const uint64_t* glwe, uint32_t count);

/* Ciphertext memory (slots addressed by cid, hpu_dop_fmt.h:79-85 / CID_W=16) and the
 * register file (rid, RID_W=7).  count = number of consecutive BLWEs. */
int luthe_write_ct(luthe_ctx* ctx, uint32_t cid, const uint64_t* blwe, uint32_t count);
int luthe_read_ct(luthe_ctx* ctx, uint32_t cid, uint64_t* blwe, uint32_t count);
int luthe_write_reg(luthe_ctx* ctx, uint32_t rid, const uint64_t* blwe, uint32_t count);
int luthe_read_reg(luthe_ctx* ctx, uint32_t rid, uint64_t* blwe, uint32_t count);

/* ---------------------------------------------------------------- DOp stream */

/* Submit a stream of 32-bit DOp words (formats of fw/ublaze/src/hpu_dop_fmt.h:50-116,
 * opcodes of hpu_common_instruction_pkg.sv:42-68).  The whole stream is validated
 * first (nothing runs if any word is invalid); memory operands must be concrete
 * (MEM_ADDR) and immediates constant (IMM_CST).  DOps run as soon as their register
 * and slot dependencies allow, PBS DOps of the same dependency level -- of this and of
 * the other submissions launched with it (luthe_set_dop_batching) -- sharing one batched
 * launch; results equal in-order execution of all submissions.  SYNC words are accepted and
 * order nothing further (completion is per ticket).  Replaces the firmware's DOp
 * stream to the scheduler (ucore_fw.c:96-110) and its ack path (ucore_fw.c:224-232). */
int luthe_submit_dops(luthe_ctx* ctx, const uint32_t* words, size_t n, uint64_t* ticket);
/* 1 = done, 0 = pending, < 0 = error.  Polling a ticket that has not been launched yet
 * launches its group. */
int luthe_poll(luthe_ctx* ctx, uint64_t ticket);
/* Block until the ticket completes; timeout_us < 0 waits forever. */
int luthe_wait(luthe_ctx* ctx, uint64_t ticket, int64_t timeout_us);

/* Continuous batching across submissions.  DOp and IOp submissions join an open group; a
 * group is scheduled as one program (the streams concatenated in submission order, so
 * results equal running the submissions one at a time) and the PBS of all its tickets that
 * are ready at one dependency level share one batched launch.  The open group is launched
 * when fewer than max_groups launched groups are still running and
 *   fill_pbs = 0 (default): at once -- submissions made while a group runs merge into the next;
 *   fill_pbs > 0: once it holds fill_pbs PBS, a PBS carries the flush bit (PBS_F, docs/dop.md:
 *     108-111), or its oldest submission has waited timeout_us (0 = no timeout);
 * and in any case when the host needs a result: luthe_poll / luthe_wait on one of its
 * tickets, luthe_flush, luthe_sync, ct / register / LUT reads and writes, graph capture.  The
 * rules are evaluated whenever the host calls into the context (there is no timer thread).
 * Replaces the BPIP batch former's start conditions (pep_sequencer.sv:584-608: full batch,
 * timeout, flush) and its opportunistic mode (bpip_use_opportunism / bpip_timeout,
 * hpu_config.toml.j2:9-11) over the instruction scheduler's DOp pool
 * (instruction_scheduler.sv:1-20). */
typedef struct luthe_dop_batching {
  uint32_t fill_pbs;    /* 0 = launch whenever a group slot is free */
  uint32_t timeout_us;  /* with fill_pbs > 0: launch once the oldest submission waited this long */
  uint32_t max_groups;  /* launched groups allowed to run at once, 1..64 (default 1) */
} luthe_dop_batching;
int luthe_set_dop_batching(luthe_ctx* ctx, const luthe_dop_batching* b);
int luthe_get_dop_batching(const luthe_ctx* ctx, luthe_dop_batching* out);
/* Launch the open group now (the host-side flush trigger). */
int luthe_flush(luthe_ctx* ctx);

/* ---------------------------------------------------------------- IOp path */

/* Register the DOp template program of IOp `opcode` for operand width `max_align`
 * blocks (the translation table at DOP_LUT_ADDR, ucore.h:40-42, ucore_fw.c:207-220).
 * Templates may use MEM_HEAP / MEM_SRC / MEM_DST slots and IMM_VAR immediates. */
int luthe_register_iop(luthe_ctx* ctx, uint8_t opcode, uint8_t max_align, const uint32_t* dops, size_t n);
/* Submit one IOp (header, destination and source operand words, optional immediates:
 * hpu_iop_fmt.h:19-76): parse_iop + get_lookup + patch_dop of ucore_fw.c:135-220 and
 * hpu_dop_fmt.h:148-229, then as luthe_submit_dops. */
int luthe_submit_iop(luthe_ctx* ctx, const uint32_t* words, size_t n, uint64_t* ticket);

/* Host-only utilities (no device needed).  Validate a concrete DOp stream and return
 * each DOp's dependency level (-1 for SYNC) in `levels` under register and slot hazards
 * (read-after-write, write-after-read, write-after-write), the order the HPU's
 * instruction scheduler enforces; result = number of levels or a negative error. */
int luthe_dop_schedule(const uint32_t* words, size_t n, uint32_t regfile_size, uint32_t ct_slots,
                       uint32_t lut_slots, int32_t* levels);
/* The plan luthe_submit_dops executes (host only): register writes renamed to fresh rows,
 * so levels follow true data dependencies and slot hazards only.  Row space: register
 * r = r, slot s = regfile_size + s, renamed k = regfile_size + ct_slots + k.
 * entries[i] = {level (-1 for SYNC), dst row, src0 row, src1 row}; copy_back receives
 * (register, row) pairs applied after the last level (2 * copy_back_cap words).
 * Returns the number of levels (copy-back excluded) or a negative error. */
typedef struct luthe_dop_plan_entry {
  int32_t level;
  uint32_t dst, src0, src1;
} luthe_dop_plan_entry;
int luthe_dop_plan(const uint32_t* words, size_t n, uint32_t regfile_size, uint32_t ct_slots, uint32_t lut_slots,
                   luthe_dop_plan_entry* entries, uint32_t* copy_back, size_t copy_back_cap, size_t* n_copy_back,
                   uint32_t* rename_rows);
/* Translate one IOp against a DOp template program (parse_iop + patch_dop of
 * ucore_fw.c:135-220, hpu_dop_fmt.h:148-229) without executing: out receives n_tpl
 * concrete DOp words.  Returns 0 or a negative error. */
int luthe_iop_translate(const uint32_t* iop_words, size_t n_iop, const uint32_t* tpl, size_t n_tpl,
                        uint32_t ct_slots, uint32_t* out);

/* ---------------------------------------------------------------- batched hot path */

/* Batched PBS on caller device memory: for c < n, d_out[c*lut_nb + j] = PBS of d_in[c]
 * with LUT gid d_gid[c], Many-LUT output j < lut_nb = 2^log_lut_nb (BLWE stride
 * k*N+1).  Key switch, blind rotation and sample extract in one stream-ordered
 * sequence on `stream` (a hipStream_t, NULL = the context stream).  Replaces the PBS
 * processing element pe_pbs (hw/module/pe_pbs/rtl/pe_pbs.sv:14-60) and its batch
 * command (pep_common_param_pkg.sv:95-159).  LUT ids >= lut_slots are clamped to the
 * last LUT slot (device-side ids cannot be validated without a host round trip). */
int luthe_pbs_batch(luthe_ctx* ctx, const uint64_t* d_in, uint64_t* d_out, const uint32_t* d_gid,
                    uint32_t log_lut_nb, size_t n, void* stream);

/* Several independent batches -- each with its own input, output and LUT-id arrays, sizes and
 * Many-LUT count -- as one key-switch launch and one persistent blind-rotation launch, so
 * uneven batch sizes do not leave the GPU partly idle at each batch's tail.  d_descs: `count`
 * descriptors in device memory (read by the device, so the call can be captured in a HIP graph
 * and replayed; a descriptor's log_lut_nb is clamped to 3).  total must be the sum of the
 * descriptors' n: PBS beyond total are not run, and slots of total that no descriptor fills
 * run nothing.  Results equal luthe_pbs_batch on each batch.  Replaces the HPU's batch
 * commands over its PBS pool (pep_common_param_pkg.sv:95-159), which mix PBS of any origin in
 * one batch. */
typedef struct luthe_pbs_batch_desc {
  const uint64_t* d_in;   /* n BLWEs */
  uint64_t* d_out;        /* n << log_lut_nb BLWEs */
  const uint32_t* d_gid;  /* n LUT ids (NULL: LUT 0) */
  uint32_t n;
  uint32_t log_lut_nb;
} luthe_pbs_batch_desc;
int luthe_pbs_batch_list(luthe_ctx* ctx, const luthe_pbs_batch_desc* d_descs, uint32_t count, size_t total,
                         void* stream);

/* Key switch + mod switch to 2N only: d_lwe2n[c*(n+1) + j] in [0, 2N)
 * (pep_key_switch: pep_ks_blram_format.sv:159-246, pep_ks_out_process.sv:276-341). */
int luthe_keyswitch_batch(luthe_ctx* ctx, const uint64_t* d_blwe, uint16_t* d_lwe2n, size_t n, void* stream);

/* Blind rotation + sample extract from mod-switched LWEs (pep_mono_mult_acc +
 * ntt_core_gf64 + decomposer); d_lwe2n rows of n+1 values. */
int luthe_blind_rotate_batch(luthe_ctx* ctx, const uint16_t* d_lwe2n, uint64_t* d_out, const uint32_t* d_gid,
                             uint32_t log_lut_nb, size_t n, void* stream);

/* Negacyclic NTT core, N = 2048 (ntt_core_gf64: ntt_core_gf64.sv:106-163,
 * ntt_core_gf64_middle.sv:201-475): d_out[b] = INTT(NTT(d_a[b]) (.) d_bn[b]) for
 * b < npoly; d_bn in the NTT domain (reference ntt() order, cut [5,6]); output
 * canonical in [0, p). */
int luthe_ntt_polymul_batch(luthe_ctx* ctx, const uint64_t* d_a, const uint64_t* d_bn, uint64_t* d_out,
                            size_t npoly, void* stream);
/* Forward / inverse transforms alone (NTT-domain in reference ntt() order). */
int luthe_ntt_forward_batch(luthe_ctx* ctx, const uint64_t* d_in, uint64_t* d_out, size_t npoly, void* stream);
int luthe_ntt_inverse_batch(luthe_ctx* ctx, const uint64_t* d_in, uint64_t* d_out, size_t npoly, void* stream);

/* Batched ALU DOp over contiguous BLWE arrays: d_out[c] = op(d_a[c], d_b[c]) with the
 * semantics of pea_alu_core.sv:55-160 (opcode = 6-bit DOp code, d_b unused by the
 * immediate forms).  Replaces the ALU processing element pe_alu for batch hosts. */
int luthe_alu_batch(luthe_ctx* ctx, uint32_t opcode, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_out,
                    uint32_t mul_factor, uint32_t msg_cst, size_t n, void* stream);

/* luthe_pbs_batch without any allocation: no hipMalloc, hipFree or allocating library call,
 * at any batch size it accepts.  Same results as luthe_pbs_batch (small batches run on the
 * one-CU latency kernel instead of the two-CU split kernel, whose exchange area is not part of
 * the fixed array); its scratch
 * (job descriptors and mod-switched LWEs) is a fixed device array of 16.1875 MiB that each GPU
 * holds for the life of the process, so batches are limited to n <= LUTHE_NOALLOC_MAX_BATCH and
 * n * lwe_stride <= LUTHE_NOALLOC_MAX_LWE_WORDS, with lwe_stride = lwe_dim + 1 rounded up to a
 * multiple of 8 (n <= 8192 at lwe_dim <= 1016); larger batches return LUTHE_ERR_ARG before any
 * work.  All calls on one GPU share that array: calls that overlap in time on one GPU (any
 * stream, context or thread) race and must be serialised by the caller.  Capturable in a HIP
 * graph without luthe_reserve.  No counterpart in the reference (its PBS unit has fixed
 * on-chip buffers); for callers that cannot allocate: code under graph capture, tight serving
 * loops, harnesses that forbid allocation. */
#define LUTHE_NOALLOC_MAX_BATCH 8192
#define LUTHE_NOALLOC_MAX_LWE_WORDS 8388608
int luthe_pbs_batch_noalloc(luthe_ctx* ctx, const uint64_t* d_in, uint64_t* d_out, const uint32_t* d_gid,
                            uint32_t log_lut_nb, size_t n, void* stream);

/* Pre-size the per-context batch scratch for batches of up to max_batch ciphertexts,
 * so later batch calls allocate nothing (required before graph capture). */
int luthe_reserve(luthe_ctx* ctx, size_t max_batch);

/* HIP graph capture of batch calls issued on the context stream (stream argument NULL):
 * begin, issue luthe_pbs_batch / luthe_alu_batch / NTT calls, end; then replay with
 * luthe_graph_launch.  DOp submissions cannot be captured.  Shapes and pointers are
 * fixed at capture time.  A graph keeps using the scratch buffers it was captured with:
 * while any graph of the context exists, scratch that a later call grows is kept (not
 * freed) and is released when the last graph is destroyed.  A graph replays only on the
 * context it was captured on; destroying the context first leaves the graph valid only for
 * luthe_graph_destroy. */
typedef struct luthe_graph luthe_graph;
int luthe_capture_begin(luthe_ctx* ctx);
int luthe_capture_end(luthe_ctx* ctx, luthe_graph** out);
int luthe_graph_launch(luthe_ctx* ctx, luthe_graph* g);
int luthe_graph_destroy(luthe_graph* g);

/* Synchronise the context stream.  LUTHE_ERR_HIP if a split blind rotation gave up waiting
 * for its partner workgroup since the last synchronising call (outputs of that launch are
 * undefined; see luthe_tuning.br_split_max). */
int luthe_sync(luthe_ctx* ctx);
/* The context's HIP stream (hipStream_t), on which DOp submissions and NULL-stream
 * batch calls run.  Batch calls share per-context scratch: calls that overlap in time
 * must be issued on the same stream. */
void* luthe_stream(luthe_ctx* ctx);
/* Blind-rotation launch width (workgroups of the persistent kernel); 0 = occupancy x CUs. */
int luthe_set_br_grid(luthe_ctx* ctx, int workgroups);
int luthe_get_br_grid(luthe_ctx* ctx);
/* Batches of at most max_cohort ciphertexts run the latency variant of the blind rotation
 * (512 threads per ciphertext, one GLWE polynomial per half); larger ones the throughput
 * kernel.  Results are identical.  Default: the variant's resident capacity (CUs x its
 * workgroups per CU: 512 at L=1, 256 at L=2 on MI355X); 0 disables it.  No
 * reference counterpart (the HPU's PBS latency is fixed by its pipeline). */
int luthe_set_br_latency_max(luthe_ctx* ctx, int max_cohort);
int luthe_get_br_latency_max(luthe_ctx* ctx);
/* Split variant (luthe_tuning.br_split_max): 1 = the two workgroups of a ciphertext exchange
 * their digits through the L2 of the XCC they share (workgroup-scope stores, the default when
 * luthe_create found every pair on one XCC; LUTHE_ERR_STATE otherwise), 0 = agent-scope
 * stores, correct wherever the pair runs.  Results are identical.  No reference counterpart
 * (the HPU has no multi-CU split). */
int luthe_set_br_split_local(luthe_ctx* ctx, int local);
int luthe_get_br_split_local(luthe_ctx* ctx);
/* Tuning knobs and the autotuner.  luthe_autotune times, with the context's keys and
 * synthetic inputs, (1) the persistent blind-rotation width (ciphertexts resident per CU,
 * batches of `batch`), (2) the largest cohort for which the latency variant is faster,
 * (3) polynomials per NTT workgroup (1, 2, 4), (4) the smallest batch for which the
 * matrix-core key switch beats the VALU one, and sets (5) the serving cohort size to one
 * wave of the chosen width; applies them and writes them to json_path (optional).
 * luthe_load_tuning applies a saved file (refused if tuned for another PBS level).
 * No reference counterpart: the HPU's equivalents are build-time parameters
 * (docs/hpu_parameters.md). */
typedef struct luthe_tuning {
  int32_t br_grid;          /* persistent blind-rotation workgroups, 0 = occupancy x CUs */
  int32_t br_latency_max;   /* largest cohort for the latency variant, -1 = its capacity */
  int32_t ntt_polys_per_wg; /* 1, 2 or 4 */
  int32_t cohort_max;       /* serving cohort size, 0 = blind-rotation width */
  int32_t br_pair;          /* throughput kernel: 0 = 256 threads per ciphertext (up to 4 per CU),
                               1 = 512 threads per ciphertext (2 per CU), -1 = default (1 at PBS
                               level 2, 0 at level 1) */
  int32_t br_split_max;     /* largest cohort for the split variant (one 512-thread workgroup per
                               GLWE polynomial, two CUs per ciphertext), -1 = CUs / 2, 0 = off */
  int32_t br_split_wait_us; /* split variant: longest wait of a workgroup for its partner before it
                               gives up (LUTHE_ERR_HIP at the next synchronising call, outputs of
                               that launch undefined); -1 = 2 s, 0 = give up at the first missing
                               word (for tests of that path).  Start from luthe_get_tuning. */
  int32_t ks_mfma_min;      /* smallest batch whose key switch runs on the int8 matrix cores (the
                               KSK as three balanced byte planes, exact int32 sums); smaller batches
                               take the VALU kernel.  -1 = default (CUs); a value above every
                               batch size turns it off.  Results are identical. */
} luthe_tuning;
int luthe_autotune(luthe_ctx* ctx, uint32_t batch, const char* json_path, luthe_tuning* out);
int luthe_load_tuning(luthe_ctx* ctx, const char* json_path);
int luthe_get_tuning(luthe_ctx* ctx, luthe_tuning* out);
int luthe_set_tuning(luthe_ctx* ctx, const luthe_tuning* t);

/* Diagnostics: throughput of the Goldilocks modular multiply (64x64 -> 128-bit product
 * plus reduction) at full occupancy, in 10^9 modmul/s; the VALU roofline denominator of
 * the blind rotation.  No reference counterpart (the HPU's rate is fixed by its RTL). */
int luthe_microbench_modmul(luthe_ctx* ctx, double* gmodmul_per_s);

/* ---------------------------------------------------------------- client-side helpers
 * Key generation, encryption, decryption and LUT encoding live in tfhe-rs for the
 * reference (not part of the checkout); these builder-defined, seeded host routines
 * play that role for tests and benchmarks.  Secret keys are arrays of 0/1 u64. */

int luthe_keygen_secret(const luthe_params* p, uint64_t seed, uint64_t* lwe_sk, uint64_t* glwe_sk);
/* bsk: [n][k+1][L][k+1][N] coefficient domain (LUTHE_BSK_COEF). */
int luthe_keygen_bsk(const luthe_params* p, uint64_t seed, const uint64_t* lwe_sk, const uint64_t* glwe_sk,
                     uint64_t* bsk, int threads);
int luthe_keygen_ksk(const luthe_params* p, uint64_t seed, const uint64_t* lwe_sk, const uint64_t* glwe_sk,
                     uint32_t* ksk);
/* BLWE encryption of raw plaintexts pt[c] (already scaled into the top bits) under
 * the flattened GLWE key with GLWE noise; stream `seed`. */
int luthe_encrypt_blwe(const luthe_params* p, uint64_t seed, const uint64_t* glwe_sk, const uint64_t* pt, size_t n,
                       uint64_t* blwe);
/* Phase b - <a, s> mod 2^64 of n BLWEs. */
int luthe_decrypt_blwe(const luthe_params* p, const uint64_t* glwe_sk, const uint64_t* blwe, size_t n,
                       uint64_t* phase);
/* Message scale Delta = 2^64 / 2^(msg_bits + carry_bits + 1) (padding bit). */
uint64_t luthe_delta(const luthe_params* p);
/* Decode a phase to the message in [0, 2^(msg+carry+1)). */
uint32_t luthe_decode(const luthe_params* p, uint64_t phase);
/* Trivial GLWE LUT of table[m] (m < 2^(msg+carry)), boxes of N / 2^(msg+carry)
 * coefficients, shifted by half a box (negacyclic) so noisy phases round to m. */
int luthe_lut_encode(const luthe_params* p, const uint8_t* table, uint32_t table_len, uint64_t* glwe);
/* Many-LUT GLWE: 2^log_lut_nb functions of an input restricted to
 * 2^(msg+carry-log_lut_nb) values; function j lands on output lut j. */
int luthe_lut_encode_many(const luthe_params* p, const uint8_t* tables, uint32_t table_len, uint32_t log_lut_nb,
                          uint64_t* glwe);

/* ---------------------------------------------------------------- HPU memory images (host only)
 * Conversions between this engine's layouts and the HPU's HBM images, so keys, LUTs and
 * ciphertexts laid out for the FPGA load directly (and results can be written back).
 *  - BSK: per CMUX i, output column j, stage iteration s < N/chunk, input row g, gadget
 *    level l (least significant first), `chunk` = PSI*R consecutive NTT-domain
 *    coefficients, the polynomial in the HW NTT order bsk_order(ntt_cut)
 *    (hw/module/hpu/simu/scripts/gen_stimuli.py:108-136, sw/ntt/ntt_gf64_bsk_order.py:9-35;
 *    the header of bsk_manager.sv notes the reversed (R,N) order).
 *  - LUT: body polynomials only (the mask of a trivial GLWE is zero), N u64 per gid
 *    (pep_load_glwe.sv:9,83,204).
 *  - Ciphertexts: pem_pc pseudo-channel regions of count * slot_bytes bytes; slot s at
 *    s * slot_bytes (page aligned, pem_common_param_pkg.sv:24-36); register-file word w
 *    sends coefficients [w*RC + p*RC/PC, w*RC + (p+1)*RC/PC) to channel p; the body is
 *    the first u64 of one extra AXI word of channel 0 (pem_load.sv:90-156).
 *  - KSK: see luthe_hpu_ksk_layout below. */
typedef struct luthe_hpu_layout {
  uint32_t axi_bytes;    /* AXI4 data width of the ciphertext channels in bytes (8..64) */
  uint32_t pem_pc;       /* PEM_PC: ciphertext pseudo-channels, 1 or 2 */
  uint32_t regf_coef_nb; /* REGF_COEF_NB: coefficients per register-file word */
  uint32_t bsk_chunk;    /* PSI * R of the HW NTT: coefficients per BSK image chunk */
  uint32_t ntt_ncut;     /* radix stages of the HW NTT (1..8) */
  uint32_t ntt_cut[8];   /* log2 radix per stage, summing to log2 N */
} luthe_hpu_layout;

/* order[pos] = natural NTT-domain index (evaluation at zeta^(2m+1)) held at position pos
 * by an NTT split as `cut`; [5,6] is this engine's NTT-domain order. */
int luthe_hpu_bsk_order(const uint32_t* cut, uint32_t ncut, uint32_t* order);
size_t luthe_hpu_bsk_image_words(const luthe_params* p);
/* image <-> LUTHE_BSK_NTT layout ([n][k+1][L][k+1][N], ntt() order of cut [5,6]). */
int luthe_hpu_bsk_from_image(const luthe_params* p, const luthe_hpu_layout* lay, const uint64_t* image,
                             uint64_t* bsk_ntt);
int luthe_hpu_bsk_to_image(const luthe_params* p, const luthe_hpu_layout* lay, const uint64_t* bsk_ntt,
                           uint64_t* image);
/* Upload a BSK given as an HPU image (converted on the host, then luthe_upload_bsk). */
int luthe_upload_bsk_hpu_image(luthe_ctx* ctx, const luthe_hpu_layout* lay, const uint64_t* image);
/* CT_MEM_BYTES: bytes per ciphertext slot and pseudo-channel (0 if the layout is invalid). */
size_t luthe_hpu_ct_slot_bytes(const luthe_params* p, const luthe_hpu_layout* lay);
int luthe_hpu_ct_to_image(const luthe_params* p, const luthe_hpu_layout* lay, const uint64_t* blwe, uint32_t count,
                          uint8_t* image);
int luthe_hpu_ct_from_image(const luthe_params* p, const luthe_hpu_layout* lay, const uint8_t* image, uint32_t count,
                            uint64_t* blwe);
/* LUT bodies (count x N) -> full GLWEs (count x (k+1)N, zero mask). */
int luthe_hpu_lut_from_image(const luthe_params* p, const uint64_t* bodies, uint32_t count, uint64_t* glwe);
int luthe_write_lut_hpu_image(luthe_ctx* ctx, uint32_t gid, const uint64_t* bodies, uint32_t count);

/* KSK image: one byte stream per KSK pseudo-channel, in the order ksk_if reads the key
 * (hw/module/ksk/module/ksk_if/simu/rtl/tb_ksk_if.sv:470-525): for block column c < ceil((n+1)/
 * LBX), column x < LBX, block line l < ceil(k*N/LBY), level group lg < ceil(KS_L/LBZ), line y of
 * the channel's cuts (channel pc owns y in [pc*LBY/KSK_PC, (pc+1)*LBY/KSK_PC)), level z < LBZ:
 * one ks_mod_log-bit coefficient; each (c, x) run is zero-padded to whole AXI words; an AXI word
 * of axi_bytes (little-endian) holds axi_bytes*8 / KSK_ACS_W slots of LBZ coefficients, KSK_ACS_W
 * = 64 if LBZ * ks_mod_log > 32 else 32 (ksk_if_common_param_pkg.sv:20-21).  Output column
 * j = c*LBX + x (j = n: the body), level lg*LBZ + z (this engine's level order, least
 * significant digit first), and line l*LBY + y = g*N + (N-1-m) holds mask coefficient m of
 * GLWE polynomial g ("reverse order in the BLWE_K dimension, on N basis", ksk_manager.sv:11).
 * Requires KSK_CUT_NB % KSK_PC == 0 and LBY % KSK_CUT_NB == 0 (the shipped configurations). */
typedef struct luthe_hpu_ksk_layout {
  uint32_t axi_bytes;  /* AXI4 data width of the KSK channels in bytes (8..128) */
  uint32_t ksk_pc;     /* KSK_PC: pseudo-channels */
  uint32_t ksk_cut_nb; /* KSK_CUT_NB: cuts of the LBY lines over the channels */
  uint32_t lbx, lby, lbz; /* key-switch block: columns, lines, levels (pep_ks_common_param_pkg.sv:25-35) */
} luthe_hpu_ksk_layout;
/* Bytes of one channel's image (0 if the layout is invalid for these parameters). */
size_t luthe_hpu_ksk_image_bytes(const luthe_params* p, const luthe_hpu_ksk_layout* lay);
/* ksk: this engine's [k*N][ks_level][n+1] layout; images[pc]: ksk_pc buffers of
 * luthe_hpu_ksk_image_bytes each. */
int luthe_hpu_ksk_to_image(const luthe_params* p, const luthe_hpu_ksk_layout* lay, const uint32_t* ksk,
                           uint8_t* const* images);
int luthe_hpu_ksk_from_image(const luthe_params* p, const luthe_hpu_ksk_layout* lay, const uint8_t* const* images,
                             uint32_t* ksk);
/* Upload a KSK given as an HPU image (converted on the host, then luthe_upload_ksk). */
int luthe_upload_ksk_hpu_image(luthe_ctx* ctx, const luthe_hpu_ksk_layout* lay, const uint8_t* const* images);

/* ---------------------------------------------------------------- trace
 * Instruction-scheduler trace records (isc_trace_t, isc_common_param_pkg.sv:35-48):
 * timestamp | DOp word | query command | pool state (sync_id, rd/wr/issue locks, vld,
 * rd_pdg, pdg) | pe_reserved, packed most significant field first into
 * luthe_trace_record_bits() bits and stored as 32-bit words, least significant word
 * first (trace_manager.sv:245-264).  The DOp path of a context logs REFILL when a DOp is
 * accepted, ISSUE when its dependency level is launched and RETIRE when its submission
 * completes; timestamps are microseconds since luthe_trace_enable (the HPU counts clock
 * cycles), pe_reserved of a PBS RETIRE holds the number of PBS launched with it. */
enum { LUTHE_TRC_NONE = 0, LUTHE_TRC_RDUNLOCK = 1, LUTHE_TRC_RETIRE = 2, LUTHE_TRC_REFILL = 3, LUTHE_TRC_ISSUE = 4 };
typedef struct luthe_trace_event {
  uint32_t timestamp;
  uint32_t insn;
  uint8_t cmd, sync_id, rd_lock, wr_lock, issue_lock, vld, rd_pdg, pdg;
  uint32_t pe_reserved;
} luthe_trace_event;
int luthe_trace_record_bits(uint32_t lwe_dim);
/* Returns the words per record (> 0) or an error. */
int luthe_trace_pack(uint32_t lwe_dim, const luthe_trace_event* ev, size_t n, uint32_t* words);
int luthe_trace_unpack(uint32_t lwe_dim, const uint32_t* words, size_t n, luthe_trace_event* ev);
/* Keep the last `capacity` events of the DOp path (0 = off; the HPU's trace buffer wraps). */
int luthe_trace_enable(luthe_ctx* ctx, size_t capacity);
/* Copy out up to cap events, oldest first, and drop them from the log. */
int luthe_trace_read(luthe_ctx* ctx, luthe_trace_event* out, size_t cap, size_t* n);

/* ---------------------------------------------------------------- radix IOp programs
 * Builder-written IOp -> DOp-level programs on 2-bit blocks (2 message + 2 carry bits),
 * the role of the tfhe-rs "fw" translation programs that feed the reference firmware
 * (fw/ublaze/src/ucore_fw.c:54-128; not part of the checkout).  Operands are radix
 * integers of nblocks blocks, least significant block first.  Host-only calls. */
enum {
  LUTHE_IOP_ADD = 0,        /* (a + b) mod 4^n, carry-lookahead (Sklansky prefix over carry states) */
  LUTHE_IOP_CMP_GT = 1,     /* one block: a > b (comparison tree, docs/debug.md:284-350,430-490) */
  LUTHE_IOP_MUL = 2,        /* (a * b) mod 4^n: partial products, column reduction, carry-lookahead add */
  LUTHE_IOP_ADD_RIPPLE = 3, /* (a + b) mod 4^n, ripple carry with the Many-LUT msg/carry split */
  LUTHE_IOP_COUNT = 4
};
#define LUTHE_IOP_LUTS 11 /* logical LUTs the programs use (luthe_iop_lut) */

typedef struct luthe_iop_info {
  uint32_t nblocks, rows, nodes, npbs, nalu, depth, nout, a_row0, b_row0;
} luthe_iop_info;

/* One node: is_pbs ? dst..dst+2^log_lut_nb-1 = PBS(src0, lut) : dst = ALU(opcode, src0, src1, mul_factor).
 * Rows are single-assignment indices into the request workspace. */
typedef struct luthe_iop_node {
  uint8_t is_pbs, opcode, lut, log_lut_nb;
  uint32_t dst, src0, src1, mul_factor;
} luthe_iop_node;

/* Shape of a program, and optionally its nodes (topological order) and output rows. */
int luthe_iop_program(uint32_t op, uint32_t nblocks, luthe_iop_info* info, luthe_iop_node* nodes, size_t nodes_cap,
                      uint32_t* out_rows, size_t out_cap);
/* Table of logical LUT `lut` on 4-bit inputs (Many-LUT: 2^log_lut_nb tables of 16 >> log_lut_nb entries). */
int luthe_iop_lut(uint32_t lut, uint8_t* table16, uint32_t* log_lut_nb);
/* Plaintext evaluation of a program on block values (PBS = table lookup): the host check of its logic. */
int luthe_iop_eval_plain(uint32_t op, uint32_t nblocks, const uint32_t* a, const uint32_t* b, uint32_t* out);

/* IOp -> DOp template (the translation-table entry the reference firmware patches,
 * ucore_fw.c:135-220): sources TS[0].i = a, TS[1].i = b, results TD[0].i, spills to heap
 * slots (TH), PBS on LUT ids lut_base + logical LUT, for a register file of `regs`
 * registers.  DOps are issued level by level (the last PBS of a level carries the flush
 * bit, docs/dop.md:108-111); Many-LUT destinations are aligned.  *n = words written
 * (call with words = NULL to size).  The IOp opcode is the reference's (docs/iop.md:176-196:
 * ADD 0xE0, CMP_GT 0xC0, MUL 0xE4; ripple ADD uses user opcode 0x60) with alignment
 * 2 * nblocks bits. */
typedef struct luthe_iop_dop_stats {
  uint32_t dops, ld, st, alu, pbs, heap_slots, levels;
  uint32_t iop_opcode, align_bits;
} luthe_iop_dop_stats;
int luthe_iop_program_dops(uint32_t op, uint32_t nblocks, uint32_t regs, uint32_t lut_base, uint32_t* words,
                           size_t cap, size_t* n, luthe_iop_dop_stats* stats);
/* Write the programs' LUTs into LUT slots lut_base .. lut_base + LUTHE_IOP_LUTS - 1. */
int luthe_write_iop_luts(luthe_ctx* ctx, uint32_t lut_base);
/* luthe_iop_program_dops for the context's register file + luthe_register_iop under the
 * program's IOp opcode and alignment. */
int luthe_register_iop_program(luthe_ctx* ctx, uint32_t op, uint32_t nblocks, uint32_t lut_base);

/* ---------------------------------------------------------------- serving engine
 * Continuous batching of radix IOps.  Replaces the backend-to-HPU IOp queue and ack
 * (ucore_fw.c:54-232), the instruction scheduler's out-of-order issue
 * (hw/module/instruction_scheduler/rtl/instruction_scheduler.sv) and the BPIP batch
 * former with fill threshold and timeout (pep_sequencer.sv:584-620, bpip registers at
 * hpu_regif_core_cfg_1in3.toml:240-260).  A server thread issues its steps on the context
 * stream.  While a server is attached, DOp/IOp ticket submissions keep working (their groups
 * are launched on the same stream with scratch of their own, between the server's steps; they
 * do not join its cohorts); batch calls, graph capture, the autotuner, key uploads and launch
 * settings (tuning, grid widths) return LUTHE_ERR_BUSY, and luthe_destroy refuses.  Each request runs in an HBM workspace admitted
 * from an arena; each step launches the ready PBS of all admitted requests as one cohort.
 * Results do not depend on which requests share a cohort. */
typedef struct luthe_server luthe_server;

typedef struct luthe_server_config {
  uint64_t arena_bytes;       /* HBM for request workspaces (admission budget); default 4 GiB */
  uint64_t pinned_bytes;      /* pinned host staging for operands/results; default 256 MiB */
  uint32_t cohort_max;        /* PBS per cohort; 0 = blind-rotation grid width (one wave) */
  uint32_t cohort_min;        /* launch a cohort once this many PBS are ready ... (default 1) */
  uint32_t cohort_timeout_us; /* ... or when the oldest ready PBS has waited this long */
  uint32_t pipeline_depth;    /* steps in flight on the stream, 1..4 (default 2) */
  uint32_t lut_base;          /* first of LUTHE_IOP_LUTS LUT slots written at creation;
                                 0xFFFFFFFF = the top LUTHE_IOP_LUTS slots */
  uint32_t max_inflight;      /* admitted requests at once; 0 = limited by the arena only */
  uint32_t queue_max;         /* submitted-but-not-admitted requests; beyond: LUTHE_ERR_BUSY */
  uint32_t policy;            /* which ready PBS fill a cohort: LUTHE_COHORT_* (default SHORTEST) */
  uint32_t age_limit_us;      /* SHORTEST: PBS ready this long go first whatever their height
                                 (no starvation of long requests); 0 = never (default 500 ms) */
} luthe_server_config;
/* Cohort policies.  OLDEST: ready PBS in arrival order (the HPU's PBS pool order).  SHORTEST:
 * the PBS with the fewest PBS levels left on their request's critical path first (short
 * requests finish without waiting behind the fan-out of long ones), subject to age_limit_us. */
enum { LUTHE_COHORT_OLDEST = 0, LUTHE_COHORT_SHORTEST = 1 };

/* Counters of the mockup's per-IOp report (docs/debug.md:149-170: PBS issued, batches,
 * by_timeout, usage) plus admission and latency figures since creation or the last reset. */
typedef struct luthe_server_stats {
  uint64_t submitted, admitted, completed, rejected;
  uint64_t steps, cohorts, cohorts_by_timeout, pbs, alu, alu_launches;
  uint64_t fill_hist[8];      /* cohorts by fill: bucket floor(8 * size / cohort_max), full = 7 */
  uint64_t arena_rows, arena_rows_used, arena_rows_peak;
  uint32_t cohort_max, lut_base, pending, inflight;
  double latency_p50_us, latency_p99_us, latency_max_us; /* submit -> result on the host */
} luthe_server_stats;

typedef struct luthe_request_info {
  uint32_t op, nblocks, npbs, depth, cohorts; /* cohorts = PBS cohorts the request took part in */
  int32_t status;
  double queue_us;   /* submit -> admission */
  double latency_us; /* submit -> result on the host */
} luthe_request_info;

void luthe_server_config_default(luthe_server_config* cfg);
/* Requires uploaded keys and msg_bits = carry_bits = 2; writes the program LUTs. */
int luthe_server_create(luthe_ctx* ctx, const luthe_server_config* cfg, luthe_server** out);
/* Completes all submitted requests, then stops.  Results not collected are dropped. */
int luthe_server_destroy(luthe_server* srv);
const char* luthe_server_last_error(luthe_server* srv);
/* Queue IOp `op` on a, b (host, nblocks BLWE each; copied before return).  `out` (host,
 * nblocks BLWE, or 1 for CMP_GT) is written before the request reports done and must stay
 * valid until then.  LUTHE_ERR_NOMEM: larger than the arena or staging; LUTHE_ERR_BUSY:
 * queue full. */
int luthe_server_submit(luthe_server* srv, uint32_t op, uint32_t nblocks, const uint64_t* a, const uint64_t* b,
                        uint64_t* out, uint64_t* id);
/* 1 = done (request released, info filled), 0 = pending, < 0 = error. */
int luthe_server_poll(luthe_server* srv, uint64_t id, luthe_request_info* info);
/* Wait for a request (timeout_us < 0: no limit); releases it. */
int luthe_server_wait(luthe_server* srv, uint64_t id, int64_t timeout_us, luthe_request_info* info);
/* Launch the ready PBS at the next step even below cohort_min (the PBS_F flush, docs/dop.md:108-111). */
int luthe_server_flush(luthe_server* srv);
int luthe_server_get_stats(luthe_server* srv, luthe_server_stats* out, int reset);

#ifdef __cplusplus
}
#endif
#endif
