"""Benchmark: batched PBS throughput on one MI355X.

Workload (one "step"): one batched PBS (key switch + blind rotation + sample extract)
over 8192 ciphertexts with n=724, k=1, N=2048, PBS base 2^23 level 1, KS base 2^3
level 5, w_ks=21, seeded keys, uniform 4-bit plaintexts, identity
LUT; inputs resident in HBM before the timed region.  Beside it:
  * the plain path: the same kernels with one ciphertext per launch;
  * the blind-rotation kernel's own duration (HIP events on the engine stream) and its
    roofline: algorithmic modmul-equivalents per launch / duration against the measured
    full-occupancy rate of the same Goldilocks modmul primitive;
  * the NTT product (4096 polynomials: NTT, pointwise product, INTT), uneven
    graph-captured batches (64 batches, PBS level 2) and serving (Poisson arrivals of radix
    IOps through the continuous-batching server, PBS/s with p50/p99 latency per IOp);
  * cpu_baseline: the C oracle (oracle/luthe_ref.c, OpenMP over ciphertexts) timed on a
    bounded sample of the same batch; its outputs are compared bit for bit with the GPU's.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "PBS/sec on 1 MI355X (batched vs plain), bit-exact; p50/p99 IOp latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class Timer:
    """HIP events recorded on a given stream."""

    def __init__(self, hip, stream):
        self.hip, self.stream = hip, stream
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        hip.hipEventCreate(C.byref(self.e0))
        hip.hipEventCreate(C.byref(self.e1))

    def start(self):
        self.hip.hipEventRecord(self.e0, self.stream)

    def stop_ms(self):
        self.hip.hipEventRecord(self.e1, self.stream)
        self.hip.hipEventSynchronize(self.e1)
        ms = C.c_float(0)
        self.hip.hipEventElapsedTime(C.byref(ms), self.e0, self.e1)
        return ms.value


def modmul_equiv_per_pbs(p):
    """Algorithmic Goldilocks modmul-equivalents of one blind rotation: per CMUX,
    (k+1)L forward + (k+1) inverse NTTs of 11,264 butterflies each, (k+1)^2 L pointwise
    products per NTT-domain index, and (k+1)N p->2^64 switch multiplies."""
    k1, L, N = p.glwe_k + 1, p.pbs_level, p.poly_size
    per_cmux = (k1 * L + k1) * 11264 + k1 * L * k1 * N + k1 * N
    return p.lwe_dim * per_cmux


def bsk_bytes(p):
    k1 = p.glwe_k + 1
    return p.lwe_dim * k1 * p.pbs_level * k1 * p.poly_size * 8


def run_ntt_product(la, hip, ctx, peak, npoly=4096, reps=5):
    """NTT -> pointwise multiply by an NTT-domain set -> INTT, 4096 polynomials x 2048."""
    rng = np.random.default_rng(1)
    a = rng.integers(0, la.P, (npoly, 2048), dtype=np.uint64)
    bn = rng.integers(0, la.P, (npoly, 2048), dtype=np.uint64)
    da, db = la.DeviceBuffer.from_numpy(a), la.DeviceBuffer.from_numpy(bn)
    do = la.DeviceBuffer((npoly, 2048), np.uint64)
    tm = Timer(hip, ctx.stream)
    la.lib().luthe_ntt_polymul_batch(ctx.handle, da.ptr, db.ptr, do.ptr, npoly, None)
    ctx.sync()
    tm.start()
    for _ in range(reps):
        la.lib().luthe_ntt_polymul_batch(ctx.handle, da.ptr, db.ptr, do.ptr, npoly, None)
    ms = tm.stop_ms() / reps
    work = npoly * (2 * 11264 + 2048)
    nbytes = 3 * npoly * 2048 * 8
    return {"workload": f"NTT, pointwise x NTT-domain operand, INTT: {npoly} polynomials x 2048, uniform < p",
            "ms_per_batch": round(ms, 4), "polys_per_s": round(npoly / (ms * 1e-3), 1),
            "hbm_gbps": round(nbytes / (ms * 1e-3) / 1e9, 1), "hbm_frac": round(nbytes / (ms * 1e-3) / 8e12, 4),
            "gmodmul_per_s": round(work / (ms * 1e-3) / 1e9, 1), "valu_frac": round(work / (ms * 1e-3) / 1e9 / peak, 4)}


def run_uneven_batches(la, hip, peak, seed=3):
    """64 batches, sizes uniform in [1, 4096]; per batch an ADD and a MAC then a
    PBS with a per-ciphertext LUT (random 16-entry table); PBS level 2, base 2^15; the
    whole 64-batch program captured as one HIP graph and replayed.  Two captures of the same
    program: the PBS of all 64 batches as one batch list (one key switch and one persistent
    blind-rotation launch, luthe_pbs_batch_list; the figure reported), and one PBS launch per
    batch (for comparison).  Both outputs must agree bit for bit."""
    prm = la.default_params(pbs_base_log=15, pbs_level=2, lut_slots=4096)
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(1, 4097, 64)]
    total = sum(sizes)
    log(f"bench: uneven batches, {total} PBS in 64 batches, keygen L=2")
    keys = la.Keys(prm, seed=5678)
    ctx = la.Context(prm, device=0)
    ctx.upload_bsk(keys.bsk)
    ctx.upload_ksk(keys.ksk)
    tables = np.random.default_rng(seed + 1).integers(0, 16, (4096, 16), dtype=np.uint8)
    luts = np.stack([la.lut_encode(prm, t) for t in tables])
    ctx.write_lut(0, luts)
    ma, mb = rng.integers(0, 4, total), rng.integers(0, 4, total)
    dA = la.DeviceBuffer.from_numpy(keys.encrypt_msg(ma, seed=11))
    dB = la.DeviceBuffer.from_numpy(keys.encrypt_msg(mb, seed=12))
    dT = la.DeviceBuffer((total, prm.blen), np.uint64)
    dU = la.DeviceBuffer((total, prm.blen), np.uint64)
    dO = la.DeviceBuffer((total, prm.blen), np.uint64)
    gids = np.concatenate([np.arange(s, dtype=np.uint32) for s in sizes])
    dG = la.DeviceBuffer.from_numpy(gids)
    b8 = prm.blen * 8

    dO2 = la.DeviceBuffer((total, prm.blen), np.uint64)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(int)
    descs = la.DeviceBuffer.from_numpy(la.batch_descs(
        [(dU.ptr.value + int(o) * b8, dO.ptr.value + int(o) * b8, dG.ptr.value + int(o) * 4, s, 0)
         for o, s in zip(starts, sizes)]))
    ctx.reserve(total)

    def program(listed):
        off = 0
        for s in sizes:
            a_p, b_p, t_p = dA.ptr.value + off * b8, dB.ptr.value + off * b8, dT.ptr.value + off * b8
            u_p, o_p, g_p = dU.ptr.value + off * b8, dO2.ptr.value + off * b8, dG.ptr.value + off * 4
            ctx.alu_batch_dev(0x01, a_p, b_p, t_p, s)
            ctx.alu_batch_dev(0x05, t_p, b_p, u_p, s, mul_factor=2)
            if not listed:
                ctx.pbs_batch_dev(u_p, o_p, g_p, s)
            off += s
        if listed:
            ctx.pbs_batch_list_dev(descs, len(sizes), total)

    tm = Timer(hip, ctx.stream)
    times = {}
    for listed in (True, False):
        ctx.capture_begin()
        program(listed)
        g = ctx.capture_end()
        g.launch()          # warm
        ctx.sync()
        tm.start()
        g.launch()
        times[listed] = tm.stop_ms()
        g.destroy()
    ms = times[True]
    out = dO.numpy()
    same = bool(np.array_equal(out, dO2.numpy()))
    exp = tables[gids, 2 * (ma + mb) + mb]   # MAC: (a + b) * 2 + b
    ok = float((keys.decrypt_msg(out) == exp).mean())
    ctx.close()
    work = total * modmul_equiv_per_pbs(prm)
    res = {"workload": "64 uneven batches (sizes uniform in [1,4096]), ADD+MAC then PBS with per-ciphertext "
                       "LUTs, PBS 2^15 x 2, one HIP graph",
           "pbs": total, "ms": round(ms, 2), "pbs_per_s": round(total / (ms * 1e-3), 1),
           "decrypt_success": round(ok, 4),
           "pbs_launch": "one batch list (luthe_pbs_batch_list) for the 64 batches",
           "per_batch_launches_ms": round(times[False], 2),
           "per_batch_launches_pbs_per_s": round(total / (times[False] * 1e-3), 1),
           "list_equals_per_batch": same,
           "roofline": {"bound": "valu", "achieved": round(work / (ms * 1e-3) / 1e9, 2), "peak": round(peak, 2),
                        "unit": "Gmodmul/s", "frac": round(work / (ms * 1e-3) / 1e9 / peak, 4),
                        "note": "whole graph (ALU, key switch, blind rotation) against the blind rotation's "
                                "modmul-equivalents at PBS level 2"}}
    return res


def _pct(v, q):
    return float(np.percentile(np.asarray(v), q, method="nearest")) if len(v) else None


def run_serving(la, ctx, keys, rates=(8, 20, 40, 60), seconds=5.0, burst=96, seed=4, cohort_max=0):
    """Serving.  Poisson arrivals (seeded) of radix IOps drawn uniformly from
    {ADD, CMP_GT, MUL} x {8, 16, 32} blocks of 2 bits, through the continuous-batching
    server (HBM arena admission, one PBS cohort per step); latency = submit -> result on
    the host.  Plus a burst of `burst` requests submitted at once (saturation PBS/s)."""
    prm = ctx.params
    mix = [(op, nb) for op in (la.IOP_ADD, la.IOP_CMP_GT, la.IOP_MUL) for nb in (8, 16, 32)]
    npbs = {k: la.iop_program(*k)[0].npbs for k in mix}
    rng = np.random.default_rng(seed)
    pool_a = keys.encrypt_msg(rng.integers(0, 4, 32), seed=41)
    pool_b = keys.encrypt_msg(rng.integers(0, 4, 32), seed=42)
    name = {la.IOP_ADD: "ADD", la.IOP_CMP_GT: "CMP_GT", la.IOP_MUL: "MUL"}
    mean_pbs = float(np.mean([npbs[k] for k in mix]))
    res = {"workload": "Poisson arrivals of radix IOps, uniform over {ADD, CMP_GT, MUL} x {8, 16, 32} blocks x 2 "
                       "bits, n=724 PBS 2^23 x 1 parameters, keys resident, continuous batching",
           "pbs_per_iop": {f"{name[o]}{n}": npbs[(o, n)] for o, n in mix}, "mean_pbs_per_request": mean_pbs}
    srv = la.Server(ctx, cohort_max=cohort_max)
    res["cohort_max"] = srv.stats()["cohort_max"]
    for k in mix:  # warm every program once
        srv.wait(srv.submit(k[0], pool_a[:k[1]], pool_b[:k[1]]))

    def summarize(infos, wall, st):
        lat = [i["latency_us"] / 1e3 for _, _, i in infos]
        by = {}
        for op, nb, i in infos:
            by.setdefault(f"{name[op]}{nb}", []).append(i["latency_us"] / 1e3)
        total_pbs = sum(npbs[(op, nb)] for op, nb, _ in infos)
        return {"requests": len(infos), "wall_s": round(wall, 3), "pbs": total_pbs,
                "pbs_per_s": round(total_pbs / wall, 1), "iops_per_s": round(len(infos) / wall, 2),
                "p50_ms": round(_pct(lat, 50), 2), "p99_ms": round(_pct(lat, 99), 2),
                "per_iop_ms": {k: {"n": len(v), "p50": round(_pct(v, 50), 2), "p99": round(_pct(v, 99), 2)}
                               for k, v in sorted(by.items())},
                "cohorts": st["cohorts"], "mean_cohort": round(st["pbs"] / max(1, st["cohorts"]), 1),
                "cohort_fill_hist": st["fill_hist"], "arena_rows_peak": st["arena_rows_peak"]}

    # burst: all at once
    kinds = rng.integers(0, len(mix), burst)
    srv.stats(reset=True)
    t0 = time.perf_counter()
    ids = [(mix[k][0], mix[k][1], srv.submit(mix[k][0], pool_a[:mix[k][1]], pool_b[:mix[k][1]])) for k in kinds]
    infos = [(op, nb, srv.wait(rid)[1]) for op, nb, rid in ids]
    res["burst"] = summarize(infos, time.perf_counter() - t0, srv.stats())
    log(f"bench: serving burst of {burst}: {res['burst']['pbs_per_s']} PBS/s, "
        f"p50 {res['burst']['p50_ms']} ms, p99 {res['burst']['p99_ms']} ms")
    res["rates"] = []
    for rate in rates:
        r2 = np.random.default_rng(seed * 1000 + rate)
        n = max(8, int(rate * seconds))
        gaps = r2.exponential(1.0 / rate, n)
        kinds = r2.integers(0, len(mix), n)
        srv.stats(reset=True)
        ids = []
        t0 = time.perf_counter()
        t = t0
        for i in range(n):
            t += gaps[i]
            dt = t - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            op, nb = mix[kinds[i]]
            ids.append((op, nb, srv.submit(op, pool_a[:nb], pool_b[:nb])))
        infos = [(op, nb, srv.wait(rid)[1]) for op, nb, rid in ids]
        row = summarize(infos, time.perf_counter() - t0, srv.stats())
        row["offered_iops_per_s"] = rate
        row["offered_load"] = round(rate * mean_pbs / max(1.0, res["burst"]["pbs_per_s"]), 3)
        res["rates"].append(row)
        log(f"bench: serving rate {rate}/s: {row['pbs_per_s']} PBS/s, p50 {row['p50_ms']} ms, "
            f"p99 {row['p99_ms']} ms, mean cohort {row['mean_cohort']}")
    srv.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--plain-reps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--br-grid", type=int, default=0, help="blind-rotation workgroups (0 = occupancy x CUs)")
    ap.add_argument("--no-extra", action="store_true", help="skip the NTT-product, uneven-batch and serving measurements")
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="arrival window per rate (serving)")
    ap.add_argument("--serve-cohort-max", type=int, default=0, help="PBS per cohort (serving; 0 = server default)")
    ap.add_argument("--extras", default="ntt,uneven,serving",
                    help="extra workloads to run beside the headline (ntt, uneven, serving)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library defaults)")
    ap.add_argument("--tuning-json", default="", help="where the autotuner writes its choices")
    ap.add_argument("--tuning-file", default="", help="apply a saved tuning instead of running the autotuner")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this engine shards by replicating keys and splitting "
              f"batches; no multi-GPU figure is produced): --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import luthe_amd as la
    hip = la.hip()

    if args.calib_copy_mib:
        nbytes = args.calib_copy_mib << 20
        a = la.DeviceBuffer((nbytes // 8,), np.uint64)
        b = la.DeviceBuffer((nbytes // 8,), np.uint64)
        a.zero()
        for _ in range(3):
            hip.hipMemcpy(b.ptr, a.ptr, nbytes, 3)
        hip.hipDeviceSynchronize()
        print(json.dumps({"calib_copy_bytes": nbytes}))
        return

    prm = la.default_params()
    log(f"bench: batched PBS batch={args.batch} n={prm.lwe_dim} N={prm.poly_size}: keygen")
    keys = la.Keys(prm, seed=1234)
    ctx = la.Context(prm, device=0)
    if args.br_grid:
        ctx.set_br_grid(args.br_grid)
    ctx.upload_bsk(keys.bsk, la.BSK_COEF)
    ctx.upload_ksk(keys.ksk)
    lut = la.lut_encode(prm, np.arange(16, dtype=np.uint8))
    ctx.write_lut(0, lut)
    tuning = None
    if args.tuning_file:
        ctx.load_tuning(args.tuning_file)
        tuning = ctx.tuning()
        log(f"bench: tuning from {args.tuning_file}: {tuning}")
    elif not args.no_tune and not args.br_grid:
        t0 = time.perf_counter()
        tuning = ctx.autotune(args.batch, args.tuning_json or None)
        log(f"bench: autotuned in {time.perf_counter() - t0:.1f} s: {tuning}")
    rng = np.random.default_rng(99)
    msgs = rng.integers(0, 16, args.batch)
    cts = keys.encrypt_msg(msgs, seed=4321)
    d_in = la.DeviceBuffer.from_numpy(cts)
    d_gid = la.DeviceBuffer.from_numpy(np.zeros(args.batch, np.uint32))
    d_out = la.DeviceBuffer((args.batch, prm.blen), np.uint64)
    stream = ctx.stream
    tm = Timer(hip, stream)

    log(f"bench: warmup {args.warmup}")
    for _ in range(args.warmup):
        ctx.pbs_batch_dev(d_in, d_out, d_gid, args.batch)
    ctx.sync()
    log(f"bench: timed {args.steps} steps")
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    tm.start()
    for _ in range(args.steps):
        ctx.pbs_batch_dev(d_in, d_out, d_gid, args.batch)
    ev_ms = tm.stop_ms()
    hip.hipDeviceSynchronize()
    wall = time.perf_counter() - t0
    ms_per_step = ev_ms / args.steps
    value = args.batch * args.steps / (ev_ms * 1e-3)
    gpu_out = d_out.numpy()
    dec = keys.decrypt_msg(gpu_out[:2048])
    decrypt_ok = float((dec == msgs[:2048]).mean())
    log(f"bench: {value:.1f} PBS/s ({ms_per_step:.2f} ms/step, wall {wall:.3f}s), decrypt rate {decrypt_ok:.4f}")

    # dominant kernel alone: blind rotation over the batch's key-switched LWEs
    d_lwe = la.DeviceBuffer((args.batch, prm.lwe_dim + 1), np.uint16)
    la.lib().luthe_keyswitch_batch(ctx.handle, d_in.ptr, d_lwe.ptr, args.batch, None)
    ctx.sync()
    tm.start()
    la.lib().luthe_keyswitch_batch(ctx.handle, d_in.ptr, d_lwe.ptr, args.batch, None)
    ks_ms = tm.stop_ms()
    br_ms = []
    for _ in range(2):
        tm.start()
        rc = la.lib().luthe_blind_rotate_batch(ctx.handle, d_lwe.ptr, d_out.ptr, d_gid.ptr, 0, args.batch, None)
        br_ms.append(tm.stop_ms())
        assert rc == 0
    br_ms = min(br_ms)
    for _ in range(2):  # the clock ramps during the first runs
        ctx.modmul_peak()
    peaks = sorted(ctx.modmul_peak() for _ in range(5))
    peak = peaks[2]  # median of five runs of the modmul microbenchmark
    work = modmul_equiv_per_pbs(prm) * args.batch
    achieved = work / (br_ms * 1e-3) / 1e9
    log(f"bench: blind rotation {br_ms:.2f} ms, key switch {ks_ms:.2f} ms, {achieved:.1f} Gmodmul/s of {peak:.1f}")

    # plain path: one ciphertext per launch
    plain_ms = []
    for r in range(args.plain_reps):
        tm.start()
        ctx.pbs_batch_dev(d_in, d_out, d_gid, 1)
        plain_ms.append(tm.stop_ms())
    plain = 1000.0 / float(np.median(plain_ms)) if plain_ms else None
    if plain:
        log(f"bench: plain path {np.median(plain_ms):.2f} ms per PBS -> {plain:.1f} PBS/s")
    traffic, valu_issue = None, None
    tpath = os.path.join(HERE, "profiles", "k_blind_rotate_8192_hbm_traffic.json")
    if os.path.exists(tpath) and args.batch == 8192:
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    cpath = os.path.join(HERE, "profiles", "k_blind_rotate_8192_sq_counters.json")
    if os.path.exists(cpath) and args.batch == 8192:
        with open(cpath) as f:
            sq = json.load(f)
        # wave64 VALU instructions per launch (profile) over the CU-clocks of this run's launch
        valu_issue = sq["sq_insts_valu_per_launch"] / (sq["cus"] * sq["clock_ghz"] * 1e9 * br_ms * 1e-3)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "PBS/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic: seeded keys, uniform 4-bit plaintexts, identity LUT",
        "config": {"workload": f"batched PBS+KS (key switch, blind rotation, sample extract) over {args.batch} "
                               f"ciphertexts, n=724 k=1 N=2048 PBS 2^23x1 KS 2^3x5 w_ks=21",
                   "batch": args.batch, "lwe_dim": prm.lwe_dim, "poly_size": prm.poly_size,
                   "pbs_base_log": prm.pbs_base_log, "pbs_level": prm.pbs_level,
                   "ks_base_log": prm.ks_base_log, "ks_level": prm.ks_level, "br_workgroups": ctx.br_grid(),
                   "tuning": tuning if tuning is not None else ctx.tuning()},
        "plain_pbs_per_s": round(plain, 2) if plain else None,
        "batched_over_plain": round(value / plain, 1) if plain else None,
        "decrypt_success": decrypt_ok,
        "kernels_ms": {"blind_rotate": round(br_ms, 3), "keyswitch": round(ks_ms, 3)},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": round(peak, 2),
                     "unit": "Gmodmul/s", "frac": round(achieved / peak, 4), "traffic": traffic,
                     "algorithmic_work": "modmul-equivalents per launch = batch x n x ((k+1)(L+1) x 11264 "
                                         "+ (k+1)^2 L x N + (k+1) N)",
                     "peak_runs": [round(x, 1) for x in peaks],
                     "traffic_source": "profile: rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel, calibrated "
                                       "(profiles/k_blind_rotate_8192_hbm_traffic.md), not measured in this run",
                     "valu_issue_frac": round(valu_issue, 4) if valu_issue else None,
                     "valu_issue_source": "profile SQ_INSTS_VALU per launch (profiles/k_blind_rotate_8192_sq_"
                                          "counters.md) / (256 CUs x 2.4 GHz x this run's kernel time)",
                     "bsk_consumed_gbps": round(args.batch * bsk_bytes(prm) / (br_ms * 1e-3) / 1e9, 1),
                     "traffic_gbps": round(traffic / (br_ms * 1e-3) / 1e9, 2) if traffic else None},
    }

    extras = set() if args.no_extra else {e.strip() for e in args.extras.split(",") if e.strip()}
    if "ntt" in extras:
        out["ntt_product_4096"] = run_ntt_product(la, hip, ctx, peak)
        log(f"bench: NTT product {out['ntt_product_4096']}")
    if "uneven" in extras:
        out["uneven_batches"] = run_uneven_batches(la, hip, peak)
        log(f"bench: uneven batches {out['uneven_batches']}")
    if "serving" in extras:
        out["serving"] = run_serving(la, ctx, keys, seconds=args.serve_seconds,
                                     cohort_max=args.serve_cohort_max)

    if not args.no_cpu_baseline:
        from oracle import ref  # CPU baseline leg only
        try:
            affinity = len(os.sched_getaffinity(0))
        except AttributeError:
            affinity = os.cpu_count() or 1
        # 16 host threads: the CPU share of one GPU on the benchmark host (whatever nproc shows)
        threads = args.cpu_threads or min(16, affinity)
        sample = 4 * threads
        log(f"bench: cpu baseline, oracle on {sample} ciphertexts with {threads} threads")
        rp = ref.params()
        bsk_ntt = ref.bsk_to_ntt(rp, keys.bsk.reshape(-1), threads=threads)
        t0 = time.perf_counter()
        exp = ref.pbs_batch(rp, bsk_ntt, keys.ksk, lut.reshape(1, -1), np.zeros(sample, np.uint32),
                            cts[:sample], threads=threads)
        cpu_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        exp1 = ref.pbs_batch(rp, bsk_ntt, keys.ksk, lut.reshape(1, -1), np.zeros(2, np.uint32), cts[:2], threads=1)
        cpu1_s = time.perf_counter() - t0
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
        out["cpu_baseline"] = {"value": round(sample / cpu_s, 3), "unit": "PBS/s", "cores": threads,
                               "kind": "port",
                               "sample": f"first {sample} ciphertexts of the batch, same keys and LUT "
                                         f"({cpu_s:.1f} s wall on {threads} threads)",
                               "single_thread_pbs_per_s": round(2 / cpu1_s, 3),
                               "single_thread_sample": f"first 2 ciphertexts ({cpu1_s:.1f} s)",
                               "host_nproc": os.cpu_count(), "host_affinity_cpus": affinity, "cpu_model": model,
                               "threads_note": "16 threads = the host CPU share of one GPU on the benchmark host"}
        assert np.array_equal(exp1, exp[:2])
        out["bit_exact_sample"] = bool(np.array_equal(exp, gpu_out[:sample]))
        log(f"bench: cpu {sample / cpu_s:.3f} PBS/s; bit-exact sample: {out['bit_exact_sample']}")
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
