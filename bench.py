"""Benchmark: spin-samples/s of the mixed-precision BlochHighOrder path on one MI355X.

Workload (default): 3D 128x128x64 = 1,048,576 spins, 32 Gaussian receive coils,
3rd-order fields (mask "1111", off-resonance on), 16 interleaves x (90-degree
hard RF + 8000-sample spiral readout at 2 us + idle to a 500 ms TR), synthetic
seeded inputs resident in HBM before timing starts.

One step = reset the spin state, then the whole block loop (RF blocks and
segment-fused precession with coil reception) producing the 128000 x 32
signal.  Prints one JSON line; progress goes to stderr.

Also measured: the plain GPU-FP64 drop-in (literal per-block algorithm, 32-node
blocks) on the same workload, signal NRMSE vs the f64 oracle on a spin
subsample, the oracle's CPU throughput on that subsample, and the serving
stream: Poisson arrivals (200 requests/s) of simulations with spin counts
log-uniform in 1e3..1e6, readouts uniform in 1000..20000 samples, 1/8/16/32
coils and masks 1100/1110/1111 over resident inputs (1,048,576-spin phantom,
32 coil maps, RF + 20000-sample spiral), served by the continuous-batching
server; its p50/p99 latency (arrival to signal on the host) goes in
``latency``.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
METRIC = "spin-samples/s (3rd-order, 32 coils, NRMSE≤1e-5 vs f64) + p50/p99 latency"

# MI355X peaks used by the ceiling model: 256 CUs x 4 SIMDs at the 2.4 GHz
# peak clock.
P_FP16_MFMA = 2.5e15      # dense FP16 MFMA flop/s (published)
P_FP32_VEC = 157.3e12     # FP32 vector flop/s (published: FMA = 2 flops, packed FP32 = 2 lanes)
P_FP32_LANE = 39.3e12     # FP32 VALU lane operations/s without packed FP32: 16 lanes per cycle per SIMD
P_TRANS = 19.7e12         # v_sin/v_cos/v_exp_f32 lane operations/s: 8 lanes per cycle per SIMD
                          # (measured 1.65-1.76e13/s with 4 waves per SIMD below the peak clock)


def load_package():
    if "hom_amd" in sys.modules:
        return sys.modules["hom_amd"]
    pkg = ROOT / "highordermrisim.jl_amd"
    spec = importlib.util.spec_from_file_location("hom_amd", pkg / "__init__.py",
                                                  submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hom_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = ROOT / "profiles" / "k_mix_main-traffic-3d-1M-32coil-16x8000.md"


def measured_traffic(spin_samples_per_launch: float, workload: str):
    """HBM bytes per launch of the fused kernel, scaled from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE measurement of this workload (bytes per
    spin-sample x spin-samples of this launch); None when the profile is
    absent or was taken on another workload."""
    try:
        text = TRAFFIC_PROFILE.read_text()
    except OSError:
        return None
    vals = dict(re.findall(r"^- (\w+): ([0-9.eE+-]+|[\w.-]+)$", text, flags=re.M))
    if vals.get("workload") != workload or "bytes_per_spin_sample" not in vals:
        return None
    return float(vals["bytes_per_spin_sample"]) * spin_samples_per_launch


COUNTERS_PROFILE = ROOT / "profiles" / "k_mix_main-counters-3d-1M-32coil-16x8000.md"


def measured_counters(spin_samples_per_launch: float, kernel_s: float, workload: str):
    """rocprofv3 instruction counts of the fused kernel (committed counters
    profile of this workload, per 16 x 16 tile) turned into rates with this
    run's kernel time: FP16 matrix work, transcendentals and VALU issue against
    MI355X's published peaks; None when the profile is absent or was taken on
    another workload."""
    try:
        text = COUNTERS_PROFILE.read_text()
    except OSError:
        return None
    vals = dict(re.findall(r"^- (\w+): ([0-9.eE+-]+|[\w.-]+)$", text, flags=re.M))
    if vals.get("workload") != workload or kernel_s <= 0:
        return None
    tiles = spin_samples_per_launch / 256.0
    f16 = float(vals["mfma_f16_flops_per_tile"]) * tiles / kernel_s
    trans = float(vals["trans_insts_per_tile"]) * 64 * tiles / kernel_s
    return {"source": str(COUNTERS_PROFILE.relative_to(ROOT)),
            "mfma_f16_tflops": f16 / 1e12, "mfma_f16_frac_of_peak": f16 / 2.5e15,
            "mfma_busy_frac_of_simd_cycles": float(vals["mfma_busy_frac"]),
            "transcendentals_per_s": trans, "transcendental_frac_of_peak": trans / P_TRANS,
            "valu_insts_per_tile": float(vals["valu_insts_per_tile"]),
            "simd_cycles_per_tile": float(vals["simd_cycles_per_tile"]),
            "shader_clock_ghz_under_load": float(vals["shader_clock_ghz"])}


def ceiling(nc, nt, nmh):
    """Lower bound on the fused kernel's time per spin-sample, with the matrix
    pipe and the VALU perfectly overlapped: t_min = max(t_mfma, t_valu).
    t_mfma: the FP16x3 complex coil GEMM (3 x 8 Nc flops) and the split-FP16
    phase-offset contraction (nmh 16x16x32 MFMAs per 256 spin-samples = 64 nmh
    flops) at the FP16 MFMA peak.  t_valu: 3 transcendentals (sin, cos, exp2)
    and 8 FP32 lane operations (phase add, decay argument, two magnitude
    products, FP16 hi/lo split of the two state components).  Returns
    (t_min, algorithmic flops F = 8 Nc + 2 Nt + 14)."""
    t_mfma = (3 * 8 * nc + 64 * nmh) / P_FP16_MFMA
    t_valu = 3 / P_TRANS + 8 / P_FP32_LANE
    return max(t_mfma, t_valu), 8 * nc + 2 * nt + 14


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--subsample", type=int, default=1, help="keep every k-th spin")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-subsample", type=int, default=40)
    ap.add_argument("--no-cpu-reference", action="store_true",
                    help="skip the reference-algorithm CPU timing (broadcast + zgemm restatement)")
    ap.add_argument("--cpu-reference-subsample", type=int, default=8)
    ap.add_argument("--block-length", type=int, default=32)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-budget", type=int, default=0,
                    help="spin-samples per request per step (0: server default)")
    ap.add_argument("--serve-no-graphs", action="store_true")
    ap.add_argument("--serve-autotune", action="store_true",
                    help="re-probe the serving kernel's shape instead of the persisted tuning")
    ap.add_argument("--serve-no-autotune", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU spin sharding is a design, not a measured path)"}))
        sys.exit(2)

    load_package()
    from hom_amd.runtime import Plan, device_count
    from hom_amd.types import GAMMA, BlochHighOrder, get_sim_ranges
    from hom_amd.workloads import make_config

    if device_count() < 1:
        log("no HIP device visible")
        sys.exit(3)
    if args.serve_only:
        print(json.dumps({"serving": serve_benchmark(args)}), flush=True)
        return
    t0 = time.perf_counter()
    obj, seq, cfg = make_config(args.config, spin_subsample=args.subsample if args.subsample > 1 else None)
    method = BlochHighOrder("1111")
    L = args.block_length
    parts = get_sim_ranges(seq, L, L)
    total_adc = int(np.count_nonzero(seq.adc))
    ns, nc = len(obj), obj.n_coils
    spin_samples = ns * total_adc
    max_adc = max(seq.block_adc(a, b) for a, b in zip(parts[0], parts[1]))
    max_rf_adc = max((seq.block_adc(a, b) for a, b, r in zip(*parts) if r), default=0)
    log(f"workload {cfg.name}: spins={ns} coils={nc} nodes={seq.n_nodes} adc={total_adc} "
        f"parts={len(parts[0])} built in {time.perf_counter() - t0:.1f}s")

    # ---------------- mixed path (the measured product) -----------------
    plan = Plan("mixed", ns, nc, seq.n_nodes, max_adc)
    plan.upload_phantom(obj)
    plan.upload_sequence(seq)
    for _ in range(max(args.warmup, 0)):
        plan.reset_state()
        plan.run(parts, method)
        plan.synchronize()
    log("warmup done")
    plan.set_timing(True)
    step_ms = []
    plan.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        plan.reset_state()
        plan.run(parts, method)
        plan.synchronize()
        step_ms.append((time.perf_counter() - a) * 1e3)
    plan.synchronize()
    elapsed = time.perf_counter() - t_start
    seg_ms, seg_n = plan.kernel_time("mixed_segment")
    fin_ms, fin_n = plan.kernel_time("mixed_finalize")
    prep_ms, prep_n = plan.kernel_time("mixed_prep")
    plan.set_timing(False)
    from hom_amd import tuning
    ncb, nmh = tuning.shape_class(nc, method.flags)
    tuned_plan = {"source": "highordermrisim.jl_amd/tuning.json" if tuning.load() else None,
                  "shape_class": {"ncb": ncb, "nmh": nmh}, "spin_tile_bound": plan.spin_tile,
                  "spin_tile": plan.last_spin_tile}
    log(f"tuning: shape class ncb={ncb} nmh={nmh}, spin tile {plan.last_spin_tile} groups "
        f"(bound {plan.spin_tile}) "
        f"({tuned_plan['source'] or 'library default'})")
    value = spin_samples * args.steps / elapsed
    log(f"mixed: {value:.4e} spin-samples/s, {elapsed / args.steps * 1e3:.2f} ms/step; "
        f"main kernel {seg_ms / max(seg_n, 1):.3f} ms avg over {seg_n} launches")

    # dominant kernel: per launch spin-samples = ns * (samples of one segment)
    nt = method.n_terms
    t_min, f_alg = ceiling(nc, nt, nmh)
    ss_per_launch = spin_samples / max(seg_n / max(args.steps, 1), 1)
    kern_rate = ss_per_launch / (seg_ms / max(seg_n, 1) / 1e3) if seg_n else 0.0
    # achieved / peak against the published dense FP16 MFMA peak (the pipe of
    # the coil GEMM); the path's own ceiling (transcendentals, FP64 phase,
    # VALU work) is reported beside it.
    roofline = {
        "bound": "mfma",
        "achieved": kern_rate * f_alg / 1e12,
        "peak": P_FP16_MFMA / 1e12,
        "unit": "TFLOP/s",
        "frac": kern_rate * f_alg / P_FP16_MFMA,
        "traffic": measured_traffic(ss_per_launch, cfg.name),
        "kernel": "k_mix_main",
        "kernel_ms_avg": seg_ms / max(seg_n, 1),
        "kernel_spin_samples_per_s": kern_rate,
        "flops_per_spin_sample": f_alg,
        "fp32_equiv_frac": kern_rate * f_alg / P_FP32_VEC,
        "ceiling_model": {
            "spin_samples_per_s": 1.0 / t_min,
            "frac": kern_rate * t_min,
            "formula": "t_min=max((24*Nc+64*nmh)/P_fp16mfma, 3/P_trans+8/P_fp32lane) with "
                       "P_trans=19.7e12/s, P_fp32lane=39.3e12/s; F_alg=8*Nc+2*Nt+14 per spin-sample",
        },
        "counters": measured_counters(ss_per_launch, seg_ms / max(seg_n, 1) / 1e3, cfg.name),
    }
    if roofline["counters"]:
        c = roofline["counters"]
        log(f"counters: FP16 matrix work {c['mfma_f16_tflops']:.0f} TFLOP/s "
            f"({c['mfma_f16_frac_of_peak']:.2f} of peak), transcendentals "
            f"{c['transcendentals_per_s']:.2e}/s ({c['transcendental_frac_of_peak']:.2f} of peak)")
    extra = {"finalize_ms_avg": fin_ms / max(fin_n, 1), "prep_ms_avg": prep_ms / max(prep_n, 1),
             "p50_step_ms": float(np.percentile(step_ms, 50)),
             "p99_step_ms": float(np.percentile(step_ms, 99))}
    plan.close()

    # ---------------- end to end: host inputs in, host signal out -----------------
    # A fresh plan per call: phantom/CSM and sequence uploads (H2D), the host
    # integration of the sequence tables, the run and the signal download (D2H).
    with Plan("mixed", ns, nc, seq.n_nodes, max_adc) as pe:
        pe.synchronize()
        a = time.perf_counter()
        pe.upload_phantom(obj)
        pe.synchronize()
        a1 = time.perf_counter()
        pe.upload_sequence(seq)
        pe.synchronize()
        a2 = time.perf_counter()
        pe.run(parts, method)
        pe.synchronize()
        b = time.perf_counter()
        sig_host = pe.download_signal()
        c = time.perf_counter()
    end_to_end = {"value": spin_samples / (c - a), "unit": "spin-samples/s", "seconds": c - a,
                  "phantom_upload_ms": (a1 - a) * 1e3, "sequence_upload_ms": (a2 - a1) * 1e3,
                  "first_run_ms": (b - a2) * 1e3, "download_ms": (c - b) * 1e3,
                  "signal_bytes": int(sig_host.nbytes),
                  "includes": "uploads (H2D), host table integration, run, signal download (D2H)"}
    del sig_host
    log(f"end to end: {end_to_end['value']:.4e} spin-samples/s ({c - a:.3f}s: phantom "
        f"{end_to_end['phantom_upload_ms']:.1f} ms, sequence {end_to_end['sequence_upload_ms']:.1f} ms, "
        f"first run {end_to_end['first_run_ms']:.1f} ms, download {end_to_end['download_ms']:.1f} ms)")

    # ---------------- plain GPU-FP64 drop-in on the same workload -----------------
    plain = None
    if not args.no_plain:
        def time_plain(precision, mode):
            """One full run of the per-block drop-in after a one-interleave warmup."""
            with Plan(precision, ns, nc, seq.n_nodes, max_adc, max_rf_adc=max_rf_adc,
                      coil_combination=mode) as pp:
                pp.upload_phantom(obj)
                pp.upload_sequence(seq)
                pp.reset_state()
                pp.run([arr[:2] for arr in parts], method)
                pp.reset_state()
                pp.synchronize()
                a = time.perf_counter()
                pp.run(parts, method)
                pp.synchronize()
                dt = time.perf_counter() - a
            log(f"plain {precision} {mode}: {spin_samples / dt:.4e} spin-samples/s ({dt:.2f}s)")
            return dt

        dt = time_plain("f64", "gemm")
        plain = {"value": spin_samples / dt, "unit": "spin-samples/s", "block_length": L,
                 "coil_combination": "gemm", "seconds": dt,
                 "f32_value": spin_samples / time_plain("f32", "gemm"),
                 "f64_fused_value": spin_samples / time_plain("f64", "fused_tiled"),
                 "f32_fused_value": spin_samples / time_plain("f32", "fused_tiled")}

    # ---------------- accuracy + CPU baseline on a spin subsample -----------------
    accuracy, cpu = None, None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT))
        import oracle
        sub = obj[::args.cpu_subsample]
        threads = max(1, min(16, os.cpu_count() or 1))
        log(f"oracle on {len(sub)} spins x {total_adc} samples, {threads} threads")
        a = time.perf_counter()
        ref, _, _ = oracle.run(sub, seq, parts, method.flags, GAMMA, n_threads=threads)
        cdt = time.perf_counter() - a
        cpu = {"value": len(sub) * total_adc / cdt, "unit": "spin-samples/s", "cores": threads,
               "kind": "port",
               "algorithm": "the f64 oracle: a per-spin fused restatement of the reference CPU "
                            "path (one pass per spin over the nodes, coil sum folded in); the "
                            "reference's own broadcast + zgemm algorithm is timed beside it in "
                            "reference_algorithm",
               "sample": f"every {args.cpu_subsample}th spin ({len(sub)} spins) x all {total_adc} "
                         f"samples, {nc} coils, f64, {threads} threads, {cdt:.1f}s"}
        log(f"cpu oracle: {cpu['value']:.4e} spin-samples/s")
        # one thread on about 1.5e8 spin-samples (a few seconds)
        stride1 = max(1, -(-ns * total_adc // 150_000_000))
        sub1 = obj[::stride1]
        log(f"oracle on {len(sub1)} spins x {total_adc} samples, 1 thread")
        a = time.perf_counter()
        oracle.run(sub1, seq, parts, method.flags, GAMMA, n_threads=1)
        c1 = time.perf_counter() - a
        cpu["single_thread"] = {"value": len(sub1) * total_adc / c1, "unit": "spin-samples/s",
                                "sample": f"every {stride1}th spin ({len(sub1)} "
                                          f"spins) x all {total_adc} samples, 1 thread, {c1:.1f}s"}
        log(f"cpu oracle, 1 thread: {cpu['single_thread']['value']:.4e} spin-samples/s")
        if not args.no_cpu_reference:
            cpu["reference_algorithm"] = reference_algorithm_leg(
                obj, seq, parts, method, threads, args.cpu_reference_subsample, cfg.n_adc)
        with Plan("mixed", len(sub), nc, seq.n_nodes, max_adc) as ps:
            ps.upload_phantom(sub)
            ps.upload_sequence(seq)
            # the timed run's spin tile, so each FP32 partial sum covers as
            # many spins as there (the subsample alone would get a smaller one)
            ps.spin_tile = tuned_plan["spin_tile"]
            ps.run(parts, method)
            sig = ps.download_signal()
        n = cfg.n_adc
        per = [float(np.linalg.norm(sig[i * n:(i + 1) * n] - ref[i * n:(i + 1) * n]) /
                     np.linalg.norm(ref[i * n:(i + 1) * n])) for i in range(cfg.n_interleaves)]
        accuracy = {"nrmse_vs_f64_oracle": float(np.linalg.norm(sig - ref) / np.linalg.norm(ref)),
                    "nrmse_max_per_interleave": max(per), "spins_checked": len(sub),
                    "spin_tile": tuned_plan["spin_tile"]}
        # the plain drop-ins on the same spins
        for prec in ("f64", "f32"):
            with Plan(prec, len(sub), nc, seq.n_nodes, max_adc) as pd:
                pd.upload_phantom(sub)
                pd.upload_sequence(seq)
                pd.run(parts, method)
                sp = pd.download_signal()
            accuracy[f"plain_{prec}_nrmse_vs_f64_oracle"] = float(np.linalg.norm(sp - ref) / np.linalg.norm(ref))
        log(f"plain drop-in NRMSE vs f64 oracle: f64 {accuracy['plain_f64_nrmse_vs_f64_oracle']:.3e}, "
            f"f32 {accuracy['plain_f32_nrmse_vs_f64_oracle']:.3e}")
        log(f"NRMSE vs f64 oracle: {accuracy['nrmse_vs_f64_oracle']:.3e} "
            f"(max per interleave {max(per):.3e})")

    serving = None if args.no_serve else serve_benchmark(args)
    latency = {"source": "serving stream", **{k: serving[k] for k in ("p50_ms", "p99_ms")}} \
        if serving else {}
    latency.update({"plan_step_p50_ms": extra["p50_step_ms"],
                    "plan_step_p99_ms": extra["p99_step_ms"]})

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "spin-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed(f64 phase anchors, split-fp16 offsets and fp16x3 coil mfma, f32 tile acc, "
                 "f64 signal)",
        "data": "synthetic",
        "config": {"workload": cfg.name, "spins": ns, "coils": nc,
                   "interleaves": cfg.n_interleaves, "samples_per_interleave": cfg.n_adc,
                   "mask": method.name, "precision": "mixed", "block_length": L},
        "roofline": roofline,
        "tuning": tuned_plan,
        "end_to_end": end_to_end,
        "cpu_baseline": cpu,
        "plain_f64_dropin": plain,
        "speedup_vs_plain_f64": (value / plain["value"]) if plain else None,
        "speedup_vs_fastest_plain_f64": (value / max(plain["value"], plain["f64_fused_value"]))
        if plain else None,
        "accuracy": accuracy,
        "latency": latency,
        "serving": serving,
        "kernels": {"finalize_ms_avg": extra["finalize_ms_avg"], "prep_ms_avg": extra["prep_ms_avg"]},
    }
    print(json.dumps(out), flush=True)


def reference_algorithm_leg(obj, seq, parts, method, threads, stride, per_readout,
                            interleaves=2) -> dict:
    """Times the reference's own CPU algorithm (oracle/broadcast.py: per-thread
    contiguous spin parts, per-block broadcast temporaries of the fifteen
    field terms, cumtrapz, complex exp and a zgemm coil combination) on every
    ``stride``-th spin over the blocks of the first ``interleaves`` readouts,
    an exact prefix of the run."""
    from oracle import broadcast
    from hom_amd.types import GAMMA
    pf, pl, rf = (np.asarray(p) for p in parts)
    adc_after = np.cumsum([int(np.count_nonzero(seq.adc[f + 1:l + 1])) for f, l in zip(pf, pl)])
    nb = int(np.searchsorted(adc_after, per_readout * interleaves) + 1)
    prefix = (pf[:nb], pl[:nb], rf[:nb])
    samples = int(adc_after[nb - 1])
    sub = obj[::stride]
    log(f"reference CPU algorithm on {len(sub)} spins x {samples} samples, {threads} threads")
    a = time.perf_counter()
    broadcast.run(sub, seq, prefix, method.flags, GAMMA, n_threads=threads)
    dt = time.perf_counter() - a
    out = {"value": len(sub) * samples / dt, "unit": "spin-samples/s", "cores": threads,
           "kind": "port",
           "algorithm": "the reference CPU algorithm restated in numpy: contiguous spin parts per "
                        "thread, per 32-node block the fifteen field terms as spins x nodes "
                        "broadcast arrays, cumtrapz, complex exp, then one zgemm (OpenBLAS, one "
                        "thread per worker) for the coil combination; numpy evaluates each "
                        "broadcast expression through temporaries where Julia fuses it",
           "sample": f"every {stride}th spin ({len(sub)} spins) x the first {samples} samples "
                     f"({interleaves} of the run's readouts), f64, {threads} threads, {dt:.1f}s"}
    log(f"reference CPU algorithm: {out['value']:.4e} spin-samples/s")
    return out


def serve_benchmark(args) -> dict:
    """Serving stream over resident inputs; returns latency and throughput."""
    from hom_amd import tuning
    from hom_amd.serving import SimulationServer, request_stream, run_stream
    from hom_amd.workloads import CONFIG_SEEDS, serving_inputs

    t0 = time.perf_counter()
    sobj, sseq = serving_inputs()
    log(f"serving inputs: spins={len(sobj)} coils={sobj.n_coils} nodes={sseq.n_nodes} "
        f"built in {time.perf_counter() - t0:.1f}s")
    with SimulationServer(memory_bytes=64 << 30, result_bytes=4 << 30, threaded=True,
                          step_budget=args.serve_budget or None,
                          use_graphs=not args.serve_no_graphs) as srv:
        ph, sq = srv.put_phantom(sobj), srv.put_sequence(sseq)
        tuned = None
        if args.serve_autotune:
            cus = srv.config()["main_grid"] // 3
            gpt, grid, secs = srv.autotune(ph, sq, spins=(0, 1 << 18), coils=32, n_samples=20000,
                                           gpts=[128, 256, 512], grids=[2 * cus, 3 * cus, 4 * cus])
            tuned = {"gpt": gpt, "main_grid": grid, "probe_kernel_ms": secs * 1e3,
                     "probe": "262144 spins x 20000 samples x 32 coils, mask 1111",
                     "candidates": "gpt {128,256,512} x grid {2,3,4} workgroups per CU"}
            log(f"autotune: gpt {gpt}, grid {grid} ({secs * 1e3:.2f} ms on the probe)")
        warm = request_stream(CONFIG_SEEDS[3] + 1, 40, rate=args.serve_rate,
                              n_spins_max=len(sobj))
        run_stream(srv, ph, sq, warm)
        srv.drain()
        before = srv.stats()
        stream = request_stream(CONFIG_SEEDS[3], args.serve_requests, rate=args.serve_rate,
                                n_spins_max=len(sobj))
        log(f"serving {len(stream)} requests at {args.serve_rate:g}/s "
            f"(offered {sum(r.spin_count * r.n_samples for r in stream) / stream[-1].arrival:.3e} "
            "spin-samples/s)")
        res = run_stream(srv, ph, sq, stream)
        srv.drain()
        after = srv.stats()
        cfg = srv.config()
    res.pop("requests_list")
    res.update({
        "rate_req_per_s": args.serve_rate,
        "steps": after["steps"] - before["steps"],
        "graph_launches": after["graph_launches"] - before["graph_launches"],
        "graphs_captured": after["graphs_captured"],
        "max_active": after["max_active_seen"],
        "device_busy_s": after["busy_seconds"] - before["busy_seconds"],
        "step_budget": cfg["step_budget"], "gpt": cfg["gpt"], "main_grid": cfg["main_grid"],
        "use_graphs": bool(cfg["use_graphs"]),
        "autotune": tuned,
        "shape_source": "probed at start-up" if tuned else
                        ("highordermrisim.jl_amd/tuning.json" if tuning.server_shape() else "library default"),
        "inputs": f"resident {len(sobj)} spins x {sobj.n_coils} coils, RF + {sseq.adc.sum()}-sample "
                  "spiral; requests: spins log-U[1e3,1e6], samples U[1000,20000], "
                  "coils {1,8,16,32}, masks {1100,1110,1111}",
    })
    log(f"serving: p50 {res['p50_ms']:.1f} ms, p99 {res['p99_ms']:.1f} ms, "
        f"throughput {res['throughput']:.3e} spin-samples/s over {res['wall_s']:.2f}s, "
        f"{res['steps']} steps, max active {res['max_active']}")
    return res


if __name__ == "__main__":
    main()
