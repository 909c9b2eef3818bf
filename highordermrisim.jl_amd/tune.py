"""Autotuner of the mixed-precision kernels' launch shapes; writes tuning.json.

    python highordermrisim.jl_amd/tune.py [--out PATH] [--spins N] [--samples S]

For each shape class of the plan path (coil blocks 1..4 of a 32-coil pass x
offset MFMAs 1..2, probed at 8/16/24/32 coils and field orders "1100" /
"1111") it times ``k_mix_main`` plus ``k_mix_finalize`` (a smaller spin tile
means more spin-tile partials to sum) over one RF + readout run of a probe
phantom for every candidate spin tile (HIP events on the plan's stream,
median of three runs after one warm-up) and keeps the fastest; for the
server it times the persistent kernel over (spin tile, grid) pairs through
``hom_server_autotune``.  Run it on the GPU the library will serve; the file
it writes is what :mod:`tuning` loads.
"""
from __future__ import annotations

import argparse
import datetime
import importlib.util
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def _package():
    if "hom_amd" in sys.modules:
        return sys.modules["hom_amd"]
    spec = importlib.util.spec_from_file_location("hom_amd", HERE / "__init__.py",
                                                  submodule_search_locations=[str(HERE)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hom_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def time_plan(obj, seq, parts, method, tile, runs=3):
    from hom_amd.runtime import Plan
    nadc = max(seq.block_adc(a, b) for a, b in zip(parts[0], parts[1]))
    with Plan("mixed", len(obj), obj.n_coils, seq.n_nodes, nadc, max_rf_adc=0) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        plan.spin_tile = tile
        plan.run(parts, method)          # warm-up: tables, operands
        plan.synchronize()
        times = []
        for _ in range(runs):
            plan.set_timing(True)
            plan.reset_state()
            plan.run(parts, method)
            plan.synchronize()
            ms, n = plan.kernel_time("mixed_segment")
            fms, fn = plan.kernel_time("mixed_finalize")
            plan.set_timing(False)
            times.append(ms / max(n, 1) + fms / max(fn, 1))
        return sorted(times)[len(times) // 2]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(HERE / "tuning.json"))
    ap.add_argument("--spins", type=int, default=1 << 20)
    ap.add_argument("--samples", type=int, default=8000)
    ap.add_argument("--tiles", default="64,128,256,512")
    ap.add_argument("--no-server", action="store_true")
    args = ap.parse_args(argv)
    _package()
    from hom_amd import tuning, workloads
    from hom_amd.runtime import load_library
    from hom_amd.serving import SimulationServer
    from hom_amd.types import BlochHighOrder, get_sim_ranges
    load_library()
    tiles = [int(t) for t in args.tiles.split(",")]
    plan_entries = []
    for nc in (8, 16, 24, 32):
        obj, seq = workloads.serving_inputs(n_spins=args.spins, n_coils=nc, n_samples=args.samples)
        parts = get_sim_ranges(seq, 10 ** 9, 10 ** 9)
        for order in ("1100", "1111"):
            method = BlochHighOrder(order)
            ncb, nmh = tuning.shape_class(nc, method.flags)
            times = {str(t): round(time_plan(obj, seq, parts, method, t), 4) for t in tiles}
            best = min(times, key=lambda k: times[k])
            plan_entries.append({"ncb": ncb, "nmh": nmh, "spin_tile": int(best), "kernel_ms": times,
                                 "probe": f"{args.spins} spins x {args.samples} samples x {nc} coils, "
                                          f"mask {order}"})
            print(f"plan ncb={ncb} nmh={nmh}: {times} -> {best}", flush=True)
    doc = {"schema": tuning.SCHEMA,
           "created": datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"),
           "plan": plan_entries}
    if not args.no_server:
        obj, seq = workloads.serving_inputs(n_spins=1 << 20, n_coils=32, n_samples=20_000)
        with SimulationServer(memory_bytes=16 << 30, threaded=False, gpt=256, main_grid=768) as srv:
            ph, sq = srv.put_phantom(obj), srv.put_sequence(seq)
            cus = srv.config()["main_grid"] // 3
            gpt, grid, secs = srv.autotune(ph, sq, spins=(0, 1 << 18), coils=32, n_samples=20_000,
                                           gpts=[128, 256, 512], grids=[2 * cus, 3 * cus, 4 * cus])
        doc["server"] = {"spin_tile": int(gpt), "main_grid": int(grid),
                         "kernel_ms": round(secs * 1e3, 4),
                         "probe": "262144 spins x 20000 samples x 32 coils, mask 1111"}
        print(f"server: spin tile {gpt}, grid {grid} ({secs * 1e3:.3f} ms)", flush=True)
    tuning.validate(doc)
    Path(args.out).write_text(json.dumps(doc, indent=1) + "\n")
    print(f"wrote {args.out}")


if __name__ == "__main__":
    main()
