#!/usr/bin/env python3
"""Throughput of the DSMC step on one MI355X.

Workload (default "c3"): argon VHS (m = 6.63e-26 kg, d_ref = 4.17e-10 m,
omega = 0.81, T_ref = 273 K), freestream n = 1e21 m^-3, T = 200 K,
U = 3000 m/s along +x, domain 0.4 x 0.2 x 0.2 m on a 256 x 128 x 128 grid,
a closed sphere-cone body (nose radius 0.02 m, 30 deg half-angle, base radius
0.05 m, ~50k triangles) with a diffuse 300 K wall, 2^25 simulated particles,
dt = 2e-7 s, FP32 particle state. A "step" is one pass of the full DSMC step
(counting sort, NTC collisions, move with walls and faces, freestream
injection) over all particles. Inputs are synthetic and generated on the GPU.

Prints one JSON line: particle-steps/s, the whole-step HBM fraction under
the algorithmic byte model (144 B per particle-step + 32 B per cell-step),
the roofline of the dominant kernel (the cell-tile kernel) timed with HIP
events over the timed steps, and the CPU oracle timed on a bounded sample on
this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
B_P, B_P_SAMPLE = 144.0, 156.0  # algorithmic bytes per particle-step
B_C, B_C_SAMPLE = 32.0, 120.0  # per cell-step
# cell-tile kernel: perm 4 + gathered pos/vel 24 + written pos/vel 24 + key 4 per
# particle; cell offset 4 + NTC state 8 read and 8 written per cell
TILE_B_P, TILE_B_C = 56.0, 20.0

METRIC = "particle-steps/s + %HBM peak on 1 MI355X; ensemble time-to-steady warm vs cold"

from cudadsmc_amd.cases import WORKLOADS, make_case  # noqa: E402


def make_solver_for(name: str, cg_refine: int | None = None, tile_cells: int | None = None,
                    scatter_grid: int | None = None):
    """A solver of workload `name` with the tuned launch layout (RNG streams,
    cells per tile, scatter grid), geometry set, no particles loaded."""
    import cudadsmc_amd as dm
    from cudadsmc_amd import ensemble as ens

    s, mesh, cgrid = make_case(name, cg_refine=cg_refine)
    s.n_streams = int(ens.tuned(name, 0))  # 0: the engine default
    sol = dm.Solver()
    sol.Init(s)
    tc = tile_cells or int(ens.tuned(name, 0, "tile_cells"))
    if tc:
        sol.set_tile_cells(tc)
    sg = scatter_grid or int(ens.tuned(name, 0, "scatter_grid"))
    if sg:
        sol.set_scatter_grid(sg)
    sol.SetModel(mesh, grid=cgrid)
    return sol, s, mesh


def timed(sol, steps: int, dt: float) -> dict:
    """`steps` steps with per-kernel HIP-event timing: wall time, particles
    (mean of before and after) and the kernel classes' average ms."""
    import ctypes as C

    import cudadsmc_amd as dm

    _, n_before, _ = sol.counts()
    sol.set_timing(True)
    sol.synchronize()
    t0 = time.perf_counter()
    sol.DoBenchmark(steps, dt)
    sol.synchronize()
    el = time.perf_counter() - t0
    tm = np.zeros(7, np.float32)
    dm.check(dm.lib().dsmc_get_timings(sol._h, tm.ctypes.data_as(C.POINTER(C.c_float)), 7), sol._h)
    sol.set_timing(False)
    _, n_after, _ = sol.counts()
    return {"elapsed": el, "n_avg": 0.5 * (n_before + n_after), "tm": tm}


def throughput_line(sol, s, steps: int, dt: float, sampling: bool = False) -> dict:
    r = timed(sol, steps, dt)
    cells = int(np.prod(s.grid_dim))
    bp, bc = (B_P_SAMPLE, B_C_SAMPLE) if sampling else (B_P, B_C)
    tile_ms = float(r["tm"][2])
    tile_bytes = TILE_B_P * r["n_avg"] + TILE_B_C * cells
    return {"steps": steps, "ms_per_step": 1e3 * r["elapsed"] / steps,
            "particle_steps_per_s": r["n_avg"] * steps / r["elapsed"], "particles_avg": r["n_avg"],
            "pct_hbm_peak": 100.0 * (bp * r["n_avg"] + bc * cells) * steps / r["elapsed"] / HBM_PEAK,
            "k_tile_ms": tile_ms, "k_tile_frac": tile_bytes / (tile_ms * 1e-3) / HBM_PEAK,
            "k_walls_ms": float(r["tm"][6])}


def other_workload(name: str, warmup: int, steps: int) -> dict:
    sol, s, mesh = make_solver_for(name)
    n0 = sol.InitSystemState(tuple(s.stream_velocity))
    sol.DoBenchmark(warmup, s.dt)
    sol.synchronize()
    out = throughput_line(sol, s, steps, s.dt)
    out.update({"particles_loaded": n0, "warmup": warmup, "grid": list(s.grid_dim),
                "triangles": int(len(mesh.triangles))})
    sol.close()
    return out


def reference_case():
    """The reference's own benchmark case (test.bat:1, `-f data/config_500k.xml`;
    kernels/dsmc.cu:398-405: ares.raw, CreateGrid(5,5,20), MoveModel(0,0,-3))
    in REFERENCE mode with the MT streams (`-mt`), dt = 0.25e-4 s: a solver
    with the model set, no particles loaded, and the moved mesh."""
    import cudadsmc_amd as dm

    g = ROOT / "tests" / "golden"  # the reference's data files, kept as test fixtures
    s = dm.settings_from_file(str(g / "config_500k.xml"))
    s.benchmark, s.geometry, s.mt = 1, 1, 1
    s.dt = 0.25e-4
    sol = dm.Solver()
    sol.Init(s)
    mesh = dm.load_raw(str(g / "ares.raw"))
    sol.SetModel(mesh, grid=(5, 5, 20), move=(0.0, 0.0, -3.0))
    return sol, s, mesh.translate((0.0, 0.0, -3.0))


def reference_benchmark(warmup: int = 200, steps: int = 1000) -> dict:
    sol, s, mesh = reference_case()
    n0 = sol.InitSystemState(tuple(s.stream_velocity))
    sol.DoBenchmark(warmup, s.dt)
    sol.synchronize()
    out = throughput_line(sol, s, steps, s.dt)
    out.update({"particles_loaded": n0, "warmup": warmup, "grid": list(s.grid_dim),
                "triangles": int(len(mesh.triangles)), "mode": "reference, MT streams, samples every 100th step",
                "case": "config_500k.xml + ares.raw (test.bat:1, dsmc.cu:398-405)"})
    sol.close()
    return out


def measured_traffic(workload: str, phase: str):
    """HBM bytes per k_tile launch of this workload in the early ("early":
    steps shortly after the freestream start) or developed ("developed":
    after 1000 steps) flow, as measured by rocprofv3 (FETCH_SIZE +
    WRITE_SIZE, scaled by the same counters on a plain copy of known size)
    and stored next to the profile summaries; None if absent."""
    p = ROOT / "profiles" / f"k_tile_{workload}_traffic.txt"
    try:
        for line in p.read_text().splitlines():
            if line.startswith(f"{phase}_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


def stored_warm_start():
    """The second half of the metric, warm- against cold-start time to
    steady on 32 held-out C5 cases, as measured by the committed evaluation
    run (training takes minutes, so it is not repeated here); None if
    absent."""
    p = ROOT / "profiles" / "warmstart_c5_eval.txt"
    out = {}
    try:
        for line in p.read_text().splitlines():
            if line.startswith("#") or ":" not in line:
                continue
            k, v = (x.strip() for x in line.split(":", 1))
            out[k] = None if v == "null" else float(v)
    except (OSError, ValueError):
        return None
    keys = ("steps_ratio_median", "steps_ratio_median_dense", "steps_ratio_median_rarefied", "median_steps_cold",
            "median_steps_warm", "cold_reached", "warm_reached", "median_pred_err", "wall_ratio")
    res = {k: out.get(k) for k in keys}
    res.update({"cases": out.get("cases"), "particles_per_case": out.get("N"), "target": ">= 3x",
                "source": "stored: profiles/warmstart_c5_eval.txt (python -m cudadsmc_amd.warmstart data/train/eval "
                          "on one MI355X; steady = accumulated block fields within 2% L2 of a long cold "
                          "reference and C_d within 3 sigma)"})
    return res


def traffic_phase(first_step: int) -> str:
    """Which stored traffic figure fits timed steps that start at `first_step`."""
    return "early" if first_step < 500 else "developed"


def _oracle_case(name: str):
    """The oracle mirrored from a solver of workload `name` created on the
    GPU (derived scalars, collision grid, RNG layout; no particles there)."""
    import oracle

    if name == "reference":
        sol, s, mesh = reference_case()
        o = oracle.OracleSim.from_solver(sol)
        o.set_geometry(mesh, sol.collision_grid(), None)
    else:
        sol, s, mesh = make_solver_for(name)
        o = oracle.OracleSim.from_solver(sol)
        o.set_geometry(mesh, sol.collision_grid(), sol.fluid_fraction())
        o.set_near_cells(sol.near_cells())
    o.set_rng(sol.rng_state())
    o.set_ntc(sol.ntc_state())
    sol.close()
    return o, s


def cpu_baseline(name: str = "c3", seconds: float = 15.0, max_steps: int = 50, threads: int = 0,
                 sampling_from: int | None = None) -> dict:
    """The CPU oracle on workload `name`: after loading, steps until `seconds`
    have passed or `max_steps` are done (steps from `sampling_from` on are
    sampled); `threads` > 0 sets the OpenMP thread count."""
    import oracle

    L = oracle.lib()
    t_all = L.oracle_threads()
    if threads:
        L.oracle_set_threads(threads)
    try:
        o, s = _oracle_case(name)
        o.load(tuple(s.stream_velocity))
        n0 = o.state()["pos"].shape[1]
        steps, psteps, t0 = 0, 0, time.perf_counter()
        while True:
            n = int(o.cell_start()[-1]) if steps else n0
            o.step(sample=sampling_from is not None and steps >= sampling_from)
            psteps += n
            steps += 1
            if time.perf_counter() - t0 > seconds or steps >= max_steps:
                break
        dt = time.perf_counter() - t0
        cores = int(L.oracle_threads())
    finally:
        L.oracle_set_threads(t_all)
    done = "to completion" if steps >= max_steps else "bounded sample"
    return {"value": psteps / dt, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"oracle (dsmc_cpu.cpp, OpenMP, {cores} threads) on the {name} workload ({n0} particles "
                      f"loaded), {steps} steps after loading ({dt:.1f} s, {done})",
            "cpu_model": _cpu_model(), "hardware_concurrency": os.cpu_count(),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def ensemble_bench(steps: int, warmup: int, instances: int = 64) -> dict:
    """The batched ensemble: `instances` argon sphere cases on 64x32x32 grids,
    particle counts log-uniform in [1e4, 5e5], U in [600, 3000] m/s, n in
    [3e19, 1e21] m^-3, r in [0.015, 0.035] m, seeds 0..instances-1, advanced
    by the batched step with HIP graph replay; RNG streams per instance from
    the autotuner's table, and cells per tile likewise."""
    import cudadsmc_amd as dm
    from cudadsmc_amd import ensemble as ens

    ns = int(ens.tuned("ensemble64", 16384))
    cases = ens.draw_cases(instances, seed=2024, temperatures=False)
    t0 = time.perf_counter()
    sols = [ens.make_solver(c, n_streams=ns) for c in cases]
    tcs = [s.layout()["tile_cells"] for s in sols]
    b = dm.Batch(sols)
    b.synchronize()
    setup = time.perf_counter() - t0
    b.run(warmup)
    b.synchronize()
    n0 = sum(s.counts()[1] for s in sols)
    t0 = time.perf_counter()
    b.run(steps)
    b.synchronize()
    el = time.perf_counter() - t0
    n1 = sum(s.counts()[1] for s in sols)
    navg = 0.5 * (n0 + n1)
    cells = instances * int(np.prod(ens.GRID))
    b.close()
    for s in sols:
        s.close()
    ps = navg * steps / el
    return {"instances": instances, "particles_total": navg, "cells_total": cells, "steps": steps,
            "warmup": warmup, "ms_per_step": 1e3 * el / steps, "particle_steps_per_s": ps,
            "pct_hbm_peak": 100.0 * (B_P * navg + B_C * cells) * steps / el / HBM_PEAK,
            "n_streams_per_instance": ns, "tile_cells": sorted({int(t) for t in tcs}), "setup_s": setup,
            "N_range": [min(c.N for c in cases), max(c.N for c in cases)]}


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized buffer 5 times (calibrates HBM byte counters)")
    ap.add_argument("--ensemble-steps", type=int, default=300)
    ap.add_argument("--no-developed", action="store_true",
                    help="skip the developed-flow and sampling-step measurements")
    ap.add_argument("--developed-warmup", type=int, default=1000,
                    help="steps after loading before the developed-flow measurement")
    ap.add_argument("--developed-steps", type=int, default=100)
    ap.add_argument("--no-other-workloads", action="store_true", help="skip the c1 and c2 throughput lines")
    ap.add_argument("--scatter-grid", type=int, default=None,
                    help="workgroups of the scatter kernel (default: tuned, else the engine's)")
    ap.add_argument("--cg-refine", type=int, default=None,
                    help="collision-grid cells per flow cell along each axis (default: tuned, else 1)")
    ap.add_argument("--tile-cells", type=int, default=None,
                    help="cells per tile of the tile kernel (default: tuned, else the engine's sizing)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU decomposition is "
              "design-only in this build", file=sys.stderr)
        return 2

    if args.copy_probe:
        import torch

        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(4):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 4
        print(json.dumps({"copy_probe_bytes": 4 * n, "copies": 5, "copy_ms": ms,
                          "copy_GBps": 2 * 4 * n / (ms * 1e-3) / 1e9}))
        return 0

    t_setup = time.perf_counter()
    sol, s, mesh = make_solver_for(args.workload, cg_refine=args.cg_refine, tile_cells=args.tile_cells,
                                   scatter_grid=args.scatter_grid)
    cgrid = [int(x) for x in sol.collision_grid()["dims"]]
    n0 = sol.InitSystemState(tuple(s.stream_velocity))
    sol.synchronize()
    t_setup = time.perf_counter() - t_setup
    print(f"[bench] {args.workload}: {n0} particles loaded, {len(mesh.triangles)} triangles, setup {t_setup:.1f} s",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    sol.DoBenchmark(args.warmup, s.dt)
    sol.synchronize()
    print(f"[bench] warmup {args.warmup} steps in {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    try:
        import torch

        torch.cuda.synchronize()
    except Exception:  # noqa: BLE001
        pass
    r = timed(sol, args.steps, s.dt)
    elapsed, n_avg, tm = r["elapsed"], r["n_avg"], r["tm"]
    psteps = n_avg * args.steps / elapsed
    cells = int(np.prod(s.grid_dim))
    step_bytes = B_P * n_avg + B_C * cells
    hbm_step = step_bytes * args.steps / elapsed
    tile_ms = float(tm[2])
    tile_bytes = TILE_B_P * n_avg + TILE_B_C * cells
    tile_gbs = tile_bytes / (tile_ms * 1e-3)
    bad = sol.validate()
    line = {
        "metric": METRIC,
        "value": psteps,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (particles generated on the GPU from the freestream Maxwellian)",
        "config": {
            "workload": f"{args.workload}: argon VHS, {n0} particles loaded, grid {'x'.join(map(str, s.grid_dim))}, "
                        f"{len(mesh.triangles)}-triangle body, U={s.stream_velocity[0]:.0f} m/s, n={s.density:.0e}, "
                        f"dt={s.dt:.1e}",
            "particles_avg": n_avg,
            "cells": cells,
            "collision_grid": list(cgrid),
            "tile_cells": int(sol.layout()["tile_cells"]),
            "parallelism": "single GPU",
        },
        "flow": f"timed steps {args.warmup + 1}-{args.warmup + args.steps} after a uniform freestream start "
                f"(the shock layer and wake form over the first ~1000 steps; see developed_flow)",
        "pct_hbm_peak": 100.0 * hbm_step / HBM_PEAK,
        "hbm_model": "144 B/particle-step + 32 B/cell-step (algorithmic, whole step)",
        "kernel_ms_per_step": {"scan": float(tm[0]), "scatter": float(tm[1]), "tile": float(tm[2]),
                               "walls": float(tm[6]), "inject": float(tm[3]), "total": float(tm[4]),
                               "timed_steps": int(tm[5])},
        "roofline": {"bound": "hbm", "achieved": tile_gbs / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": tile_gbs / HBM_PEAK,
                     "traffic": measured_traffic(args.workload, traffic_phase(args.warmup + 1)),
                     "traffic_source": f"stored: profiles/k_tile_{args.workload}_traffic.txt (rocprofv3 --pmc "
                                       f"FETCH_SIZE + WRITE_SIZE per k_tile launch, "
                                       f"{traffic_phase(args.warmup + 1)} flow, copy-calibrated)",
                     "kernel": "k_tile (56 B/particle + 20 B/cell algorithmic)"},
        "validate_bad_particles": int(bad),
        "warm_vs_cold_time_to_steady": stored_warm_start(),
    }
    if not args.no_developed:
        # the same solver continued into the developed flow, then timed again
        done = args.warmup + args.steps
        if done < args.developed_warmup:
            sol.DoBenchmark(args.developed_warmup - done, s.dt)
            sol.synchronize()
        dev = throughput_line(sol, s, args.developed_steps, s.dt)
        dev["after_steps"] = max(done, args.developed_warmup)
        dev["k_tile_traffic"] = measured_traffic(args.workload, "developed")
        line["developed_flow"] = dev
        sol.set_sampling(True)
        smp = throughput_line(sol, s, max(10, args.developed_steps // 2), s.dt, sampling=True)
        sol.set_sampling(False)
        smp["hbm_model"] = "156 B/particle-step + 120 B/cell-step (sampling steps, FP64 accumulators)"
        line["sampling_steps"] = smp
        print(f"[bench] developed flow {dev['ms_per_step']:.3f} ms/step, sampling {smp['ms_per_step']:.3f} ms/step",
              file=sys.stderr, flush=True)
    sol.close()
    if not args.no_ensemble:
        t0 = time.perf_counter()
        line["ensemble"] = ensemble_bench(args.ensemble_steps, 500)
        print(f"[bench] ensemble {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    if not args.no_other_workloads and args.workload == "c3":
        t0 = time.perf_counter()
        line["other_workloads"] = {"c1": other_workload("c1", 1000, 1000), "c2": other_workload("c2", 1000, 500),
                                   "reference_benchmark": reference_benchmark()}
        print(f"[bench] c1/c2 {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        line["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds)
        if args.workload == "c3":
            line["cpu_baseline_more"] = {
                # c1 to completion: 1000 transient + 1000 sampled steps
                "c1_full": cpu_baseline("c1", 600.0, 2000, sampling_from=1000),
                "c1_one_thread": cpu_baseline("c1", 10.0, 2000, threads=1),
                "c3_one_thread": cpu_baseline("c3", 10.0, 50, threads=1),
                # the reference's own benchmark case, REFERENCE mode
                "reference_benchmark": cpu_baseline("reference", 10.0, 1200),
            }
        print(f"[bench] cpu baseline {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
