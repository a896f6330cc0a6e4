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

WORKLOADS = {
    # name: (grid, target particles, n, U, dt, body)
    "c3": dict(grid=(256, 128, 128), n_part=2**25, density=1.0e21, U=3000.0, dt=2.0e-7, body="spherecone"),
    "c2": dict(grid=(128, 64, 64), n_part=2**20, density=1.0e20, U=1500.0, dt=5.0e-7, body="ico5"),
    "c1": dict(grid=(64, 32, 32), n_part=100000, density=1.0e20, U=1500.0, dt=1.0e-6, body="ico3"),
}


def make_case(name: str, n_part: int | None = None, cg_refine: int | None = None):
    """Settings, body mesh and collision-grid dimensions of a workload. The
    collision grid covers the body's bounding box with cells `cg_refine`
    times finer than the flow cells (default: the tuned value, else 1)."""
    import cudadsmc_amd as dm
    from cudadsmc_amd import ensemble as ens

    w = WORKLOADS[name]
    s = dm.settings_default()
    s.mode = dm.MODE_PHYSICS
    s.bmin[:] = (0.0, 0.0, 0.0)
    s.bmax[:] = (0.4, 0.2, 0.2)
    s.grid_dim[:] = w["grid"]
    for a in range(3):
        s.cell_size[a] = (s.bmax[a] - s.bmin[a]) / s.grid_dim[a]
    s.molmass, s.diameter, s.omega, s.Tref = 6.63e-26, 4.17e-10, 0.81, 273.0
    s.temp, s.density, s.wall_temp = 200.0, w["density"], 300.0
    s.stream_velocity[:] = (w["U"], 0.0, 0.0)
    s.face_bc[:] = [dm.FACE_FREESTREAM] * 6
    s.target_particles = n_part or w["n_part"]
    s.dt = w["dt"]
    if w["body"] == "spherecone":
        mesh = dm.sphere_cone(0.1, 0.1, 0.1, 0.02, 30.0, 0.05, 200, 124)
    elif w["body"] == "ico5":
        mesh = dm.icosphere((0.1, 0.1, 0.1), 0.025, 5)
    else:
        mesh = dm.icosphere((0.1, 0.1, 0.1), 0.025, 3)
    if cg_refine is None:
        cg_refine = int(ens.tuned(name, 1, "cg_refine"))
    lo, hi = mesh.vertices.min(0), mesh.vertices.max(0)
    cgrid = tuple(int(max(1, np.ceil(cg_refine * (hi[k] - lo[k]) / s.cell_size[k]))) for k in range(3))
    return s, mesh, cgrid


def measured_traffic(workload: str):
    """HBM bytes per k_tile launch of this workload as measured by rocprofv3
    (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a plain copy of
    known size) and stored next to the profile summaries; None if absent."""
    p = ROOT / "profiles" / f"k_tile_{workload}_traffic.txt"
    try:
        for line in p.read_text().splitlines():
            if line.startswith("bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


def cpu_baseline(seconds: float = 15.0) -> dict:
    """The CPU oracle on a bounded sample of the same workload: the full c3 case
    (2^25 particles, same grid, body and flow), as many steps as fit in
    `seconds` after loading (at most 50)."""
    import oracle

    import cudadsmc_amd as dm

    s, mesh, cgrid = make_case("c3")
    # the oracle needs the engine's derived scalars and RNG layout: mirror a
    # solver created on the GPU (no particles loaded there)
    sol = dm.Solver()
    sol.Init(s)
    sol.SetModel(mesh, grid=cgrid)
    o = oracle.OracleSim.from_solver(sol)
    o.set_geometry(mesh, sol.collision_grid(), sol.fluid_fraction())
    o.set_near_cells(sol.near_cells())
    o.set_rng(sol.rng_state())
    o.set_ntc(sol.ntc_state())
    sol.close()
    o.load(tuple(s.stream_velocity))
    n0 = o.state()["pos"].shape[1]
    steps, t0 = 0, time.perf_counter()
    psteps = 0
    while True:
        n = int(o.cell_start()[-1]) if steps else n0
        o.step(sample=False)
        psteps += n
        steps += 1
        if time.perf_counter() - t0 > seconds or steps >= 50:
            break
    dt = time.perf_counter() - t0
    return {"value": psteps / dt, "unit": "particle-steps/s", "cores": int(oracle.lib().oracle_threads()),
            "kind": "port",
            "sample": f"oracle (dsmc_cpu.cpp, OpenMP) on the c3 workload ({n0} particles loaded), "
                      f"{steps} steps after loading ({dt:.1f} s)",
            "cpu_model": _cpu_model()}


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

    import cudadsmc_amd as dm
    from cudadsmc_amd import ensemble as ens

    s, mesh, cgrid = make_case(args.workload, cg_refine=args.cg_refine)
    s.n_streams = int(ens.tuned(args.workload, 0))  # 0: the engine default
    t_setup = time.perf_counter()
    sol = dm.Solver()
    sol.Init(s)
    tc = args.tile_cells or int(ens.tuned(args.workload, 0, "tile_cells"))
    if tc:
        sol.set_tile_cells(tc)
    sg = args.scatter_grid or int(ens.tuned(args.workload, 0, "scatter_grid"))
    if sg:
        sol.set_scatter_grid(sg)
    sol.SetModel(mesh, grid=cgrid)
    n0 = sol.InitSystemState(tuple(s.stream_velocity))
    sol.synchronize()
    t_setup = time.perf_counter() - t_setup
    print(f"[bench] {args.workload}: {n0} particles loaded, {len(mesh.triangles)} triangles, setup {t_setup:.1f} s",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    sol.DoBenchmark(args.warmup, s.dt)
    sol.synchronize()
    print(f"[bench] warmup {args.warmup} steps in {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)

    _, n_before, _ = sol.counts()
    sol.set_timing(True)
    try:
        import torch

        torch.cuda.synchronize()
    except Exception:  # noqa: BLE001
        pass
    t0 = time.perf_counter()
    sol.DoBenchmark(args.steps, s.dt)
    sol.synchronize()
    elapsed = time.perf_counter() - t0
    tm = np.zeros(7, np.float32)
    import ctypes as C

    dm.check(dm.lib().dsmc_get_timings(sol._h, tm.ctypes.data_as(C.POINTER(C.c_float)), 7), sol._h)
    sol.set_timing(False)
    _, n_after, _ = sol.counts()
    n_avg = 0.5 * (n_before + n_after)
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
        "pct_hbm_peak": 100.0 * hbm_step / HBM_PEAK,
        "hbm_model": "144 B/particle-step + 32 B/cell-step (algorithmic, whole step)",
        "kernel_ms_per_step": {"scan": float(tm[0]), "scatter": float(tm[1]), "tile": float(tm[2]),
                               "walls": float(tm[6]), "inject": float(tm[3]), "total": float(tm[4]),
                               "timed_steps": int(tm[5])},
        "roofline": {"bound": "hbm", "achieved": tile_gbs / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": tile_gbs / HBM_PEAK, "traffic": measured_traffic(args.workload),
                     "kernel": "k_tile (56 B/particle + 20 B/cell algorithmic)"},
        "validate_bad_particles": int(bad),
    }
    sol.close()
    if not args.no_ensemble:
        t0 = time.perf_counter()
        line["ensemble"] = ensemble_bench(args.ensemble_steps, 500)
        print(f"[bench] ensemble {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        line["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        print(f"[bench] cpu baseline {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
