"""Autotuner of the DSMC step's launch layout.

Launch parameters are tuned by timing the real case on the GPU:

* ``n_streams``: the number of RNG streams, which fixes the tile kernel's
  grid (n_streams / 256 workgroups, each walking tiles vwg, vwg + G, ...).
  More workgroups give more parallelism; fewer give longer tile walks and
  less RNG-state traffic. It fixes the stream assignment, so it is chosen
  before particles are loaded and must be recorded with the results (a run
  is reproducible for a given value).
* ``scatter_grid``: workgroups of the grid-stride scatter kernel (does not
  change results).

For the ensemble it also picks the cells per tile: its instances are small
(65536 cells), so the engine's sizing (tiles of about 1000 expected
particles) leaves about one tile per workgroup and no next tile to prefetch;
smaller tiles give each workgroup a walk of several tiles. Like n_streams,
the tile size fixes the stream assignment.

For the large single case it also picks the collision-grid refinement (cells
of the triangle grid per flow cell along each axis): a finer grid holds
fewer triangles per cell and marks fewer flow cells as near the body, at the
price of more cells to traverse. Nearest-hit results do not depend on it.

The tuner builds one solver per candidate (``make(n_streams) -> Solver``,
loaded and ready), warms it up, times steps between device
synchronisations, and keeps the fastest. Results can be cached in a JSON
file keyed by a case signature.
"""
from __future__ import annotations

import json
import time
from pathlib import Path
from typing import Callable, Iterable

DEFAULT_STREAMS = (16384, 32768, 65536, 131072, 262144, 524288)
DEFAULT_SCATTER = (1024, 2048, 4096, 8192, 16384)


def _time_steps(run: Callable[[int], None], sync: Callable[[], None], warmup: int, steps: int) -> float:
    run(warmup)
    sync()
    t0 = time.perf_counter()
    run(steps)
    sync()
    return (time.perf_counter() - t0) / steps


def tune_solver(make: Callable[[int], "object"], streams: Iterable[int] = DEFAULT_STREAMS,
                scatter: Iterable[int] = DEFAULT_SCATTER, warmup: int = 20, steps: int = 40,
                cache: str | Path | None = None, key: str | None = None) -> dict:
    """Best {n_streams, scatter_grid} for one solver case."""
    if cache and key:
        db = _load(cache)
        if key in db:
            return db[key]
    table = []
    best = None
    for ns in streams:
        sol = make(int(ns))
        dt = sol.settings.dt
        t = _time_steps(lambda k: sol.DoBenchmark(k, dt), sol.synchronize, warmup, steps)
        table.append({"n_streams": int(ns), "scatter_grid": None, "ms_per_step": 1e3 * t})
        print("solver", json.dumps(table[-1]), flush=True)
        if best is None or t < best[0]:
            best = (t, int(ns))
        sol.close()
    sol = make(best[1])
    dt = sol.settings.dt
    sbest = None
    for g in scatter:
        sol.set_scatter_grid(int(g))
        t = _time_steps(lambda k: sol.DoBenchmark(k, dt), sol.synchronize, warmup // 2 + 2, steps)
        table.append({"n_streams": best[1], "scatter_grid": int(g), "ms_per_step": 1e3 * t})
        print("solver", json.dumps(table[-1]), flush=True)
        if sbest is None or t < sbest[0]:
            sbest = (t, int(g))
    sol.close()
    out = {"n_streams": best[1], "scatter_grid": sbest[1], "ms_per_step": 1e3 * sbest[0], "table": table}
    if cache and key:
        db = _load(cache)
        db[key] = out
        Path(cache).write_text(json.dumps(db, indent=1))
    return out


def tune_batch(make_members: Callable[[int, int], list], streams: Iterable[int] = (4096, 8192, 16384, 32768, 65536),
               tiles: Iterable[int] = (0,), warmup: int = 10, steps: int = 20) -> dict:
    """Best per-instance (n_streams, tile_cells) for an ensemble stepped as
    one batch; ``make_members(n_streams, tile_cells)`` builds the loaded
    members (tile_cells 0: the engine's sizing)."""
    import cudadsmc_amd as dm

    table = []
    best = None
    for tc in tiles:
        for ns in streams:
            members = make_members(int(ns), int(tc))
            b = dm.Batch(members)
            t = _time_steps(b.run, b.synchronize, warmup, steps)
            table.append({"n_streams": int(ns), "tile_cells": int(tc), "ms_per_step": 1e3 * t})
            print("ensemble", json.dumps(table[-1]), flush=True)
            if best is None or t < best[0]:
                best = (t, int(ns), int(tc))
            b.close()
            for m in members:
                m.close()
    return {"n_streams": best[1], "tile_cells": best[2], "ms_per_step": 1e3 * best[0], "table": table}


def _load(path) -> dict:
    p = Path(path)
    if p.exists():
        try:
            return json.loads(p.read_text())
        except json.JSONDecodeError:
            return {}
    return {}


def main(argv=None) -> None:
    """Tune the ensemble (64 instances of the 64x32x32 sphere case) and the
    large single case; write cudadsmc_amd/tuned.json."""
    import argparse

    from . import ensemble as ens

    ap = argparse.ArgumentParser(description="tune launch parameters on the GPU")
    ap.add_argument("--what", default="ensemble64,c3", help="comma list of ensemble64, c3")
    ap.add_argument("--out", default=str(ens.TUNED))
    a = ap.parse_args(argv)
    out = Path(a.out)
    db = _load(out)
    for what in a.what.split(","):
        if what == "ensemble64":
            cases = ens.draw_cases(64, seed=2024, temperatures=False)
            res = tune_batch(lambda ns, tc: [ens.make_solver(c, n_streams=ns, tile_cells=tc) for c in cases],
                             streams=(32768, 65536, 131072), tiles=(32, 64, 128, 0), warmup=300, steps=60)
        elif what == "c3":
            import sys

            sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
            from bench import make_case  # the benchmark's workload definition

            from . import Solver

            def make(ns, refine):
                s, mesh, cg = make_case("c3", cg_refine=refine)
                s.n_streams = ns
                sol = Solver()
                sol.Init(s)
                sol.SetModel(mesh, grid=cg)
                sol.InitSystemState(tuple(s.stream_velocity))
                return sol

            # collision-grid refinement first (engine-default streams), then
            # the launch layout on the chosen grid
            # timed on the developed flow (shock layer and wake formed)
            rtab = []
            for refine in (1, 2, 3, 4):
                sol = make(0, refine)
                dt = sol.settings.dt
                t = _time_steps(lambda k: sol.DoBenchmark(k, dt), sol.synchronize, 600, 40)
                sol.close()
                rtab.append({"cg_refine": refine, "ms_per_step": 1e3 * t})
                print("c3", json.dumps(rtab[-1]), flush=True)
            refine = min(rtab, key=lambda r: r["ms_per_step"])["cg_refine"]
            res = tune_solver(lambda ns: make(ns, refine), streams=(262144, 524288, 1048576, 2097152),
                              scatter=(2048, 4096, 8192, 16384, 32768), warmup=600, steps=40)
            res["cg_refine"] = refine
            res["table"] = rtab + res["table"]
        else:
            raise SystemExit(f"unknown tuning target {what}")
        db[what] = res
        print(what, json.dumps({k: v for k, v in res.items() if k != "table"}), flush=True)
        out.write_text(json.dumps(db, indent=1))


if __name__ == "__main__":
    main()
