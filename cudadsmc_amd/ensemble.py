"""Ensembles of the argon sphere case: parameter draws, solver set-up,
batched runs with sampled fields and batch-means drag statistics.

Every instance is the 0.4 x 0.2 x 0.2 m box on a 64 x 32 x 32 cell grid with
an icosphere (1280 triangles, diffuse wall) centred at (0.1, 0.1, 0.1), in
PHYSICS mode (argon VHS, freestream faces). Instances differ in freestream
speed U, density n, temperature T_inf, wall temperature T_w, sphere radius r,
simulated particle count N and seed; a batch of them advances through the
engine's batched step (one launch per step kernel for all instances, HIP
graph replay).
"""
from __future__ import annotations

import json
import math
import time
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np

from . import Batch, Solver, icosphere, settings_default
from ._capi import FACE_FREESTREAM, MODE_PHYSICS

BOX_MIN = (0.0, 0.0, 0.0)
BOX_MAX = (0.4, 0.2, 0.2)
GRID = (64, 32, 32)  # X, Y, Z
CENTER = (0.1, 0.1, 0.1)
ARGON = {"molmass": 6.63e-26, "diameter": 4.17e-10, "omega": 0.81, "Tref": 273.0}
ICO_SUBDIV = 3  # 1280 triangles
# Launch parameters chosen by cudadsmc_amd.autotune on the GPU (written by
# `python -m cudadsmc_amd.autotune`); absent entries fall back to defaults.
TUNED = Path(__file__).resolve().parent / "tuned.json"


def tuned(key: str, default, field: str = "n_streams"):
    try:
        v = json.loads(TUNED.read_text()).get(key, {}).get(field)
    except (OSError, ValueError, AttributeError):
        return default
    return default if v is None else v


@dataclass
class Case:
    U: float
    n: float
    T_inf: float
    T_w: float
    r: float
    N: int
    seed: int

    def sphere(self) -> tuple[float, float, float, float]:
        return (*CENTER, self.r)

    def cd_scale(self) -> float:
        """Dynamic pressure times the frontal area: F_x / this = C_d."""
        return 0.5 * self.n * ARGON["molmass"] * self.U ** 2 * math.pi * self.r ** 2


def draw_cases(k: int, seed: int, N_range=(1.0e4, 5.0e5), temperatures=True, seed0: int = 0) -> list[Case]:
    """k cases: U ~ U[600, 3000] m/s, n log-uniform in [3e19, 1e21] m^-3,
    r ~ U[0.015, 0.035] m, N log-uniform in N_range; T_inf ~ U[150, 300] K and
    T_w ~ U[250, 600] K (or 200 K / 300 K with temperatures=False); instance
    seeds seed0, seed0 + 1, ..."""
    g = np.random.default_rng(seed)
    out = []
    for i in range(k):
        U = g.uniform(600.0, 3000.0)
        n = 10.0 ** g.uniform(math.log10(3e19), math.log10(1e21))
        r = g.uniform(0.015, 0.035)
        N = int(round(10.0 ** g.uniform(math.log10(N_range[0]), math.log10(N_range[1]))))
        Ti, Tw = (g.uniform(150.0, 300.0), g.uniform(250.0, 600.0)) if temperatures else (200.0, 300.0)
        out.append(Case(float(U), float(n), float(Ti), float(Tw), float(r), N, seed0 + i))
    return out


def settings(case: Case, dt: float = 1.0e-6, n_streams: int = 0):
    s = settings_default()
    s.mode = MODE_PHYSICS
    s.bmin[:] = BOX_MIN
    s.bmax[:] = BOX_MAX
    s.grid_dim[:] = GRID
    for a in range(3):
        s.cell_size[a] = (BOX_MAX[a] - BOX_MIN[a]) / GRID[a]
    s.molmass = ARGON["molmass"]
    s.diameter = ARGON["diameter"]
    s.omega = ARGON["omega"]
    s.Tref = ARGON["Tref"]
    s.temp = case.T_inf
    s.density = case.n
    s.stream_velocity[:] = (case.U, 0.0, 0.0)
    s.wall_temp = case.T_w
    s.face_bc[:] = [FACE_FREESTREAM] * 6
    s.target_particles = int(case.N)
    s.seed = int(case.seed)
    s.dt = dt
    s.n_streams = int(n_streams)
    return s


def collision_grid_dims(r: float) -> tuple[int, int, int]:
    cell = (BOX_MAX[0] - BOX_MIN[0]) / GRID[0]
    g = max(4, int(math.ceil(2.0 * r / cell)) + 1)
    return (g, g, g)


def make_solver(case: Case, dt: float = 1.0e-6, n_streams: int | None = None, fields: dict | None = None,
                tile_cells: int | None = None) -> Solver:
    """A loaded solver: cold (uniform freestream) or warm (fields n, u, v,
    w, T per cell, X fastest). ``n_streams`` and ``tile_cells`` default to
    the tuned ensemble layout (0 tile cells: the engine's sizing); both fix
    the stream assignment, so results are reproducible for given values."""
    if n_streams is None:
        n_streams = tuned("ensemble64", 16384)
    if tile_cells is None:
        tile_cells = int(tuned("ensemble64", 0, "tile_cells"))
    sol = Solver()
    sol.Init(settings(case, dt, n_streams))
    if tile_cells:
        sol.set_tile_cells(tile_cells)
    sol.SetModel(icosphere(CENTER, case.r, ICO_SUBDIV), grid=collision_grid_dims(case.r))
    if fields is None:
        sol.InitSystemState((case.U, 0.0, 0.0))
    else:
        sol.InitSystemStateFromFields(*(fields[k] for k in ("n", "u", "v", "w", "T")))
    return sol


class DragWindows:
    """Batch means of the drag coefficient from successive cumulative reads
    of the solver's sampled force (each read averages since sampling began)."""

    def __init__(self, case: Case, dt: float):
        self.case, self.dt = case, dt
        self.prev_impulse = 0.0
        self.prev_steps = 0
        self.means: list[float] = []

    def update(self, sol: Solver, steps_sampled: int) -> None:
        fx = float(sol.drag()[0])
        impulse = fx * steps_sampled
        k = steps_sampled - self.prev_steps
        if k > 0:
            self.means.append((impulse - self.prev_impulse) / k / self.case.cd_scale())
        self.prev_impulse, self.prev_steps = impulse, steps_sampled

    def cd(self) -> tuple[float, float]:
        """Mean and standard error of the mean over the windows."""
        m = np.asarray(self.means)
        if len(m) == 0:
            return float("nan"), float("nan")
        if len(m) == 1:
            return float(m[0]), float("inf")
        return float(m.mean()), float(m.std(ddof=1) / math.sqrt(len(m)))


def run_batch(cases: list[Case], transient: int, sampling: int, windows: int = 20, dt: float = 1.0e-6,
              n_streams: int | None = None, fields: list[dict] | None = None, graphs: bool = True,
              keep_stats: bool = False) -> dict:
    """Run the cases as one batch: transient steps, then sampling steps in
    `windows` equal windows. Returns per-case fields, drag C_d (mean,
    batch-means sigma), particle counts and timings; with `keep_stats` also
    the per-cell moment sums (C, 5) of the sampling window."""
    t0 = time.perf_counter()
    sols = [make_solver(c, dt, n_streams, None if fields is None else fields[i]) for i, c in enumerate(cases)]
    t_setup = time.perf_counter() - t0
    b = Batch(sols)
    b.set_graphs(graphs)
    b.synchronize()
    t1 = time.perf_counter()
    b.run(transient)
    b.synchronize()
    t_transient = time.perf_counter() - t1
    b.set_sampling(True)
    drags = [DragWindows(c, dt) for c in cases]
    per = max(1, sampling // windows)
    done = 0
    t2 = time.perf_counter()
    t_read = 0.0
    while done < sampling:
        k = min(per, sampling - done)
        b.run(k)
        done += k
        tr = time.perf_counter()
        b.synchronize()
        for d, s in zip(drags, sols):
            d.update(s, done)
        t_read += time.perf_counter() - tr
    t_sampling = time.perf_counter() - t2
    out = {"cases": [asdict(c) for c in cases], "fields": [], "counts": [], "cd": [], "cd_sigma": [],
           "n_particles": [],
           "fluid_fraction": None, "timing": {"setup_s": t_setup, "transient_s": t_transient,
                                              "sampling_s": t_sampling, "readout_s": t_read,
                                              "steps": transient + sampling}}
    total_np = 0
    for d, s in zip(drags, sols):
        f = s.fields()
        out["fields"].append(f)
        cs = s.cell_stats()
        out["counts"].append(cs[:, 0].astype(np.float32))
        if keep_stats:
            out.setdefault("stats", []).append(cs)
        m, e = d.cd()
        out["cd"].append(m)
        out["cd_sigma"].append(e)
        npart = s.counts()[1]
        out["n_particles"].append(npart)
        total_np += npart
    out["fluid_fraction"] = [s.fluid_fraction() for s in sols]
    b.close()
    for s in sols:
        s.close()
    return out


def stack_fields(f: dict) -> np.ndarray:
    """Fields dict -> (Z, Y, X, 5) array in the order n, u, v, w, T."""
    X, Y, Z = GRID
    return np.stack([np.asarray(f[k], np.float32).reshape(Z, Y, X) for k in ("n", "u", "v", "w", "T")], -1)
