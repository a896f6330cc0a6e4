"""GPU parity at the measured sizes, bit-exact against the oracle.

* The reference's own benchmark case: data/config_500k.xml (32^3 cells,
  17 particles per cell, periodic x) with data/ares.raw on a 5 x 5 x 20
  collision grid moved by (0, 0, -3) and MT collisions (test.bat:1,
  kernels/dsmc.cu:398-405), REFERENCE mode.
* The c1 (100k particles, 64x32x32, 1280-triangle icosphere), c2 (2^20,
  128x64x64, 20 480 triangles) and c3 (2^25, 256x128x128, 48 800-triangle
  sphere-cone) workloads, PHYSICS mode, with sampling.
* The 64-member ensemble of the benchmark (draw_cases(64, seed=2024)), run
  as one batch, every member against its own oracle.

Each run prints what it is about to do before the GPU work starts.
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np
import pytest

import cudadsmc_amd as dm
from cudadsmc_amd import _capi, cases
from cudadsmc_amd import ensemble as ens
import oracle

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def note(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def mirror(sol: dm.Solver, mesh=None) -> oracle.OracleSim:
    o = oracle.OracleSim.from_solver(sol)
    if mesh is not None:
        phys = sol.settings.mode == dm.MODE_PHYSICS
        o.set_geometry(mesh, sol.collision_grid(), sol.fluid_fraction() if phys else None)
        if phys:
            o.set_near_cells(sol.near_cells())
    o.set_rng(sol.rng_state())
    o.set_ntc(sol.ntc_state())
    return o


def same_state(sol: dm.Solver, o: oracle.OracleSim, tag: str) -> None:
    g, w = sol.particles(), o.state()
    assert g["pos"].shape == w["pos"].shape, f"{tag}: particle count"
    assert np.array_equal(g["key"], w["key"]), f"{tag}: keys"
    assert np.array_equal(g["pos"], w["pos"]), f"{tag}: positions"
    assert np.array_equal(g["vel"], w["vel"]), f"{tag}: velocities"


def steps_bit_exact(sol, o, steps: int, sample: bool, tag: str, perm: bool = True) -> None:
    for k in range(steps):
        sol.RunBenchmarkSimulationStep(sol.settings.dt)
        o.step(sample=sample)
        t = f"{tag} step {k + 1}"
        assert np.array_equal(sol.cell_start(), o.cell_start()), f"{t}: cell offsets"
        if perm:
            npart = int(o.cell_start()[-1])
            assert np.array_equal(sol.perm(npart), o.perm(npart)), f"{t}: permutation"
        same_state(sol, o, t)
        assert np.array_equal(sol.ntc_state(), o.ntc()), f"{t}: ntc"
    assert np.array_equal(sol.rng_state(), o.rng()), f"{tag}: rng state"


# ------------------------------------------------- the reference's benchmark
def test_gpu_reference_benchmark_case_500k_ares():
    note("reference benchmark case: config_500k.xml + ares.raw (5x5x20 grid moved by (0,0,-3)), mt, 40 steps")
    s = dm.settings_from_file(str(GOLDEN / "config_500k.xml"))
    s.benchmark, s.geometry, s.mt = 1, 1, 1
    s.dt = 0.25e-4  # test.bat:1
    s.sample_every = 7  # the reference samples every 100th step; sample within the window
    assert s.part_per_cell == 17 and s.num_particles == 557056
    sol = dm.Solver()
    sol.Init(s)
    mesh = dm.load_raw(str(GOLDEN / "ares.raw"))
    sol.SetModel(mesh, grid=(5, 5, 20), move=(0.0, 0.0, -3.0))
    moved = mesh.translate((0.0, 0.0, -3.0))
    assert int(sol.collision_grid()["flags"].astype(bool).sum()) > 0
    o = mirror(sol, moved)
    U = tuple(s.stream_velocity)
    n = sol.InitSystemState(U)
    o.load(U)
    assert n == 557056
    same_state(sol, o, "load")
    # particles of cells that touch a flagged collision-grid cell are parked
    # at the origin with zero velocity (dsmc_kernel.cu:669-673)
    parked = int((np.abs(sol.particles()["vel"]).sum(0) == 0).sum())
    assert parked > 0
    sol.set_debug(_capi.FLAG_KEEP_PERM | _capi.FLAG_HITS)
    hits = 0
    # the particles start at least about one flow cell (0.6 m) from the
    # model (cells touching it are parked): the first wall hits take a few
    # tens of 25 us steps at 1000 m/s
    for k in range(40):
        sol.RunBenchmarkSimulationStep(s.dt)
        o.step(sample=((k + 1) % 7 == 0))
        npart = int(o.cell_start()[-1])
        assert np.array_equal(sol.cell_start(), o.cell_start()), f"step {k + 1}: cell offsets"
        h = sol.hit_counts(npart)
        assert np.array_equal(h, o.hits(npart)), f"step {k + 1}: wall hits"
        hits += int(h.sum())
        if k % 8 == 7 or k == 39:
            assert np.array_equal(sol.perm(npart), o.perm(npart)), f"step {k + 1}: permutation"
            same_state(sol, o, f"step {k + 1}")
            assert np.array_equal(sol.ntc_state(), o.ntc()), f"step {k + 1}: ntc"
    assert np.array_equal(sol.rng_state(), o.rng())
    assert np.array_equal(sol.cell_stats(), o.samples())
    coll = sol.collision_stats()
    ref = o.cstat().astype(np.float64).sum(0)
    assert np.allclose(coll[:3], ref[:3], rtol=1e-12)
    assert coll[1] > 0, "no collisions"
    assert hits > 0, "no wall hits"


# ------------------------------------------------------- c1, c2, c3 workloads
@pytest.mark.parametrize("name,steps", [("c1", 6), ("c2", 5), ("c3", 2)])
def test_gpu_workload_bit_exact(name, steps):
    note(f"{name} workload: load + {steps} steps with sampling, bit-exact against the oracle")
    sol, s, mesh = cases.make_solver(name)
    o = mirror(sol, mesh)
    U = tuple(s.stream_velocity)
    n = sol.InitSystemState(U)
    o.load(U)
    note(f"{name}: {n} particles loaded")
    assert n == o.state()["pos"].shape[1]
    same_state(sol, o, f"{name} load")
    sol.set_debug(_capi.FLAG_KEEP_PERM)
    sol.set_sampling(True)
    steps_bit_exact(sol, o, steps, True, name)
    assert np.array_equal(sol.cell_stats(), o.samples()), f"{name}: samples"
    acc = o.acc()
    coll = sol.collision_stats()
    assert np.allclose(coll[:3], acc[:3], rtol=1e-12), f"{name}: collision totals"
    assert coll[1] > 0
    F = sol.drag()
    want = acc[3:6] * float(sol.derived()["fnum"]) / (steps * s.dt)
    assert np.allclose(F[:3], want, rtol=1e-9, atol=1e-12 * np.abs(want).max()), (F, want)
    assert sol.validate() == 0
    sol.close()


# ------------------------------------------------------- 64-member ensemble
def test_gpu_ensemble64_members_match_oracle():
    note("64-member ensemble (draw_cases(64, seed=2024)) as one batch, 4 steps, every member against its oracle")
    cs = ens.draw_cases(64, seed=2024, temperatures=False)
    ns = int(ens.tuned("ensemble64", 16384))
    sols = [ens.make_solver(c, n_streams=ns) for c in cs]
    orcs = []
    for c, sol in zip(cs, sols):
        o = oracle.OracleSim.from_solver(sol)
        mesh = dm.icosphere(ens.CENTER, c.r, ens.ICO_SUBDIV)
        o.set_geometry(mesh, sol.collision_grid(), sol.fluid_fraction())
        o.set_near_cells(sol.near_cells())
        o.set_ntc(sol.ntc_state())
        o.set_rng(sol.rng_state())  # after loading: make_solver loads the particles
        o.set_state(sol.particles()["pos"], sol.particles()["vel"])
        orcs.append(o)
    note(f"ensemble: {sum(s.counts()[1] for s in sols)} particles in {len(sols)} members")
    b = dm.Batch(sols)
    b.set_sampling(True)
    steps = 4
    b.run(steps)
    b.synchronize()
    for o in orcs:
        for _ in range(steps):
            o.step(sample=True)
    for i, (sol, o) in enumerate(zip(sols, orcs)):
        same_state(sol, o, f"member {i}")
        assert np.array_equal(sol.cell_start(), o.cell_start()), f"member {i}: cell offsets"
        assert np.array_equal(sol.ntc_state(), o.ntc()), f"member {i}: ntc"
        assert np.array_equal(sol.rng_state(), o.rng()), f"member {i}: rng"
        assert np.array_equal(sol.cell_stats(), o.samples()), f"member {i}: samples"
    b.close()
    for sol in sols:
        sol.close()
