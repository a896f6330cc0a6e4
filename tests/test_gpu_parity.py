"""GPU parity: the HIP path against the CPU oracle and the golden vectors.

Bit-exact: MT outputs, cell keys, the stable sort permutation and cell
offsets, particle state after loading and after every step, wall-hit counts,
NTC state, per-cell counters and samples, RNG state. The PHYSICS
accumulators (double sums reduced in a different order) are compared to
1e-12 relative.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import cudadsmc_amd as dm
from cudadsmc_amd import _capi
import oracle

from helpers import small_physics_settings, small_reference_settings

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a: np.ndarray):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def note(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------------ RNG
def test_gpu_mt_matches_reference_genrand(golden):
    d = np.load(golden / "mt_dcmt_rows.npz")
    rows, seeds, outs = d["rows"], d["seeds"], d["outputs"]
    n = outs.shape[1]
    a, b, c, s = (dev(x.astype(np.int32)) for x in (rows[:, 0], rows[:, 1], rows[:, 2], seeds))
    out = torch.zeros(n * len(seeds), dtype=torch.int32, device="cuda")
    _capi.check(_capi.lib().dsmc_op_mt_generate(ptr(a), ptr(b), ptr(c), ptr(s), len(seeds), n, ptr(out), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint32).reshape(n, len(seeds)).T
    assert np.array_equal(got, outs)


def test_gpu_mt_streams_vs_oracle():
    rng = np.random.default_rng(7)
    ns, nd = 4096, 300
    text = (oracle.HERE.parent / "cudadsmc_amd" / "csrc" / "host" / "mt_params_table.inc").read_text()
    rows = np.array([int(t.rstrip("u,"), 16) for t in text.split() if t.startswith("0x")], np.uint32)
    a = rows[np.arange(ns) % len(rows)]
    b = np.full(ns, 0x9D2C5680, np.uint32)
    c = np.full(ns, 0xEFC60000, np.uint32)
    seeds = rng.integers(0, 2**32, ns, dtype=np.uint64).astype(np.uint32)
    ta, tb, tc, ts = (dev(x.view(np.int32)) for x in (a, b, c, seeds))
    out = torch.zeros(ns * nd, dtype=torch.int32, device="cuda")
    _capi.check(_capi.lib().dsmc_op_mt_generate(ptr(ta), ptr(tb), ptr(tc), ptr(ts), ns, nd, ptr(out), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint32).reshape(nd, ns)
    assert np.array_equal(got, oracle.mt_streams(a, b, c, seeds, nd))


# ----------------------------------------------------------------- keys
@pytest.mark.parametrize("n", [0, 1, 1000, 300000])
def test_gpu_cell_ids(n):
    s = dm.settings_default()
    s.bmin[:] = (-0.1, 0.0, 0.2)
    s.bmax[:] = (0.3, 0.2, 0.5)
    s.grid_dim[:] = (64, 32, 48)
    rng = np.random.default_rng(n)
    # inside, on faces, outside on both sides (saturating keys), NaN
    x = rng.uniform(-0.2, 0.4, n).astype(np.float32)
    y = rng.uniform(-0.1, 0.3, n).astype(np.float32)
    z = rng.uniform(0.1, 0.6, n).astype(np.float32)
    if n >= 8:
        x[:4] = [-0.1, 0.3, np.nextafter(np.float32(0.3), np.float32(1)), np.nan]
    want = oracle.cell_ids(s, x, y, z)
    pad = lambda a: np.concatenate([a, np.zeros(1, np.float32)])  # noqa: E731  (non-null buffers for n = 0)
    tx, ty, tz = dev(pad(x)), dev(pad(y)), dev(pad(z))
    out = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    _capi.check(_capi.lib().dsmc_op_cell_ids(C.byref(s), ptr(tx), ptr(ty), ptr(tz), n, ptr(out), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint32)[:n]
    assert np.array_equal(got, want)
    assert (got < 64 * 32 * 48).all()


# ----------------------------------------------------------------- sort
def _gpu_sort(keys: np.ndarray, n_cells: int):
    n = len(keys)
    tk = dev(keys.view(np.int32)) if n else torch.zeros(1, dtype=torch.int32, device="cuda")
    perm = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    cs = torch.zeros(n_cells + 1, dtype=torch.int32, device="cuda")
    work = torch.zeros(int(_capi.lib().dsmc_op_sort_work_words(n, n_cells)), dtype=torch.int32, device="cuda")
    _capi.check(_capi.lib().dsmc_op_sort(ptr(tk), n, n_cells, ptr(perm), ptr(cs), ptr(work), None))
    torch.cuda.synchronize()
    cs_ = cs.cpu().numpy().view(np.uint32)
    return perm.cpu().numpy().view(np.uint32)[: cs_[-1]], cs_


@pytest.mark.parametrize("case", ["random", "nearly_sorted", "one_cell", "one_cell_large", "invalid", "out_of_range",
                                  "empty", "single", "many_cells", "one_pass_cells"])
def test_gpu_stable_sort(case):
    rng = np.random.default_rng(11)
    C_ = 65536
    if case == "random":
        keys = rng.integers(0, C_, 500000).astype(np.uint32)
    elif case == "nearly_sorted":
        keys = np.sort(rng.integers(0, C_, 500000)).astype(np.uint32)
        mv = rng.random(len(keys)) < 0.3
        keys[mv] = np.clip(keys[mv].astype(np.int64) + rng.choice([-1, 1, -256, 256], mv.sum()), 0, C_ - 1)
    elif case == "one_cell":
        keys = np.full(20000, 123, np.uint32)
    elif case == "one_cell_large":  # a single segment of 2M keys (linear work per radix pass)
        keys = np.full(2_000_000, 65535, np.uint32)
    elif case == "invalid":
        keys = rng.integers(0, C_, 100000).astype(np.uint32)
        keys[rng.random(len(keys)) < 0.2] = 0xFFFFFFFF
    elif case == "out_of_range":  # any key >= n_cells is dropped, not only 0xFFFFFFFF
        keys = rng.integers(0, 2 * C_, 100000).astype(np.uint32)
    elif case == "empty":
        keys = np.zeros(0, np.uint32)
    elif case == "single":
        keys = np.array([7], np.uint32)
    elif case == "one_pass_cells":  # n_cells < 256: one radix pass
        C_ = 200
        keys = rng.integers(0, 210, 300000).astype(np.uint32)
    else:
        C_ = 4 * 1024 * 1024
        keys = rng.integers(0, C_, 2_000_000).astype(np.uint32)
    note(f"sort case {case}: {len(keys)} keys, {C_} cells")
    want_p, want_c = oracle.stable_sort(keys, C_)
    got_p, got_c = _gpu_sort(keys, C_)
    assert np.array_equal(got_c, want_c)
    assert np.array_equal(got_p, want_p)
    valid = keys < C_
    assert np.array_equal(want_p, np.argsort(np.where(valid, keys.astype(np.int64), 2**32),
                                             kind="stable")[: valid.sum()])


# ------------------------------------------------------------ full steps
def _compare_state(sol: dm.Solver, o: oracle.OracleSim, tag: str) -> None:
    g = sol.particles()
    w = o.state()
    assert g["pos"].shape == w["pos"].shape, tag
    assert np.array_equal(g["key"], w["key"]), f"{tag}: keys"
    np.testing.assert_array_equal(g["pos"], w["pos"], err_msg=f"{tag}: positions")
    np.testing.assert_array_equal(g["vel"], w["vel"], err_msg=f"{tag}: velocities")


def _mirror(sol: dm.Solver, mesh=None) -> oracle.OracleSim:
    o = oracle.OracleSim.from_solver(sol)
    if mesh is not None:
        ff = sol.fluid_fraction() if sol.settings.mode == dm.MODE_PHYSICS else None
        o.set_geometry(mesh, sol.collision_grid(), ff)
        if sol.settings.mode == dm.MODE_PHYSICS:
            o.set_near_cells(sol.near_cells())
    o.set_rng(sol.rng_state())
    o.set_ntc(sol.ntc_state())
    return o


@pytest.mark.parametrize("periodic,geom,mt", [("x", True, 1), ("z", True, 1), ("n", False, 1), ("y", True, 1),
                                               ("x", True, 0), ("n", False, 0)])
def test_gpu_reference_steps_bit_exact(periodic, geom, mt):
    """mt = 1: MT collision streams; mt = 0: the reference's default Hybrid-Taus
    per-cell streams (ComputeBirdCollisionsNoMTDev)."""
    note(f"reference step parity periodic={periodic} geometry={geom} mt={mt}")
    s = small_reference_settings(periodic=periodic, geometry=geom)
    s.sample_every = 3
    s.mt = mt
    sol = dm.Solver()
    sol.Init(s)
    mesh = None
    if geom:
        mesh = dm.icosphere((0.07, 0.06, 0.05), 0.022, 2)
        sol.SetModel(mesh, grid=(6, 6, 6))
    o = _mirror(sol, mesh)
    if not mt:
        o.set_taus(sol.taus_state())
    U = (900.0, 150.0, -120.0)
    n = sol.InitSystemState(U)
    o.load(U)
    assert n == o.state()["pos"].shape[1]
    _compare_state(sol, o, "load")
    assert np.array_equal(sol.rng_state(), o.rng()), "rng after load"
    sol.set_debug(_capi.FLAG_KEEP_PERM | _capi.FLAG_HITS)
    total_hits = 0
    for k in range(8):
        sol.RunBenchmarkSimulationStep(s.dt)
        o.step(sample=((k + 1) % 3 == 0))
        tag = f"step {k + 1}"
        assert np.array_equal(sol.cell_start(), o.cell_start()), f"{tag}: cell offsets"
        npart = int(o.cell_start()[-1])
        assert np.array_equal(sol.perm(npart), o.perm(npart)), f"{tag}: permutation"
        hits = sol.hit_counts(npart)
        assert np.array_equal(hits, o.hits(npart)), f"{tag}: wall hits"
        total_hits += int(hits.sum())
        _compare_state(sol, o, tag)
        assert np.array_equal(sol.ntc_state(), o.ntc()), f"{tag}: ntc"
        if not mt:
            assert np.array_equal(sol.taus_state(), o.taus()), f"{tag}: taus state"
    assert np.array_equal(sol.rng_state(), o.rng()), "rng after steps"
    cs = sol.cell_stats()
    assert np.array_equal(cs, o.samples()), "samples"
    coll = sol.collision_stats()
    ref = o.cstat().astype(np.float64).sum(0)
    assert np.allclose(coll[:3], ref[:3], rtol=1e-12)
    assert coll[1] > 0, "no collisions happened"
    if geom:
        assert total_hits > 0, "no wall hits happened"


def test_gpu_physics_steps_bit_exact():
    note("physics step parity")
    s = small_physics_settings(dims=(24, 12, 12), n_target=30000)
    sol = dm.Solver()
    sol.Init(s)
    mesh = dm.icosphere((0.07, 0.05, 0.05), 0.02, 2)
    sol.SetModel(mesh, grid=(8, 8, 8))
    o = _mirror(sol, mesh)
    n = sol.InitSystemState(tuple(s.stream_velocity))
    o.load(tuple(s.stream_velocity))
    assert n == o.state()["pos"].shape[1]
    _compare_state(sol, o, "load")
    sol.set_debug(_capi.FLAG_KEEP_PERM | _capi.FLAG_HITS)
    sol.set_sampling(True)
    total_hits = 0
    for k in range(10):
        sol.RunBenchmarkSimulationStep(s.dt)
        o.step(sample=True)
        tag = f"step {k + 1}"
        assert np.array_equal(sol.cell_start(), o.cell_start()), f"{tag}: cell offsets"
        npart = int(o.cell_start()[-1])
        assert np.array_equal(sol.perm(npart), o.perm(npart)), f"{tag}: permutation"
        hits = sol.hit_counts(npart)
        assert np.array_equal(hits, o.hits(npart)), f"{tag}: wall hits"
        total_hits += int(hits.sum())
        _compare_state(sol, o, tag)
        assert np.array_equal(sol.ntc_state(), o.ntc()), f"{tag}: ntc"
    assert np.array_equal(sol.rng_state(), o.rng())
    assert np.array_equal(sol.cell_stats(), o.samples())
    acc = o.acc()
    coll = sol.collision_stats()
    assert np.allclose(coll[:3], acc[:3], rtol=1e-12)
    assert coll[1] > 0 and total_hits > 0
    # wall force of the wall pass: same per-hit terms, summed in another order
    F = sol.drag()
    want = acc[3:6] * float(sol.derived()["fnum"]) / (10 * s.dt)
    assert np.allclose(F[:3], want, rtol=1e-9, atol=1e-12 * np.abs(want).max()), (F, want)


def test_gpu_overflow_tiles_bit_exact():
    # dense cells: tiles above the LDS capacity take the global-memory path
    note("overflow tile parity")
    s = small_reference_settings(periodic="x", dims=(16, 4, 4), ppc=40, geometry=False)
    sol = dm.Solver()
    sol.Init(s)
    sol.set_tile_cells(256)  # one 256-cell tile of 10240 particles (the default would be 32-cell tiles)
    assert sol.layout()["tile_cells"] == 256
    o = _mirror(sol)
    n = sol.InitSystemState((300.0, 0.0, 0.0))
    o.load((300.0, 0.0, 0.0))
    assert n == 16 * 4 * 4 * 40 and n > 2048
    sol.set_debug(_capi.FLAG_KEEP_PERM)
    for k in range(4):
        sol.RunBenchmarkSimulationStep(s.dt)
        o.step(sample=False)
        _compare_state(sol, o, f"step {k + 1}")
        npart = int(o.cell_start()[-1])
        assert np.array_equal(sol.perm(npart), o.perm(npart))
    assert np.array_equal(sol.rng_state(), o.rng())


def test_gpu_graph_replay_matches_eager():
    # dsmc_run replays two-step HIP graphs; several replays, interleaved with
    # eager steps and host synchronisation, must equal a fully eager run
    note("graph replay vs eager")
    states = []
    for graphs in (True, False):
        s = small_physics_settings(dims=(24, 12, 12), n_target=40000)
        sol = dm.Solver()
        sol.Init(s)
        sol.SetModel(dm.icosphere((0.07, 0.05, 0.05), 0.02, 2), grid=(8, 8, 8))
        sol.InitSystemState((1500.0, 0.0, 0.0))
        sol.set_graphs(graphs)
        sol.set_sampling(True)
        for chunk in (7, 1, 10, 5):
            sol.DoBenchmark(chunk, s.dt)
            sol.synchronize()
            sol.RunBenchmarkSimulationStep(s.dt)
        states.append((sol.particles(), sol.cell_stats(), sol.rng_state(), sol.counts()))
        sol.close()
    (p0, c0, r0, n0), (p1, c1, r1, n1) = states
    for k in ("pos", "vel", "key"):
        assert np.array_equal(p0[k], p1[k]), k
    assert np.array_equal(c0, c1) and np.array_equal(r0, r1) and n0 == n1


@pytest.mark.parametrize("grid", [64, 256, 512, 4096])
def test_gpu_scatter_pass_sizes_bit_exact(grid):
    # 1024 tiles of 64 cells with about 640 particles each (the tile queues
    # and the tile-local scatter): the scatter takes 16, 4, 2 and 1 tiles per
    # workgroup pass at these grids; every pass size gives the oracle's state
    note(f"scatter passes, grid {grid}")
    s = small_physics_settings(dims=(64, 32, 32), n_target=655360)
    sol = dm.Solver()
    sol.Init(s)
    assert sol.layout()["tile_cells"] == 64
    sol.set_scatter_grid(grid)
    mesh = dm.icosphere((0.07, 0.05, 0.05), 0.02, 2)
    sol.SetModel(mesh, grid=(8, 8, 8))
    o = _mirror(sol, mesh)
    sol.InitSystemState(tuple(s.stream_velocity))
    o.load(tuple(s.stream_velocity))
    sol.set_debug(_capi.FLAG_KEEP_PERM)
    for k in range(4):
        sol.RunBenchmarkSimulationStep(s.dt)
        o.step(sample=False)
        tag = f"grid {grid} step {k + 1}"
        assert np.array_equal(sol.cell_start(), o.cell_start()), f"{tag}: cell offsets"
        npart = int(o.cell_start()[-1])
        assert np.array_equal(sol.perm(npart), o.perm(npart)), f"{tag}: permutation"
        _compare_state(sol, o, tag)
    sol.close()


def test_gpu_capacity_limit():
    # the kernels address particle arrays by 32-bit byte offsets, so a
    # capacity above 2^30 particles is refused at init, before any allocation
    note("capacity limit")
    s = small_physics_settings()
    s.capacity = (1 << 30) + 1
    sol = dm.Solver()
    with pytest.raises(dm.DsmcError, match="capacity"):
        sol.Init(s)
    sol.close()
    s.capacity = 1 << 20
    sol = dm.Solver()
    sol.Init(s)
    sol.SetModel(dm.icosphere((0.07, 0.05, 0.05), 0.02, 2), grid=(8, 8, 8))
    assert sol.InitSystemState((1500.0, 0.0, 0.0)) > 0
    sol.DoBenchmark(2, s.dt)
    sol.close()
