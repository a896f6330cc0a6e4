"""Batched ensembles: instances stepped together by shared launches must be
bit-identical to the same instances stepped alone (GPU)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import cudadsmc_amd as dm

from helpers import small_physics_settings

pytestmark = pytest.mark.gpu


def make(seed: int, n_target: int, U: float, r: float, dims=(24, 12, 12)):
    s = small_physics_settings(dims=dims, n_target=n_target, U=U)
    s.seed = seed
    sol = dm.Solver()
    sol.Init(s)
    sol.SetModel(dm.icosphere((0.07, 0.05, 0.05), r, 2), grid=(8, 8, 8))
    sol.InitSystemState((U, 0.0, 0.0))
    return sol


CASES = [(1, 20000, 900.0, 0.015), (1000, 60000, 2500.0, 0.025), (77, 35000, 1600.0, 0.02)]


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_equals_individual_runs(graphs):
    print(f"batch vs individual graphs={graphs}", flush=True)
    sys.stdout.flush()
    alone = []
    for seed, n, U, r in CASES:
        sol = make(seed, n, U, r)
        sol.set_sampling(True)
        sol.set_graphs(False)
        sol.DoBenchmark(13, sol.settings.dt)
        alone.append((sol.particles(), sol.cell_stats(), sol.rng_state(), sol.drag()))
        sol.close()
    members = [make(seed, n, U, r) for seed, n, U, r in CASES]
    batch = dm.Batch(members)
    batch.set_graphs(graphs)
    batch.set_sampling(True)
    batch.run(6)
    batch.run(7)
    batch.synchronize()
    for sol, (p, cs, rng, drag) in zip(members, alone):
        q = sol.particles()
        for k in ("pos", "vel", "key"):
            assert np.array_equal(p[k], q[k]), k
        assert np.array_equal(cs, sol.cell_stats())
        assert np.array_equal(rng, sol.rng_state())
        assert np.allclose(drag, sol.drag(), rtol=1e-12, atol=0)
        assert sol.counts()[0] == 13
    batch.close()


def test_batch_rejects_mismatched_members():
    a = make(1, 10000, 1000.0, 0.02)
    b = make(2, 10000, 1000.0, 0.02)
    a.RunBenchmarkSimulationStep(a.settings.dt)
    with pytest.raises(dm.DsmcError, match="same step"):
        dm.Batch([a, b])
    with pytest.raises(dm.DsmcError, match="twice"):
        dm.Batch([b, b])


def test_batch_members_sample_independently():
    """Only the members whose sampling is on accumulate cell samples."""
    print("batch per-member sampling", flush=True)
    cases = CASES[:2]
    alone = []
    for i, (seed, n, U, r) in enumerate(cases):
        sol = make(seed, n, U, r)
        sol.set_graphs(False)
        sol.DoBenchmark(4, sol.settings.dt)
        if i == 0:
            sol.set_sampling(True)
        sol.DoBenchmark(6, sol.settings.dt)
        alone.append((sol.particles(), sol.cell_stats(), sol.counts()[2], sol.drag()))
        sol.close()
    members = [make(seed, n, U, r) for seed, n, U, r in cases]
    batch = dm.Batch(members)
    batch.run(4)
    members[0].set_sampling(True)
    batch.run(6)
    batch.synchronize()
    for sol, (p, cs, ns, drag) in zip(members, alone):
        q = sol.particles()
        for k in ("pos", "vel", "key"):
            assert np.array_equal(p[k], q[k]), k
        assert np.array_equal(cs, sol.cell_stats())
        assert sol.counts()[2] == ns
        assert np.allclose(drag, sol.drag(), rtol=1e-12, atol=0)
    assert alone[0][2] == 6 and alone[1][2] == 0
    assert alone[1][1].sum() == 0.0 and alone[0][1].sum() > 0.0
    batch.close()
