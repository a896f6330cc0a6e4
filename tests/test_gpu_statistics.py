"""Statistical agreement of independent GPU and CPU runs (3 sigma, batch means).

The c1 workload (100k argon particles, 64x32x32 cells, 1280-triangle
icosphere, diffuse wall) runs on the GPU and in the CPU oracle from
different seeds, so the two are independent realisations of the same flow.
After a transient, both sample 20 windows; per window they give the body
force F_x and, on 16x16x16-cell blocks (16 of them, about 6000 particles
each per step), the particle count (density), mean x
velocity and temperature. Batch means over the windows give each arm's
mean and standard error, the variance widened by (1 + rho) / (1 - rho) for
the lag-one correlation rho of consecutive windows (pooled over the
blocks); the combined sigma is sqrt(s_gpu^2 + s_cpu^2).

Gate (the reference checks T and n of its own runs against the set
values at the end of ProcessAndOutputStatistics, kernels/dsmc_base.cu:
149-162; here the two engines are checked against each other):
  * drag: |F_gpu - F_cpu| < 3 sigma;
  * blocks: at least 95% of the 3 x 16 block z-scores within 3, none
    beyond 4 (48 comparisons: about a 0.3% chance under agreement), and
    mean z^2 below 1.6 (a chi-square per degree of freedom).
A run with the freestream speed changed by 5% must fail the same gate, so
the gate has the power to see a physics difference of that size.
Seeds are fixed, so the outcome is reproducible.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import cudadsmc_amd as dm
from cudadsmc_amd import cases
import oracle

pytestmark = pytest.mark.gpu

TRANSIENT, WINDOWS, PER = 1000, 20, 50
BLOCK = 16
KB = 1.380649e-23


def note(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def block_sums(stats: np.ndarray, grid) -> np.ndarray:
    """(C, 5) per-cell sums -> (nblocks, 5) sums over BLOCK^3 cells."""
    X, Y, Z = grid
    a = stats.reshape(Z, Y, X, 5)
    a = a.reshape(Z // BLOCK, BLOCK, Y // BLOCK, BLOCK, X // BLOCK, BLOCK, 5).sum(axis=(1, 3, 5))
    return a.reshape(-1, 5)


def window_quantities(d: np.ndarray, mass: float) -> np.ndarray:
    """Block sums of one window -> (nblocks, 3): count, mean u_x, temperature."""
    n = np.maximum(d[:, 0], 1.0)
    u = d[:, 1:4] / n[:, None]
    T = mass * (d[:, 4] / n - (u ** 2).sum(1)) / (3.0 * KB)
    return np.stack([d[:, 0], u[:, 0], T], 1)


def gpu_run(seed: int):
    s, mesh, cg = cases.make_case("c1")
    s.seed = seed
    sol = dm.Solver()
    sol.Init(s)
    sol.SetModel(mesh, grid=cg)
    sol.InitSystemState(tuple(s.stream_velocity))
    sol.DoBenchmark(TRANSIENT, s.dt)
    sol.set_sampling(True)
    prev = np.zeros((int(np.prod(s.grid_dim)), 5))
    q, fx = [], []
    imp_prev = 0.0
    for w in range(WINDOWS):
        sol.DoBenchmark(PER, s.dt)
        cs = sol.cell_stats()
        q.append(window_quantities(block_sums(cs - prev, s.grid_dim), s.molmass))
        prev = cs
        imp = float(sol.drag()[0]) * PER * (w + 1)  # drag() averages over the sampled steps
        fx.append((imp - imp_prev) / PER)
        imp_prev = imp
    sol.close()
    return np.stack(q), np.asarray(fx)


def cpu_run(seed: int, U_scale: float = 1.0):
    s, mesh, cg = cases.make_case("c1")
    s.seed = seed
    s.stream_velocity[0] = s.stream_velocity[0] * U_scale
    sol = dm.Solver()  # supplies the derived scalars, grid and stream layout; holds no particles
    sol.Init(s)
    sol.SetModel(mesh, grid=cg)
    o = oracle.OracleSim.from_solver(sol)
    o.set_geometry(mesh, sol.collision_grid(), sol.fluid_fraction())
    o.set_near_cells(sol.near_cells())
    o.set_rng(sol.rng_state())
    o.set_ntc(sol.ntc_state())
    fnum = float(sol.derived()["fnum"])
    sol.close()
    o.load(tuple(s.stream_velocity))
    for _ in range(TRANSIENT):
        o.step(sample=False)
    prev = o.samples()
    acc_prev = o.acc()[3]
    q, fx = [], []
    for _ in range(WINDOWS):
        for _ in range(PER):
            o.step(sample=True)
        cs = o.samples()
        q.append(window_quantities(block_sums(cs - prev, s.grid_dim), s.molmass))
        prev = cs
        acc = o.acc()[3]
        fx.append((acc - acc_prev) * fnum / (PER * s.dt))  # mean force over the window
        acc_prev = acc
    return np.stack(q), np.asarray(fx)


def lag1(x: np.ndarray) -> float:
    """Lag-one correlation of consecutive windows, pooled over the trailing
    axes, clipped to [0, 0.8]."""
    d = x - x.mean(0)
    num = (d[1:] * d[:-1]).sum()
    den = (d * d).sum()
    return float(np.clip(num / den if den > 0 else 0.0, 0.0, 0.8))


def sem(x: np.ndarray) -> np.ndarray:
    """Standard error of the mean of the windows, widened for their
    correlation."""
    rho = lag1(x)
    return x.std(0, ddof=1) / np.sqrt(len(x)) * np.sqrt((1 + rho) / (1 - rho))


def gate(g, c) -> dict:
    (qg, fg), (qc, fc) = g, c
    sig_f = float(np.hypot(sem(fg), sem(fc)))
    z_f = (fg.mean() - fc.mean()) / sig_f
    mg, mc = qg.mean(0), qc.mean(0)
    sig = np.stack([np.hypot(sem(qg[:, :, k]), sem(qc[:, :, k])) for k in range(3)], 1)
    ok = (sig > 0) & (mg[:, 0:1] > 100).repeat(3, 1)  # blocks holding gas (not inside the body)
    z = np.where(ok, (mg - mc) / np.where(sig > 0, sig, 1.0), 0.0)[ok]
    zall = np.where(ok, (mg - mc) / np.where(sig > 0, sig, 1.0), 0.0)
    worst = np.unravel_index(int(np.argmax(np.abs(zall))), zall.shape)
    return {"z_drag": float(z_f), "F_gpu": float(fg.mean()), "F_cpu": float(fc.mean()),
            "frac3": float((np.abs(z) < 3).mean()), "zmax": float(np.abs(z).max()), "chi2": float((z ** 2).mean()),
            "n": int(z.size), "worst_block_quantity": [int(worst[0]), int(worst[1])]}


def passes(r: dict) -> bool:
    return abs(r["z_drag"]) < 3 and r["frac3"] >= 0.95 and r["zmax"] < 4.0 and r["chi2"] < 1.6


def test_gpu_cpu_batch_means_agree_within_3_sigma():
    note(f"c1 GPU run (seed 11): {TRANSIENT} + {WINDOWS}x{PER} steps")
    g = gpu_run(11)
    note(f"c1 oracle run (seed 22): {TRANSIENT} + {WINDOWS}x{PER} steps")
    c = cpu_run(22)
    r = gate(g, c)
    note(f"gate: {r}")
    assert r["n"] >= 40
    assert r["F_gpu"] > 0 and r["F_cpu"] > 0
    assert passes(r), r
    note("oracle run with U + 5% (seed 22) against the same GPU run")
    r5 = gate(g, cpu_run(22, U_scale=1.05))
    note(f"gate (U + 5%): {r5}")
    assert not passes(r5), r5
