"""Pins of the PHYSICS-mode NTC in rounds of disjoint pairs (CPU only).

* The keyed pairing permutation is a bijection of [0, n) for every n
  (brute force over n = 2..300 and several keys).
* Its matchings are uniform: over many keys, every unordered pair of a small
  cell is the k-th pair of a round equally often (chi-square), and the
  single-pair draw is uniform over ordered pairs of distinct particles.
* The oracle's cell NTC conserves momentum exactly up to FP32 rounding and
  kinetic energy likewise, draws M = floor(n (n - 1) / 2 F max dt / V + rem)
  candidates, and its number of collisions matches the expectation
    E = sum over candidate slots of E_pair[min(1, sigma c_r / max)]
  computed by brute force over all pairs of the loaded velocities; the
  velocities depend on the particle index, so a pairing that favoured some
  indices (adjacent ones, or the first few) would miss the expectation.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle


def perm(n: int, keys) -> np.ndarray:
    k = np.ascontiguousarray(keys, np.uint32)
    L = oracle.lib()
    return np.array([L.oracle_pair_perm(n, oracle.up(k), x) for x in range(n)], np.uint32)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pair_permutation_is_a_bijection(seed):
    rng = np.random.default_rng(seed)
    for n in list(range(2, 70)) + [127, 128, 129, 255, 256, 257, 300]:
        keys = rng.integers(0, 2**32, 4, dtype=np.uint64).astype(np.uint32)
        p = perm(n, keys)
        assert np.array_equal(np.sort(p), np.arange(n)), n


@pytest.mark.parametrize("n", [4, 5, 8, 16, 17, 33, 64])
def test_pair_matchings_are_uniform(n):
    """Pair k of a round, for the first, a middle and the last k, is every
    unordered pair equally often over random keys (chi-square z < 4)."""
    rng = np.random.default_rng(n)
    L = oracle.lib()
    npairs = n * (n - 1) // 2
    trials = 40 * npairs
    keys = rng.integers(0, 2**32, (trials, 4), dtype=np.uint64).astype(np.uint32)
    for k in sorted({0, n // 4, n // 2 - 1}):
        counts = np.zeros((n, n), np.int64)
        for t in range(trials):
            kk = keys[t]
            a = L.oracle_pair_perm(n, oracle.up(kk), 2 * k)
            b = L.oracle_pair_perm(n, oracle.up(kk), 2 * k + 1)
            counts[min(a, b), max(a, b)] += 1
        obs = counts[np.triu_indices(n, 1)]
        exp = trials / npairs
        chi2 = float(((obs - exp) ** 2 / exp).sum())
        dof = npairs - 1
        z = (chi2 - dof) / math.sqrt(2 * dof)
        assert z < 4.0, (n, k, z)
        assert obs.min() > 0


def _sim_physics(vhs_exp: float = 1.0, vhs_k: float = 1.0e-18, fnum: float = 1.0e12, dt: float = 1.0e-6):
    import cudadsmc_amd as dm

    s = dm.settings_default()
    s.mode = dm.MODE_PHYSICS
    s.bmin[:] = (0.0, 0.0, 0.0)
    s.bmax[:] = (0.1, 0.1, 0.1)
    s.grid_dim[:] = (4, 4, 4)
    s.seed = 7
    s.dt = dt
    s.fnum = fnum
    layout = {"tile_cells": 32, "threads": 256, "n_streams": 256}
    z = np.zeros(256, np.uint32)
    o = oracle.OracleSim(s, layout, (z, z, z))
    lam = np.zeros(6, np.float32)
    assert oracle.lib().oracle_set_scalars(o._o, fnum, vhs_k, vhs_exp, oracle.fp(lam)) == 0
    return o


def ntc_cell(o, cell: int, step: int, v: np.ndarray, vol: float, nt: np.ndarray):
    vx, vy, vz = (np.array(v[k], np.float32) for k in range(3))  # copies: the call updates them
    nt = np.ascontiguousarray(nt, np.float32)
    acc = np.zeros(3, np.float64)
    assert oracle.lib().oracle_ntc_cell(o._o, cell, step, oracle.fp(vx), oracle.fp(vy), oracle.fp(vz), len(vx),
                                        float(vol), oracle.fp(nt), oracle.dp(acc)) == 0
    return np.stack([vx, vy, vz]), nt, acc


def test_ntc_rounds_conserve_momentum_and_energy():
    rng = np.random.default_rng(3)
    n, vol, F, dt = 40, 1.0e-6, 1.0e12, 1.0e-6
    v = rng.normal(0.0, 300.0, (3, n)).astype(np.float32)
    # max chosen so that M = 200 candidates: 10 rounds of 20 pairs; the
    # cross-section makes a typical acceptance about one half
    mx = 200.0 * vol / (0.5 * n * (n - 1) * F * dt)
    o = _sim_physics(vhs_exp=1.0, vhs_k=mx / 900.0, fnum=F, dt=dt)
    v1, nt, acc = ntc_cell(o, 5, 1, v, vol, np.array([mx, 0.25], np.float32))
    assert acc[0] == 200 and acc[1] > 20
    p0, p1 = v.astype(np.float64).sum(1), v1.astype(np.float64).sum(1)
    e0, e1 = (v.astype(np.float64) ** 2).sum(), (v1.astype(np.float64) ** 2).sum()
    assert np.allclose(p0, p1, atol=1e-2 * n)
    assert abs(e1 - e0) < 1e-5 * e0
    assert nt[0] >= np.float32(mx) and abs(nt[1] - 0.25) < 1e-6
    # the same cell, step and state give the same draws; another step others
    v2, _, acc2 = ntc_cell(o, 5, 1, v, vol, np.array([mx, 0.25], np.float32))
    assert np.array_equal(v1, v2) and np.array_equal(acc, acc2)
    v3, _, _ = ntc_cell(o, 5, 2, v, vol, np.array([mx, 0.25], np.float32))
    assert not np.array_equal(v1, v3)


def test_ntc_rounds_collision_count_matches_expectation():
    """One round per cell (M <= n/2): every candidate slot is a uniformly
    random pair of the loaded velocities, accepted with sigma c_r / max."""
    n, cells, vol, F, dt = 24, 4000, 1.0e-6, 1.0e12, 1.0e-6
    idx = np.arange(n, dtype=np.float32)
    # velocities grow with the index: the relative speed of a pair is about
    # 20 |i - j| m/s, so the pairing's index structure shows in the count
    v = np.zeros((3, n), np.float32)
    v[0] = 20.0 * idx
    v[1] = 3.0 * np.sin(idx)
    cr = np.sqrt(((v[:, :, None].astype(np.float64) - v[:, None, :]) ** 2).sum(0))
    M = 9  # candidates per cell: one round of at most 12 pairs
    mx = M * vol / (0.5 * n * (n - 1) * F * dt)
    k = mx / (1.05 * float(cr.max()))  # sigma c_r = k c_r < max for every pair
    o = _sim_physics(vhs_exp=1.0, vhs_k=k, fnum=F, dt=dt)
    total = 0.0
    for c in range(cells):
        _, _, acc = ntc_cell(o, c, 1, v, vol, np.array([mx, 0.5], np.float32))
        assert acc[0] == M
        total += acc[1]
    iu = np.triu_indices(n, 1)
    p = (np.float32(k) * cr[iu]) / np.float32(mx)  # acceptance of each pair
    mean_acc = float(p.mean())
    expect = cells * M * mean_acc
    sd = math.sqrt(cells * M * mean_acc * (1 - mean_acc))
    assert abs(total - expect) < 4 * sd, (total, expect, sd)
    # adjacent-index pairing would give a far smaller count
    adj = float(((np.float32(k) * cr[np.arange(n - 1), np.arange(1, n)]) / np.float32(mx)).mean())
    assert abs(cells * M * adj - expect) > 20 * sd
