"""Statistical GPU tests of the physics against closed forms.

* Equilibrium collision rate in a closed box: collisions per step =
  N nu dt / 2 with the VHS rate nu = 4 d_ref^2 n sqrt(pi k T_ref / m)
  (T/T_ref)^(1-omega) (Bird 1994, eq. 4.64; hard spheres at omega = 1/2).
* Free-molecular sphere drag with diffuse re-emission at T_w (Bird 1994,
  eq. 7.52; Schaaf & Chambre):
    C_D = (2s^2+1)/(sqrt(pi) s^3) exp(-s^2) + (4s^4+4s^2-1)/(2s^4) erf(s)
          + 2 sqrt(pi)/(3s) sqrt(T_w/T_inf),   s = U / sqrt(2 k T_inf / m)
* Open freestream faces with inflow injection keep the sampled n, u, T of an
  empty domain at their freestream values and the particle count steady.
* Determinism: two runs with the same seeds give identical bits.
Tolerances are several standard errors of the sampled quantity.
"""
from __future__ import annotations

import math
import sys

import numpy as np
import pytest

import cudadsmc_amd as dm

from helpers import small_physics_settings, small_reference_settings

pytestmark = pytest.mark.gpu

K = 1.380649e-23


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def vhs_rate(n, T, m, d, Tref, omega):
    return 4 * d * d * n * math.sqrt(math.pi * K * Tref / m) * (T / Tref) ** (1 - omega)


@pytest.mark.parametrize("omega", [0.81, 0.5])
def test_equilibrium_collision_rate(omega):
    note(f"equilibrium collision rate omega={omega}")
    s = small_physics_settings(dims=(24, 24, 24), n_target=300000, omega=omega, U=0.0, density=1.0e21,
                               faces="specular", wall_temp=0.0)
    s.bmax[:] = (0.05, 0.05, 0.05)
    for a in range(3):
        s.cell_size[a] = (s.bmax[a] - s.bmin[a]) / s.grid_dim[a]
    s.dt = 2.0e-7
    sol = dm.Solver()
    sol.Init(s)
    n = sol.InitSystemState((0.0, 0.0, 0.0))
    sol.DoBenchmark(20, s.dt)  # let (sigma c_r)max settle
    c0 = sol.collision_stats()
    steps = 200
    sol.DoBenchmark(steps, s.dt)
    c1 = sol.collision_stats()
    coll = (c1[1] - c0[1]) / steps
    nu = vhs_rate(1.0e21, 200.0, s.molmass, s.diameter, s.Tref, omega)
    want = 0.5 * n * nu * s.dt
    note(f"collisions/step {coll:.1f} expected {want:.1f} ratio {coll / want:.4f}")
    assert abs(coll / want - 1) < 0.02
    # mean relative speed of colliding pairs: <c_r sigma c_r>/<sigma c_r> for VHS
    # = <c_r^(3-2w)>/<c_r^(2-2w)> of the Maxwellian relative-speed distribution
    mean_cr = (c1[2] - c0[2]) / (c1[1] - c0[1])
    cm = math.sqrt(2 * K * 200.0 / (s.molmass / 2))  # most probable relative speed
    a = 3 - 2 * omega
    want_cr = cm * math.gamma((a + 3) / 2) / math.gamma((a + 2) / 2)
    assert abs(mean_cr / want_cr - 1) < 0.01


def test_free_molecular_sphere_drag():
    note("free-molecular sphere drag")
    U, Tinf, Tw, r = 1500.0, 200.0, 300.0, 0.02
    s = small_physics_settings(dims=(40, 20, 20), n_target=600000, omega=0.81, wall_temp=Tw, U=U, density=1.0e17)
    s.dt = 1.0e-6
    sol = dm.Solver()
    sol.Init(s)
    mesh = dm.icosphere((0.08, 0.05, 0.05), r, 4)
    sol.SetModel(mesh, grid=(10, 10, 10))
    sol.InitSystemState((U, 0.0, 0.0))
    sol.DoBenchmark(300, s.dt)
    sol.set_sampling(True)
    sol.DoBenchmark(700, s.dt)
    F = sol.drag()
    m = s.molmass
    sr = U / math.sqrt(2 * K * Tinf / m)
    cd_fm = ((2 * sr**2 + 1) / (math.sqrt(math.pi) * sr**3) * math.exp(-sr**2)
             + (4 * sr**4 + 4 * sr**2 - 1) / (2 * sr**4) * math.erf(sr)
             + 2 * math.sqrt(math.pi) / (3 * sr) * math.sqrt(Tw / Tinf))
    # projected area of the tessellated sphere (its silhouette is slightly smaller)
    v = mesh.vertices.astype(np.float64) - np.array([0.08, 0.05, 0.05])
    area = math.pi * np.max(np.hypot(v[:, 1], v[:, 2])) ** 2
    cd =F[0] / (0.5 * 1.0e17 * m * U * U * area)
    note(f"C_D {cd:.4f} free-molecular {cd_fm:.4f}")
    assert abs(cd / cd_fm - 1) < 0.04
    assert abs(F[1]) < 0.05 * abs(F[0]) and abs(F[2]) < 0.05 * abs(F[0])
    # no particle got through the wall: none inside the inscribed sphere of the mesh
    p = sol.particles()["pos"].astype(np.float64)
    rr = np.sqrt(((p - np.array([[0.08], [0.05], [0.05]])) ** 2).sum(0))
    assert int((rr < 0.99 * r).sum()) == 0


def test_freestream_preserved_by_open_faces():
    note("freestream preservation")
    s = small_physics_settings(dims=(32, 16, 16), n_target=400000, omega=0.81, wall_temp=300.0, U=1500.0,
                               density=1.0e20)
    sol = dm.Solver()
    sol.Init(s)
    n0 = sol.InitSystemState((1500.0, 0.0, 0.0))
    sol.DoBenchmark(200, s.dt)
    sol.set_sampling(True)
    sol.DoBenchmark(300, s.dt)
    f = sol.fields()
    _, n1, _ = sol.counts()
    assert abs(n1 / n0 - 1) < 0.02
    # interior cells (away from the faces' first cell)
    g = np.s_[1:-1, 1:-1, 1:-1]
    dims = tuple(s.grid_dim)
    shape = (dims[2], dims[1], dims[0])
    dens = f["n"].reshape(shape)[g].mean()
    u = f["u"].reshape(shape)[g].mean()
    v = f["v"].reshape(shape)[g].mean()
    T = f["T"].reshape(shape)[g].mean()
    note(f"n {dens:.4e} u {u:.2f} v {v:.2f} T {T:.2f}")
    assert abs(dens / 1.0e20 - 1) < 0.01
    assert abs(u - 1500.0) < 3.0 and abs(v) < 3.0
    assert abs(T / 200.0 - 1) < 0.01


def test_reference_mode_equilibrium_rate():
    note("reference mode hard-sphere rate")
    s = small_reference_settings(periodic="x", dims=(24, 24, 24), ppc=20, geometry=False)
    s.stream_velocity[:] = (0.0, 0.0, 0.0)
    s.bmax[:] = (0.05, 0.05, 0.05)
    for a in range(3):
        s.cell_size[a] = (s.bmax[a] - s.bmin[a]) / s.grid_dim[a]
    vol = s.cell_size[0] * s.cell_size[1] * s.cell_size[2]
    s.density = 1.0e21
    s.fnum = np.float32(s.density * vol / 20 * 0.999)
    s.dt = 2.0e-7
    s.part_per_cell = 0
    s.num_particles = 0
    sol = dm.Solver()
    sol.Init(s)
    n = sol.InitSystemState((0.0, 0.0, 0.0))
    sol.DoBenchmark(200, s.dt)  # relax the loading to Poisson cells and a Maxwellian
    c0 = sol.collision_stats()
    steps = 200
    sol.DoBenchmark(steps, s.dt)
    c1 = sol.collision_stats()
    coll = (c1[1] - c0[1]) / steps
    # hard spheres, cross-section pi d^2; the reference's Boltzmann constant
    k = 1.3806e-23
    T = s.temp
    nu = 4 * s.diameter**2 * (n * s.fnum / (0.05**3)) * math.sqrt(math.pi * k * T / s.molmass)
    # The reference draws both partners with replacement (i == j possible,
    # never accepted) and scales the candidate count by the global particles
    # per cell, so a cell of N_c particles collides at (N_c - 1) / N_c of the
    # kinetic rate: sum over cells of (N_c - 1) replaces N.
    cnt = np.diff(sol.cell_start().astype(np.int64))
    eff = float(np.sum(np.maximum(cnt - 1, 0))) / float(cnt.sum())
    want = 0.5 * n * nu * s.dt * eff
    note(f"collisions/step {coll:.1f} expected {want:.1f} (self-pair factor {eff:.4f}) ratio {coll / want:.4f}")
    assert abs(coll / want - 1) < 0.02


def test_deterministic_replay():
    note("deterministic replay")
    runs = []
    for _ in range(2):
        s = small_physics_settings(dims=(24, 12, 12), n_target=50000)
        sol = dm.Solver()
        sol.Init(s)
        sol.SetModel(dm.icosphere((0.07, 0.05, 0.05), 0.02, 2), grid=(8, 8, 8))
        sol.InitSystemState((1500.0, 0.0, 0.0))
        sol.DoBenchmark(50, s.dt)
        runs.append(sol.particles())
        sol.close()
    for k in ("pos", "vel", "key"):
        assert np.array_equal(runs[0][k], runs[1][k])


def test_near_body_codes():
    """Near-body codes against a brute-force distance from every cell box to
    the occupied collision-grid cells: a code never overstates the distance
    (a particle that could reach the body is always deferred), 255 only
    where nothing occupied lies within one cell, and the codes are the
    rounded-down distances."""
    note("near-body codes")
    s = small_physics_settings(dims=(24, 16, 16), n_target=20000, omega=0.81, wall_temp=300.0, U=1500.0,
                               density=1.0e20)
    sol = dm.Solver()
    sol.Init(s)
    mesh = dm.icosphere((0.06, 0.05, 0.05), 0.018, 2)
    sol.SetModel(mesh, grid=(13, 13, 13))
    code = sol.near_cells().astype(np.int64)
    g = sol.collision_grid()
    b, dims = g["bounds"].astype(np.float64), g["dims"].astype(np.int64)
    gmin, gcell = b[0:3], b[6:9]
    X, Y, Z = (int(d) for d in s.grid_dim)
    cell = np.array([s.cell_size[0], s.cell_size[1], s.cell_size[2]], np.float64)
    unit = float(np.float32(np.float32(min(s.cell_size)) / np.float32(64.0)))
    occ = np.nonzero(g["flags"] != 0)[0]
    ox, oy, oz = occ % dims[0], (occ // dims[0]) % dims[1], occ // (dims[0] * dims[1])
    q0 = np.stack([gmin[0] + ox * gcell[0], gmin[1] + oy * gcell[1], gmin[2] + oz * gcell[2]], 1)
    q1 = q0 + gcell
    cz, cy, cx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    b0 = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], 1) * cell + np.array(s.bmin[:3], np.float64)
    b1 = b0 + cell
    gap = np.maximum(0.0, np.maximum(q0[None] - b1[:, None], b0[:, None] - q1[None]))  # (cells, occ, 3)
    d = np.sqrt((gap ** 2).sum(-1))
    within = (gap <= cell * (1 + 1e-6)).all(-1)  # inside the cell's box widened by one cell
    d_all = d.min(1)
    d_win = np.where(within, d, np.inf).min(1)
    near = code != 255
    assert near.sum() > 0
    # "far" only where nothing occupied lies within the widened box (a particle
    # would have to move a cell or more along some axis, which always defers it)
    assert np.all(~np.isfinite(d_win[~near]))
    assert np.all(np.isfinite(d_win[near]) | (d_all[near] > 0))
    # never more than the distance to an occupied cell of the widened box, so a
    # particle whose flight reaches one is deferred; never less than the
    # rounded-down distance to the nearest occupied cell overall
    fin = near & np.isfinite(d_win)
    assert np.all(code[fin] * unit <= d_win[fin] * (1 + 1e-6) + 1e-12)
    assert np.all(code[near] >= np.minimum(254, np.floor(d_all[near] / unit * (1 - 1e-6))))
