"""Known-answer pins of the oracle's REFERENCE-mode functions (CPU only).

Each test states the expected result from the reference's own definition,
independently of the oracle's code, and chooses its cases so that a
plausible misreading (wrong axis, a wrap taken for a clamp, a missing
velocity flip, nearest hit instead of first-in-list, a wrong stream or draw
order) fails:

* cell keys (GetCellID3D, kernels/dsmc_kernel.cu:1701-1710): numpy brute
  force with FP32 IEEE division, saturating float->uint conversion and the
  clamp to the last cell; for interior points, the cell's box must contain
  the point;
* box faces (PeriodicX/Y/ZBoundaryCondition, NoPeriodicBoundaryCondition,
  dsmc_kernel.cu:1316-1555): hand-derived positions after one step for
  particles crossing each face, corners included;
* triangle walls (CollideWithGeometry / ProcessCollisionsDev,
  dsmc_kernel.cu:1256-1314, 1596-1657): axis rays into two parallel walls,
  where the first triangle in the cell's list wins even when it is farther;
* loading (GenerateInitialVelocitiesDevBlocks, dsmc_kernel.cu:629-680,
  GetRandomVelocity/GetRandomPosition :400-427): particle positions and
  velocities from the golden genrand_mt outputs of the reference's own
  MersenneTwister.dat rows and seed chain (stream X*Y*i + y*X + x, seven
  draws per particle).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle

BOUNCE = 0.001  # BOUNCE_DELTA, dsmc_kernel.cu:53 (metres, absolute)


def _settings(periodic="n", dims=(16, 12, 10), bmax=(0.16, 0.12, 0.10), ppc=4, dt=2.0e-6):
    import cudadsmc_amd as dm

    s = dm.settings_default()
    s.mode = dm.MODE_REFERENCE
    s.bmin[:] = (0.0, 0.0, 0.0)
    s.bmax[:] = bmax
    s.grid_dim[:] = dims
    for a in range(3):
        s.cell_size[a] = np.float32(np.float32(bmax[a]) / np.float32(dims[a]))
    s.periodic_condition = periodic.encode()
    s.temp = 300.0
    s.density = 1.0e20
    s.fnum = 1.0
    s.dt = dt
    s.part_per_cell = ppc
    s.num_particles = ppc * int(np.prod(dims))
    s.geometry = 0
    s.mt = 1
    return s


def _sim(s, n_streams=256, rows=None):
    if rows is None:
        rows = np.zeros((n_streams, 3), np.uint32)
    layout = {"tile_cells": 32, "threads": 256, "n_streams": n_streams}
    return oracle.OracleSim(s, layout, (rows[:, 0], rows[:, 1], rows[:, 2]))


def _f2u_sat(q: np.ndarray) -> np.ndarray:
    """Saturating float32 -> uint32 conversion (round toward zero)."""
    q = np.asarray(q, np.float64)
    out = np.zeros(q.shape, np.uint64)
    big = q >= 4294967296.0
    ok = np.isfinite(q) & (q > 0) & ~big
    out[ok] = np.floor(q[ok]).astype(np.uint64)
    out[big] = 0xFFFFFFFF
    return out


def _keys(s, x, y, z) -> np.ndarray:
    dims = [int(d) for d in s.grid_dim]
    ids = []
    for a, p in enumerate((x, y, z)):
        mn = np.float32(s.bmin[a])
        cell = np.float32(np.float32(np.float32(s.bmax[a]) - mn) / np.float32(dims[a]))
        with np.errstate(invalid="ignore", over="ignore"):
            q = (np.asarray(p, np.float32) - mn) / cell  # FP32 IEEE ops
        ids.append(np.minimum(_f2u_sat(q), dims[a] - 1))
    return (ids[2] * dims[0] * dims[1] + ids[1] * dims[0] + ids[0]).astype(np.uint32)


# ------------------------------------------------------------------ cell keys
def test_cell_keys_brute_force():
    s = _settings(dims=(7, 5, 3), bmax=(0.7, 0.31, 0.123))
    rng = np.random.default_rng(7)
    n = 20000
    x = rng.uniform(-0.1, 0.8, n).astype(np.float32)
    y = rng.uniform(-0.1, 0.4, n).astype(np.float32)
    z = rng.uniform(-0.05, 0.2, n).astype(np.float32)
    special = np.array([0.0, -0.0, 0.7, 0.31, 0.123, 1e-30, -1e-30, np.nan, np.inf, -np.inf, 1e30, -1e30, 0.1, 0.2,
                        0.3, 0.6999999], np.float32)
    sx, sy, sz = np.meshgrid(special, special, special, indexing="ij")
    x = np.concatenate([x, sx.ravel()])
    y = np.concatenate([y, sy.ravel()])
    z = np.concatenate([z, sz.ravel()])
    got = oracle.cell_ids(s, x, y, z)
    assert np.array_equal(got, _keys(s, x, y, z))
    # NaN and negative coordinates saturate to cell 0 on that axis, +inf and
    # values past the box to the last cell
    k = oracle.cell_ids(s, np.array([np.nan, -5.0, np.inf, 0.75], np.float32), np.zeros(4, np.float32),
                        np.zeros(4, np.float32))
    assert list(k) == [0, 0, 6, 6]
    k = oracle.cell_ids(s, np.zeros(3, np.float32), np.array([np.nan, 0.4, 0.0], np.float32),
                        np.array([0.0, 0.0, 1.0], np.float32))
    assert list(k) == [0, 4 * 7, 2 * 7 * 5]


def test_cell_keys_interior_points_lie_in_their_cell():
    s = _settings(dims=(7, 5, 3), bmax=(0.7, 0.31, 0.123))
    rng = np.random.default_rng(11)
    n = 5000
    p = np.stack([rng.uniform(0.0, b, n) for b in (0.7, 0.31, 0.123)]).astype(np.float32)
    k = oracle.cell_ids(s, *p).astype(np.int64)
    ix, iy, iz = k % 7, (k // 7) % 5, k // 35
    for a, (i, b, d) in enumerate(((ix, 0.7, 7), (iy, 0.31, 5), (iz, 0.123, 3))):
        h = b / d
        lo, hi = i * h, (i + 1) * h
        # FP32 division may round a point within an ulp of a cell boundary
        # into the neighbour; everything else lies inside its cell's box
        assert np.all(p[a] >= lo - 1e-6) and np.all(p[a] < hi + 1e-6)


# ------------------------------------------------------------------ box faces
def _one_step(s, pos, vel):
    o = _sim(s)
    o.set_state(np.asarray(pos, np.float32), np.asarray(vel, np.float32))
    keys0 = _keys(s, *np.asarray(pos, np.float32))
    order = np.argsort(keys0, kind="stable")
    o.step(sample=False)
    st = o.state()
    return order, st


def _after(x0, v, dt=2.0e-6):
    """Free flight in FP32: x + v*dt."""
    return np.float32(np.float32(x0) + np.float32(np.float32(v) * np.float32(dt)))


@pytest.mark.parametrize("periodic", ["x", "y", "z", "n"])
def test_reference_box_faces(periodic):
    s = _settings(periodic=periodic)
    lo = (0.0, 0.0, 0.0)
    hi = (0.16, 0.12, 0.10)
    inside = (0.08, 0.06, 0.05)
    cases = []  # (pos, vel)
    for a in range(3):
        for side in (+1, -1):
            p = list(inside)
            v = [10.0, -20.0, 30.0]
            p[a] = hi[a] - 0.0001 if side > 0 else lo[a] + 0.0001
            v[a] = 200.0 * side  # 0.4 mm per step: crosses the face
            cases.append((p, v))
    # a corner: crosses the high x and high y faces in the same step
    cases.append(([0.1599, 0.1199, 0.05], [200.0, 200.0, 0.0]))
    pos = np.array([c[0] for c in cases], np.float32).T.copy()
    vel = np.array([c[1] for c in cases], np.float32).T.copy()
    order, st = _one_step(s, pos, vel)
    wrap_axis = {"x": 0, "y": 1, "z": 2, "n": -1}[periodic]
    for j, i in enumerate(order):
        p0, v0 = pos[:, i], vel[:, i]
        for a in range(3):
            x1 = _after(p0[a], v0[a])
            # the bounds are FP32 settings; the reference evaluates the face
            # expressions in double (BOUNCE_DELTA is a double literal)
            lo_a, hi_a = np.float64(np.float32(lo[a])), np.float64(np.float32(hi[a]))
            if x1 > np.float32(hi[a]):
                if a == wrap_axis:  # re-enters at the low face, 1 mm in, same velocity
                    want_x = np.float32((lo_a + BOUNCE + np.float64(x1)) - hi_a)
                    want_v = v0[a]
                    assert abs(want_x - (lo[a] + BOUNCE + 0.0003)) < 1e-7
                else:  # clamped 1 mm inside the high face, normal velocity flipped
                    want_x = np.float32(hi_a - BOUNCE)
                    want_v = -v0[a]
            elif x1 < np.float32(lo[a]):
                if a == wrap_axis:
                    want_x = np.float32((hi_a - BOUNCE + np.float64(x1)) - lo_a)
                    want_v = v0[a]
                    assert abs(want_x - (hi[a] - BOUNCE - 0.0003)) < 1e-7
                else:
                    want_x = np.float32(lo_a + BOUNCE)
                    want_v = -v0[a]
            else:
                want_x, want_v = x1, v0[a]
            assert st["pos"][a, j] == want_x, (periodic, i, a)
            assert st["vel"][a, j] == np.float32(want_v), (periodic, i, a)
    assert np.array_equal(st["key"], _keys(s, *st["pos"]))


# ---------------------------------------------------------------- walls
def _two_walls(far_first: bool):
    """Two square walls normal to x at x = 0.10 and x = 0.12 (two triangles
    each), normals facing -x (towards particles arriving from low x)."""
    def square(x):
        v = np.array([[x, 0.02, 0.02], [x, 0.02, 0.08], [x, 0.10, 0.08], [x, 0.10, 0.02]], np.float32)
        return v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)

    vn, tn = square(0.10)
    vf, tf = square(0.12)
    parts = [(vf, tf), (vn, tn)] if far_first else [(vn, tn), (vf, tf)]
    verts = np.concatenate([parts[0][0], parts[1][0]])
    tris = np.concatenate([parts[0][1], parts[1][1] + 4])
    normals = np.tile(np.array([[-1.0, 0.0, 0.0]], np.float32), (4, 1))
    # one collision-grid cell over the walls, listing the triangles in order
    grid = {"bounds": np.array([0.05, 0.0, 0.0, 0.15, 0.12, 0.10, 0.10, 0.12, 0.10], np.float32),
            "dims": np.array([1, 1, 1], np.uint32), "flags": np.array([1], np.uint32),
            "start": np.array([0], np.uint32), "end": np.array([4], np.uint32),
            "entries": np.arange(4, dtype=np.uint32)}

    class M:
        pass

    m = M()
    m.vertices, m.triangles, m.normals = verts, tris, normals
    return m, grid


@pytest.mark.parametrize("far_first", [True, False])
def test_reference_first_hit_in_list_order(far_first):
    s = _settings(periodic="n", dt=1.0e-4)
    s.geometry = 1
    m, grid = _two_walls(far_first)
    o = _sim(s)
    o.set_geometry(m, grid)
    # x = 0.09 moving +x at 350 m/s: 35 mm in one step, through both walls
    pos = np.array([[0.09], [0.05], [0.05]], np.float32)
    vel = np.array([[350.0], [0.0], [0.0]], np.float32)
    o.set_state(pos, vel)
    o.step(sample=False)
    st = o.state()
    x1, v1 = float(st["pos"][0, 0]), st["vel"][:, 0]
    # reflected specularly: x velocity flips, |v| unchanged
    assert v1[0] == np.float32(-350.0) and v1[1] == 0.0 and v1[2] == 0.0
    if far_first:
        # the far wall is first in the cell's list: hit at 0.12 after 30 mm,
        # the last 5 mm go back towards -x (the near wall is behind the
        # re-emission point 0.12 - 1e-4 and 20 mm away, beyond the 5 mm left)
        assert abs(x1 - (0.12 - 0.005)) < 2e-6
    else:
        # the near wall is first: hit at 0.10 after 10 mm, 25 mm back
        assert abs(x1 - (0.10 - 0.025)) < 2e-6
    assert int(o.hits(1)[0]) >= 1
    # the re-emission point is the barycentric hit point (FP32 rounding)
    assert abs(st["pos"][1, 0] - 0.05) < 1e-7 and abs(st["pos"][2, 0] - 0.05) < 1e-7


def test_reference_wall_centroid_ray_and_miss():
    s = _settings(periodic="n", dt=1.0e-4)
    s.geometry = 1
    m, grid = _two_walls(False)
    o = _sim(s)
    o.set_geometry(m, grid)
    # particle 0: through the centroid of triangle 0 of the near wall
    # (x = 0.1, y = (0.02 + 0.02 + 0.10) / 3, z = (0.02 + 0.08 + 0.08) / 3)
    cy, cz = np.float32((0.02 + 0.02 + 0.10) / 3.0), np.float32((0.02 + 0.08 + 0.08) / 3.0)
    # particle 1: passes beside the walls (y = 0.11 > 0.10): no hit
    # particle 2: stops 1 mm short of the near wall: no hit
    pos = np.array([[0.095, 0.095, 0.089], [cy, 0.11, 0.05], [cz, 0.05, 0.05]], np.float32)
    vel = np.array([[100.0, 100.0, 100.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]], np.float32)
    keys0 = _keys(s, *pos)
    order = np.argsort(keys0, kind="stable")
    o.set_state(pos, vel)
    o.step(sample=False)
    st = o.state()
    hits = o.hits(3)
    res = {int(i): (st["pos"][:, j], st["vel"][:, j], int(hits[j])) for j, i in enumerate(order)}
    p, v, h = res[0]
    assert h >= 1 and v[0] == np.float32(-100.0)
    assert abs(p[0] - (0.10 - 0.005)) < 2e-6  # 5 mm to the wall, 5 mm back
    assert abs(p[1] - cy) < 1e-6 and abs(p[2] - cz) < 1e-6
    for i, x_end in ((1, 0.105), (2, 0.099)):
        p, v, h = res[i]
        assert h == 0 and v[0] == np.float32(100.0)
        assert abs(p[0] - x_end) < 2e-6


# ---------------------------------------------------------------- loading
def test_reference_loading_from_golden_streams(golden):
    d = np.load(golden / "mt_dcmt_rows.npz")
    rows, seeds, outs = d["rows"], d["seeds"], d["outputs"].astype(np.uint64)
    X, Y, Z, P = 4, 2, 3, 8  # X*Y*P = 64 streams, 7*Z = 21 draws each
    s = _settings(dims=(X, Y, Z), bmax=(0.04, 0.03, 0.06), ppc=P)
    s.stream_velocity[:] = (400.0, -50.0, 25.0)
    S = 256
    allrows = np.zeros((S, 3), np.uint32)
    allrows[:64] = rows[:, :3]
    o = _sim(s, S, allrows)
    win = np.zeros((S, 20), np.uint32)
    for t in range(64):
        win[t] = oracle.mt_window_after(int(rows[t, 0]), int(rows[t, 1]), int(rows[t, 2]), int(seeds[t]), 0)
    o.set_rng(win)
    o.load(tuple(s.stream_velocity))
    st = o.state()
    n = X * Y * Z * P
    assert st["pos"].shape[1] == n
    # expected from the reference's definitions and the golden draws
    R = ((outs.astype(np.float32) + np.float32(1.0)) / np.float32(4294967296.0)).astype(np.float32)  # (0, 1]
    m, k, T = np.float32(s.molmass), np.float32(1.3806e-23), np.float32(300.0)
    cell = [np.float32(np.float32(b) / np.float32(g)) for b, g in zip((0.04, 0.03, 0.06), (X, Y, Z))]
    twopi = np.float32(2.0) * np.float32(3.14159265358979)
    U = [np.float32(u) for u in s.stream_velocity]
    want_p = np.zeros((3, n), np.float32)
    want_v = np.zeros((3, n), np.float32)
    for i in range(P):
        for cy in range(Y):
            for cx in range(X):
                t = X * Y * i + cy * X + cx  # the loading thread's stream
                for cz in range(Z):
                    r = R[t, 7 * cz: 7 * cz + 7]  # R0, Rx, Ry, Rz, px, py, pz
                    idx = (X * Y * cz + cy * X + cx) * P + i
                    vv = np.sqrt(np.float32(2) * k * T / m * -np.log(r[0]))
                    for a in range(3):
                        want_v[a, idx] = vv * np.sin(r[1 + a] * twopi) + U[a]
                    base = (cx * cell[0], cy * cell[1], cz * cell[2])
                    for a in range(3):
                        want_p[a, idx] = np.float32(base[a]) + r[4 + a] * cell[a]
    # positions need no transcendental: bit-exact; velocities go through
    # log/sin (libm here, the shared series in the oracle): a few ulp
    assert np.array_equal(st["pos"], want_p)
    np.testing.assert_allclose(st["vel"], want_v, rtol=2e-6, atol=2e-3)
    # a planted mistake: the streams of a cell-major mapping (t = cell * P + i)
    # give different positions
    wrong = np.zeros(n, np.float32)
    for c in range(X * Y):
        for i in range(P):
            wrong[c * P + i] = R[c * P + i, 4] * cell[0] + np.float32((c % X) * cell[0])
    assert not np.array_equal(wrong, want_p[0, : X * Y * P])
