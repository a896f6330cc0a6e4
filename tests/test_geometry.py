"""Host geometry: meshes, the triangle/box test and collision grids (CPU).

The SAT and the REFERENCE grid are checked against the reference's own
compiled code (verdicts and grids in tests/golden written by
make_golden.py from include/tritest.h, src/aabb.cpp, src/reggrid.cpp); the
generators against closed forms (radius, closedness, enclosed volume).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import cudadsmc_amd as dm
from cudadsmc_amd import _capi


def mesh_volume(m: dm.Mesh) -> float:
    v = m.vertices.astype(np.float64)[m.triangles]
    return float(np.einsum("ij,ij->i", v[:, 0], np.cross(v[:, 1], v[:, 2])).sum() / 6.0)


def closed_edges(m: dm.Mesh) -> bool:
    t = m.triangles
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    fwd = set(map(tuple, e))
    return all((b, a) in fwd for a, b in fwd) and len(fwd) == len(e)


@pytest.mark.parametrize("subdiv,nt", [(2, 320), (3, 1280), (5, 20480)])
def test_icosphere(subdiv, nt):
    c = np.array([0.1, 0.1, 0.1])
    m = dm.icosphere(c, 0.025, subdiv)
    assert len(m.triangles) == nt
    r = np.linalg.norm(m.vertices - c, axis=1)
    assert np.allclose(r, 0.025, atol=1e-7)
    assert closed_edges(m)
    cen = m.vertices[m.triangles].mean(axis=1) - c
    assert (np.einsum("ij,ij->i", cen, m.normals) > 0).all()
    assert np.allclose(np.linalg.norm(m.normals, axis=1), 1.0, atol=1e-6)
    vol = mesh_volume(m.translate(-c))
    exact = 4 / 3 * np.pi * 0.025**3
    assert 0.9 * exact < vol < exact


def test_sphere_cone():
    rn, rb, th = 0.02, 0.05, 30.0
    m = dm.sphere_cone(0.08, 0.1, 0.1, rn, th, rb, 200, 124)
    assert 45000 < len(m.triangles) < 55000
    assert closed_edges(m)
    v = m.vertices.astype(np.float64)
    rad = np.hypot(v[:, 1] - 0.1, v[:, 2] - 0.1)
    assert np.isclose(rad.max(), rb, rtol=1e-5)
    assert np.isclose(v[:, 0].min(), 0.08, atol=1e-7)
    # closed-form volume: spherical cap (to the tangency point) + frustum
    a = np.deg2rad(90 - th)
    xt, rt = rn - rn * np.cos(a), rn * np.sin(a)
    h_cap = xt
    cap = np.pi * h_cap**2 * (3 * rn - h_cap) / 3
    h_fr = (rb - rt) / np.tan(np.deg2rad(th))
    frustum = np.pi * h_fr / 3 * (rt**2 + rt * rb + rb**2)
    vol = mesh_volume(m.translate((-0.08, -0.1, -0.1)))
    assert np.isclose(vol, cap + frustum, rtol=2e-3)
    cen = m.vertices[m.triangles].mean(axis=1).astype(np.float64)
    inside = np.array([0.08 + (xt + xt + h_fr) / 2, 0.1, 0.1])
    # outward: normals point away from an interior point for this convex body
    assert (np.einsum("ij,ij->i", cen - inside, m.normals) > 0).mean() > 0.999


def test_load_raw_ares(golden):
    m = dm.load_raw(str(golden / "ares.raw"))
    assert m.vertices.shape == (4148, 3) and m.triangles.shape == (7744, 3)
    used = m.vertices[np.unique(m.triangles)]
    lo, hi = used.min(0), used.max(0)
    assert np.allclose(lo, [-0.435, -0.350, 0.202], atol=2e-3)
    assert np.allclose(hi, [0.263, 0.348, 9.862], atol=2e-3)
    with pytest.raises(dm.DsmcError):
        dm.load_raw(str(golden / "config.xml"))


def test_tri_box_matches_reference_tritest(golden):
    d = np.load(golden / "tribox.npz")
    L = _capi.lib()
    fp = lambda x: np.ascontiguousarray(x, np.float32).ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    got = np.array([L.dsmc_tri_box_overlap(fp(d["center"][k]), fp(d["half"][k]), fp(d["tri"][k][0]),
                                           fp(d["tri"][k][1]), fp(d["tri"][k][2]))
                    for k in range(len(d["verdict"]))], np.int32)
    bad = np.nonzero(got != d["verdict"])[0]
    assert len(bad) == 0, bad[:10]


def test_tri_box_closed_form():
    # a triangle in the plane z = 1 against the unit box: touching face counts
    L = _capi.lib()
    fp = lambda x: np.ascontiguousarray(x, np.float32).ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    c, h = [0, 0, 0], [1, 1, 1]
    for z, want in [(0.5, 1), (1.0, 1), (1.01, 0), (-1.0, 1), (-1.5, 0)]:
        t = [[-0.2, -0.2, z], [0.3, -0.1, z], [0.0, 0.4, z]]
        assert L.dsmc_tri_box_overlap(fp(c), fp(h), fp(t[0]), fp(t[1]), fp(t[2])) == want
    # a large triangle whose plane cuts a box corner but whose edges miss it
    t = [[3, 0, 0], [0, 3, 0], [0, 0, 3]]
    assert L.dsmc_tri_box_overlap(fp(c), fp(h), fp(t[0]), fp(t[1]), fp(t[2])) == 1
    t = [[6, 0, 0], [0, 6, 0], [0, 0, 6]]
    assert L.dsmc_tri_box_overlap(fp(c), fp(h), fp(t[0]), fp(t[1]), fp(t[2])) == 0


def test_reference_grid_ares(golden):
    g = np.load(golden / "grid_ares.npz")
    m = dm.load_raw(str(golden / "ares.raw"))
    # the reference builds the grid over the bbox stored in the file; our mesh
    # bbox from the referenced vertices must be that bbox
    used = m.vertices[np.unique(m.triangles)]
    lo, hi = used.min(0), used.max(0)
    assert np.array_equal(np.concatenate([lo, hi]).astype(np.float32), g["bbox"])
    got = dm.cgrid_build(dm.MODE_REFERENCE, m, (5, 5, 20))
    assert np.array_equal(got["bounds"][6:], g["cell"])
    counts = np.where(got["start"] == 0xFFFFFFFF, 0, got["end"] - got["start"])
    assert np.array_equal(counts, g["counts"])
    assert np.array_equal(got["entries"], g["entries"])
    assert np.array_equal((got["flags"] != 0).astype(np.uint32), g["flags"])


def test_reference_grid_icosphere(golden):
    g = np.load(golden / "grid_icosphere.npz")
    m = dm.icosphere((0.1, 0.1, 0.1), 0.025, 2)
    got = dm.cgrid_build(dm.MODE_REFERENCE, m, (6, 6, 6))
    counts = np.where(got["start"] == 0xFFFFFFFF, 0, got["end"] - got["start"])
    assert np.array_equal(counts, g["counts"])
    assert np.array_equal(got["entries"], g["entries"])
    assert np.array_equal((got["flags"] != 0).astype(np.uint32), g["flags"])


def test_physics_grid_interior_and_lists():
    c = np.array([0.1, 0.1, 0.1])
    m = dm.icosphere(c, 0.025, 3)
    g = dm.cgrid_build(dm.MODE_PHYSICS, m, (12, 12, 12))
    d = g["dims"]
    assert list(d) == [14, 14, 14]
    flags = g["flags"].reshape(d[2], d[1], d[0])
    # the outer layer holds no interior cells (the conservative SAT may list the
    # extreme triangles there), the centre is filled interior
    for face in (flags[0], flags[-1], flags[:, 0], flags[:, -1], flags[:, :, 0], flags[:, :, -1]):
        assert (face != 2).all()
    assert flags[d[2] // 2, d[1] // 2, d[0] // 2] == 2
    # every triangle is listed in the cell that holds its centroid
    bmin, cell = g["bounds"][:3], g["bounds"][6:]
    cen = m.vertices[m.triangles].mean(axis=1)
    idx = np.floor((cen - bmin) / cell).astype(int)
    lin = idx[:, 2] * d[0] * d[1] + idx[:, 1] * d[0] + idx[:, 0]
    for t in range(len(cen)):
        s, e = g["start"][lin[t]], g["end"][lin[t]]
        assert t in g["entries"][s:e]
