"""The measured workloads: argon sphere and sphere-cone flows at three sizes.

All three are argon VHS (m = 6.63e-26 kg, d_ref = 4.17e-10 m, omega = 0.81,
T_ref = 273 K), freestream T = 200 K along +x, a 0.4 x 0.2 x 0.2 m box with
freestream faces, a diffuse 300 K body at (0.1, 0.1, 0.1), PHYSICS mode:

  c1  64 x 32 x 32 cells, 100 000 particles, icosphere r = 0.025 m
      (1280 triangles), n = 1e20 m^-3, U = 1500 m/s, dt = 1e-6 s
  c2  128 x 64 x 64 cells, 2^20 particles, icosphere r = 0.025 m
      (20 480 triangles), n = 1e20 m^-3, U = 1500 m/s, dt = 5e-7 s
  c3  256 x 128 x 128 cells, 2^25 particles, closed sphere-cone (nose
      radius 0.02 m, 30 deg half-angle, base radius 0.05 m, 48 800
      triangles), n = 1e21 m^-3, U = 3000 m/s, dt = 2e-7 s
"""
from __future__ import annotations

import numpy as np

WORKLOADS = {
    # name: (grid, target particles, n, U, dt, body)
    "c3": dict(grid=(256, 128, 128), n_part=2**25, density=1.0e21, U=3000.0, dt=2.0e-7, body="spherecone"),
    "c2": dict(grid=(128, 64, 64), n_part=2**20, density=1.0e20, U=1500.0, dt=5.0e-7, body="ico5"),
    "c1": dict(grid=(64, 32, 32), n_part=100000, density=1.0e20, U=1500.0, dt=1.0e-6, body="ico3"),
}


def make_case(name: str, n_part: int | None = None, cg_refine: int | None = None):
    """Settings, body mesh and collision-grid dimensions of a workload. The
    collision grid covers the body's bounding box with cells `cg_refine`
    times finer than the flow cells (default: the tuned value, else 1)."""
    from . import FACE_FREESTREAM, MODE_PHYSICS, icosphere, settings_default, sphere_cone
    from .ensemble import tuned

    w = WORKLOADS[name]
    s = settings_default()
    s.mode = MODE_PHYSICS
    s.bmin[:] = (0.0, 0.0, 0.0)
    s.bmax[:] = (0.4, 0.2, 0.2)
    s.grid_dim[:] = w["grid"]
    for a in range(3):
        s.cell_size[a] = (s.bmax[a] - s.bmin[a]) / s.grid_dim[a]
    s.molmass, s.diameter, s.omega, s.Tref = 6.63e-26, 4.17e-10, 0.81, 273.0
    s.temp, s.density, s.wall_temp = 200.0, w["density"], 300.0
    s.stream_velocity[:] = (w["U"], 0.0, 0.0)
    s.face_bc[:] = [FACE_FREESTREAM] * 6
    s.target_particles = n_part or w["n_part"]
    s.dt = w["dt"]
    if w["body"] == "spherecone":
        mesh = sphere_cone(0.1, 0.1, 0.1, 0.02, 30.0, 0.05, 200, 124)
    elif w["body"] == "ico5":
        mesh = icosphere((0.1, 0.1, 0.1), 0.025, 5)
    else:
        mesh = icosphere((0.1, 0.1, 0.1), 0.025, 3)
    if cg_refine is None:
        cg_refine = int(tuned(name, 1, "cg_refine"))
    lo, hi = mesh.vertices.min(0), mesh.vertices.max(0)
    cgrid = tuple(int(max(1, np.ceil(cg_refine * (hi[k] - lo[k]) / s.cell_size[k]))) for k in range(3))
    return s, mesh, cgrid


def make_solver(name: str, layout_from_tuning: bool = True):
    """A loaded solver of a workload with the tuned launch layout (RNG
    streams, cells per tile, scatter grid) unless told otherwise."""
    from . import Solver
    from .ensemble import tuned

    s, mesh, cgrid = make_case(name)
    if layout_from_tuning:
        s.n_streams = int(tuned(name, 0))
    sol = Solver()
    sol.Init(s)
    if layout_from_tuning:
        tc = int(tuned(name, 0, "tile_cells"))
        if tc:
            sol.set_tile_cells(tc)
        sg = int(tuned(name, 0, "scatter_grid"))
        if sg:
            sol.set_scatter_grid(sg)
    sol.SetModel(mesh, grid=cgrid)
    return sol, s, mesh
