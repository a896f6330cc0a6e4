"""Helpers shared by the tests (settings factories, GPU detection)."""
from __future__ import annotations

import ctypes as C

import numpy as np


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:  # noqa: BLE001
        return False


def small_reference_settings(periodic: str = "x", dims=(16, 12, 10), ppc: int = 6, geometry: bool = True):
    """A config.xml-like REFERENCE case sized for oracle comparisons."""
    import cudadsmc_amd as dm

    s = dm.settings_default()
    s.bmin[:] = (0.0, 0.0, 0.0)
    s.bmax[:] = (0.16, 0.12, 0.10)
    s.grid_dim[:] = dims
    for a in range(3):
        s.cell_size[a] = (s.bmax[a] - s.bmin[a]) / s.grid_dim[a]
    s.periodic_condition = periodic.encode()
    s.stream_velocity[:] = (400.0, 0.0, 0.0)
    s.temp = 300.0
    s.density = 1.0e20
    vol = s.cell_size[0] * s.cell_size[1] * s.cell_size[2]
    s.fnum = np.float32(s.density * vol / ppc * 0.999)
    s.dt = 2.0e-6
    s.geometry = int(geometry)
    s.mt = 1
    dm.check(dm.lib().dsmc_settings_derive_particles(C.byref(s)))
    return s


def small_physics_settings(dims=(24, 12, 12), n_target=40000, omega=0.81, wall_temp=300.0, U=1500.0,
                           density=1.0e20, faces="freestream"):
    """A scaled-down version of the argon sphere case (PHYSICS mode)."""
    import cudadsmc_amd as dm

    s = dm.settings_default()
    s.mode = dm.MODE_PHYSICS
    s.bmin[:] = (0.0, 0.0, 0.0)
    s.bmax[:] = (0.2, 0.1, 0.1)
    s.grid_dim[:] = dims
    for a in range(3):
        s.cell_size[a] = (s.bmax[a] - s.bmin[a]) / s.grid_dim[a]
    s.molmass = 6.63e-26
    s.diameter = 4.17e-10
    s.omega = omega
    s.Tref = 273.0
    s.temp = 200.0
    s.density = density
    s.stream_velocity[:] = (U, 0.0, 0.0)
    s.wall_temp = wall_temp
    kind = {"freestream": dm.FACE_FREESTREAM, "specular": dm.FACE_SPECULAR, "periodic": dm.FACE_PERIODIC}[faces]
    s.face_bc[:] = [kind] * 6
    s.target_particles = n_target
    s.dt = 1.0e-6
    s.n_streams = 4096
    return s
