"""ctypes binding of the engine's C ABI (include/dsmc/dsmc_c.h).

The shared library is built in-tree (cudadsmc_amd/lib/libdsmc_hip.so); if it
is missing this module raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(os.environ.get("DSMC_LIB") or Path(__file__).resolve().parent / "lib" / "libdsmc_hip.so")

DSMC_OK = 0
ERRORS = {-1: "INVALID", -2: "IO", -3: "DEVICE", -4: "STATE", -5: "CAPACITY", -6: "UNSUPPORTED", -7: "NUMERIC"}
SOLVER_GPU, SOLVER_GPU_VBO, SOLVER_CPU = 0, 1, 2
MODE_REFERENCE, MODE_PHYSICS = 0, 1
FACE_SPECULAR, FACE_PERIODIC, FACE_FREESTREAM = 0, 1, 2
FLAG_KEEP_PERM, FLAG_HITS = 2, 4


class Settings(C.Structure):
    """dsmc_settings: settings_t (include/sim_settings.h:32-69) + extensions."""

    _fields_ = [
        ("benchmark", C.c_int32), ("cpu", C.c_int32), ("geometry", C.c_int32), ("mt", C.c_int32),
        ("num_particles", C.c_uint32), ("part_per_cell", C.c_uint32),
        ("max_run_time", C.c_float), ("dt", C.c_float),
        ("molmass", C.c_float), ("diameter", C.c_float), ("Tref", C.c_float),
        ("gas_name", C.c_char * 64),
        ("density", C.c_float), ("fnum", C.c_float), ("temp", C.c_float),
        ("stream_velocity", C.c_float * 3),
        ("periodic_condition", C.c_char), ("pad_", C.c_char * 3),
        ("bmin", C.c_float * 3), ("bmax", C.c_float * 3), ("cell_size", C.c_float * 3),
        ("grid_dim", C.c_int32 * 3),
        ("mode", C.c_int32), ("omega", C.c_float), ("wall_temp", C.c_float),
        ("face_bc", C.c_int32 * 6),
        ("target_particles", C.c_uint32), ("seed", C.c_uint32), ("sample_every", C.c_uint32),
        ("capacity", C.c_uint32), ("n_streams", C.c_uint32),
        ("rng_param_file", C.c_char * 256),
    ]


class CellStats(C.Structure):
    _fields_ = [("num_in_sample", C.c_double), ("xyz_sum", C.c_double * 3), ("vel_mag", C.c_double)]


class SimState(C.Structure):
    """dsmc_sim_state: the counters of simulation_state_t (include/draw_helper.h:40-78)."""

    _fields_ = [("sim_time", C.c_float), ("real_time", C.c_float), ("step_count", C.c_uint32),
                ("samples_count", C.c_uint32)]


assert C.sizeof(Settings) == 496, C.sizeof(Settings)

_lib = None


def lib() -> C.CDLL:
    """Load libdsmc_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python cudadsmc_amd/build.py` "
                           "(the engine has no non-HIP fallback)")
    L = C.CDLL(str(LIB_PATH))
    P, U32, F, I32, D = C.c_void_p, C.c_uint32, C.c_float, C.c_int, C.c_double
    pU = C.POINTER(C.c_uint32)
    pF = C.POINTER(C.c_float)
    pD = C.POINTER(C.c_double)
    pS = C.POINTER(Settings)
    sig = {
        "dsmc_version": ([], C.c_char_p),
        "dsmc_global_last_error": ([], C.c_char_p),
        "dsmc_settings_default": ([pS], None),
        "dsmc_settings_read_file": ([pS, C.c_char_p], I32),
        "dsmc_settings_read_cmdline": ([pS, I32, C.POINTER(C.c_char_p)], I32),
        "dsmc_settings_derive_particles": ([pS], I32),
        "dsmc_create": ([I32, C.POINTER(P)], I32),
        "dsmc_destroy": ([P], None),
        "dsmc_last_error": ([P], C.c_char_p),
        "dsmc_init": ([P, pS], I32),
        "dsmc_init_from_config": ([P, C.c_char_p], I32),
        "dsmc_set_geometry": ([P, pF, U32, pU, U32, pF, U32, U32, U32], I32),
        "dsmc_set_geometry_moved": ([P, pF, U32, pU, U32, pF, U32, U32, U32, pF], I32),
        "dsmc_init_state": ([P, F, F, F, pU], I32),
        "dsmc_init_state_from_fields": ([P, pF, pF, pF, pF, pF, U32, pU], I32),
        "dsmc_step": ([P, F], I32),
        "dsmc_run": ([P, U32, F], I32),
        "dsmc_set_sampling": ([P, I32, I32], I32),
        "dsmc_sample": ([P], I32),
        "dsmc_get_cell_stats": ([P, C.POINTER(CellStats), U32], I32),
        "dsmc_get_fields": ([P, pF, pF, pF, pF, pF, U32], I32),
        "dsmc_get_collision_stats": ([P, pD], I32),
        "dsmc_get_drag": ([P, pD], I32),
        "dsmc_dump_statistics": ([P, C.c_char_p], I32),
        "dsmc_sample_statistics": ([P, pF, pF, U32], I32),
        "dsmc_format_statistics": ([P, C.POINTER(SimState), C.c_char_p, C.c_uint64, C.POINTER(C.c_uint64)], I32),
        "dsmc_get_sim_state": ([P, C.POINTER(SimState)], I32),
        "dsmc_get_counts": ([P, C.POINTER(C.c_uint64), pU, pU], I32),
        "dsmc_synchronize": ([P], I32),
        "dsmc_validate": ([P, pU], I32),
        "dsmc_get_particles": ([P, pF, pF, pU, U32, pU], I32),
        "dsmc_set_particles": ([P, pF, pF, U32], I32),
        "dsmc_get_cell_start": ([P, pU, U32], I32),
        "dsmc_get_perm": ([P, pU, U32], I32),
        "dsmc_get_hit_counts": ([P, pU, U32], I32),
        "dsmc_get_ntc_state": ([P, pF, U32], I32),
        "dsmc_get_rng_state": ([P, pU, U32], I32),
        "dsmc_set_rng_state": ([P, pU, U32], I32),
        "dsmc_set_tile_cells": ([P, U32], I32),
        "dsmc_get_taus_state": ([P, pU, U32], I32),
        "dsmc_set_taus_state": ([P, pU, U32], I32),
        "dsmc_get_near_cells": ([P, C.POINTER(C.c_uint8), U32], I32),
        "dsmc_get_fluid_fraction": ([P, pF, U32], I32),
        "dsmc_get_collision_grid": ([P, pU, U32], I32),
        "dsmc_get_timings": ([P, pF, U32], I32),
        "dsmc_set_timing": ([P, I32], I32),
        "dsmc_set_debug": ([P, U32], I32),
        "dsmc_set_graphs": ([P, I32], I32),
        "dsmc_set_scatter_grid": ([P, U32], I32),
        "dsmc_get_layout": ([P, pU, pU, pU, pU], I32),
        "dsmc_get_derived": ([P, pF, U32], I32),
        "dsmc_get_rng_params": ([P, pU, pU, pU, U32], I32),
        "dsmc_get_collision_grid_full": ([P, pF, pU, pU, pU, pU, pU, U32, U32, pU], I32),
        "dsmc_mesh_icosphere": ([F, F, F, F, U32, pF, U32, pU, U32, pF, pU, pU], I32),
        "dsmc_mesh_sphere_cone": ([F, F, F, F, F, F, U32, U32, pF, U32, pU, U32, pF, pU, pU], I32),
        "dsmc_mesh_load_raw": ([C.c_char_p, pF, U32, pU, U32, pF, pU, pU], I32),
        "dsmc_tri_box_overlap": ([pF, pF, pF, pF, pF], I32),
        "dsmc_cgrid_build":([I32, pF, U32, pU, U32, pF, U32, U32, U32, pF, pU, pU, pU, pU, pU, U32, U32, pU], I32),
        "dsmc_batch_create": ([C.POINTER(P), U32, C.POINTER(P)], I32),
        "dsmc_batch_destroy": ([P], None),
        "dsmc_batch_last_error": ([P], C.c_char_p),
        "dsmc_batch_run": ([P, U32], I32),
        "dsmc_batch_set_sampling": ([P, I32, I32], I32),
        "dsmc_batch_set_graphs": ([P, I32], I32),
        "dsmc_batch_synchronize": ([P], I32),
        "dsmc_op_mt_generate":([P, P, P, P, U32, U32, P, P], I32),
        "dsmc_op_cell_ids": ([pS, P, P, P, U32, P, P], I32),
        "dsmc_op_sort": ([P, U32, U32, P, P, P, P], I32),
        "dsmc_op_sort_work_words": ([U32, U32], C.c_uint64),
        "dsmc_op_sort_static": ([P, U32, U32, P, P, P], I32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


class DsmcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{ERRORS.get(code, code)}] {msg}")
        self.code = code


def check(rc: int, handle=None) -> None:
    if rc == DSMC_OK:
        return
    L = lib()
    msg = (L.dsmc_last_error(handle) if handle else L.dsmc_global_last_error()) or b""
    raise DsmcError(rc, msg.decode(errors="replace"))


def exported_symbols() -> list[str]:
    """Names of the dsmc_* functions the library exports (reads the .so's dynamic symbols)."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB_PATH)], capture_output=True, text=True, check=True)
    return sorted(line.split()[-1] for line in out.stdout.splitlines() if " T dsmc_" in line)


def header_symbols() -> list[str]:
    """Function names declared in include/dsmc/dsmc_c.h."""
    import re
    hdr = Path(__file__).resolve().parent.parent / "include" / "dsmc" / "dsmc_c.h"
    text = hdr.read_text()
    return sorted(set(re.findall(r"\b(dsmc_[a-z0-9_]+)\s*\(", text)))


def settings_default() -> Settings:
    s = Settings()
    lib().dsmc_settings_default(C.byref(s))
    return s


if os.environ.get("DSMC_PRELOAD"):
    lib()
