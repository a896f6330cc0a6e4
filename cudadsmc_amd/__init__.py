"""MI355X-native DSMC engine (HIP kernels for gfx950 behind a C ABI).

Python view of the engine for tests, benchmarks and tooling. The solver
mirrors the reference's dsmc::DSMCSolver method set (include/dsmc_base.h:40-80):
Init, SetModel, InitSystemState, RunBenchmarkSimulationStep / RunSimulationStep,
DoBenchmark, SampleStatistics, DumpStatistics; CreateSolver mirrors
dsmc::CreateSolver (kernels/dsmc_base.cu:221-233).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import (FACE_FREESTREAM, FACE_PERIODIC, FACE_SPECULAR, MODE_PHYSICS, MODE_REFERENCE, SOLVER_CPU,
                    SOLVER_GPU, SOLVER_GPU_VBO, CellStats, DsmcError, Settings, check, lib)

__all__ = [
    "Settings", "CellStats", "DsmcError", "Solver", "CreateSolver", "Mesh", "icosphere", "sphere_cone", "load_raw",
    "settings_default", "settings_from_file", "cgrid_build", "MODE_REFERENCE", "MODE_PHYSICS", "FACE_SPECULAR",
    "FACE_PERIODIC", "FACE_FREESTREAM", "SOLVER_GPU", "SOLVER_GPU_VBO", "SOLVER_CPU",
]


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def settings_default() -> Settings:
    """settings_t::settings_t (src/sim_settings.cpp:34-44)."""
    return _capi.settings_default()


def settings_from_file(path: str, derive: bool = True) -> Settings:
    """settings_t::readFromFile (src/sim_settings.cpp:46-132) + main's particle count."""
    s = settings_default()
    check(lib().dsmc_settings_read_file(C.byref(s), path.encode()))
    if derive and s.mode == MODE_REFERENCE:
        check(lib().dsmc_settings_derive_particles(C.byref(s)))
    return s


def settings_from_cmdline(argv: list[str]) -> Settings:
    """settings_t::readFromCmdLine (src/sim_settings.cpp:134-186)."""
    s = settings_default()
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    check(lib().dsmc_settings_read_cmdline(C.byref(s), len(argv), arr))
    return s


@dataclass
class Mesh:
    vertices: np.ndarray  # (nv, 3) float32
    triangles: np.ndarray  # (nt, 3) uint32
    normals: np.ndarray  # (nt, 3) float32

    def translate(self, d) -> "Mesh":
        return Mesh((self.vertices + np.asarray(d, np.float32)).astype(np.float32), self.triangles, self.normals)


def _mesh_call(fn, *args) -> Mesh:
    nv, nt = C.c_uint32(), C.c_uint32()
    check(fn(*args, None, 0, None, 0, None, C.byref(nv), C.byref(nt)))
    v = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    n = np.zeros((nt.value, 3), np.float32)
    check(fn(*args, _fp(v), nv.value, _up(t), nt.value, _fp(n), C.byref(nv), C.byref(nt)))
    return Mesh(v, t, n)


def icosphere(center, r: float, subdiv: int) -> Mesh:
    return _mesh_call(lib().dsmc_mesh_icosphere, *map(float, center), float(r), int(subdiv))


def sphere_cone(x0: float, cy: float, cz: float, rn: float, half_angle_deg: float, rb: float, n_az: int,
                n_profile: int) -> Mesh:
    return _mesh_call(lib().dsmc_mesh_sphere_cone, float(x0), float(cy), float(cz), float(rn), float(half_angle_deg),
                      float(rb), int(n_az), int(n_profile))


def load_raw(path: str) -> Mesh:
    """Model::LoadRaw (src/model.cpp:348-401)."""
    return _mesh_call(lib().dsmc_mesh_load_raw, path.encode())


def cgrid_build(mode: int, mesh: Mesh, dims) -> dict:
    """Collision grid (Model::CreateGrid / reg_grid_t) on the host."""
    L = lib()
    v = np.ascontiguousarray(mesh.vertices, np.float32)
    t = np.ascontiguousarray(mesh.triangles, np.uint32)
    n = np.ascontiguousarray(mesh.normals, np.float32)
    bounds = np.zeros(9, np.float32)
    d = np.zeros(3, np.uint32)
    ne = C.c_uint32()
    check(L.dsmc_cgrid_build(mode, _fp(v), len(v), _up(t), len(t), _fp(n), *map(int, dims), _fp(bounds), _up(d),
                             None, None, None, None, 0, 0, C.byref(ne)))
    cells = int(np.prod(d))
    flags = np.zeros(cells, np.uint32)
    start = np.zeros(cells, np.uint32)
    end = np.zeros(cells, np.uint32)
    ent = np.zeros(max(ne.value, 1), np.uint32)
    check(L.dsmc_cgrid_build(mode, _fp(v), len(v), _up(t), len(t), _fp(n), *map(int, dims), _fp(bounds), _up(d),
                             _up(flags), _up(start), _up(end), _up(ent), cells, len(ent), C.byref(ne)))
    return {"bounds": bounds, "dims": d, "flags": flags, "start": start, "end": end, "entries": ent[:ne.value]}


class Solver:
    """One simulation instance on the GPU (the DSMCSolver interface)."""

    def __init__(self, solver_type: int = SOLVER_GPU):
        self._h = C.c_void_p()
        check(lib().dsmc_create(solver_type, C.byref(self._h)))
        self.settings: Settings | None = None

    # -- lifetime
    def close(self) -> None:
        if self._h:
            lib().dsmc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc: int) -> None:
        check(rc, self._h)

    # -- DSMCSolver
    def Init(self, s: Settings) -> None:
        self._c(lib().dsmc_init(self._h, C.byref(s)))
        self.settings = s

    def SetModel(self, mesh: Mesh, grid=(5, 5, 20), move=None) -> None:
        """SetModel with the collision grid of `grid` cells; `move` shifts the
        mesh and its grid after the grid is built (Model::CreateGrid then
        Model::MoveModel, as the reference's driver does)."""
        v = np.ascontiguousarray(mesh.vertices, np.float32)
        t = np.ascontiguousarray(mesh.triangles, np.uint32)
        n = np.ascontiguousarray(mesh.normals, np.float32)
        if move is None:
            self._c(lib().dsmc_set_geometry(self._h, _fp(v), len(v), _up(t), len(t), _fp(n), *map(int, grid)))
        else:
            mv = np.ascontiguousarray(move, np.float32)
            self._c(lib().dsmc_set_geometry_moved(self._h, _fp(v), len(v), _up(t), len(t), _fp(n), *map(int, grid),
                                                  _fp(mv)))

    def InitSystemState(self, stream_vel=(0.0, 0.0, 0.0)) -> int:
        n = C.c_uint32()
        self._c(lib().dsmc_init_state(self._h, *map(float, stream_vel), C.byref(n)))
        return n.value

    def InitSystemStateFromFields(self, n, u, v, w, T) -> int:
        arrs = [np.ascontiguousarray(a, np.float32).ravel() for a in (n, u, v, w, T)]
        out = C.c_uint32()
        self._c(lib().dsmc_init_state_from_fields(self._h, *[_fp(a) for a in arrs], len(arrs[0]), C.byref(out)))
        return out.value

    def RunBenchmarkSimulationStep(self, dt: float) -> None:
        self._c(lib().dsmc_step(self._h, float(dt)))

    RunSimulationStep = RunBenchmarkSimulationStep

    def DoBenchmark(self, itn: int, dt: float) -> None:
        self._c(lib().dsmc_run(self._h, int(itn), float(dt)))

    def DumpStatistics(self, path: str) -> None:
        self._c(lib().dsmc_dump_statistics(self._h, path.encode()))

    def SampleStatistics(self) -> tuple[np.ndarray, np.ndarray]:
        """GPUDSMCSolver::SampleStatistics (include/gpu_dsmc.inl:277-282): the
        float4 collision totals {sel, coll, sum c_r, 0} and, after one more
        sample that the samples counter does not include, the per-cell
        statistics_t records (C, 8) float32 {numInSample, xyzSum[3], velMag, pad}."""
        coll = np.zeros(4, np.float32)
        st = np.zeros((self.n_cells, 8), np.float32)
        self._c(lib().dsmc_sample_statistics(self._h, _fp(coll), _fp(st), self.n_cells))
        return coll, st

    def sim_state(self) -> dict:
        """The engine's simulation_state_t counters (simTime, realTime, stepCount, samplesCount)."""
        st = _capi.SimState()
        self._c(lib().dsmc_get_sim_state(self._h, C.byref(st)))
        return {"sim_time": st.sim_time, "real_time": st.real_time, "step_count": st.step_count,
                "samples_count": st.samples_count}

    def format_statistics(self, sim_time=None, real_time=None, samples=None) -> str:
        """ProcessAndOutputStatistics text of the current sums (no sampling);
        counters default to the engine's own."""
        cur = self.sim_state()
        st = _capi.SimState(cur["sim_time"] if sim_time is None else sim_time,
                            cur["real_time"] if real_time is None else real_time, cur["step_count"],
                            cur["samples_count"] if samples is None else samples)
        n = C.c_uint64()
        self._c(lib().dsmc_format_statistics(self._h, C.byref(st), None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        self._c(lib().dsmc_format_statistics(self._h, C.byref(st), buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    # -- extensions
    def set_sampling(self, enabled: bool, reset: bool = True) -> None:
        self._c(lib().dsmc_set_sampling(self._h, int(enabled), int(reset)))

    def sample(self) -> None:
        self._c(lib().dsmc_sample(self._h))

    def synchronize(self) -> None:
        self._c(lib().dsmc_synchronize(self._h))

    @property
    def n_cells(self) -> int:
        s = self.settings
        return int(s.grid_dim[0]) * int(s.grid_dim[1]) * int(s.grid_dim[2])

    def cell_stats(self) -> np.ndarray:
        arr = (CellStats * self.n_cells)()
        self._c(lib().dsmc_get_cell_stats(self._h, arr, self.n_cells))
        return np.frombuffer(arr, dtype=np.float64).reshape(self.n_cells, 5).copy()

    def fields(self) -> dict:
        C_ = self.n_cells
        out = {k: np.zeros(C_, np.float32) for k in ("n", "u", "v", "w", "T")}
        self._c(lib().dsmc_get_fields(self._h, *[_fp(out[k]) for k in ("n", "u", "v", "w", "T")], C_))
        return out

    def collision_stats(self) -> np.ndarray:
        o = np.zeros(4, np.float64)
        self._c(lib().dsmc_get_collision_stats(self._h, o.ctypes.data_as(C.POINTER(C.c_double))))
        return o

    def drag(self) -> np.ndarray:
        o = np.zeros(4, np.float64)
        self._c(lib().dsmc_get_drag(self._h, o.ctypes.data_as(C.POINTER(C.c_double))))
        return o

    def counts(self) -> tuple[int, int, int]:
        st, n, sm = C.c_uint64(), C.c_uint32(), C.c_uint32()
        self._c(lib().dsmc_get_counts(self._h, C.byref(st), C.byref(n), C.byref(sm)))
        return st.value, n.value, sm.value

    def particles(self) -> dict:
        n = C.c_uint32()
        self._c(lib().dsmc_get_particles(self._h, None, None, None, 0, C.byref(n)))
        cap = n.value
        pos = np.zeros(3 * max(cap, 1), np.float32)
        vel = np.zeros(3 * max(cap, 1), np.float32)
        keys = np.zeros(max(cap, 1), np.uint32)
        self._c(lib().dsmc_get_particles(self._h, _fp(pos), _fp(vel), _up(keys), cap, C.byref(n)))
        return {"pos": pos.reshape(3, -1)[:, :cap].copy(), "vel": vel.reshape(3, -1)[:, :cap].copy(),
                "key": keys[:cap].copy()}

    def set_particles(self, pos: np.ndarray, vel: np.ndarray) -> None:
        p = np.ascontiguousarray(pos, np.float32)
        v = np.ascontiguousarray(vel, np.float32)
        self._c(lib().dsmc_set_particles(self._h, _fp(p.ravel()), _fp(v.ravel()), p.shape[1]))

    def _u32(self, fn, n: int) -> np.ndarray:
        a = np.zeros(max(n, 1), np.uint32)
        self._c(fn(self._h, _up(a), n))
        return a[:n]

    def cell_start(self) -> np.ndarray:
        return self._u32(lib().dsmc_get_cell_start, self.n_cells + 1)

    def perm(self, n: int) -> np.ndarray:
        return self._u32(lib().dsmc_get_perm, n)

    def hit_counts(self, n: int) -> np.ndarray:
        return self._u32(lib().dsmc_get_hit_counts, n)

    def layout(self) -> dict:
        v = [C.c_uint32() for _ in range(4)]
        self._c(lib().dsmc_get_layout(self._h, *[C.byref(x) for x in v]))
        return dict(zip(("tile_cells", "threads", "n_streams", "n_vwg"), (x.value for x in v)))

    def rng_state(self) -> np.ndarray:
        ns = self.layout()["n_streams"]
        return self._u32(lib().dsmc_get_rng_state, 20 * ns).reshape(ns, 20)

    def rng_params(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        ns = self.layout()["n_streams"]
        a, b, c = (np.zeros(ns, np.uint32) for _ in range(3))
        self._c(lib().dsmc_get_rng_params(self._h, _up(a), _up(b), _up(c), ns))
        return a, b, c

    def set_rng_state(self, words: np.ndarray) -> None:
        w = np.ascontiguousarray(words, np.uint32).ravel()
        self._c(lib().dsmc_set_rng_state(self._h, _up(w), len(w)))

    def set_tile_cells(self, n: int) -> None:
        """Cells per tile of the tile kernel (power of two in [32, 256]), before loading."""
        self._c(lib().dsmc_set_tile_cells(self._h, int(n)))

    def taus_state(self) -> np.ndarray:
        """Hybrid-Taus collision RNG state (REFERENCE mode with mt = 0), (cells, 4)."""
        return self._u32(lib().dsmc_get_taus_state, 4 * self.n_cells).reshape(self.n_cells, 4)

    def set_taus_state(self, words: np.ndarray) -> None:
        w = np.ascontiguousarray(words, np.uint32).ravel()
        self._c(lib().dsmc_set_taus_state(self._h, _up(w), len(w)))

    def ntc_state(self) -> np.ndarray:
        a = np.zeros(2 * self.n_cells, np.float32)
        self._c(lib().dsmc_get_ntc_state(self._h, _fp(a), self.n_cells))
        return a.reshape(-1, 2)

    def near_cells(self) -> np.ndarray:
        """PHYSICS with a body: cells whose particles go through the wall pass."""
        a = np.zeros(self.n_cells, np.uint8)
        self._c(lib().dsmc_get_near_cells(self._h, a.ctypes.data_as(C.POINTER(C.c_uint8)), self.n_cells))
        return a

    def fluid_fraction(self) -> np.ndarray:
        a = np.zeros(self.n_cells, np.float32)
        self._c(lib().dsmc_get_fluid_fraction(self._h, _fp(a), self.n_cells))
        return a

    def derived(self) -> dict:
        a = np.zeros(12, np.float32)
        self._c(lib().dsmc_get_derived(self._h, _fp(a), 12))
        return {"fnum": a[0], "vhs_k": a[1], "vhs_exp": a[2], "ntc_init": a[3], "cell_volume": a[4],
                "sigma_hs": a[5], "inj_lambda": a[6:12].copy()}

    def collision_grid(self) -> dict:
        bounds = np.zeros(9, np.float32)
        d = np.zeros(3, np.uint32)
        ne = C.c_uint32()
        self._c(lib().dsmc_get_collision_grid_full(self._h, _fp(bounds), _up(d), None, None, None, None, 0, 0,
                                                   C.byref(ne)))
        cells = int(np.prod(d))
        flags, start, end = (np.zeros(cells, np.uint32) for _ in range(3))
        ent = np.zeros(max(ne.value, 1), np.uint32)
        self._c(lib().dsmc_get_collision_grid_full(self._h, _fp(bounds), _up(d), _up(flags), _up(start), _up(end),
                                                   _up(ent), cells, len(ent), C.byref(ne)))
        return {"bounds": bounds, "dims": d, "flags": flags, "start": start, "end": end, "entries": ent[:ne.value]}

    def set_debug(self, flags: int) -> None:
        self._c(lib().dsmc_set_debug(self._h, int(flags)))

    def set_graphs(self, enabled: bool) -> None:
        self._c(lib().dsmc_set_graphs(self._h, int(enabled)))

    def set_scatter_grid(self, blocks: int) -> None:
        self._c(lib().dsmc_set_scatter_grid(self._h, int(blocks)))

    def set_timing(self, enabled: bool) -> None:
        self._c(lib().dsmc_set_timing(self._h, int(enabled)))

    def timings(self) -> np.ndarray:
        a = np.zeros(5, np.float32)
        self._c(lib().dsmc_get_timings(self._h, _fp(a), 5))
        return a

    def validate(self) -> int:
        n = C.c_uint32()
        rc = lib().dsmc_validate(self._h, C.byref(n))
        if rc not in (0, -7):
            self._c(rc)
        return n.value


class Batch:
    """Independent instances advanced by shared kernel launches (the ensemble step)."""

    def __init__(self, solvers: list[Solver]):
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(solvers))(*[s._h.value for s in solvers])
        self._b = C.c_void_p()
        check(lib().dsmc_batch_create(arr, len(solvers), C.byref(self._b)))

    def _c(self, rc: int) -> None:
        if rc:
            msg = lib().dsmc_batch_last_error(self._b) or b""
            raise DsmcError(rc, msg.decode(errors="replace"))

    def run(self, steps: int) -> None:
        self._c(lib().dsmc_batch_run(self._b, int(steps)))

    def set_sampling(self, enabled: bool, reset: bool = True) -> None:
        self._c(lib().dsmc_batch_set_sampling(self._b, int(enabled), int(reset)))

    def set_graphs(self, enabled: bool) -> None:
        self._c(lib().dsmc_batch_set_graphs(self._b, int(enabled)))

    def synchronize(self) -> None:
        self._c(lib().dsmc_batch_synchronize(self._b))

    def close(self) -> None:
        if self._b:
            lib().dsmc_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def CreateSolver(solver_type: int = SOLVER_GPU) -> Solver:
    """dsmc::CreateSolver (kernels/dsmc_base.cu:221-233)."""
    return Solver(solver_type)
