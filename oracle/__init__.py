"""CPU oracle of the DSMC step -- TEST INFRASTRUCTURE.

ctypes view of oracle/_build/libdsmc_oracle.so (the C++ restatement in
dsmc_cpu.cpp) and, when built, of oracle/_ref/libdsmc_ref.so (the
reference's own CPU sources). Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline use it, as the checker; the engine never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libdsmc_oracle.so"
REF_LIB = HERE / "_ref" / "libdsmc_ref.so"

_lib = None
_ref = None


def build(with_ref: bool = True) -> None:
    """Compile the oracle (and the reference CPU sources when present)."""
    subprocess.run(["make", "-B", "-C", str(HERE)], check=True, capture_output=True)
    if with_ref and Path("/root/reference/src/genmtrand.c").exists():
        subprocess.run(["make", "-B", "-C", str(HERE), "ref"], check=True, capture_output=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build(with_ref=False)
    from cudadsmc_amd._capi import Settings

    L = C.CDLL(str(LIB))
    P, U32, F, I32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
    pU, pF, pD = C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.POINTER(C.c_double)
    pS = C.POINTER(Settings)
    sig = {
        "oracle_mt_generic": ([U32] * 12 + [pU], None),
        "oracle_mt_streams": ([pU, pU, pU, pU, U32, U32, pU], None),
        "oracle_cell_ids": ([pS, pF, pF, pF, U32, pU], None),
        "oracle_stable_sort": ([pU, U32, U32, pU, pU], None),
        "oracle_intersect": ([pF, pF, pF, pF, pF, pF], I32),
        "oracle_detmath": ([I32, pF, pF, U32, pF], None),
        "oracle_create": ([pS, U32, U32, U32, pU, pU, pU], P),
        "oracle_destroy": ([P], None),
        "oracle_set_geometry": ([P, pF, U32, pU, U32, pF, pF, pU, pU, pU, pU, pU, U32, pF], I32),
        "oracle_set_scalars": ([P, F, F, F, pF], I32),
        "oracle_set_rng": ([P, pU, U32], I32),
        "oracle_get_rng": ([P, pU, U32], I32),
        "oracle_set_ntc": ([P, pF, U32], I32),
        "oracle_get_ntc": ([P, pF, U32], I32),
        "oracle_tile_order": ([P, pU, U32], I32),
        "oracle_tile_walk": ([P, U32, pU, U32], I32),
        "oracle_set_state": ([P, pF, pF, U32], I32),
        "oracle_n_stored": ([P], U32),
        "oracle_get_state": ([P, pF, pF, pU, U32], I32),
        "oracle_load": ([P, F, F, F], I32),
        "oracle_step": ([P, I32], I32),
        "oracle_ntc_cell": ([P, U32, U32, pF, pF, pF, U32, F, pF, pD], I32),
        "oracle_pair_perm": ([U32, pU, U32], U32),
        "oracle_get_perm": ([P, pU, U32], I32),
        "oracle_get_cell_start": ([P, pU, U32], I32),
        "oracle_get_hits": ([P, pU, U32], I32),
        "oracle_get_samples": ([P, pD, U32], I32),
        "oracle_get_cstat": ([P, pF, U32], I32),
        "oracle_get_acc": ([P, pD, U32], I32),
        "oracle_threads": ([], I32),
        "oracle_set_threads": ([I32], None),
        "oracle_set_taus": ([P, pU, U32], I32),
        "oracle_get_taus": ([P, pU, U32], I32),
        "oracle_taus_draws": ([pU, I32, U32, pU], None),
        "oracle_taus_period": ([I32, U32, C.c_uint64], C.c_uint64),
        "oracle_mt_window_after": ([U32, U32, U32, U32, U32, pU], None),
        "oracle_philox": ([pU, pU, pU], None),
        "oracle_set_near_cells": ([P, C.POINTER(C.c_uint8)], I32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def ref() -> C.CDLL | None:
    """The reference's compiled CPU sources, or None where they were not built."""
    global _ref
    if _ref is not None:
        return _ref
    if not REF_LIB.exists():
        return None
    R = C.CDLL(str(REF_LIB))
    U32, pU, pF = C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_float)
    R.ref_genrand.argtypes = [U32, U32, U32, U32, U32, pU]
    R.ref_genrand.restype = None
    R.ref_tri_box.argtypes = [pF, pF, pF, pF, pF]
    R.ref_tri_box.restype = C.c_int
    R.ref_create_grid.argtypes = [pF, pU, U32, pF, C.c_int, C.c_int, C.c_int, pF, pU, pU, pU, U32, pU]
    R.ref_create_grid.restype = C.c_int
    _ref = R
    return R


def fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def up(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ----------------------------------------------------------------- primitives
def mt_generic(n, m, a, lower, u, s, b, t, c, l, seed, n_out) -> np.ndarray:
    out = np.zeros(n_out, np.uint32)
    lib().oracle_mt_generic(n, m, a, lower, u, s, b, t, c, l, seed, n_out, up(out))
    return out


def mt_streams(a, b, c, seeds, n_draws) -> np.ndarray:
    a, b, c, seeds = (np.ascontiguousarray(x, np.uint32) for x in (a, b, c, seeds))
    out = np.zeros((n_draws, len(a)), np.uint32)
    lib().oracle_mt_streams(up(a), up(b), up(c), up(seeds), len(a), n_draws, up(out))
    return out


def cell_ids(settings, x, y, z) -> np.ndarray:
    x, y, z = (np.ascontiguousarray(v, np.float32) for v in (x, y, z))
    out = np.zeros(len(x), np.uint32)
    lib().oracle_cell_ids(C.byref(settings), fp(x), fp(y), fp(z), len(x), up(out))
    return out


def stable_sort(keys, n_cells) -> tuple[np.ndarray, np.ndarray]:
    keys = np.ascontiguousarray(keys, np.uint32)
    perm = np.zeros(max(len(keys), 1), np.uint32)
    cs = np.zeros(n_cells + 1, np.uint32)
    lib().oracle_stable_sort(up(keys), len(keys), n_cells, up(perm), up(cs))
    return perm[: int(cs[-1])], cs


def taus_draws(seed, n: int, component: int = 0) -> np.ndarray:
    """Hybrid-Taus outputs from state seed (x, y, z, w); component 1-3 = one
    Tausworthe component, 4 = the LCG, 0 = the combined generator."""
    sd = np.ascontiguousarray(seed, np.uint32)
    out = np.zeros(n, np.uint32)
    lib().oracle_taus_draws(up(sd), component, n, up(out))
    return out


def philox(ctr, key) -> np.ndarray:
    """Philox4x32-10 of a 4-word counter under a 2-word key."""
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().oracle_philox(up(c), up(k), up(out))
    return out


def mt_window_after(a: int, b: int, c: int, seed: int, d: int) -> np.ndarray:
    """Sliding-window state (19 words + a zero) of one stream after d draws."""
    out = np.zeros(20, np.uint32)
    lib().oracle_mt_window_after(a, b, c, seed, d, up(out))
    return out


def taus_period(component: int, v: int, limit: int) -> int:
    return int(lib().oracle_taus_period(component, v, limit))


def intersect(o, d, v0, v1, v2):
    arrs = [np.ascontiguousarray(x, np.float32) for x in (o, d, v0, v1, v2)]
    out = np.zeros(3, np.float32)
    hit = lib().oracle_intersect(*[fp(a) for a in arrs], fp(out))
    return bool(hit), out


def detmath(f: int, x, y=None) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y if y is not None else x, np.float32)
    out = np.zeros_like(x)
    lib().oracle_detmath(f, fp(x), fp(y), len(x), fp(out))
    return out


# ----------------------------------------------------------------- simulation
class OracleSim:
    """The oracle's simulation object, mirrored from an engine Solver."""

    def __init__(self, settings, layout: dict, rng_params):
        a, b, c = (np.ascontiguousarray(x, np.uint32) for x in rng_params)
        self._keep = (a, b, c)
        self.settings = settings
        self.C = int(settings.grid_dim[0]) * int(settings.grid_dim[1]) * int(settings.grid_dim[2])
        self.S = int(layout["n_streams"])
        self._o = lib().oracle_create(C.byref(settings), layout["tile_cells"], layout["threads"], self.S, up(a),
                                      up(b), up(c))
        if not self._o:
            raise RuntimeError("oracle_create failed")

    def __del__(self):
        if getattr(self, "_o", None):
            lib().oracle_destroy(self._o)
            self._o = None

    @classmethod
    def from_solver(cls, solver) -> "OracleSim":
        o = cls(solver.settings, solver.layout(), solver.rng_params())
        d = solver.derived()
        lam = np.ascontiguousarray(d["inj_lambda"], np.float32)
        lib().oracle_set_scalars(o._o, float(d["fnum"]), float(d["vhs_k"]), float(d["vhs_exp"]), fp(lam))
        return o

    def set_near_cells(self, flags) -> None:
        f = np.ascontiguousarray(flags, np.uint8)
        self._near = f
        assert lib().oracle_set_near_cells(self._o, f.ctypes.data_as(C.POINTER(C.c_uint8))) == 0

    def set_geometry(self, mesh, grid: dict, fluid_frac=None) -> None:
        v = np.ascontiguousarray(mesh.vertices, np.float32)
        t = np.ascontiguousarray(mesh.triangles, np.uint32)
        n = np.ascontiguousarray(mesh.normals, np.float32)
        b = np.ascontiguousarray(grid["bounds"], np.float32)
        d = np.ascontiguousarray(grid["dims"], np.uint32)
        fl, st, en = (np.ascontiguousarray(grid[k], np.uint32) for k in ("flags", "start", "end"))
        ent = np.ascontiguousarray(grid["entries"], np.uint32)
        if len(ent) == 0:
            ent = np.zeros(1, np.uint32)
        ff = None if fluid_frac is None else np.ascontiguousarray(fluid_frac, np.float32)
        self._geom = (v, t, n, b, d, fl, st, en, ent, ff)
        rc = lib().oracle_set_geometry(self._o, fp(v), len(v), up(t), len(t), fp(n), fp(b), up(d), up(fl), up(st),
                                       up(en), up(ent), len(grid["entries"]), fp(ff) if ff is not None else None)
        assert rc == 0

    def set_rng(self, words: np.ndarray) -> None:
        w = np.ascontiguousarray(words, np.uint32).ravel()
        assert lib().oracle_set_rng(self._o, up(w), len(w)) == 0

    def rng(self) -> np.ndarray:
        w = np.zeros(20 * self.S, np.uint32)
        assert lib().oracle_get_rng(self._o, up(w), len(w)) == 0
        return w.reshape(self.S, 20)

    def set_taus(self, words: np.ndarray) -> None:
        w = np.ascontiguousarray(words, np.uint32).ravel()
        assert lib().oracle_set_taus(self._o, up(w), len(w)) == 0

    def taus(self) -> np.ndarray:
        w = np.zeros(4 * self.C, np.uint32)
        assert lib().oracle_get_taus(self._o, up(w), len(w)) == 0
        return w.reshape(self.C, 4)

    def set_ntc(self, ntc: np.ndarray) -> None:
        a = np.ascontiguousarray(ntc, np.float32).ravel()
        assert lib().oracle_set_ntc(self._o, fp(a), self.C) == 0

    def tile_order(self) -> np.ndarray:
        a = np.zeros(self.C, np.uint32)
        k = lib().oracle_tile_order(self._o, up(a), self.C)
        assert k >= 0
        return a[:k].copy()

    def tile_walk(self, vw: int) -> np.ndarray:
        a = np.zeros(self.C, np.uint32)
        k = lib().oracle_tile_walk(self._o, vw, up(a), self.C)
        assert k >= 0
        return a[:k].copy()

    def ntc(self) -> np.ndarray:
        a = np.zeros(2 * self.C, np.float32)
        assert lib().oracle_get_ntc(self._o, fp(a), self.C) == 0
        return a.reshape(-1, 2)

    def set_state(self, pos: np.ndarray, vel: np.ndarray) -> None:
        p = np.ascontiguousarray(pos, np.float32)
        v = np.ascontiguousarray(vel, np.float32)
        assert lib().oracle_set_state(self._o, fp(p.ravel()), fp(v.ravel()), p.shape[1]) == 0

    def state(self) -> dict:
        n = lib().oracle_n_stored(self._o)
        pos = np.zeros(3 * max(n, 1), np.float32)
        vel = np.zeros(3 * max(n, 1), np.float32)
        key = np.zeros(max(n, 1), np.uint32)
        assert lib().oracle_get_state(self._o, fp(pos), fp(vel), up(key), max(n, 1)) == 0
        cap = max(n, 1)
        return {"pos": pos.reshape(3, cap)[:, :n].copy(), "vel": vel.reshape(3, cap)[:, :n].copy(),
                "key": key[:n].copy()}

    def load(self, u=(0.0, 0.0, 0.0)) -> None:
        assert lib().oracle_load(self._o, *map(float, u)) == 0

    def step(self, sample: bool = False) -> None:
        assert lib().oracle_step(self._o, int(sample)) == 0

    def cell_start(self) -> np.ndarray:
        a = np.zeros(self.C + 1, np.uint32)
        assert lib().oracle_get_cell_start(self._o, up(a), len(a)) == 0
        return a

    def perm(self, n: int) -> np.ndarray:
        a = np.zeros(max(n, 1), np.uint32)
        lib().oracle_get_perm(self._o, up(a), n)
        return a[:n]

    def hits(self, n: int) -> np.ndarray:
        a = np.zeros(max(n, 1), np.uint32)
        lib().oracle_get_hits(self._o, up(a), n)
        return a[:n]

    def samples(self) -> np.ndarray:
        a = np.zeros(5 * self.C, np.float64)
        assert lib().oracle_get_samples(self._o, dp(a), self.C) == 0
        return a.reshape(5, self.C).T.copy()

    def cstat(self) -> np.ndarray:
        a = np.zeros(4 * self.C, np.float32)
        assert lib().oracle_get_cstat(self._o, fp(a), self.C) == 0
        return a.reshape(-1, 4)

    def acc(self) -> np.ndarray:
        a = np.zeros(7, np.float64)
        lib().oracle_get_acc(self._o, dp(a), 7)
        return a
