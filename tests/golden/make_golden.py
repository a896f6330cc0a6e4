"""Write the golden fixtures of tests/golden from the reference's own code.

Runs in the build container only (needs /root/reference and oracle/_ref,
built by `make -C oracle ref` from the reference's CPU sources). Writes:
  mt_dcmt_rows.npz   first 64 rows of data/MersenneTwister.dat, the
                     InitGPUTwisters seed chain (mtwister.cu:41-64) and 256
                     outputs per row of Matsumoto-Nishimura genrand_mt
                     (src/genmtrand.c:45-80)
  tribox.npz         random triangle/box pairs and triBoxOverlap verdicts
                     (include/tritest.h:180-326 via src/aabb.cpp:38-50)
  grid_ares.npz      the collision grid of Model::CreateGrid({5,5,20})
                     (src/model.cpp:733-769, src/reggrid.cpp) for data/ares.raw
  grid_icosphere.npz the same loop for an icosphere (subdivision 2, 6x6x6)
and copies the small reference data files used as fixtures (config*.xml,
gases.xml, ares.raw).
"""
from __future__ import annotations

import ctypes as C
import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
REF = Path("/root/reference")
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
import cudadsmc_amd as dm  # noqa: E402


def lcg_chain(n: int, seed: int = 999) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    s = seed
    for i in range(n):
        s = (s * 279470273) % 4294967291
        out[i] = s
    return out


def main() -> None:
    oracle.build(with_ref=True)
    R = oracle.ref()
    assert R is not None, "oracle/_ref not built"
    # ---- MT golden vectors
    rows = np.fromfile(REF / "data" / "MersenneTwister.dat", dtype="<u4").reshape(-1, 4)[:64].copy()
    seeds = lcg_chain(64)
    n = 256
    outs = np.zeros((64, n), np.uint32)
    for i in range(64):
        buf = np.zeros(n, np.uint32)
        R.ref_genrand(int(rows[i, 0]), int(rows[i, 1]), int(rows[i, 2]), int(seeds[i]), n, oracle.up(buf))
        outs[i] = buf
    np.savez(HERE / "mt_dcmt_rows.npz", rows=rows, seeds=seeds, outputs=outs)
    # ---- triangle/box verdicts (boxes of half-size ~1, triangles around them,
    # including near-touching configurations)
    rng = np.random.default_rng(20240601)
    m = 4000
    center = rng.uniform(-1, 1, (m, 3)).astype(np.float32)
    half = rng.uniform(0.2, 1.0, (m, 3)).astype(np.float32)
    tri = (center[:, None, :] + rng.uniform(-3.0, 3.0, (m, 3, 3)) * half[:, None, :]).astype(np.float32)
    # snap some vertices onto box faces to exercise the boundary cases
    snap = rng.random(m) < 0.25
    ax = rng.integers(0, 3, m)
    sgn = np.where(rng.random(m) < 0.5, -1.0, 1.0).astype(np.float32)
    for k in np.nonzero(snap)[0]:
        tri[k, 0, ax[k]] = center[k, ax[k]] + sgn[k] * half[k, ax[k]]
    verdict = np.zeros(m, np.int32)
    for k in range(m):
        c = np.ascontiguousarray(center[k])
        h = np.ascontiguousarray(half[k])
        a, b, cc = (np.ascontiguousarray(tri[k, j]) for j in range(3))
        verdict[k] = R.ref_tri_box(oracle.fp(c), oracle.fp(h), oracle.fp(a), oracle.fp(b), oracle.fp(cc))
    np.savez(HERE / "tribox.npz", center=center, half=half, tri=tri, verdict=verdict)
    # ---- reference collision grids
    def ref_grid(mesh, bbox6, dims):
        v = np.ascontiguousarray(mesh.vertices, np.float32)
        t = np.ascontiguousarray(mesh.triangles, np.uint32)
        cells = int(np.prod(dims))
        cap = 4 * len(t) * 8 + 1024
        flags = np.zeros(cells, np.uint32)
        counts = np.zeros(cells, np.uint32)
        ent = np.zeros(cap, np.uint32)
        cell = np.zeros(3, np.float32)
        ne = C.c_uint32()
        bb = np.ascontiguousarray(bbox6, np.float32)
        rc = R.ref_create_grid(oracle.fp(v), oracle.up(t), len(t), oracle.fp(bb), int(dims[0]), int(dims[1]),
                               int(dims[2]), oracle.fp(cell), oracle.up(flags), oracle.up(counts), oracle.up(ent), cap,
                               C.byref(ne))
        assert rc == 0
        return dict(bbox=bb, dims=np.asarray(dims, np.uint32), cell=cell, flags=flags, counts=counts,
                    entries=ent[:ne.value].copy())

    ares = dm.load_raw(str(REF / "data" / "ares.raw"))
    raw = (REF / "data" / "ares.raw").read_bytes()
    bbox6 = np.frombuffer(raw[10:34], dtype="<f4")  # bbox stored after magic and versions
    np.savez(HERE / "grid_ares.npz", **ref_grid(ares, bbox6, (5, 5, 20)))
    ico = dm.icosphere((0.1, 0.1, 0.1), 0.025, 2)
    v = ico.vertices
    np.savez(HERE / "grid_icosphere.npz", **ref_grid(ico, np.concatenate([v.min(0), v.max(0)]), (6, 6, 6)))
    # ---- reference data files used as fixtures
    for f in ["config.xml", "config_500k.xml", "config_mln.xml", "ares.raw"]:
        shutil.copyfile(REF / "data" / f, HERE / f)
    shutil.copyfile(REF / "gases.xml", HERE / "gases.xml")
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
