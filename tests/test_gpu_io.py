"""statistics.txt and the command-line driver on the GPU.

* DumpStatistics' table (ProcessAndOutputStatistics, kernels/dsmc_base.cu:
  104-163) is recomputed line by line from the per-cell sums the engine
  holds after the dump (which samples once more, uncounted, as
  SampleStatistics does, include/gpu_dsmc.inl:277-282).
* cudadsmc_amd/bin/dsmc_main runs the reference's benchmark flow
  (kernels/dsmc.cu:308-470) on data/config.xml and writes statistics.txt
  with the benchmark's wall time as the running time.
"""
from __future__ import annotations

import math
import re
import shutil
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import cudadsmc_amd as dm

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
BOLZ, PI = np.float32(1.3806e-23), np.float32(3.14159265358979)  # DSMC_BOLZ, DSMC_PI (include/dsmc.h)
NUM = r"(-?(?:nan|inf|[0-9.]+(?:e[+-]?[0-9]+)?))"


def note(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def fnum(x: str) -> float:
    return float(x.replace("-nan", "nan"))


def close(a: float, b: float, rtol: float = 2e-5) -> bool:
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return abs(a - b) <= rtol * max(abs(a), abs(b)) + 1e-30


def reference_settings(periodic_z=True):
    s = dm.settings_from_file(str(GOLDEN / "config.xml"))
    s.benchmark, s.geometry, s.mt = 1, 0, 1
    s.dt = 2.0e-6
    return s


def test_gpu_statistics_table_line_by_line(tmp_path):
    note("statistics.txt of config.xml after 201 steps, checked line by line")
    s = reference_settings()
    sol = dm.Solver()
    sol.Init(s)
    sol.InitSystemState(tuple(s.stream_velocity))
    sol.DoBenchmark(201, s.dt)
    st = sol.sim_state()
    assert st["step_count"] == 201 and st["samples_count"] == 2  # steps 100 and 200
    t = np.float32(0.0)
    for _ in range(201):
        t = np.float32(t + np.float32(s.dt))  # advanceSimTime in FP32
    assert st["sim_time"] == t
    path = tmp_path / "statistics.txt"
    sol.DumpStatistics(str(path))
    lines = path.read_text().splitlines()
    cs = sol.cell_stats()  # the sums the table was made from (the dump's extra sample included)
    coll = sol.collision_stats()
    C = int(np.prod(s.grid_dim))
    assert (cs[:, 0].sum() / 3.0) == pytest.approx(s.num_particles, rel=1e-6)  # 2 counted + 1 extra sample
    sel, col, cr = (np.float32(x) for x in coll[:3])
    m = re.fullmatch(r"Selections: " + NUM, lines[0])
    assert m and close(fnum(m[1]), float(sel))
    m = re.fullmatch(r"Collisions: " + NUM, lines[1])
    assert m and close(fnum(m[1]), float(col)) and col > 0
    m = re.fullmatch(r"Mean separation: " + NUM, lines[2])
    assert m and close(fnum(m[1]), float(cr / col))
    m = re.fullmatch(r"Collision ratio: " + NUM, lines[3])
    assert m and close(fnum(m[1]), float(col / sel))
    m = re.fullmatch(r"Simulation time: " + NUM + r" Running time: " + NUM, lines[4])
    assert m and close(fnum(m[1]), float(t)) and fnum(m[2]) == 0.0  # benchmark path: realTime stays 0
    assert lines[5] == "Cell N  ||  SampNum  ||  Dens  ||          Velocity          ||  Temperature"
    cell_volume = np.float32(np.float32(s.cell_size[0]) * np.float32(s.cell_size[1])) * np.float32(s.cell_size[2])
    a = np.float32(np.float32(s.fnum) / np.float32(cell_volume * np.float32(2)))
    row = re.compile(r"(\d+)  \|\|  (\d+)  \|\|  " + NUM + r"  \|\|  " + NUM + "," + NUM + "," + NUM + r"  \|\|  " + NUM)
    tot = 0.0
    with np.errstate(invalid="ignore", divide="ignore"):
        for i in range(C):
            m = row.fullmatch(lines[6 + i])
            assert m, lines[6 + i]
            n = np.float32(cs[i, 0])
            v = [np.float32(np.float32(cs[i, 1 + k]) / n) for k in range(3)]
            uu = np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
            tt = np.float32(np.float32(s.molmass) * np.float32(np.float32(cs[i, 4]) / n - uu)) / np.float32(3 * BOLZ)
            tot += float(tt)
            assert int(m[1]) == i and int(m[2]) == int(n)
            assert close(fnum(m[3]), float(a * n)), (i, m[3])
            for k in range(3):
                assert close(fnum(m[4 + k]), float(v[k]), 5e-5), (i, k, m[4 + k])
            assert close(fnum(m[7]), float(tt), 5e-5), (i, m[7])
    rest = "\n".join(lines[6 + C:])
    m = re.search(r"Expected collisions: " + NUM + r" Ratio: " + NUM, rest)
    assert m and fnum(m[1]) > 0
    m = re.search(r"Average temperature: " + NUM + r" Reference temperature is: " + NUM, rest)
    assert m and close(fnum(m[1]), tot / C, 1e-4) and fnum(m[2]) == 300.0
    # an equilibrium box at rest keeps the loading temperature: from the
    # moments summed over all cells it is 300 K; the table's average of the
    # per-cell estimates (a handful of samples each, no Bessel correction)
    # lies below it, as the reference's formula gives
    tot_s = cs.sum(0)
    u = tot_s[1:4] / tot_s[0]
    T_all = s.molmass * (tot_s[4] / tot_s[0] - (u ** 2).sum()) / (3 * 1.380649e-23)
    assert abs(T_all - 300.0) < 3.0, T_all
    assert 230.0 < tot / C < 300.0
    sol.close()


def test_gpu_driver_writes_statistics(tmp_path):
    exe = ROOT / "cudadsmc_amd" / "bin" / "dsmc_main"
    assert exe.exists()
    for f in ("config.xml", "gases.xml"):
        shutil.copyfile(GOLDEN / f, tmp_path / f)
    cmd = [str(exe), "-f", "config.xml", "-benchmark", "-no-geometry", "-mt", "-dt", "2e-6", "-run-for", "4e-4"]
    note("dsmc_main " + " ".join(cmd[1:]))
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"Will run (\d+) iterations", r.stdout)
    assert m, r.stdout
    iters = int(m[1])
    assert iters in (200, 201)
    m = re.search(r"Benchmark is done in ([0-9.e+-]+) ms", r.stdout)
    assert m and float(m[1]) > 0
    lines = (tmp_path / "statistics.txt").read_text().splitlines()
    m = re.fullmatch(r"Simulation time: " + NUM + r" Running time: " + NUM, lines[4])
    assert m and abs(fnum(m[1]) - iters * 2e-6) < 1e-8
    assert abs(fnum(m[2]) - float(re.search(r"done in ([0-9.e+-]+) ms", r.stdout)[1]) / 1e3) < 1e-3
    assert sum(1 for ln in lines if "  ||  " in ln) == 1 + 30 * 18 * 18
    m = re.search(r"Average temperature: " + NUM, "\n".join(lines))
    assert m and 230.0 < fnum(m[1]) < 300.0  # mean of per-cell estimates (biased low, few samples)
