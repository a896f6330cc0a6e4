"""Settings reader, particle-count KATs and the C ABI surface (CPU).

KATs are derived from main's FP32 arithmetic (kernels/dsmc.cu:324-330) for
the reference's own config files (fixtures copied from data/ and gases.xml).
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
from pathlib import Path

import numpy as np
import pytest

import cudadsmc_amd as dm
from cudadsmc_amd import _capi

from helpers import gpu_available


@pytest.fixture()
def cfgdir(tmp_path, golden):
    for f in ["config.xml", "config_500k.xml", "config_mln.xml", "gases.xml"]:
        shutil.copyfile(golden / f, tmp_path / f)
    return tmp_path


@pytest.mark.parametrize("name,ppc,cells", [("config.xml", 5, 9720), ("config_500k.xml", 17, 32768),
                                            ("config_mln.xml", 32, 32768)])
def test_particle_count_kats(cfgdir, name, ppc, cells, monkeypatch):
    monkeypatch.chdir(cfgdir)  # <gas file="gases.xml"> is relative to the working directory
    s = dm.settings_from_file(name)
    assert s.part_per_cell == ppc
    assert s.num_particles == ppc * cells
    # recompute in numpy FP32 exactly as main does
    cv = np.float32(s.cell_size[0]) * np.float32(s.cell_size[1]) * np.float32(s.cell_size[2])
    assert int(np.float32(np.float32(s.density) * cv) / np.float32(s.fnum)) == ppc


def test_config_fields(cfgdir, monkeypatch):
    monkeypatch.chdir(cfgdir)
    s = dm.settings_from_file("config_500k.xml")
    assert list(s.bmin) == [-10, -10, -10] and list(s.bmax) == [10, 10, 10]
    assert list(s.grid_dim) == [32, 32, 32]
    assert s.periodic_condition == b"x"
    assert np.isclose(s.stream_velocity[0], 1000.0)
    assert np.isclose(s.molmass, 5.0e-26) and np.isclose(s.diameter, 3.5e-10) and s.Tref == 273.0
    assert s.cell_size[0] == np.float32(20.0) / np.float32(32)


def test_gas_file_relative_to_config_dir(cfgdir, tmp_path_factory, monkeypatch):
    monkeypatch.chdir(tmp_path_factory.mktemp("elsewhere"))
    s = dm.settings_from_file(str(cfgdir / "config.xml"))
    assert s.part_per_cell == 5


def test_defaults_match_reference_constants():
    s = dm.settings_default()
    assert list(s.grid_dim) == [30, 18, 18] and s.num_particles == 4096 * 256
    assert np.isclose(s.density, 1e20) and np.isclose(s.fnum, 0.2e14) and s.temp == 300.0
    assert s.periodic_condition == b"n" and s.geometry == 1 and s.mt == 0 and s.max_run_time == -1.0


def test_cmdline(cfgdir, monkeypatch):
    monkeypatch.chdir(cfgdir)
    s = dm.settings_from_cmdline(["prog", "-f", "config.xml", "-benchmark", "-mt", "-no-geometry", "-dt", "2.5e-5",
                                  "-run-for", "0.01", "-density", "2e20"])
    assert s.benchmark == 1 and s.mt == 1 and s.geometry == 0
    assert np.isclose(s.dt, 2.5e-5) and np.isclose(s.max_run_time, 0.01) and np.isclose(s.density, 2e20)


def test_errors(tmp_path):
    with pytest.raises(dm.DsmcError, match="IO"):
        dm.settings_from_file(str(tmp_path / "missing.xml"))
    (tmp_path / "bad.xml").write_text("<config><grid><min>0 0 0</min></grid>")
    with pytest.raises(dm.DsmcError, match="IO"):
        dm.settings_from_file(str(tmp_path / "bad.xml"))
    (tmp_path / "nogas.xml").write_text((Path(__file__).parent / "golden" / "config.xml").read_text()
                                        .replace("default", "argon"))
    shutil.copyfile(Path(__file__).parent / "golden" / "gases.xml", tmp_path / "gases.xml")
    with pytest.raises(dm.DsmcError, match="not found"):
        dm.settings_from_file(str(tmp_path / "nogas.xml"))


def test_physics_extensions(tmp_path):
    (tmp_path / "gases.xml").write_text(
        "<gases><gas name='argon'><mass>6.63e-26</mass><diam>4.17e-10</diam><temp>273</temp>"
        "<omega>0.81</omega></gas></gases>")
    (tmp_path / "c.xml").write_text(
        "<config><grid><min>0 0 0</min><max>0.4 0.2 0.2</max><dim>64 32 32</dim></grid>"
        "<density>1e20</density><fnum>1e13</fnum><temp>200</temp><streamvel>1500 0 0</streamvel>"
        "<mode>physics</mode><wall_temp>300</wall_temp><particles>100000</particles>"
        "<boundary>freestream</boundary><gas file='gases.xml'>argon</gas></config>")
    s = dm.settings_from_file(str(tmp_path / "c.xml"))
    assert s.mode == dm.MODE_PHYSICS and np.isclose(s.omega, 0.81) and s.wall_temp == 300.0
    assert s.target_particles == 100000 and list(s.face_bc) == [dm.FACE_FREESTREAM] * 6


def test_abi_exports_every_declared_symbol():
    declared = set(_capi.header_symbols())
    exported = set(_capi.exported_symbols())
    assert declared, "no declarations parsed"
    missing = declared - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    # every binding in the ctypes layer names a real export
    L = _capi.lib()
    for name in declared:
        assert hasattr(L, name)


def test_abi_settings_layout():
    assert C.sizeof(_capi.Settings) == 496
    assert _capi.Settings.rng_param_file.offset == 240


def test_solver_types():
    with pytest.raises(dm.DsmcError, match="UNSUPPORTED"):
        dm.Solver(dm.SOLVER_CPU)
    with pytest.raises(dm.DsmcError, match="UNSUPPORTED"):
        dm.Solver(dm.SOLVER_GPU_VBO)
    if not gpu_available():
        with pytest.raises(dm.DsmcError, match="DEVICE"):
            dm.Solver(dm.SOLVER_GPU)


REFERENCE_STYLE_CLIENT = r"""
// Written against the reference's dsmc::DSMCSolver (include/dsmc_base.h:40-80)
#include <fstream>
#include <iostream>
#include <sstream>
#include "dsmc/dsmc_solver.hpp"

class Probe : public dsmc::DSMCSolver {  // the protected members are reachable from a subclass
 public:
  using dsmc::DSMCSolver::DSMCSolver;
  dsmc::statistics_t* Sample(dsmc::float4& c) { return SampleStatistics(c); }
  void BenchStep(float dt) { RunBenchmarkSimulationStep(dt); }
};

int main(int argc, char** argv) {
  dsmc::settings_t s;
  s.readFromCmdLine(argc, argv);
  s.deriveParticles();
  dsmc::DSMCSolver* solver = dsmc::CreateSolver(dsmc::SolverGPU);
  if (!solver) { std::cout << "no device\n"; return 0; }
  solver->Init(s);
  dsmc::Model model;
  if (s.geometry && model.LoadRaw("ares.raw", false)) {
    model.CreateGrid(dsmc::vec3i_t(5, 5, 20));
    model.MoveModel(dsmc::vec3_t(0, 0, -3));
    solver->SetModel(&model);
  }
  dsmc::simulation_state_t state;
  solver->RegisterSimState(&state);
  unsigned n = solver->InitSystemState(dsmc::float3{s.stream_velocity[0], s.stream_velocity[1], 0.0f});
  solver->DoBenchmark(3, s.dt);
  solver->RunSimulationStep(s.dt);
  std::ostringstream out;
  solver->DumpStatistics(out);
  std::cout << n << " " << state.stepCount << " " << out.str().size() << "\n";
  delete solver;
  return 0;
}
"""


def test_reference_style_client_compiles_and_links(tmp_path):
    """The C++ mirror provides the reference's solver members with their
    signatures: a client written against them builds and links with g++."""
    import shutil as sh
    import subprocess

    if not sh.which("g++"):
        pytest.skip("g++ not available")
    src = tmp_path / "client.cpp"
    src.write_text(REFERENCE_STYLE_CLIENT)
    root = Path(__file__).resolve().parent.parent
    lib = root / "cudadsmc_amd" / "lib"
    exe = tmp_path / "client"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", f"-I{root / 'include'}", str(src), f"-L{lib}",
                        "-ldsmc_hip", f"-Wl,-rpath,{lib}", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    if not gpu_available():
        r = subprocess.run([str(exe), "-f", str(root / "tests" / "golden" / "config.xml"), "-benchmark"],
                           capture_output=True, text=True, cwd=tmp_path, timeout=120)
        assert r.returncode == 0 and "no device" in r.stdout, (r.stdout, r.stderr)


def test_driver_binary_is_built():
    root = Path(__file__).resolve().parent.parent
    exe = root / "cudadsmc_amd" / "bin" / "dsmc_main"
    assert exe.exists(), "cudadsmc_amd/bin/dsmc_main is built by cudadsmc_amd/build.py"
