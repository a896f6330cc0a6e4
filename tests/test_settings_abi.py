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
