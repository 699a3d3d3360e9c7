"""Environment knobs: an explicit argument wins, else the environment, else
the default; malformed values raise (the reference's rule,
_backend/array_api.py:137-175)."""

import pytest

from magnelio_amd import config, kernels
from magnelio_amd import mesh as mm
from magnelio_amd.solver import FITTimeDomainSolver

KNOBS = ("MAGNELIO_PRECISION", "MAGNELIO_GPU_GRAPHS", "MAGNELIO_TILE_SKIP",
         "MAGNELIO_HIP_LAUNCH_SHAPE")


@pytest.fixture(autouse=True)
def clean_env(monkeypatch):
    for k in KNOBS:
        monkeypatch.delenv(k, raising=False)


def test_defaults_without_environment():
    assert config.precision(None) == "single"
    assert config.graphs(None) is False and config.graphs(None, default=True) is True
    assert config.tile_skip(None) == "auto"
    assert config.launch_shape(None) == kernels.DEFAULT_KCHUNK


def test_environment_then_explicit(monkeypatch):
    monkeypatch.setenv("MAGNELIO_PRECISION", "double")
    monkeypatch.setenv("MAGNELIO_GPU_GRAPHS", "1")
    monkeypatch.setenv("MAGNELIO_TILE_SKIP", "0")
    monkeypatch.setenv("MAGNELIO_HIP_LAUNCH_SHAPE", "3,2,16")
    assert config.precision(None) == "double" and config.precision("single") == "single"
    assert config.graphs(None) is True and config.graphs(False) is False
    assert config.tile_skip(None) == "off" and config.tile_skip("on") == "on"
    assert config.launch_shape(None) == kernels.launch_shape(3, 2, 16)
    assert config.launch_shape(4) == 4
    monkeypatch.setenv("MAGNELIO_HIP_LAUNCH_SHAPE", "auto")
    assert config.launch_shape(None) == "auto"
    monkeypatch.setenv("MAGNELIO_HIP_LAUNCH_SHAPE", "2")
    assert config.launch_shape(None) == 2


@pytest.mark.parametrize("name,value", [("MAGNELIO_PRECISION", "half"),
                                        ("MAGNELIO_GPU_GRAPHS", "yes"),
                                        ("MAGNELIO_TILE_SKIP", "maybe"),
                                        ("MAGNELIO_HIP_LAUNCH_SHAPE", "1,2"),
                                        ("MAGNELIO_HIP_LAUNCH_SHAPE", "fast"),
                                        ("MAGNELIO_HIP_LAUNCH_SHAPE", "1,9,8")])
def test_malformed_values_raise(monkeypatch, name, value):
    monkeypatch.setenv(name, value)
    fn = {"MAGNELIO_PRECISION": config.precision, "MAGNELIO_GPU_GRAPHS": config.graphs,
          "MAGNELIO_TILE_SKIP": config.tile_skip,
          "MAGNELIO_HIP_LAUNCH_SHAPE": config.launch_shape}[name]
    with pytest.raises(ValueError):
        fn(None)


def test_solver_fields_resolve(monkeypatch):
    grid = mm.GridLines(*(([0.0, 1e-3, 2e-3, 3e-3],) * 3))
    mesh = mm.Mesh.from_grid(grid)
    monkeypatch.setenv("MAGNELIO_PRECISION", "double")
    monkeypatch.setenv("MAGNELIO_TILE_SKIP", "0")
    s = FITTimeDomainSolver(mesh=mesh, total_time_steps=2, dt=1e-12)
    assert (s.precision, s.use_graphs, s.tile_skip, s.kchunk) == ("double", False, "off", 2)
    s = FITTimeDomainSolver(mesh=mesh, total_time_steps=2, dt=1e-12, precision="single",
                            tile_skip="on", use_graphs=True, kchunk="auto")
    assert (s.precision, s.use_graphs, s.tile_skip, s.kchunk) == ("single", True, "on", "auto")
