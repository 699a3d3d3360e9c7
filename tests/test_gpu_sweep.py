"""GPU sweep: all (port, mode) excitations of a mesh as field sets of one
solver (the reference's per-excitation loop, scattering_td.py:1252-1273).

On the C1 guide (two 3-mode DTBC ports, K = 6 excitations):
* the batched sweep's S-matrix equals the one-excitation-per-solver sweep's
  bit for bit (every kernel treats the sets independently);
* FP32 against an FP64 sweep of the same kernels: every S entry within 1e-4;
* a sweep with the float64 check of every column reports, per column, the
  largest |S_fp32 - S_fp64| of the separate FP64 sweep (the check runs the
  same kernels), within 1e-4;
* physics of the empty guide: |S_{2m,1m}| = 1 and S_{1m,1m} ~ 0 for every
  propagating mode, reciprocity S_{ij} = S_{ji}, no mode conversion.  Every
  channel is driven with a modulated Gaussian on 18-25 GHz, clear of the
  three cut-offs (6.56, 13.1, 14.75 GHz): content at a cut-off rings with
  zero group velocity and would need a far longer record.
"""

import numpy as np
import pytest

import cases
from magnelio_amd.signals import Excitation
from magnelio_amd.sweep import ScatteringSweep

pytestmark = pytest.mark.gpu

F = np.linspace(19e9, 24e9, 21)  # inside the 18-25 GHz drive band; all three modes propagate


@pytest.fixture(scope="module")
def sweeps():
    mesh, ports, dt = cases.c1_mesh_ports()
    wf = {(p.name, m): Excitation(18e9, 25e9, m) for p in ports for m in range(p.n_modes)}
    out = {}
    for key, kw in {"batched": {}, "single": {"batch": 1}, "fp64": {"precision": "double"},
                    "checked": {"fp64_check": "all"}}.items():
        print(f"C1 sweep {key}", flush=True)
        sw = ScatteringSweep(mesh, ports, dt=dt, n_steps=4000, f_axis=F, f_max=cases.C1_FMAX, waveforms=wf, **kw)
        out[key] = sw.run()
    return out


def test_batched_equals_single(sweeps, gpu):
    b, s = sweeps["batched"].s_params, sweeps["single"].s_params
    assert b.excitations == s.excitations and len(b.excitations) == 6
    assert len(sweeps["batched"].batches) == 1 and len(sweeps["single"].batches) == 6
    assert np.array_equal(b.matrix, s.matrix, equal_nan=True)


def test_fp32_within_1e4_of_fp64(sweeps, gpu):
    b, d = sweeps["batched"].s_params.matrix, sweeps["fp64"].s_params.matrix
    assert np.isfinite(d).all()
    assert np.abs(b - d).max() < 1e-4


def test_fp64_check_reports_the_fp64_sweep(sweeps, gpu):
    chk = sweeps["checked"].fp64
    b, d = sweeps["checked"].s_params, sweeps["fp64"].s_params
    assert np.array_equal(b.matrix, sweeps["batched"].s_params.matrix, equal_nan=True)
    assert chk["mode"] == "all" and len(chk["columns"]) == 6
    for col in chk["columns"]:
        k = b.excitations.index(tuple(col["excited"]))
        assert col["max_abs_dS"] == float(np.abs(b.matrix[:, :, k] - d.matrix[:, :, k]).max())
    assert 0.0 < chk["max_abs_dS"] < 1e-4
    print(f"C1 sweep fp32 vs fp64: {chk['db']:.1f} dB")
    assert sweeps["batched"].fp64 is None


def test_empty_guide_physics(sweeps, gpu):
    r = sweeps["fp64"].s_params
    for m in range(3):
        t = np.abs(np.abs(r.S("port2", "port1", m, m)) - 1.0).max()
        refl = np.abs(r.S("port1", "port1", m, m)).max()
        recip = np.abs(r.S("port1", "port2", m, m) - r.S("port2", "port1", m, m)).max()
        print(f"mode {m}: ||S21|-1| {t:.2e}  |S11| {refl:.2e}  |S12-S21| {recip:.2e}")
        assert t < 1e-3 and refl < 1e-3 and recip < 1e-3
        for n in range(3):
            if n != m:
                assert np.abs(r.S("port2", "port1", n, m)).max() < 1e-8


def test_auto_batch_times_caches_and_matches(tmp_path, monkeypatch, gpu):
    """batch="auto": each candidate K is timed once (a throw-away solver),
    the choice is cached, a second sweep reads it, and the S-matrix equals
    the all-in-one-solver sweep's bit for bit."""
    from magnelio_amd import autotune

    monkeypatch.setenv("MAGNELIO_HIP_AUTOTUNE_CACHE", str(tmp_path / "autotune.json"))
    mesh, ports, dt = cases.c1_mesh_ports()
    wf = {(p.name, m): Excitation(18e9, 25e9, m) for p in ports for m in range(p.n_modes)}
    kw = dict(dt=dt, n_steps=600, f_axis=F, f_max=cases.C1_FMAX, waveforms=wf)
    print("C1 sweep, all excitations in one solver", flush=True)
    ref = ScatteringSweep(mesh, ports, **kw).run()
    print("C1 sweep, batch='auto' (times K = 6, 3, 2, 1)", flush=True)
    sw = ScatteringSweep(mesh, ports, batch="auto", **kw)
    res = sw.run()
    assert sw.batch in autotune.batch_candidates(6)
    assert [len(b) for b in res.batches] == [len(sw.excited[i:i + sw.batch]) for i in range(0, 6, sw.batch)]
    (entry,) = [v for k, v in autotune._load(tmp_path / "autotune.json").items() if "|batch|" in k]
    assert entry["batch"] == sw.batch and set(entry["step_ms"]) == {"6", "3", "2", "1"}
    print(f"C1 batch choice {sw.batch}: step ms {entry['step_ms']}", flush=True)
    assert np.array_equal(res.s_params.matrix, ref.s_params.matrix, equal_nan=True)

    def no_timing(*args, **kwargs):
        raise AssertionError("cached batch timed again")

    monkeypatch.setattr(autotune, "time_steps", no_timing)
    again = ScatteringSweep(mesh, ports, batch="auto", **kw)
    again.run()
    assert again.batch == sw.batch
