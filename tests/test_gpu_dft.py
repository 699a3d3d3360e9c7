"""GPU parity of the DFT monitors (sampling kernel + MFMA block contraction).

* the block contraction against numpy on random inputs: ragged F (37) and S
  (70, H samples from 48), two field sets, a full and a partial block, plain
  and compensated bins, blocks of 64-512 steps (one and two LDS passes) — a
  wrong MFMA operand or result lane layout gives O(1) errors, float32
  products give ~1e-7; 2 or 4 consecutive blocks per launch (each with its
  own phase, bins updated once), full and partial fills;
* a region monitor inside the plain WR-90 run (C1, 300 steps: four full
  blocks and a partial one closed by ``finalize``) against the oracle's
  FieldDFT fed with float64 fields — FP64 and FP32 stepping;
* the same run replayed through HIP graphs == eager launches, bit for bit;
* the pipelined contraction (mgh_dft_slice): the slices of one period of
  steps, driven by the device counter, add the previous period's block
  contraction to the bins (numpy), including slices with no waves and split
  frequency groups; the explicit-range flush (mgh_dft_blocks_range) too;
* a run flushed by ``finalize`` in mid-period and then continued gives the
  bins of an uninterrupted run (each sample contracted once).
"""

import math

import numpy as np
import pytest
import torch

import cases
from magnelio_amd import _lib
from magnelio_amd.monitors import MonitorFieldFrequency, block_pad, lane_order
from oracle import monitors as omon

pytestmark = pytest.mark.gpu

FREQS = [8e9, 12e9, 16e9, 20e9, 24e9]
REGION = (slice(10, 30), slice(4, 20), slice(30, 60))
N_STEPS = 300


@pytest.mark.parametrize("B", [64, 100, 256, 300, 512])
def test_block_contraction_matches_numpy(B, gpu):
    rng = np.random.default_rng(7)
    K, S, s_h, F = 2, 70, 48, 37
    S_pad, F_pad = 80, 48
    dt = 1.3e-12
    omega = 2 * np.pi * rng.uniform(1e9, 30e9, F)
    bp = block_pad(B)
    b = np.arange(bp)
    Pc = np.zeros((F_pad, bp), np.float32)
    Ps = np.zeros((F_pad, bp), np.float32)
    Pc[:F, :B] = (np.cos(np.outer(omega, b * dt)) * dt)[:, :B]
    Ps[:F, :B] = (np.sin(np.outer(omega, b * dt)) * dt)[:, :B]
    X = rng.standard_normal((K, B, S_pad)).astype(np.float32)
    n_start = 5
    ctr_val = n_start + 3 * B + 40
    n0 = n_start + 3 * B
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev)
    Pcd, Psd = torch.from_numpy(lane_order(Pc)).to(dev), torch.from_numpy(lane_order(Ps)).to(dev)
    om = torch.from_numpy(omega).to(dev)
    blk = torch.zeros((F, 4), dtype=torch.float64, device=dev)
    ctr = torch.tensor([ctr_val], dtype=torch.int64, device=dev)
    t0 = np.where(np.arange(S) >= s_h, n0 * dt + 0.5 * dt, n0 * dt)
    rot = np.exp(1j * omega[:, None] * t0[None, :])  # (F, S)
    P = (Pc[:F].astype(np.float64) + 1j * Ps[:F].astype(np.float64))
    for nb in (B, 23):
        for comp in (False, True):
            print(f"dft block contraction B={B} nb={nb} compensated={comp}", flush=True)
            # bins already holding values of the size of one block's sum
            base = rng.standard_normal((K, F, S, 2)) * (8 * dt)
            bins = torch.from_numpy(base.copy()).to(dev)
            lo = torch.zeros(bins.shape, dtype=torch.float32, device=dev) if comp else None
            stream = torch.cuda.current_stream(dev)
            _lib.call("mgh_dft_block", K, Xd.data_ptr(), S, S_pad, B, nb, Pcd.data_ptr(),
                      Psd.data_ptr(), om.data_ptr(), F, F_pad, dt, blk.data_ptr(), s_h,
                      bins.data_ptr(), lo.data_ptr() if comp else None, ctr.data_ptr(), n_start,
                      int(stream.cuda_stream))
            torch.cuda.synchronize(dev)
            got = bins.cpu().numpy()
            if comp:
                got = got + lo.cpu().numpy().astype(np.float64)
            for k in range(K):
                want = (P[:, :nb] @ X[k, :nb, :S].astype(np.float64)) * rot
                g = (got[k, ..., 0] - base[k, ..., 0]) + 1j * (got[k, ..., 1] - base[k, ..., 1])
                err = np.abs(g - want).max() / (math.sqrt(nb) * dt)
                print(f"  set {k}: max err / (sqrt(nb) dt) = {err:.2e}", flush=True)
                assert err < 1e-5


@pytest.mark.parametrize("B,nsub", [(64, 2), (64, 4), (100, 2), (128, 2)])
def test_blocks_per_launch_match_numpy(B, nsub, gpu):
    """nsub consecutive blocks in one launch: each block's own contraction and
    phase, the bins updated once (full, one-block-and-a-bit and short fills)."""
    rng = np.random.default_rng(11)
    K, S, s_h, F = 2, 70, 48, 37
    S_pad, F_pad = 80, 48
    dt = 1.3e-12
    P_steps = B * nsub
    omega = 2 * np.pi * rng.uniform(1e9, 30e9, F)
    bp = block_pad(B)
    b = np.arange(bp)
    Pc = np.zeros((F_pad, bp), np.float32)
    Ps = np.zeros((F_pad, bp), np.float32)
    Pc[:F, :B] = (np.cos(np.outer(omega, b * dt)) * dt)[:, :B]
    Ps[:F, :B] = (np.sin(np.outer(omega, b * dt)) * dt)[:, :B]
    X = rng.standard_normal((K, P_steps, S_pad)).astype(np.float32)
    n_start = 5
    ctr_val = n_start + 3 * P_steps + 40
    n0 = n_start + 3 * P_steps
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev)
    Pcd, Psd = torch.from_numpy(lane_order(Pc)).to(dev), torch.from_numpy(lane_order(Ps)).to(dev)
    om = torch.from_numpy(omega).to(dev)
    blk = torch.zeros((nsub * F, 4), dtype=torch.float64, device=dev)
    ctr = torch.tensor([ctr_val], dtype=torch.int64, device=dev)
    P = (Pc[:F].astype(np.float64) + 1j * Ps[:F].astype(np.float64))
    for nb in (P_steps, B + 17, 23):
        for comp in (False, True):
            print(f"dft blocks B={B} nsub={nsub} nb={nb} compensated={comp}", flush=True)
            base = rng.standard_normal((K, F, S, 2)) * (8 * dt)
            bins = torch.from_numpy(base.copy()).to(dev)
            lo = torch.zeros(bins.shape, dtype=torch.float32, device=dev) if comp else None
            stream = torch.cuda.current_stream(dev)
            _lib.call("mgh_dft_blocks", K, Xd.data_ptr(), S, S_pad, B, nsub, nb, Pcd.data_ptr(),
                      Psd.data_ptr(), om.data_ptr(), F, F_pad, dt, blk.data_ptr(), s_h,
                      bins.data_ptr(), lo.data_ptr() if comp else None, ctr.data_ptr(), n_start,
                      int(stream.cuda_stream))
            torch.cuda.synchronize(dev)
            got = bins.cpu().numpy()
            if comp:
                got = got + lo.cpu().numpy().astype(np.float64)
            for k in range(K):
                want = np.zeros((F, S), complex)
                for q in range(nsub):
                    rows = np.arange(q * B, min((q + 1) * B, nb))
                    if rows.size == 0:
                        continue
                    t0 = (n0 + q * B) * dt + np.where(np.arange(S) >= s_h, 0.5 * dt, 0.0)
                    rot = np.exp(1j * omega[:, None] * t0[None, :])
                    want += (P[:, : rows.size] @ X[k, rows, :S].astype(np.float64)) * rot
                g = (got[k, ..., 0] - base[k, ..., 0]) + 1j * (got[k, ..., 1] - base[k, ..., 1])
                err = np.abs(g - want).max() / (math.sqrt(nb) * dt)
                print(f"  set {k}: max err / (sqrt(nb) dt) = {err:.2e}", flush=True)
                assert err < 1e-5


@pytest.mark.parametrize("S,nsub,fsplit,cad", [(70, 1, 1, 1), (70, 4, 2, 4), (3000, 2, 3, 8),
                                               (5, 1, 1, 64), (3000, 1, 2, 2)])
def test_pipelined_slices_match_numpy(S, nsub, fsplit, cad, gpu):
    """One period of pipelined slices (a slice every `cad` steps, indices
    from the device counter, phases from the period table) == the
    contraction of the previous period's buffer; then a flush of part of
    the other buffer through mgh_dft_blocks_range."""
    rng = np.random.default_rng(5)
    K, F, B = 2, 37, 64
    s_h = (S // 2) // 16 * 16
    S_pad, F_pad = -(-S // 16) * 16, 48
    P_steps = B * nsub
    dt = 1.1e-12
    omega = 2 * np.pi * rng.uniform(1e9, 30e9, F)
    b = np.arange(B)
    Pc = np.zeros((F_pad, B), np.float32)
    Ps = np.zeros((F_pad, B), np.float32)
    Pc[:F] = np.cos(np.outer(omega, b * dt)) * dt
    Ps[:F] = np.sin(np.outer(omega, b * dt)) * dt
    X = rng.standard_normal((2, K, P_steps, S_pad)).astype(np.float32)
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev)
    Pcd, Psd = torch.from_numpy(lane_order(Pc)).to(dev), torch.from_numpy(lane_order(Ps)).to(dev)
    om = torch.from_numpy(omega).to(dev)
    blk = torch.zeros((nsub * F, 4), dtype=torch.float64, device=dev)
    base = rng.standard_normal((K, F, S, 2)) * (8 * dt)
    bins = torch.from_numpy(base.copy()).to(dev)
    lo = torch.zeros(bins.shape, dtype=torch.float32, device=dev)
    n_start = 3
    P = Pc[:F].astype(np.float64) + 1j * Ps[:F].astype(np.float64)
    stream = int(torch.cuda.current_stream(dev).cuda_stream)
    xbuf = K * P_steps * S_pad

    def want_of(buf, n0, rows, waves=None):
        w = np.zeros((K, F, S), complex)
        cols = np.arange(S) if waves is None else np.arange(waves[0] * 16, min(waves[1] * 16, S))
        for k in range(K):
            for q in range(nsub):
                rr = np.arange(q * B, min((q + 1) * B, rows))
                if rr.size == 0:
                    continue
                t0 = (n0 + q * B) * dt + np.where(cols >= s_h, 0.5 * dt, 0.0)
                rot = np.exp(1j * omega[:, None] * t0[None, :])
                w[k][:, cols] += (P[:, : rr.size] @ X[buf, k, rr][:, cols].astype(np.float64)) * rot
        return w

    def got_minus(ref):
        g = bins.cpu().numpy() + lo.cpu().numpy().astype(np.float64)
        return (g[..., 0] - ref[..., 0]) + 1j * (g[..., 1] - ref[..., 1])

    n_per = 8
    tab = torch.zeros((n_per, nsub, F, 4), dtype=torch.float64, device=dev)
    _lib.call("mgh_dft_phase_table", om.data_ptr(), F, dt, n_start, B, nsub, n_per, tab.data_ptr(),
              stream)

    def slice_at(n):
        ctr = torch.tensor([n], dtype=torch.int64, device=dev)
        _lib.call("mgh_dft_slice", K, Xd.data_ptr(), xbuf, S, S_pad, B, nsub, Pcd.data_ptr(),
                  Psd.data_ptr(), F, F_pad, tab.data_ptr(), n_per, s_h, bins.data_ptr(),
                  lo.data_ptr(), ctr.data_ptr(), n_start, cad, fsplit, stream)

    # period 2 runs: its steps contract period 1, held in buffer 1
    for r in range(P_steps):
        if (r + 1) % cad == 0:
            slice_at(n_start + 2 * P_steps + r)
    torch.cuda.synchronize(dev)
    err = np.abs(got_minus(base) - want_of(1, n_start + P_steps, P_steps)).max() / (
        math.sqrt(P_steps) * dt)
    print(f"pipelined S={S} nsub={nsub} fsplit={fsplit}: max err {err:.2e}", flush=True)
    assert err < 1e-5
    # period 0 contracts nothing
    before = bins.cpu().numpy().copy()
    slice_at(n_start + cad - 1)
    torch.cuda.synchronize(dev)
    assert np.array_equal(bins.cpu().numpy(), before)
    # explicit flush of 23 rows of buffer 0 over a wave range
    W = -(-S // 16)
    wl, wh = W // 3, W
    n0 = n_start + 7 * P_steps
    _lib.call("mgh_dft_blocks_range", K, Xd[0].data_ptr(), S, S_pad, B, nsub, 23, wl, wh, n0,
              Pcd.data_ptr(), Psd.data_ptr(), om.data_ptr(), F, F_pad, dt, blk.data_ptr(), s_h,
              bins.data_ptr(), lo.data_ptr(), fsplit, stream)
    torch.cuda.synchronize(dev)
    err = np.abs(got_minus(before) - want_of(0, n0, 23, (wl, wh))).max() / (math.sqrt(23) * dt)
    print(f"range flush: max err {err:.2e}", flush=True)
    assert err < 1e-5


def test_blocks_per_launch_checks(gpu):
    dev = torch.device("cuda:0")
    z = torch.zeros(64, dtype=torch.float64, device=dev)
    stream = int(torch.cuda.current_stream(dev).cuda_stream)
    # 4 blocks of 100 steps (128 padded) exceed the 256 rows a lane holds
    with pytest.raises(ValueError, match="exceed 256"):
        _lib.call("mgh_dft_blocks", 1, z.data_ptr(), 1, 16, 100, 4, 10, z.data_ptr(), z.data_ptr(),
                  z.data_ptr(), 1, 16, 1e-12, z.data_ptr(), 0, z.data_ptr(), None, z.data_ptr(), 0,
                  stream)
    with pytest.raises(ValueError, match="must be 1, 2 or 4"):
        _lib.call("mgh_dft_blocks", 1, z.data_ptr(), 1, 16, 64, 3, 10, z.data_ptr(), z.data_ptr(),
                  z.data_ptr(), 1, 16, 1e-12, z.data_ptr(), 0, z.data_ptr(), None, z.data_ptr(), 0,
                  stream)


def _oracle_dft(steps):
    orc, _, dt, _ = cases.c1_oracle(np.float64)
    x, y, z = cases.c1_axes()
    mon = omon.FieldDFT(FREQS, x, y, z, REGION)
    orc.monitors.append(mon)
    orc.run(steps)
    return mon.bins


@pytest.fixture(scope="module")
def oracle_bins():
    print(f"oracle C1 float64 + region DFT, {N_STEPS} steps (CPU)", flush=True)
    return _oracle_dft(N_STEPS)


def _hip_bins(precision, use_graphs=False, block=64, per_launch=None, pipelined=True):
    mon = MonitorFieldFrequency(FREQS, region=REGION, block=block, blocks_per_launch=per_launch,
                                pipelined=pipelined)
    sol, _, _ = cases.c1_solver(n_steps=N_STEPS, precision=precision, use_graphs=use_graphs,
                                monitors=[mon])
    sol.run()
    if use_graphs:
        assert sol._graph is not None, getattr(sol, "graph_error", "graph not captured")
    return mon.result(0)


def _errors(got, want):
    """Per component: relative L2 for the TE10 components (Ey, Hx, Hz);
    absolute error over the E or H peak for the ones that stay at round-off."""
    out = {}
    peak = {f: max(np.abs(want[c]).max() for c in want if c[0] == f) for f in "EH"}
    for c in want:
        if c in ("Ey", "Hx", "Hz"):
            out[c] = float(np.linalg.norm(got[c] - want[c]) / np.linalg.norm(want[c]))
        else:
            out[c] = float(np.abs(got[c] - want[c]).max() / peak[c[0]])
    return out


@pytest.mark.parametrize("precision,gate,block,per_launch,pipelined", [
    ("double", 1e-5, 64, None, True), ("single", 2e-5, 64, None, True),
    ("single", 2e-5, 64, 1, True), ("single", 2e-5, 128, None, True),
    ("double", 1e-5, 256, None, True), ("single", 2e-5, 256, None, True),
    ("single", 2e-5, 512, None, True), ("double", 1e-5, 64, None, False),
    ("single", 2e-5, 64, 1, False), ("single", 2e-5, 512, None, False)])
def test_region_monitor_vs_oracle(oracle_bins, precision, gate, block, per_launch, pipelined, gpu):
    print(f"HIP C1 {precision} + region DFT (block {block}, {per_launch} per launch, "
          f"pipelined {pipelined}), {N_STEPS} steps", flush=True)
    got = _hip_bins(precision, block=block, per_launch=per_launch, pipelined=pipelined)
    assert set(got) == set(oracle_bins)
    for c in oracle_bins:
        assert got[c].shape == oracle_bins[c].shape
    errs = _errors(got, oracle_bins)
    print(f"region DFT {precision} vs float64 oracle: " +
          " ".join(f"{c} {v:.2e}" for c, v in errs.items()), flush=True)
    assert max(errs.values()) < gate


@pytest.mark.parametrize("pipelined", [False, True])
def test_region_monitor_graph_equals_eager(pipelined, gpu):
    # two 64-step blocks per launch: non-pipelined, a 128-step graph (steps
    # 128-255) inside the 300; pipelined, one-step graphs from step 2 on
    print(f"HIP C1 single + region DFT, graphs vs eager (pipelined {pipelined})", flush=True)
    eager = _hip_bins("single", per_launch=2, pipelined=pipelined)
    graph = _hip_bins("single", use_graphs=True, per_launch=2, pipelined=pipelined)
    for c in eager:
        assert np.array_equal(eager[c], graph[c]), c


@pytest.mark.parametrize("pipelined", [False, True])
def test_finalize_mid_period_then_continue(oracle_bins, pipelined, gpu):
    """finalize() at step 150 (mid-period: 256-step periods), then the run
    goes on to 300: every sample is contracted once — the bins match the
    oracle as closely as an uninterrupted run."""
    print(f"HIP C1 double, finalize at 150 then continue (pipelined {pipelined})", flush=True)
    mon = MonitorFieldFrequency(FREQS, region=REGION, block=64, pipelined=pipelined)
    sol, _, _ = cases.c1_solver(n_steps=N_STEPS, precision="double", monitors=[mon])
    sol.run(150)
    mon.finalize(sol)
    sol.run()
    errs = _errors(mon.result(0), oracle_bins)
    print("  " + " ".join(f"{c} {v:.2e}" for c, v in errs.items()), flush=True)
    assert max(errs.values()) < 1e-5


def _farfield_bins(structured, precision, steps=150):
    """Small CPML box (28 x 24 x 32 non-uniform cells, 40 seeded objects),
    +z TF/SF Gaussian, far-field monitor over a Huygens box, K = 2."""
    from bench import c2_mesh
    from magnelio_amd import mesh as mm
    from magnelio_amd.boundaries import CPMLBoundary
    from magnelio_amd.monitors import MonitorFarField
    from magnelio_amd.solver import FITTimeDomainSolver
    from magnelio_amd.sources import PlaneWaveSource

    length = 64e-3
    mesh = c2_mesh(32, length=length)
    bcs = {f: CPMLBoundary(f, mesh.grid, thickness_cells=4) for f in mm.FACES}
    mon = MonitorFarField(np.linspace(1e9, 20e9, 7), margin_cells=2, block=64)
    mon.structured = structured
    src = PlaneWaveSource(direction=(0, 0, 1), polarization=(1, 0, 0),
                          corners=((0.15 * length,) * 3, (0.85 * length,) * 3), f_max=20e9,
                          waveform="gaussian")
    sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, total_time_steps=steps,
                              dt=mm.courant_dt(mesh.grid, "normal"), precision=precision, nsets=2,
                              record=False, sources=[src], monitors=[mon])
    sol.run()
    assert (mon.grp_d is not None) == structured
    return mon.bins.cpu().numpy(), mon.lo.cpu().numpy()


@pytest.mark.parametrize("precision", ["single", "double"])
def test_structured_sampling_equals_generic_farfield(precision, gpu):
    """The structured sampling kernel and the per-sample (index, weight)
    gather give the same bins bit for bit (24 face groups, 8 and 4 terms)."""
    print(f"far-field DFT structured vs generic gather, {precision}", flush=True)
    a = _farfield_bins(True, precision)
    b = _farfield_bins(False, precision)
    assert np.abs(a[0]).max() > 0
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_structured_sampling_equals_generic_region(gpu):
    print("region DFT structured vs generic gather (C1, single)", flush=True)
    got = _hip_bins("single")
    mon = MonitorFieldFrequency(FREQS, region=REGION, block=64)
    mon.structured = False
    sol, _, _ = cases.c1_solver(n_steps=N_STEPS, precision="single", monitors=[mon])
    sol.run()
    assert mon.grp_d is None
    ref = mon.result(0)
    for c in ref:
        assert np.array_equal(got[c], ref[c]), c
