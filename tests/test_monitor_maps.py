"""CPU checks of the DFT monitors' sampling maps (host integer/weight work).

The product builds each monitor sample as a sparse row of (field index,
weight) terms in the reference layout; applied to random fields it must
reproduce the oracle's cell-centre interpolation (base.py:51-173) and the
Huygens-face weighting of the far-field monitor (far_field.py:230-248).
"""

import numpy as np
import pytest

from magnelio_amd import mesh as mm
from magnelio_amd.boundaries import CPMLBoundary
from magnelio_amd.monitors import (DG_W, MonitorFarField, MonitorFieldFrequency, expand_groups,
                                   lane_order, natural, structured_groups)
from oracle import monitors as omon
from oracle import yee


def _apply(idx, w, e, h):
    vals = np.where(idx >= 0, e[np.where(idx >= 0, idx, 0)], h[np.where(idx < 0, -(idx + 1), 0)])
    return np.sum(np.where(w != 0, w * vals, 0.0), axis=1)


def _grid():
    rng = np.random.default_rng(2)
    axes = [np.cumsum(np.r_[0, rng.uniform(0.5e-3, 1e-3, n)]) for n in (14, 15, 16)]
    return mm.GridLines(*axes), axes


def test_region_map_matches_oracle_interpolation():
    g, (x, y, z) = _grid()
    rng = np.random.default_rng(3)
    e = rng.standard_normal(yee.n_e(*g.dims))
    h = rng.standard_normal(yee.n_h(*g.dims))
    mon = MonitorFieldFrequency([1e9], region=(slice(2, 7), slice(0, 15), slice(4, 9)))
    idx, w = mon.sample_map(g)
    vals = _apply(idx, w, e, h)
    ref = omon.cell_centred(e, h, x, y, z, ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"], mon.slices)
    for comp, shape, start, n in mon._layout:
        scale = np.abs(ref[comp]).max()
        assert np.allclose(natural(vals[start:start + n], shape), ref[comp], rtol=0, atol=1e-13 * scale)
    assert mon.s_h_start % 16 == 0


def test_huygens_map_weights_two_layers():
    g, (x, y, z) = _grid()
    rng = np.random.default_rng(4)
    e = rng.standard_normal(yee.n_e(*g.dims))
    h = rng.standard_normal(yee.n_h(*g.dims))
    bcs = {f: CPMLBoundary(f, g, thickness_cells=3) for f in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")}
    mon = MonitorFarField([1e9], margin_cells=2)
    idx, w = mon.sample_map(g, bcs)
    vals = _apply(idx, w, e, h)
    nodes = (x, y, z)
    lo, hi = [5, 5, 5], [14 - 5, 15 - 5, 16 - 5]
    for face, axis, nn, wgt, shape in mon.faces:
        t1, t2 = [a for a in range(3) if a != axis]
        cc = 0.5 * (nodes[axis][:-1] + nodes[axis][1:])
        assert wgt == (nodes[axis][nn] - cc[nn - 1]) / (cc[nn] - cc[nn - 1])
        sl = [None] * 3
        sl[axis] = slice(nn - 1, nn + 1)
        sl[t1] = slice(lo[t1], hi[t1])
        sl[t2] = slice(lo[t2], hi[t2])
        ref = omon.cell_centred(e, h, x, y, z, ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"], sl)
        for comp, shp, start, n in mon._layout:
            if shp != shape:
                continue
            arr = np.moveaxis(ref[comp], axis, 0)
            want = (1.0 - wgt) * arr[0] + wgt * arr[1]
            got = natural(vals[start:start + n], shp)
            if np.allclose(got, want, rtol=1e-12):
                break
        else:
            raise AssertionError(f"no sample group matches face {face}")


def test_lane_order_layout():
    """element [t][q][l][c] of the lane-ordered table is P[16 t + l % 16][4 (4 q + c) + l // 16]."""
    rng = np.random.default_rng(0)
    P = rng.standard_normal((32, 128))
    L = lane_order(P).reshape(2, 8, 64, 4)
    for t in range(2):
        for q in range(8):
            for lane in range(64):
                for c in range(4):
                    assert L[t, q, lane, c] == P[16 * t + lane % 16, 4 * (4 * q + c) + lane // 16]


def test_natural_order_inverts_storage():
    a = np.arange(2 * 3 * 4 * 5).reshape(2, 3, 4, 5)  # (F, n0, n1, n2) natural
    stored = a.transpose(0, 3, 2, 1).reshape(2, -1)  # first axis fastest
    assert np.array_equal(natural(stored, (3, 4, 5)), a)


def _rows(idx, w):
    """Per sample, the (index, weight) terms with non-zero weight, in order."""
    return [[(int(i), float(v)) for i, v in zip(ri, rw) if v != 0] for ri, rw in zip(idx, w)]


def _structured_roundtrip(mon, g, bcs=None):
    dev_idx, w = mon.device_map(g, bcs)
    w32 = w.astype(np.float32)
    st = structured_groups(dev_idx, w32, mon._layout)
    assert st is not None
    desc, wtab = st
    assert desc.dtype == np.int32 and desc.shape[1] == DG_W
    idx2, w2 = expand_groups(desc, wtab, mon.S)
    assert _rows(idx2, w2) == _rows(dev_idx, w32)
    return desc


def test_structured_groups_region_monitor():
    """The region map (4 E terms, 2 H terms, non-uniform weights along the
    component axis) factors into one affine group per component, and the
    groups expand to the generic map term for term."""
    g, _ = _grid()
    mon = MonitorFieldFrequency([1e9], region=(slice(2, 7), slice(0, 15), slice(4, 9)))
    desc = _structured_roundtrip(mon, g)
    assert desc.shape[0] == 6
    assert list(desc[:, 5]) == [4, 4, 4, 2, 2, 2]


def test_structured_groups_huygens_box():
    """Far-field map: 6 faces x 4 tangential components, 8 E terms (two
    layers x four edges) and 4 H terms per sample."""
    g, _ = _grid()
    bcs = {f: CPMLBoundary(f, g, thickness_cells=3) for f in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")}
    mon = MonitorFarField([1e9], margin_cells=2)
    desc = _structured_roundtrip(mon, g, bcs)
    assert desc.shape[0] == 24
    assert sorted(set(desc[:, 5])) == [4, 8]


def test_structured_groups_reject_non_affine():
    """A map whose indices are not affine in the sample coordinates (one
    sample's term moved) or whose weights vary along two axes is refused,
    so the generic gather serves it."""
    g, _ = _grid()
    mon = MonitorFieldFrequency([1e9], region=(slice(2, 7), slice(0, 15), slice(4, 9)))
    dev_idx, w = mon.device_map(g)
    w32 = w.astype(np.float32)
    bad = dev_idx.copy()
    bad[7, 1] += 1
    assert structured_groups(bad, w32, mon._layout) is None
    w_bad = w32.copy()
    w_bad[5, 0] *= np.float32(1.5)
    assert structured_groups(dev_idx, w_bad, mon._layout) is None


def _kernel_cover(desc, S):
    """Samples written by each (work-group, thread) of k_dft_sample_groups,
    emulated on the host from the launch layout words [33..35]."""
    hits = np.zeros(S, dtype=np.int64)
    n_blocks = int(desc[:, 35].astype(np.int64).sum())
    tid = np.arange(256)
    for blk in range(n_blocks):
        g = int(np.nonzero(desc[:, 33] <= blk)[0].max())
        d = desc[g].astype(np.int64)
        n0, n1 = int(d[2]), int(d[3])
        n2 = int(d[1]) // (n0 * n1)
        wb = blk - int(d[33])
        if d[34] == 0:
            loc = wb * 256 + tid
            ok = loc < d[1]
            c0, c1, c2 = loc % n0, (loc // n0) % n1, loc // (n0 * n1)
        else:
            p0, p1 = -(-n0 // 16), -(-n1 // 16)
            c0 = (wb % p0) * 16 + (tid & 15)
            c1 = ((wb // p0) % p1) * 16 + (tid >> 4)
            c2 = np.full(256, wb // (p0 * p1))
            ok = (c0 < n0) & (c1 < n1) & (c2 < n2)
        s = d[0] + c0 + n0 * (c1 + n1 * c2)
        np.add.at(hits, s[ok], 1)
    return hits


def test_structured_launch_covers_each_sample_once():
    """The work-group layout (256 consecutive samples, or 16 x 16 patches for
    the x-normal faces whose fastest coordinate is strided) writes every
    sample of every group exactly once and nothing else (the E/H padding is
    left to the buffer's zeros)."""
    g, _ = _grid()
    bcs = {f: CPMLBoundary(f, g, thickness_cells=3) for f in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")}
    for mon, b in ((MonitorFarField([1e9], margin_cells=2), bcs),
                   (MonitorFieldFrequency([1e9], region=(slice(2, 7), slice(0, 15), slice(4, 9))), None)):
        dev_idx, w = mon.device_map(g, b)
        desc, _ = structured_groups(dev_idx, w.astype(np.float32), mon._layout)
        if isinstance(mon, MonitorFarField):
            assert set(desc[:, 34]) == {0, 1}  # x faces tiled, y/z faces linear
        hits = _kernel_cover(desc, mon.S)
        want = np.zeros(mon.S, dtype=np.int64)
        for d in desc:
            want[d[0]:d[0] + d[1]] = 1
        assert np.array_equal(hits, want)


def test_blocks_per_launch_rule():
    """Consecutive blocks per contraction: as many as fit 256 sample rows per
    lane (blocks padded to 64 steps), or the requested count when it fits."""
    from magnelio_amd.monitors import blocks_for

    assert [blocks_for(b) for b in (1, 64, 65, 100, 128, 129, 256, 300, 512)] == [4, 4, 2, 2, 2, 1, 1,
                                                                                  1, 1]
    assert blocks_for(64, 1) == 1 and blocks_for(64, 2) == 2 and blocks_for(128, 2) == 2
    for b, n in ((64, 3), (100, 4), (256, 2), (512, 4)):
        with pytest.raises(ValueError):
            blocks_for(b, n)


def test_graph_length_covers_bin_update_period():
    """A captured graph spans whole launch cadences of every monitor — the
    period (block x blocks per launch) of a non-pipelined one — and replays
    only from steps congruent to its capture step."""
    from magnelio_amd.solver import FITTimeDomainSolver

    g = mm.GridLines(np.linspace(0, 1e-2, 9), np.linspace(0, 1e-2, 9), np.linspace(0, 1e-2, 9))
    mesh = mm.Mesh.from_grid(g)
    dt = mm.courant_dt(g, "normal")
    region = (slice(0, 2), slice(0, 2), slice(0, 2))
    for per_launch, graph_block, period, steps in ((None, 1, 256, 256), (1, 1, 64, 64),
                                                   (2, 1, 128, 128), (None, 100, 256, 6400)):
        mon = MonitorFieldFrequency([1e9], region=region, block=64, blocks_per_launch=per_launch,
                                    pipelined=False)
        sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions={}, total_time_steps=10, dt=dt,
                                  monitors=[mon], graph_block=graph_block)
        assert mon.period == period == mon.cadence
        assert sol.graph_steps == steps
        assert sol._graph_ok(5, steps) and not sol._graph_ok(5, steps + 1)
        sol._graph, sol._graph_len, sol._graph_start = object(), steps, 5  # a captured graph
        assert sol._graph_ok(5 + period, steps) and not sol._graph_ok(6 + period, steps)


@pytest.mark.parametrize("S,F,K,block,want", [(1336704, 1001, 1, 256, (4, 6)),
                                              (55296, 2001, 6, 64, (8, 9)),
                                              (900, 5, 1, 64, (256, 1)),
                                              (1336704, 7, 2, 64, (32, 1))])
def test_pipelined_slicing_choice(S, F, K, block, want):
    """Slice cadence and frequency split of a pipelined monitor: >= 8
    frequency tiles per work-group, the shortest power-of-two cadence whose
    launch holds >= 1.5 rounds of resident work-groups (256 CUs x waves per
    SIMD of the contraction variant) and fills >= 90 % of them; a small
    monitor contracts once per 256-step period."""
    mon = MonitorFieldFrequency(np.linspace(1e9, 2e9, F), region=(slice(0, 1),) * 3, block=block)
    mon.S, mon.F_pad = S, -(-F // 16) * 16
    c, fs = mon._slicing(K)
    assert (c, fs) == want
    tiles = mon.F_pad // 16
    per = -(-tiles // fs)
    assert fs == 1 or per >= mon.MIN_TILES
    if c < mon.period:
        slots = 256 * mon._occupancy()
        wgs = -(-(-(-(-(-S // 16)) * c // mon.period)) // 4) * K * (-(-tiles // per))
        assert wgs / (-(-wgs // slots) * slots) >= mon.MIN_FILL
        assert wgs >= mon.MIN_ROUNDS * slots


def _slices(W, P):
    """Wave range of each pipelined slice r (csrc/mgh_dft.hip DftRange)."""
    return [(r * W // P, (r + 1) * W // P) for r in range(P)]


@pytest.mark.parametrize("W,P", [(1, 256), (7, 64), (83544, 64), (3456, 32), (255, 256),
                                 (257, 256), (100, 128)])
def test_pipelined_slices_cover_every_wave_once(W, P):
    """The P slices of a period (here P: slices per period) cover waves
    [0, W) exactly once, each within the launch's grid (ceil(ceil(W / P) /
    4) work-groups of 4 waves)."""
    hits = np.zeros(W, dtype=np.int64)
    gx = -(-(-(-W // P)) // 4)
    for lo, hi in _slices(W, P):
        assert 0 <= hi - lo <= 4 * gx
        hits[lo:hi] += 1
    assert np.all(hits == 1)


@pytest.mark.parametrize("steps", [1, 5, 63, 64, 65, 255, 256, 257, 300, 511, 512, 700])
@pytest.mark.parametrize("cad", [1, 4, 64])
def test_pipelined_flush_accounts_every_row_once(steps, cad):
    """Replay of the pipelined schedule on the host: a slice every `cad`
    steps over the previous period plus the flush at the run's end
    (monitors.finalize) contract every (row, wave) of every period once."""
    W, P = 37, 64
    ns = P // cad
    count = {}

    def add(period, rows, lo, hi):
        for r in range(rows):
            for w in range(lo, hi):
                count[(period, r, w)] = count.get((period, r, w), 0) + 1

    for n in range(steps):
        p, r = divmod(n, P)
        if p >= 1 and (r + 1) % cad == 0:
            lo, hi = _slices(W, ns)[r // cad]
            add(p - 1, P, lo, hi)
    p, r = divmod(steps - 1, P)
    done = (r + 1) // cad
    if p >= 1 and done < ns:
        add(p - 1, P, done * W // ns, W)
    add(p, r + 1, 0, W)
    want = {(n // P, n % P, w) for n in range(steps) for w in range(W)}
    assert set(count) == want
    assert all(v == 1 for v in count.values())


def test_stall_detector_on_synthetic_envelopes():
    """The port-signal stall watchdog (fit_td.py:76-132): an exponential
    ring-down that crosses the threshold before the cap is never stopped, a
    plateau or a rising envelope is, once armed and the window is full, and
    an algebraic decay whose extrapolated crossing lies past the cap is."""
    from magnelio_amd.solver import SignalStallDetector

    def first_stop(env_db, cap, window=10, thr=-60.0):
        det = SignalStallDetector(40.0, window, cap)
        for k, db in enumerate(env_db):
            if db <= thr:
                return None  # the criterion itself fires first
            if det.observe(100 * k, db, thr):
                return 100 * k
        return None

    steps = np.arange(200)
    assert first_stop(-0.5 * steps, cap=20000) is None  # 0.5 dB per check: crosses at k = 120
    flat = np.concatenate([-0.5 * steps[:90], np.full(110, -45.0)])
    # armed at k = 81 (-40.5 dB), window full at k = 90; the fitted slope
    # flattens as the plateau fills the window and the projected crossing
    # passes the cap before the window is wholly flat (k = 99)
    stop = first_stop(flat, cap=20000)
    assert stop is not None and 100 * 90 < stop <= 100 * 99
    algebraic = -40.0 - 20 * np.log10(1 + steps / 50.0)
    stop = first_stop(algebraic, cap=20000)
    assert stop is not None and stop < 100 * 200
    # the same decay with a cap beyond its projected crossing is left alone
    assert first_stop(algebraic[:40], cap=10 ** 9) is None
    # below the arming level nothing is collected
    det = SignalStallDetector(40.0, 10, 100)
    assert not any(det.observe(k, -30.0, -60.0) for k in range(50))
