"""Frequency-domain (DFT) monitors on the device.

Reference counterparts (relative to the reference repository root):

* ``DFTAccumulator``          src/magnelio/monitors/_dft.py:18-61
* cell-centre interpolation   src/magnelio/monitors/base.py:51-173
  E: 0.25 x (4 edge voltages) / primal length; H: 0.5 x sum of the two face
  values divided by their solver dual widths (d[0], averages, d[N-1])
* region snapping             src/magnelio/monitors/base.py:207-330
* ``MonitorFieldFrequency``   src/magnelio/monitors/field_frequency.py:249-294
* ``MonitorFarField``         src/magnelio/monitors/far_field.py:104-248
  (Huygens box ``margin_cells`` inside the CPML faces; tangential E/H of each
  face linearly weighted between the two cell layers around the box plane)

The per-step work is one sampling kernel (cell-centre samples of every
monitored component into a block buffer) and an MFMA phasor contraction
fused with the float64 bin update (csrc/mgh_dft.hip).  Pipelined (the
default): the samples of period p (P = block x blocks_per_launch steps) go
to one of two buffers while each step contracts 1/P of the samples of period
p - 1 from the other, so every step carries the same share of the
contraction; the run's end flushes what is left.  Otherwise one contraction
of the whole buffer runs every P steps.
E samples carry the time stamp t = n dt, H samples t + dt/2, where n is the
index of the step whose H update has just finished — the reference's stamps
(fit_td.py:996-1002), kept as they are.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib, layout
from .farfield import far_field_from_dump, source_spectrum, wall_position

NT = 8
BMAX = 512  # longest block (steps per MFMA contraction)
BLOCK = 64  # default block
TANGENTIALS = {0: ("Ey", "Ez", "Hy", "Hz"), 1: ("Ex", "Ez", "Hx", "Hz"), 2: ("Ex", "Ey", "Hx", "Hy")}
COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def _dual(d: np.ndarray) -> np.ndarray:
    n = d.size
    out = np.empty(n + 1)
    out[0] = d[0]
    out[n] = d[-1]
    if n > 1:
        out[1:n] = 0.5 * (d[:-1] + d[1:])
    return out


def natural(vals: np.ndarray, shape) -> np.ndarray:
    """Samples of one group (last axis, stored with the first grid axis
    fastest) -> (..., *shape) in natural axis order."""
    lead = vals.shape[:-1]
    nd = len(shape)
    arr = vals.reshape(*lead, *shape[::-1])
    return arr.transpose(*range(len(lead)), *range(len(lead) + nd - 1, len(lead) - 1, -1))


def unnatural(arr: np.ndarray) -> np.ndarray:
    """Inverse of :func:`natural`: (F, *shape) -> (F, n) samples, first
    grid axis fastest."""
    nd = arr.ndim - 1
    return np.ascontiguousarray(np.moveaxis(arr, tuple(range(1, nd + 1)), tuple(range(nd, 0, -1)))
                                ).reshape(arr.shape[0], -1)


def _corners_array(corners) -> np.ndarray:
    """(2, 3) corners, None as -inf / +inf, axes sorted (the stored form)."""
    if corners is None:
        return np.array([[-np.inf] * 3, [np.inf] * 3])
    a = np.array([[(-np.inf if v is None else float(v)) for v in corners[0]],
                  [(np.inf if v is None else float(v)) for v in corners[1]]])
    return np.sort(a, axis=0)


def blocks_for(block: int, requested: int | None = None) -> int:
    """Consecutive blocks contracted per launch (the bins are read and
    written once for all of them): ``requested`` (1, 2 or 4, checked), or by
    default as many as fit 256 sample rows per lane."""
    fits = [n for n in (4, 2, 1) if n == 1 or block_pad(block) * n <= 256]
    if requested is None:
        return fits[0]
    if int(requested) not in fits:
        raise ValueError(f"blocks_per_launch must be one of {sorted(fits)} for {block}-step blocks")
    return int(requested)


def block_pad(block: int) -> int:
    """Phasor-table row length of a block: whole 64-step register tiles up to
    256 steps, 512 beyond (the contraction kernel's variants)."""
    return (block + 63) // 64 * 64 if block <= 256 else 512


def lane_order(P: np.ndarray) -> np.ndarray:
    """(F_pad, Bp) phasor table -> the contraction kernel's lane order, float4
    groups: element [t][q][l][c] = P[16 t + l % 16][4 (4 q + c) + l // 16]."""
    fp, bp = P.shape
    # P[16 t + col][16 q + 4 c + krow] -> (t, q, krow, col, c); lane l = 16 krow + col
    return np.ascontiguousarray(P.reshape(fp // 16, 16, bp // 16, 4, 4).transpose(0, 2, 4, 1, 3)).ravel()


def snap_range(cc: np.ndarray, lo: float, hi: float) -> slice:
    """Cell indices whose centres fall within [lo, hi] (base.py:207-220)."""
    if lo == hi:
        i = int(np.argmin(np.abs(cc - lo)))
        return slice(i, i + 1)
    idx = np.nonzero((cc >= lo) & (cc <= hi))[0]
    if idx.size == 0:
        i = int(np.argmin(np.abs(cc - 0.5 * (lo + hi))))
        return slice(i, i + 1)
    return slice(int(idx[0]), int(idx[-1]) + 1)


class _SampleBuilder:
    """Sparse (index, weight) rows of cell-centred samples, reference layout."""

    def __init__(self, grid):
        self.grid = grid
        self.dims = grid.dims
        nx, ny, nz = self.dims
        self.d = (grid.dx, grid.dy, grid.dz)
        self.da = tuple(_dual(a) for a in self.d)
        shp = layout.component_shapes(nx, ny, nz)
        self.shape = shp
        self.e_off = {"Ex": 0, "Ey": int(np.prod(shp["Ex"])), "Ez": int(np.prod(shp["Ex"]) + np.prod(shp["Ey"]))}
        self.h_off = {"Hx": 0, "Hy": int(np.prod(shp["Hx"])), "Hz": int(np.prod(shp["Hx"]) + np.prod(shp["Hy"]))}

    def _flat(self, comp, i, j, k):
        n0, n1, n2 = self.shape[comp]
        off = self.e_off.get(comp, self.h_off.get(comp))
        return off + (i * n1 + j) * n2 + k

    def centred(self, comp: str, I, J, K):
        """(idx list, weight list) of the cell-centre value at cells (I, J, K) (broadcast arrays)."""
        I, J, K = np.broadcast_arrays(I, J, K)
        terms = []
        if comp in ("Ex", "Ey", "Ez"):
            ax = "xyz".index(comp[1])
            other = [a for a in range(3) if a != ax]
            L = self.d[ax][(I, J, K)[ax]]
            for oa in (0, 1):
                for ob in (0, 1):
                    ijk = [I, J, K]
                    ijk[other[0]] = ijk[other[0]] + oa
                    ijk[other[1]] = ijk[other[1]] + ob
                    terms.append((self._flat(comp, *ijk), 0.25 / L))
        else:
            ax = "xyz".index(comp[1])
            for o in (0, 1):
                ijk = [I, J, K]
                ijk[ax] = ijk[ax] + o
                terms.append((-(self._flat(comp, *ijk)) - 1, 0.5 / self.da[ax][ijk[ax]]))
        return terms


def _pack(rows: list[list]) -> tuple[np.ndarray, np.ndarray]:
    """rows: list over sample groups of [(idx array, w array), ...] -> (S, 8)."""
    idx_blocks, w_blocks = [], []
    for terms in rows:
        n = terms[0][0].size
        idx = np.zeros((n, NT), dtype=np.int64)
        w = np.zeros((n, NT), dtype=np.float64)
        for t, (ii, ww) in enumerate(terms):
            idx[:, t] = ii.ravel()
            w[:, t] = ww.ravel()
        idx_blocks.append(idx)
        w_blocks.append(w)
    return np.concatenate(idx_blocks), np.concatenate(w_blocks)


DG_W = 40  # int32 words per structured group descriptor (csrc/mgh_dft.hip)


def structured_groups(dev_idx: np.ndarray, w32: np.ndarray, groups: list):
    """Factor a device-layout sample map into affine groups.

    ``dev_idx`` (S, 8) int64 (E >= 0, H at -(i + 1)), ``w32`` (S, 8) float32,
    ``groups``: [(comp, shape, start, count)] with samples numbered c0
    fastest.  Returns (descriptors (G, DG_W) int32, weight tables float32)
    whose expansion (``expand_groups``) equals the map exactly, or None when
    some group does not factor (the generic gather then serves the map)."""
    desc, tabs, n_tab = [], [], 0
    for comp, shape, start, count in sorted(groups, key=lambda g: g[2]):
        if count == 0:
            continue
        shp = tuple(int(v) for v in shape) + (1,) * (3 - len(shape))
        if len(shape) > 3 or int(np.prod(shp)) != count:
            return None
        n0, n1, n2 = shp
        rows = slice(start, start + count)
        w = w32[rows]
        used = [t for t in range(w.shape[1]) if np.any(w[:, t] != 0)]
        if not used or len(used) > 8:
            return None
        is_h = comp.startswith("H")
        field = 1 if is_h else 0
        d = np.zeros(DG_W, dtype=np.int64)
        d[0:6] = (start, count, n0, n1, field, len(used))
        c2, c1, c0 = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
        c0, c1, c2 = c0.ravel(), c1.ravel(), c2.ravel()
        strides = None
        for q, t in enumerate(used):
            wt = w[:, t]
            live = wt != 0
            ix = dev_idx[rows, t].astype(np.int64)
            if is_h:
                if np.any(ix[live] >= 0):
                    return None
                ix = -(ix + 1)
            elif np.any(ix[live] < 0):
                return None
            base = int(ix[0]) if live[0] else None
            if strides is None or base is None:
                # strides from the first sample and its neighbours along each axis
                st = []
                for step, n_ax in ((1, n0), (n0, n1), (n0 * n1, n2)):
                    st.append(int(ix[step] - ix[0]) if n_ax > 1 else 0)
                if strides is None:
                    strides = st
            if base is None:
                return None
            pred = base + c0 * strides[0] + c1 * strides[1] + c2 * strides[2]
            if not np.array_equal(pred[live], ix[live]):
                return None
            # weight: constant, or a function of one natural coordinate
            grid = wt.reshape(n2, n1, n0)
            ax, tab = None, None
            if np.all(wt == wt[0]):
                ax, tab = -1, wt[:1]
            else:
                for a, line in ((0, grid[0, 0, :]), (1, grid[0, :, 0]), (2, grid[:, 0, 0])):
                    cand = (line[c0] if a == 0 else line[c1] if a == 1 else line[c2])
                    if np.array_equal(cand, wt):
                        ax, tab = a, line
                        break
            if ax is None:
                return None
            d[9 + q] = base
            d[17 + q] = ax
            d[25 + q] = n_tab
            tabs.append(np.asarray(tab, dtype=np.float32))
            n_tab += tab.size
        d[6:9] = strides
        desc.append(d)
    if not desc:
        return None
    arr = np.stack(desc)
    # work-groups of each group: 256 consecutive samples where c0 is
    # unit-stride in memory, 16 x 16 patches of (c0, c1) where it is strided
    first = 0
    for d in arr:
        n0, n1 = int(d[2]), int(d[3])
        n2 = int(d[1]) // (n0 * n1)
        tiled = abs(int(d[6])) > 1
        nblk = (-(-n0 // 16)) * (-(-n1 // 16)) * n2 if tiled else -(-int(d[1]) // 256)
        d[33], d[34], d[35] = first, int(tiled), nblk
        first += nblk
    if np.abs(arr).max() >= 2**31 - 1:
        return None
    return arr.astype(np.int32), np.concatenate(tabs).astype(np.float32)


def expand_groups(desc: np.ndarray, wtab: np.ndarray, S: int):
    """(idx, w) (S, 8) of structured groups, zero-weight terms dropped and
    the rest packed first — the rows the sampling kernel sums."""
    idx = np.zeros((S, 8), dtype=np.int64)
    w = np.zeros((S, 8), dtype=np.float32)
    for d in desc.astype(np.int64):
        start, count, n0, n1, field, nt = (int(v) for v in d[:6])
        loc = np.arange(count)
        c = (loc % n0, (loc // n0) % n1, loc // (n0 * n1))
        off = c[0] * d[6] + c[1] * d[7] + c[2] * d[8]
        for t in range(nt):
            ax, wo = int(d[17 + t]), int(d[25 + t])
            wt = wtab[wo + (c[ax] if ax >= 0 else 0)]
            ix = d[9 + t] + off
            idx[start:start + count, t] = -(ix + 1) if field else ix
            w[start:start + count, t] = wt
    return idx, w


class DFTMonitorBase:
    """Shared device machinery: sample map, block buffer, phasors, bins."""

    block = BLOCK
    compensated = True
    # sample through the structured groups when the map factors into them
    # (False: always the generic per-sample (index, weight) gather)
    structured = True

    def __init__(self, freqs, name: str = "", block: int = BLOCK, compensated: bool = True,
                 blocks_per_launch: int | None = None, pipelined: bool = True,
                 fsplit: int | None = None, slice_steps: int | None = None):
        self.freqs = np.asarray(freqs, dtype=float).ravel()
        self.pipelined = bool(pipelined)
        if fsplit is not None and int(fsplit) < 1:
            raise ValueError("fsplit must be a positive int")
        self.fsplit_req = fsplit
        self.slice_steps_req = slice_steps
        if self.freqs.size == 0:
            raise ValueError("freqs must be a non-empty 1D array")
        if not (1 <= block <= BMAX):
            raise ValueError(f"block must be 1..{BMAX}")
        self.block = int(block)
        self.blocks_per_launch = blocks_for(self.block, blocks_per_launch)
        self.compensated = compensated
        self.name = name or f"dft_{id(self):x}"
        self._attached = False

    # subclasses fill self._groups = [(comp, shape, terms)]
    def _build(self, grid, boundary_conditions) -> None:  # pragma: no cover - abstract
        raise NotImplementedError

    def sample_map(self, grid, boundary_conditions=None):
        """(idx, w) of every sample in the reference flat layout, (S, 8):
        idx >= 0 indexes e, idx < 0 indexes h at -(idx + 1); E samples
        first, H samples from ``s_h_start`` (a multiple of 16)."""
        self._build(grid, boundary_conditions or {})
        e_rows, h_rows = [], []
        self._layout = []  # (comp, shape, start, count)
        s = 0
        for comp, shape, terms in self._groups:
            if comp.startswith("E"):
                e_rows.append(terms)
        for comp, shape, terms in self._groups:
            if comp.startswith("E"):
                n = int(np.prod(shape))
                self._layout.append((comp, shape, s, n))
                s += n
        s_e = s
        self.s_h_start = int(math.ceil(s_e / 16) * 16)
        s = self.s_h_start
        for comp, shape, terms in self._groups:
            if comp.startswith("H"):
                h_rows.append(terms)
                n = int(np.prod(shape))
                self._layout.append((comp, shape, s, n))
                s += n
        self.S = s
        self.S_pad = int(math.ceil(self.S / 16) * 16)
        idx = np.zeros((self.S, NT), dtype=np.int64)
        w = np.zeros((self.S, NT), dtype=np.float64)
        if e_rows:
            ie, we = _pack(e_rows)
            idx[: ie.shape[0]] = ie
            w[: ie.shape[0]] = we
        if h_rows:
            ih, wh = _pack(h_rows)
            idx[self.s_h_start: self.s_h_start + ih.shape[0]] = ih
            w[self.s_h_start: self.s_h_start + ih.shape[0]] = wh
        return idx, w

    def device_map(self, grid, boundary_conditions=None):
        """``sample_map`` with the indices in the device layout (E >= 0, H at
        -(i + 1))."""
        idx, w = self.sample_map(grid, boundary_conditions)
        dims = grid.dims
        pos = idx >= 0
        dev_idx = np.zeros_like(idx)
        if pos.any():
            dev_idx[pos] = layout.index_to_device(idx[pos], *dims, "E")
        neg = (~pos) & (w != 0)
        if neg.any():
            dev_idx[neg] = -(layout.index_to_device(-(idx[neg] + 1), *dims, "H")) - 1
        if np.abs(dev_idx).max() >= 2**31 - 1:
            raise ValueError("field vectors too large for 32-bit sample indices")
        return dev_idx, w

    def footprint(self, grid, boundary_conditions, K: int) -> int:
        """Device bytes ``attach`` allocates for K field sets: the float64 bins
        and their float32 compensation terms, the float32 sample buffers, the
        sample map (int32 index + float32 weight per term) and the phasor
        tables — from the host-side sample map, before any device work."""
        idx, _ = self.sample_map(grid, boundary_conditions)
        F = int(self.freqs.size)
        bins = K * F * self.S * 2 * (8 + (4 if self.compensated else 0))
        nbuf = 2 if self.pipelined else 1
        X = nbuf * K * self.period * self.S_pad * 4
        tables = 2 * int(math.ceil(F / 16) * 16) * block_pad(self.block) * 4 + 8 * idx.size
        return int(bins + X + tables)

    def attach(self, solver) -> None:
        dev_idx, w = self.device_map(solver.mesh.grid, solver.boundary_conditions)
        dev = solver.e.device
        K = solver.nsets
        self.F = self.freqs.size
        self.F_pad = int(math.ceil(self.F / 16) * 16)
        dt = solver.dt
        self.dt = dt
        omega = 2.0 * math.pi * self.freqs
        bp = block_pad(self.block)
        b = np.arange(bp)
        ph = np.outer(omega, b * dt)
        Pc = np.zeros((self.F_pad, bp))
        Ps = np.zeros((self.F_pad, bp))
        valid = b < self.block
        Pc[: self.F, valid] = (np.cos(ph) * dt)[:, valid]
        Ps[: self.F, valid] = (np.sin(ph) * dt)[:, valid]
        f32 = torch.float32
        w32 = w.astype(np.float32)
        self.idx_d = torch.from_numpy(dev_idx.astype(np.int32)).to(dev)
        self.w_d = torch.from_numpy(w32).to(dev)
        self.grp_d = self.wtab_d = None
        if self.structured:
            st = structured_groups(dev_idx, w32, self._layout)
            if st is not None:
                self.grp_d = torch.from_numpy(st[0]).to(dev)
                self.wtab_d = torch.from_numpy(st[1]).to(dev)
                self.grp_blocks = int(st[0][:, 35].astype(np.int64).sum())
        self.Pc = torch.from_numpy(lane_order(Pc)).to(dev, f32)
        self.Ps = torch.from_numpy(lane_order(Ps)).to(dev, f32)
        self.omega = torch.from_numpy(omega).to(dev, torch.float64)
        self.blk = torch.zeros((self.blocks_per_launch * self.F, 4), dtype=torch.float64, device=dev)
        # pipelined: two sample buffers, period p in buffer p mod 2
        nbuf = 2 if self.pipelined else 1
        self.X = torch.zeros((nbuf, K, self.period, self.S_pad), dtype=f32, device=dev)
        self.x_buf = K * self.period * self.S_pad if self.pipelined else 0
        props = torch.cuda.get_device_properties(dev)
        self.slice_steps, self.fsplit = self._slicing(K, int(props.multi_processor_count))
        self.nsets = K
        self.n_start = solver.n_completed
        self._horizon = solver.horizon
        import weakref  # noqa: PLC0415

        self._solver = weakref.ref(solver)
        self.phase_tab = None
        if self.pipelined:
            self._phase_table(torch.cuda.current_stream(dev))
        self.bins = torch.zeros((K, self.F, self.S, 2), dtype=torch.float64, device=dev)
        # TwoSum compensation terms of the bins, float32 (csrc/mgh_dft.hip two_sum_acc)
        self.lo = torch.zeros(self.bins.shape, dtype=f32, device=dev) if self.compensated else None
        self._pending = 0
        self._attached = True

    # contraction slices: each wave keeps at least MIN_TILES frequency tiles
    # of 16 and a launch should hold at least MIN_ROUNDS rounds of resident
    # work-groups (so their load and matrix phases drift apart) and fill at
    # least MIN_FILL of them.  Measured on the 256^3 far-field monitor
    # (1001 frequencies, 256-step periods), ms per step: slices every 1 / 2 /
    # 4 / 8 / 16 / 32 steps 0.597 / 0.584 / 0.577 / 0.574 / 0.572 / 0.569,
    # one contraction per period 0.565.
    MIN_TILES = 8
    MIN_ROUNDS = 1.5
    MIN_FILL = 0.9

    def _occupancy(self) -> int:
        """Resident waves per SIMD of the contraction variant (its
        __launch_bounds__ in csrc/mgh_dft.hip)."""
        if self.blocks_per_launch > 1:
            return 3
        return 4 if block_pad(self.block) <= 256 else 2

    def _slicing(self, K: int, cus: int = 256) -> tuple[int, int]:
        """(steps between contraction slices c, frequency groups fsplit).

        A pipelined slice covers c / P of the previous period's sample waves
        (4 per work-group) for all frequencies, split into fsplit groups of
        work-groups.  Taken: the smallest power of two c dividing P for
        which some split keeps MIN_TILES tiles per work-group and gives
        MIN_ROUNDS rounds of resident work-groups (cus x waves per SIMD of
        them), MIN_FILL of them full; of those splits the one with the
        fewest groups (each group reads the samples once, and a wave's
        sample operand serves more tiles).  Without one, the fullest rounds."""
        tiles = self.F_pad // 16
        if not self.pipelined:
            return self.period, 1
        P = self.period
        waves = -(-self.S // 16)
        slots = cus * self._occupancy()
        cands = []
        cs = [c for c in (1 << k for k in range(12)) if c <= P and P % c == 0]
        if self.slice_steps_req is not None:
            cs = [int(self.slice_steps_req)]
            if cs[0] < 1 or P % cs[0]:
                raise ValueError(f"slice_steps must divide the {P}-step period, got {cs[0]}")
        splits = range(1, tiles + 1) if self.fsplit_req is None else [min(int(self.fsplit_req), tiles)]
        for c in cs:
            best = None
            for fs in splits:
                per = -(-tiles // fs)
                z = -(-tiles // per)
                if per < self.MIN_TILES and fs > 1 and self.fsplit_req is None:
                    continue
                wgs = -(-(-(-waves * c // P)) // 4) * K * z
                if wgs < self.MIN_ROUNDS * slots and c < P:
                    continue
                fill = wgs / (-(-wgs // slots) * slots)
                key = (True, -z, fill) if fill >= self.MIN_FILL else (False, fill, -z)
                if best is None or key > best[0]:
                    best = (key, z)
            if best is None:
                continue
            cands.append((c, best[1], best[0]))
            if best[0][0]:
                return c, best[1]
        c, z, _ = max(cands, key=lambda t: (t[2][1], -t[0]))
        return c, z

    def _phase_table(self, stream) -> None:
        """Block phases of every period from n_start to past the horizon."""
        P = self.period
        n_per = max(1, -(-(self._horizon - self.n_start) // P) + 1)
        self.n_periods = n_per
        self.phase_tab = torch.empty((n_per, self.blocks_per_launch, self.F, 4), dtype=torch.float64,
                                     device=self.omega.device)
        _lib.call("mgh_dft_phase_table", self.omega.data_ptr(), self.F, float(self.dt), self.n_start,
                  self.block, self.blocks_per_launch, n_per, self.phase_tab.data_ptr(),
                  int(stream.cuda_stream))

    @property
    def cadence(self) -> int:
        """Steps between contraction launches (a captured graph spans a
        multiple of it and replays at starts congruent to its capture)."""
        return self.slice_steps if self.pipelined else self.period

    @property
    def waves(self) -> int:
        return -(-self.S // 16)

    @property
    def period(self) -> int:
        """Steps between bin updates (sample rows of the block buffer)."""
        return self.block * self.blocks_per_launch

    def bytes(self) -> int:
        n = self.bins.numel() * (12 if self.compensated else 8)
        return int(n + self.X.numel() * 4)

    def record(self, solver, stream, n: int) -> None:
        sh = int(stream.cuda_stream)
        code = 4 if solver.e.dtype == torch.float32 else 8
        if self.grp_d is not None:
            _lib.call("mgh_dft_sample_groups", code, self.nsets, solver.e.data_ptr(), solver.e.shape[1],
                      solver.h.data_ptr(), solver.h.shape[1], self.grp_d.data_ptr(),
                      self.grp_d.shape[0], self.grp_blocks, self.wtab_d.data_ptr(), self.S,
                      self.X.data_ptr(), self.S_pad, self.period, self.x_buf,
                      solver.step_ctr.data_ptr(), self.n_start, sh)
        else:
            _lib.call("mgh_dft_sample", code, self.nsets, solver.e.data_ptr(), solver.e.shape[1],
                      solver.h.data_ptr(), solver.h.shape[1], self.idx_d.data_ptr(),
                      self.w_d.data_ptr(), self.S, self.X.data_ptr(), self.S_pad, self.period,
                      self.x_buf, solver.step_ctr.data_ptr(), self.n_start, sh)
        P = self.period
        if self.pipelined:
            c = self.slice_steps
            if (n - self.n_start) % c == c - 1:
                _lib.call("mgh_dft_slice", self.nsets, self.X.data_ptr(), self.x_buf, self.S,
                          self.S_pad, self.block, self.blocks_per_launch, self.Pc.data_ptr(),
                          self.Ps.data_ptr(), self.F, self.F_pad, self.phase_tab.data_ptr(),
                          self.n_periods, self.s_h_start, self.bins.data_ptr(), self._lo_ptr(),
                          solver.step_ctr.data_ptr(), self.n_start, c, self.fsplit, sh)
        elif (n - self.n_start) % P == P - 1:
            self._contract(solver, sh, P)

    def _lo_ptr(self):
        return self.lo.data_ptr() if self.lo is not None else None

    def _contract(self, solver, sh: int, nb: int) -> None:
        _lib.call("mgh_dft_blocks", self.nsets, self.X.data_ptr(), self.S, self.S_pad, self.block,
                  self.blocks_per_launch, nb,
                  self.Pc.data_ptr(), self.Ps.data_ptr(), self.omega.data_ptr(), self.F, self.F_pad,
                  float(self.dt), self.blk.data_ptr(), self.s_h_start, self.bins.data_ptr(),
                  self._lo_ptr(), solver.step_ctr.data_ptr(), self.n_start, sh)

    def _contract_range(self, sh: int, buf: int, nb: int, w_lo: int, w_hi: int, n0: int) -> None:
        _lib.call("mgh_dft_blocks_range", self.nsets, self.X[buf].data_ptr(), self.S, self.S_pad,
                  self.block, self.blocks_per_launch, nb, w_lo, w_hi, n0, self.Pc.data_ptr(),
                  self.Ps.data_ptr(), self.omega.data_ptr(), self.F, self.F_pad, float(self.dt),
                  self.blk.data_ptr(), self.s_h_start, self.bins.data_ptr(), self._lo_ptr(),
                  self.fsplit, sh)

    def finalize(self, solver) -> None:
        """Contract every recorded sample not yet in the bins; the sample
        buffers are then cleared, so a run that goes on from here adds each
        sample once (the counter stands at the number of completed steps)."""
        stream = torch.cuda.current_stream(solver.e.device)
        sh = int(stream.cuda_stream)
        P = self.period
        rel = solver.n_completed - 1 - self.n_start  # last recorded step
        if rel < 0:
            return
        p, r = divmod(rel, P)
        if self.pipelined:
            W = self.waves
            ns = P // self.slice_steps
            done = (r + 1) // self.slice_steps  # slices of period p - 1 already contracted
            if p >= 1 and done < ns:
                self._contract_range(sh, (p - 1) % 2, P, done * W // ns, W,
                                     self.n_start + (p - 1) * P)
            self._contract_range(sh, p % 2, r + 1, 0, W, self.n_start + p * P)
        elif r + 1 < P:  # a full period was contracted by its last step
            self._contract(solver, sh, r + 1)
        else:
            return
        self.X.zero_()

    def result(self, s: int = 0) -> dict:
        """{component: complex128 (F, *shape)} of field set s (host)."""
        b = self.bins[s].cpu().numpy()
        if self.lo is not None:
            b = b + self.lo[s].cpu().numpy().astype(np.float64)
        z = b[..., 0] + 1j * b[..., 1]
        out = {}
        for comp, shape, start, n in self._layout:
            out.setdefault(comp, []).append(natural(z[:, start:start + n], shape))
        return {k: (v[0] if len(v) == 1 else v) for k, v in out.items()}

    def state_dict(self) -> dict:
        sd = {"X": self.X.cpu().numpy(), "bins": self.bins.cpu().numpy(), "n_start": self.n_start}
        if self.lo is not None:
            sd["lo"] = self.lo.cpu().numpy()
        return sd

    def load_state_dict(self, sd: dict) -> None:
        self.X.copy_(torch.from_numpy(np.asarray(sd["X"])))
        self.bins.copy_(torch.from_numpy(np.asarray(sd["bins"])))
        if self.lo is not None:
            self.lo.copy_(torch.from_numpy(np.asarray(sd["lo"])))
        self._set_n_start(int(sd["n_start"]))

    def _set_n_start(self, n_start: int) -> None:
        if n_start != self.n_start:
            self.n_start = n_start
            if self.pipelined:
                self._phase_table(torch.cuda.current_stream(self.X.device))

    # per field set -------------------------------------------------------
    def set_state(self, s: int) -> dict:
        """Field set s's sample buffers and bins (exact resume)."""
        d = {"X": self.X[:, s].cpu().numpy(), "bins": self.bins[s].cpu().numpy(), "n_start": self.n_start}
        if self.lo is not None:
            d["lo"] = self.lo[s].cpu().numpy()
        return d

    def load_set_state(self, s: int, d: dict) -> None:
        self.X[:, s].copy_(torch.from_numpy(np.asarray(d["X"])))
        self.bins[s].copy_(torch.from_numpy(np.asarray(d["bins"])))
        if self.lo is not None:
            self.lo[s].copy_(torch.from_numpy(np.asarray(d["lo"])))
        self._set_n_start(int(d["n_start"]))

    def complete_bins(self, s: int = 0) -> np.ndarray:
        """(F, S) complex128 bins of field set s with every recorded sample
        in them: pending samples are contracted into a copy (the monitor's
        own state is left as it was)."""
        solver = self._solver()
        rel = solver.n_completed - self.n_start if solver is not None else 0
        if rel > 0 and solver is not None:
            keep = (self.X.clone(), self.bins.clone(), self.lo.clone() if self.lo is not None else None)
            self.finalize(solver)
            b = self.bins[s].cpu().numpy()
            lo = self.lo[s].cpu().numpy() if self.lo is not None else None
            self.X.copy_(keep[0])
            self.bins.copy_(keep[1])
            if self.lo is not None:
                self.lo.copy_(keep[2])
        else:
            b = self.bins[s].cpu().numpy()
            lo = self.lo[s].cpu().numpy() if self.lo is not None else None
        if lo is not None:
            b = b + lo.astype(np.float64)
        return b[..., 0] + 1j * b[..., 1]

    def _load_bins(self, s: int, z: np.ndarray) -> None:
        """Field set s's bins from (F, S) complex values; no samples pending."""
        b = np.zeros((self.F, self.S, 2))
        b[..., 0], b[..., 1] = z.real, z.imag
        self.bins[s].copy_(torch.from_numpy(b))
        if self.lo is not None:
            self.lo[s].zero_()
        self.X[:, s].zero_()

    def _component_bins(self, s: int) -> list:
        """[(comp, shape, (F, *shape) complex)] in sample-group order."""
        z = self.complete_bins(s)
        return [(comp, shape, natural(z[:, start:start + n], shape)) for comp, shape, start, n in self._layout]


class MonitorFieldFrequency(DFTMonitorBase):
    """DFT of cell-centred fields in a box region (field_frequency.py)."""

    def __init__(self, freqs, corners=None, fields=("E", "H"), name: str = "", region=None,
                 block: int = BLOCK, compensated: bool = True, blocks_per_launch: int | None = None,
                 pipelined: bool = True, fsplit: int | None = None, slice_steps: int | None = None):
        super().__init__(freqs, name, block, compensated, blocks_per_launch, pipelined, fsplit,
                         slice_steps)
        self.corners = corners
        self.region = region  # optional explicit (slice, slice, slice) of cells
        comps = []
        for f in fields:
            comps += [f + a for a in "xyz"] if f in ("E", "H") else [f]
        self.components = [c for c in COMPS if c in comps]

    def _build(self, grid, boundary_conditions) -> None:
        if self.region is not None:
            sl = tuple(self.region)
        else:
            cc = grid.centres()
            sl = []
            for a in range(3):
                if self.corners is None:
                    sl.append(slice(0, cc[a].size))
                else:
                    lo = min(self.corners[0][a], self.corners[1][a])
                    hi = max(self.corners[0][a], self.corners[1][a])
                    sl.append(snap_range(cc[a], lo, hi))
            sl = tuple(sl)
        self.slices = sl
        cc = grid.centres()
        self._centres = tuple(np.asarray(cc[a][sl[a]], float) for a in range(3))
        sb = _SampleBuilder(grid)
        I, J, K = np.meshgrid(np.arange(sl[0].start, sl[0].stop), np.arange(sl[1].start, sl[1].stop),
                              np.arange(sl[2].start, sl[2].stop), indexing="ij")
        shape = I.shape
        # samples numbered x fastest, like the device field layout, so the
        # per-step gathers of neighbouring samples coalesce
        I, J, K = (a.transpose(2, 1, 0) for a in (I, J, K))
        self._groups = [(c, shape, sb.centred(c, I, J, K)) for c in self.components]


    def result_dump(self, s: int = 0) -> dict:
        """The reference's dump of this monitor (field_frequency.py:304-331):
        region cell centres, frequencies and the raw complex bins per
        component of field set s, every recorded sample included."""
        bins = {comp: b for comp, _, b in self._component_bins(s)}
        return {"components": list(self.components), "fields": sorted({c[0] for c in self.components}),
                "interval": None, "symmetry": [], "freqs": np.asarray(self.freqs, float),
                "corners": _corners_array(self.corners), "grid_x": self._centres[0],
                "grid_y": self._centres[1], "grid_z": self._centres[2], "bins": bins}

    def load_result_dump(self, dump: dict, s: int = 0) -> None:
        """Field set s's bins from the reference's dump (resume)."""
        z = np.zeros((self.F, self.S), complex)
        for comp, shape, start, n in self._layout:
            z[:, start:start + n] = unnatural(np.asarray(dump["bins"][comp]))
        self._load_bins(s, z)


class MonitorFarField(DFTMonitorBase):
    """Huygens-box DFT of the tangential fields (far_field.py:104-248).

    The box sits ``margin_cells`` inside every CPML face (absorber depth from
    the solver's CPML boundaries).  Per open face and tangential component
    the sample is (1 - w) v[plane-1] + w v[plane] of the cell-centred values
    of the two cell layers around the box node plane.
    """

    def __init__(self, freqs, margin_cells: int = 3, name: str = "", block: int = BLOCK,
                 compensated: bool = True, blocks_per_launch: int | None = None,
                 pipelined: bool = True, fsplit: int | None = None, slice_steps: int | None = None):
        super().__init__(freqs, name, block, compensated, blocks_per_launch, pipelined, fsplit,
                         slice_steps)
        if margin_cells < 1:
            raise ValueError("margin_cells must be at least 1")
        self.margin_cells = margin_cells
        self._source_spectrum = None
        self._accepted_power = None

    def _build(self, grid, boundary_conditions) -> None:
        from .boundaries import CPMLBoundary  # noqa: PLC0415

        n_cells = grid.dims
        nodes = (grid.x, grid.y, grid.z)
        lo_n = [0, 0, 0]
        hi_n = list(n_cells)
        open_faces = []
        for face, bc in boundary_conditions.items():
            if isinstance(bc, CPMLBoundary):
                axis = "xyz".index(face[0])
                idx = bc.thickness_cells + self.margin_cells
                if face.endswith("min"):
                    lo_n[axis] = idx
                else:
                    hi_n[axis] = n_cells[axis] - idx
                open_faces.append(face)
        if not open_faces:
            raise ValueError("far-field monitor needs at least one CPML face")
        sb = _SampleBuilder(grid)
        centres = grid.centres()
        self._groups = []
        self.faces = []
        self._geometry = {}
        # walls handled by image theory in the far-field post (far_field.py:150-162)
        self._image_planes = []
        for face, bc in boundary_conditions.items():
            kind = getattr(bc, "kind", type(bc).__name__.replace("Boundary", ""))
            if str(kind).upper() in ("PEC", "PMC"):
                axis = "xyz".index(face[0])
                at_low = face.endswith("min")
                self._image_planes.append({"axis": axis, "kind": str(kind).upper(), "at_low": at_low,
                                           "position": wall_position(face, str(kind).upper(), nodes[axis]),
                                           "physical_halfspace": True})
        for face in open_faces:
            axis = "xyz".index(face[0])
            nn = lo_n[axis] if face.endswith("min") else hi_n[axis]
            t1, t2 = [a for a in range(3) if a != axis]
            cc = centres[axis]
            wgt = float((nodes[axis][nn] - cc[nn - 1]) / (cc[nn] - cc[nn - 1]))
            r1 = np.arange(lo_n[t1], hi_n[t1])
            r2 = np.arange(lo_n[t2], hi_n[t2])
            A, Bm = np.meshgrid(r1, r2, indexing="ij")
            shape = A.shape
            A, Bm = A.T, Bm.T  # samples numbered along the lower axis fastest
            self.faces.append((face, axis, nn, wgt, shape))
            widths = (grid.dx, grid.dy, grid.dz)
            self._geometry[face] = dict(axis=axis, sign=-1.0 if face.endswith("min") else 1.0,
                                        plane=float(nodes[axis][nn]),
                                        c1=centres[t1][r1].copy(), c2=centres[t2][r2].copy(),
                                        w1=np.asarray(widths[t1])[r1].copy(),
                                        w2=np.asarray(widths[t2])[r2].copy())
            for comp in TANGENTIALS[axis]:
                terms = []
                for layer, lw in ((nn - 1, 1.0 - wgt), (nn, wgt)):
                    ijk = [None, None, None]
                    ijk[axis] = np.full_like(A, layer)
                    ijk[t1], ijk[t2] = A, Bm
                    for ii, ww in sb.centred(comp, *ijk):
                        terms.append((ii, ww * lw))
                self._groups.append((comp, shape, terms))

    def result_dump(self, s: int = 0) -> dict:
        """The reference's dump of this monitor (far_field.py:403-452): per
        Huygens face its geometry and the raw complex bins of field set s,
        every recorded sample included — the input of the far-field post."""
        per_face: dict = {}
        k = {}
        comps = self._component_bins(s)
        by_comp: dict = {}
        for comp, _, b in comps:
            by_comp.setdefault(comp, []).append(b)
        for face, axis, _, _, _ in self.faces:
            for comp in TANGENTIALS[axis]:
                i = k.get(comp, 0)
                k[comp] = i + 1
                per_face.setdefault(face, {})[comp] = by_comp[comp][i]
        faces = []
        for face, axis, _, _, _ in self.faces:
            g = self._geometry[face]
            faces.append({"name": face, "axis": int(g["axis"]), "sign": g["sign"], "plane": g["plane"],
                          "c1": g["c1"], "c2": g["c2"], "w1": g["w1"], "w2": g["w2"],
                          "bins": per_face[face]})
        dump = {"name": self.name, "freqs": np.asarray(self.freqs, float),
                "margin_cells": int(self.margin_cells), "faces": faces,
                "image_planes": [dict(p) for p in self._image_planes]}
        if self._source_spectrum is not None:
            dump["source_spectrum"] = np.asarray(self._source_spectrum)
        if self._accepted_power is not None:
            dump["accepted_power"] = np.asarray(self._accepted_power)
        return dump

    def renormalize(self, source_signal, dt: float | None = None) -> None:
        """Per-watt normalisation (monitors/far_field.py:250-267): store the
        source's DFT in the bins' own convention as the divisor of the
        far-field patch sets; the bins stay as recorded.  source_signal is
        an object with ``values`` and ``dt`` or the sample array (with dt)."""
        vals = getattr(source_signal, "values", source_signal)
        step = getattr(source_signal, "dt", dt)
        if step is None:
            raise ValueError("renormalize needs the source's time step")
        self._source_spectrum = source_spectrum(vals, float(step), self.freqs)

    @property
    def is_renormalized(self) -> bool:
        return self._source_spectrum is not None

    def set_accepted_power(self, f_axis, accepted) -> None:
        """Accepted input power of the run on the monitor frequencies (for gain)."""
        self._accepted_power = np.interp(self.freqs, np.asarray(f_axis, float), np.asarray(accepted, float))

    def far_field(self, f: float | None = None, *, f_index: int | None = None, theta=None, phi=None,
                  s: int = 0):
        """Far-field pattern of field set s at one recorded frequency, the
        transform summed on the device (the reference monitor's ``result``,
        monitors/far_field.py:357-391)."""
        solver = self._solver()
        dev = solver.e.device if solver is not None else None
        return far_field_from_dump(self.result_dump(s), f, f_index=f_index, theta=theta, phi=phi,
                                   device=dev)

    def load_result_dump(self, dump: dict, s: int = 0) -> None:
        """Field set s's bins from the reference's dump (resume)."""
        by_face = {str(f["name"]): f["bins"] for f in (dump["faces"].values()
                                                      if isinstance(dump["faces"], dict)
                                                      else dump["faces"])}
        z = np.zeros((self.F, self.S), complex)
        k = {}
        order = []
        for face, axis, _, _, _ in self.faces:
            for comp in TANGENTIALS[axis]:
                order.append((face, comp))
        # groups were laid out E components first, then H, each in face order
        e_first = [fc for fc in order if fc[1].startswith("E")] + [fc for fc in order if fc[1].startswith("H")]
        for (face, comp), (c2, shape, start, n) in zip(e_first, self._layout):
            assert comp == c2
            z[:, start:start + n] = unnatural(np.asarray(by_face[face][comp]))
        self._load_bins(s, z)
