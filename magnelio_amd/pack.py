"""Packed stepping of small solvers: several jobs' steps in shared launches.

A small grid fills a sliver of the device: a 32^3 job's curl launch is a
few hundred waves, and each of its steps is about seven launches (E curl,
the four port launches, H curl, step counter) whose fixed cost outweighs
their work.  ``PackedGroup`` advances several set-up solvers together, each
step as one E-curl launch over all of their grids, the port steps of all of
their ports in the four port launches (eight ports per launch group), one
H-curl launch and one launch advancing every step counter
(``mgh_update_e_packed``, ``mgh_port_step_packed``, ``mgh_update_h_packed``,
``mgh_step_advance_packed``).  Every kernel reads each solver's own buffers
and step counter, so a packed solver's fields and port signals equal the
same solver stepped alone, bit for bit.

A solver can be packed when its step is exactly those launches: modal ports
(or none), no CPML, sources, lumped elements, ADE/SIBC completions or
monitors, the stencil's own operation order (no float64 curl accumulator).
Dead-unit skipping is dropped (dense launches give the same values).
Members share one precision.  Steps are captured as a HIP graph per
membership once the first two steps ran eagerly; a member leaves a group by
the group being rebuilt without it.

The reference has no counterpart: it runs one solver at a time.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib, kernels

MAX_JOBS = 16  # grids per packed curl launch
PORTS_PER_LAUNCH = 8  # ports per packed port step
MAX_COUNTERS = 64


def packable(solver) -> bool:
    """Whether a set-up solver's step is the plain curl + port + counter
    sequence that ``PackedGroup`` reproduces."""
    return (solver._setup_done and solver.cpml is None and not solver.sources
            and solver.dlumped is None and solver._dispersion is None
            and solver._dispersion_mu is None and solver._sibc is None
            and not solver.monitors and not solver.curl64)


class PackedGroup:
    """Steps the given set-up solvers together (see the module docstring).

    ``shape``: launch shape word of the packed curl launches (its sets part
    must be 0 or 1); ``use_graphs``: replay captured steps.
    """

    def __init__(self, solvers, shape: int = kernels.DEFAULT_KCHUNK, use_graphs: bool = True):
        self.solvers = list(solvers)
        n = len(self.solvers)
        if not 1 <= n <= MAX_JOBS:
            raise ValueError(f"a packed group holds 1..{MAX_JOBS} solvers, got {n}")
        for s in self.solvers:
            if not packable(s):
                raise ValueError("solver cannot be packed (set it up first; CPML, sources, lumped "
                                 "elements, dispersion, SIBC, monitors and curl64 are not packed)")
        dtypes = {s.e.dtype for s in self.solvers}
        devices = {s.e.device for s in self.solvers}
        if len(dtypes) != 1 or len(devices) != 1:
            raise ValueError("packed solvers must share one precision and one device")
        if kernels.shape_parts(shape)[1] > 1:
            raise ValueError("packed curl launches take one field set per thread")
        self.shape = int(shape)
        self.code = 4 if dtypes.pop() == torch.float32 else 8
        self.device = devices.pop()
        self.use_graphs = bool(use_graphs)
        self._graph = None
        self._graph_len = 0
        self._build_tables()

    # ------------------------------------------------------------------
    def _build_tables(self) -> None:
        """Host int64 tables the launches read (kept alive with the group:
        the entry points copy them into the launch arguments)."""
        rows_e, rows_h = [], []
        for s in self.solvers:
            nx, ny, nz = s.dims
            base = [nx, ny, nz, s.nsets, s.e.data_ptr(), s.e.shape[1], s.h.data_ptr(), s.h.shape[1]]
            rows_e.append(base + [s.alpha_e.data_ptr(), s.beta_e.data_ptr()])
            rows_h.append(base + [s.alpha_h.data_ptr(), s.beta_h.data_ptr()])
        self._curl_e = np.ascontiguousarray(rows_e, dtype=np.int64)
        self._curl_h = np.ascontiguousarray(rows_h, dtype=np.int64)
        descs, jobs = [], []
        for s in self.solvers:
            if s.dports is None:
                continue
            words = s.dports.WORDS
            for pi in range(len(s.ports)):
                descs.append(s.dports.desc[pi * words:(pi + 1) * words])
                jobs.append([s.e.data_ptr(), s.e.shape[1], s.h.data_ptr(), s.h.shape[1], s.nsets,
                             s.step_ctr.data_ptr()])
        self._port_groups = []
        for lo in range(0, len(descs), PORTS_PER_LAUNCH):
            d = np.ascontiguousarray(np.concatenate(descs[lo:lo + PORTS_PER_LAUNCH]), dtype=np.int64)
            j = np.ascontiguousarray(jobs[lo:lo + PORTS_PER_LAUNCH], dtype=np.int64)
            self._port_groups.append((len(j), d, j))
        ctrs = [s.step_ctr.data_ptr() for s in self.solvers]
        self._ctrs = [np.ascontiguousarray(ctrs[lo:lo + MAX_COUNTERS], dtype=np.int64)
                      for lo in range(0, len(ctrs), MAX_COUNTERS)]

    def _step(self) -> None:
        """Enqueue one step of every member on the current stream."""
        sh = int(torch.cuda.current_stream(self.device).cuda_stream)
        n = len(self.solvers)
        _lib.call("mgh_update_e_packed", self.code, n, self._curl_e.ctypes.data, self.shape, sh)
        for k, d, j in self._port_groups:
            _lib.call("mgh_port_step_packed", self.code, k, d.ctypes.data, j.ctypes.data, sh)
        _lib.call("mgh_update_h_packed", self.code, n, self._curl_h.ctypes.data, self.shape, sh)
        for c in self._ctrs:
            _lib.call("mgh_step_advance_packed", c.size, c.ctypes.data, 1, sh)

    @property
    def remaining(self) -> int:
        """Steps every member can still take (the nearest horizon)."""
        return min(s.horizon - s.n_completed for s in self.solvers)

    def launch(self, n_steps: int) -> int:
        """Enqueue up to ``n_steps`` steps of every member on the current
        stream (no more than the nearest member's horizon allows) and return
        the steps enqueued; the caller synchronises."""
        n = max(0, min(int(n_steps), self.remaining))
        first = min(s.n_completed for s in self.solvers)
        k = 0
        while k < n:
            G = self._graph_len
            if self.use_graphs and first + k >= 2 and self._graph is None and n - k >= 8:
                G = self._capture(min(n - k, 64))
            if self._graph is not None and n - k >= G:
                self._graph.replay()
                k += G
                continue
            self._step()
            k += 1
        for s in self.solvers:
            s._resume_step += n
            if s._resume_step >= s.horizon:
                s._stop_reason = "steps"
        return n

    def _capture(self, G: int) -> int:
        """Capture G steps as one graph (recorded only; replayed next)."""
        try:
            cur = torch.cuda.current_stream(self.device)
            cur.synchronize()
            g = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream(device=self.device)
            side.wait_stream(cur)
            with torch.cuda.stream(side):
                g.capture_begin()
                for _ in range(G):
                    self._step()
                g.capture_end()
            cur.wait_stream(side)
            side.synchronize()
            self._graph, self._graph_len = g, G
        except Exception as exc:  # pragma: no cover - depends on the runtime
            import warnings  # noqa: PLC0415

            warnings.warn(f"packed graph capture failed ({type(exc).__name__}: {exc}); eager steps")
            self.use_graphs = False
            self._graph, self._graph_len = None, 0
        return self._graph_len
