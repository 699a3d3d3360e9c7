"""S-parameter sweeps: every excited (port, mode) channel of a mesh as one
field set of a batched solver.

Reference: ``AnalysisScatteringTD`` runs one full FIT-TD simulation per
excited channel, sharing dt, coefficients and port models
(src/magnelio/analysis/scattering_td.py:1252-1273, ``_run_one_excitation``
:1490-1582), and merges the K single-column results into the S-matrix
(``SParameterResult.merge``, src/magnelio/post/sparameter_result.py:449-495).

Here the K runs are K field sets of one :class:`FITTimeDomainSolver`
(``nsets = K``): every curl launch reads the shared coefficients once for
all K sets, the port step handles all (port, set) pairs in one launch, and
the steps are replayed from HIP graphs.  ``batch`` caps the sets per solver
(``batch=1`` is the one-excitation-per-launch path, kept for comparison);
results are independent of it, bit for bit, because every kernel treats the
sets independently.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import autotune, config
from .boundaries import PECBoundary
from .signals import resolve_excitation, s_parameters
from .solver import ChannelExcitation, FITTimeDomainSolver


@dataclass
class SParameterResult:
    """S[f, observed, excited] with (port, mode) channel labels
    (sparameter_result.py:32-140)."""

    f_axis: np.ndarray
    channels: tuple
    excitations: tuple
    matrix: np.ndarray

    def __post_init__(self) -> None:
        f = np.asarray(self.f_axis, dtype=float)
        if f.ndim != 1 or f.size == 0:
            raise ValueError(f"f_axis must be 1D with at least one frequency; got shape {f.shape}.")
        if np.any(f <= 0.0):
            raise ValueError("f_axis must contain only positive frequencies.")
        self.channels = tuple(tuple(c) for c in self.channels)
        self.excitations = tuple(tuple(c) for c in self.excitations)
        if not self.channels or not self.excitations:
            raise ValueError("channels and excitations must be non-empty.")
        if len(set(self.channels)) != len(self.channels):
            raise ValueError("channels must be unique.")
        if len(set(self.excitations)) != len(self.excitations):
            raise ValueError("excitations must be unique.")
        missing = [e for e in self.excitations if e not in self.channels]
        if missing:
            raise ValueError(f"excitation {missing[0]!r} not in channels {sorted(self.channels)}.")
        expected = (f.size, len(self.channels), len(self.excitations))
        if np.shape(self.matrix) != expected:
            raise ValueError(f"matrix shape {np.shape(self.matrix)} does not match expected {expected} "
                             "= (n_frequencies, n_channels, n_excitations).")
        self.f_axis = f

    @property
    def n_frequencies(self) -> int:
        return int(self.f_axis.size)

    @property
    def n_channels(self) -> int:
        return len(self.channels)

    @property
    def n_excitations(self) -> int:
        return len(self.excitations)

    @property
    def is_complete(self) -> bool:
        return set(self.excitations) == set(self.channels)

    def S(self, observed: str, excited: str, observed_mode: int = 0, excited_mode: int = 0) -> np.ndarray:
        i = self.channels.index((observed, observed_mode))
        j = self.excitations.index((excited, excited_mode))
        return self.matrix[:, i, j]

    def db(self, observed: str, excited: str, observed_mode: int = 0, excited_mode: int = 0) -> np.ndarray:
        return 20.0 * np.log10(np.abs(self.S(observed, excited, observed_mode, excited_mode)))

    @classmethod
    def merge(cls, results: list["SParameterResult"]) -> "SParameterResult":
        """Columns of K single-excitation results, in list order."""
        if not results:
            raise ValueError("results list is empty.")
        first = results[0]
        for k, r in enumerate(results):
            if not np.array_equal(r.f_axis, first.f_axis):
                raise ValueError(f"results[{k}].f_axis differs from results[0].f_axis.")
            if r.channels != first.channels:
                raise ValueError(f"results[{k}].channels {r.channels} differs from "
                                 f"results[0].channels {first.channels}.")
        exc = [e for r in results for e in r.excitations]
        if len(set(exc)) != len(exc):
            raise ValueError(f"merge: duplicate excitations across inputs: {exc}")
        return cls(first.f_axis.copy(), first.channels, tuple(exc),
                   np.concatenate([r.matrix for r in results], axis=2))


@dataclass
class SweepResult:
    s_params: SParameterResult
    signals: dict  # {excited channel: {(port, mode): (V, I)}}
    dt: float
    n_steps: int
    seconds: float  # device wall time of the marching (set-up excluded)
    batches: list = field(default_factory=list)  # excited channels per solver
    n_cells: int = 0
    # error of the single-precision S-parameters against a float64 run of
    # the same kernels (ScatteringSweep fp64_check), or None
    fp64: dict | None = None

    @property
    def cell_updates(self) -> float:
        """Cell-updates of the sweep: cells x steps x excitations."""
        return float(self.n_cells) * self.n_steps * sum(len(b) for b in self.batches)


FP64_CHECKS = ("off", "column", "all")


class ScatteringSweep:
    """Batched S-parameter sweep over a mesh with modal ports.

    ``excited``: channels to excite (default: every (port, mode)); each gets
    the reference's auto waveform for its mode (``resolve_excitation``,
    scattering_td.py:2322-2366) unless ``waveforms`` maps it to a callable.
    ``batch``: field sets per solver (None = all channels in one solver;
    "auto": timed at the first run, ``autotune.batch_for``).
    ``kchunk``: the solvers' curl launch shape (a ``kernels.launch_shape``
    word or "auto"; None: the solver's default).
    ``fp64_check``: "off", "column" (the first excited channel) or "all" —
    those excitations are run again in float64 (same ports, waveforms and
    kernels) and the result carries max |S_fp32 - S_fp64| over every
    frequency and observed channel of those columns (``SweepResult.fp64``).
    """

    def __init__(self, mesh, ports, *, dt: float, n_steps: int, f_axis, f_max: float,
                 f_min: float = 0.0, excited=None, waveforms=None, boundary_conditions=None,
                 precision: str | None = None, batch: int | str | None = None,
                 use_graphs: bool | None = None,
                 monitors_factory=None, device: str = "cuda:0", graph_block: int = 32,
                 fp64_check: str = "off", kchunk=None):
        self.mesh, self.ports = mesh, list(ports)
        self.dt, self.n_steps = float(dt), int(n_steps)
        self.f_axis = np.asarray(f_axis, dtype=float)
        self.f_max, self.f_min = float(f_max), float(f_min)
        self.channels = tuple((p.name, m) for p in self.ports for m in range(p.n_modes))
        self.excited = tuple(excited) if excited is not None else self.channels
        for ch in self.excited:
            if ch not in self.channels:
                raise ValueError(f"excited channel {ch!r} is not a port mode of this mesh")
        self.waveforms = dict(waveforms or {})
        if boundary_conditions is None:
            port_faces = {p.face for p in self.ports}
            boundary_conditions = {f: PECBoundary(f) for f in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")
                                   if f not in port_faces}
        self.boundary_conditions = boundary_conditions
        # explicit arguments win; else MAGNELIO_PRECISION / MAGNELIO_GPU_GRAPHS,
        # else single precision with graphs
        self.precision = config.precision(precision)
        self.batch = len(self.excited) if batch is None else batch if batch == "auto" else int(batch)
        if self.batch != "auto" and self.batch < 1:
            raise ValueError("batch must be >= 1 or 'auto'")
        self.use_graphs = config.graphs(use_graphs, default=True)
        self.monitors_factory = monitors_factory
        self.device = device
        self.graph_block = int(graph_block)
        if fp64_check not in FP64_CHECKS:
            raise ValueError(f"fp64_check must be one of {FP64_CHECKS}, got {fp64_check!r}")
        if fp64_check != "off" and self.precision != "single":
            raise ValueError("fp64_check applies to single-precision sweeps")
        self.fp64_check = fp64_check
        self.kchunk = kchunk  # the solvers' curl launch shape (None: the solver's default)
        self.solvers: list[FITTimeDomainSolver] = []

    def waveform(self, ch):
        if ch in self.waveforms:
            return self.waveforms[ch]
        port = next(p for p in self.ports if p.name == ch[0])
        return resolve_excitation(port.modes[ch[1]].omega_c, self.f_min, self.f_max, ch[1])

    def make_solver(self, chans, precision: str | None = None,
                    monitors: bool = True) -> FITTimeDomainSolver:
        exc = [ChannelExcitation(p, m, self.waveform((p, m))) for p, m in chans]
        mons = (list(self.monitors_factory())
                if monitors and self.monitors_factory is not None else [])
        return FITTimeDomainSolver(mesh=self.mesh, boundary_conditions=self.boundary_conditions,
                                   ports=self.ports, total_time_steps=self.n_steps, dt=self.dt,
                                   precision=precision or self.precision, nsets=len(chans),
                                   excitations=exc, use_graphs=self.use_graphs, monitors=mons,
                                   device=self.device, graph_block=self.graph_block,
                                   **({} if self.kchunk is None else {"kchunk": self.kchunk}))

    def check_channels(self) -> tuple:
        """Excitations the float64 check runs again (empty when off)."""
        if self.fp64_check == "off":
            return ()
        return self.excited[:1] if self.fp64_check == "column" else self.excited

    def make_check_solver(self) -> FITTimeDomainSolver | None:
        """The float64 solver of the check channels (no DFT monitors: only
        the port signals are compared), or None when the check is off."""
        chans = self.check_channels()
        return self.make_solver(chans, precision="double", monitors=False) if chans else None

    def fp64_error(self, s32: SParameterResult, check: FITTimeDomainSolver) -> dict:
        """max |S_fp32 - S_fp64| per checked column and over all of them."""
        cols = []
        worst = 0.0
        for k, ch in enumerate(self.check_channels()):
            c64 = self.s_column(check.recorded(k), ch).matrix[:, :, 0]
            c32 = s32.matrix[:, :, s32.excitations.index(ch)]
            d = float(np.max(np.abs(c32 - c64)))
            cols.append({"excited": list(ch), "max_abs_dS": d})
            worst = max(worst, d)
        return {"mode": self.fp64_check, "columns": cols, "max_abs_dS": worst,
                "db": 20.0 * math.log10(worst) if worst > 0 else -math.inf}

    def s_column(self, signals: dict, ch) -> SParameterResult:
        modes = {p.name: p.modes for p in self.ports}
        lines = {(p.name, m): lp for p in self.ports for m, lp in p.dtbc_line_params.items()}
        ndx = {p.name: p.plane.normal_dx for p in self.ports}
        s = s_parameters(signals, modes, ch, self.f_axis, self.dt, port_normal_dx=ndx,
                         port_line_params=lines)
        mat = np.stack([s[c] for c in self.channels], axis=1)[:, :, None]
        return SParameterResult(self.f_axis, self.channels, (ch,), mat)

    def run(self) -> SweepResult:
        if self.batch == "auto":
            self.batch = autotune.batch_for(self)
        batches = [self.excited[i:i + self.batch] for i in range(0, len(self.excited), self.batch)]
        solvers = [self.make_solver(b) for b in batches]
        for s in solvers:
            s.setup()
        dev = torch.device(self.device)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        for s in solvers:
            s.run()
        torch.cuda.synchronize(dev)
        seconds = time.perf_counter() - t0
        self.solvers = solvers
        cols, sigs = [], {}
        for s, b in zip(solvers, batches):
            for k, ch in enumerate(b):
                rec = s.recorded(k)
                sigs[ch] = rec
                cols.append(self.s_column(rec, ch))
        res = SweepResult(SParameterResult.merge(cols), sigs, self.dt, self.n_steps, seconds,
                          [list(b) for b in batches], int(np.prod(self.mesh.dims)))
        check = self.make_check_solver()
        if check is not None:
            check.setup()
            check.run()
            torch.cuda.synchronize(dev)
            res.fp64 = self.fp64_error(res.s_params, check)
        return res
