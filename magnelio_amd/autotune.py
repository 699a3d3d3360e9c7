"""Launch-shape tuning of the curl launches, cached on disk.

The shape of a curl launch is three numbers (``kernels.launch_shape``): the
z-chunk (planes a thread marches, keeping the z neighbour in registers —
shorter chunks give more work-groups, longer ones fewer re-read chunk-start
planes), the field sets a thread advances (each coefficient word it loads
serves them all; more sets cost registers and occupancy) and the XCD band
(consecutive launch units a chiplet takes, which share its L2).  The best
shape depends on the grid, K and the CPML the launch carries, so it is
measured per configuration.

``FITTimeDomainSolver(kchunk="auto")`` calls :func:`shape_for` at setup,
before any state is loaded: the curl launches (with the solver's CPML, as a
step runs them) are timed on the solver's own zeroed buffers — zero fields
and zero psi stay zero, so the timing changes nothing — and the fastest
candidate is stored in a JSON cache (``MAGNELIO_HIP_AUTOTUNE_CACHE``, default
``~/.cache/magnelio_amd/autotune.json``) keyed by device, the library's
content hash, grid, K, dtype and CPML faces; later solvers of the same key
read it.  Every shape gives the same fields bit for bit.

``ScatteringSweep(batch="auto")`` calls :func:`batch_for`: the number of
excitations per solver (K) is picked the same way.  For each batch size
``ceil(n / m)`` (the smallest largest batch of m solvers) one solver of that
many excitations is set up and its full step (curl, ports, monitors, as a
captured graph when the sweep uses graphs) is timed; the sweep's step costs
``ceil(n / k)`` such steps, and the cheapest k is cached under the grid, n,
precision, CPML faces, port channels and monitors.  Every batch size gives
the same S-parameters bit for bit (the kernels treat the sets
independently).

The reference's CUDA tier has a fixed block shape (``numba_kernels.py:446``)
and runs one excitation at a time (``scattering_td.py:1252-1273``); this
module has no counterpart there.
"""

from __future__ import annotations

import hashlib
import json
import os
import tempfile
from functools import lru_cache
from pathlib import Path

import torch

from . import _lib, kernels

KCHUNKS = (1, 2, 3, 4)
SETS = (1, 2, 3)  # for K >= 2; one set is always 1
BANDS = (4, 8, 16)
REPS = 5


def candidates(nsets: int) -> list[int]:
    """Shape words tried for K = nsets field sets."""
    sets = SETS if nsets >= 2 else (1,)
    return [kernels.launch_shape(kc, s, b) for kc in KCHUNKS for s in sets for b in BANDS]


def cache_path() -> Path:
    env = os.environ.get("MAGNELIO_HIP_AUTOTUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "magnelio_amd" / "autotune.json"


def _load(path: Path) -> dict:
    try:
        with open(path) as fh:
            data = json.load(fh)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(path: Path, data: dict) -> None:
    """Write the cache atomically (a concurrent writer loses only its own entry)."""
    path.parent.mkdir(parents=True, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=path.parent, prefix=".autotune-", suffix=".json")
    with os.fdopen(fd, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


@lru_cache(maxsize=1)
def library_id() -> str:
    """Content hash of the loaded kernel library: a rebuilt library is a new
    cache key."""
    return hashlib.sha1(Path(_lib.LIB_PATH).read_bytes()).hexdigest()[:12]


def cache_key(device: torch.device, dims, nsets: int, dtype, cpml_faces) -> str:
    props = torch.cuda.get_device_properties(device)
    arch = getattr(props, "gcnArchName", props.name)
    faces = ",".join(cpml_faces) if cpml_faces else "-"
    return (f"{arch}|lib:{library_id()}|{dims[0]}x{dims[1]}x{dims[2]}|K{nsets}|"
            f"{str(dtype).replace('torch.', '')}|cpml:{faces}")


def time_curl(e, h, ae, be, ah, bh, dims, kchunk: int, cpml=None, curl64: bool = False,
              stream=None, reps: int = REPS) -> float:
    """Mean milliseconds of one E + H curl step (with ``cpml`` as the step
    launches it) at launch shape ``kchunk``, after one untimed step."""
    stream = stream or torch.cuda.current_stream(e.device)

    def step():
        kernels.update_e(e, h, ae, be, dims, curl64=curl64, kchunk=kchunk, cpml=cpml,
                         stream=stream)
        kernels.update_h(e, h, ah, bh, dims, kchunk=kchunk, cpml=cpml, stream=stream)

    step()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        step()
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / reps


def choose(timings: dict) -> int:
    """Fastest shape; ties go to the shorter chunk, then fewer sets per
    thread, then the smaller band."""
    return min(sorted(timings), key=lambda w: (timings[w], kernels.shape_parts(w)))


def shape_for(solver, path: Path | None = None) -> int:
    """Launch shape word for ``solver`` (set up, fields still zero): from the
    cache, or timed now and stored."""
    path = path or cache_path()
    cpml = solver.cpml if (solver.cpml is not None and solver._dispersion is None
                          and solver._dispersion_mu is None and solver._sibc is None) else None
    faces = [f.face for f in solver.cpml.faces] if solver.cpml is not None else []
    key = cache_key(solver.e.device, solver.dims, solver.nsets, solver.e.dtype, faces)
    cands = candidates(solver.nsets)
    data = _load(path)
    hit = data.get(key)
    if isinstance(hit, dict) and hit.get("shape") in cands:
        return int(hit["shape"])
    timings = {w: time_curl(solver.e, solver.h, solver.alpha_e, solver.beta_e, solver.alpha_h,
                            solver.beta_h, solver.dims, w, cpml=cpml, curl64=solver.curl64)
               for w in cands}
    best = choose(timings)
    kc, sets, band = kernels.shape_parts(best)
    data = _load(path)  # re-read: another process may have added entries meanwhile
    data[key] = {"shape": best, "kchunk": kc, "sets": sets, "band": band,
                 "ms": {str(k): round(v, 4) for k, v in timings.items()}}
    _store(path, data)
    return best


BATCH_STEPS = 64  # steps timed per candidate batch (whole graphs where graphs are used)


def batch_candidates(n: int) -> list[int]:
    """Batch sizes timed for n excitations, largest first: ceil(n / m) for
    m = 1..n solvers."""
    return sorted({-(-n // m) for m in range(1, n + 1)}, reverse=True)


def time_steps(solver, steps: int = BATCH_STEPS) -> float:
    """Mean milliseconds of one full step of a set-up solver (its own
    launches; graph replays where the solver uses graphs), after two eager
    steps and one graph's capture and replay.  Advances the solver."""
    dev = solver.e.device
    stream = torch.cuda.current_stream(dev)
    left = solver.horizon - solver.n_completed - 1  # stay short of the monitors' finalize
    warm = min(2, left)
    solver.launch(warm)
    left -= warm
    G = solver.graph_steps if solver.use_graphs else 1
    if solver.use_graphs and left >= 2 * G:
        solver.launch(G)  # capture and first replay, untimed
        left -= G
        n = G * max(1, min(left // G, -(-steps // G)))
    else:
        n = min(left, steps)
    if n < 1:
        raise ValueError("the sweep is too short to time its steps")
    stream.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    done = solver.launch(n)
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / done


def choose_batch(n: int, per_step: dict) -> tuple[int, dict]:
    """Cheapest batch for n excitations from per-step times of one solver of
    k excitations: a sweep step costs ceil(n / k) of them (the last solver
    may hold fewer sets; it is costed as a full one).  Ties go to the larger
    batch (fewer solvers).  Returns the batch and every candidate's cost."""
    cost = {k: -(-n // k) * t for k, t in per_step.items()}
    return min(cost, key=lambda k: (cost[k], -k)), cost


def batch_key(sweep, n_monitors: int) -> str:
    from .boundaries import CPMLBoundary  # noqa: PLC0415

    faces = sorted(f for f, bc in sweep.boundary_conditions.items() if isinstance(bc, CPMLBoundary))
    dtype = torch.float32 if sweep.precision == "single" else torch.float64
    base = cache_key(torch.device(sweep.device), sweep.mesh.dims, len(sweep.excited), dtype, faces)
    return f"{base}|batch|ch{len(sweep.channels)}|mon{n_monitors}|g{int(sweep.use_graphs)}"


def batch_for(sweep, path: Path | None = None) -> int:
    """Excitations per solver for ``sweep`` (a ScatteringSweep): from the
    cache, or timed now (one throw-away solver per candidate) and stored.
    A sweep of fewer than 8 steps keeps every excitation in one solver."""
    n = len(sweep.excited)
    if n <= 1 or sweep.n_steps < 8:  # nothing to choose, or too short a run to time
        return max(n, 1)
    path = path or cache_path()
    n_mon = len(list(sweep.monitors_factory())) if sweep.monitors_factory is not None else 0
    key = batch_key(sweep, n_mon)
    cands = batch_candidates(n)
    hit = _load(path).get(key)
    if isinstance(hit, dict) and hit.get("batch") in cands:
        return int(hit["batch"])
    per_step = {}
    for k in cands:
        s = sweep.make_solver(sweep.excited[:k])
        s.setup()
        per_step[k] = time_steps(s)
        del s
    best, cost = choose_batch(n, per_step)
    data = _load(path)
    data[key] = {"batch": best, "step_ms": {str(k): round(v, 4) for k, v in per_step.items()},
                 "sweep_step_ms": {str(k): round(v, 4) for k, v in cost.items()}}
    _store(path, data)
    return best
