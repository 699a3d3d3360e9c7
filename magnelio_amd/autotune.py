"""Launch-shape tuning of the curl launches, cached on disk.

The tunable is the z-chunk of the curl kernels (planes a thread marches,
keeping the z neighbour in registers): shorter chunks give more work-groups,
longer ones fewer re-read chunk-start planes.  The sets per thread (two, or
three for multiples of 3) and the XCD band mapping are fixed in the
library from measurements; the chunk depends on the grid, K and the CPML the
launch carries, so it is measured per configuration.

``FITTimeDomainSolver(kchunk="auto")`` calls :func:`kchunk_for` at setup,
before any state is loaded: the curl launches (with the solver's CPML, as a
step runs them) are timed on the solver's own zeroed buffers — zero fields
and zero psi stay zero, so the timing changes nothing — and the fastest
candidate is stored in a JSON cache (``MAGNELIO_HIP_AUTOTUNE_CACHE``, default
``~/.cache/magnelio_amd/autotune.json``) keyed by device, the library's
content hash, grid, K, dtype and CPML faces; later solvers of the same key
read it.

The reference's CUDA tier has a fixed block shape (``numba_kernels.py:446``);
this module has no counterpart there.
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

CANDIDATES = (1, 2, 3, 4)
REPS = 5


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
    launches it) at z-chunk ``kchunk``, after one untimed step."""
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
    """Fastest candidate; ties go to the smaller chunk."""
    return min(sorted(timings), key=lambda kc: (timings[kc], kc))


def kchunk_for(solver, candidates=CANDIDATES, path: Path | None = None) -> int:
    """z-chunk for ``solver`` (set up, fields still zero): from the cache, or
    timed now and stored."""
    path = path or cache_path()
    cpml = solver.cpml if (solver.cpml is not None and solver._dispersion is None
                          and solver._dispersion_mu is None and solver._sibc is None) else None
    faces = [f.face for f in solver.cpml.faces] if solver.cpml is not None else []
    key = cache_key(solver.e.device, solver.dims, solver.nsets, solver.e.dtype, faces)
    data = _load(path)
    hit = data.get(key)
    if isinstance(hit, dict) and hit.get("kchunk") in candidates:
        return int(hit["kchunk"])
    timings = {kc: time_curl(solver.e, solver.h, solver.alpha_e, solver.beta_e, solver.alpha_h,
                             solver.beta_h, solver.dims, kc, cpml=cpml, curl64=solver.curl64)
               for kc in candidates}
    best = choose(timings)
    data = _load(path)  # re-read: another process may have added entries meanwhile
    data[key] = {"kchunk": best, "ms": {str(k): round(v, 4) for k, v in timings.items()}}
    _store(path, data)
    return best
