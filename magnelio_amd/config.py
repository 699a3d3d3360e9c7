"""Environment knobs of the HIP backend.

An explicit argument always wins; an argument left at ``None`` takes the
environment, then the built-in default — the reference's rule for its own
knobs (src/magnelio/_backend/array_api.py:137-175, ``MAGNELIO_PRECISION``;
``MAGNELIO_GPU_GRAPHS``, solver/_gpu_graphs.py:53-55;
``MAGNELIO_TILE_SKIP``, solver/_tile_skip.py:58-60).  The launch shape of
the curl kernels has a knob of its own, ``MAGNELIO_HIP_LAUNCH_SHAPE``:
"auto" (timed at setup), a shape word, or "kchunk,sets,band".  A malformed
value raises rather than being ignored.
"""

from __future__ import annotations

import os

from . import kernels

PRECISIONS = ("single", "double")


def _env(name: str) -> str:
    return os.environ.get(name, "").strip().lower()


def precision(value: str | None) -> str:
    if value is None:
        value = _env("MAGNELIO_PRECISION") or "single"
        src = "MAGNELIO_PRECISION"
    else:
        src = "precision"
    if value not in PRECISIONS:
        raise ValueError(f"{src} must be one of {PRECISIONS}, got {value!r}")
    return value


def graphs(value: bool | None, default: bool = False) -> bool:
    """HIP-graph replay: the argument, else ``MAGNELIO_GPU_GRAPHS`` ("0"
    off, "1" on), else ``default``."""
    if value is not None:
        return bool(value)
    env = _env("MAGNELIO_GPU_GRAPHS")
    if env in ("", None):
        return default
    if env not in ("0", "1"):
        raise ValueError(f"MAGNELIO_GPU_GRAPHS must be 0 or 1, got {env!r}")
    return env == "1"


def tile_skip(value: str | None) -> str:
    """Dead-unit skipping: the argument ("auto", "on", "off"), else
    ``MAGNELIO_TILE_SKIP`` ("0": off, "1": auto), else "auto"."""
    if value is not None:
        return value
    env = _env("MAGNELIO_TILE_SKIP")
    if env == "":
        return "auto"
    if env not in ("0", "1"):
        raise ValueError(f"MAGNELIO_TILE_SKIP must be 0 or 1, got {env!r}")
    return "off" if env == "0" else "auto"


def launch_shape(value: int | str | None) -> int | str:
    """Curl launch shape: the argument, else ``MAGNELIO_HIP_LAUNCH_SHAPE``,
    else the library default (two z planes per thread)."""
    if value is not None:
        return value
    env = _env("MAGNELIO_HIP_LAUNCH_SHAPE")
    if env == "":
        return kernels.DEFAULT_KCHUNK
    if env == "auto":
        return "auto"
    try:
        parts = [int(p) for p in env.split(",")]
    except ValueError:
        raise ValueError(f"MAGNELIO_HIP_LAUNCH_SHAPE must be 'auto', a shape word or "
                         f"'kchunk,sets,band', got {env!r}") from None
    if len(parts) == 1:
        word = parts[0]
        if kernels.shape_parts(word)[0] < 1:
            raise ValueError(f"MAGNELIO_HIP_LAUNCH_SHAPE: no z planes in {env!r}")
        return word
    if len(parts) != 3:
        raise ValueError(f"MAGNELIO_HIP_LAUNCH_SHAPE must be 'kchunk,sets,band', got {env!r}")
    return kernels.launch_shape(*parts)
