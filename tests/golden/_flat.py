"""Nested dicts of arrays <-> flat npz keys ("a/b/c"), for checkpoint fixtures."""

import numpy as np


def flatten(d, prefix=""):
    """Dicts and lists of dicts become key paths (list items "0", "1", ...);
    None is left out; everything else must be a plain (non-object) array."""
    out = {}
    items = d.items() if isinstance(d, dict) else enumerate(d)
    for k, v in items:
        key = f"{prefix}/{k}" if prefix else str(k)
        if v is None:
            continue
        if isinstance(v, dict) or (isinstance(v, (list, tuple)) and v and isinstance(v[0], dict)):
            out.update(flatten(v, key))
        else:
            a = np.array(v, copy=True)  # a live accumulator's view must not follow the run
            if a.dtype == object:
                raise TypeError(f"{key}: not a plain array")
            out[key] = a
    return out


def unflatten(flat, prefix):
    """The nested dict stored under ``prefix`` (scalars back to Python)."""
    out = {}
    for k, v in flat.items():
        if not k.startswith(prefix + "/"):
            continue
        parts = k[len(prefix) + 1:].split("/")
        node = out
        for p in parts[:-1]:
            node = node.setdefault(p, {})
        v = np.asarray(v)
        node[parts[-1]] = v.item() if v.ndim == 0 and v.dtype.kind in "iufb" else v
    return out
