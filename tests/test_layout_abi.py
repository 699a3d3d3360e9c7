"""CPU checks of the host boundary: layout permutations and the C-ABI exports.

* reference (z-fastest) <-> device (x-fastest) permutations are exact
  inverses, and the index map agrees element for element with the value
  permutation (bit-exact integer work);
* libmgh.so loads without a GPU and exports every function declared in
  include/mgh.h (no compute call is made here).
"""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from magnelio_amd import layout

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("side", ["E", "H"])
@pytest.mark.parametrize("dims", [(3, 4, 5), (1, 1, 1), (7, 2, 3)])
def test_layout_roundtrip_and_index_map(side, dims):
    nx, ny, nz = dims
    n = layout.vec_len(nx, ny, nz, side)
    ref = np.arange(n, dtype=np.int64)
    dev = layout.to_device(ref, nx, ny, nz, side)
    assert np.array_equal(layout.to_reference(dev, nx, ny, nz, side), ref)
    # the value at device slot p is the reference index of that element
    idx = layout.index_to_device(ref, nx, ny, nz, side)
    assert np.array_equal(dev[idx], ref)
    assert np.array_equal(layout.index_to_reference(idx, nx, ny, nz, side), ref)
    assert np.array_equal(np.sort(idx), ref)  # a permutation


def test_layout_component_order_is_x_fastest():
    nx, ny, nz = 3, 4, 5
    shp = layout.component_shapes(nx, ny, nz)["Ex"]
    ref = np.arange(layout.vec_len(nx, ny, nz, "E"))
    dev = layout.to_device(ref, nx, ny, nz, "E")
    ex_ref = ref[: np.prod(shp)].reshape(shp)
    # device Ex block: element (i,j,k) at k*(n1*n0) + j*n0 + i
    n0, n1, _ = shp
    for (i, j, k) in [(0, 0, 0), (2, 3, 4), (1, 2, 3)]:
        assert dev[k * n1 * n0 + j * n0 + i] == ex_ref[i, j, k]


def test_batched_layout_leading_axis():
    nx, ny, nz = 2, 3, 4
    n = layout.vec_len(nx, ny, nz, "H")
    ref = np.arange(3 * n).reshape(3, n)
    dev = layout.to_device(ref, nx, ny, nz, "H")
    for s in range(3):
        assert np.array_equal(dev[s], layout.to_device(ref[s], nx, ny, nz, "H"))


def _declared_functions() -> list[str]:
    text = (ROOT / "include" / "mgh.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mgh_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    from magnelio_amd import _lib

    lib = _lib.load()
    declared = _declared_functions()
    assert len(declared) >= 7
    for name in declared:
        assert hasattr(lib, name), f"libmgh.so does not export {name}"
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    # no undeclared exports either
    assert set(_lib.SIGNATURES) == set(declared)
    assert lib.mgh_version() >= 10000


def test_library_rejects_bad_arguments_without_a_gpu():
    """Argument validation happens on the host before any HIP call."""
    from magnelio_amd import _lib

    lib = _lib.load()
    rc = lib.mgh_update_e(3, 4, 4, 4, 1, None, 0, None, 0, None, None, 0, 16, None)
    assert rc == 1
    assert "dtype" in _lib.last_error()
    rc = lib.mgh_update_h(4, 0, 4, 4, 1, None, 0, None, 0, None, None, 16, None)
    assert rc == 1
    with pytest.raises(ValueError):
        _lib.check(rc, "mgh_update_h")
