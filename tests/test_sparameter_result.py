"""SParameterResult bookkeeping (CPU): merge order and the reference's
validation errors (sparameter_result.py:75-101, :449-495)."""

import numpy as np
import pytest

from magnelio_amd.sweep import SParameterResult

CH = (("p1", 0), ("p1", 1), ("p2", 0))
F = np.array([1e9, 2e9, 3e9])


def _col(exc, val):
    m = np.full((F.size, len(CH), 1), val, dtype=complex)
    return SParameterResult(F, CH, (exc,), m)


def test_merge_orders_columns():
    r = SParameterResult.merge([_col(("p2", 0), 2.0), _col(("p1", 0), 1.0)])
    assert r.excitations == (("p2", 0), ("p1", 0))
    assert r.matrix.shape == (3, 3, 2)
    assert np.all(r.S("p1", "p2") == 2.0) and np.all(r.S("p1", "p1") == 1.0)
    assert not r.is_complete


def test_validation():
    with pytest.raises(ValueError):
        SParameterResult.merge([_col(("p1", 0), 1.0), _col(("p1", 0), 1.0)])
    with pytest.raises(ValueError):
        SParameterResult(F, CH, (("p3", 0),), np.zeros((3, 3, 1)))
    with pytest.raises(ValueError):
        SParameterResult(-F, CH, (("p1", 0),), np.zeros((3, 3, 1)))
    with pytest.raises(ValueError):
        SParameterResult(F, CH, (("p1", 0),), np.zeros((3, 2, 1)))
    other = SParameterResult(F * 2, CH, (("p1", 1),), np.zeros((3, 3, 1)))
    with pytest.raises(ValueError):
        SParameterResult.merge([_col(("p1", 0), 1.0), other])
