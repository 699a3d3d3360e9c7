"""Pins of oracle/mesh.py by brute force, and of the CPU baseline kernels.

* PEC edge rule (src/magnelio/mesh/indexing.py:175-257): an edge is PEC when
  ANY cell around it (index-clamped at the grid boundary) is PEC — checked
  edge by edge with explicit loops over its (up to four) cells on random
  small meshes with random PEC cells;
* graded-mesh masses (material_matrices.py:55-73, :386-407): M_eps of an
  edge = EPS0 eps_r(owning cell) * (dual width) * (dual width) / (primal
  length), M_mu of a face = MU0 mu_r * primal area / dual width, dual widths
  (d[0], (d[i-1] + d[i]) / 2, d[N-1]) — per element with scalar loops on a
  random graded grid with random materials;
* port-plane flattening (material_matrices.py:106-179): the boundary slab
  of the two tangential E components (or of the normal H component) takes
  the values of the first interior slab, nothing else changes;
* the C/OpenMP restatement of the Numba kernels (oracle/fit_omp.c, the CPU
  baseline) equals the oracle's Numba-order update bit for bit.
"""

import numpy as np
import pytest

from oracle import build as ob
from oracle import mesh as om
from oracle import yee

SHAPES_E = lambda nx, ny, nz: [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)]  # noqa: E731
SHAPES_H = lambda nx, ny, nz: [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]  # noqa: E731


def _split(flat, shapes):
    offs = np.cumsum([0] + [int(np.prod(s)) for s in shapes])
    return [flat[offs[c]:offs[c + 1]].reshape(shapes[c]) for c in range(3)]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pec_or_rule_brute_force(seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(v) for v in rng.integers(2, 6, 3))
    pec_cell = rng.random((nx, ny, nz)) < 0.25
    got = _split(om.pec_edge_mask(pec_cell), SHAPES_E(nx, ny, nz))
    n = (nx, ny, nz)
    for c, shape in enumerate(SHAPES_E(nx, ny, nz)):
        others = [a for a in range(3) if a != c]
        for idx in np.ndindex(*shape):
            cells = []
            for da in (-1, 0):
                for db in (-1, 0):
                    cell = list(idx)
                    cell[others[0]] = min(max(idx[others[0]] + da, 0), n[others[0]] - 1)
                    cell[others[1]] = min(max(idx[others[1]] + db, 0), n[others[1]] - 1)
                    cells.append(tuple(cell))
            assert got[c][idx] == any(pec_cell[cc] for cc in cells), (c, idx)


def _dual(d, i):
    n = len(d)
    if i == 0:
        return d[0]
    if i == n:
        return d[n - 1]
    return 0.5 * (d[i - 1] + d[i])


def test_graded_masses_brute_force():
    rng = np.random.default_rng(5)
    nx, ny, nz = 4, 3, 5
    axes = [np.concatenate([[0.0], np.cumsum(rng.uniform(0.5e-3, 1.5e-3, k))]) for k in (nx, ny, nz)]
    d = [np.diff(a) for a in axes]
    mid = rng.integers(0, 3, (nx, ny, nz)).astype(np.int32)
    eps = [[1.0, 2.5, 7.0], [1.0, 3.5, 5.0], [1.0, 4.5, 9.0]]
    mu = [[1.0, 1.5, 2.0], [1.0, 1.25, 3.0], [1.0, 1.75, 4.0]]
    me = _split(om.mass_e(*axes, mid, eps, om.EPS0), SHAPES_E(nx, ny, nz))
    mh = _split(om.mass_h(*axes, mid, mu, om.MU0), SHAPES_H(nx, ny, nz))
    n = (nx, ny, nz)
    for c in range(3):
        a1, a2 = [a for a in range(3) if a != c]
        for idx in np.ndindex(*SHAPES_E(nx, ny, nz)[c]):
            cell = tuple(min(idx[a], n[a] - 1) for a in range(3))
            want = om.EPS0 * eps[c][mid[cell]] * (_dual(d[a1], idx[a1]) * _dual(d[a2], idx[a2])
                                                  / d[c][idx[c]])
            assert me[c][idx] == pytest.approx(want, rel=1e-14)
        for idx in np.ndindex(*SHAPES_H(nx, ny, nz)[c]):
            cell = tuple(min(idx[a], n[a] - 1) for a in range(3))
            want = om.MU0 * mu[c][mid[cell]] * (d[a1][idx[a1]] * d[a2][idx[a2]] / _dual(d[c], idx[c]))
            assert mh[c][idx] == pytest.approx(want, rel=1e-14)


@pytest.mark.parametrize("face", ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
@pytest.mark.parametrize("side", ["E", "H"])
def test_flatten_port_brute_force(face, side):
    nx, ny, nz = 3, 4, 5
    shapes = SHAPES_E(nx, ny, nz) if side == "E" else SHAPES_H(nx, ny, nz)
    total = sum(int(np.prod(s)) for s in shapes)
    vec = np.random.default_rng(7).standard_normal(total)
    got = _split(om.flatten_port(vec, nx, ny, nz, face, side), shapes)
    src = _split(vec, shapes)
    axis = "xyz".index(face[0])
    n_ax = (nx, ny, nz)[axis]
    plane, inner = (0, 1) if face.endswith("min") else (n_ax, n_ax - 1)
    touched = [c for c in range(3) if c != axis] if side == "E" else [axis]
    for c in range(3):
        for idx in np.ndindex(*shapes[c]):
            if c in touched and idx[axis] == plane:
                j = list(idx)
                j[axis] = inner
                assert got[c][idx] == src[c][tuple(j)]
            else:
                assert got[c][idx] == src[c][idx]


@pytest.mark.parametrize("dims", [(5, 4, 6), (9, 3, 2), (1, 1, 1)])
def test_cpu_baseline_kernels_equal_oracle_numba_order(dims):
    """oracle/fit_omp.c == oracle/yee.py (order="numba"), float32, bit for bit."""
    rng = np.random.default_rng(sum(dims))
    ne, nh = yee.n_e(*dims), yee.n_h(*dims)
    e = rng.standard_normal(ne).astype(np.float32)
    h = rng.standard_normal(nh).astype(np.float32)
    ae, be = (rng.uniform(0.3, 1.0, ne).astype(np.float32) for _ in range(2))
    ah, bh = (rng.uniform(0.3, 1.0, nh).astype(np.float32) for _ in range(2))
    e2, h2 = e.copy(), h.copy()
    lib = ob.load()
    lib.fit_omp_set_threads(3)
    for _ in range(4):
        yee.update_e(e, h, ae, be, *dims, order="numba")
        yee.update_h(e, h, ah, bh, *dims)
        ob.update_e_f32(e2, h2, ae, be, *dims)
        ob.update_h_f32(e2, h2, ah, bh, *dims)
        assert np.array_equal(e, e2) and np.array_equal(h, h2)
