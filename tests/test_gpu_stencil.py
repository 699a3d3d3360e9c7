"""GPU parity of the fused curl/material kernels against the oracle.

The HIP kernels reproduce the reference's operation order with FMA
contraction off, so every comparison here is bit-exact (np.array_equal):
one step and many steps, float32 (array-stencil and Numba orders) and
float64, K field sets advanced in one launch, ragged grid sizes that do not
fill the 64 x 4 thread tile, and every z-chunking of the march.
"""

import numpy as np
import pytest
import torch

from magnelio_amd import kernels, layout
from oracle import yee

pytestmark = pytest.mark.gpu


def _state(rng, nx, ny, nz, dtype, nsets):
    ne, nh = yee.n_e(nx, ny, nz), yee.n_h(nx, ny, nz)
    e = rng.standard_normal((nsets, ne)).astype(dtype)
    h = rng.standard_normal((nsets, nh)).astype(dtype)
    ae = rng.uniform(0.5, 1.0, ne).astype(dtype)
    be = rng.uniform(0.01, 0.3, ne).astype(dtype)
    ah = rng.uniform(0.5, 1.0, nh).astype(dtype)
    bh = rng.uniform(0.01, 0.3, nh).astype(dtype)
    # a sprinkle of frozen (PEC-like) edges
    frozen = rng.random(ne) < 0.05
    ae[frozen] = 0
    be[frozen] = 0
    e[:, frozen] = 0
    return e, h, ae, be, ah, bh


def _to_dev(a, dims, side, gpu):
    return torch.from_numpy(np.ascontiguousarray(layout.to_device(a, *dims, side))).to(gpu)


def _to_ref(t, dims, side):
    return layout.to_reference(t.cpu().numpy(), *dims, side)


def _run(dims, dtype, nsets, order, steps, kchunk, gpu, seed=0):
    rng = np.random.default_rng(seed)
    e, h, ae, be, ah, bh = _state(rng, *dims, dtype, nsets)
    de, dh = _to_dev(e, dims, "E", gpu), _to_dev(h, dims, "H", gpu)
    dae, dbe = _to_dev(ae, dims, "E", gpu), _to_dev(be, dims, "E", gpu)
    dah, dbh = _to_dev(ah, dims, "H", gpu), _to_dev(bh, dims, "H", gpu)
    for _ in range(steps):
        kernels.update_e(de, dh, dae, dbe, dims, curl64=(order == "numba"), kchunk=kchunk)
        kernels.update_h(de, dh, dah, dbh, dims, kchunk=kchunk)
    torch.cuda.synchronize()
    for s in range(nsets):
        for _ in range(steps):
            yee.update_e(e[s], h[s], ae, be, *dims, order=order)
            yee.update_h(e[s], h[s], ah, bh, *dims)
    return _to_ref(de, dims, "E"), _to_ref(dh, dims, "H"), e, h


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims", [(5, 3, 4), (70, 9, 6), (1, 1, 1), (64, 4, 17)])
def test_single_set_bit_exact(dtype, dims, gpu):
    print(f"stencil dtype={np.dtype(dtype).name} dims={dims}", flush=True)
    ge, gh, oe, oh = _run(dims, dtype, 1, "stencil", 3, 4, gpu)
    assert np.array_equal(ge, oe)
    assert np.array_equal(gh, oh)


def test_numba_order_float32_bit_exact(gpu):
    dims = (33, 12, 10)
    ge, gh, oe, oh = _run(dims, np.float32, 1, "numba", 4, 3, gpu, seed=4)
    assert np.array_equal(ge, oe)
    assert np.array_equal(gh, oh)


@pytest.mark.parametrize("nsets", [2, 3, 6, 9, 11])
def test_k_field_sets_bit_exact(nsets, gpu):
    print(f"K sets nsets={nsets}", flush=True)
    dims = (20, 7, 9)
    ge, gh, oe, oh = _run(dims, np.float32, nsets, "stencil", 2, 5, gpu, seed=nsets)
    assert np.array_equal(ge, oe)
    assert np.array_equal(gh, oh)


@pytest.mark.parametrize("kchunk", [1, 2, 7, 64])
def test_result_independent_of_z_chunking(kchunk, gpu):
    dims = (40, 10, 13)
    ge, gh, oe, oh = _run(dims, np.float32, 1, "stencil", 2, kchunk, gpu, seed=9)
    assert np.array_equal(ge, oe)
    assert np.array_equal(gh, oh)


def test_many_steps_bit_exact(gpu):
    dims = (24, 16, 20)
    ge, gh, oe, oh = _run(dims, np.float32, 1, "stencil", 40, 16, gpu, seed=1)
    assert np.array_equal(ge, oe)
    assert np.array_equal(gh, oh)


def test_energy_sum_matches_float64_reference(gpu):
    rng = np.random.default_rng(3)
    dims = (30, 20, 10)
    ne = yee.n_e(*dims)
    e = rng.standard_normal((3, ne)).astype(np.float32)
    w = rng.uniform(0.1, 1.0, ne).astype(np.float32)
    red = kernels.SqSum(3, gpu)
    out = torch.zeros(3, dtype=torch.float64, device=gpu)
    red(torch.from_numpy(e).to(gpu), torch.from_numpy(w).to(gpu), out, scale=0.5)
    got = out.cpu().numpy()
    for s in range(3):
        want = 0.5 * float(np.sum((w * e[s]) * e[s], dtype=np.float64))
        assert abs(got[s] - want) <= 1e-12 * abs(want)


def test_type_errors_match_reference_contract(gpu):
    dims = (2, 2, 2)
    e = torch.zeros(yee.n_e(*dims), dtype=torch.float16, device=gpu)
    h = torch.zeros(yee.n_h(*dims), dtype=torch.float16, device=gpu)
    with pytest.raises(TypeError):
        kernels.update_e(e, h, e, e, dims)
