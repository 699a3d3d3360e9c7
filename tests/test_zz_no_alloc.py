"""GPU test of the allocation-free energy reduction (mgh_weighted_sq_sum_noalloc).

It must give what mgh_weighted_sq_sum gives with a caller workspace — the
same kernels over the same partial-sum layout, so bit for bit — on two calls
in a row, the second with a smaller input (fewer sets and elements: the
static scratch keeps the first call's partials and must not leak them),
float32 and float64, with and without accumulation; and it must refuse more
field sets than its scratch holds before launching anything.
"""

import numpy as np
import pytest
import torch

from magnelio_amd import _lib

pytestmark = pytest.mark.gpu


def _reference(code, x, w, k, n, out, accumulate, scale):
    work = torch.empty(int(_lib.load().mgh_weighted_sq_sum_workspace(k)), dtype=torch.float64,
                       device=x.device)
    _lib.call("mgh_weighted_sq_sum", code, n, k, x.data_ptr(), n, w.data_ptr(), out.data_ptr(),
              int(accumulate), scale, work.data_ptr(), work.numel(),
              int(torch.cuda.current_stream().cuda_stream))


def _noalloc(code, x, w, k, n, out, accumulate, scale):
    _lib.call("mgh_weighted_sq_sum_noalloc", code, n, k, x.data_ptr(), n, w.data_ptr(),
              out.data_ptr(), int(accumulate), scale, int(torch.cuda.current_stream().cuda_stream))


@pytest.mark.parametrize("dtype,code", [(torch.float32, 4), (torch.float64, 8)])
def test_noalloc_equals_workspace_form(dtype, code, gpu):
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(3)
    for k, n, accumulate in ((5, 300_007, False), (2, 1_001, True)):
        print(f"noalloc reduction {dtype} sets {k} n {n} accumulate {accumulate}", flush=True)
        x = torch.randn((k, n), generator=g, device=dev, dtype=dtype)
        w = torch.rand(n, generator=g, device=dev, dtype=dtype)
        start = torch.randn(k, generator=g, device=dev, dtype=torch.float64)
        want, got = start.clone(), start.clone()
        _reference(code, x, w, k, n, want, accumulate, 0.5)
        _noalloc(code, x, w, k, n, got, accumulate, 0.5)
        torch.cuda.synchronize(dev)
        assert torch.equal(got, want)
        # and against a float64 sum of the field-precision products
        prod = ((w * x) * x).double().sum(dim=1) * 0.5
        exp = prod + (start if accumulate else 0)
        assert torch.allclose(got, exp, rtol=1e-12, atol=0)


def test_noalloc_rejects_too_many_sets(gpu):
    dev = torch.device("cuda:0")
    x = torch.zeros((65, 16), device=dev)
    w = torch.zeros(16, device=dev)
    out = torch.zeros(65, dtype=torch.float64, device=dev)
    with pytest.raises(ValueError, match="exceed the static scratch"):
        _noalloc(4, x, w, 65, 16, out, False, 1.0)
