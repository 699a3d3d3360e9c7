"""The HBM streaming probe (mgh_stream_probe) computes what it states:
out[w] = sum of in[r] over r = w mod n_write, in ascending r, float32."""

import pytest

torch = pytest.importorskip("torch")

from magnelio_amd import _lib  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nr,nw,n", [(12, 3, 1000), (42, 18, 777), (12, 3, 256)])
def test_stream_probe_sums(nr, nw, n):
    g = torch.Generator().manual_seed(nr + n)
    src = torch.randn(nr * n, generator=g, dtype=torch.float32)
    dst = torch.full((nw * n,), float("nan"), dtype=torch.float32, device="cuda:0")
    d_src = src.to("cuda:0")
    _lib.call("mgh_stream_probe", nr, nw, d_src.data_ptr(), dst.data_ptr(), n,
              int(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    x = src.view(nr, n)
    for w in range(nw):
        acc = torch.zeros(n, dtype=torch.float32)
        for r in range(w, nr, nw):
            acc = acc + x[r]
        assert torch.equal(dst.view(nw, n)[w].cpu(), acc)


def test_stream_probe_rejects_other_mixes():
    buf = torch.zeros(64, device="cuda:0")
    with pytest.raises(ValueError, match="unsupported mix"):
        _lib.call("mgh_stream_probe", 5, 2, buf.data_ptr(), buf.data_ptr(), 8,
                  int(torch.cuda.current_stream().cuda_stream))
