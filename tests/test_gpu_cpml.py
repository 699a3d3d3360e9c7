"""GPU parity of the CPML recursion fused into the curl kernels.

All six faces at once (corner and edge overlaps exercise the face order),
PEC cells inside the slabs (air masks), K field sets (one, two and three per
thread), float32 and float64, CPML with the curl (y/z faces in the curl
launch, the x-slab columns after it, composed or through the two entry
points) and as separate slab launches; compared bit-exactly with the
oracle's separate slab passes after several leapfrog steps.
"""

import numpy as np
import pytest
import torch

from magnelio_amd import boundaries, kernels, layout
from magnelio_amd.mesh import GridLines
from oracle import cpml as oc
from oracle import mesh as om
from oracle import yee

pytestmark = pytest.mark.gpu

FACES = ["zmin", "xmax", "ymin", "zmax", "xmin", "ymax"]


def _setup(rng, dims, dtype, with_pec):
    nx, ny, nz = dims
    x, y, z = (np.cumsum(np.r_[0, rng.uniform(1e-3, 1.4e-3, n)]) for n in dims)
    mid = np.zeros(dims, np.int32)
    if with_pec:
        mid[1:4, 2:5, 0:3] = 1  # PEC brick reaching into the zmin/xmin/ymin slabs
    pec_cell = mid == 1
    me = om.mass_e(x, y, z, mid, [[1.0, 1.0]] * 3, om.EPS0)
    mm = om.mass_h(x, y, z, mid, [[1.0, 1.0]] * 3, om.MU0)
    pec = om.pec_edge_mask(pec_cell)
    dt = om.courant_dt(x, y, z, 0.95)
    ae, be = om.alpha_beta(me, np.zeros_like(me), dt, pec)
    ah, bh = om.alpha_beta(mm, np.zeros_like(mm), dt)
    return (x, y, z), pec, pec_cell, dt, [a.astype(dtype) for a in (ae, be, ah, bh)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("with_pec", [False, True])
@pytest.mark.parametrize("faces", [FACES, ["zmin", "xmax", "ymin"]], ids=["six", "lone_x"])
@pytest.mark.parametrize("mode,nsets", [("with_curl", 3), ("with_curl", 2), ("with_curl", 1),
                                        ("split", 3), ("separate", 3)])
def test_fused_cpml_bit_exact(dtype, with_pec, faces, mode, nsets, gpu):
    print(f"cpml dtype={np.dtype(dtype).name} pec={with_pec} faces={faces} mode={mode} K={nsets}",
          flush=True)
    rng = np.random.default_rng(12)
    dims = (13, 11, 15)
    axes, pec, pec_cell, dt, (ae, be, ah, bh) = _setup(rng, dims, dtype, with_pec)
    nl = 4
    ofaces_sets = [[oc.CPMLFace(f, *axes, dt, dtype, pec_e_flat=pec if with_pec else None,
                                pec_cell=pec_cell if with_pec else None, n_layers=nl)
                    for f in faces] for _ in range(nsets)]
    g = GridLines(*axes)
    hfaces = []
    for f in faces:
        b = boundaries.CPMLBoundary(f, g, thickness_cells=nl)
        b.initialize(dt)
        hfaces.append(b)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    dev_pml = boundaries.DeviceCPML(hfaces, dims, nsets, tdt, gpu,
                                    pec_e_flat=pec if with_pec else None,
                                    pec_cell=pec_cell if with_pec else None)
    e = (rng.standard_normal((nsets, yee.n_e(*dims))) * (~pec)).astype(dtype)
    h = rng.standard_normal((nsets, yee.n_h(*dims))).astype(dtype)

    def d(a, side):
        return torch.from_numpy(np.ascontiguousarray(layout.to_device(a, *dims, side))).to(gpu)

    de, dh = d(e, "E"), d(h, "H")
    dae, dbe, dah, dbh = d(ae, "E"), d(be, "E"), d(ah, "H"), d(bh, "H")
    for _ in range(6):
        if mode == "with_curl":  # y/z faces in the curl launch, then the x-slab columns
            kernels.update_e(de, dh, dae, dbe, dims, cpml=dev_pml, kchunk=4)
            kernels.update_h(de, dh, dah, dbh, dims, cpml=dev_pml, kchunk=4)
        elif mode == "split":  # the same two launches through their own entry points
            kernels.update_e(de, dh, dae, dbe, dims, cpml=dev_pml, kchunk=4, xcols=False)
            kernels.cpml_xcols_e(de, dh, dbe, dims, dev_pml)
            kernels.update_h(de, dh, dah, dbh, dims, cpml=dev_pml, kchunk=4, xcols=False)
            kernels.cpml_xcols_h(de, dh, dbh, dims, dev_pml)
        else:  # curl launch, then one slab launch per face
            kernels.update_e(de, dh, dae, dbe, dims, kchunk=4)
            kernels.cpml_e(de, dh, dbe, dims, dev_pml)
            kernels.update_h(de, dh, dah, dbh, dims, kchunk=4)
            kernels.cpml_h(de, dh, dbh, dims, dev_pml)
    torch.cuda.synchronize()
    for s in range(nsets):
        for _ in range(6):
            yee.update_e(e[s], h[s], ae, be, *dims)
            for f in ofaces_sets[s]:
                f.update_e(e[s], h[s], be)
            yee.update_h(e[s], h[s], ah, bh, *dims)
            for f in ofaces_sets[s]:
                f.update_h(e[s], h[s], bh)
    ge = layout.to_reference(de.cpu().numpy(), *dims, "E")
    gh = layout.to_reference(dh.cpu().numpy(), *dims, "H")
    assert np.array_equal(ge, e)
    assert np.array_equal(gh, h)
    sd = dev_pml.state_dict()
    for s in range(nsets):
        for f in ofaces_sets[s]:
            for t in range(2):
                assert np.array_equal(sd[f"{f.face}_psi_e{t}"][s], f.psi_e[t])
                assert np.array_equal(sd[f"{f.face}_psi_h{t}"][s], f.psi_h[t])
