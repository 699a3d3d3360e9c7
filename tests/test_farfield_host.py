"""Host side of the far-field post (CPU): the pieces around the device sums.

* direction frames equal the oracle's (theta outer, phi inner) bit for bit,
  and are orthonormal right-handed (r x theta = phi);
* device patch records carry area (n x H) and -area (n x E);
* the patch sets of a monitor dump equal the oracle's, spectrum division
  and geometry included; a dump without a spectrum is refused;
* image-plane walls: PEC on the outer node plane, PMC half a cell out, and
  the product far-field monitor books its walls that way;
* frequency picking mirrors the reference monitor's rules;
* grid faces (a monitor face and its images) are recognised with their
  currents laid out for the face kernel; a set of several faces is not.
"""

import math

import numpy as np
import pytest

from magnelio_amd import farfield as pf
from magnelio_amd import mesh as mm
from magnelio_amd.boundaries import CPMLBoundary, PECBoundary, PMCBoundary
from magnelio_amd.monitors import MonitorFarField
from oracle import farfield as off


def test_direction_frames_equal_oracle_and_are_orthonormal():
    th, ph = pf.default_angular_grid()
    fr = pf.direction_frames(th, ph)
    r, t, p = off.directions(th, ph)
    assert np.array_equal(fr[:, 0:3], r) and np.array_equal(fr[:, 3:6], t) and np.array_equal(fr[:, 6:9], p)
    assert np.allclose(np.cross(fr[:, 0:3], fr[:, 3:6]), fr[:, 6:9], atol=1e-15)
    for a in (r, t):
        assert np.allclose(np.sum(a * a, 1), 1.0)


def test_patch_records():
    rng = np.random.default_rng(2)
    n = 7
    s = pf.SurfacePatchSet(rng.random((n, 3)), np.eye(3)[rng.integers(0, 3, n)], rng.random(n),
                           rng.random((n, 3)) + 1j * rng.random((n, 3)), rng.random((n, 3)) + 1j * rng.random((n, 3)))
    rec = pf.patch_array(s)
    wJ = rec[:, 3:9:2] + 1j * rec[:, 4:9:2]
    wM = rec[:, 9:15:2] + 1j * rec[:, 10:15:2]
    for i in range(n):
        assert np.allclose(wJ[i], s.areas[i] * np.cross(s.normals[i], s.H[i]))
        assert np.allclose(wM[i], -s.areas[i] * np.cross(s.normals[i], s.E[i]))
    assert np.array_equal(rec[:, :3], s.centers)


def _dump(rng, spectrum=True):
    faces = []
    for name, axis, sign in (("xmin", 0, -1.0), ("zmax", 2, 1.0)):
        c1, c2 = np.sort(rng.random(3)), np.sort(rng.random(4))
        comps = pf.TANGENTIALS[axis]
        faces.append(dict(name=name, axis=axis, sign=sign, plane=0.3, c1=c1, c2=c2, w1=rng.random(3),
                          w2=rng.random(4), bins={c: rng.random((2, 3, 4)) + 1j * rng.random((2, 3, 4))
                                                  for c in comps}))
    d = dict(freqs=np.array([1e9, 2e9]), faces=faces, image_planes=[])
    if spectrum:
        d["source_spectrum"] = np.array([0.5 + 0.25j, 1e-320 + 0j])
    return d


def test_dump_patch_sets_equal_oracle():
    rng = np.random.default_rng(5)
    d = _dump(rng)
    for idx in (0, 1):
        got = pf.patch_sets_from_dump(d, idx)
        want = off.patch_sets_from_dump(d, idx)
        for g, w in zip(got, want):
            for k in ("centers", "normals", "areas", "E", "H"):
                assert np.array_equal(getattr(g, k), w[k]), k
    assert np.all(pf.patch_sets_from_dump(d, 1)[0].E == 0)  # underflowed spectrum -> zero
    with pytest.raises(ValueError, match="per watt"):
        pf.patch_sets_from_dump(_dump(rng, spectrum=False), 0)


def test_wall_positions_and_monitor_image_planes():
    nodes = np.array([0.0, 1.0, 3.0])
    for face in ("xmin", "xmax"):
        for kind in ("PEC", "PMC"):
            assert pf.wall_position(face, kind, nodes) == off.wall_position(face, kind, nodes)
    for fld in ("E", "H"):
        for kind in ("PEC", "PMC"):
            for c in range(3):
                assert pf.mirror_sign(fld, c, 1, kind) == off.mirror_sign(fld, c, 1, kind)
    g = mm.GridLines(np.linspace(0, 1e-2, 21), np.linspace(0, 1e-2, 21), np.linspace(0, 2e-2, 41))
    bcs = {f: CPMLBoundary(f, g, thickness_cells=4) for f in ("xmin", "xmax", "ymax", "zmin", "zmax")}
    for wall, kind in ((PECBoundary("ymin"), "PEC"), (PMCBoundary("ymin"), "PMC")):
        mon = MonitorFarField([1e9], margin_cells=2)
        mon._build(g, dict(bcs, ymin=wall))
        (p,) = mon._image_planes
        assert p["axis"] == 1 and p["kind"] == kind and p["at_low"]
        assert p["position"] == off.wall_position("ymin", kind, g.y)


def test_frequency_index():
    fr = [1e9, 2e9, 3e9]
    assert pf.frequency_index(fr, 2e9, None) == 1
    assert pf.frequency_index(fr, None, 2) == 2
    assert pf.frequency_index([5e9], None, None) == 0
    with pytest.raises(ValueError, match="recorded"):
        pf.frequency_index(fr, 2.5e9, None)
    with pytest.raises(ValueError, match="pass f="):
        pf.frequency_index(fr, None, None)


def test_source_spectrum_matches_oracle():
    v = np.sin(np.arange(300) * 0.1)
    assert np.allclose(pf.source_spectrum(v, 1e-12, [1e9, 2e10]), off.source_spectrum(v, 1e-12, [1e9, 2e10]),
                       rtol=1e-14, atol=0)
    assert math.isfinite(abs(pf.source_spectrum([1.0], 1.0, [1.0])[0]))


def test_tensor_face_detection_keeps_images_and_rejects_other_sets():
    rng = np.random.default_rng(9)
    d = _dump(rng)
    sets = pf.patch_sets_from_dump(d, 0)
    for s in sets:
        axis, plane, c1, c2, W = pf.tensor_face(s)
        n1, n2 = W.shape[:2]
        assert plane == s.centers[0, axis] and c1.size == n1 and c2.size == n2
        rec = pf.patch_array(s)
        t1, t2 = [b for b in range(3) if b != axis]
        wj = (rec[:, 3:9:2] + 1j * rec[:, 4:9:2]).reshape(n1, n2, 3)
        assert np.array_equal(W[..., 0], wj[..., t1]) and np.array_equal(W[..., 1], wj[..., t2])
        for plane_ in (pf.ImagePlane(axis, 0.0, "PEC", True), pf.ImagePlane(t1, -0.2, "PMC", True)):
            m = pf.tensor_face(pf._mirror(s, plane_))
            assert m is not None and m[3].size == n2
    assert pf.tensor_face(pf._concat(sets)) is None  # two faces: no single grid
    s = sets[0]
    shuffled = pf.SurfacePatchSet(s.centers[::-1].copy(), s.normals, s.areas, s.E, s.H)
    assert pf.tensor_face(shuffled) is None or pf.tensor_face(shuffled)[2][0] == s.centers[-1, 1]
