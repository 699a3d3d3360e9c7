"""The device near-to-far-field transform (mgh_ntff / mgh_ntff_faces) against the oracle.

* seeded random patch sets with a PEC and a PMC image plane (one physical,
  one a symmetry plane): E_theta / E_phi within 1e-12 of the largest
  amplitude of the oracle's transform (FP64 both; the sums associate
  differently), the physical mask equal;
* the exact near fields of a Hertzian dipole on a closed box: the dipole's
  closed-form far field (2e-3, the box quadrature), its radiated power and
  peak directivity 1.5 through ``FarFieldResult``;
* an empty patch set gives a zero pattern;
* seeded graded grid faces with a PEC and a PMC image plane: the grid-face
  kernel (mgh_ntff_faces, phase factored per grid axis) and the per-patch
  kernel (mgh_ntff) each within 1e-12 of the oracle's transform;
* the golden CPML / TF-SF box run (tests/golden/box_tfsf_dft.npz): the
  far-field monitor's pattern from its own device bins equals the oracle's
  transform of the reference's bins within 1e-5 of the peak (the bins agree
  to ~1e-7).
"""

import math

import numpy as np
import pytest

import cases
from magnelio_amd import farfield as pf
from oracle import farfield as off

pytestmark = pytest.mark.gpu

THETA = np.linspace(0.0, math.pi, 37)
PHI = np.linspace(0.0, 2 * math.pi, 73)


def _random_sets(rng, n_sets=2, n=1500, size=0.2):
    out = []
    for _ in range(n_sets):
        cen = rng.uniform(0.0, size, (n, 3))
        nrm = np.zeros((n, 3))
        ax = rng.integers(0, 3, n)
        nrm[np.arange(n), ax] = rng.choice([-1.0, 1.0], n)
        E = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
        H = (rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))) / 377.0
        out.append(dict(centers=cen, normals=nrm, areas=rng.uniform(1e-6, 4e-6, n), E=E, H=H))
    return out


def test_ntff_matches_oracle_with_image_planes(gpu):
    print("NTFF random patches + PEC/PMC images", flush=True)
    rng = np.random.default_rng(17)
    sets = _random_sets(rng)
    planes = [dict(axis=2, position=0.0, kind="PEC", at_low=True, physical_halfspace=True),
              dict(axis=0, position=-0.01, kind="PMC", at_low=True, physical_halfspace=False)]
    f = 4.5e9
    oth, oph, omask = off.ntff(sets, planes, f, THETA, PHI)
    res = pf.ntff_transform([pf.SurfacePatchSet(**s) for s in sets],
                            [pf.ImagePlane(**p) for p in planes], f, THETA, PHI)
    scale = max(np.abs(oth).max(), np.abs(oph).max())
    err = max(np.abs(res.E_theta - oth).max(), np.abs(res.E_phi - oph).max()) / scale
    print(f"  max |dE| / peak {err:.2e}", flush=True)
    assert err <= 1e-12
    assert np.array_equal(res.physical_mask, omask)
    assert res.physical_sphere_fraction == 0.5


def test_ntff_dipole_closed_form(gpu):
    import test_oracle_farfield as tf  # noqa: PLC0415

    h = 0.35 * tf.LAM
    faces = tf.box_faces((-h, -h, -h), (h, h, h), 60)
    r0 = (0.05 * tf.LAM, -0.03 * tf.LAM, 0.02 * tf.LAM)
    sets = [pf.SurfacePatchSet(**s) for s in tf.sampled(faces, (0, 0, 1), np.asarray(r0))]
    theta, phi = np.linspace(0.0, math.pi, 61), np.linspace(0.0, 2 * math.pi, 121)
    res = pf.ntff_transform(sets, [], tf.FREQ, theta, phi)
    wth, wph = tf.dipole_far(theta, phi, (0, 0, 1), r0)
    scale = np.abs(wth).max()
    assert np.abs(res.E_theta - wth).max() <= 2e-3 * scale
    assert np.abs(res.E_phi - wph).max() <= 2e-3 * scale
    want = off.ETA0 * tf.K**2 * tf.IL**2 / (6 * math.pi)
    assert res.P_rad == pytest.approx(want, rel=5e-3)
    assert res.directivity.max() == pytest.approx(1.5, rel=5e-3)
    with pytest.raises(ValueError, match="accepted"):
        _ = res.gain
    res.accepted_power = 2.0 * want
    assert res.radiation_efficiency == pytest.approx(0.5, rel=5e-3)


def test_ntff_empty_patch_set(gpu):
    z = np.zeros((0, 3))
    res = pf.ntff_transform([pf.SurfacePatchSet(z, z, np.zeros(0), z.astype(complex), z.astype(complex))],
                            [], 1e9, THETA, PHI)
    assert np.all(res.E_theta == 0) and np.all(res.E_phi == 0)


def test_golden_box_far_field_matches_reference_bins(gpu):
    import test_gpu_golden_box as gb  # noqa: PLC0415

    print("golden box far field (double)", flush=True)
    f, mk, sol, region, far = gb._run("double", True)
    far.renormalize(np.array([1.0]), 1.0)  # unit spectrum: the bins as recorded
    theta, phi = np.linspace(0.0, math.pi, 31), np.linspace(0.0, 2 * math.pi, 61)
    dump = far.result_dump(0)
    ref = dict(dump, faces=[dict(fc, bins={c: f[f"double_far_{fc['name']}_{c}"] for c in fc["bins"]})
                            for fc in dump["faces"]])
    for idx in range(len(mk.FREQS)):
        got = far.far_field(f_index=idx, theta=theta, phi=phi)
        oth, oph, _ = off.ntff(off.patch_sets_from_dump(ref, idx), [], float(mk.FREQS[idx]), theta, phi)
        scale = max(np.abs(oth).max(), np.abs(oph).max())
        err = max(np.abs(got.E_theta - oth).max(), np.abs(got.E_phi - oph).max()) / scale
        print(f"  f = {mk.FREQS[idx] / 1e9:.2f} GHz: max |dE| / peak {err:.2e}", flush=True)
        assert err <= 1e-5
    assert cases.golden_box()[0]["dt"] == f["dt"]


def _tensor_dump(rng):
    """A seeded dump of four graded grid faces of a box (no xmin / ymin face:
    those are image planes), one frequency, unit spectrum."""
    faces = []
    for name, axis, sign, plane in (("xmax", 0, 1.0, 0.07), ("ymax", 1, 1.0, 0.05), ("zmin", 2, -1.0, 0.01),
                                    ("zmax", 2, 1.0, 0.09)):
        n1, n2 = int(rng.integers(20, 40)), int(rng.integers(17, 50))
        e1 = np.cumsum(rng.uniform(1e-3, 2e-3, n1 + 1))
        e2 = np.cumsum(rng.uniform(1e-3, 2e-3, n2 + 1))
        shp = (1, n1, n2)
        faces.append(dict(name=name, axis=axis, sign=sign, plane=plane, c1=0.5 * (e1[1:] + e1[:-1]),
                          c2=0.5 * (e2[1:] + e2[:-1]), w1=np.diff(e1), w2=np.diff(e2),
                          bins={c: rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
                                for c in pf.TANGENTIALS[axis]}))
    planes = [dict(axis=0, position=0.0, kind="PEC", at_low=True, physical_halfspace=True),
              dict(axis=1, position=-0.0005, kind="PMC", at_low=True, physical_halfspace=False)]
    return dict(freqs=np.array([7e9]), faces=faces, image_planes=planes, source_spectrum=np.array([1.0 + 0j]))


def test_tensor_face_path_matches_oracle_and_generic_path(gpu):
    print("NTFF tensor faces + images vs oracle and the generic kernel", flush=True)
    dump = _tensor_dump(np.random.default_rng(23))
    theta, phi = np.linspace(0.0, math.pi, 29), np.linspace(0.0, 2 * math.pi, 57)
    sets = pf.patch_sets_from_dump(dump, 0)
    assert all(pf.tensor_face(s) is not None for s in sets)
    planes = pf.image_planes_from_dump(dump)
    res = pf.ntff_transform(sets, planes, 7e9, theta, phi)
    oth, oph, omask = off.ntff(off.patch_sets_from_dump(dump, 0), dump["image_planes"], 7e9, theta, phi)
    scale = max(np.abs(oth).max(), np.abs(oph).max())
    err = max(np.abs(res.E_theta - oth).max(), np.abs(res.E_phi - oph).max()) / scale
    # one set holding every face is no grid: the per-patch kernel
    merged = pf._concat(sets)
    assert pf.tensor_face(merged) is None
    gen = pf.ntff_transform([merged], planes, 7e9, theta, phi)
    err_g = max(np.abs(gen.E_theta - oth).max(), np.abs(gen.E_phi - oph).max()) / scale
    print(f"  face path {err:.2e}, per-patch path {err_g:.2e} of the peak", flush=True)
    assert err <= 1e-12 and err_g <= 1e-12
    assert np.array_equal(res.physical_mask, omask)
