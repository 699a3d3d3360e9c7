"""Pins for the oracle's near-to-far-field transform (CPU).

The transform is checked against what the mathematics fixes, not against a
restatement of itself:

* the exact near fields of a Hertzian dipole (vector form, any orientation
  and position) sampled on a closed box give the dipole's closed-form far
  field — amplitude, shape, the e^{jk r.r0} position phase and the library's
  e^{-jwt} phasor convention (the conjugate of the textbook amplitude);
* a horizontal dipole over a PEC and over a PMC ground plane, the box open
  at the plane and the plane booked as an image plane, gives the
  two-element array of the dipole and its image (reversed over PEC, equal
  over PMC), with the directions below the plane masked;
* the radiated power and the peak directivity 1.5 of the dipole pattern;
* the source-spectrum division of a monitor dump (a sampled sinusoid's
  spectrum divides its own bins to 1).
"""

import math

import numpy as np
import pytest

from oracle import farfield as off

FREQ = 3e9
LAM = off.C0 / FREQ
K = 2 * math.pi / LAM
IL = 1e-3  # current moment I l [A m]


def dipole_fields(pts, p_hat, r0):
    """Textbook (e^{+jwt}) E and H of a z'-dipole of moment IL along p_hat at r0."""
    d = pts - r0
    r = np.linalg.norm(d, axis=1)[:, None]
    rh = d / r
    p = np.asarray(p_hat, float)[None, :]
    g = np.exp(-1j * K * r)
    pr = np.sum(rh * p, 1)[:, None]
    E = off.ETA0 * IL / (4 * math.pi) * g * ((1j * K / r) * (1 + 1 / (1j * K * r) - 1 / (K * r) ** 2)
                                             * (pr * rh - p) + (2 / r**2) * (1 + 1 / (1j * K * r)) * pr * rh)
    H = 1j * K * IL / (4 * math.pi * r) * (1 + 1 / (1j * K * r)) * g * np.cross(p, rh)
    return E, H


def dipole_far(theta, phi, p_hat, r0):
    """Library-convention far amplitudes conj(jk eta IL/(4 pi) (-p.th) e^{jk r.r0})."""
    r, th, ph = off.directions(theta, phi)
    p = np.asarray(p_hat, float)
    ph0 = np.exp(1j * K * (r @ np.asarray(r0, float)))
    a = 1j * K * off.ETA0 * IL / (4 * math.pi) * ph0
    shape = (theta.size, phi.size)
    return np.conj(-a * (th @ p)).reshape(shape), np.conj(-a * (ph @ p)).reshape(shape)


def box_faces(lo, hi, n, omit=()):
    """Patch sets (centres, normals, areas) of the closed box [lo, hi], n x n
    midpoint patches per face; faces named in omit are left out."""
    out = []
    for a in range(3):
        t1, t2 = [b for b in range(3) if b != a]
        e1 = np.linspace(lo[t1], hi[t1], n + 1)
        e2 = np.linspace(lo[t2], hi[t2], n + 1)
        c1, c2 = 0.5 * (e1[1:] + e1[:-1]), 0.5 * (e2[1:] + e2[:-1])
        g1, g2 = np.meshgrid(c1, c2, indexing="ij")
        for side, pos in ((-1.0, lo[a]), (1.0, hi[a])):
            if ("xyz"[a] + ("min" if side < 0 else "max")) in omit:
                continue
            cen = np.zeros((g1.size, 3))
            cen[:, a], cen[:, t1], cen[:, t2] = pos, g1.ravel(), g2.ravel()
            nrm = np.zeros_like(cen)
            nrm[:, a] = side
            out.append(dict(centers=cen, normals=nrm, areas=np.full(g1.size, np.diff(e1)[0] * np.diff(e2)[0])))
    return out


def sampled(faces, p_hat, r0, images=()):
    sets = []
    for fc in faces:
        E, H = dipole_fields(fc["centers"], p_hat, r0)
        for pi, ri in images:
            Ei, Hi = dipole_fields(fc["centers"], pi, ri)
            E, H = E + Ei, H + Hi
        sets.append(dict(fc, E=np.conj(E), H=np.conj(H)))  # library phasors
    return sets


THETA = np.linspace(0.0, math.pi, 19)
PHI = np.linspace(0.0, 2 * math.pi, 25)


@pytest.mark.parametrize("p_hat,r0", [((0, 0, 1), (0, 0, 0)), ((1, 0, 0), (0.05 * LAM, -0.03 * LAM, 0.02 * LAM)),
                                      ((0, 0.6, 0.8), (0.0, 0.04 * LAM, 0.0))])
def test_dipole_far_field_from_box(p_hat, r0):
    h = 0.35 * LAM
    faces = box_faces((-h, -h, -h), (h, h, h), 60)
    eth, eph, mask = off.ntff(sampled(faces, p_hat, np.asarray(r0, float)), [], FREQ, THETA, PHI)
    wth, wph = dipole_far(THETA, PHI, p_hat, r0)
    scale = max(np.abs(wth).max(), np.abs(wph).max())
    assert mask is None
    # midpoint quadrature of a smooth integrand over a 0.7-wavelength box
    assert np.abs(eth - wth).max() <= 2e-3 * scale
    assert np.abs(eph - wph).max() <= 2e-3 * scale


@pytest.mark.parametrize("kind,image_sign", [("PEC", -1.0), ("PMC", 1.0)])
def test_ground_plane_images(kind, image_sign):
    """Horizontal dipole at height hd over a ground plane at z = 0."""
    hd = 0.15 * LAM
    h = 0.35 * LAM
    p = (1.0, 0.0, 0.0)
    r0 = np.array([0.0, 0.0, hd])
    ri = np.array([0.0, 0.0, -hd])
    pi_ = tuple(image_sign * np.asarray(p))
    faces = box_faces((-h, -h, 0.0), (h, h, 0.5 * LAM), 60, omit=("zmin",))
    # the fields on the box are those of the dipole with its image (the
    # physical fields above the plane)
    sets = sampled(faces, p, r0, images=[(pi_, ri)])
    plane = dict(axis=2, position=0.0, kind=kind, at_low=True, physical_halfspace=True)
    eth, eph, mask = off.ntff(sets, [plane], FREQ, THETA, PHI)
    w1 = dipole_far(THETA, PHI, p, r0)
    w2 = dipole_far(THETA, PHI, pi_, ri)
    wth, wph = w1[0] + w2[0], w1[1] + w2[1]
    up = mask
    assert up.shape == (THETA.size, PHI.size) and up[0].all() and not up[-1].any()
    scale = max(np.abs(wth).max(), np.abs(wph).max())
    assert np.abs(eth - wth)[up].max() <= 3e-3 * scale
    assert np.abs(eph - wph)[up].max() <= 3e-3 * scale


def test_dipole_power_and_directivity():
    theta = np.linspace(0.0, math.pi, 91)
    phi = np.linspace(0.0, 2 * math.pi, 181)
    eth, eph = dipole_far(theta, phi, (0, 0, 1), (0, 0, 0))
    P = off.radiated_power(eth, eph, theta, phi)
    want = off.ETA0 * K**2 * IL**2 / (6 * math.pi)
    assert P == pytest.approx(want, rel=2e-4)
    U = (np.abs(eth) ** 2 + np.abs(eph) ** 2) / off.ETA0
    assert 4 * math.pi * U.max() / P == pytest.approx(1.5, rel=2e-4)
    assert off.radiated_power(eth, eph, theta, phi, 0.5) == pytest.approx(0.5 * P, rel=1e-15)


def test_wall_positions_and_signs():
    nodes = np.array([0.0, 1.0, 3.0])
    assert off.wall_position("xmin", "PEC", nodes) == 0.0
    assert off.wall_position("xmax", "PEC", nodes) == 3.0
    assert off.wall_position("xmin", "PMC", nodes) == -0.5
    assert off.wall_position("xmax", "PMC", nodes) == 4.0
    # tangential E odd / normal E even across PEC; H the other way round
    assert off.mirror_sign("E", 0, 2, "PEC") == -1.0 and off.mirror_sign("E", 2, 2, "PEC") == 1.0
    assert off.mirror_sign("H", 0, 2, "PEC") == 1.0 and off.mirror_sign("H", 2, 2, "PEC") == -1.0
    assert off.mirror_sign("E", 0, 2, "PMC") == 1.0 and off.mirror_sign("E", 2, 2, "PMC") == -1.0


def test_dump_patch_sets_divide_by_the_spectrum():
    dt = 1e-12
    n = np.arange(4000)
    wave = np.sin(2 * math.pi * FREQ * n * dt) * np.exp(-(((n - 2000) / 600.0) ** 2))
    spec = off.source_spectrum(wave, dt, [FREQ])
    bins = {c: (2.0 * spec).reshape(1, 1, 1) * np.ones((1, 2, 3)) for c in ("Ex", "Ey", "Hx", "Hy")}
    dump = dict(faces=[dict(axis=2, sign=1.0, plane=0.5, c1=np.array([0.1, 0.2]),
                            c2=np.array([0.0, 0.1, 0.2]), w1=np.array([0.1, 0.1]),
                            w2=np.array([0.1, 0.1, 0.1]), bins=bins)], source_spectrum=spec)
    (ps,) = off.patch_sets_from_dump(dump, 0)
    assert np.allclose(ps["E"][:, :2], 2.0) and np.all(ps["E"][:, 2] == 0)
    assert np.all(ps["centers"][:, 2] == 0.5) and np.all(ps["normals"][:, 2] == 1.0)
    assert np.allclose(ps["areas"], 0.01)
