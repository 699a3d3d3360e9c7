"""Oracle: near-to-far-field transform on a closed Huygens surface (CPU).

Test infrastructure: only the tests, the smoke check and the benchmark's CPU
baseline use this module, as the checker of the device transform.

Restates (paths relative to the reference repository root):

* the transform          src/magnelio/post/far_field.py:139-271: library
                         phasors (e^{-jwt}) are conjugated on entry, the
                         image planes double the patch set (each plane
                         mirrors every patch present so far), J = n x H,
                         M = -n x E weighted by the patch area, the
                         radiation vectors N = sum J e^{+jk r.r'},
                         L = sum M e^{+jk r.r'} over directions theta-outer /
                         phi-inner, A_th = -jk/(4 pi) (eta0 N_th + L_ph),
                         A_ph = -jk/(4 pi) (eta0 N_ph - L_th), conjugated
                         back; directions behind a physical image plane
                         masked
* the image signs        src/magnelio/post/_symmetry.py:71-84 (E: normal
                         component odd across PMC, tangential odd across
                         PEC; H the other way round) and the wall position
                         :53-68 (PEC on the outer node plane, PMC half a
                         cell outside it)
* the result quantities  post/far_field.py:274-390 (U = (|E_th|^2 +
                         |E_ph|^2)/eta0, P_rad by the trapezoid rule over
                         theta and phi times the physical sphere fraction,
                         directivity, realized gain, gain)
* the patch sets of a    monitors/far_field.py:325-355 (bins divided by the
  monitor dump           source spectrum, zero where it underflows;
                         patch centres on the face node plane, area
                         w1 w2), monitors/_dft.py:64-86

Pinned by tests/test_oracle_farfield.py against what the mathematics fixes:
the exact near fields of a Hertzian dipole sampled on a closed box give the
dipole's closed-form far field (amplitude, the sin(theta) shape and the
phasor's phase); the same dipole over a PEC and over a PMC ground plane,
box open at the plane, gives the closed-form two-element array; the
dipole's radiated power and peak directivity 1.5 follow from the pattern.
"""

from __future__ import annotations

import math

import numpy as np

C0 = 299_792_458.0
MU0 = 1.25663706212e-6
EPS0 = 1.0 / (MU0 * C0 * C0)
ETA0 = math.sqrt(MU0 / EPS0)


def mirror_sign(field: str, comp_axis: int, mirror_axis: int, kind: str) -> float:
    is_normal = comp_axis == mirror_axis
    flips_normal = (kind == "PMC") if field == "E" else (kind == "PEC")
    return -1.0 if is_normal == flips_normal else 1.0


def wall_position(face: str, kind: str, axis_nodes) -> float:
    n = np.asarray(axis_nodes, float)
    low = face.endswith("min")
    if kind == "PEC":
        return float(n[0] if low else n[-1])
    return float(n[0] - 0.5 * (n[1] - n[0]) if low else n[-1] + 0.5 * (n[-1] - n[-2]))


def mirror(p: dict, plane: dict) -> dict:
    """The image copy of patch set p (dict centers, normals, areas, E, H)."""
    a = plane["axis"]
    c = p["centers"].copy()
    c[:, a] = 2.0 * plane["position"] - c[:, a]
    nrm = p["normals"].copy()
    nrm[:, a] = -nrm[:, a]
    E, H = p["E"].copy(), p["H"].copy()
    for k in range(3):
        E[:, k] *= mirror_sign("E", k, a, plane["kind"])
        H[:, k] *= mirror_sign("H", k, a, plane["kind"])
    return dict(centers=c, normals=nrm, areas=p["areas"].copy(), E=E, H=H)


def concat(sets):
    return {k: np.concatenate([s[k] for s in sets], axis=0) for k in ("centers", "normals", "areas", "E", "H")}


def directions(theta, phi):
    st, ct = np.sin(theta), np.cos(theta)
    sp, cp = np.sin(phi), np.cos(phi)
    r = np.stack([np.outer(st, cp).ravel(), np.outer(st, sp).ravel(), np.outer(ct, np.ones_like(sp)).ravel()], 1)
    th = np.stack([np.outer(ct, cp).ravel(), np.outer(ct, sp).ravel(), np.outer(-st, np.ones_like(sp)).ravel()], 1)
    ph = np.stack([np.outer(np.ones_like(ct), -sp).ravel(), np.outer(np.ones_like(ct), cp).ravel(),
                   np.zeros(theta.size * phi.size)], 1)
    return r, th, ph


def ntff(patch_sets, image_planes, f, theta, phi):
    """(E_theta, E_phi, mask) on the (theta, phi) grid; patch sets are dicts
    with centers, normals, areas, E, H (library phasors)."""
    theta = np.atleast_1d(np.asarray(theta, float))
    phi = np.atleast_1d(np.asarray(phi, float))
    s = concat(list(patch_sets))
    s = dict(s, E=np.conj(s["E"]), H=np.conj(s["H"]))
    sets = [s]
    for pl in image_planes:
        sets += [mirror(q, pl) for q in sets]
    s = concat(sets)
    wJ = np.cross(s["normals"], s["H"]) * s["areas"][:, None]
    wM = -np.cross(s["normals"], s["E"]) * s["areas"][:, None]
    k = 2.0 * math.pi * float(f) / C0
    r, th, ph = directions(theta, phi)
    N = np.empty((r.shape[0], 3), complex)
    L = np.empty((r.shape[0], 3), complex)
    for lo in range(0, r.shape[0], 128):
        hi = min(lo + 128, r.shape[0])
        phase = np.exp(1j * k * (r[lo:hi] @ s["centers"].T))
        N[lo:hi] = phase @ wJ
        L[lo:hi] = phase @ wM
    pre = -1j * k / (4.0 * math.pi)
    a_th = pre * (ETA0 * np.sum(N * th, 1) + np.sum(L * ph, 1))
    a_ph = pre * (ETA0 * np.sum(N * ph, 1) - np.sum(L * th, 1))
    mask = None
    phys = [p for p in image_planes if p.get("physical_halfspace", True)]
    if phys:
        mask = np.ones(r.shape[0], bool)
        for p in phys:
            inward = 1.0 if p["at_low"] else -1.0
            mask &= inward * r[:, p["axis"]] >= -1e-12
        mask = mask.reshape(theta.size, phi.size)
    shape = (theta.size, phi.size)
    return np.conj(a_th).reshape(shape), np.conj(a_ph).reshape(shape), mask


def radiated_power(E_th, E_ph, theta, phi, fraction=1.0):
    """P_rad = fraction * trapz_theta trapz_phi (U sin theta)."""
    U = (np.abs(E_th) ** 2 + np.abs(E_ph) ** 2) / ETA0
    return float(np.trapezoid(np.trapezoid(U * np.sin(theta)[:, None], phi, axis=1), theta)) * fraction


def source_spectrum(values, dt, freqs):
    t = np.arange(len(values)) * dt
    w = 2.0 * math.pi * np.asarray(freqs, float)
    return (np.exp(1j * w[:, None] * t[None, :]) * dt) @ np.asarray(values, float)


def patch_sets_from_dump(dump: dict, idx: int):
    """Patch sets of frequency index idx of a far-field monitor dump."""
    spec = np.asarray(dump["source_spectrum"])[idx]
    tang = {0: ("Ey", "Ez", "Hy", "Hz"), 1: ("Ex", "Ez", "Hx", "Hz"), 2: ("Ex", "Ey", "Hx", "Hy")}
    out = []
    for fc in dump["faces"]:
        a = int(fc["axis"])
        t1, t2 = [b for b in range(3) if b != a]
        comps = tang[a]
        v = {c: (np.asarray(fc["bins"][c])[idx] / spec if abs(spec) > 1e-300
                 else np.zeros_like(np.asarray(fc["bins"][c])[idx])) for c in comps}
        g1, g2 = np.meshgrid(fc["c1"], fc["c2"], indexing="ij")
        n = g1.size
        cen = np.zeros((n, 3))
        cen[:, a] = fc["plane"]
        cen[:, t1], cen[:, t2] = g1.ravel(), g2.ravel()
        nrm = np.zeros((n, 3))
        nrm[:, a] = fc["sign"]
        E = np.zeros((n, 3), complex)
        H = np.zeros((n, 3), complex)
        E[:, t1], E[:, t2] = v[comps[0]].ravel(), v[comps[1]].ravel()
        H[:, t1], H[:, t2] = v[comps[2]].ravel(), v[comps[3]].ravel()
        out.append(dict(centers=cen, normals=nrm, areas=np.outer(fc["w1"], fc["w2"]).ravel(), E=E, H=H))
    return out
