"""Far-field (NTFF) post on the device: Huygens patch sets -> antenna patterns.

Mirrors the reference's post API (src/magnelio/post/far_field.py):
``SurfacePatchSet``, ``ImagePlane``, ``default_angular_grid``,
``ntff_transform(patches, image_planes, f, theta, phi, accepted_power)`` and
``FarFieldResult`` with its derived quantities (:274-390), plus the
monitor-side helpers of monitors/far_field.py — the patch sets of a
``result_dump`` divided by the source spectrum (:325-355,
monitors/_dft.py:64-86) and the image-plane booking of wall faces
(:150-171, post/_symmetry.py:53-84).

The host does the O(patches) work: the image expansion (each plane mirrors
every patch present so far, with the E/H continuation signs), the currents
area (n x H) and -area (n x E), and the direction frames (theta outer, phi
inner).  The O(patches x directions) sums run on the device in FP64: patch
sets that are one grid on a plane (every monitor face and its images) in
``mgh_ntff_faces``, whose phase factors per grid axis, any other set in
``mgh_ntff`` (csrc/mgh_ntff.hip).  No CPU path: without the HIP library the
call raises.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .mesh import C0, EPS0, MU0

ETA0 = math.sqrt(MU0 / EPS0)
TANGENTIALS = {0: ("Ey", "Ez", "Hy", "Hz"), 1: ("Ex", "Ez", "Hx", "Hz"), 2: ("Ex", "Ey", "Hx", "Hy")}


@dataclass
class SurfacePatchSet:
    """Patch centres (n, 3), outward normals (n, 3), areas (n,) and E, H
    (n, 3) library phasors of one piece of a Huygens surface."""

    centers: np.ndarray
    normals: np.ndarray
    areas: np.ndarray
    E: np.ndarray
    H: np.ndarray


@dataclass(frozen=True)
class ImagePlane:
    """A wall the surface does not cross, handled by image theory."""

    axis: int
    position: float
    kind: str  # "PEC" | "PMC"
    at_low: bool
    physical_halfspace: bool = True


def mirror_sign(field: str, comp_axis: int, mirror_axis: int, kind: str) -> float:
    """+-1 continuation of a component across a mirror (post/_symmetry.py:71-84):
    E's normal component is odd across PMC and its tangential ones across
    PEC; H the other way round."""
    is_normal = comp_axis == mirror_axis
    flips_normal = (kind == "PMC") if field == "E" else (kind == "PEC")
    return -1.0 if is_normal == flips_normal else 1.0


def wall_position(face: str, kind: str, axis_nodes) -> float:
    """Mirror plane of a wall face (post/_symmetry.py:53-68): the outer node
    plane for PEC, half a cell beyond it for PMC."""
    n = np.asarray(axis_nodes, float)
    low = face.endswith("min")
    if kind == "PEC":
        return float(n[0] if low else n[-1])
    return float(n[0] - 0.5 * (n[1] - n[0]) if low else n[-1] + 0.5 * (n[-1] - n[-2]))


def default_angular_grid() -> tuple[np.ndarray, np.ndarray]:
    """2 degree (theta, phi) grid over the full sphere."""
    return np.linspace(0.0, np.pi, 91), np.linspace(0.0, 2.0 * np.pi, 181)


def _mirror(p: SurfacePatchSet, plane: ImagePlane) -> SurfacePatchSet:
    a = plane.axis
    c = p.centers.copy()
    c[:, a] = 2.0 * plane.position - c[:, a]
    n = p.normals.copy()
    n[:, a] = -n[:, a]
    E, H = p.E.copy(), p.H.copy()
    for k in range(3):
        E[:, k] *= mirror_sign("E", k, a, plane.kind)
        H[:, k] *= mirror_sign("H", k, a, plane.kind)
    return SurfacePatchSet(c, n, p.areas.copy(), E, H)


def _concat(sets: Sequence[SurfacePatchSet]) -> SurfacePatchSet:
    return SurfacePatchSet(*(np.concatenate([getattr(s, k) for s in sets], axis=0)
                             for k in ("centers", "normals", "areas", "E", "H")))


def direction_frames(theta: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """(n_theta * n_phi, 9): r, theta and phi unit vectors, theta outer."""
    st, ct = np.sin(theta), np.cos(theta)
    sp, cp = np.sin(phi), np.cos(phi)
    one_p, one_t = np.ones_like(sp), np.ones_like(ct)
    return np.stack([np.outer(st, cp).ravel(), np.outer(st, sp).ravel(), np.outer(ct, one_p).ravel(),
                     np.outer(ct, cp).ravel(), np.outer(ct, sp).ravel(), np.outer(-st, one_p).ravel(),
                     np.outer(one_t, -sp).ravel(), np.outer(one_t, cp).ravel(),
                     np.zeros(theta.size * phi.size)], axis=1)


def patch_array(surf: SurfacePatchSet) -> np.ndarray:
    """(n, 15) device patch records: centre, area (n x H), -area (n x E)."""
    wJ = np.cross(surf.normals, surf.H) * surf.areas[:, None]
    wM = -np.cross(surf.normals, surf.E) * surf.areas[:, None]
    out = np.empty((surf.centers.shape[0], 15))
    out[:, 0:3] = surf.centers
    out[:, 3:9:2], out[:, 4:9:2] = wJ.real, wJ.imag
    out[:, 9:15:2], out[:, 10:15:2] = wM.real, wM.imag
    return out


def ntff_chunks(n_patches: int, n_dirs: int) -> int:
    """Patch ranges per launch: enough work-groups to fill the chip (about
    4096 of 256 directions) with at least 512 patches each."""
    tiles = (n_dirs + 255) // 256
    want = max(1, -(-4096 // tiles))
    return int(max(1, min(want, -(-n_patches // 512), 65535)))


FACE_JT = 16  # columns per tile job of mgh_ntff_faces


def tensor_face(surf: SurfacePatchSet):
    """(axis, plane, c1, c2, W) when the patch set is one grid c1 x c2 on a
    plane x_axis = const with one normal, patches i-major (i n2 + j) — the
    layout of a monitor face and of its images; else None.  W (n1, n2, 4)
    holds area (n x H)_t1, _t2 and -area (n x E)_t1, _t2."""
    n = surf.centers.shape[0]
    if n == 0:
        return None
    nrm = surf.normals
    axis = int(np.argmax(np.abs(nrm[0])))
    unit = np.zeros(3)
    unit[axis] = nrm[0, axis]
    if abs(nrm[0, axis]) != 1.0 or not np.all(nrm == unit):
        return None
    t1, t2 = [b for b in range(3) if b != axis]
    cen = surf.centers
    if not np.all(cen[:, axis] == cen[0, axis]):
        return None
    col = cen[:, t1]
    n2 = int(np.argmax(col != col[0])) if np.any(col != col[0]) else n
    if n2 < 1 or n % n2:
        return None
    n1 = n // n2
    g1 = col.reshape(n1, n2)
    g2 = cen[:, t2].reshape(n1, n2)
    if not (np.all(g1 == g1[:, :1]) and np.all(g2 == g2[:1, :])):
        return None
    wJ = np.cross(surf.normals, surf.H) * surf.areas[:, None]
    wM = -np.cross(surf.normals, surf.E) * surf.areas[:, None]
    W = np.stack([wJ[:, t1], wJ[:, t2], wM[:, t1], wM[:, t2]], axis=1).reshape(n1, n2, 4)
    return axis, float(cen[0, axis]), g1[:, 0].copy(), g2[0, :].copy(), W


def _to(dev, a, dtype):
    import torch  # noqa: PLC0415

    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(dev)


class DeviceTransform:
    """Device-resident inputs of one transform — tensor faces (mgh_ntff_faces)
    or a generic patch set (mgh_ntff) — on fixed direction frames; ``launch``
    enqueues the sums for a frequency, ``result`` reads E_theta, E_phi."""

    def __init__(self, frames: np.ndarray, faces=None, patches: Optional[np.ndarray] = None, device=None):
        import torch  # noqa: PLC0415

        self.dev = torch.device(device if device is not None else "cuda")
        self.nd = int(frames.shape[0])
        self.D = _to(self.dev, frames, np.float64)
        self.out = torch.empty((self.nd, 4), dtype=torch.float64, device=self.dev)
        self.faces = faces is not None
        if self.faces:
            desc, planes, coords, ws, jobs = [], [], [], [], []
            off_c = off_w = 0
            for fi, (axis, plane, c1, c2, W) in enumerate(faces):
                n1, n2 = W.shape[0], W.shape[1]
                desc.append([axis, n1, n2, off_c, off_c + n1, off_w])
                planes.append(plane)
                coords += [c1, c2]
                off_c += n1 + n2
                w = np.empty((n1, n2, 8))
                w[..., 0::2], w[..., 1::2] = W.real, W.imag
                ws.append(w.ravel())
                off_w += w.size
                jobs += [[fi, j0, min(FACE_JT, n2 - j0)] for j0 in range(0, n2, FACE_JT)]
            self.n_faces, self.n_jobs = len(faces), len(jobs)
            self.desc = _to(self.dev, np.array(desc), np.int64)
            self.planes = _to(self.dev, np.array(planes), np.float64)
            self.coords = _to(self.dev, np.concatenate(coords), np.float64)
            self.W = _to(self.dev, np.concatenate(ws), np.float64)
            self.jobs = _to(self.dev, np.array(jobs), np.int32)
            self.work = torch.empty(self.n_jobs * 8 * self.nd, dtype=torch.float64, device=self.dev)
            self.pairs = int(sum(W.shape[0] * W.shape[1] for *_, W in faces)) * self.nd
        else:
            self.np_ = int(patches.shape[0])
            self.chunks = ntff_chunks(self.np_, self.nd)
            self.P = _to(self.dev, patches if self.np_ else np.zeros((1, 15)), np.float64)
            self.work = torch.empty(self.chunks * 12 * self.nd, dtype=torch.float64, device=self.dev)
            self.pairs = self.np_ * self.nd

    def launch(self, f: float, stream=None) -> None:
        import torch  # noqa: PLC0415

        st = stream if stream is not None else torch.cuda.current_stream(self.dev)
        k = 2.0 * math.pi * float(f) / C0
        pre = k / (4.0 * math.pi)
        if self.faces:
            _lib.call("mgh_ntff_faces", self.n_faces, self.desc.data_ptr(), self.planes.data_ptr(),
                      self.coords.data_ptr(), self.W.data_ptr(), self.n_jobs, self.jobs.data_ptr(), self.nd,
                      self.D.data_ptr(), k, 0.0, pre, ETA0, self.work.data_ptr(), self.out.data_ptr(),
                      int(st.cuda_stream))
        else:
            _lib.call("mgh_ntff", self.np_, self.P.data_ptr(), self.nd, self.D.data_ptr(), k, 0.0, pre, ETA0,
                      self.chunks, self.work.data_ptr(), self.out.data_ptr(), int(st.cuda_stream))

    def result(self) -> tuple[np.ndarray, np.ndarray]:
        o = self.out.cpu().numpy()
        return o[:, 0] + 1j * o[:, 1], o[:, 2] + 1j * o[:, 3]


def ntff_transform(patches: Sequence[SurfacePatchSet], image_planes: Sequence[ImagePlane], f: float,
                   theta=None, phi=None, accepted_power: Optional[float] = None, *,
                   device=None) -> "FarFieldResult":
    """Far-field pattern of the patch sets at one frequency (the reference's
    ``ntff_transform``, post/far_field.py:139-271), summed on the device."""
    if theta is None or phi is None:
        th_d, ph_d = default_angular_grid()
        theta = th_d if theta is None else theta
        phi = ph_d if phi is None else phi
    theta = np.atleast_1d(np.asarray(theta, float))
    phi = np.atleast_1d(np.asarray(phi, float))
    # the image expansion per input set keeps each set's own structure; the
    # result is the same sum over the same patches as expanding the union
    sets = [s for s in patches if s.centers.shape[0] > 0]
    for plane in image_planes:
        sets += [_mirror(s, plane) for s in sets]
    frames = direction_frames(theta, phi)
    faces, rest = [], []
    for s in sets:
        tf = tensor_face(s)
        (faces if tf is not None else rest).append(tf if tf is not None else s)
    e_th = np.zeros(frames.shape[0], complex)
    e_ph = np.zeros(frames.shape[0], complex)
    parts = []
    if faces:
        parts.append(DeviceTransform(frames, faces=faces, device=device))
    if rest or not faces:
        recs = patch_array(_concat(rest)) if rest else np.zeros((0, 15))
        parts.append(DeviceTransform(frames, patches=recs, device=device))
    for t in parts:
        t.launch(f)
        a, b = t.result()
        e_th, e_ph = e_th + a, e_ph + b
    mask = None
    physical = [p for p in image_planes if p.physical_halfspace]
    if physical:
        mask = np.ones(frames.shape[0], bool)
        for p in physical:
            inward = 1.0 if p.at_low else -1.0
            mask &= inward * frames[:, p.axis] >= -1e-12
        mask = mask.reshape(theta.size, phi.size)
    shape = (theta.size, phi.size)
    return FarFieldResult(f=float(f), theta=theta, phi=phi, E_theta=e_th.reshape(shape),
                          E_phi=e_ph.reshape(shape), accepted_power=accepted_power,
                          physical_mask=mask, physical_sphere_fraction=0.5 ** len(physical))


@dataclass
class FarFieldResult:
    """Far-zone amplitudes of one frequency (post/far_field.py:274-390): the
    field at distance r is (E_theta theta + E_phi phi) times the outgoing
    spherical factor over r; effective (RMS) phasors per sqrt(W) of incident
    power."""

    f: float
    theta: np.ndarray
    phi: np.ndarray
    E_theta: np.ndarray
    E_phi: np.ndarray
    accepted_power: Optional[float] = None
    physical_mask: Optional[np.ndarray] = None
    physical_sphere_fraction: float = 1.0

    @property
    def _U_unmasked(self) -> np.ndarray:
        return (np.abs(self.E_theta) ** 2 + np.abs(self.E_phi) ** 2) / ETA0

    @property
    def U(self) -> np.ndarray:
        """Radiation intensity [W/sr], zero behind a ground plane."""
        u = self._U_unmasked
        return np.where(self.physical_mask, u, 0.0) if self.physical_mask is not None else u

    @property
    def P_rad(self) -> float:
        """Radiated power [W] over the physical sphere (trapezoid rule)."""
        integrand = self._U_unmasked * np.sin(self.theta)[:, None]
        full = float(np.trapezoid(np.trapezoid(integrand, self.phi, axis=1), self.theta))
        return full * self.physical_sphere_fraction

    @property
    def directivity(self) -> np.ndarray:
        return 4.0 * np.pi * self.U / self.P_rad

    @property
    def realized_gain(self) -> np.ndarray:
        """4 pi U per incident watt (mismatch included)."""
        return 4.0 * np.pi * self.U

    @property
    def gain(self) -> np.ndarray:
        if self.accepted_power is None:
            raise ValueError("gain needs the accepted input power; use realized_gain or directivity, "
                             "or set accepted_power on the result")
        return self.realized_gain / self.accepted_power

    @property
    def radiation_efficiency(self) -> float:
        if self.accepted_power is None:
            raise ValueError("radiation_efficiency needs the accepted input power; set accepted_power "
                             "on the result")
        return self.P_rad / self.accepted_power

    def _quantity(self, quantity: str) -> np.ndarray:
        table = {"realized_gain": lambda: self.realized_gain, "gain": lambda: self.gain,
                 "directivity": lambda: self.directivity, "U": lambda: self.U}
        if quantity not in table:
            raise ValueError(f"unknown quantity {quantity!r}; expected one of {sorted(table)}")
        return table[quantity]()

    def cut(self, *, plane: str = "phi", angle: float = 0.0, quantity: str = "realized_gain"):
        """One pattern cut: fixed azimuth over theta (and the back half at
        angle + pi), or fixed theta over phi."""
        values = self._quantity(quantity)
        if plane == "theta":
            i = int(np.argmin(np.abs(self.theta - angle)))
            return self.phi, values[i, :]
        if plane != "phi":
            raise ValueError(f"plane must be 'phi' or 'theta', got {plane!r}")
        jf = int(np.argmin(np.abs(self.phi - (angle % (2.0 * np.pi)))))
        jb = int(np.argmin(np.abs(self.phi - ((angle + np.pi) % (2.0 * np.pi)))))
        angles = np.concatenate([self.theta, 2.0 * np.pi - self.theta[-2::-1]])
        return angles, np.concatenate([values[:, jf], values[-2::-1, jb]])


def source_spectrum(values, dt: float, freqs) -> np.ndarray:
    """sum_n v[n] exp(+j w n dt) dt — the monitors' own DFT of the source
    (monitors/_dft.py:64-74), the divisor of the 1 W renormalisation."""
    values = np.asarray(values, float)
    t = np.arange(values.size) * dt
    w = 2.0 * np.pi * np.asarray(freqs, float)
    return (np.exp(1j * w[:, None] * t[None, :]) * dt) @ values


def patch_sets_from_dump(dump: dict, idx: int) -> list[SurfacePatchSet]:
    """Patch sets of frequency index idx of a far-field monitor dump (the
    reference's schema, monitors/far_field.py:403-452), divided by its
    source spectrum (zero where the spectrum underflows)."""
    if dump.get("source_spectrum") is None:
        raise ValueError("far-field data is only meaningful per watt of incident power; renormalize "
                         "the monitor with the run's source signal first")
    spec = complex(np.asarray(dump["source_spectrum"])[idx])
    faces = dump["faces"].values() if isinstance(dump["faces"], dict) else dump["faces"]
    out = []
    for fc in faces:
        a = int(fc["axis"])
        t1, t2 = [b for b in range(3) if b != a]
        comps = TANGENTIALS[a]
        vals = {}
        for c in comps:
            b = np.asarray(fc["bins"][c])[idx]
            vals[c] = b / spec if abs(spec) > 1e-300 else np.zeros_like(b)
        g1, g2 = np.meshgrid(np.asarray(fc["c1"], float), np.asarray(fc["c2"], float), indexing="ij")
        n = g1.size
        cen = np.zeros((n, 3))
        cen[:, a], cen[:, t1], cen[:, t2] = float(fc["plane"]), g1.ravel(), g2.ravel()
        nrm = np.zeros((n, 3))
        nrm[:, a] = float(fc["sign"])
        E = np.zeros((n, 3), complex)
        H = np.zeros((n, 3), complex)
        E[:, t1], E[:, t2] = vals[comps[0]].ravel(), vals[comps[1]].ravel()
        H[:, t1], H[:, t2] = vals[comps[2]].ravel(), vals[comps[3]].ravel()
        out.append(SurfacePatchSet(cen, nrm, np.outer(fc["w1"], fc["w2"]).ravel(), E, H))
    return out


def image_planes_from_dump(dump: dict) -> list[ImagePlane]:
    return [ImagePlane(int(p["axis"]), float(p["position"]), str(p["kind"]), bool(p["at_low"]),
                       bool(p.get("physical_halfspace", True))) for p in dump.get("image_planes", [])]


def frequency_index(freqs, f: Optional[float], f_index: Optional[int]) -> int:
    """The recorded frequency picked by f or f_index (monitors/far_field.py:287-305)."""
    freqs = np.asarray(freqs, float)
    if f_index is not None:
        return int(f_index)
    if f is None:
        if freqs.size == 1:
            return 0
        raise ValueError(f"this monitor recorded {freqs.size} frequencies; pass f= or f_index=")
    idx = int(np.argmin(np.abs(freqs - f)))
    if abs(freqs[idx] - f) > 1e-6 * max(abs(f), 1.0):
        raise ValueError(f"frequency {f:.6g} Hz was not recorded")
    return idx


def far_field_from_dump(dump: dict, f: Optional[float] = None, *, f_index: Optional[int] = None,
                        theta=None, phi=None, device=None) -> FarFieldResult:
    """The far-field pattern at one recorded frequency of a monitor dump —
    the reference monitor's ``result`` (monitors/far_field.py:357-391)."""
    idx = frequency_index(dump["freqs"], f, f_index)
    accepted = None
    if dump.get("accepted_power") is not None:
        accepted = float(np.asarray(dump["accepted_power"])[idx])
    return ntff_transform(patch_sets_from_dump(dump, idx), image_planes_from_dump(dump),
                          float(np.asarray(dump["freqs"])[idx]), theta=theta, phi=phi,
                          accepted_power=accepted, device=device)
