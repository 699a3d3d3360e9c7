"""Boundary conditions of the bbox faces: PEC, PMC and CFS-PML (CPML).

Reference: src/magnelio/boundaries/ (pec.py:50-90, pmc.py:34-68,
cpml.py:37-548).

* PEC faces are folded into the update coefficients (their tangential
  edges carry alpha_E = beta_E = 0, src/magnelio/solver/fit_td.py:499-513),
  so a PEC face needs no device work; its ``apply_E`` zeroing is then a
  no-op and is not launched.
* PMC is the natural (no-op) boundary of the FIT update.
* CPML keeps the auxiliary psi fields of its slab on the device; their
  recursion runs inside the fused curl kernels (one descriptor per run,
  faces applied in the order the boundary dict lists them).  Like the
  reference, the CPML outer-face PEC ``apply`` is not part of the step
  (the solver never calls it, fit_td.py:870-882).
"""

from __future__ import annotations

import numpy as np

from .mesh import C0, EPS0

FACE_AXIS = {"xmin": (0, "min"), "xmax": (0, "max"), "ymin": (1, "min"),
             "ymax": (1, "max"), "zmin": (2, "min"), "zmax": (2, "max")}

# Field components of the two psi terms per normal axis (E side, H side);
# must match the term order of the HIP kernels (csrc/mgh_cpml.h).
E_TERM_COMPS = {0: (1, 2), 1: (0, 2), 2: (0, 1)}
H_TERM_COMPS = {0: (1, 2), 1: (0, 2), 2: (0, 1)}


class PECBoundary:
    """Tangential E = 0 on a bbox face (folded into alpha/beta)."""

    kind = "PEC"

    def __init__(self, face: str):
        if face not in FACE_AXIS:
            raise ValueError(f"Unknown face: {face!r}")
        self.face = face


class PMCBoundary:
    """Tangential H = 0 on a bbox face: the FIT update's natural boundary."""

    kind = "PMC"

    def __init__(self, face: str):
        if face not in FACE_AXIS:
            raise ValueError(f"Unknown face: {face!r}")
        self.face = face


class CPMLBoundary:
    """CFS-PML on one bbox face (reference defaults, cpml.py:66-75).

    Profiles (cpml.py:119-145), sampled at cell centres of the PML layers:
      sigma = sigma_max rho^m, kappa = 1 + (kappa_max - 1) rho^m,
      alpha = alpha_max (1 - rho), sigma_max = -(m+1) C0 EPS0 ln(R) / (2 d)
      b = exp(-(sigma/kappa + alpha) dt/EPS0)
      c = sigma / (kappa sigma + kappa^2 alpha) (1 - b)   (0 where sigma = 0)
      ck = 1 - 1/kappa
    """

    kind = "CPML"

    def __init__(self, face: str, grid, thickness_cells: int = 8, m: int = 3,
                 R_target: float = 1e-8, kappa_max: float = 7.0, alpha_max: float = 0.02):
        if face not in FACE_AXIS:
            raise ValueError(f"Unknown face: {face!r}")
        self.face = face
        self.grid = grid
        self.thickness_cells = thickness_cells
        self.m = m
        self.R_target = R_target
        self.kappa_max = kappa_max
        self.alpha_max = alpha_max
        self.axis, self.side = FACE_AXIS[face]
        self._initialized = False

    def initialize(self, dt: float) -> None:
        d_arr = (self.grid.dx, self.grid.dy, self.grid.dz)[self.axis]
        n_axis = d_arr.size
        npml = min(self.thickness_cells, n_axis)
        cells = np.arange(npml) if self.side == "min" else np.arange(n_axis - npml, n_axis)
        # Scalar, in-order accumulation: the profile values must round exactly
        # like the reference's per-layer loop (a pairwise array sum or an
        # array power would differ in the last bit).
        d_phys = 0.0
        for c_idx in cells:
            d_phys += float(d_arr[c_idx])
        d_phys = d_phys or 1.0
        sigma_max = -(self.m + 1) * C0 * EPS0 * np.log(self.R_target) / (2 * d_phys)
        sigma, kappa, alpha = np.zeros(npml), np.ones(npml), np.zeros(npml)
        cum = 0.0
        for li, c_idx in enumerate(cells):
            half = float(d_arr[c_idx]) / 2.0
            rho = 1.0 - (cum + half) / d_phys if self.side == "min" else (cum + half) / d_phys
            cum += float(d_arr[c_idx])
            sigma[li] = sigma_max * rho**self.m
            kappa[li] = 1.0 + (self.kappa_max - 1.0) * rho**self.m
            alpha[li] = self.alpha_max * (1.0 - rho)
        self.b = np.exp(-(sigma / kappa + alpha) * dt / EPS0)
        with np.errstate(divide="ignore", invalid="ignore"):
            self.c = np.where(sigma > 0.0, sigma / (kappa * sigma + kappa**2 * alpha) * (1.0 - self.b), 0.0)
        self.ck = 1.0 - 1.0 / kappa
        self.sigma = sigma
        self.cells = cells
        self.k0 = int(cells[0])
        self.n = int(npml)
        self._initialized = True

    @property
    def sigma_per_cell(self) -> np.ndarray:
        if not self._initialized:
            raise RuntimeError("CPMLBoundary.initialize() must be called first.")
        return self.sigma.copy()

    def psi_shapes(self, dims) -> tuple[list, list]:
        """(E-term shapes, H-term shapes) of the psi slabs, (n0, n1, n2) order
        (known before ``initialize``: the slab is min(thickness, cells) deep)."""
        nx, ny, nz = dims
        e_sh = [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)]
        h_sh = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]

        n = self.n if self._initialized else min(self.thickness_cells, dims[self.axis])

        def cut(s):
            s = list(s)
            s[self.axis] = n
            return tuple(s)

        return ([cut(e_sh[c]) for c in E_TERM_COMPS[self.axis]],
                [cut(h_sh[c]) for c in H_TERM_COMPS[self.axis]])

    def air_masks(self, pec_e_flat: np.ndarray, pec_cell: np.ndarray, dims):
        """0/1 uint8 masks (E0, E1, H0, H1) in slab (n0, n1, n2) order, or None
        when the slab holds no PEC (cpml.py:213-310)."""
        nx, ny, nz = dims
        n_ex = nx * (ny + 1) * (nz + 1)
        n_ey = (nx + 1) * ny * (nz + 1)
        pe = [pec_e_flat[:n_ex].reshape(nx, ny + 1, nz + 1),
              pec_e_flat[n_ex:n_ex + n_ey].reshape(nx + 1, ny, nz + 1),
              pec_e_flat[n_ex + n_ey:].reshape(nx + 1, ny + 1, nz)]
        ph = [np.zeros((nx + 1, ny, nz), bool), np.zeros((nx, ny + 1, nz), bool),
              np.zeros((nx, ny, nz + 1), bool)]
        if pec_cell is not None and pec_cell.any():
            ph[0][:-1] |= pec_cell
            ph[0][1:] |= pec_cell
            ph[1][:, :-1] |= pec_cell
            ph[1][:, 1:] |= pec_cell
            ph[2][:, :, :-1] |= pec_cell
            ph[2][:, :, 1:] |= pec_cell
        out = []
        any_pec = False
        for comps, src in ((E_TERM_COMPS[self.axis], pe), (H_TERM_COMPS[self.axis], ph)):
            for c in comps:
                m = np.take(src[c], self.cells, axis=self.axis)
                any_pec |= bool(m.any())
                out.append((~m).astype(np.uint8))
        return out if any_pec else None


class DeviceCPML:
    """Device-resident psi state of every CPML face + the kernel descriptor.

    ``faces`` are initialized :class:`CPMLBoundary` objects in application
    order.  psi slabs are laid out x-fastest like the fields, one slab per
    field set, so the fused kernels index them from the element (i, j, k).
    """

    WORDS = 9

    def __init__(self, faces, dims, nsets: int, dtype, device, pec_e_flat=None, pec_cell=None):
        import torch  # noqa: PLC0415

        self.faces = list(faces)
        if len(self.faces) > 6:
            raise ValueError("at most six CPML faces")
        self.dims = dims
        self.nsets = nsets
        self.dtype = dtype
        self.psi_e: list[list] = []
        self.psi_h: list[list] = []
        self._air_bytes = [0, 0]  # uint8 air-mask words of the E / H terms
        self._keep = []
        desc = np.zeros(1 + self.WORDS * len(self.faces), dtype=np.int64)
        desc[0] = len(self.faces)
        for fi, f in enumerate(self.faces):
            if not f._initialized:
                raise RuntimeError(f"CPML face {f.face} is not initialized")
            coef = torch.tensor(np.concatenate([f.b, f.c, f.ck]), dtype=dtype, device=device)
            e_sh, h_sh = f.psi_shapes(dims)
            # one buffer per side: term 0 of every set, then term 1
            views = []
            for shapes in (e_sh, h_sh):
                sizes = [int(np.prod(sh)) for sh in shapes]
                buf = torch.zeros(nsets * sum(sizes), dtype=dtype, device=device)
                views.append((buf, [buf[: nsets * sizes[0]].view(nsets, sizes[0]),
                                    buf[nsets * sizes[0]:].view(nsets, sizes[1])]))
            (pe_buf, pe), (ph_buf, ph) = views
            masks = f.air_masks(pec_e_flat, pec_cell, dims) if pec_e_flat is not None else None
            air = [None, None]
            if masks is not None:
                flat = [np.ascontiguousarray(np.moveaxis(m, (0, 1, 2), (2, 1, 0)).ravel()) for m in masks]
                air = [torch.from_numpy(np.concatenate(flat[0:2])).to(device),
                       torch.from_numpy(np.concatenate(flat[2:4])).to(device)]
            self.psi_e.append(pe)
            self.psi_h.append(ph)
            self._air_bytes[0] += air[0].numel() if air[0] is not None else 0
            self._air_bytes[1] += air[1].numel() if air[1] is not None else 0
            self._keep += [coef, pe_buf, ph_buf] + [a for a in air if a is not None]
            r = desc[1 + fi * self.WORDS: 1 + (fi + 1) * self.WORDS]
            r[0], r[1], r[2] = f.axis, f.k0, f.n
            r[3] = coef.data_ptr()
            r[4], r[5] = pe_buf.data_ptr(), ph_buf.data_ptr()
            r[6] = air[0].data_ptr() if air[0] is not None else 0
            r[7] = air[1].data_ptr() if air[1] is not None else 0
            r[8] = nsets
        self.desc = desc

    @property
    def desc_ptr(self) -> int:
        return int(self.desc.ctypes.data)

    def x_split(self) -> bool:
        """Both x slabs at the domain ends with columns between them: the
        curl launches split their planes and carry the x slabs themselves."""
        nx = self.dims[0]
        xs = [f for f in self.faces if f.axis == 0]
        lo = [f for f in xs if f.k0 == 0]
        hi = [f for f in xs if f.k0 + f.n == nx]
        return bool(lo) and bool(hi) and lo[0].n < hi[0].k0

    def slab_bytes(self, side: str, in_curl: bool = False) -> int:
        """Algorithmic bytes of the CPML of ``side`` ("E" or "H"): every psi
        word of every set read and written and the air masks read once (the
        field, coefficient and neighbour words are the curl's own when the
        recursion rides in the curl launch; the layer coefficients are a few
        words).  ``in_curl``: only the words the curl launch itself updates —
        all of them when the two x slabs sit at the domain ends (the launch
        splits its planes and takes the x-slab columns), else the y and z
        faces' elements outside every x slab (the x columns take a pass of
        their own)."""
        lst = self.psi_e if side == "E" else self.psi_h
        if not lst:
            return 0
        item = lst[0][0].element_size()
        if not in_curl or self.x_split():
            return int(2 * item * sum(t.numel() for terms in lst for t in terms)
                       + self._air_bytes[0 if side == "E" else 1])
        x_layers = sum(f.n for f in self.faces if f.axis == 0)
        words = 0.0
        for f, terms in zip(self.faces, lst):
            if f.axis == 0:
                continue
            shapes = f.psi_shapes(self.dims)[0 if side == "E" else 1]
            for sh, t in zip(shapes, terms):
                words += t.numel() * (sh[0] - x_layers) / sh[0]
        return int(2 * item * words)

    def zero_(self) -> None:
        for lst in self.psi_e + self.psi_h:
            for t in lst:
                t.zero_()

    def state_dict(self) -> dict:
        """psi slabs per face in the reference's (n0, n1, n2) C order, per set."""
        out = {}
        for f, pe, ph in zip(self.faces, self.psi_e, self.psi_h):
            e_sh, h_sh = f.psi_shapes(self.dims)
            for name, lst, shs in (("e", pe, e_sh), ("h", ph, h_sh)):
                for t, (ten, sh) in enumerate(zip(lst, shs)):
                    a = ten.cpu().numpy().reshape(self.nsets, sh[2], sh[1], sh[0])
                    out[f"{f.face}_psi_{name}{t}"] = np.ascontiguousarray(np.moveaxis(a, (1, 2, 3), (3, 2, 1)))
        return out

    def reference_state(self, s: int) -> dict:
        """{face: {"_psi_Ey": ..., ...}} of field set s: the reference's
        CPMLBoundary.state_dict (cpml.py:174-187, :197-204) — each term's psi
        slab in C order with the normal extent cut to the layers."""
        out = {}
        for f, pe, ph in zip(self.faces, self.psi_e, self.psi_h):
            e_sh, h_sh = f.psi_shapes(self.dims)
            d = {}
            for side, lst, shs, comps in (("E", pe, e_sh, E_TERM_COMPS[f.axis]),
                                          ("H", ph, h_sh, H_TERM_COMPS[f.axis])):
                for ten, sh, c in zip(lst, shs, comps):
                    a = ten[s].cpu().numpy().reshape(sh[2], sh[1], sh[0])
                    d[f"_psi_{side}{'xyz'[c]}"] = np.ascontiguousarray(np.moveaxis(a, (0, 1, 2), (2, 1, 0)))
            out[f.face] = d
        return out

    def load_reference_state(self, s: int, faces: dict) -> None:
        """Field set s from the reference's per-face psi dicts."""
        import torch  # noqa: PLC0415

        for f, pe, ph in zip(self.faces, self.psi_e, self.psi_h):
            d = faces[f.face]
            for side, lst, comps in (("E", pe, E_TERM_COMPS[f.axis]), ("H", ph, H_TERM_COMPS[f.axis])):
                for ten, c in zip(lst, comps):
                    a = np.asarray(d[f"_psi_{side}{'xyz'[c]}"])
                    dev = np.ascontiguousarray(np.moveaxis(a, (0, 1, 2), (2, 1, 0)).ravel())
                    ten[s].copy_(torch.from_numpy(dev).to(ten.device, ten.dtype))

    def load_state_dict(self, sd: dict) -> None:
        import torch  # noqa: PLC0415

        for f, pe, ph in zip(self.faces, self.psi_e, self.psi_h):
            for name, lst in (("e", pe), ("h", ph)):
                for t, ten in enumerate(lst):
                    a = np.asarray(sd[f"{f.face}_psi_{name}{t}"])
                    dev = np.moveaxis(a, (1, 2, 3), (3, 2, 1)).reshape(self.nsets, -1)
                    ten.copy_(torch.from_numpy(np.ascontiguousarray(dev)).to(ten.device, ten.dtype))
