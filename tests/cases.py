"""Shared workloads for the parity tests (built independently on both sides).

``c1_*``: the plain one-excitation WR-90 case — empty air-filled PEC guide
22.86 x 10.16 x 40 mm on a uniform 48 x 24 x 96 grid, DTBC TE ports with
3 modes on z_min / z_max, modulated Gaussian on port 1 mode 0 from the
TE10 cut-off up to 25 GHz, dt = courant_dt(grid, "normal").
"""

from __future__ import annotations

import math

import numpy as np

from oracle import fdtd as ofd
from oracle import mesh as om
from oracle import ports as op

C1_DIMS = (48, 24, 96)
C1_SIZE = (22.86e-3, 10.16e-3, 40e-3)
C1_FMAX = 25e9
C1_MODES = [(1, 0), (2, 0), (0, 1)]  # TE10, TE20, TE01 (ascending cut-off)


def c1_axes(dims=C1_DIMS, size=C1_SIZE):
    return tuple(np.linspace(0.0, L, n + 1) for L, n in zip(size, dims))


def gaussian(t: float, f_max: float) -> float:
    """Plain Gaussian pulse, peak 1 at 4 / f_max (the TEM waveform)."""
    sigma = 2.0 / (math.pi * f_max)
    x = (t - 4.0 / f_max) / sigma
    return math.exp(-x * x)


def modulated_gaussian(t: float, f_max: float, f_min: float) -> float:
    bw = f_max - f_min
    sigma = 2.0 / (math.pi * bw)
    t0 = 4.0 / bw
    x = (t - t0) / sigma
    fc = 0.5 * (f_min + f_max)
    return math.exp(-x * x) * math.cos(2.0 * math.pi * fc * (t - t0))


def oracle_complement(pl, x, y, z, mid, me_flat, pec_mesh, dt):
    """The complement-absorber arrays of an oracle port, as the reference's
    factory configures them on every port (factory.py:687-735): the plane's
    face flattened in a vacuum M_eps, the mesh's own PEC mask."""
    dims = (x.size - 1, y.size - 1, z.size - 1)
    nmat = int(np.max(mid)) + 1 if np.size(mid) else 1
    mvac = om.flatten_port(om.mass_e(x, y, z, mid, [[1.0] * nmat] * 3, om.EPS0), *dims, pl["face"], "E")
    return op.complement_arrays(pl, me_flat, mvac, pec_mesh, dt)


def c1_oracle(dtype=np.float64, dims=C1_DIMS, size=C1_SIZE, excite=("port1", 0), order="stencil",
              termination="auto", band=None, kernel_init=None):
    """Oracle side: coefficients, two ports, excitation; returns (Oracle, ports, dt)."""
    x, y, z = c1_axes(dims, size)
    nx, ny, nz = dims
    mid = np.zeros(dims, np.int32)
    me = om.mass_e(x, y, z, mid, [[1.0]] * 3, om.EPS0)
    mm = om.mass_h(x, y, z, mid, [[1.0]] * 3, om.MU0)
    dt = om.courant_dt(x, y, z, 0.95)
    pec = om.bbox_walls(np.zeros(me.size, bool), nx, ny, nz, ["xmin", "xmax", "ymin", "ymax"])
    pec_mesh = pec
    for f in ("zmin", "zmax"):
        me = om.flatten_port(me, nx, ny, nz, f, "E")
        mm = om.flatten_port(mm, nx, ny, nz, f, "H")
        pec = om.flatten_port(pec, nx, ny, nz, f, "E")
    ae, be = om.alpha_beta(me, np.zeros_like(me), dt, pec)
    ah, bh = om.alpha_beta(mm, np.zeros_like(mm), dt)
    omega_calc = 2 * math.pi * C1_FMAX
    ports = []
    for name, face in (("port1", "zmin"), ("port2", "zmax")):
        pl = op.plane(face, x, y, z)
        modes = [op.rect_te_mode(pl, x, y, z, m, n, me) for m, n in C1_MODES]
        ports.append(op.Port(name, pl, modes, me, mm, dt, omega_calc, termination=termination,
                             complement=oracle_complement(pl, x, y, z, mid, me, pec_mesh, dt),
                             kernel_init=kernel_init))
    if excite is not None:
        p = ports[0] if excite[0] == "port1" else ports[1]
        f_lo, f_hi = band if band is not None else (p.modes[excite[1]]["wc"] / (2 * math.pi), C1_FMAX)
        p.set_excitation(excite[1], lambda t, lo=f_lo, hi=f_hi: modulated_gaussian(t, hi, lo))
    orc = ofd.Oracle(dims, ae, be, ah, bh, dt, ports=ports, dtype=dtype, order=order)
    return orc, ports, dt, (me, mm)


def c1_mesh_ports(dims=C1_DIMS, size=C1_SIZE, termination="auto"):
    """Product side: the C1 mesh, its two 3-mode modal ports and dt."""
    from magnelio_amd import mesh as mm
    from magnelio_amd import ports as pp

    g = mm.GridLines(*c1_axes(dims, size))
    mesh = mm.Mesh.from_grid(g, pec_faces=("xmin", "xmax", "ymin", "ymax"))
    dt = mm.courant_dt(g, "normal")
    me, mu = mm.build_m_eps(mesh), mm.build_m_mu(mesh)
    ports = [pp.build_port(name, face, mesh, me, mu, dt=dt, f_calc=C1_FMAX, n_modes=3,
                           termination=termination)
             for name, face in (("port1", "zmin"), ("port2", "zmax"))]
    return mesh, ports, dt


def c1_solver(dims=C1_DIMS, size=C1_SIZE, excitations=(("port1", 0),), n_steps=4000,
              precision="single", use_graphs=False, termination="auto", curl64=False, monitors=(),
              kernel_init=None):
    """Product side: mesh, ports, batched HIP solver (one set per excitation)."""
    from magnelio_amd import mesh as mm
    from magnelio_amd import ports as pp
    from magnelio_amd.boundaries import PECBoundary
    from magnelio_amd.signals import resolve_excitation
    from magnelio_amd.solver import ChannelExcitation, FITTimeDomainSolver

    g = mm.GridLines(*c1_axes(dims, size))
    mesh = mm.Mesh.from_grid(g, pec_faces=("xmin", "xmax", "ymin", "ymax"))
    dt = mm.courant_dt(g, "normal")
    me, mu = mm.build_m_eps(mesh), mm.build_m_mu(mesh)
    ports = [pp.build_port("port1", "zmin", mesh, me, mu, dt=dt, f_calc=C1_FMAX, n_modes=3,
                           termination=termination),
             pp.build_port("port2", "zmax", mesh, me, mu, dt=dt, f_calc=C1_FMAX, n_modes=3,
                           termination=termination)]
    by = {p.name: p for p in ports}
    exc = []
    for name, m in excitations:
        wf = resolve_excitation(by[name].modes[m].omega_c, 0.0, C1_FMAX, m)
        exc.append(ChannelExcitation(name, m, wf))
    bcs = {f: PECBoundary(f) for f in ("xmin", "xmax", "ymin", "ymax")}
    sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, ports=ports,
                              total_time_steps=n_steps, dt=dt, precision=precision,
                              nsets=len(exc), excitations=exc, use_graphs=use_graphs,
                              curl64=curl64, monitors=list(monitors))
    if kernel_init is not None:
        sol.dtbc_kernel_init = kernel_init
    return sol, ports, dt


# ---------------------------------------------------------------------------
# Reduced C1 cases of the reference's own runs (tests/golden/c1_*.npz,
# written by tests/golden/make_c1_fixtures.py)

GOLDEN_BLOCK = (0.25 * 22.86e-3, 0.0, 8e-3, 0.75 * 22.86e-3, 0.6 * 10.16e-3, 12e-3)
GOLDEN_POST = (0.45 * 22.86e-3, 0.0, 9e-3, 0.55 * 22.86e-3, 10.16e-3, 11e-3)
GOLDEN_EPS, GOLDEN_SIGMA = 4.0, 0.05


def golden_c1(name):
    import os  # noqa: PLC0415

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", f"{name}.npz")
    return dict(np.load(path))


def _c1_fixture_module():
    import os  # noqa: PLC0415
    import sys  # noqa: PLC0415

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    if here not in sys.path:
        sys.path.insert(0, here)
    import make_c1_fixtures as mk  # noqa: PLC0415

    return mk


def golden_n_modes(name) -> int:
    """Modes per port of a golden case (3 unless the fixture says)."""
    return int(golden_c1(name).get("n_modes", 3))


def golden_excite(name) -> int:
    """The excited channel of port 1 in a golden case (0 unless the fixture says)."""
    return int(golden_c1(name).get("excite", 0))


def golden_c1_oracle(name, dtype=np.float64):
    """Oracle of a golden case: masses, PEC and alpha/beta from oracle.mesh,
    ports from the reference's numerical mode profiles (the oracle's Port
    calibrates them itself), each with its complement absorber; c1_mur's
    port 1 terminates on Mur; c1_tm's corner block is PEC and its ports carry a
    TM channel.  Port 1's channel ``golden_excite(name)`` is excited.  Returns
    (Oracle, ports, dt, (me, mu))."""
    f = golden_c1(name)
    x, y, z = f["x"], f["y"], f["z"]
    dims = (x.size - 1, y.size - 1, z.size - 1)
    boxes = []
    eps_block = GOLDEN_EPS
    if name == "c1_loaded":
        boxes = [(1, GOLDEN_BLOCK), (2, GOLDEN_POST)]
    elif name == "c1_mur":
        mk = _c1_fixture_module()
        boxes, eps_block = [(1, mk.MUR_BLOCK)], mk.MUR_EPS
    elif name == "c1_tm":
        boxes = [(2, _c1_fixture_module().TM_CORNER)]
    elif name == "c1_coax":
        boxes = [(2, _c1_fixture_module().COAX_INNER)]
    mid = om.fill_aabb(x, y, z, boxes)
    eps = [[1.0, eps_block, 1.0]] * 3
    sig = [[0.0, GOLDEN_SIGMA if name == "c1_loaded" else 0.0, 0.0]] * 3
    me = om.mass_e(x, y, z, mid, eps, om.EPS0)
    mvac = om.mass_e(x, y, z, mid, [[1.0, 1.0, 1.0]] * 3, om.EPS0)
    mm = om.mass_h(x, y, z, mid, [[1.0, 1.0, 1.0]] * 3, om.MU0)
    ms = om.mass_e(x, y, z, mid, sig, 1.0)
    pec_mesh = om.bbox_walls(om.pec_edge_mask(mid == 2), *dims, ["xmin", "xmax", "ymin", "ymax"])
    pec = pec_mesh
    for face in ("zmin", "zmax"):
        me = om.flatten_port(me, *dims, face, "E")
        mvac = om.flatten_port(mvac, *dims, face, "E")
        mm = om.flatten_port(mm, *dims, face, "H")
        pec = om.flatten_port(pec, *dims, face, "E")
    dt = float(f["dt"])
    ae, be = om.alpha_beta(me, ms, dt, pec)
    ah, bh = om.alpha_beta(mm, np.zeros_like(mm), dt)
    ports = []
    for label, face in (("port1", "zmin"), ("port2", "zmax")):
        pl = op.plane(face, x, y, z)
        modes = [(f[f"{label}_e_u{m}"], f[f"{label}_e_v{m}"], float(f[f"{label}_omega_c{m}"]),
                  str(f.get(f"{label}_type{m}", "TE")),
                  float(f[f"{label}_z_line{m}"]) if f"{label}_z_line{m}" in f else None)
                 for m in range(golden_n_modes(name))]
        term = "mur" if (name == "c1_mur" and label == "port1") else "auto"
        comp = op.complement_arrays(pl, me, mvac, pec_mesh, dt)
        ports.append(op.Port(label, pl, modes, me, mm, dt, 2 * math.pi * C1_FMAX, termination=term,
                             complement=comp))
    exc = golden_excite(name)
    f_lo = ports[0].modes[exc]["wc"] / (2 * math.pi)
    if f_lo > 0.0:
        ports[0].set_excitation(exc, lambda t: modulated_gaussian(t, C1_FMAX, f_lo))
    else:  # a TEM channel: the plain Gaussian
        ports[0].set_excitation(exc, lambda t: gaussian(t, C1_FMAX))
    orc = ofd.Oracle(dims, ae, be, ah, bh, dt, ports=ports, dtype=dtype)
    return orc, ports, dt, (me, mm)


def golden_c1_mesh(name):
    """Product mesh of a golden case."""
    from magnelio_amd import mesh as pm  # noqa: PLC0415

    f = golden_c1(name)
    g = pm.GridLines(f["x"], f["y"], f["z"])
    regions = []
    if name == "c1_loaded":
        regions = [(pm.Material(name="block", epsilon=(GOLDEN_EPS,) * 3, sigma=(GOLDEN_SIGMA,) * 3),
                    GOLDEN_BLOCK), (pm.Material.pec(), GOLDEN_POST)]
    elif name == "c1_mur":
        mk = _c1_fixture_module()
        regions = [(pm.Material(name="feed_block", epsilon=(mk.MUR_EPS,) * 3), mk.MUR_BLOCK)]
    elif name == "c1_tm":
        regions = [(pm.Material.pec(), _c1_fixture_module().TM_CORNER)]
    elif name == "c1_coax":
        regions = [(pm.Material.pec(), _c1_fixture_module().COAX_INNER)]
    return pm.Mesh.from_grid(g, regions=regions, pec_faces=("xmin", "xmax", "ymin", "ymax"))


def golden_port(name, label, face, mesh, me, mu, dt, n_modes):
    """Product port of a golden case: coaxial (TEM) for c1_coax, else numerical."""
    from magnelio_amd import ports as pp  # noqa: PLC0415

    if name == "c1_coax":
        w = _c1_fixture_module().COAX_W
        return pp.build_coax_port(label, face, mesh, me, mu, dt=dt, f_calc=C1_FMAX, center=(0.5 * w, 0.5 * w),
                                  n_modes=n_modes)
    return pp.build_port(label, face, mesh, me, mu, dt=dt, f_calc=C1_FMAX, n_modes=n_modes)


def golden_c1_solver(name, precision="double", n_steps=300):
    """Product HIP solver of a golden case (port 1's golden channel excited)."""
    from magnelio_amd import mesh as pm  # noqa: PLC0415
    from magnelio_amd import ports as pp  # noqa: PLC0415
    from magnelio_amd.boundaries import PECBoundary  # noqa: PLC0415
    from magnelio_amd.signals import resolve_excitation  # noqa: PLC0415
    from magnelio_amd.solver import ChannelExcitation, FITTimeDomainSolver  # noqa: PLC0415

    mesh = golden_c1_mesh(name)
    dt = pm.courant_dt(mesh.grid, "normal")
    me, mu = pm.build_m_eps(mesh), pm.build_m_mu(mesh)
    n_modes, exc = golden_n_modes(name), golden_excite(name)
    ports = [golden_port(name, lbl, face, mesh, me, mu, dt, n_modes)
             for lbl, face in (("port1", "zmin"), ("port2", "zmax"))]
    wf = resolve_excitation(ports[0].modes[exc].omega_c, 0.0, C1_FMAX, exc)
    bcs = {f: PECBoundary(f) for f in ("xmin", "xmax", "ymin", "ymax")}
    sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, ports=ports,
                              total_time_steps=n_steps, dt=dt, precision=precision, nsets=1,
                              excitations=[ChannelExcitation("port1", exc, wf)])
    return sol, ports, dt


# ---------------------------------------------------------------------------
# CPML box with TF/SF and DFT monitors (tests/golden/box_tfsf_dft.npz,
# written by tests/golden/make_box_fixtures.py)

def golden_box():
    import os  # noqa: PLC0415
    import sys  # noqa: PLC0415

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    if here not in sys.path:
        sys.path.insert(0, here)
    import make_box_fixtures as mk  # noqa: PLC0415

    return dict(np.load(os.path.join(here, "box_tfsf_dft.npz"))), mk


def golden_box_oracle(dtype=np.float64):
    """Oracle of the golden box: (Oracle, region monitor, far-field monitor)."""
    from oracle import cpml as oc  # noqa: PLC0415
    from oracle import monitors as omon  # noqa: PLC0415
    from oracle import tfsf as ot  # noqa: PLC0415

    f, mk = golden_box()
    x, y, z = f["x"], f["y"], f["z"]
    dims = (x.size - 1, y.size - 1, z.size - 1)
    mid = om.fill_aabb(x, y, z, [(i + 1, box) for i, (_, box) in enumerate(mk.BRICKS)])
    eps = [[1.0] + [e for e, _ in mk.BRICKS]] * 3
    me = om.mass_e(x, y, z, mid, eps, om.EPS0)
    mmu = om.mass_h(x, y, z, mid, [[1.0] * (len(mk.BRICKS) + 1)] * 3, om.MU0)
    dt = float(f["dt"])
    ae, be = om.alpha_beta(me, np.zeros_like(me), dt, np.zeros(me.size, bool))
    ah, bh = om.alpha_beta(mmu, np.zeros_like(mmu), dt)
    faces = [oc.CPMLFace(fc, x, y, z, dt, dtype, n_layers=mk.LAYERS) for fc in mk.FACES]
    src = ot.PlaneWave(x, y, z, be, bh, dt, direction=(0, 0, 1), polarization=(1, 0, 0),
                       box=ot.snap_box(x, y, z, mk.TFSF), f_max=20e9, dtype=dtype)
    rs = f["region_slices"]
    region = omon.FieldDFT(mk.FREQS, x, y, z, tuple(slice(int(a), int(b)) for a, b in rs))
    far = omon.FarFieldDFT(mk.FREQS, x, y, z, mk.FACES, {fc: mk.LAYERS + 2 for fc in mk.FACES})
    orc = ofd.Oracle(dims, ae, be, ah, bh, dt, cpml=faces, sources=[src], monitors=[region, far],
                     dtype=dtype)
    return orc, region, far


# ---------------------------------------------------------------------------
# SIBC cavity (tests/golden/sibc_cavity.npz): a graded PEC cavity with a
# lossy-metal brick and a plain PEC brick, walls and fits of the reference


def golden_sibc():
    """(fixture, its surfaces as oracle.sibc.Surface, {tag: oracle.sibc.Fit})."""
    from oracle import sibc as osb  # noqa: PLC0415

    import os  # noqa: PLC0415

    f = dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sibc_cavity.npz")))
    surfs, fits = [], {}
    for i in range(int(f["n_surfaces"])):
        t = str(f[f"surf{i}_tag"])
        tag = int(t) if t.isdigit() else t
        surfs.append(osb.Surface(tag, f[f"surf{i}_comp"], f[f"surf{i}_flat_idx"], f[f"surf{i}_weight"],
                                 f[f"surf{i}_inv_l_dual"]))
        fits[tag] = osb.Fit(float(f[f"surf{i}_c0"]), [tuple(r) for r in f[f"surf{i}_branches"]])
    return f, surfs, fits


def golden_sibc_oracle(dtype=np.float64, surfaces=None, fits=None):
    """Oracle of the golden SIBC cavity seeded as the reference's run:
    (Oracle, SIBC, (m_eps, m_mu))."""
    from oracle import sibc as osb  # noqa: PLC0415

    f, surfs, fts = golden_sibc()
    surfaces = surfs if surfaces is None else surfaces
    fits = fts if fits is None else fits
    x, y, z = f["x"], f["y"], f["z"]
    dims = (x.size - 1, y.size - 1, z.size - 1)
    mid = f["material_id"]
    nmat = int(mid.max()) + 1
    me = om.mass_e(x, y, z, mid, [[1.0] * nmat] * 3, om.EPS0)
    mmu = om.mass_h(x, y, z, mid, [[1.0] * nmat] * 3, om.MU0)
    dt = float(f["dt"])
    pec = f["pec_edges"]
    sib = osb.SIBC.build(surfaces, fits, dims, dt, frozen=mmu <= 0)
    ae, be = om.alpha_beta(me, np.zeros_like(me), dt, pec)
    ah, bh = om.alpha_beta(mmu, sib.W if sib is not None else np.zeros_like(mmu), dt)
    if sib is not None:
        sib.bind(bh, dtype)
    orc = ofd.Oracle(dims, ae, be, ah, bh, dt, dtype=dtype, sibc=sib)
    e = np.random.default_rng(11).standard_normal(me.size)
    e[pec] = 0.0
    orc.e[:] = e.astype(dtype)
    return orc, sib, (me, mmu)


# ---------------------------------------------------------------------------
# C3 (batched multi-excitation) at reduced size: WR-90 guide, two eps_r = 30
# posts, a PEC iris with a seeded window, 3-mode ports on z_min / z_max

C3_SMALL = (48, 24, 72)
C3_SIZE = (22.86e-3, 10.16e-3, 60e-3)


def c3_oracle(excite, dims=C3_SMALL, dtype=np.float64):
    """Oracle of the reduced C3 case with one excited (port, mode): cells
    filled from the seeded layout (bench.c3_geometry) with the cell-centre
    tests (iris box, window box, then the posts), closed-form TE modes on
    the uniform feed.  Returns (Oracle, ports, dt, (me, mm))."""
    from bench import c3_geometry  # noqa: PLC0415

    W, H, L = C3_SIZE
    x, y, z = (np.linspace(0.0, s, n + 1) for s, n in zip(C3_SIZE, dims))
    posts, (zi, dz), (wx, wy, ww, wh) = c3_geometry(dims, 3, C3_SIZE)
    AIR, PEC, DIEL = 0, 1, 2
    mid = om.fill_aabb(x, y, z, [(PEC, (0.0, 0.0, zi, W, H, zi + dz)),
                                 (AIR, (wx, wy, zi, wx + ww, wy + wh, zi + dz))])
    xc_, zc_ = 0.5 * (x[:-1] + x[1:]), 0.5 * (z[:-1] + z[1:])
    for px, pz, r in posts:
        inside = ((xc_[:, None] - px) ** 2 + (zc_[None, :] - pz) ** 2) < r * r
        mid[np.broadcast_to(inside[:, None, :], mid.shape)] = DIEL
    me = om.mass_e(x, y, z, mid, [[1.0, 1.0, 30.0]] * 3, om.EPS0)
    mmu = om.mass_h(x, y, z, mid, [[1.0, 1.0, 1.0]] * 3, om.MU0)
    nx, ny, nz = dims
    pec = om.bbox_walls(om.pec_edge_mask(mid == PEC), nx, ny, nz, ["xmin", "xmax", "ymin", "ymax"])
    pec_mesh = pec
    for f in ("zmin", "zmax"):
        me = om.flatten_port(me, nx, ny, nz, f, "E")
        mmu = om.flatten_port(mmu, nx, ny, nz, f, "H")
        pec = om.flatten_port(pec, nx, ny, nz, f, "E")
    dt = om.courant_dt(x, y, z, 0.95)
    ae, be = om.alpha_beta(me, np.zeros_like(me), dt, pec)
    ah, bh = om.alpha_beta(mmu, np.zeros_like(mmu), dt)
    ports = []
    for name, face in (("port1", "zmin"), ("port2", "zmax")):
        pl = op.plane(face, x, y, z)
        modes = [op.rect_te_mode(pl, x, y, z, m, n, me) for m, n in C1_MODES]
        ports.append(op.Port(name, pl, modes, me, mmu, dt, 2 * math.pi * 25e9,
                             complement=oracle_complement(pl, x, y, z, mid, me, pec_mesh, dt)))
    p = ports[0] if excite[0] == "port1" else ports[1]
    f_lo = p.modes[excite[1]]["wc"] / (2 * math.pi)
    p.set_excitation(excite[1], lambda t, lo=f_lo: modulated_gaussian(t, 25e9, lo))
    orc = ofd.Oracle(dims, ae, be, ah, bh, dt, ports=ports, dtype=dtype)
    return orc, ports, dt, (me, mmu)


# ---------------------------------------------------------------------------
# C5 serving jobs: the oracle of one job from its seeded geometry

def c5_job_oracle(job, excite, dtype=np.float64):
    """Oracle of a serving job (bench.c5_jobs) with one excited (port,
    mode): the job's grid — a WR-90 guide (1-2 ports on its z ends) or a
    cross of two (3-4 ports on z and x ends, PEC corner blocks) — its bricks
    (clipped to the region clear of the feeds) and spheres (cell-centre test)
    filled in the job's order, PEC walls, closed-form TE ports on each
    port's rectangular window.  Returns (Oracle, ports, dt)."""
    import bench  # noqa: PLC0415

    dims, n_ports, objs = job.build.tasks["mesh"][1].args
    n_modes = job.build.tasks["port1"][0][3]
    size = bench._c5_size(n_ports)
    x, y, z = (np.linspace(0.0, s, n + 1) for s, n in zip(size, dims))
    klo, khi = bench._c5_keepout(n_ports)
    boxes, spheres, eps = [], [], [1.0]
    if n_ports > 2:
        eps.append(1.0)
        boxes += [(1, c) for c in bench._c5_corners()]
    pec_id = 1 if n_ports > 2 else -1
    for e, sphere, ctr, half in objs:
        eps.append(e)
        if sphere:
            spheres.append((len(eps) - 1, ctr, float(half.min())))
        else:
            lo, hi = np.maximum(ctr - half, klo), np.minimum(ctr + half, khi)
            boxes.append((len(eps) - 1, tuple(lo) + tuple(hi)))
    mid = om.fill_aabb(x, y, z, boxes)
    cx, cy, cz = np.meshgrid(*(0.5 * (a[:-1] + a[1:]) for a in (x, y, z)), indexing="ij")
    for m, c, r in spheres:
        inside = (((cx - c[0]) ** 2 + (cy - c[1]) ** 2 + (cz - c[2]) ** 2 < r * r) & (cx > klo[0])
                  & (cx < khi[0]) & (cz > klo[2]) & (cz < khi[2]))
        mid[inside] = m
    me = om.mass_e(x, y, z, mid, [eps] * 3, om.EPS0)
    mmu = om.mass_h(x, y, z, mid, [[1.0] * len(eps)] * 3, om.MU0)
    nx, ny, nz = dims
    pec = om.bbox_walls(om.pec_edge_mask(mid == pec_id), nx, ny, nz, list(bench._c5_walls(n_ports)))
    pec_mesh = pec
    faces = bench.C5_FACES[:n_ports]
    for f in faces:
        me = om.flatten_port(me, nx, ny, nz, f, "E")
        mmu = om.flatten_port(mmu, nx, ny, nz, f, "H")
        pec = om.flatten_port(pec, nx, ny, nz, f, "E")
    dt = om.courant_dt(x, y, z, 0.95)
    ae, be = om.alpha_beta(me, np.zeros_like(me), dt, pec)
    ah, bh = om.alpha_beta(mmu, np.zeros_like(mmu), dt)
    ports = []
    for k, face in enumerate(faces):
        pl = op.plane(face, x, y, z)
        if face[0] == "z":  # global tangential axes (x, y): the guide is x-wide
            window = ((nx // 4, 3 * nx // 4), (0, ny)) if n_ports > 2 else None
            mns = C1_MODES
        else:  # (y, z): the arm is z-wide
            window = ((0, ny), (nz // 4, 3 * nz // 4))
            mns = [(n, m) for m, n in C1_MODES]
        modes = [op.rect_te_mode(pl, x, y, z, m, n, me, window=window) for m, n in mns[:n_modes]]
        ports.append(op.Port(f"port{k + 1}", pl, modes, me, mmu, dt, 2 * math.pi * job.f_max,
                             complement=oracle_complement(pl, x, y, z, mid, me, pec_mesh, dt)))
    p = next(q for q in ports if q.name == excite[0])
    f_lo = p.modes[excite[1]]["wc"] / (2 * math.pi)
    p.set_excitation(excite[1], lambda t, lo=f_lo, hi=job.f_max: modulated_gaussian(t, hi, lo))
    return ofd.Oracle(dims, ae, be, ah, bh, dt, ports=ports, dtype=dtype), ports, dt


# ---------------------------------------------------------------------------
# C4 (precision study) at reduced size: Lorentz slab with sigma, a DTBC port
# on z_min, CPML on z_max

def lowest_te_modes(x, y, n_modes, eps_r=1.0):
    """(m, n) of the n_modes lowest discrete TE cut-offs of a uniform PEC
    rectangle (closed form k_c^2 = sum (2/d sin(m pi / 2N))^2)."""
    nx, ny = x.size - 1, y.size - 1
    dx, dy = float(x[1] - x[0]), float(y[1] - y[0])
    cands = [(m, n) for m in range(0, 4) for n in range(0, 4) if (m, n) != (0, 0)]
    kc = {mn: (2 / dx * math.sin(mn[0] * math.pi / (2 * nx))) ** 2
          + (2 / dy * math.sin(mn[1] * math.pi / (2 * ny))) ** 2 for mn in cands}
    return sorted(cands, key=lambda mn: kc[mn])[:n_modes]


def c4_oracle(n, excite, dtype=np.float64, n_modes=2, seed=4, f_max=20e9):
    """Oracle of the reduced C4 case (bench.c4_case on an n^3 grid) with one
    excited mode of port 1.  Returns (Oracle, ports, dt, (me, mm))."""
    from bench import C4_LAYERS, C4_SIZE, c4_layout  # noqa: PLC0415
    from oracle import cpml as oc  # noqa: PLC0415
    from oracle import dispersion as od  # noqa: PLC0415

    W, H, L = C4_SIZE
    x, y, z = (np.linspace(0.0, s, n + 1) for s in C4_SIZE)
    dims = (n, n, n)
    eps_inf, sigma, poles = c4_layout(seed)
    mid = om.fill_aabb(x, y, z, [(1, (0.0, 0.0, 0.4 * L, W, H, 0.6 * L))])
    me = om.mass_e(x, y, z, mid, [[1.0, eps_inf]] * 3, om.EPS0)
    ms = om.mass_e(x, y, z, mid, [[0.0, sigma]] * 3, 1.0)
    mmu = om.mass_h(x, y, z, mid, [[1.0, 1.0]] * 3, om.MU0)
    dt = om.courant_dt(x, y, z, 0.95)
    walls = om.bbox_walls(np.zeros(me.size, bool), *dims, ["xmin", "xmax", "ymin", "ymax"])
    disp = od.Dispersion(x, y, z, mid, {1: list(poles)}, dt, pec_e=walls)
    me = om.flatten_port(me, *dims, "zmin", "E")
    mmu = om.flatten_port(mmu, *dims, "zmin", "H")
    pec = om.flatten_port(walls, *dims, "zmin", "E")
    ae, be = om.alpha_beta(me, ms, dt, pec, w=disp.w_half)
    ah, bh = om.alpha_beta(mmu, np.zeros_like(mmu), dt)
    disp.bind(be, dtype)
    pl = op.plane("zmin", x, y, z)
    modes = [op.rect_te_mode(pl, x, y, z, m, k, me) for m, k in lowest_te_modes(x, y, n_modes)]
    port = op.Port("port1", pl, modes, me, mmu, dt, 2 * math.pi * f_max,
                   complement=oracle_complement(pl, x, y, z, mid, me, walls, dt))
    f_lo = port.modes[excite[1]]["wc"] / (2 * math.pi)
    port.set_excitation(excite[1], lambda t, lo=f_lo: modulated_gaussian(t, f_max, lo))
    face = oc.CPMLFace("zmax", x, y, z, dt, dtype, pec_e_flat=pec, pec_cell=np.zeros(dims, bool),
                       n_layers=C4_LAYERS)
    orc = ofd.Oracle(dims, ae, be, ah, bh, dt, cpml=[face], ports=[port], dtype=dtype, dispersion=disp)
    return orc, [port], dt, (me, mmu)
