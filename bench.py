"""FIT-TD stepping benchmark on one MI355X.

Prints ONE JSON line: Yee Gcell-updates/s of the stepping engine on the
single-large-job workload — 256^3 seeded non-uniform grid with 40 seeded
dielectric bricks/spheres (eps_r ~ U[1, 10]), 8-layer CPML on all six faces,
a TF/SF plane wave (+z, x-polarised Gaussian) and a 1001-frequency far-field
(Huygens box) DFT monitor with compensated float64 bins, every part of the
step inside the timed region.  The timed steps run from a seeded mid-run
state (step 1024 of the job: random fields, full DFT sample buffers), where
each step carries its 1/256 share of the pipelined DFT contraction.  Beside
it: the HBM-roofline fraction of the dominant kernel measured live with HIP
events, the batched K = 6 sweep of the 192 x 96 x 288 dielectric-resonator
guide against its one-excitation-per-launch path, and the CPU baseline (the
oracle's C/OpenMP restatement of the reference's Numba curl kernels plus the
NumPy CPML restatement) timed on the host's cores on a bounded sample of the
same grid.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per k_update_e launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (bench.py --pmc-probe, at the tuned launch shape), FETCH scaled x2.0 and WRITE x1.0 from
# a 1 GiB device copy: {(n, K, CPML fused into the launch): bytes}
K_UPDATE_E_TRAFFIC = {(256, 1, True): 1.384e9}


def stream_ceiling(n: int, n_read: int, n_write: int, reps: int = 20) -> dict:
    """The device's streaming rate for a curl launch's access mix: n_read
    input and n_write output arrays of n float32 words, one element per lane
    and nothing else in the kernel (mgh_stream_probe), median of `reps`
    launches timed with HIP events on the current stream."""
    import torch

    from magnelio_amd import _lib

    src = torch.ones(n_read * n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty(n_write * n, dtype=torch.float32, device="cuda:0")
    st = torch.cuda.current_stream()
    sh = int(st.cuda_stream)

    def launch():
        _lib.call("mgh_stream_probe", n_read, n_write, src.data_ptr(), dst.data_ptr(), n, sh)

    for _ in range(3):
        launch()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for a, b in ev:
        a.record(st)
        launch()
        b.record(st)
    torch.cuda.synchronize()
    us = float(np.median([a.elapsed_time(b) * 1e3 for a, b in ev]))
    byt = 4.0 * n * (n_read + n_write)
    del src, dst
    torch.cuda.empty_cache()
    return {"reads": n_read, "writes": n_write, "elements": n, "bytes": byt, "us": us,
            "GBps": byt / us / 1e3, "frac_of_peak": byt / us / 1e3 / HBM_PEAK_GBS}


def curl_launch_times(sol, reps: int = 20) -> tuple[float, float]:
    """Median HIP-event times (us) of the solver's E and H curl launches on
    its current state (the fields change; call after the timed steps)."""
    import torch

    from magnelio_amd import kernels

    st = torch.cuda.current_stream()
    ev = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(reps)]
    for a, b, c in ev:
        a.record(st)
        kernels.update_e(sol.e, sol.h, sol.alpha_e, sol.beta_e, sol.dims, curl64=sol.curl64,
                         kchunk=sol.kchunk, stream=st)
        b.record(st)
        kernels.update_h(sol.e, sol.h, sol.alpha_h, sol.beta_h, sol.dims, kchunk=sol.kchunk,
                         stream=st)
        c.record(st)
    torch.cuda.synchronize()
    return (float(np.median([a.elapsed_time(b) * 1e3 for a, b, _ in ev])),
            float(np.median([b.elapsed_time(c) * 1e3 for _, b, c in ev])))


def c2_mesh(n: int = 256, seed: int = 2024, n_objects: int = 40, length: float = 64e-3):
    """Seeded non-uniform n^3 grid (adjacent spacing ratio <= 1.5) + objects."""
    from magnelio_amd.mesh import GridLines, Material, Mesh

    rng = np.random.default_rng(seed)

    def axis():
        logs = np.cumsum(rng.uniform(math.log(1 / 1.5), math.log(1.5), n - 1))
        w = np.exp(np.concatenate([[0.0], logs]))
        # keep every adjacent ratio within [1/1.5, 1.5]: the walk already is
        w = w / w.sum() * length
        return np.concatenate([[0.0], np.cumsum(w)])

    grid = GridLines(axis(), axis(), axis())
    lo, hi = 0.15 * length, 0.85 * length
    boxes, shapes = [], []
    for _ in range(n_objects):
        mat = Material.dielectric(float(rng.uniform(1.0, 10.0)))
        c = rng.uniform(lo, hi, 3)
        s = rng.uniform(length / 20, length / 6, 3)
        if rng.random() < 0.5:
            boxes.append((mat, (*(c - s / 2), *(c + s / 2))))
        else:
            r = float(s[0] / 2)
            shapes.append((mat, lambda X, Y, Z, c=c, r=r: (X - c[0]) ** 2 + (Y - c[1]) ** 2
                           + (Z - c[2]) ** 2 < r * r))
    return Mesh.from_grid(grid, regions=boxes, shapes=shapes, pec_faces=())


C3_SIZE = (22.86e-3, 10.16e-3, 60e-3)
C3_FMAX = 25e9


def c3_geometry(dims=(192, 96, 288), seed=3, size=C3_SIZE):
    """Seeded layout of the batched workload: two eps_r = 30 posts (axis
    along y) at z = 0.3 L and 0.7 L, [(xc, zc, r)], and the PEC iris layer
    (z in [zi, zi + dz)) with its rectangular window (wx, wy, ww, wh)."""
    rng = np.random.default_rng(seed)
    W, H, L = size
    posts = []
    for zc in (0.3 * L, 0.7 * L):
        xc = W / 2 + rng.uniform(-0.15, 0.15) * W
        r = rng.uniform(0.08, 0.12) * W
        posts.append((float(xc), float(zc), float(r)))
    ww, wh = rng.uniform(0.3, 0.6) * W, rng.uniform(0.4, 0.8) * H
    wx = rng.uniform(0.0, W - ww)
    wy = rng.uniform(0.0, H - wh)
    dz = L / dims[2]
    zi = round(0.5 * dims[2]) * dz
    return posts, (float(zi), float(dz)), (float(wx), float(wy), float(ww), float(wh))


def c3_case(dims=(192, 96, 288), seed=3, size=C3_SIZE, n_modes=3, f_max=C3_FMAX):
    """Batched multi-excitation workload: a WR-90 guide with two seeded eps_r = 30
    dielectric posts (axis along y) and a PEC iris wall at mid-length with a
    seeded rectangular window; modal ports with n_modes modes on zmin and zmax
    (homogeneous, uniformly gridded feeds).  Returns (mesh, ports, dt)."""
    from magnelio_amd import mesh as mm
    from magnelio_amd import ports as pp

    W, H, L = size
    axes = [np.linspace(0.0, s, n + 1) for s, n in zip(size, dims)]
    grid = mm.GridLines(*axes)
    posts, (zi, dz), (wx, wy, ww, wh) = c3_geometry(dims, seed, size)
    diel = mm.Material.dielectric(30.0, name="resonator")
    shapes = [(diel, lambda X, Y, Z, xc=xc, zc=zc, r=r: (X - xc) ** 2 + (Z - zc) ** 2 < r * r)
              for xc, zc, r in posts]
    regions = [(mm.Material.pec(), (0.0, 0.0, zi, W, H, zi + dz)),
               (mm.Material.air(), (wx, wy, zi, wx + ww, wy + wh, zi + dz))]
    # the shapes fill after the boxes; keep the posts clear of the iris layer
    mesh = mm.Mesh.from_grid(grid, regions=regions, shapes=shapes,
                             pec_faces=("xmin", "xmax", "ymin", "ymax"))
    dt = mm.courant_dt(grid, "normal")
    ports = [pp.build_port_local(name, face, mesh, dt=dt, f_calc=f_max, n_modes=n_modes)
             for name, face in (("port1", "zmin"), ("port2", "zmax"))]
    return mesh, ports, dt


C5_SIZE = (22.86e-3, 10.16e-3, 45.72e-3)  # WR-90 cross-section, 2a long
# 3-4 ports: a cross of two WR-90 guides, the main one along z over
# x in [a/2, 3a/2], the arm along x over z in [a/2, 3a/2]; PEC blocks fill the
# four corners (with 3 ports the x_max end of the arm is a PEC wall)
C5_CROSS = (2 * 22.86e-3, 10.16e-3, 2 * 22.86e-3)
C5_FACES = ("zmin", "zmax", "xmin", "xmax")


def _c5_size(n_ports):
    return C5_SIZE if n_ports <= 2 else C5_CROSS


def _c5_walls(n_ports):
    if n_ports <= 2:
        return ("xmin", "xmax", "ymin", "ymax") + (("zmax",) if n_ports == 1 else ())
    return ("ymin", "ymax") + (("xmax",) if n_ports == 3 else ())


def _c5_corners():
    """PEC corner blocks of the cross guide."""
    a, b = C5_SIZE[0], C5_SIZE[1]
    lo, hi, end = 0.5 * a, 1.5 * a, 2 * a
    return [(x0, 0.0, z0, x1, b, z1) for x0, x1 in ((0.0, lo), (hi, end)) for z0, z1 in ((0.0, lo), (hi, end))]


def _c5_dims(c, n_ports):
    """Grid of a job of about c cells: uniform, 8-cell multiples across the guide."""
    a, b = C5_SIZE[0], C5_SIZE[1]
    if n_ports <= 2:
        n = max(16, int(round((c / 0.889) ** (1 / 3) / 8.0)) * 8)
        return (n, max(8, int(round(n * b / a))), 2 * n)
    m = max(16, int(round((c / 1.778) ** (1 / 3) / 8.0)) * 8)
    return (2 * m, max(8, int(round(m * b / a))), 2 * m)


def _c5_keepout(n_ports):
    """(lo, hi) corners of the region the dielectric objects are clipped to:
    clear of every port's feed slab."""
    if n_ports <= 2:
        return np.array([0.0, 0.0, 0.2 * C5_SIZE[2]]), np.array([C5_SIZE[0], C5_SIZE[1], 0.8 * C5_SIZE[2]])
    a = C5_SIZE[0]
    return np.array([0.5 * a, 0.0, 0.5 * a]), np.array([1.5 * a, C5_SIZE[1], 1.5 * a])


def _c5_mesh(dims, n_ports, objs):
    """Mesh and dt of one serving job (module level: a process pool runs it)."""
    from magnelio_amd import mesh as mm

    axes = [np.linspace(0.0, L, k + 1) for L, k in zip(_c5_size(n_ports), dims)]
    grid = mm.GridLines(*axes)
    boxes = [(mm.Material.pec(), c) for c in _c5_corners()] if n_ports > 2 else []
    shapes = []
    klo, khi = _c5_keepout(n_ports)
    for eps, sphere, ctr, half in objs:
        mat = mm.Material.dielectric(eps)
        if sphere:
            r = float(half.min())
            shapes.append((mat, lambda X, Y, Z, c=ctr, r=r: ((X - c[0]) ** 2 + (Y - c[1]) ** 2
                           + (Z - c[2]) ** 2 < r * r) & (X > klo[0]) & (X < khi[0])
                           & (Z > klo[2]) & (Z < khi[2])))
        else:
            lo, hi = np.maximum(ctr - half, klo), np.minimum(ctr + half, khi)
            boxes.append((mat, (lo[0], lo[1], lo[2], hi[0], hi[1], hi[2])))
    mesh = mm.Mesh.from_grid(grid, regions=boxes, shapes=shapes, pec_faces=_c5_walls(n_ports))
    return mesh, mm.courant_dt(grid, "normal")


def _c5_port(dims, n_ports, name, face, n_modes, f_max):
    """Modal port of a serving job.  The feed slab holds no dielectric in any
    job, so the model depends on (grid, walls, face, modes) only and is built
    on the empty guide of that grid (the cross's PEC corners included),
    identical to the build on the job's own mesh."""
    from magnelio_amd import mesh as mm
    from magnelio_amd import ports as pp

    axes = [np.linspace(0.0, L, k + 1) for L, k in zip(_c5_size(n_ports), dims)]
    grid = mm.GridLines(*axes)
    regions = [(mm.Material.pec(), c) for c in _c5_corners()] if n_ports > 2 else []
    mesh = mm.Mesh.from_grid(grid, regions=regions, pec_faces=_c5_walls(n_ports))
    return pp.build_port_local(name, face, mesh, dt=mm.courant_dt(grid, "normal"), f_calc=f_max,
                               n_modes=n_modes)


def _c5_worker_init():
    """A build worker: one BLAS/OpenMP thread and a lower CPU priority, so
    that the builds neither oversubscribe the host nor delay the loop that
    enqueues the device work."""
    from threadpoolctl import threadpool_limits

    threadpool_limits(1)
    if "torch" in sys.modules:
        sys.modules["torch"].set_num_threads(1)
    os.nice(10)


def _c5_assemble(parts):
    mesh, dt = parts["mesh"]
    ports = [parts[k] for k in sorted(k for k in parts if k.startswith("port"))]
    return mesh, ports, dt


def c5_jobs(n_jobs=200, seed=5, rate=1.5, cells=(32 ** 3, 384 ** 3), steps=(2000, 20000),
            step_scale=1.0, f_max=25e9, n_freq=101, fp64_check="off"):
    """Serving workload: Poisson arrivals (rate jobs/s), cell counts log-uniform
    in `cells`, 1-4 seeded dielectric bricks/spheres (eps_r U[1, 10]) clear of
    1-4 modal ports x 1-3 modes — a WR-90 guide with ports on its z ends for
    1-2 ports, a cross of two WR-90 guides with ports on z and x ends for 3-4
    — steps U{steps} x step_scale."""
    import functools

    from magnelio_amd.serve import BuildPlan, SweepJob

    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    t -= t[0]
    jobs = []
    for i in range(n_jobs):
        c = math.exp(rng.uniform(math.log(cells[0]), math.log(cells[1])))
        n_ports = int(rng.integers(1, 5))
        dims = _c5_dims(c, n_ports)
        n_modes = int(rng.integers(1, 4))
        n_steps = max(64, int(round(int(rng.integers(steps[0], steps[1] + 1)) * step_scale)))
        size = np.array(_c5_size(n_ports))
        objs = []
        for _ in range(int(rng.integers(1, 5))):
            if n_ports <= 2:
                ctr = rng.uniform([0.2, 0.2, 0.25], [0.8, 0.8, 0.75]) * size
                half = rng.uniform(0.05, 0.15, 3) * size
            else:
                ctr = rng.uniform([0.3, 0.2, 0.3], [0.7, 0.8, 0.7]) * size
                half = rng.uniform(0.05, 0.15, 3) * np.array(C5_SIZE)
            objs.append((float(rng.uniform(1.0, 10.0)), bool(rng.random() < 0.5), ctr, half))
        tasks = {"mesh": (None, functools.partial(_c5_mesh, dims, n_ports, objs))}
        for k, face in enumerate(C5_FACES[:n_ports]):
            name = f"port{k + 1}"
            tasks[name] = ((tuple(dims), n_ports, face, n_modes),
                           functools.partial(_c5_port, dims, n_ports, name, face, n_modes, f_max))
        build = BuildPlan(tasks, _c5_assemble)
        f_c = 3e8 / (2 * C5_SIZE[0])
        jobs.append(SweepJob(i, float(t[i]), build, n_steps, np.linspace(1.1 * f_c, 0.96 * f_max, n_freq),
                             f_max, fp64_check=fp64_check))
    return jobs


C4_SIZE = (32e-3, 20e-3, 48e-3)


def lorentz_pole(d_eps, f0, gamma):
    """Pole-residue pair (stored once) of d_eps w0^2 / (w0^2 - w^2 + j gamma w)."""
    w0 = 2 * math.pi * f0
    wd = math.sqrt(w0 * w0 - 0.25 * gamma * gamma)
    return complex(-0.5 * gamma, wd), complex(0.0, -d_eps * w0 * w0 / (2.0 * wd))


def c4_layout(seed=4):
    """Seeded precision-study material: (eps_inf, sigma, poles) of the slab,
    poles as pole-residue pairs of two Lorentz terms (d_eps U[0.5, 2],
    f0 U[5, 20] GHz, damping f0 / 10)."""
    rng = np.random.default_rng(seed)
    poles = tuple(lorentz_pole(float(rng.uniform(0.5, 2.0)), f0, 2 * math.pi * f0 / 10)
                  for f0 in rng.uniform(5e9, 20e9, 2))
    eps_inf = float(rng.uniform(2.0, 4.0))
    return eps_inf, 0.05, poles


C4_LAYERS = 8


def c4_case(n=320, seed=4, f_max=20e9, n_modes=2):
    """Precision study: a 32 x 20 x 48 mm guide with PEC x/y walls on an
    n^3 grid, a seeded slab (z in [0.4, 0.6] L) of two-pole Lorentz
    material (eps_inf U[2, 4], pole frequencies U[5, 20] GHz, seeded
    residues, damping f0/10) with bulk conductivity 0.05 S/m, a modal DTBC
    port with n_modes modes on the open end z_min and an 8-layer CPML
    absorbing the open end z_max.  Returns (mesh, ports, dt, boundary
    conditions, info)."""
    from magnelio_amd import mesh as mm
    from magnelio_amd import ports as pp
    from magnelio_amd.boundaries import CPMLBoundary, PECBoundary

    W, H, L = C4_SIZE
    grid = mm.GridLines(*(np.linspace(0.0, s, n + 1) for s in C4_SIZE))
    eps_inf, sigma, poles = c4_layout(seed)
    slab = mm.Material(name="lorentz", epsilon=(eps_inf,) * 3, sigma=(sigma,) * 3, poles=poles)
    walls = ("xmin", "xmax", "ymin", "ymax")
    mesh = mm.Mesh.from_grid(grid, regions=[(slab, (0.0, 0.0, 0.4 * L, W, H, 0.6 * L))], pec_faces=walls)
    dt = mm.courant_dt(grid, "normal")
    ports = [pp.build_port_local("port1", "zmin", mesh, dt=dt, f_calc=f_max, n_modes=n_modes)]
    bcs = {f: PECBoundary(f) for f in walls}
    bcs["zmax"] = CPMLBoundary("zmax", grid, thickness_cells=C4_LAYERS)
    return mesh, ports, dt, bcs, {"eps_inf": eps_inf, "poles_hz": [p[0].imag / (2 * math.pi) for p in poles]}


def main_c4(args) -> int:
    """FP32 fields and curls (FP64 DTBC history, port accumulators and
    compensated DFT) against a full FP64 run of the same kernels."""
    import torch

    from magnelio_amd.monitors import MonitorFieldFrequency
    from magnelio_amd.sweep import ScatteringSweep

    t_setup = time.perf_counter()
    mesh, ports, dt, bcs, info = c4_case(args.n if args.n != 256 else 320)
    nx, ny, nz = mesh.dims
    steps = args.steps if args.steps != 512 else 3000
    f_axis = np.linspace(6e9, 19e9, 131)
    kz = round(0.75 * nz)

    def monitors():
        return [MonitorFieldFrequency(np.linspace(6e9, 19e9, 64),
                                      region=(slice(0, nx), slice(0, ny), slice(kz, kz + 1)),
                                      fields=("E",), block=64, name="plane")]

    excited = [("port1", 0), ("port1", 1)]
    out_runs = {}
    for prec in ("single", "double"):
        sw = ScatteringSweep(mesh, ports, dt=dt, n_steps=steps, f_axis=f_axis, f_max=20e9,
                             excited=excited, precision=prec, monitors_factory=monitors,
                             boundary_conditions=bcs)
        res = sw.run()
        sol = sw.solvers[0]
        out_runs[prec] = {"S": res.s_params.matrix, "seconds": res.seconds,
                          "e": sol.e.double().cpu(), "h": sol.h.double().cpu(),
                          "dft": [sol.monitors[0].result(k) for k in range(len(excited))]}
        del sw, sol, res
        torch.cuda.empty_cache()
    a, b = out_runs["single"], out_runs["double"]

    def rel(x, y):
        return float(torch.linalg.vector_norm(x - y) / torch.linalg.vector_norm(y))

    s_err = np.abs(a["S"] - b["S"])
    finite = np.isfinite(s_err)
    def stack(r):
        return np.concatenate([np.asarray(r[c]).ravel() for c in sorted(r)])

    dft_err = max(float(np.linalg.norm(stack(x) - stack(y)) / np.linalg.norm(stack(y)))
                  for x, y in zip(a["dft"], b["dft"]))
    cells = nx * ny * nz
    out = {
        "metric": "precision study: FP32 engine vs FP64 run of the same kernels (S-parameter and field error)",
        "value": float(s_err[finite].max()),
        "unit": "max |S_fp32 - S_fp64|",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": a["seconds"] / steps * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{nx}x{ny}x{nz} guide 32x20x48 mm with PEC x/y walls, seeded two-pole Lorentz "
                         "slab with sigma 0.05 S/m (eps_inf %.2f, poles %.1f/%.1f GHz), DTBC port on the "
                         "open end zmin x 2 modes, 8-layer CPML on the open end zmax, 2 excitations, %d "
                         "steps, field-plane DFT 64 freqs compensated"
                         % (info["eps_inf"], info["poles_hz"][0] / 1e9, info["poles_hz"][1] / 1e9, steps)),
            "setup_s": round(time.perf_counter() - t_setup - a["seconds"] - b["seconds"], 1),
        },
        "s_err_db": float(20 * np.log10(s_err[finite].max())),
        "field_rel_l2": {"e": rel(a["e"], b["e"]), "h": rel(a["h"], b["h"])},
        "dft_rel_l2": dft_err,
        "gcell_per_s": {"fp32": len(excited) * cells * steps / a["seconds"] / 1e9,
                        "fp64": len(excited) * cells * steps / b["seconds"] / 1e9},
    }
    print(json.dumps(out))
    return 0


COPY_BYTES = 1 << 30


FP64_PEAK_TFLOPS = 78.6
# per (direction, patch) of a grid face: the four tangential currents' complex
# multiply-adds against the column phase (the row and plane phases are per
# row and per face)
NTFF_FLOPS_PER_PAIR = 32


def main_ntff(args) -> int:
    """Far-field post of the C2 Huygens box: one pattern (all default
    directions, every patch) per step, on device-resident inputs — the
    grid-face kernel the monitor path uses, the per-patch kernel beside it."""
    import torch

    from magnelio_amd import farfield as pf
    from magnelio_amd import mesh as mm
    from magnelio_amd.boundaries import CPMLBoundary
    from magnelio_amd.monitors import MonitorFarField

    mesh = c2_mesh(args.n, length=64e-3)
    bcs = {f: CPMLBoundary(f, mesh.grid, thickness_cells=8) for f in mm.FACES}
    mon = MonitorFarField([10e9], margin_cells=2)
    mon._build(mesh.grid, bcs)
    rng = np.random.default_rng(11)
    faces = []
    for face, axis, _, _, _ in mon.faces:
        g = mon._geometry[face]
        shp = (1, g["c1"].size, g["c2"].size)
        faces.append(dict(name=face, axis=axis, sign=g["sign"], plane=g["plane"], c1=g["c1"], c2=g["c2"],
                          w1=g["w1"], w2=g["w2"],
                          bins={c: rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
                                for c in pf.TANGENTIALS[axis]}))
    dump = dict(freqs=np.array([10e9]), faces=faces, image_planes=[], source_spectrum=np.array([1.0 + 0j]))
    sets = pf.patch_sets_from_dump(dump, 0)
    theta, phi = pf.default_angular_grid()
    frames = pf.direction_frames(theta, phi)
    n_d = frames.shape[0]
    tens = pf.DeviceTransform(frames, faces=[pf.tensor_face(s) for s in sets])
    gen = pf.DeviceTransform(frames, patches=pf.patch_array(pf._concat(sets)))
    n_p = gen.np_
    st = torch.cuda.current_stream()

    def timed(t, steps, warm):
        for _ in range(warm):
            t.launch(10e9, st)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t0 = time.perf_counter()
        for a, b in ev:
            a.record(st)
            t.launch(10e9, st)
            b.record(st)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, float(np.mean([a.elapsed_time(b) for a, b in ev]))

    steps, warm = max(1, min(args.steps, 100)), max(1, min(args.warmup, 5))
    wall, ms = timed(tens, steps, warm)
    _, ms_gen = timed(gen, max(1, steps // 10), 1)
    a, b = tens.result()
    gen.launch(10e9, st)
    c, d = gen.result()
    agree = float(max(np.abs(a - c).max(), np.abs(b - d).max()) / max(np.abs(c).max(), np.abs(d).max()))
    flops = NTFF_FLOPS_PER_PAIR * float(n_p) * n_d
    achieved = flops / (ms * 1e-3) / 1e12
    out = {
        "metric": "far-field patterns/s",
        "value": steps / wall,
        "unit": "patterns/s",
        "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"near-to-far-field transform of the C2 Huygens box ({args.n}^3 grid, "
                                f"8-layer CPML + 2 margin cells): {len(sets)} grid faces, {n_p} surface "
                                f"patches with seeded bins, {n_d} directions (2-degree theta x phi grid), "
                                "one frequency per step, device-resident inputs"),
                   "patches": n_p, "directions": n_d, "face_jobs": tens.n_jobs,
                   "per_patch_kernel_ms": ms_gen, "per_patch_vs_face_max_rel": agree},
        "roofline": {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "k_ntff_face_tile",
                     "kernel_ms": ms, "alg_flops_per_launch": flops,
                     "flops_note": "32 per (direction, patch): four complex multiply-adds of the tangential "
                                   "currents; phases (one sincos per direction and row or column) not counted"},
    }
    if not args.no_cpu:
        from oracle import farfield as off

        surf = pf._concat(sets)
        sub = max(1, int(round(n_d / 64)))
        idx = np.arange(0, n_d, sub)
        host = dict(centers=surf.centers, normals=surf.normals, areas=surf.areas, E=surf.E, H=surf.H)
        th_s = np.arccos(np.clip(frames[idx, 2], -1, 1))
        t1 = time.perf_counter()
        off.ntff([host], [], 10e9, th_s, np.array([0.0]))
        t_cpu = time.perf_counter() - t1
        per_pattern = t_cpu * n_d / idx.size
        out["cpu_baseline"] = {"value": 1.0 / per_pattern, "unit": "patterns/s", "cores": _host_threads(),
                               "kind": "port",
                               "sample": (f"oracle NumPy transform (the reference's algorithm) over all "
                                          f"{n_p} patches for {idx.size} of the {n_d} directions "
                                          f"({t_cpu:.1f} s), scaled to the full grid")}
    print(json.dumps(out))
    return 0


def pmc_probe(args) -> int:
    """Launches whose counters calibrate and measure HBM traffic: a device
    copy of COPY_BYTES (1 GiB read + 1 GiB written) and `--steps` E/H curl
    launches on the c2 grid with its 8-layer CPML on six faces inside them
    (sources and monitors off)."""
    import torch

    from magnelio_amd import mesh as mm
    from magnelio_amd.solver import FITTimeDomainSolver

    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    from magnelio_amd.boundaries import CPMLBoundary

    mesh = c2_mesh(args.n)
    dt = mm.courant_dt(mesh.grid, "normal")
    bcs = {f: CPMLBoundary(f, mesh.grid, thickness_cells=8) for f in mm.FACES}
    from magnelio_amd import kernels

    # the launch shape the default line tunes to (run once without the
    # profiler first: the tuner's timing launches then come from its cache)
    sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, total_time_steps=args.steps + 2,
                              dt=dt, record=False, tile_skip="off", kchunk="auto")
    sol.setup()
    sol.e.normal_().mul_((sol.beta_e != 0).float())
    for i in range(args.steps):
        sol._step(i)
    torch.cuda.synchronize()
    cells = int(np.prod(mesh.dims))
    print(json.dumps({"copy_bytes": COPY_BYTES, "cells": cells,
                      "launch_shape": dict(zip(("kchunk", "sets_per_thread", "band"),
                                               kernels.shape_parts(sol.kchunk))),
                      "alg_bytes_update_e": 4.0 * cells * (6 + 9) + sol.cpml.slab_bytes("E", True),
                      "alg_bytes_update_h": 4.0 * cells * (6 + 9) + sol.cpml.slab_bytes("H", True),
                      "launches": args.steps}))
    return 0


def _host_threads() -> int:
    """CPU threads this process may use: OMP_NUM_THREADS when set (the
    host's share per GPU), else every CPU of its affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(mesh, coeffs, bcs, dt, nsamp=(10, 20, 40), cpml_steps=3):
    """Oracle C/OpenMP restatement of the Numba kernels on the same arrays,
    plus the NumPy restatement of the CPML slab updates (the reference's
    CPML is NumPy array code on the CPU path)."""
    from oracle import build as ob
    from oracle import cpml as oc

    lib = ob.load()
    threads = _host_threads()
    lib.fit_omp_set_threads(threads)
    nx, ny, nz = mesh.dims
    ae = coeffs.alpha_e.astype(np.float32)
    be = coeffs.beta_e.astype(np.float32)
    ah = coeffs.alpha_h.astype(np.float32)
    bh = coeffs.beta_h.astype(np.float32)
    rng = np.random.default_rng(1)
    e = rng.standard_normal(ae.size).astype(np.float32) * 1e-3
    h = rng.standard_normal(ah.size).astype(np.float32) * 1e-6
    times = []
    for n in nsamp:
        t0 = time.perf_counter()
        for _ in range(n):
            ob.update_e_f32(e, h, ae, be, nx, ny, nz)
            ob.update_h_f32(e, h, ah, bh, nx, ny, nz)
        times.append(time.perf_counter() - t0)
    # t(N) = setup + N * cost (least squares over the sample points)
    cost = float(np.polyfit(np.array(nsamp, float), np.array(times), 1)[0])
    cost = max(cost, min(t / n for t, n in zip(times, nsamp)) * 0.5)
    g = mesh.grid
    faces = [oc.CPMLFace(f, g.x, g.y, g.z, dt, np.float32, n_layers=bc.thickness_cells)
             for f, bc in bcs.items()]
    t0 = time.perf_counter()
    for _ in range(cpml_steps):
        for fc in faces:
            fc.update_e(e, h, be)
        for fc in faces:
            fc.update_h(e, h, bh)
    cpml_cost = (time.perf_counter() - t0) / cpml_steps
    rate = mesh.grid.n_cells / (cost + cpml_cost) / 1e9
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": rate,
        "unit": "Gcell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{nx}x{ny}x{nz} same grid and coefficients: E+H curl kernels (Numba-order "
                   f"C/OpenMP, {threads} threads, {'/'.join(map(str, nsamp))} steps, fit t=a+N*c: "
                   f"{cost * 1e3:.1f} ms/step) + NumPy CPML on {len(faces)} faces ({cpml_steps} steps, "
                   f"{cpml_cost * 1e3:.1f} ms/step, one thread as the reference's array code); "
                   f"TF/SF and DFT not included; {cpu}"),
        "curl_ms_per_step": cost * 1e3,
        "cpml_ms_per_step": cpml_cost * 1e3,
    }


def seed_mid_run(sol, n0: int, seed: int = 7) -> None:
    """Put a set-up solver into a seeded state at step n0 of its run:
    random fields (zero where beta_E = 0) and DFT sample buffers full of
    random samples, the step counter at n0.  Steps from there do what the
    steps of a long run do (a pipelined monitor contracts its previous
    period's samples from the first of them on)."""
    import torch

    g = torch.Generator(device=sol.e.device).manual_seed(seed)
    sol.e.normal_(generator=g).mul_(1e-3).mul_((sol.beta_e != 0).to(sol.e.dtype))
    for mon in sol.monitors:
        mon.X.normal_(generator=g).mul_(1e-3)
    sol.step_ctr.fill_(n0)
    sol._resume_step = n0
    torch.cuda.synchronize()


MID_RUN_STEP = 1024


def c3_batched(warmup: int, steps: int, freqs: int = 2001, block: int = 64, plain: bool = False,
               single_path: bool = True, batch_tuner: bool = False) -> dict:
    """The batched K = 6 sweep of the C3 guide (2 ports x 3 modes as six
    field sets of one solver, graph replays, pipelined 2001-frequency
    field-plane DFT) and the same six excitations run one per solver, each
    timed over `steps` steps after `warmup` from the mid-run state."""
    import torch

    from magnelio_amd.monitors import MonitorFieldFrequency
    from magnelio_amd.sweep import ScatteringSweep

    mesh, ports, dt = c3_case()
    nx, ny, nz = mesh.dims
    f_axis = np.linspace(1e9, C3_FMAX, max(freqs, 1))
    kz = round(0.3 * nz)

    def monitors():
        if freqs <= 0:
            return []
        return [MonitorFieldFrequency(f_axis, region=(slice(0, nx), slice(0, ny), slice(kz, kz + 1)),
                                      fields=("E",), block=block, compensated=not plain,
                                      name="post_plane")]

    n0 = MID_RUN_STEP
    G_MAX = 64  # the warm-up also replays one graph of at most this many steps
    total = n0 + warmup + G_MAX + steps
    # short graphs (a multiple of the DFT slice cadence), so that a short
    # timed window replays graphs too
    sweep = ScatteringSweep(mesh, ports, dt=dt, n_steps=total, f_axis=f_axis, f_max=C3_FMAX,
                            monitors_factory=monitors, graph_block=8, kchunk="auto")
    K = len(sweep.excited)

    extra = {"steps": 0}

    def timed(sol):
        sol.setup()
        seed_mid_run(sol, n0)
        sol.run(warmup)
        if sol.use_graphs:
            # capture (records only) and one replay (its first launch uploads
            # the graph) inside the warm-up, outside the timed region
            G = sol.graph_steps
            assert G <= G_MAX
            if sol._graph is None:
                sol._capture(sol.n_completed, G)
            sol.run(G)
            extra["steps"] = G
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol.run(steps)
        torch.cuda.synchronize()
        ok = bool(torch.isfinite(sol.e).all())
        return time.perf_counter() - t0, ok

    sol = sweep.make_solver(sweep.excited)
    t_b, ok = timed(sol)
    samples = sol.monitors[0].S if sol.monitors else 0
    bins_gb = sol.monitors[0].bytes() / 1e9 if sol.monitors else 0.0
    e_us, h_us = curl_launch_times(sol)
    from magnelio_amd import kernels

    shape = dict(zip(("kchunk", "sets_per_thread", "band"), kernels.shape_parts(sol.kchunk)))
    del sol
    torch.cuda.empty_cache()
    t_s = 0.0
    if single_path:
        for ch in sweep.excited:
            s1 = sweep.make_solver([ch])
            dt_s, ok1 = timed(s1)
            t_s += dt_s
            ok = ok and ok1
            del s1
            torch.cuda.empty_cache()
    if not ok:
        raise RuntimeError("non-finite field after the timed C3 steps")
    cells = nx * ny * nz
    rate_b = K * cells * steps / t_b / 1e9
    roof = HBM_PEAK_GBS * 1e9 / (4.0 * (18 + 12 / K)) / 1e9
    out = {"workload": (f"{nx}x{ny}x{nz} WR-90 guide, two seeded eps_r=30 posts, PEC iris with a "
                        f"seeded window, modal DTBC ports on zmin/zmax x 3 modes = K={K} field sets, "
                        "HIP graphs" + (f", pipelined field-plane DFT {f_axis.size} freqs "
                                        f"({samples} samples/step per set, block {block})"
                                        if freqs > 0 else "")),
           "K": K, "value": rate_b, "unit": "Gcell-updates/s", "ms_per_step": t_b / steps * 1e3,
           "steps": steps, "warmup": warmup + extra["steps"],
           "roofline_gcell": roof, "frac_of_roofline": rate_b / roof, "dft_bins_gb": round(bins_gb, 2)}
    # the K-set curl launches alone: algorithmic bytes 4 (6 + 9 K) per cell
    alg = 4.0 * cells * (6 + 9 * K)
    curl = {"e_us": e_us, "h_us": h_us, "launch_shape": shape, "alg_bytes_per_launch": alg,
            "frac_of_peak_e": alg / e_us / 1e3 / HBM_PEAK_GBS,
            "frac_of_peak_h": alg / h_us / 1e3 / HBM_PEAK_GBS}
    if K == 6:
        ceil = stream_ceiling(cells, 42, 18)
        curl["stream_ceiling"] = ceil
        curl["frac_of_stream_ceiling_e"] = ceil["us"] / e_us
        curl["frac_of_stream_ceiling_h"] = ceil["us"] / h_us
    out["curl_launches"] = curl
    if single_path:
        rate_s = K * cells * steps / t_s / 1e9
        out["single_excitation_path"] = {"value": rate_s, "ms_per_step_all_K": t_s / steps * 1e3,
                                         "speedup_batched": rate_b / rate_s}
    if batch_tuner:
        out["batch_tuner"] = c3_batch_tuner(sweep)
    return out


def c3_batch_tuner(sweep) -> dict:
    """The sweep's batch="auto" choice on this workload: one throw-away
    solver per candidate K timed over whole graph replays (a cache of its
    own, so the choice is measured here)."""
    import tempfile

    from magnelio_amd import autotune

    with tempfile.TemporaryDirectory() as td:
        path = Path(td) / "autotune.json"
        best = autotune.batch_for(sweep, path)
        (entry,) = [v for k, v in autotune._load(path).items() if "|batch|" in k]
    return {"batch": best, "step_ms": entry["step_ms"], "sweep_step_ms": entry["sweep_step_ms"]}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=512)
    ap.add_argument("--warmup", type=int, default=256)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--k", type=int, default=1, help="field sets advanced per launch")
    ap.add_argument("--freqs", type=int, default=1001, help="DFT monitor frequencies (0: none)")
    ap.add_argument("--dft-block", type=int, default=256, help="steps per DFT contraction")
    ap.add_argument("--dft-plain", action="store_true", help="uncompensated float64 bins")
    ap.add_argument("--dft-slice-steps", type=int, default=None,
                    help="steps between pipelined contraction slices (default: chosen for occupancy)")
    ap.add_argument("--dft-fsplit", type=int, default=None,
                    help="frequency groups of a contraction slice (default: chosen for occupancy)")
    ap.add_argument("--dft-batched", action="store_true",
                    help="one contraction of the whole sample buffer per period instead of slices")
    ap.add_argument("--no-tfsf", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="leave the batched K=6 sweep out of the line")
    ap.add_argument("--batch-tuner", action="store_true",
                    help="c3: also time the sweep's batch=\"auto\" candidates (K = 6, 3, 2, 1) and report the choice")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="a short stencil-only run of the c2 grid plus a 1 GiB device copy, "
                         "for HBM traffic counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--workload", choices=("c2", "c3", "c4", "c5", "ntff"), default="c2",
                    help="c2: single large job (default line); c3: batched K=6 sweep vs one "
                         "excitation per launch; c4: FP32 vs FP64 study; c5: serving loop vs one "
                         "job at a time; ntff: far-field post of the C2 Huygens box")
    ap.add_argument("--jobs", type=int, default=200, help="c5: number of sweep jobs")
    ap.add_argument("--rate", type=float, default=1.5, help="c5: Poisson arrival rate [jobs/s]")
    ap.add_argument("--step-scale", type=float, default=0.05,
                    help="c5: fraction of the 2000-20000 steps each job runs")
    ap.add_argument("--max-active", type=int, default=8, help="c5: resident jobs at once")
    ap.add_argument("--cells-max", type=int, default=384 ** 3,
                    help="c5: largest job (cells log-uniform from 32^3 to this)")
    ap.add_argument("--pack-cells", type=int, default=0,
                    help="c5: jobs of at most this many cells step in shared launches (0: none)")
    ap.add_argument("--fp64-check", choices=("off", "column", "all"), default="off",
                    help="c5: float64 run of each job's first excitation (column) or all, "
                         "its S-parameter error reported per job")
    ap.add_argument("--c5-mode", choices=("both", "continuous", "sequential"), default="both",
                    help="c5: serving modes to run (both: continuous, then one job at a time)")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this engine advances one job per device "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        return 2
    if args.pmc_probe:
        return pmc_probe(args)
    if args.workload == "c3":
        return main_c3(args)
    if args.workload == "c5":
        return main_c5(args)
    if args.workload == "c4":
        return main_c4(args)
    if args.workload == "ntff":
        return main_ntff(args)

    import torch

    from magnelio_amd import mesh as mm
    from magnelio_amd.boundaries import CPMLBoundary
    from magnelio_amd.monitors import MonitorFarField
    from magnelio_amd.solver import FITTimeDomainSolver
    from magnelio_amd.sources import PlaneWaveSource

    t_setup = time.perf_counter()
    length = 64e-3
    mesh = c2_mesh(args.n, length=length)
    nx, ny, nz = mesh.dims
    dt = mm.courant_dt(mesh.grid, "normal")
    bcs = {f: CPMLBoundary(f, mesh.grid, thickness_cells=8) for f in mm.FACES}
    n0 = MID_RUN_STEP
    total = n0 + args.warmup + args.steps
    sources, monitors = [], []
    if not args.no_tfsf:
        sources.append(PlaneWaveSource(direction=(0, 0, 1), polarization=(1, 0, 0),
                                       corners=((0.15 * length,) * 3, (0.85 * length,) * 3),
                                       f_max=20e9, waveform="gaussian"))
    if args.freqs > 0:
        # Huygens box 10 cells in from every face (8 CPML layers + 2)
        monitors.append(MonitorFarField(np.linspace(1e9, 20e9, args.freqs), margin_cells=2,
                                        block=args.dft_block, name="far_field",
                                        compensated=not args.dft_plain,
                                        pipelined=not args.dft_batched, fsplit=args.dft_fsplit,
                                        slice_steps=args.dft_slice_steps))
    sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, total_time_steps=total, dt=dt,
                              precision="single", nsets=args.k, record=False, sources=sources,
                              monitors=monitors, kchunk="auto")
    sol.setup()
    from magnelio_amd import kernels

    shape = dict(zip(("kchunk", "sets_per_thread", "band"), kernels.shape_parts(sol.kchunk)))
    seed_mid_run(sol, n0)
    setup_s = time.perf_counter() - t_setup

    for i in range(args.warmup):
        sol._step(n0 + i)
    torch.cuda.synchronize()
    sol.timing_events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                         for _ in range(args.steps)]
    sol.timed_events = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        sol._step(n0 + args.warmup + i)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    e_ms = float(np.mean([a.elapsed_time(b) for a, b in sol.timed_events]))
    if not torch.isfinite(sol.e).all():
        print("non-finite field after the timed steps", file=sys.stderr)
        return 1

    cells = nx * ny * nz
    K = args.k
    rate = K * cells * args.steps / wall / 1e9
    # E phase, algorithmic bytes: 6 coefficient words shared by the K sets +
    # K x (read E, read H, write E) per cell; the CPML psi words the launch
    # reads and writes are counted apart
    fused = sol.cpml is not None and sol.tile_plan is None and sol._dispersion is None
    bytes_e = 4.0 * cells * (6 + 9 * K)
    psi_e = sol.cpml.slab_bytes("E", in_curl=True) if fused else 0.0
    achieved = bytes_e / (e_ms * 1e-3) / 1e9
    mon = monitors[0] if monitors else None
    out = {
        "metric": "Yee Gcell-updates/s & % HBM roofline",
        "value": rate,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"single large job {nx}x{ny}x{nz} seeded non-uniform grid (ratio <= 1.5), "
                         "40 seeded dielectric bricks/spheres eps_r U[1,10], 8-layer CPML on six faces"
                         + ("" if args.no_tfsf else
                            ", TF/SF +z x-polarised Gaussian (f_max 20 GHz) on [0.15L, 0.85L]^3")
                         + ("" if mon is None else
                            f", far-field Huygens-box DFT {args.freqs} freqs 1-20 GHz, 10 cells in, "
                            f"{mon.S} samples/step, pipelined contraction over {mon.period}-step "
                            "periods (each step contracts 1/" f"{mon.period} of the previous period), "
                            + ("plain" if args.dft_plain else "compensated") + " FP64 bins")
                         + f"; timed steps from a seeded mid-run state at step {n0} "
                         "(random fields, full sample buffers)"),
            "K": K,
            "launch_shape": shape,  # autotuned at setup (0: the library's choice)
            "dft_bins_gb": round(mon.bytes() / 1e9, 2) if mon is not None else 0.0,
            "dft_slice": ({"steps_between": mon.cadence, "frequency_groups": mon.fsplit,
                           "pipelined": mon.pipelined} if mon is not None else None),
            "dft_contraction_launches_timed": (sum(1 for i in range(args.steps)
                                                   if (n0 + args.warmup + i - mon.n_start) % mon.cadence
                                                   == mon.cadence - 1) if mon is not None else 0),
            "dft_contractions_timed": (args.steps / mon.period if mon is not None and mon.pipelined
                                       else None),
            "setup_s": round(setup_s, 2),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": K_UPDATE_E_TRAFFIC.get((args.n, K, fused)),
            "traffic_source": ("stored: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this launch "
                               "(bench.py --pmc-probe), profiles/k_update_traffic_c2_256cube_k1.md"),
            "kernel": "k_update_e",
            "kernel_ms": e_ms,
            "alg_bytes_per_launch": bytes_e,
            "cpml_psi_bytes_per_launch": psi_e,
        },
    }
    if K == 1:
        # what the device streams for this launch's algorithmic words alone
        ceil = stream_ceiling(cells, 12, 3)
        out["roofline"]["stream_ceiling"] = ceil
        out["roofline"]["frac_of_stream_ceiling"] = achieved / ceil["GBps"]
    coeffs = sol._coeffs  # float64 host alpha/beta (host-built: the mesh has a TF/SF source)
    del sol, monitors, mon
    torch.cuda.empty_cache()
    if not args.no_c3:
        out["c3_batched"] = c3_batched(args.warmup, args.steps)
    if not args.no_cpu:
        if coeffs is None:
            from magnelio_amd import mesh as _m

            coeffs = _m.update_coefficients(_m.build_m_eps(mesh), _m.build_m_mu(mesh),
                                            _m.build_m_sigma(mesh), _m.build_m_sigma_m(mesh), dt,
                                            _m.pec_flat(mesh.pec_mask_edges, mesh.dims))
        out["cpu_baseline"] = cpu_baseline(mesh, coeffs, bcs, dt)
    print(json.dumps(out))
    return 0


def main_c3(args) -> int:
    """Batched multi-excitation sweep: K = 6 (port, mode) excitations as six
    field sets of one solver against the same six run one per solver."""
    t_setup = time.perf_counter()
    freqs = 2001 if args.freqs == 1001 else args.freqs
    block = 64 if args.dft_block == 256 else args.dft_block
    r = c3_batched(args.warmup, args.steps, freqs=freqs, block=block, plain=args.dft_plain,
                   batch_tuner=args.batch_tuner)
    out = {
        "metric": "batched sweep Gcell-updates/s (K excitations per launch) vs one excitation per launch",
        "value": r["value"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched multi-excitation sweep " + r["workload"]
                   + f"; timed steps from a seeded mid-run state at step {MID_RUN_STEP}",
                   "K": r["K"], "dft_bins_gb": r["dft_bins_gb"],
                   "wall_s": round(time.perf_counter() - t_setup, 1)},
        "roofline_gcell": r["roofline_gcell"],
        "frac_of_roofline": r["frac_of_roofline"],
        "curl_launches": r["curl_launches"],
        "single_excitation_path": r["single_excitation_path"],
    }
    if "batch_tuner" in r:
        out["batch_tuner"] = r["batch_tuner"]
    print(json.dumps(out))
    return 0


def main_c5(args) -> int:
    """Serving: seeded Poisson stream of sweep jobs through the continuous-
    batching loop, then the same stream one job at a time."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor

    from magnelio_amd.serve import SweepServer

    # host builds (mesh fill, port mode solves) run in worker processes forked
    # before this process touches the GPU; they never use it
    workers = 10
    pool = ProcessPoolExecutor(max_workers=workers, mp_context=multiprocessing.get_context("fork"),
                               initializer=_c5_worker_init)
    list(pool.map(abs, range(4 * workers)))  # start every worker now

    def stream_of_jobs():
        return c5_jobs(n_jobs=args.jobs, rate=args.rate, step_scale=args.step_scale,
                       fp64_check=args.fp64_check, cells=(32 ** 3, args.cells_max))

    # warm-up: one small job loads the library and the runtime
    warm = c5_jobs(n_jobs=1, seed=99, cells=(24 ** 3, 24 ** 3), steps=(200, 200))
    SweepServer(build_pool=pool).run(warm)
    reports = {}
    modes = ("continuous", "sequential") if args.c5_mode == "both" else (args.c5_mode,)
    for mode in modes:
        print(f"c5 {mode}: {args.jobs} jobs", file=sys.stderr, flush=True)
        srv = SweepServer(max_active=args.max_active, build_pool=pool, progress=max(1, args.jobs // 60),
                          pack_cells=args.pack_cells)
        reports[mode] = srv.run(stream_of_jobs(), mode=mode)
        r = reports[mode]
        print(f"c5 {mode}: {r.jobs_per_s:.3f} jobs/s p50 {r.p50:.2f} s p99 {r.p99:.2f} s",
              file=sys.stderr, flush=True)
    c = reports.get("continuous", reports.get("sequential"))
    q = reports.get("sequential")
    same = (all(np.array_equal(a.result.matrix, b.result.matrix, equal_nan=True)
                for a, b in zip(reports["continuous"].records, q.records))
            if len(reports) == 2 else None)
    recs = c.records
    out = {
        "metric": ("serving sweep jobs/s with p50/p99 job latency (continuous batching)"
                   if "continuous" in reports else "serving sweep jobs/s with p50/p99 job latency "
                   "(one job at a time)"),
        "value": c.jobs_per_s,
        "unit": "jobs/s",
        "n_gpus": 1,
        "steps": int(sum(r.n_steps for r in recs)),
        "warmup": 1,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.jobs} sweep jobs, Poisson arrivals {args.rate}/s (seed 5), cells "
                         f"log-uniform 32^3-{round(args.cells_max ** (1 / 3))}^3 (a WR-90 guide with z-end ports for 1-2 ports, a "
                         "cross of two WR-90 guides with z- and x-end ports for 3-4; 1-4 dielectric "
                         "bricks/spheres eps_r U[1,10]), 1-4 modal DTBC ports x 1-3 modes, steps "
                         "U{2000..20000} x "
                         f"{args.step_scale}; admission by HBM footprint, up to {args.max_active} "
                         "resident jobs on separate streams, graph-replayed 256-step chunks, mesh and "
                         f"port builds in {workers} host processes"),
            "cells_min": int(min(r.cells for r in recs)),
            "cells_max": int(max(r.cells for r in recs)),
            "cell_updates": float(sum(r.cell_updates for r in recs)),
        },
        "latency_s": {"p50": c.p50, "p99": c.p99, **c.breakdown()},
        "gcell_per_s": c.gcell_per_s,
        "peak_resident": c.peak_resident,
        "peak_footprint_gb": round(c.peak_bytes / 1e9, 2),
        "one_job_at_a_time": ({"jobs_per_s": q.jobs_per_s, "p50": q.p50, "p99": q.p99,
                               "gcell_per_s": q.gcell_per_s, **q.breakdown(),
                               # device time of the jobs run alone: n / this is the
                               # throughput ceiling of one-at-a-time service
                               "device_s": float(sum(r.finished - r.admitted for r in q.records))}
                              if q is not None else None),
        "results_identical": same,
        "pack_cells": args.pack_cells,
        "packed_jobs": int(sum(r.packed for r in recs)),
        # each job's single-precision S-matrix against a float64 run of its
        # first excitation (or all) on the same kernels, when asked for
        "fp64_check": ({"mode": args.fp64_check,
                        "max_abs_dS_max": max(r.fp64["max_abs_dS"] for r in recs),
                        "max_abs_dS_p50": float(np.median([r.fp64["max_abs_dS"] for r in recs])),
                        "db_max": max(r.fp64["db"] for r in recs),
                        "check_cell_updates": float(sum(r.cells * r.check_sets * r.n_steps
                                                        for r in recs))}
                       if args.fp64_check != "off" else None),
        "ports_per_job": {str(k): int(v) for k, v in zip(*np.unique(
            [len([t for t in j.build.tasks if t.startswith("port")]) for j in stream_of_jobs()],
            return_counts=True))},
        "host_build_workers": workers,
    }
    pool.shutdown()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
