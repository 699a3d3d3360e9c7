"""FIT-TD leapfrog driver of the HIP backend (``backend="hip"``).

Mirrors ``FITTimeDomainSolver`` (src/magnelio/solver/fit_td.py:135-1386):
same constructor fields (mesh, boundary_conditions, ports, total_time_steps,
dt, energy_stop_db, port_signal_stop_db, precision) and the same
``setup`` / ``run`` / ``state_dict`` / ``load_state_dict`` surface, plus
``nsets`` / ``excitations``: K field sets of one mesh — one per excited
(port, mode) channel of an S-parameter sweep — advance together, every launch
reading the shared material coefficients once for all K.

Step order (fit_td.py:950-1007), t = n dt:
  E phase    ADE field stash                           mgh_ade_save
             curl update                               mgh_update_e
             ADE history subtraction + pole advance    mgh_ade_update
             CPML faces                                mgh_cpml_e
  sources    TF/SF E injection at t + dt                mgh_tfsf_apply
  ports      modal DTBC/Mur step, write-back, V/I record mgh_port_step
             lumped RLC elements                       mgh_lumped_step
  H phase    curl update, CPML faces                   mgh_update_h, mgh_cpml_h
Without dispersion and listed (tile-skipping) launches, the CPML runs with
the curl: the curl launch (mgh_update_e_pml_core / _h_pml_core) applies the
faces' terms, in face order, to each slab element before its store (its
planes split so that the x-slab elements fill waves of their own); only when
the x slabs are not both at the domain ends does a launch over the x-slab
columns (mgh_cpml_xcols_e/h) apply theirs — the same values as the separate
slab passes.
  sources    TF/SF H injection at t + 3dt/2
  monitors   DFT accumulation                          mgh_dft_*
  counter    n <- n + 1 on the device                  mgh_step_advance
Nothing in the loop synchronises the host except the optional energy /
port-signal stop checks every ``check_interval`` steps.  With
``use_graphs=True`` one step is captured as a HIP graph (through torch's
graph API) after two eager warm-up steps and replayed.

All state lives on the device; ``state_dict`` exports it to host NumPy in
the reference's layout and keys (fit_td.py:1264-1353).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, autotune, config, dtbc, kernels, layout
from . import mesh as _mesh
from .dispersion import DeviceDispersion
from .sibc import DeviceSIBC
from .lumped import DeviceLumped
from .tileskip import build_tile_plan
from .boundaries import CPMLBoundary, DeviceCPML, PECBoundary, PMCBoundary
from .ports import PortModel

# split port step geometry (csrc/mgh_ports.h)
PORT_CHUNK, PORT_MAX_MODES = 1024, 8
PORT_PQ = 4 * PORT_MAX_MODES



def _face_bits(faces) -> int:
    """xmin 1, xmax 2, ymin 4, ymax 8, zmin 16, zmax 32 (mgh_build_coefficients)."""
    return sum(1 << _mesh.FACES.index(f) for f in set(faces))


PRECISIONS = {"single": (np.float32, torch.float32), "double": (np.float64, torch.float64)}

# port-signal stall watchdog: arming level below the run peak and the
# shortest fit window, in checks (fit_td.py:62-67)
STALL_ARM_DB = 40.0
STALL_MIN_WINDOW_CHECKS = 10


class SignalStallDetector:
    """Decides when the port-signal threshold cannot be reached before the
    runtime cap (fit_td.py:76-132): once the envelope sits ``arm_db`` below
    its peak, a least-squares line through the last ``window`` check
    samples (dB against step) is extrapolated to the threshold; a flat or
    rising envelope, or a crossing beyond ``cap_step``, ends the run now —
    the outcome the cap would give, without the steps in between."""

    def __init__(self, arm_db: float, window: int, cap_step: int) -> None:
        self.arm_db = float(arm_db)
        self.window = max(int(window), 2)
        self.cap_step = int(cap_step)
        self.slope_db_per_step: float | None = None
        self._samples: list[tuple[int, float]] = []

    def reset(self) -> None:
        self._samples.clear()

    def observe(self, n: int, sig_db: float, threshold_db: float) -> bool:
        if sig_db > -self.arm_db:
            self._samples.clear()
            return False
        self._samples.append((n, float(sig_db)))
        if len(self._samples) > self.window:
            del self._samples[0]
        elif len(self._samples) < self.window:
            return False
        steps, vals = (np.array(c, dtype=np.float64) for c in zip(*self._samples))
        slope = float(np.polyfit(steps, vals, 1)[0])
        self.slope_db_per_step = slope
        remaining = threshold_db - float(vals[-1])
        if remaining >= 0.0:
            return False
        if slope >= 0.0:
            return True
        return n + remaining / slope > self.cap_step


@dataclass
class ChannelExcitation:
    """Set s of the batch excites ``mode`` of port ``port`` with ``waveform``."""

    port: str
    mode: int
    waveform: object  # callable t -> amplitude [sqrt(W)]


class DevicePorts:
    """Device arrays + descriptors of the modal ports for K field sets."""

    WORDS = 32

    def __init__(self, ports: list[PortModel], dims, nsets: int, n_steps: int, device,
                 excitations: list[ChannelExcitation | None], dt: float, record: bool = True,
                 kernel_init: int = dtbc.N_KERNEL_INIT):
        """kernel_init: length of the first DTBC kernel version; version v
        (length kernel_init * 2^v) is in force while the step index is below
        it (the reference's schedule, dtbc.py:317-363)."""
        if kernel_init < 2:
            raise ValueError(f"kernel_init must be >= 2, got {kernel_init}")
        self.ports = ports
        self.nsets = nsets
        self.n_steps = n_steps
        self.device = device
        self._keep = []
        self.desc = np.zeros(self.WORDS * len(ports), dtype=np.int64)
        self.V_rec, self.I_rec, self.state, self.w_hist, self.s_hist, self.absmax = [], [], [], [], [], []
        self.comp = []  # per port: (nsets, 2, n_u + n_v) complement states or None
        f64 = torch.float64
        for pi, P in enumerate(ports):
            pl = P.plane
            n_u, n_v, M = pl.e_u_indices.size, pl.e_v_indices.size, P.n_modes

            def dev(a, dtype=f64):
                t = torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype)
                self._keep.append(t)
                return t

            idx_port = layout.index_to_device(np.concatenate([pl.e_u_indices, pl.e_v_indices]), *dims, "E")
            idx_int = layout.index_to_device(
                np.concatenate([pl.e_u_indices_interior, pl.e_v_indices_interior]), *dims, "E")
            idx_h = layout.index_to_device(np.concatenate([pl.h_v_indices, pl.h_u_indices]), *dims, "H")
            params = np.zeros((M, 8))
            for m in range(M):
                r = P.dtbc_r[m]
                if r is not None:
                    q = P.dtbc_q[m]
                    r2, q2 = r * r, q * q
                    params[m, 0] = 1.0
                    params[m, 1] = r2
                    params[m, 2] = q2
                    params[m, 5] = 2.0 - 2.0 * r2 - q2
                params[m, 3] = P.mur_r[m]
                params[m, 4] = P.record_scale[m]
            kern = P.packed_dtbc_kernels(max(n_steps, 1), kernel_init)
            state = dev(np.zeros((nsets, M, 4)))
            cap = max(n_steps, 1)
            any_dtbc = any(r is not None for r in P.dtbc_r)
            w_hist = dev(np.zeros((nsets, M, cap if any_dtbc else 1)))
            s_hist = dev(np.zeros((nsets, M, cap if any_dtbc else 1)))
            exc = np.full(nsets, -1, dtype=np.int32)
            src_tab = np.zeros((nsets, n_steps))
            mur_tab = None
            for s, ex in enumerate(excitations):
                if ex is None or ex.port != P.name:
                    continue
                exc[s] = ex.mode
                scale = float(P.source_scale[ex.mode])
                if P.dtbc_r[ex.mode] is not None:
                    # DTBC ghost source s^n = scale * waveform(t - dt), t = n dt + dt
                    for n in range(n_steps):
                        t = n * dt
                        src_tab[s, n] = scale * ex.waveform((t + dt) - dt)
                else:
                    if mur_tab is None:
                        mur_tab = np.zeros((nsets, n_steps, 4))
                    mur_tab[s] = _mur_incident_table(ex.waveform, scale, float(P.tau[ex.mode]), dt, n_steps)
            rec_cap = n_steps if record else 0
            V = dev(np.zeros((nsets, M, max(rec_cap, 1))))
            I = dev(np.zeros((nsets, M, max(rec_cap, 1))))
            absmax = dev(np.zeros(nsets))
            t_idx = [dev(idx_port, torch.int64), dev(idx_int, torch.int64), dev(idx_h, torch.int64)]
            t_me = dev(np.concatenate([P.me_u_port, P.me_v_port]))
            t_mi = dev(np.concatenate([P.me_u_int, P.me_v_int]))
            t_mh = dev(np.concatenate([P.mh_v, P.mh_u]))
            t_pe, t_ph = dev(P.prof_e()), dev(P.prof_h())
            t_par = dev(params)
            t_kern = dev(kern) if kern.size else None
            t_exc = dev(exc, torch.int32)
            t_src = dev(src_tab if n_steps > 0 else np.zeros((nsets, 1)))
            t_mur = dev(mur_tab) if mur_tab is not None else None
            w = self.desc[pi * self.WORDS:(pi + 1) * self.WORDS]
            w[0], w[1], w[2] = n_u, n_v, M
            w[3], w[4], w[5] = (t.data_ptr() for t in t_idx)
            w[6], w[7], w[8] = t_me.data_ptr(), t_mi.data_ptr(), t_mh.data_ptr()
            w[9], w[10], w[11] = t_pe.data_ptr(), t_ph.data_ptr(), t_par.data_ptr()
            w[12] = t_kern.data_ptr() if t_kern is not None else 0
            w[13] = kern.shape[1] if kern.size else 0
            w[14] = kernel_init
            w[15], w[16], w[17] = state.data_ptr(), w_hist.data_ptr(), s_hist.data_ptr()
            w[18] = w_hist.shape[2] if any_dtbc else 0
            w[19] = t_exc.data_ptr()
            w[20], w[21] = t_src.data_ptr(), n_steps
            w[22] = t_mur.data_ptr() if t_mur is not None else 0
            w[23], w[24], w[25] = (V.data_ptr(), I.data_ptr(), rec_cap) if record else (0, 0, 0)
            w[26] = absmax.data_ptr()
            # scratch of the split port step: projection / record partials per
            # (set, chunk of 1024 plane points or history terms), voltages
            n_pts = n_u + n_v
            nch = max(-(-n_pts // PORT_CHUNK), -(-int(w[18]) // PORT_CHUNK), 1)
            work_len = nsets * (2 * nch * PORT_PQ + PORT_MAX_MODES)
            work = dev(np.zeros(work_len))
            w[27], w[28] = work.data_ptr(), work_len
            comp = None
            if P.complement_active:
                # per-edge Mur of the port-unrepresented remainder
                # (operator.py:1395-1427, :1499-1522): r, live, and per set
                # [interior remainder | written remainder] of the last step
                t_cr, t_cl = dev(P.comp_r), dev(P.comp_live)
                comp = dev(np.zeros((nsets, 2, n_pts)))
                w[29], w[30], w[31] = t_cr.data_ptr(), t_cl.data_ptr(), comp.data_ptr()
            self.comp.append(comp)
            self.V_rec.append(V)
            self.I_rec.append(I)
            self.state.append(state)
            self.w_hist.append(w_hist)
            self.s_hist.append(s_hist)
            self.absmax.append(absmax)

    @property
    def desc_ptr(self) -> int:
        return int(self.desc.ctypes.data)


def _mur_incident_table(waveform, scale: float, tau: float, dt: float, n_steps: int) -> np.ndarray:
    """Per step: src(t^{n+1}), and the ring-buffer linear look-backs at tau,
    tau + dt and dt (operator.py:1435-1463, _interp_delayed :1539-1562)."""
    buf_len = max(int(math.ceil((tau + dt) / dt)) + 3, 5)
    out = np.zeros((n_steps, 4))
    buf: list[float] = []

    def look(delay: float) -> float:
        n = len(buf)
        if n == 0 or dt <= 0.0:
            return 0.0
        frac = delay / dt
        i0 = int(math.floor(frac))
        a = frac - i0
        if n - 1 - i0 < 0:
            return 0.0
        v0 = buf[n - 1 - i0]
        v1 = buf[n - 2 - i0] if n - 2 - i0 >= 0 else 0.0
        return (1.0 - a) * v0 + a * v1

    for n in range(n_steps):
        t = n * dt + dt
        v = scale * waveform(t)
        buf.append(v)
        if len(buf) > buf_len:
            buf.pop(0)
        out[n] = (v, look(tau), look(tau + dt), look(dt))
    return out


@dataclass
class FITTimeDomainSolver:
    """HIP FIT-TD solver (see module docstring)."""

    mesh: _mesh.Mesh
    boundary_conditions: dict = field(default_factory=dict)
    ports: list = field(default_factory=list)
    sources: list = field(default_factory=list)
    total_time_steps: int | None = 1000  # None: unbounded, ended by a stop criterion or the cap
    dt: float = 1e-12
    verbose: bool = False
    energy_stop_db: float | None = None
    port_signal_stop_db: float | None = None
    port_signal_min_steps: int | None = None  # the port-signal criterion arms at this step
    max_time_steps: int | None = None  # runtime cap of an unbounded run (absolute step)
    energy_check_interval: int | None = None
    monitors: list = field(default_factory=list)
    lumped: list = field(default_factory=list)
    precision: str | None = None  # "single" / "double"; None: MAGNELIO_PRECISION, else single
    nsets: int = 1
    excitations: list | None = None  # per set: ChannelExcitation | None
    record: bool = True
    curl64: bool = False
    # curl launch shape word (kernels.launch_shape: z planes per thread, sets
    # per thread, XCD band), or "auto" (timed at setup, autotune.py)
    # (None: MAGNELIO_HIP_LAUNCH_SHAPE, else two z planes per thread)
    kchunk: int | str | None = None
    use_graphs: bool | None = None  # None: MAGNELIO_GPU_GRAPHS, else off
    graph_block: int = 1  # minimum steps per captured graph
    dtbc_kernel_init: int = dtbc.N_KERNEL_INIT  # first DTBC kernel length (then doubled)
    coefficients: str = "auto"  # alpha/beta built on the "device", the "host", or "auto"
    # dead-unit skipping: "auto" (when it skips >= 2 %), "on", "off"; None:
    # MAGNELIO_TILE_SKIP ("0" off), else auto
    tile_skip: str | None = None
    device: str = "cuda:0"
    # (m_eps, m_mu, m_sigma, m_sigma_m) flat float64 vectors in the reference
    # layout, unflattened — e.g. the reference's own build_M_* of a conformal
    # mesh; None: built here from the material ids
    mass_matrices: tuple | None = None
    # surface-impedance walls: a spec with ``surfaces`` and ``fits`` (the
    # reference's SIBCSpec or magnelio_amd.sibc.SIBCSpec), None for plain PEC
    sibc: object | None = None

    def __post_init__(self):
        # an explicit argument wins over the environment (config.py)
        self.precision = config.precision(self.precision)
        self.use_graphs = config.graphs(self.use_graphs)
        self.tile_skip = config.tile_skip(self.tile_skip)
        self.kchunk = config.launch_shape(self.kchunk)
        if self.kchunk != "auto" and not (isinstance(self.kchunk, int)
                                          and kernels.shape_parts(self.kchunk)[0] >= 1):
            raise ValueError(f"kchunk must be a launch shape word or 'auto', got {self.kchunk!r}")
        if self.excitations is None:
            self.excitations = [None] * self.nsets
        if len(self.excitations) != self.nsets:
            raise ValueError("one excitation entry (or None) per field set")
        self._setup_done = False
        # optional (start, end) event pairs recorded around the E-phase kernel
        # of eager steps, on the launch stream (used by the benchmark)
        self.timing_events: list = []
        self.timed_events: list = []
        self._resume_step = 0
        self._stop_reason = None
        self._final_signal_db = None
        self._stop_requested = False
        self._aborted = False
        self._graph = None
        self._graph_len = 0
        self._graph_start = 0
        self.energy_trace: list = []  # (step, energy per set) at every check
        self.set_stops: list = [None] * self.nsets  # per set: (reason, step) once its criterion fired

    # ------------------------------------------------------------------
    def prepare_host(self) -> None:
        """Host half of ``setup`` (material matrices, PEC masks, ADE blocks,
        alpha/beta in float64, device-layout host copies).  Touches no device
        and may run in a worker thread; ``setup`` calls it when not done."""
        if getattr(self, "_host", None) is not None:
            return
        mesh = self.mesh
        dims = mesh.dims
        self.dims = dims
        dt = self.dt
        np_dt, t_dt = PRECISIONS[self.precision]
        self._np_dtype, self._t_dtype = np_dt, t_dt
        if self._device_coefficients():
            self._prepare_device_coefficients()
            return
        if self.mass_matrices is not None:
            ne, nh = kernels.n_e(*dims), kernels.n_h(*dims)
            m_eps, m_mu, m_sig, m_sig_m = (np.asarray(a, dtype=np.float64).copy()
                                           for a in self.mass_matrices)
            if m_eps.size != ne or m_sig.size != ne or m_mu.size != nh or m_sig_m.size != nh:
                raise ValueError("mass_matrices must be (m_eps, m_mu, m_sigma, m_sigma_m) of lengths "
                                 f"({ne}, {nh}, {ne}, {nh})")
        else:
            m_eps = _mesh.build_m_eps(mesh)
            m_mu = _mesh.build_m_mu(mesh)
            m_sig = _mesh.build_m_sigma(mesh)
            m_sig_m = _mesh.build_m_sigma_m(mesh)
        pec = mesh.pec_mask_edges
        pec_raw = _mesh.pec_flat(pec, dims)
        port_faces = []
        for P in self.ports:
            port_faces.append(P.face)
            m_eps = _mesh.flatten_port_mass(m_eps, dims, P.face)
            m_mu = _mesh.flatten_port_mu(m_mu, dims, P.face)
        for f in port_faces:
            pec = _mesh.flatten_port_pec(pec, dims, f)
        self._port_faces = set(port_faces)
        pec_e = _mesh.pec_flat(pec, dims)
        self._pec_e = pec_e
        w_e = w_h = None
        self._dispersion = self._dispersion_mu = self._sibc = None
        for src in self.sources:
            if hasattr(src, "prepare"):
                src.prepare(self)
        # ADE: blocks over the unflattened PEC mask (_dispersion.py:296-306)
        self._dispersion = DeviceDispersion.from_mesh(mesh, dt, pec_raw, "E")
        if self._dispersion is not None:
            w_e = self._dispersion.w_half
        # SIBC walls: the instantaneous part multiplies h_mid like sigma*, so
        # W folds into M_sigma_m (fit_td.py:419-429); frozen faces excluded
        if self.sibc is not None:
            self._sibc = DeviceSIBC.from_spec(self.sibc, dims, dt, frozen=m_mu <= 0)
            if self._sibc is not None:
                m_sig_m = m_sig_m + self._sibc.W
        # mu(w): faces with M_mu = 0 (after the port flattening) stay frozen
        # (fit_td.py:436-441)
        self._dispersion_mu = DeviceDispersion.from_mesh(mesh, dt, m_mu <= 0, "H")
        if self._dispersion_mu is not None:
            w_h = self._dispersion_mu.w_half
        co = _mesh.update_coefficients(m_eps, m_mu, m_sig, m_sig_m, dt, pec_e, w_e, w_h)
        self._coeffs = co
        self._m_eps_host, self._m_mu_host = m_eps, m_mu

        def lay(a, side):
            return np.ascontiguousarray(layout.to_device(np.asarray(a).astype(np_dt), *dims, side))

        self._host = {"alpha_e": lay(co.alpha_e, "E"), "beta_e": lay(co.beta_e, "E"),
                      "alpha_h": lay(co.alpha_h, "H"), "beta_h": lay(co.beta_h, "H"),
                      "m_eps": lay(m_eps, "E"), "m_mu": lay(m_mu, "H")}

    def release_host(self) -> None:
        """Drop the float64 host arrays that ``prepare_host`` keeps after
        ``setup`` (coefficients and masses; the set-up's binds have read
        them).  For callers that hold many solvers, such as the serving
        loop; the device state is untouched."""
        if not self._setup_done:
            raise RuntimeError("release_host() before setup()")
        self._coeffs = None
        self._m_eps_host = self._m_mu_host = None

    def _device_coefficients(self) -> bool:
        """Whether alpha/beta and the masses are built on the device: not for
        given mass matrices, SIBC walls, ADE materials or TF/SF sources
        (their host set-up reads the float64 coefficients) nor for a
        hand-edited PEC mask."""
        if self.coefficients not in ("auto", "host", "device"):
            raise ValueError(f"coefficients must be 'auto', 'host' or 'device', got {self.coefficients!r}")
        if self.coefficients == "host":
            return False
        why = None
        if self.mass_matrices is not None:
            why = "given mass matrices"
        elif self.sibc is not None:
            why = "surface-impedance walls"
        elif any(getattr(m, "poles", ()) or getattr(m, "poles_mu", ())
                 for m in self.mesh.material_library.values()):
            why = "dispersive materials"
        elif self.sources:
            why = "sources"
        elif not getattr(self.mesh, "pec_rule", False):
            why = "a hand-edited PEC mask"
        if why is not None and self.coefficients == "device":
            raise ValueError(f"device-built coefficients do not cover {why}")
        return why is None

    def _prepare_device_coefficients(self) -> None:
        """Host inputs of mgh_build_coefficients: material table, 1-D widths,
        face bits; the PEC mask on the host only where the CPML needs it."""
        mesh, dims = self.mesh, self.mesh.dims
        lib = mesh.material_library
        tab = np.zeros((max(lib) + 1, 16))
        for m, mat in lib.items():
            tab[m, 0:3] = mat.epsilon
            tab[m, 3:6] = mat.mu
            tab[m, 6:9] = (0.0, 0.0, 0.0) if mat.is_pec else mat.sigma
            tab[m, 9:12] = (0.0, 0.0, 0.0) if mat.is_pec else mat.sigma_m
            tab[m, 12] = 1.0 if mat.is_pec else 0.0
        g = mesh.grid
        geo = np.concatenate([g.dx, g.dy, g.dz, _mesh._avg_d(g.dx), _mesh._avg_d(g.dy), _mesh._avg_d(g.dz)])
        self._port_faces = {P.face for P in self.ports}
        self._dispersion = self._dispersion_mu = self._sibc = None
        self._coeffs = None
        self._m_eps_host = self._m_mu_host = None
        if any(isinstance(bc, CPMLBoundary) for bc in self.boundary_conditions.values()):
            pec = mesh.pec_mask_edges
            for f in self._port_faces:
                pec = _mesh.flatten_port_pec(pec, dims, f)
            self._pec_e = _mesh.pec_flat(pec, dims)
        else:
            self._pec_e = None
        self._host = {"device_coefficients": True, "tab": tab, "geo": geo,
                      "walls": _face_bits(mesh.pec_faces), "ports": _face_bits(self._port_faces)}

    def setup(self) -> None:
        _lib.load()
        self.prepare_host()
        mesh = self.mesh
        dims = self.dims
        dt = self.dt
        t_dt = self._t_dtype
        pec_e = self._pec_e
        dev = torch.device(self.device)
        self._dev = dev
        hst = self._host
        ne, nh = kernels.n_e(*dims), kernels.n_h(*dims)
        if hst.get("device_coefficients"):
            mid = torch.from_numpy(np.ascontiguousarray(mesh.material_id, dtype=np.int32)).to(dev)
            tab = torch.from_numpy(hst["tab"]).to(dev)
            geo = torch.from_numpy(hst["geo"]).to(dev)
            self.alpha_e, self.beta_e, self.m_eps_dev = (torch.empty(ne, dtype=t_dt, device=dev)
                                                         for _ in range(3))
            self.alpha_h, self.beta_h, self.m_mu_dev = (torch.empty(nh, dtype=t_dt, device=dev)
                                                        for _ in range(3))
            sh = int(torch.cuda.current_stream(dev).cuda_stream)
            _lib.call("mgh_build_coefficients", 4 if t_dt == torch.float32 else 8, *dims,
                      mid.data_ptr(), tab.shape[0], tab.data_ptr(), geo.data_ptr(), float(dt),
                      float(_mesh.EPS0), float(_mesh.MU0), hst["walls"], hst["ports"],
                      self.alpha_e.data_ptr(), self.beta_e.data_ptr(), self.m_eps_dev.data_ptr(),
                      self.alpha_h.data_ptr(), self.beta_h.data_ptr(), self.m_mu_dev.data_ptr(), sh)
            torch.cuda.current_stream(dev).synchronize()  # inputs are freed on return
        else:
            def up(key):
                return torch.from_numpy(hst[key]).to(dev, t_dt)

            self.alpha_e, self.beta_e = up("alpha_e"), up("beta_e")
            self.alpha_h, self.beta_h = up("alpha_h"), up("beta_h")
            self.m_eps_dev, self.m_mu_dev = up("m_eps"), up("m_mu")
        self._host = {}  # uploaded: drop the host copies, keep prepare_host done
        K = self.nsets
        self.e = torch.zeros((K, ne), dtype=t_dt, device=dev)
        self.h = torch.zeros((K, nh), dtype=t_dt, device=dev)
        for disp in (self._dispersion, self._dispersion_mu, self._sibc):
            if disp is not None:
                disp.bind(self)

        cpml_faces = []
        for face, bc in self.boundary_conditions.items():
            if isinstance(bc, CPMLBoundary):
                if face in self._port_faces:
                    raise ValueError(f"face {face} carries both a port and a CPML")
                bc.initialize(dt)
                cpml_faces.append(bc)
            elif not isinstance(bc, (PECBoundary, PMCBoundary)):
                raise TypeError(f"unsupported boundary type on {face}: {type(bc).__name__}")
        pec_cell = np.zeros(dims, dtype=bool)
        pec_ids = [m for m, mat in mesh.material_library.items() if mat.is_pec]
        if pec_ids:
            pec_cell = np.isin(mesh.material_id, pec_ids)
        self.cpml = (DeviceCPML(cpml_faces, dims, K, t_dt, dev, pec_e, pec_cell)
                     if cpml_faces else None)
        n_steps = self.horizon
        self.step_ctr = torch.zeros(1, dtype=torch.int64, device=dev)
        self.dports = (DevicePorts(self.ports, dims, K, n_steps, dev, self.excitations, dt, self.record,
                                   self.dtbc_kernel_init)
                       if self.ports else None)
        for src in self.sources:
            src.attach(self)
        for mon in self.monitors:
            mon.attach(self)
        self.dlumped = DeviceLumped(self.lumped, self) if self.lumped else None
        self._sqsum = kernels.SqSum(K, dev)
        self._energy = torch.zeros(K, dtype=torch.float64, device=dev)
        if self.kchunk == "auto":  # timed on the zeroed buffers, before any state is loaded
            self.kchunk = autotune.shape_for(self)
        self.tile_plan = self._build_tile_plan()
        self._setup_done = True

    def _build_tile_plan(self):
        """Live-unit lists of the curl launches (tileskip), or None for dense
        launches: off on request and with TF/SF sources (their H injection
        reaches faces inside metal)."""
        if self.tile_skip not in ("auto", "on", "off"):
            raise ValueError(f"tile_skip must be 'auto', 'on' or 'off', got {self.tile_skip!r}")
        if self.tile_skip == "off" or self.sources:
            return None
        force = self.dlumped.edges if self.dlumped is not None else None
        plan = build_tile_plan(self.alpha_e, self.beta_e, self.alpha_h, self.beta_h, self.dims,
                               kernels.shape_parts(self.kchunk)[0], force_live_e=force)
        if self.tile_skip == "auto" and max(plan.skipped_e, plan.skipped_h) < 0.02:
            return None
        return plan

    # ------------------------------------------------------------------
    def _step(self, n: int) -> None:
        """Launch one leapfrog step (step index n) on the current stream."""
        dims, K = self.dims, self.nsets
        stream = torch.cuda.current_stream(self._dev)
        sh = int(stream.cuda_stream)
        ev = self.timing_events.pop() if self.timing_events else None
        if self._dispersion is not None:
            self._dispersion.save(self, stream)
        if self._dispersion_mu is not None:
            self._dispersion_mu.save(self, stream)
        if self._sibc is not None:
            self._sibc.save(self, stream)
        if ev is not None:
            ev[0].record(stream)
        tp = self.tile_plan
        # CPML rides in the curl launches unless a listed launch or a
        # dispersion completion (which sits between curl and CPML) forbids it
        fused = self.cpml if (self.cpml is not None and tp is None and self._dispersion is None
                              and self._dispersion_mu is None and self._sibc is None) else None
        kernels.update_e(self.e, self.h, self.alpha_e, self.beta_e, dims, curl64=self.curl64,
                         kchunk=self.kchunk, live=tp.live_e if tp is not None else None, cpml=fused,
                         xcols=False, stream=stream)
        if ev is not None:
            ev[1].record(stream)
            self.timed_events.append(ev)
        if fused is not None:
            kernels.cpml_xcols_e(self.e, self.h, self.beta_e, dims, fused, stream=stream)
        # reference E phase order: curl, dispersion completion, CPML faces
        if self._dispersion is not None:
            self._dispersion.update(self, stream)
        if self.cpml is not None and fused is None:
            kernels.cpml_e(self.e, self.h, self.beta_e, dims, self.cpml, stream=stream)
        for src in self.sources:
            src.inject_e(self, stream)
        if self.dports is not None:
            code = 4 if self._t_dtype == torch.float32 else 8
            _lib.call("mgh_port_step", code, K, self.e.data_ptr(), self.e.shape[1], self.h.data_ptr(),
                      self.h.shape[1], self.dports.desc_ptr, len(self.ports), self.step_ctr.data_ptr(), sh)
        if self.dlumped is not None:
            self.dlumped.step(self, stream)
        kernels.update_h(self.e, self.h, self.alpha_h, self.beta_h, dims, kchunk=self.kchunk,
                         live=tp.live_h if tp is not None else None, cpml=fused, xcols=False,
                         stream=stream)
        if fused is not None:
            kernels.cpml_xcols_h(self.e, self.h, self.beta_h, dims, fused, stream=stream)
        # reference H phase order: curl, mu(w) completion, SIBC completion,
        # CPML faces (fit_td.py:891-921)
        if self._dispersion_mu is not None:
            self._dispersion_mu.update(self, stream)
        if self._sibc is not None:
            self._sibc.update(self, stream)
        if self.cpml is not None and fused is None:
            kernels.cpml_h(self.e, self.h, self.beta_h, dims, self.cpml, stream=stream)
        for src in self.sources:
            src.inject_h(self, stream)
        for mon in self.monitors:
            mon.record(self, stream, n)
        _lib.call("mgh_step_advance", self.step_ctr.data_ptr(), 1, sh)

    def energy(self) -> np.ndarray:
        """0.5 (sum M_eps e^2 + sum M_mu h^2) per field set (float64, syncs)."""
        self._sqsum(self.e, self.m_eps_dev, self._energy, scale=0.5)
        self._sqsum(self.h, self.m_mu_dev, self._energy, accumulate=True, scale=0.5)
        return self._energy.cpu().numpy().copy()

    @property
    def horizon(self) -> int:
        """Steps the device buffers (port histories, recorders, source
        tables) are sized for: ``total_time_steps``, or the runtime cap of an
        unbounded run."""
        if self.total_time_steps is not None:
            return int(self.total_time_steps)
        if self.max_time_steps is None:
            raise ValueError("an unbounded run (total_time_steps=None) needs max_time_steps here: "
                             "the port histories and recorders live on the device, sized for it")
        return int(self.max_time_steps)

    @property
    def graph_steps(self) -> int:
        """Steps per captured graph: a multiple of ``graph_block`` and of
        every monitor's launch cadence (a non-pipelined monitor contracts
        once per period, a pipelined one every few steps)."""
        g = max(1, int(self.graph_block))
        return g * self._cadence() // math.gcd(g, self._cadence())

    def _cadence(self) -> int:
        L = 1
        for mon in self.monitors:
            c = int(getattr(mon, "cadence", 1))
            L = L * c // math.gcd(L, c)
        return L

    def request_stop(self) -> None:
        """Ask the marching loop to stop at the next step boundary
        (fit_td.py:1252-1262): a plain flag, safe to set from a signal
        handler; the run ends with ``stop_reason == "aborted"`` on a
        consistent leapfrog state that ``state_dict`` can checkpoint."""
        self._stop_requested = True

    @property
    def stop_reason(self) -> str | None:
        """Why the last run ended: "steps", "energy", "port_signal",
        "port_signal_stall", "runtime_cap" or "aborted"."""
        return self._stop_reason

    def _graph_ok(self, n: int, G: int) -> bool:
        """A G-step graph may run from step n: G is a multiple of the
        monitors' cadence and, once a graph exists, n is congruent to its
        capture step (the host decides per step which contraction launches
        it records, from the step index)."""
        L = self._cadence()
        if G % L:
            return False
        return self._graph is None or self._graph_len != G or (n - self._graph_start) % L == 0

    def _replay(self, n: int, G: int) -> bool:
        """Replay (capturing first if needed) a graph of G steps from step n."""
        if self._graph is not None and self._graph_len != G:
            self._graph = None
        if self._graph is None:
            self._capture(n, G)
        if self._graph is None:
            return False
        self._graph.replay()
        return True

    def run(self, n_steps: int | None = None) -> None:
        """March to ``total_time_steps`` (or ``n_steps`` more), with the
        reference's stop criteria (fit_td.py:682-1246): every
        ``check_interval`` steps the stored energy and the port-signal
        envelope of each field set are compared with that set's peak, and a
        set whose criterion fires is done; the run stops once every set is.
        An unbounded run (``total_time_steps=None``) ends at a criterion or
        at ``max_time_steps`` ("runtime_cap"), and the port-signal stall
        watchdog ends it early when the fitted envelope slope shows the
        threshold unreachable before the cap ("port_signal_stall").  With
        ``use_graphs`` the steps between two checks replay as one graph."""
        if not self._setup_done:
            self.setup()
        unbounded = self.total_time_steps is None
        checks = self.energy_stop_db is not None or self.port_signal_stop_db is not None
        if unbounded and not checks:
            raise ValueError("total_time_steps=None (unbounded run) needs energy_stop_db or "
                             "port_signal_stop_db to provide a stop criterion")
        horizon = self.horizon
        if unbounded and horizon <= self._resume_step:
            raise ValueError(f"max_time_steps={horizon} does not advance past the current step "
                             f"{self._resume_step}")
        if self.port_signal_stop_db is not None and self.dports is None:
            raise ValueError("port_signal_stop_db needs at least one modal port")
        total = horizon if n_steps is None else self._resume_step + int(n_steps)
        if total > horizon:
            raise ValueError("cannot run past total_time_steps (recorder and DTBC histories are sized for it)")
        interval = self.energy_check_interval or (100 if unbounded else max(1, min(100, horizon // 20)))
        K = self.nsets
        self._peak_e = np.broadcast_to(np.asarray(getattr(self, "_peak_e", 0.0), float), (K,)).copy()
        self._peak_s = np.broadcast_to(np.asarray(getattr(self, "_peak_s", 0.0), float), (K,)).copy()
        stall = None
        if self.port_signal_stop_db is not None and unbounded:
            window = max(STALL_MIN_WINDOW_CHECKS, (self.port_signal_min_steps or 0) // (2 * interval))
            stall = [SignalStallDetector(STALL_ARM_DB, window, horizon) for _ in range(K)]
        self._stop_reason = None
        self._stop_requested = False
        self._aborted = False
        G = interval if checks else self.graph_steps
        n = self._resume_step
        while n < total:
            if self._stop_requested:
                break
            # graphs replay whole check intervals: from just after a check
            # (or any step, without checks) to the next one
            if (self.use_graphs and n >= 2 and n + G <= total and self._graph_ok(n, G)
                    and (not checks or (n - 1) % interval == 0) and self._replay(n, G)):
                n += G
            else:
                self._step(n)
                n += 1
            self._resume_step = n
            if checks and (n - 1) % interval == 0:
                reason = self._check(n - 1, stall)
                if reason is not None:
                    self._stop_reason = reason
                    break
        if self._stop_requested and self._stop_reason is None:
            self._aborted = True
            self._stop_reason = "aborted"
        elif self._stop_reason is None and self._resume_step >= horizon:
            self._stop_reason = "runtime_cap" if unbounded else "steps"
            if unbounded:
                import warnings  # noqa: PLC0415

                warnings.warn(f"unbounded run hit the runtime cap of {horizon} steps before any stop "
                              "criterion fired; the recorded signals are truncated at this step",
                              RuntimeWarning, stacklevel=2)
        if self._stop_reason is not None:
            for mon in self.monitors:
                mon.finalize(self)
        torch.cuda.synchronize(self._dev)

    def _check(self, n: int, stall) -> str | None:
        """One stop check after step n (fit_td.py:1014-1153), per field set;
        returns the run's stop reason once every set is done."""
        dbe, dbs = self.energy_stop_db, self.port_signal_stop_db
        en = self.energy()
        self.energy_trace.append((n, en))
        for s in range(self.nsets):
            if self.set_stops[s] is not None:
                continue
            if en[s] > self._peak_e[s]:
                self._peak_e[s] = en[s]
            elif dbe is not None and self._peak_e[s] > 0 and en[s] < self._peak_e[s] * 10 ** (-dbe / 10):
                self.set_stops[s] = ("energy", n + 1)
        if dbs is not None:
            env = np.max(np.stack([a.cpu().numpy() for a in self.dports.absmax]), axis=0)
            for a in self.dports.absmax:
                a.zero_()
            armed = self.port_signal_min_steps is None or n + 1 >= self.port_signal_min_steps
            for s in range(self.nsets):
                if self.set_stops[s] is not None:
                    continue
                if env[s] > self._peak_s[s]:
                    self._peak_s[s] = env[s]
                    if stall is not None:
                        stall[s].reset()
                elif self._peak_s[s] > 0:
                    sig_db = 20.0 * math.log10(max(float(env[s]), 1e-300) / self._peak_s[s])
                    self._final_signal_db = sig_db
                    if armed and env[s] < self._peak_s[s] * 10.0 ** (-dbs / 20.0):
                        self.set_stops[s] = ("port_signal", n + 1)
                    elif armed and stall is not None and stall[s].observe(n, sig_db, -dbs):
                        self.set_stops[s] = ("port_signal_stall", n + 1)
        if all(st is not None for st in self.set_stops):
            return max(self.set_stops, key=lambda st: st[1])[0]
        return None

    def launch(self, n_steps: int) -> int:
        """Enqueue up to ``n_steps`` steps on the current stream and return
        without waiting for them (no stop checks; graph replays where a whole
        graph fits).  Returns the steps enqueued; the caller synchronises.
        A capture (first use of graphs) waits for the device once."""
        if not self._setup_done:
            self.setup()
        total = min(self.horizon, self._resume_step + int(n_steps))
        G = self.graph_steps
        n = self._resume_step
        while n < total:
            if self.use_graphs and n + G <= total and n >= 2 and self._graph_ok(n, G) and self._replay(n, G):
                n += G
                continue
            self._step(n)
            n += 1
        done = n - self._resume_step
        self._resume_step = n
        if n >= self.horizon:
            self._stop_reason = "steps"
            for mon in self.monitors:
                mon.finalize(self)
        return done

    def _capture(self, n: int, G: int) -> None:
        """Capture G steps as one HIP graph; fall back to eager launches on failure.

        Capturing records the launches without executing them, so the step
        counter and every state buffer are untouched; the graph is replayed
        from the same step on.  Device pointers baked into the graph belong
        to buffers this solver keeps alive for its lifetime.
        """
        try:
            # stream-level waits only: other solvers' streams keep running
            cur = torch.cuda.current_stream(self._dev)
            cur.synchronize()
            g = torch.cuda.CUDAGraph()
            s = torch.cuda.Stream(device=self._dev)
            s.wait_stream(cur)
            with torch.cuda.stream(s):
                g.capture_begin()
                for k in range(G):
                    self._step(n + k)
                g.capture_end()
            cur.wait_stream(s)
            s.synchronize()
            self._graph = g
            self._graph_len = G
            self._graph_start = n
        except Exception as exc:  # pragma: no cover - depends on the runtime
            import warnings  # noqa: PLC0415

            self.graph_error = f"{type(exc).__name__}: {exc}"
            warnings.warn(f"graph capture failed ({self.graph_error}); continuing with eager launches")
            self.use_graphs = False
            self._graph = None

    # ------------------------------------------------------------------
    @property
    def n_completed(self) -> int:
        return self._resume_step

    def fields(self, s: int = 0) -> tuple[np.ndarray, np.ndarray]:
        """(e, h) of field set s in the reference layout (host copies)."""
        dims = self.dims
        e = layout.to_reference(self.e[s].cpu().numpy(), *dims, "E")
        h = layout.to_reference(self.h[s].cpu().numpy(), *dims, "H")
        return e, h

    def recorded(self, s: int = 0, n: int | None = None) -> dict:
        """{(port, mode): (V, I)} float64 time series of field set s
        (modal ports, then lumped ports)."""
        n = self._resume_step if n is None else n
        out = {}
        if self.dports is not None:
            for P, V, I in zip(self.ports, self.dports.V_rec, self.dports.I_rec):
                v = V[s, :, :n].cpu().numpy()
                i = I[s, :, :n].cpu().numpy()
                for m in range(P.n_modes):
                    out[(P.name, m)] = (v[m].copy(), i[m].copy())
        if self.dlumped is not None:
            out.update(self.dlumped.recorded(s, n))
        return out

    def state_dict(self, set_index: int | None = None) -> dict:
        """Checkpoint in the reference's schema (fit_td.py:1264-1311).

        One field set: ``n_completed``, ``peak_energy``, ``peak_signal``,
        ``e`` / ``h`` (reference flat layout), ``boundaries`` {face: CPML
        psi as cpml.py:197-204 names them}, ``ports`` {name: V_port_prev,
        V_interior_prev, src_buffer, src_maxlen, dtbc {mode: u, u_prev, n,
        w_hist, s_hist}} (operator.py:993-1017, dtbc.py:365-381),
        ``monitors`` {name: result_dump} and ``dispersion`` {mat_id: J
        per pole} (_dispersion.py:571-576); what only this engine keeps —
        recorded V/I, sample buffers not yet contracted, lumped companion
        states — under ``"hip"``, which the reference does not read.  With
        K > 1 sets and no ``set_index``: {"nsets": K, "sets": [per set]}."""
        torch.cuda.synchronize(self._dev)
        K = self.nsets
        if set_index is None and K > 1:
            return {"nsets": K, "sets": [self.state_dict(s) for s in range(K)]}
        s = 0 if set_index is None else int(set_index)
        if not 0 <= s < K:
            raise ValueError(f"set_index {s} out of range [0, {K})")
        n = self._resume_step
        dims = self.dims
        peak_e = np.broadcast_to(np.asarray(getattr(self, "_peak_e", 0.0), float), (K,))
        peak_s = np.broadcast_to(np.asarray(getattr(self, "_peak_s", 0.0), float), (K,))
        sd = {"n_completed": int(n), "peak_energy": float(peak_e[s]), "peak_signal": float(peak_s[s]),
              "e": layout.to_reference(self.e[s].cpu().numpy(), *dims, "E"),
              "h": layout.to_reference(self.h[s].cpu().numpy(), *dims, "H"),
              "boundaries": self.cpml.reference_state(s) if self.cpml is not None else {},
              "ports": {}, "monitors": {}}
        hip = {"ports": {}, "monitors": {}}
        if self.dports is not None:
            for pi, P in enumerate(self.ports):
                st = self.dports.state[pi][s].cpu().numpy()
                wh, shh = self.dports.w_hist[pi], self.dports.s_hist[pi]
                d = {"V_port_prev": st[:, 2].copy(), "V_interior_prev": st[:, 3].copy(), "dtbc": {}}
                exc = self.excitations[s]
                excited = exc is not None and exc.port == P.name
                buf, maxlen = [], 0
                if excited:
                    m = exc.mode
                    tau = float(P.tau[m])
                    maxlen = max(int(math.ceil((tau + self.dt) / self.dt)) + 3, 5)
                    if P.dtbc_r[m] is None:  # the Mur TF/SF ring buffer: scaled source samples
                        scale = float(P.source_scale[m])
                        buf = [scale * exc.waveform((j + 1) * self.dt) for j in range(max(0, n - maxlen), n)]
                d["src_buffer"] = np.array(buf, dtype=float)
                d["src_maxlen"] = int(maxlen)
                for m in range(P.n_modes):
                    if P.dtbc_r[m] is None:
                        continue
                    hist = wh.shape[2] >= n and n > 0
                    d["dtbc"][str(m)] = {"u": float(st[m, 0]), "u_prev": float(st[m, 1]), "n": int(n),
                                         "w_hist": wh[s, m, :n].cpu().numpy().copy() if hist else np.zeros(n),
                                         "s_hist": shh[s, m, :n].cpu().numpy().copy() if hist else np.zeros(n)}
                if P.comp_r is not None:  # the absorber's states (operator.py:1010-1016)
                    n_u = P.plane.e_u_indices.size
                    cs = (self.dports.comp[pi][s].cpu().numpy() if self.dports.comp[pi] is not None
                          else np.zeros((2, P.comp_r.size)))
                    d["complement"] = {"int_u": cs[0, :n_u].copy(), "int_v": cs[0, n_u:].copy(),
                                       "port_u": cs[1, :n_u].copy(), "port_v": cs[1, n_u:].copy()}
                sd["ports"][P.name] = d
                hip["ports"][P.name] = {"V_rec": self.dports.V_rec[pi][s, :, :n].cpu().numpy().copy(),
                                        "I_rec": self.dports.I_rec[pi][s, :, :n].cpu().numpy().copy(),
                                        "absmax": float(self.dports.absmax[pi][s].item())}
        for i, mon in enumerate(self.monitors):
            name = getattr(mon, "name", "") or f"monitor{i}"
            if hasattr(mon, "result_dump"):
                sd["monitors"][name] = mon.result_dump(s)
            if hasattr(mon, "set_state"):
                hip["monitors"][name] = mon.set_state(s)
        if self.dlumped is not None:
            hip["lumped"] = self.dlumped.set_state(s, n)
        if self._dispersion is not None:
            sd["dispersion"] = self._dispersion.reference_state(s)
        if self._dispersion_mu is not None:
            sd["dispersion_mu"] = self._dispersion_mu.reference_state(s)
        if self._sibc is not None:
            sd["sibc"] = self._sibc.reference_state(s)
        if self.set_stops[s] is not None:
            hip["set_stop"] = {"reason": self.set_stops[s][0], "step": int(self.set_stops[s][1])}
        sd["hip"] = hip
        return sd

    def load_state_dict(self, sd: dict, set_index: int | None = None) -> None:
        """Restore a checkpoint of :meth:`state_dict` — or one the reference's
        solver wrote (fit_td.py:1313-1353): a reference checkpoint has no
        ``"hip"`` key, so recorded V/I before it start at zero and DFT
        monitors take the bins of its ``monitors`` dumps, when present.  The
        solver must be built as the one that wrote it; call before run()."""
        if not self._setup_done:
            self.setup()
        if "sets" in sd:
            if len(sd["sets"]) != self.nsets:
                raise ValueError(f"checkpoint holds {len(sd['sets'])} field sets, the solver {self.nsets}")
            for s, sub in enumerate(sd["sets"]):
                self._load_set(s, sub)
            n = int(sd["sets"][0]["n_completed"])
        else:
            if set_index is None and self.nsets > 1:
                raise ValueError("a one-set checkpoint needs set_index on a solver with several sets")
            self._load_set(0 if set_index is None else int(set_index), sd)
            n = int(sd["n_completed"])
        if self.tile_plan is not None:
            # dead E edges: zero, as the dense launch leaves them at its next E update
            self.e.masked_fill_(self.tile_plan.zero_e.unsqueeze(0), 0)
        self.step_ctr.fill_(n)
        self._resume_step = n
        self._graph = None
        torch.cuda.synchronize(self._dev)

    def _load_set(self, s: int, sd: dict) -> None:
        dims = self.dims
        n = int(sd["n_completed"])
        if n > self.horizon:
            raise ValueError(f"checkpoint at step {n} lies past this solver's {self.horizon} steps")
        K = self.nsets
        self._peak_e = np.broadcast_to(np.asarray(getattr(self, "_peak_e", 0.0), float), (K,)).copy()
        self._peak_s = np.broadcast_to(np.asarray(getattr(self, "_peak_s", 0.0), float), (K,)).copy()
        self._peak_e[s] = float(sd.get("peak_energy", 0.0))
        self._peak_s[s] = float(sd.get("peak_signal", 0.0))
        self.e[s].copy_(torch.from_numpy(layout.to_device(np.asarray(sd["e"]), *dims, "E")).to(self._dev, self._t_dtype))
        self.h[s].copy_(torch.from_numpy(layout.to_device(np.asarray(sd["h"]), *dims, "H")).to(self._dev, self._t_dtype))
        if self.cpml is not None:
            self.cpml.load_reference_state(s, sd["boundaries"])
        hip = sd.get("hip", {})
        if self.dports is not None:
            for pi, P in enumerate(self.ports):
                p = sd["ports"][P.name]
                comp = p.get("complement")
                if comp is not None:
                    parts = [np.asarray(comp[k], float) for k in ("int_u", "int_v", "port_u", "port_v")]
                    if self.dports.comp[pi] is not None:
                        cs = np.stack([np.concatenate(parts[:2]), np.concatenate(parts[2:])])
                        self.dports.comp[pi][s].copy_(torch.from_numpy(cs))
                    elif any(np.any(a != 0) for a in parts):
                        raise ValueError(f"port {P.name!r}: the checkpoint carries complement-absorber "
                                         "state but the port's absorber is not live (every channel on "
                                         "the DTBC, or no absorber configured)")
                st = self.dports.state[pi]
                stn = st[s].cpu().numpy()
                stn[:, 2] = np.asarray(p["V_port_prev"], float)
                stn[:, 3] = np.asarray(p["V_interior_prev"], float)
                stn[:, 0:2] = 0.0
                wh, shh = self.dports.w_hist[pi], self.dports.s_hist[pi]
                for m_str, t in p.get("dtbc", {}).items():
                    m = int(m_str)
                    if P.dtbc_r[m] is None:
                        raise ValueError(f"checkpoint carries DTBC state for mode {m} but port "
                                         f"{P.name!r} terminates it on Mur")
                    if int(t["n"]) != n:
                        raise ValueError(f"port {P.name!r} mode {m}: DTBC at step {t['n']}, fields at {n}")
                    stn[m, 0], stn[m, 1] = float(t["u"]), float(t["u_prev"])
                    if n:
                        wh[s, m, :n].copy_(torch.from_numpy(np.asarray(t["w_hist"], float)))
                        shh[s, m, :n].copy_(torch.from_numpy(np.asarray(t["s_hist"], float)))
                st[s].copy_(torch.from_numpy(stn))
                hp = hip.get("ports", {}).get(P.name)
                V, I = self.dports.V_rec[pi], self.dports.I_rec[pi]
                if hp is not None and n and V.shape[2] >= n:
                    V[s, :, :n].copy_(torch.from_numpy(np.asarray(hp["V_rec"])))
                    I[s, :, :n].copy_(torch.from_numpy(np.asarray(hp["I_rec"])))
                if hp is not None:
                    self.dports.absmax[pi][s] = float(hp.get("absmax", 0.0))
        for i, mon in enumerate(self.monitors):
            name = getattr(mon, "name", "") or f"monitor{i}"
            if name in hip.get("monitors", {}) and hasattr(mon, "load_set_state"):
                mon.load_set_state(s, hip["monitors"][name])
            elif name in sd.get("monitors", {}) and hasattr(mon, "load_result_dump"):
                mon.load_result_dump(sd["monitors"][name], s)
        if self.dlumped is not None and "lumped" in hip:
            self.dlumped.load_set_state(s, hip["lumped"], n)
        if self._dispersion is not None and "dispersion" in sd:
            self._dispersion.load_reference_state(s, sd["dispersion"])
        if self._dispersion_mu is not None and "dispersion_mu" in sd:
            self._dispersion_mu.load_reference_state(s, sd["dispersion_mu"])
        if self._sibc is not None and "sibc" in sd:
            self._sibc.load_reference_state(s, sd["sibc"])
        stop = hip.get("set_stop")
        self.set_stops[s] = (stop["reason"], int(stop["step"])) if stop else None
