"""Oracle: the FIT-TD leapfrog loop in the reference's step order (CPU).

Follows src/magnelio/solver/fit_td.py:950-1007 for the parts this build
covers, with t = n dt:

  stash    ADE f_prev of the eps(w) edges and mu(w) faces,
           SIBC h_prev of the wall faces                     :826-841
  E phase  curl update (stencil or Numba order), eps(w) completion,
           CPML E corrections in boundary order              :837-889
  sources  TF/SF E injection at t + dt                       :965-967
  ports    update_e(fields, t + dt) in port order            :975-976
           (modal ports, then lumped ports / elements)
  record   V/I of every port (e at t^{n+1}, h at t^{n+1/2})  :981-982
  H phase  curl update, mu(w) completion, SIBC completion,
           CPML H corrections                                :891-927
  sources  TF/SF H injection at t + 3dt/2                    :991-993
  monitors record(fields, n, t, dt)                          :996-1002
"""

from __future__ import annotations

import numpy as np

from . import yee


class Oracle:
    def __init__(self, dims, alpha_e, beta_e, alpha_h, beta_h, dt, *, cpml=(), ports=(),
                 sources=(), monitors=(), lumped=(), dtype=np.float64, order="stencil",
                 dispersion=None, dispersion_mu=None, sibc=None):
        self.dims = dims
        self.dt = dt
        self.dtype = dtype
        self.ae, self.be = alpha_e.astype(dtype), beta_e.astype(dtype)
        self.ah, self.bh = alpha_h.astype(dtype), beta_h.astype(dtype)
        self.e = np.zeros(yee.n_e(*dims), dtype)
        self.h = np.zeros(yee.n_h(*dims), dtype)
        self.cpml = list(cpml)
        self.ports = list(ports)
        self.sources = list(sources)
        self.monitors = list(monitors)
        self.lumped = list(lumped)
        self.order = order
        self.dispersion = dispersion
        self.dispersion_mu = dispersion_mu
        self.sibc = sibc
        self.n = 0

    def step(self) -> None:
        n, dt = self.n, self.dt
        t = n * dt
        dims = self.dims
        if self.dispersion is not None:
            self.dispersion.save_field(self.e)
        if self.dispersion_mu is not None:
            self.dispersion_mu.save_field(self.h)
        if self.sibc is not None:
            self.sibc.save_field(self.h)
        yee.update_e(self.e, self.h, self.ae, self.be, *dims, order=self.order)
        if self.dispersion is not None:
            self.dispersion.update_field(self.e)
        for f in self.cpml:
            f.update_e(self.e, self.h, self.be)
        for s in self.sources:
            s.inject_e(self.e, t + dt)
        for p in self.ports:
            p.update_e(self.e, t + dt)
        for el in self.lumped:
            el.update_e(self.e, t + dt)
        for p in self.ports:
            p.record(self.e, self.h)
        for el in self.lumped:
            if getattr(el, "recorded", False):
                el.record(self.e, self.h)
        yee.update_h(self.e, self.h, self.ah, self.bh, *dims)
        if self.dispersion_mu is not None:
            self.dispersion_mu.update_field(self.h)
        if self.sibc is not None:
            self.sibc.update_field(self.h)
        for f in self.cpml:
            f.update_h(self.e, self.h, self.bh)
        for s in self.sources:
            s.inject_h(self.h, t + dt + dt / 2)
        for m in self.monitors:
            m.record(self.e, self.h, n, t, dt)
        self.n = n + 1

    def run(self, n_steps: int) -> None:
        for _ in range(n_steps):
            self.step()

    def run_with_stops(self, n_steps: int, m_eps, m_mu, interval: int, energy_stop_db=None,
                       port_signal_stop_db=None, port_signal_min_steps=None):
        """March with the reference's stop checks (fit_td.py:1008-1153, the
        bounded-run branch): after step n with n % interval == 0 the stored
        energy 0.5 (sum M_eps e^2 + sum M_mu h^2) (float64 sums) and then the
        port |V| envelope polled since the previous check are compared with
        their running peaks.  Returns (reason, completed steps, energies)."""
        peak_e = peak_s = 0.0
        trace = []
        while self.n < n_steps:
            n = self.n
            self.step()
            if n % interval:
                continue
            e64, h64 = self.e.astype(np.float64), self.h.astype(np.float64)
            en = 0.5 * (float(np.sum(m_eps * e64 * e64)) + float(np.sum(m_mu * h64 * h64)))
            trace.append((n, en))
            if en > peak_e:
                peak_e = en
            elif peak_e > 0.0 and energy_stop_db is not None and en < peak_e * 10 ** (-energy_stop_db / 10):
                return "energy", n + 1, trace
            if port_signal_stop_db is not None:
                sig = max(p.poll_signal_absmax() for p in self.ports)
                if sig > peak_s:
                    peak_s = sig
                elif peak_s > 0.0:
                    armed = port_signal_min_steps is None or n + 1 >= port_signal_min_steps
                    if armed and sig < peak_s * 10.0 ** (-port_signal_stop_db / 20.0):
                        return "port_signal", n + 1, trace
        return "steps", self.n, trace
