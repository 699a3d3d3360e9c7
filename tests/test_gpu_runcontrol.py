"""GPU tests of run control on the plain WR-90 case (C1, float64).

* energy and port-signal stop criteria end the HIP run at the step where the
  oracle's restatement of the reference's check loop (fit_td.py:1008-1153)
  stops: energy 10 dB -> step 601, port signal 40 dB -> step 1501;
* with graphs the steps between two checks replay as one graph: same stop
  step, fields identical to the eager run bit for bit;
* an unbounded run (total_time_steps=None) with a 60 dB port-signal
  criterion, which the cut-off ringing never reaches, ends through the
  stall watchdog or at the runtime cap exactly where the oracle's envelope
  fed to the same watchdog says;
* request_stop from inside the run ends it on a consistent state that
  resumes bit-exactly.
"""

import numpy as np
import pytest
import torch

import cases
from magnelio_amd.solver import STALL_ARM_DB, STALL_MIN_WINDOW_CHECKS, SignalStallDetector

pytestmark = pytest.mark.gpu


def _oracle_stop(**kw):
    orc, _, _, (me, mm) = cases.c1_oracle(np.float64)
    return orc.run_with_stops(4000, me, mm, 100, **kw)


def _solver(use_graphs=False, **kw):
    sol, _, _ = cases.c1_solver(n_steps=kw.pop("n_steps", 4000), precision="double",
                                use_graphs=use_graphs)
    for k, v in kw.items():
        setattr(sol, k, v)
    return sol


@pytest.mark.parametrize("crit", [{"energy_stop_db": 10.0}, {"port_signal_stop_db": 40.0}])
def test_stop_criteria_match_oracle(crit, gpu):
    print(f"oracle C1 float64 with stop checks {crit}", flush=True)
    reason, n_stop, _ = _oracle_stop(**crit)
    assert reason != "steps"
    runs = {}
    for graphs in (False, True):
        print(f"HIP C1 float64 {crit} graphs={graphs}", flush=True)
        sol = _solver(use_graphs=graphs, **crit)
        sol.run()
        assert sol.stop_reason == reason and sol.n_completed == n_stop, (sol.stop_reason, sol.n_completed)
        if graphs:
            assert sol._graph is not None and sol._graph_len == 100
        runs[graphs] = sol
    assert torch.equal(runs[False].e, runs[True].e) and torch.equal(runs[False].h, runs[True].h)


def test_unbounded_run_stall_or_cap_matches_oracle(gpu):
    cap = 2500
    print("oracle C1 float64 envelope every 100 steps", flush=True)
    orc, _, _, _ = cases.c1_oracle(np.float64)
    det = SignalStallDetector(STALL_ARM_DB, STALL_MIN_WINDOW_CHECKS, cap)
    peak, want = 0.0, ("runtime_cap", cap)
    while orc.n < cap:
        n = orc.n
        orc.step()
        if n % 100:
            continue
        sig = max(p.poll_signal_absmax() for p in orc.ports)
        if sig > peak:
            peak = sig
            det.reset()
        elif peak > 0:
            db = 20 * np.log10(max(sig, 1e-300) / peak)
            assert sig >= peak * 1e-3  # 60 dB never reached here
            if det.observe(n, db, -60.0):
                want = ("port_signal_stall", n + 1)
                break
    print(f"expected {want}", flush=True)
    sol = _solver(n_steps=None, max_time_steps=cap, port_signal_stop_db=60.0)
    sol.total_time_steps = None
    with pytest.warns(RuntimeWarning) if want[0] == "runtime_cap" else _nothing():
        sol.run()
    assert (sol.stop_reason, sol.n_completed) == want


class _nothing:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


class _StopAt:
    """A monitor that asks the solver to stop after step `n`."""

    pipelined = True

    def __init__(self, n):
        self.n = n

    def attach(self, solver):
        pass

    def record(self, solver, stream, n):
        if n == self.n:
            solver.request_stop()

    def finalize(self, solver):
        pass


def test_request_stop_then_resume_is_bit_exact(gpu):
    full, _, _ = cases.c1_solver(n_steps=400, precision="double")
    full.run()
    part, _, _ = cases.c1_solver(n_steps=400, precision="double", monitors=[_StopAt(150)])
    part.run()
    assert part.stop_reason == "aborted" and part._aborted and part.n_completed == 151
    sd = part.state_dict()
    fresh, _, _ = cases.c1_solver(n_steps=400, precision="double")
    fresh.load_state_dict(sd)
    fresh.run()
    assert fresh.stop_reason == "steps"
    assert torch.equal(fresh.e, full.e) and torch.equal(fresh.h, full.h)
