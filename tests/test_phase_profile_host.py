"""The per-phase profile tool (python -m uclchem_amd.phase_profile) on the CPU with a
stand-in network: per-step and per-call times are the timer sums divided by the matching
solver counters (steps, RHS evaluations, Newton-matrix setups, Newton iterations)."""
import sys

import numpy as np


def test_per_call_arithmetic(monkeypatch, capsys):
    from uclchem_amd import gpu, phase_profile
    from uclchem_amd import makerates_hip as mh

    class FakeNet:
        def __init__(self, tables):
            pass

        def solve_profiled(self, grid, lu_bits=64, n_slots=0):
            n = len(grid)
            stats = np.zeros((n, 8), dtype=np.int64)
            stats[:, 0], stats[:, 1], stats[:, 2], stats[:, 3] = 100, 150, 30, 140     # steps rhs setups iters
            timers = np.zeros((n, 16))
            timers[:, 1] = 150 * 26e-6      # RHS seconds
            timers[:, 2] = 30 * 155e-6      # Jacobian
            timers[:, 3] = 30 * 236e-6      # factorisation
            timers[:, 4] = 140 * 46e-6      # solves
            timers[:, 5] = 100 * 230e-6     # total
            timers[:, 6] = 2400.0           # shader MHz
            return dict(timers=timers, stats=stats, y=None, status=np.zeros(n, np.int32))

    monkeypatch.setattr(gpu, "Network", FakeNet)
    monkeypatch.setattr(mh, "load_network", lambda path: None)
    monkeypatch.setattr(sys, "argv", ["phase_profile", "--points", "4"])
    phase_profile.main()
    out = capsys.readouterr().out
    assert "ms/step 0.230" in out
    assert "rhs 26.0" in out and "jacobian 155.0" in out and "factorisation 236.0" in out and "solve 46.0" in out
    assert out.count("lu64") == 1 and out.count("lu32") == 1
