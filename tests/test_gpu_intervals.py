"""GPU parity of single output intervals (ucg_interval, one updateChemistry call of
src/fortran_src/chemistry.f90:145-239) against the reference Fortran on the same states.

The fixture (tests/golden/ref_intervals.npz, written by make_interval_fixtures.py from
oracle/_ref) holds states of the reference's own trajectories: hot-core warm-up points
(temp_indx 4, n_H ~1e5..1e8) including their desorption temperatures, and the hard late
intervals of the 4096-point static-cloud grid's tail points (594, 380, 1203, 2549), where the
reference needs up to ~1e5 steps.  Gates: the interval's end is reached, every abundance above
1e-14 within 1% of the reference, element totals conserved to 1e-10 over the interval.

On the "first-attempt" rows (stretches where the reference's DVODE spends its whole step budget)
the reference's smallest abundances move by tens of percent when the stretch is integrated at
tighter tolerances (y_tight: the reference itself at 100x tighter tolerances, or, where that
does not finish, the host BDF core on the reference's own F at the same; tight_src 1 or
2); there an abundance passes within 1% of the reference plus twice the reference's own
integration error, |y_ref - y_tight|.  Where the reference or the GPU needs more than one
integrator call for a row, the rates are recomputed mid-interval at a place set by the step
count, and the row is not a call-for-call comparison (test_intervals_match_reference).
"""
import os

import numpy as np
import pytest

from helpers import conserved_elements, element_totals

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture():
    path = os.path.join(GOLD, "ref_intervals.npz")
    if not os.path.exists(path):
        pytest.skip("ref_intervals.npz not generated")
    return np.load(path)


# Rows on which the comparison is not call for call (see the module docstring), at most: FP64:
# the two storm stretches on which the GPU's own call runs out of the reference's step budget
# (hot-core point 108 at 7432 yr, tail point 594 at 1.13e5 yr); FP32: those and the storm
# stretches on which the FP32 factors do not reach 1% (the FP32 form is validated against FP64
# on whole models, where 99.6% of the grid's points agree within 1%).
MAX_UNGATED = {64: 2, 32: 4}


@pytest.mark.parametrize("lu_bits", [64, 32])
def test_intervals_match_reference(gpu_net, tables, lu_bits):
    """Each row is one interval of the reference's own trajectory, compared call for call: a
    reference row made of one integrator call (rates frozen at its start) against the GPU's
    interval from the same state to the same end.  Gate: status 0, the end reached, element
    totals conserved to 1e-10, and every abundance above 1e-14 within 1% of the reference plus
    twice its own integration error |y_ref - y_tight| where the fixture has y_tight.

    Not call for call, so reported and counted but not held to that gate:
    * rows the reference itself ends in several calls (its first runs out of MXSTEP and the
      rates are recomputed where it stopped, a place set by its step count); their single
      calls are rows of their own ("<how> call 1", "<how> call 2") and are gated;
    * rows on which the GPU's call runs out of MXSTEP (it needs more steps than the
      reference's budget, so it recomputes the rates at another time);
    * in FP32, storm stretches ("first-attempt" rows) beyond the gate.
    At most MAX_UNGATED[lu_bits] rows of the last two kinds."""
    f = _fixture()
    n = len(f["target"])
    how = [str(h) for h in f["how"]]
    for q in range(n):
        print(f"row {q}: {f['kind'][q]} point {f['idx'][q]} {how[q]} T={f['phys_in'][q, 1]:.1f} K "
              f"t={f['phys_in'][q, 7] / 3.16e7:.4g} -> {f['target'][q] / 3.16e7:.4g} yr "
              f"(reference: {f['steps'][q]} steps, {f['attempts'][q]} attempts)", flush=True)
    from uclchem_amd.params import widen
    out = gpu_net.interval(widen(f["params"]), f["phys_in"], f["y_in"], f["target"], lu_bits=lu_bits)
    keep = [i for i in conserved_elements(tables)]
    bad, ungated, multi = [], [], []
    for q in range(n):
        assert out["status"][q] == 0, (q, out["status"][q])
        assert out["phys"][q, 7] >= f["target"][q] * (1 - 1e-12), (q, out["phys"][q, 7], f["target"][q])
        e0, e1 = element_totals(tables, f["y_in"][q]), element_totals(tables, out["y"][q])
        k = [i for i in keep if e0[i] > 0]
        cons = float((np.abs(e1[k] - e0[k]) / e0[k]).max())
        assert cons <= 1e-10, (q, cons)
        ref = f["y_ref"][q, :tables.nspec]
        m = ref > 1e-14
        allow = 0.01 * ref[m]
        if np.all(np.isfinite(f["y_tight"][q])):
            allow = allow + 2.0 * np.abs(ref[m] - f["y_tight"][q, :tables.nspec][m])
        excess = float((np.abs(out["y"][q, :tables.nspec][m] - ref[m]) / allow).max())   # <= 1 passes
        if f["attempts"][q] > 1 and how[q] != "first-attempt":
            multi.append((q, how[q], round(excess, 3)))
            continue
        gpu_calls = 1 + int(out["stats"][q, 7])
        if excess > 1.0:
            if gpu_calls > 1 or (lu_bits == 32 and how[q] == "first-attempt"):
                ungated.append((q, how[q], gpu_calls, round(excess, 3)))
                continue
            bad.append((q, how[q], excess))
        # phi and abstol_factor leave the interval as the reference's do
        assert out["phys"][q, 5] == pytest.approx(f["phys_ref"][q, 5], rel=1e-12)
        assert out["phys"][q, 6] == f["phys_ref"][q, 6] or how[q] == "first-attempt" or gpu_calls > 1
    print(f"[intervals] reference multi-call rows (q, how, excess): {multi}", flush=True)
    print(f"[intervals] not call for call (q, how, GPU calls, excess): {ungated}", flush=True)
    assert not bad, bad
    assert len(ungated) <= MAX_UNGATED[lu_bits], ungated
    # every multi-call row of the reference has its single calls in the fixture
    for q, h, _ in multi:
        assert any(how[r] == h + " call 1" and f["idx"][r] == f["idx"][q] for r in range(n)), (q, h)


def test_interval_continues_the_full_solve(gpu_net, tables):
    """Chaining ucg_interval with the loop's physics update reproduces ucg_solve_batch's
    own intervals (cloud schedule to 1e3 yr): the same kernel path either way."""
    from uclchem_amd import workloads
    from uclchem_amd.params import INDEX
    p = workloads.static_cloud_grid(4)
    p[:, INDEX["finaltime"]] = 1e3
    full = gpu_net.solve(p, max_rec=40)
    # restart every point from its output row k (time t_k) and integrate to row k+1
    for k in (3, 8):
        rows = full["rec"][:, k]
        nxt = full["rec"][:, k + 1]
        y = np.concatenate([rows[:, 4:], rows[:, 1:2]], axis=1)
        phys = np.zeros((len(p), 8))
        phys[:, 0] = rows[:, 1]
        phys[:, 1] = phys[:, 2] = rows[:, 2]
        phys[:, 3] = rows[:, 3]
        phys[:, 4] = (p[:, INDEX["rout"]] - p[:, INDEX["rin"]]) * 3.086e18
        # phi as the previous interval's rates left it (rates.f90:321-325, from that interval's
        # start state: row k-1; every interval here ends in one attempt)
        prev = full["rec"][:, k - 1]
        ne = prev[:, 4 + tables.scalars["nelec"]]
        phi = p[:, INDEX["radfield"]] * np.exp(-2.5 * prev[:, 3]) * np.sqrt(prev[:, 2]) / (prev[:, 1] * ne)
        phys[:, 5] = np.clip(phi, 1e2, 1e6)
        phys[:, 6] = p[:, INDEX["abstol_factor"]]
        phys[:, 7] = rows[:, 0] * 3.16e7
        out = gpu_net.interval(p, phys, y, nxt[:, 0] * 3.16e7)
        assert np.all(out["status"] == 0)
        m = nxt[:, 4:] > 1e-14
        rel = np.abs(out["y"][:, :tables.nspec][m] - nxt[:, 4:][m]) / nxt[:, 4:][m]
        assert rel.max() < 1e-6, rel.max()
