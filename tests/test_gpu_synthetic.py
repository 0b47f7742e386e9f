"""GPU parity on the seeded synthetic networks (150/1500 and 500/6000).

The 500-species network's static plan has a 252-wide dense tail, too large to stage in
LDS, so it runs the unstaged plan (factors and solve program read from the workspace);
the 150-species network runs the staged plan.  Gates:
* rates and RHS against the oracle (C restatement, CPU) as on the default network:
  rates <= 1e-12 relative, RHS <= 1e-12 of each species' term magnitude;
* full free-fall solves: status 0, elements conserved to 1e-10 relative;
* the sparse plan (staged or not) against the dense LU, and the FP32 factorisation
  against FP64: every abundance above 1e-14 within 1%.
The trajectories on these networks are pinned against the reference itself elsewhere: the
reference compiled on the same networks (their network.f90/odes.f90 written by
uclchem_amd/reference_format.py) gives the fixtures of test_gpu_config3 (500/6000 free-fall
grid) and test_gpu_serving (requests over both networks).
"""
import numpy as np
import pytest

from helpers import conserved_elements, element_totals, phys_of, random_state, term_magnitude

pytestmark = pytest.mark.gpu

NETS = ["synthetic_150_1500", "synthetic_500_6000"]


@pytest.fixture(scope="module", params=NETS)
def syn(request):
    from oracle import Oracle
    from uclchem_amd import gpu
    from uclchem_amd import synthetic as sy
    t = sy.load_or_build(request.param)
    print(f"[synthetic] {request.param}: loading", flush=True)
    return t, gpu.Network(t), Oracle(t)


def _states(tables, oracle, seed, n):
    from uclchem_amd.params import make_params
    rng = np.random.default_rng(seed)
    P, PH, Y, pts = [], [], [], []
    for trial in range(n):
        p = make_params(initialdens=10 ** rng.uniform(2, 8), initialtemp=rng.uniform(10, 300),
                        zeta=10 ** rng.uniform(0, 2), radfield=10 ** rng.uniform(0, 3), baseav=rng.uniform(0, 10),
                        finaltime=1e6)
        pt, y = random_state(tables, oracle, rng, p, bare_grains=(trial % 3 == 0))
        P.append(p)
        PH.append(phys_of(pt))
        Y.append(y)
        pts.append(pt)
    return np.array(P), np.array(PH), np.array(Y), pts


def test_rates_and_rhs_match_oracle(syn):
    t, net, orc = syn
    P, PH, Y, pts = _states(t, orc, 4, 6)
    rg, yg, _ = net.rates(P, PH, Y)
    R, Yr, YD, MAG = [], [], [], []
    for q in range(len(P)):
        ro, yo, _ = orc.rates(pts[q], Y[q])
        d = np.abs(rg[q] - ro) / np.maximum(np.abs(ro), 1e-300)
        d[(rg[q] == 0) & (ro == 0)] = 0
        assert d.max() <= 1e-12, (q, int(np.argmax(d)))
        yd, r2 = orc.rhs(pts[q], ro, yo)
        R.append(ro)
        Yr.append(yo)
        YD.append(yd)
        MAG.append(term_magnitude(t, r2, yo))
    ydg, _ = net.rhs(P, PH, np.array(R), np.array(Yr))
    assert np.all(np.abs(ydg - np.array(YD)) <= 1e-12 * np.array(MAG))


def _rel(a, b):
    m = b > 1e-14
    return np.abs(a[m] - b[m]) / b[m]


def test_collapse_solves_conserve_and_agree(syn):
    t, net, _ = syn
    from uclchem_amd import workloads
    params = workloads.collapse_grid(4, final_time=1e5)
    base = net.solve(params, max_rec=80, lu_bits=64)
    print(f"[synthetic] fp64 sparse steps {base['stats'][:, 0].tolist()}", flush=True)
    assert np.all(base["status"] == 0), base["status"]
    keep = [i for i in conserved_elements(t)]
    for q in range(len(params)):
        y0 = np.append(base["rec"][q, 0, 4:], 0.0)
        e0, e1 = element_totals(t, y0), element_totals(t, base["y"][q])
        k = [i for i in keep if e0[i] > 0]
        assert (np.abs(e1[k] - e0[k]) / e0[k]).max() < 1e-10
    f32 = net.solve(params, lu_bits=32)
    assert np.all(f32["status"] == 0)
    net.set_linear_solver(net.DENSE)
    try:
        dense = net.solve(params, lu_bits=64)
    finally:
        net.set_linear_solver(net.SPARSE)
    assert np.all(dense["status"] == 0)
    for q in range(len(params)):
        yb = base["y"][q, :t.nspec]
        assert _rel(f32["y"][q, :t.nspec], yb).max() < 0.01
        assert _rel(dense["y"][q, :t.nspec], yb).max() < 0.01
