"""The free-fall workload on the 500-species synthetic network at its stated shape: all 8192
points of the seeded collapse grid (n_H,0 log-uniform 1e2..1e6 cm^-3 towards 1e6 cm^-3) to
1 Myr in one device launch with the FP32 factorisation and FP64 residuals.

Gates: every point status 0 with its element totals conserved to 1e-10 relative; the 16
points of tests/golden/ref_collapse500.npz (one per initial-density stratum) within 1% of the
reference Fortran compiled on the same network (oracle/_ref/net_synthetic_500_6000, whose
network.f90/odes.f90 uclchem_amd/reference_format.py writes) on every abundance above 1e-14;
a stratified 32-point sample within 1% of the FP64 factorisation, and 8 of those within 1% of
the dense LU.
"""
import os

import numpy as np
import pytest

from helpers import conserved_elements, element_totals

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
N_POINTS = 8192


@pytest.fixture(scope="module")
def grid500():
    import torch
    from uclchem_amd import gpu, synthetic as sy, workloads
    t = sy.load_or_build("synthetic_500_6000")
    net = gpu.Network(t)
    params = workloads.collapse_grid(N_POINTS)
    dev = torch.device("cuda:0")
    d = dict(params=torch.tensor(params, device=dev),
             y_out=torch.zeros((N_POINTS, t.neq), dtype=torch.float64, device=dev),
             status=torch.zeros(N_POINTS, dtype=torch.int32, device=dev),
             stats=torch.zeros((N_POINTS, 8), dtype=torch.int32, device=dev))
    stream = torch.cuda.current_stream(dev)
    print(f"[config3] launching {N_POINTS} points, FP32 factorisation", flush=True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    net.solve_device(d["params"], d["y_out"], d["status"], d["stats"], lu_bits=32, stream=stream.cuda_stream,
                     order="densest")
    ev1.record(stream)
    torch.cuda.synchronize()
    print(f"[config3] done in {ev0.elapsed_time(ev1) / 1e3:.1f} s", flush=True)
    return t, net, params, d["y_out"].cpu().numpy(), d["status"].cpu().numpy(), d["stats"].cpu().numpy()


def _rel(a, b):
    m = b > 1e-14
    return np.abs(a[m] - b[m]) / b[m]


def test_every_point_finishes_and_conserves(grid500, oracle_of):
    t, net, params, y, status, stats = grid500
    assert np.all(status == 0), np.unique(status, return_counts=True)
    orc = oracle_of(t)
    _, y0 = orc.init_point(params[0])        # the elemental start: same abundances for every point
    e0 = element_totals(t, y0)
    keep = [i for i in conserved_elements(t) if e0[i] > 0]
    e1 = np.array([element_totals(t, y[q]) for q in range(N_POINTS)])[:, keep]
    dev = np.abs(e1 - e0[keep]) / e0[keep]
    assert dev.max() < 1e-10, (int(np.argmax(dev.max(axis=1))), dev.max())


def test_reference_points_match(grid500):
    t, net, params, y, status, stats = grid500
    path = os.path.join(GOLD, "ref_collapse500.npz")
    if not os.path.exists(path):
        pytest.skip("ref_collapse500.npz not generated")
    f = np.load(path)
    from uclchem_amd.params import widen
    assert np.array_equal(params[f["index"]], widen(f["params"]))
    bad = []
    for j, q in enumerate(f["index"]):
        assert f["rc"][j] == 0
        r = _rel(y[q, :t.nspec], f["final"][j, 4:])
        if r.max() > 0.01:
            bad.append((int(q), float(r.max())))
    assert not bad, bad


def test_fp64_and_dense_agree_on_a_sample(grid500):
    t, net, params, y, status, stats = grid500
    from uclchem_amd.params import INDEX
    order = np.argsort(params[:, INDEX["initialdens"]], kind="stable")
    sample = np.array([s[len(s) // 2] for s in np.array_split(order, 32)])
    print(f"[config3] FP64 sample of {len(sample)} points", flush=True)
    f64 = net.solve(params[sample], lu_bits=64, order="densest")
    assert np.all(f64["status"] == 0)
    for j, q in enumerate(sample):
        assert _rel(y[q, :t.nspec], f64["y"][j, :t.nspec]).max() < 0.01, int(q)
    sub = sample[::4]
    print(f"[config3] dense LU on {len(sub)} points", flush=True)
    net.set_linear_solver(net.DENSE)
    try:
        dense = net.solve(params[sub], lu_bits=64)
    finally:
        net.set_linear_solver(net.SPARSE)
    assert np.all(dense["status"] == 0)
    for j, q in enumerate(sub):
        assert _rel(dense["y"][j, :t.nspec], f64["y"][4 * j, :t.nspec]).max() < 0.01, int(q)
