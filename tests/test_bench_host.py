"""Host-side arithmetic of bench.py's roofline figures (no GPU): the algorithmic FLOP counts
per Newton-matrix setup, per Newton iteration and per RHS evaluation that turn the solver
counters into achieved FLOP/s."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def test_flop_counts_follow_the_plan():
    import bench
    from uclchem_amd import makerates_hip as mh
    from uclchem_amd import newton_plan as npl
    t = mh.load_network(mh.network_path("default"))
    s = npl.plan_summary(t.arrays)
    lu, solve = bench.newton_flop_counts(t)
    ndiv = int((t.arrays["nw_tg_div"] >= 0).sum())
    assert npl.plan_is_staged(t)                     # the default network's plan fits LDS: Gauss-Jordan tail
    assert lu == 2 * s["npair"] + ndiv + 2 * s["T"] ** 3
    assert solve == 2 * s["nsp"] + 2 * s["T"] ** 2
    assert (s["T"], s["npair"]) == (104, 54116)
    assert abs(lu - 2.36e6) < 0.01e6 and abs(solve - 3.37e4) < 0.01e4
    rhs = bench.rhs_flop_count(t)
    assert rhs == int(t.arrays["nfac"].sum()) + len(t.arrays["loss_idx"]) + len(t.arrays["prod_idx"])


def test_algorithmic_flops_from_counters():
    import bench
    stats = np.zeros((3, 8), dtype=np.int64)
    stats[:, 1] = [10, 20, 30]      # RHS evaluations
    stats[:, 2] = [1, 2, 3]         # Newton-matrix setups
    stats[:, 3] = [4, 5, 6]         # Newton iterations
    assert bench.algorithmic_flops(stats, 100.0, 10.0, 1.0) == 6 * 100 + 15 * 10 + 60 * 1


def test_committed_traffic_summary_parses():
    """bench.py reads the HBM bytes per launch from the committed counter summary."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    path = os.path.join(root, "profiles", "solve_points_kernel_grid4096_traffic.txt")
    vals = [float(ln.split()[1]) for ln in open(path) if ln.startswith("traffic_bytes_per_launch")]
    assert len(vals) == 1 and 1e12 < vals[0] < 1e15
