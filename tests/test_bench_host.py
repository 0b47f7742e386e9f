"""Host-side arithmetic of bench.py (no GPU): the algorithmic FLOP counts per Newton-matrix
setup, per Newton iteration and per RHS evaluation that turn the solver counters into
achieved FLOP/s, the traffic summary's build check, and the CPU baseline's sample."""
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
    # algorithmic work: an LU of the dense tail (2/3 T^3), whatever the kernel does with it
    assert lu == 2 * s["npair"] + ndiv + (2.0 / 3.0) * s["T"] ** 3
    assert solve == 2 * s["nsp"] + 2 * s["T"] ** 2
    assert (s["T"], s["npair"]) == (104, 54116)
    assert abs(lu - 0.861e6) < 0.01e6 and abs(solve - 3.37e4) < 0.01e4
    rhs = bench.rhs_flop_count(t)
    assert rhs == int(t.arrays["nfac"].sum()) + len(t.arrays["loss_idx"]) + len(t.arrays["prod_idx"])


def test_algorithmic_flops_from_counters():
    import bench
    stats = np.zeros((3, 8), dtype=np.int64)
    stats[:, 1] = [10, 20, 30]      # RHS evaluations
    stats[:, 2] = [1, 2, 3]         # Newton-matrix setups
    stats[:, 3] = [4, 5, 6]         # Newton iterations
    assert bench.algorithmic_flops(stats, 100.0, 10.0, 1.0) == 6 * 100 + 15 * 10 + 60 * 1


def test_traffic_summary_is_used_only_for_its_own_build(tmp_path):
    """The HBM bytes per launch come from a counter summary only when it names the kernel
    sources built here."""
    import bench
    from uclchem_amd.build import source_hash
    good = tmp_path / "good.txt"
    good.write_text(f"source_sha {source_hash()}\ntraffic_bytes_per_launch 1.5e12\n")
    assert bench.committed_traffic(str(good))[0] == 1.5e12
    stale = tmp_path / "stale.txt"
    stale.write_text("source_sha 0123456789abcdef\ntraffic_bytes_per_launch 1.5e12\n")
    v, why = bench.committed_traffic(str(stale))
    assert v is None and "other kernel sources" in why
    assert bench.committed_traffic(str(tmp_path / "missing.txt"))[0] is None


def test_cpu_sample_is_stratified_by_density():
    import bench
    from uclchem_amd import workloads
    from uclchem_amd.params import INDEX
    p = workloads.static_cloud_grid(256)
    idx = bench.stratified_sample(p, 16)
    assert len(idx) == 16 and len(set(idx.tolist())) == 16
    ranks = np.argsort(np.argsort(p[:, INDEX["initialdens"]], kind="stable"), kind="stable")
    assert sorted((ranks[idx] // 16).tolist()) == list(range(16))   # one point per density stratum
    assert np.array_equal(idx, bench.stratified_sample(p, 16))      # seeded
