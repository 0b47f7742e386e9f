"""Selection logic of the measured autotuner (uclchem_amd/autotune.py) on the CPU, with a
stand-in measurement: the reference configuration is measured first, a faster configuration
whose results disagree or that fails points is never chosen, and the fastest qualifying one
is."""
import numpy as np

from uclchem_amd import autotune as at


class _Net:
    nspec = 3


def _fake(results):
    def m(cfg):
        key = (cfg["linear_solver"], cfg["lu_bits"], cfg["slots_per_cu"])
        return results[key]
    return m


def test_fastest_agreeing_configuration_is_chosen():
    y = np.array([[1e-4, 2e-6, 3e-10, 1.0], [5e-5, 1e-6, 2e-9, 1.0]])
    ok = np.zeros(2, dtype=np.int32)
    bad = y.copy()
    bad[0, 2] *= 1.5                                   # one point off by 50%
    res = {}
    for c in at.CANDIDATES:
        res[(c["linear_solver"], c["lu_bits"], c["slots_per_cu"])] = dict(seconds=10.0, y=y, status=ok)
    res[("sparse", 32, 2)] = dict(seconds=4.0, y=y * (1 + 1e-4), status=ok)    # fastest that agrees
    res[("sparse", 32, 1)] = dict(seconds=2.0, y=bad, status=ok)               # faster but wrong
    res[("dense", 64, 2)] = dict(seconds=1.0, y=y, status=np.array([0, -5], dtype=np.int32))  # fails a point
    order = []
    fake = _fake(res)
    out = at.autotune(_Net(), np.zeros((2, 64)), n_cu=256, measure_fn=lambda cfg: (order.append(cfg), fake(cfg))[1])
    assert order[0] == dict(linear_solver="sparse", lu_bits=64, slots_per_cu=1)
    assert len(order) == len(at.CANDIDATES)
    assert out["choice"] == dict(linear_solver="sparse", lu_bits=32, slots_per_cu=2)
    rows = {(r["linear_solver"], r["lu_bits"], r["slots_per_cu"]): r for r in out["table"]}
    assert rows[("sparse", 32, 1)]["agreement"] == 0.5
    assert rows[("dense", 64, 2)]["failed"] == 1


def test_stratified_sample_covers_the_range():
    p = np.zeros((100, 64))
    p[:, 0] = np.logspace(2, 7, 100)[::-1]
    idx = at.stratified(p, 10)
    assert len(idx) == 10 and len(set(idx.tolist())) == 10
    d = np.sort(p[idx, 0])
    assert d[0] < 1e3 and d[-1] > 1e6
