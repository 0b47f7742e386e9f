"""The measured autotuner on the device: every configuration (sparse plan or dense LU, FP64
or FP32 factors, one or two resident workgroups per CU) solves the same stratified sample of
the static-cloud grid (to 1e3 yr), each is timed and checked against the FP64 sparse-plan
solve, and the choice is the fastest configuration that agrees."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_measures_every_configuration(gpu_net):
    import torch
    from uclchem_amd import autotune, workloads
    from uclchem_amd.params import INDEX
    pool = workloads.static_cloud_grid(4096)
    sample = pool[autotune.stratified(pool, 64, INDEX["initialdens"])].copy()
    sample[:, INDEX["finaltime"]] = 1e3
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    res = autotune.autotune(gpu_net, sample, n_cu)
    for row in res["table"]:
        print(f"[autotune] {row}", flush=True)
    assert len(res["table"]) == len(autotune.CANDIDATES)
    assert res["table"][0]["agreement"] == 1.0
    assert all(r["seconds"] > 0 and r["failed"] == 0 for r in res["table"])
    assert all(r["agreement"] >= 0.99 for r in res["table"])     # every form solves these points alike
    best = min(res["table"], key=lambda r: r["seconds"])
    assert res["choice"] == {k: best[k] for k in ("linear_solver", "lu_bits", "slots_per_cu")}
