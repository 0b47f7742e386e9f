"""Measured choice of the solve configuration for a network and a workload.

The kernel's shapes that can change at run time are: the linear solver of the Newton matrix
(the static-pattern sparse plan or the dense LU, ``ucg_set_linear_solver``), the precision of
its factors (FP64, or FP32 with FP64 residuals, ``lu_bits``) and the number of resident
workgroups a launch asks for (``n_slots``: one or two per CU; LDS holds one workgroup per CU,
the second one queues behind it and takes the next point the moment a point retires).
Which of them is fastest depends on the network (its plan's dense tail, its size) and on
the workload (how many points are stiff, and how stiff), so it is measured: a stratified
sample of the workload is solved in every configuration, each timed with device events, each
checked against the FP64 sparse-plan solve of the same points (every abundance above 1e-14
within 1% on at least 99% of the points), and the fastest configuration that passes is kept.

    python -m uclchem_amd.autotune [--network default] [--points 512] [--final-time 1e5]
"""
from __future__ import annotations

import argparse
import json
import time
from typing import Dict, List, Optional

import numpy as np

CANDIDATES: List[Dict] = [dict(linear_solver=ls, lu_bits=b, slots_per_cu=s)
                          for ls in ("sparse", "dense") for b in (64, 32) for s in (1, 2)]


def stratified(params: np.ndarray, n: int, column: int = 0) -> np.ndarray:
    """n points of `params`, one from the middle of each stratum of `column` (the initial
    density by default: a point's cost grows with it)."""
    order = np.argsort(params[:, column], kind="stable")
    parts = np.array_split(order, min(n, len(order)))
    return np.array([p[len(p) // 2] for p in parts if len(p)])


def _agree(y: np.ndarray, ref: np.ndarray, nspec: int, floor: float = 1e-14, rtol: float = 0.01) -> float:
    """Fraction of points whose abundances above `floor` all agree with `ref` within rtol."""
    ok = 0
    for q in range(len(ref)):
        m = ref[q, :nspec] > floor
        ok += int(np.all(np.abs(y[q, :nspec][m] - ref[q, :nspec][m]) <= rtol * ref[q, :nspec][m]))
    return ok / max(len(ref), 1)


def measure(net, params: np.ndarray, cfg: Dict, n_cu: int) -> Dict:
    """Solve `params` once in configuration `cfg` on the device; returns its wall time (device
    events around the launch), results and per-point status."""
    import torch
    dev = torch.device("cuda:0")
    n = len(params)
    d = dict(params=torch.tensor(params, device=dev),
             y_out=torch.zeros((n, net.neq), dtype=torch.float64, device=dev),
             status=torch.zeros(n, dtype=torch.int32, device=dev),
             stats=torch.zeros((n, 8), dtype=torch.int32, device=dev))
    net.set_linear_solver(net.SPARSE if cfg["linear_solver"] == "sparse" else net.DENSE)
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    try:
        e0.record(stream)
        net.solve_device(d["params"], d["y_out"], d["status"], d["stats"], lu_bits=cfg["lu_bits"],
                         n_slots=cfg["slots_per_cu"] * n_cu, stream=stream.cuda_stream, order="densest")
        e1.record(stream)
        torch.cuda.synchronize()
    finally:
        net.set_linear_solver(net.SPARSE)
    return dict(seconds=e0.elapsed_time(e1) / 1e3, y=d["y_out"].cpu().numpy(), status=d["status"].cpu().numpy())


def autotune(net, params: np.ndarray, n_cu: int, candidates: Optional[List[Dict]] = None,
             min_agreement: float = 0.99, measure_fn=None) -> Dict:
    """Measure every candidate on `params` (the sample) and return {"choice": cfg,
    "table": [cfg + seconds, points_per_s, agreement, failed]}.  The first candidate
    measured is the reference (FP64 sparse plan, one slot per CU); a candidate qualifies when
    no point fails that the reference solved and at least `min_agreement` of the points agree
    with the reference within 1%."""
    measure_fn = measure_fn or (lambda cfg: measure(net, params, cfg, n_cu))
    cands = list(candidates or CANDIDATES)
    ref_cfg = dict(linear_solver="sparse", lu_bits=64, slots_per_cu=1)
    if ref_cfg in cands:
        cands.remove(ref_cfg)
    cands.insert(0, ref_cfg)
    table, ref = [], None
    for cfg in cands:
        r = measure_fn(cfg)
        if ref is None:
            ref = r
        failed = int(np.sum((r["status"] != 0) & (ref["status"] == 0)))
        ok = ref["status"] == 0
        agree = _agree(r["y"][ok], ref["y"][ok], net.nspec) if ok.any() else 0.0
        table.append(dict(cfg, seconds=r["seconds"], points_per_s=len(params) / r["seconds"], agreement=agree,
                          failed=failed))
    good = [t for t in table if t["failed"] == 0 and t["agreement"] >= min_agreement]
    best = min(good, key=lambda t: t["seconds"])
    choice = {k: best[k] for k in ("linear_solver", "lu_bits", "slots_per_cu")}
    return dict(choice=choice, table=table)


def main() -> None:
    from . import gpu, makerates_hip as mh, synthetic as sy, workloads
    from .params import INDEX
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--network", default="default")
    ap.add_argument("--points", type=int, default=512)
    ap.add_argument("--final-time", type=float, default=1e5)
    ap.add_argument("--workload", default="grid", choices=["grid", "collapse"])
    a = ap.parse_args()
    import torch
    t = (mh.load_network(mh.network_path(a.network)) if a.network == "default" else sy.load_or_build(a.network))
    net = gpu.Network(t)
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    pool = workloads.static_cloud_grid(4096) if a.workload == "grid" else workloads.collapse_grid(8192)
    sample = pool[stratified(pool, a.points, INDEX["initialdens"])].copy()
    sample[:, INDEX["finaltime"]] = a.final_time
    t0 = time.time()
    res = autotune(net, sample, n_cu)
    for row in res["table"]:
        print(json.dumps(row), flush=True)
    print(json.dumps(dict(network=a.network, points=len(sample), final_time=a.final_time, choice=res["choice"],
                          wall_s=round(time.time() - t0, 1))), flush=True)


if __name__ == "__main__":
    main()
