"""Benchmark: grid-point solves/s to 1 Myr with p50/p99 solve latency (one MI355X).

Workload (one "step"): the 4096-point static-cloud grid on the default UCLCHEM network
(seeded: log n_H ~ U[2,7], T ~ U[10,100] K, log zeta ~ U[0,2], log G0 ~ U[0,3],
Av ~ U[1,20]), every point integrated 0 -> 1e6 yr with UCLCHEM's interval semantics,
all points in one persistent-kernel launch, inputs resident in HBM, the workgroups taking
the points densest first (--order index: in index order).  FP64 throughout (plain batched
form); the FP32-factorisation form is reported beside it.

Usage: python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--points 4096] [--lu 64]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12     # MI355X dense FP64 (vector and matrix), FLOP/s
PEAK_HBM = 8.0e12       # bytes/s


def newton_flop_counts(tables) -> tuple:
    """FLOPs of one factorisation and one solve of the static-pattern sparse plan:
    2 per Crout multiply-add pair plus the division of every L entry, 2 T^3 for the
    Gauss-Jordan inverse of the dense tail (2/3 T^3 for its LU when the plan is not staged);
    2 per stored factor entry (L and U, both triangular solves) plus the tail's 2 T^2."""
    from uclchem_amd import newton_plan as npl
    s = npl.plan_summary(tables.arrays)
    a = tables.arrays
    ndiv = int((a["nw_tg_div"] >= 0).sum())
    # the staged plans invert the trailing block by Gauss-Jordan (2 T^3), unstaged ones factor it
    staged = npl.plan_is_staged(tables)
    lu = 2.0 * s["npair"] + ndiv + (2.0 if staged else 2.0 / 3.0) * s["T"] ** 3
    solve = 2.0 * s["nsp"] + 2.0 * s["T"] ** 2
    return lu, solve


def algorithmic_flops(stats: np.ndarray, lu_flop: float, solve_flop: float, rhs_flop: float) -> float:
    """FLOPs implied by the solver counters: factorisations per Newton-matrix setup,
    solves per Newton iteration, generated multiply/adds per RHS evaluation."""
    setups = stats[:, 2].sum()
    iters = stats[:, 3].sum()
    rhs = stats[:, 1].sum()
    return setups * lu_flop + iters * solve_flop + rhs * rhs_flop


def rhs_flop_count(tables) -> float:
    a = tables.arrays
    mults = int((a["nfac"]).sum())                 # one multiply per factor of every term
    adds = len(a["loss_idx"]) + len(a["prod_idx"])  # one add per term occurrence
    return float(mults + adds)


def cpu_baseline(params: np.ndarray, cores: int, cap_s: float, n_sample: int) -> dict:
    """The reference Fortran path (oracle/_ref) on the box's host cores, one process per
    core, each solving one grid point with a wall-time cap."""
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libucref.so")
    if not os.path.exists(ref_so):
        return None
    sample = params[:n_sample]
    procs = []
    t0 = time.time()
    code = ("import sys,time,numpy as np; sys.path.insert(0,%r); from oracle import RefLib; "
            "p=np.array([float(x) for x in sys.argv[1].split(',')]); t=time.time(); "
            "rc,rows,st=RefLib().model(p,maxrec=2000); print('%%d %%.6f'%%(rc,time.time()-t))") % os.path.join(ROOT, "oracle")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    running, times, capped = [], [], 0
    queue = list(range(len(sample)))
    while queue or running:
        while queue and len(running) < cores:
            i = queue.pop(0)
            arg = ",".join(repr(float(x)) for x in sample[i])
            pr = subprocess.Popen([sys.executable, "-c", code, arg], stdout=subprocess.PIPE,
                                  stderr=subprocess.DEVNULL, text=True, env=env)
            running.append((pr, time.time()))
        still = []
        for pr, ts in running:
            if pr.poll() is not None:
                out = pr.stdout.read().split()
                times.append(float(out[1]) if len(out) == 2 else time.time() - ts)
            elif time.time() - ts > cap_s:
                pr.kill()
                pr.wait()
                times.append(cap_s)
                capped += 1
            else:
                still.append((pr, ts))
        running = still
        time.sleep(0.05)
    wall = time.time() - t0
    per_point = float(np.mean(times))
    return {"value": cores / per_point, "unit": "grid-point solves/s", "cores": cores, "kind": "reference",
            "sample": (f"first {len(sample)} points of the same 4096-point grid through the reference Fortran "
                       f"(oracle/_ref, amdflang -O2), one process per core; {capped} point(s) hit the "
                       f"{cap_s:.0f} s cap and count as {cap_s:.0f} s (a lower bound); mean {per_point:.2f} s/point; "
                       f"sample wall {wall:.1f} s")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=4096)
    ap.add_argument("--lu", type=int, default=64)
    ap.add_argument("--also-fp32", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=16)
    ap.add_argument("--cpu-cap", type=float, default=25.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--dump", default="", help="write per-point counters and times to <prefix>_fp<bits>.npz")
    ap.add_argument("--workload", default="grid", choices=["grid", "collapse500", "hotcore", "serving"],
                    help="grid: 4096-point static cloud grid (default); collapse500: free-fall grid on the "
                         "500-species synthetic network; hotcore: hot-core warm-up grid; serving: Poisson "
                         "request stream over three networks")
    ap.add_argument("--max-steps", type=float, default=100000.0,
                    help="per-point total BDF step bound (UCG_P_MAXSTEPS; 0 = unbounded, the reference's "
                         "semantics; the default is the work the reference allows one output interval, "
                         "10 attempts x MXSTEP); points reaching it report status -5")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="only run a plain device-to-device copy of this many bytes (HBM counter calibration)")
    ap.add_argument("--requests", type=int, default=192, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=20.0, help="serving: Poisson arrival rate, requests/s")
    ap.add_argument("--order", default="densest", choices=["densest", "index"],
                    help="order in which the workgroups take the points: densest first (the costliest points "
                         "start first, so the batch's tail is shorter) or index order")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; grid points are independent and would shard "
                          "by contiguous point ranges with no collective", "n_gpus": a.gpus}))
        sys.exit(2)
    if a.copy_calibration > 0:
        import torch
        n = int(a.copy_calibration) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8}), flush=True)
        return
    if a.workload == "serving":
        return bench_serving(a)
    import torch
    from uclchem_amd import gpu, workloads
    from uclchem_amd import makerates_hip as mh
    from uclchem_amd.params import INDEX

    if a.workload == "collapse500":
        from uclchem_amd import synthetic as sy
        tables = sy.load_or_build("synthetic_500_6000")
        if a.points == 4096:
            a.points = 8192
        params = workloads.collapse_grid(a.points)
        a.no_cpu = True              # the reference is compiled for its default network only
    else:
        tables = mh.load_network(mh.network_path("default"))
        params = (workloads.hot_core_grid(a.points) if a.workload == "hotcore"
                  else workloads.static_cloud_grid(a.points))
    if a.max_steps > 0:
        params[:, INDEX["maxsteps"]] = a.max_steps
    net = gpu.Network(tables)
    dev = torch.device("cuda:0")
    n, neq = len(params), tables.neq
    d_params = torch.tensor(params, dtype=torch.float64, device=dev)
    d_y = torch.zeros((n, neq), dtype=torch.float64, device=dev)
    d_status = torch.zeros(n, dtype=torch.int32, device=dev)
    d_stats = torch.zeros((n, 8), dtype=torch.int32, device=dev)
    d_t0 = torch.zeros(n, dtype=torch.float64, device=dev)
    d_t1 = torch.zeros(n, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)

    side = torch.cuda.Stream(dev)                 # default priority, like every stream here
    host_done = torch.empty(n, dtype=torch.float64, pin_memory=True)

    def run(lu_bits, npts, label):
        d_t1.zero_()
        net.solve_device(d_params[:npts], d_y[:npts], d_status[:npts], d_stats[:npts], t_start=d_t0[:npts],
                         t_done=d_t1[:npts], lu_bits=lu_bits, stream=stream.cuda_stream,
                         order=None if a.order == "index" else a.order)
        done_ev = torch.cuda.Event()
        done_ev.record(stream)
        t_last = t_launch = time.time()
        while not done_ev.query():                # progress from the kernel's own retirement stamps
            time.sleep(0.5)
            if time.time() - t_last > 20.0:
                if a.order == "index":
                    with torch.cuda.stream(side):
                        host_done[:npts].copy_(d_t1[:npts], non_blocking=True)
                    side.synchronize()
                    print(f"[bench] {label}: {int((host_done[:npts] > 0).sum())}/{npts} points retired",
                          file=sys.stderr, flush=True)
                else:                             # the stamps land in d_t1 when the launch ends
                    print(f"[bench] {label}: running, {time.time() - t_launch:.0f} s", file=sys.stderr, flush=True)
                t_last = time.time()

    results = {}
    for lu_bits in ([a.lu] + ([32] if (a.also_fp32 and a.lu == 64) else [])):
        for _ in range(a.warmup):
            run(lu_bits, min(n, 16), f"fp{lu_bits} warmup")
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t_wall = time.time()
        ev0.record(stream)
        for _ in range(a.steps):
            run(lu_bits, n, f"fp{lu_bits}")
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.time() - t_wall
        ms = ev0.elapsed_time(ev1) / a.steps
        st = d_stats.cpu().numpy()
        status = d_status.cpu().numpy()
        t0 = d_t0.cpu().numpy()
        t1 = d_t1.cpu().numpy()
        lat = (t1 - t0.min()) * 1e-8          # launch to retirement, device wall clock at 100 MHz
        solve = (t1 - t0) * 1e-8              # one point's own solve time on its workgroup
        results[lu_bits] = dict(ms=ms, wall=wall, stats=st, status=status, lat=lat, solve=solve,
                                y=d_y.cpu().numpy().copy())
        if a.dump:
            np.savez_compressed(a.dump + f"_fp{lu_bits}.npz", stats=st, status=status, solve=solve, lat=lat)
    main_r = results[a.lu]
    value = n / (main_r["ms"] / 1e3)
    lu_flop, solve_flop = newton_flop_counts(tables)
    flops = algorithmic_flops(main_r["stats"], lu_flop, solve_flop, rhs_flop_count(tables))
    achieved = flops / (main_r["ms"] / 1e3)
    traffic = None      # HBM bytes per launch from the committed FETCH_SIZE/WRITE_SIZE summary
    tpath = os.path.join(ROOT, "profiles", "solve_points_kernel_grid4096_traffic.txt")
    if a.workload == "grid" and a.lu == 64 and n == 4096 and a.max_steps == 100000.0 and os.path.exists(tpath):
        for ln in open(tpath):
            if ln.startswith("traffic_bytes_per_launch"):
                traffic = float(ln.split()[1])
    wl = {"grid": (f"static cloud grid, default network ({tables.nspec} species, {tables.nreac} reactions), "
                   f"{n} points to 1e6 yr, FP{a.lu} factorisation" + (" (plain batched FP64)" if a.lu == 64 else "")),
          "collapse500": (f"free-fall collapse grid (n_H,0 log-uniform 1e2..1e6 towards 1e6 cm^-3, 1 Myr), "
                          f"seeded synthetic network ({tables.nspec} species, {tables.nreac} reactions), {n} points, "
                          f"FP{a.lu} factorisation"),
          "hotcore": (f"hot-core warm-up grid (temp_indx 4, 10->300 K, log n_H U[5,8]), default network, {n} "
                      f"points to 1e6 yr, FP{a.lu} factorisation")}[a.workload]
    line = {
        "metric": "grid-point solves/s to 1 Myr (p50/p99 solve latency, 1% abundance agreement)",
        "value": round(value, 3), "unit": "grid-point solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(main_r["ms"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": f"synthetic (seeded {n}-point {a.workload} workload)",
        "config": {"workload": wl, "points": n, "network": tables.name, "lu": f"fp{a.lu}",
                   "schedule": "densest first" if a.order == "densest" else "index order"},
        "latency_s": {"p50": float(np.percentile(main_r["lat"], 50)), "p99": float(np.percentile(main_r["lat"], 99)),
                      "solve_p50": float(np.percentile(main_r["solve"], 50)),
                      "solve_p99": float(np.percentile(main_r["solve"], 99))},
        "status_nonzero": int(np.count_nonzero(main_r["status"])),
        "step_bound": {"max_steps_per_point": a.max_steps,
                       "points_at_bound": int(np.sum((main_r["status"] == -5) &
                                                     (main_r["stats"][:, 0] >= a.max_steps)))
                       if a.max_steps > 0 else 0},
        "solver_counters": {"steps": int(main_r["stats"][:, 0].sum()), "rhs": int(main_r["stats"][:, 1].sum()),
                            "newton_setups": int(main_r["stats"][:, 2].sum()),
                            "newton_iters": int(main_r["stats"][:, 3].sum())},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": traffic,
                     "note": (f"solve_points_kernel (one launch = the whole grid); FP64 FLOPs of the sparse-plan "
                              f"factorisation ({lu_flop:.3g} per Newton-matrix setup), solves ({solve_flop:.3g} per "
                              f"Newton iteration) and RHS ({rhs_flop_count(tables):.3g} per evaluation) from the "
                              f"solver counters, over the launch time; traffic: HBM bytes per launch from "
                              f"rocprofv3 FETCH_SIZE/WRITE_SIZE scaled by a 1 GiB copy (profiles/)")},
    }
    if 32 in results and a.lu == 64:
        r32 = results[32]
        line["fp32_lu"] = {"value": round(n / (r32["ms"] / 1e3), 3), "ms_per_step": round(r32["ms"], 3),
                           "latency_p50": float(np.percentile(r32["lat"], 50)),
                           "latency_p99": float(np.percentile(r32["lat"], 99)),
                           "status_nonzero": int(np.count_nonzero(r32["status"]))}
        ok = (main_r["status"] == 0) & (r32["status"] == 0)
        y64, y32 = main_r["y"][ok, :tables.nspec], r32["y"][ok, :tables.nspec]
        rel = np.where(y64 > 1e-14, np.abs(y32 - y64) / np.maximum(y64, 1e-300), 0.0).max(axis=1)
        line["fp32_lu"].update({"points_compared": int(ok.sum()),
                                "points_within_1pct_of_fp64": int((rel <= 0.01).sum()),
                                "median_max_rel_vs_fp64": float(np.median(rel)) if len(rel) else None})
    if not a.no_cpu:
        cores = min(16, len(os.sched_getaffinity(0)))
        line["cpu_baseline"] = cpu_baseline(params, cores, a.cpu_cap, a.cpu_sample)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def bench_serving(a):
    """Poisson stream of single-point requests over three resident networks (default,
    150/1500 and 500/6000 synthetic), t_end log-uniform in [1e3, 1e7] yr, request queues
    with continuous admission; throughput = requests / wall, latency = arrival -> result."""
    from uclchem_amd import gpu, serving
    from uclchem_amd import makerates_hip as mh
    from uclchem_amd import synthetic as sy
    names = ["synthetic_150_1500", "default", "synthetic_500_6000"]
    nets = {"default": gpu.Network(mh.load_network(mh.network_path("default")))}
    for nm in names:
        if nm != "default":
            nets[nm] = gpu.Network(sy.load_or_build(nm))
    import torch
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    budget = 64e9                                   # HBM the serving state may take
    slots = serving.plan_slots(nets, n_cu, budget, lu_bits=a.lu)
    reqs = serving.poisson_requests(a.requests, a.rate, names)
    srv = serving.Server(nets, slots, capacity=a.requests + 8, lu_bits=a.lu)
    try:
        warm = [(0.0, nm, serving.make_params(initialdens=1e4, initialtemp=10.0, finaltime=1e3)) for nm in names]
        serving.run_stream(srv, warm, timeout_s=120.0)
        t0 = time.time()
        done = serving.run_stream(srv, reqs, timeout_s=1100.0,
                                  progress=lambda d, s, t: print(f"[bench] serving: {d}/{s} done at {t:.0f} s",
                                                                 file=sys.stderr, flush=True))
        wall = time.time() - t0
    finally:
        srv.close()
    lat = np.array([t.t_done - t.t_arrival for t in done])
    by_net = {nm: np.array([t.t_done - t.t_arrival for t in done if t.network == nm]) for nm in names}
    line = {
        "metric": "serving: completed single-point solves/s with p50/p99 latency (arrival -> result)",
        "value": round(len(done) / wall, 3), "unit": "requests/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(wall * 1e3, 1), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": f"synthetic (seeded Poisson stream, lambda={a.rate}/s, {a.requests} requests)",
        "config": {"workload": "serving loop: Poisson single-point requests over the default network and the "
                               "150/1500 and 500/6000 synthetic networks, t_end log-uniform 1e3..1e7 yr, "
                               "request queues with continuous admission, slots by HBM budget",
                   "slots": slots, "hbm_budget_gb": budget / 1e9, "lu": f"fp{a.lu}"},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "per_network_p50": {k: float(np.percentile(v, 50)) for k, v in by_net.items() if len(v)},
                      "per_network_p99": {k: float(np.percentile(v, 99)) for k, v in by_net.items() if len(v)}},
        "status_nonzero": int(sum(t.status != 0 for t in done)),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
