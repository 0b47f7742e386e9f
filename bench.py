"""Benchmark: grid-point solves/s to 1 Myr with p50/p99 solve latency (one MI355X).

Workload (one "step"): the 4096-point static-cloud grid on the default UCLCHEM network
(seeded: log n_H ~ U[2,7], T ~ U[10,100] K, log zeta ~ U[0,2], log G0 ~ U[0,3],
Av ~ U[1,20]), every point integrated 0 -> 1e6 yr with UCLCHEM's interval semantics,
all points in one persistent-kernel launch, inputs resident in HBM, the workgroups taking
the points densest first (--order index: in index order).  FP64 throughout (plain batched
form); the FP32-factorisation form is reported beside it.  Every point carries a total
step bound (--max-steps, default 100,000 BDF steps; points reaching it stop with -5 and
are counted in the line); the CPU baseline runs the reference under the same bound.

A step is a whole grid (tens of seconds), so the timed launches are capped by a wall-time
budget measured from process start (--budget): the line reports the steps actually
timed, and the FP32 and CPU legs run only if they fit (the line says when they did not).

Usage: python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--points 4096] [--lu 64]
Prints one JSON line.
"""
from __future__ import annotations

import time

T_PROC0 = time.time()

import argparse  # noqa: E402
import json  # noqa: E402
import os  # noqa: E402
import signal  # noqa: E402
import subprocess  # noqa: E402
import sys  # noqa: E402

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12     # MI355X dense FP64, FLOP/s
PEAK_HBM = 8.0e12       # bytes/s
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "solve_points_kernel_grid4096_traffic.txt")


def elapsed() -> float:
    return time.time() - T_PROC0


def newton_flop_counts(tables) -> tuple:
    """Algorithmic FLOPs of one Newton-matrix factorisation and one Newton solve of the
    static-pattern sparse plan: 2 per Crout multiply-add pair plus the division of every L
    entry, and (2/3) T^3 for an LU of the dense T x T tail (the staged plans invert the tail
    by Gauss-Jordan, 2 T^3, but only the LU's work is counted); a solve is 2 per stored
    factor entry (both triangular sweeps) plus 2 T^2 for the tail."""
    from uclchem_amd import newton_plan as npl
    s = npl.plan_summary(tables.arrays)
    a = tables.arrays
    ndiv = int((a["nw_tg_div"] >= 0).sum())
    lu = 2.0 * s["npair"] + ndiv + (2.0 / 3.0) * s["T"] ** 3
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


def committed_traffic(path: str = TRAFFIC_FILE):
    """HBM bytes per launch of the benchmarked grid launch from the committed counter
    summary, or (None, reason) when the summary was measured on other kernel sources than
    the ones built here (its source_sha line against uclchem_amd.build.source_hash())."""
    from uclchem_amd.build import source_hash
    if not os.path.exists(path):
        return None, "no counter summary"
    vals = {}
    for ln in open(path):
        parts = ln.split()
        if len(parts) >= 2:
            vals[parts[0]] = parts[1]
    want = source_hash()
    if vals.get("source_sha") != want:
        return None, f"counter summary is of other kernel sources ({vals.get('source_sha')} != {want})"
    return float(vals["traffic_bytes_per_launch"]), "rocprofv3 FETCH_SIZE/WRITE_SIZE of this build"


def stratified_sample(params: np.ndarray, n: int, seed: int = 20261202) -> np.ndarray:
    """n grid-point indices, one drawn at random from each of n equal-count strata of
    initial density (a point's integration cost grows with density)."""
    from uclchem_amd.params import INDEX
    order = np.argsort(params[:, INDEX["initialdens"]], kind="stable")
    rng = np.random.Generator(np.random.PCG64(seed))
    strata = np.array_split(order, n)
    return np.array(sorted(int(rng.choice(s)) for s in strata if len(s)))


def host_cpu() -> tuple:
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(params: np.ndarray, cores: int, cap_s: float, n_sample: int, max_steps: float) -> dict:
    """The reference Fortran path (oracle/_ref, compiled -O3) on the box's host cores, one
    single-threaded process per core pulling points from a queue, under the same total step
    bound as the GPU run and a wall-time cap per point."""
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libucref.so")
    if not os.path.exists(ref_so):
        return None
    from uclchem_amd.params import INDEX
    idx = stratified_sample(params, n_sample)
    sample = params[idx].copy()
    sample[:, INDEX["maxsteps"]] = max_steps
    t0 = time.time()
    code = ("import sys,time,numpy as np; sys.path.insert(0,%r); from oracle import RefLib; "
            "p=np.array([float(x) for x in sys.argv[1].split(',')]); t=time.time(); "
            "rc,rows,st=RefLib().model(p,maxrec=2000); print('%%d %%.6f %%d'%%(rc,time.time()-t,st[4]))"
            ) % os.path.join(ROOT, "oracle")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    running, times, capped, bounded = [], [], 0, 0
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
                times.append(float(out[1]) if len(out) == 3 else time.time() - ts)
                bounded += int(len(out) == 3 and int(out[0]) == -5)
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
    model, nproc = host_cpu()
    return {"value": cores / per_point, "unit": "grid-point solves/s", "cores": cores, "kind": "reference",
            "sample": (f"{len(sample)} points of the same grid, one per density stratum (seeded), through the "
                       f"reference Fortran hot path (oracle/_ref, amdflang -O3), one single-threaded process per "
                       f"core on {cores} cores (the box's CPU share for one GPU; nproc {nproc}, {model}); same "
                       f"step bound {max_steps:.0f} ({bounded} point(s) reached it); {capped} point(s) hit the "
                       f"{cap_s:.0f} s wall cap and count as {cap_s:.0f} s (a lower bound on their time); mean "
                       f"{per_point:.2f} s/point; throughput = cores / mean; sample wall {wall:.1f} s")}


PARTIAL = {}


def _emit_partial(signum, frame):
    """Print what has been measured if the run is stopped from outside (SIGTERM)."""
    if PARTIAL:
        line = dict(PARTIAL)
        line["partial"] = f"stopped by signal {signum} at {elapsed():.0f} s"
        print(json.dumps(line), flush=True)
    sys.exit(1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=4096)
    ap.add_argument("--lu", type=int, default=64)
    ap.add_argument("--also-fp32", type=int, default=1)
    ap.add_argument("--budget", type=float, default=480.0,
                    help="seconds from process start by which every leg must have ended; timed launches "
                         "beyond the first, the FP32 leg and the CPU leg run only if they fit")
    ap.add_argument("--cpu-sample", type=int, default=0, help="CPU baseline points (0: one per core)")
    ap.add_argument("--cpu-cap", type=float, default=30.0, help="wall-time cap per CPU baseline point, s")
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
    signal.signal(signal.SIGTERM, _emit_partial)
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

    def run(lu_bits, npts, label):
        """One launch on `stream`; returns its event-timed duration in ms."""
        d_t1.zero_()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        net.solve_device(d_params[:npts], d_y[:npts], d_status[:npts], d_stats[:npts], t_start=d_t0[:npts],
                         t_done=d_t1[:npts], lu_bits=lu_bits, stream=stream.cuda_stream,
                         order=None if a.order == "index" else a.order)
        ev1.record(stream)
        t_last = t_launch = time.time()
        while not ev1.query():
            time.sleep(0.2)
            if time.time() - t_last > 20.0:
                print(f"[bench] {label}: running, {time.time() - t_launch:.0f} s", file=sys.stderr, flush=True)
                t_last = time.time()
        ev1.synchronize()
        return ev0.elapsed_time(ev1)

    def collect(ms_list, lu_bits):
        st = d_stats.cpu().numpy()
        status = d_status.cpu().numpy()
        t0 = d_t0.cpu().numpy()
        t1 = d_t1.cpu().numpy()
        lat = (t1 - t0.min()) * 1e-8          # launch to retirement, device wall clock at 100 MHz
        solve = (t1 - t0) * 1e-8              # one point's own solve time on its workgroup
        r = dict(ms=float(np.mean(ms_list)), steps=len(ms_list), launches_ms=[round(m, 1) for m in ms_list],
                 stats=st, status=status, lat=lat, solve=solve, y=d_y.cpu().numpy().copy())
        if a.dump:
            np.savez_compressed(a.dump + f"_fp{lu_bits}.npz", stats=st, status=status, solve=solve, lat=lat)
        return r

    cpu_reserve = 0.0 if a.no_cpu else a.cpu_cap + 20.0
    skipped = {}
    # warm-up: W untimed passes over the grid's first 16 points (a whole grid is a long launch)
    for _ in range(a.warmup):
        run(a.lu, min(n, 16), f"fp{a.lu} warmup")
    torch.cuda.synchronize()
    # timed launches of the main form: the first always, the others while they fit
    results = {}
    ms_list = []
    want32 = bool(a.also_fp32 and a.lu == 64)
    while len(ms_list) < a.steps:
        if ms_list:
            est = max(ms_list) / 1e3
            after = (1.15 * est if want32 else 0.0) + cpu_reserve + 10.0
            if elapsed() + 1.1 * est + after > a.budget:
                break
        torch.cuda.synchronize()
        ms_list.append(run(a.lu, n, f"fp{a.lu} step {len(ms_list) + 1}"))
        print(f"[bench] fp{a.lu} step {len(ms_list)}: {ms_list[-1] / 1e3:.1f} s (t={elapsed():.0f} s)",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    results[a.lu] = collect(ms_list, a.lu)
    main_r = results[a.lu]
    value = n / (main_r["ms"] / 1e3)
    lu_flop, solve_flop = newton_flop_counts(tables)
    rhs_flop = rhs_flop_count(tables)
    flops = algorithmic_flops(main_r["stats"], lu_flop, solve_flop, rhs_flop)
    achieved = flops / (main_r["ms"] / 1e3)
    traffic, traffic_src = (committed_traffic() if (a.workload == "grid" and a.lu == 64 and n == 4096 and
                                                    a.max_steps == 100000.0 and a.order == "densest")
                            else (None, "no counter summary for this workload"))
    wl = {"grid": (f"static cloud grid, default network ({tables.nspec} species, {tables.nreac} reactions), "
                   f"{n} points to 1e6 yr, FP{a.lu} factorisation" + (" (plain batched FP64)" if a.lu == 64 else "")),
          "collapse500": (f"free-fall collapse grid (n_H,0 log-uniform 1e2..1e6 towards 1e6 cm^-3, 1 Myr), "
                          f"seeded synthetic network ({tables.nspec} species, {tables.nreac} reactions), {n} points, "
                          f"FP{a.lu} factorisation"),
          "hotcore": (f"hot-core warm-up grid (temp_indx 4, 10->300 K, log n_H U[5,8]), default network, {n} "
                      f"points to 1e6 yr, FP{a.lu} factorisation")}[a.workload]
    if a.max_steps > 0:
        wl += f", total step bound {a.max_steps:.0f} BDF steps per point"
    line = {
        "metric": "grid-point solves/s to 1 Myr (p50/p99 solve latency, 1% abundance agreement)",
        "value": round(value, 3), "unit": "grid-point solves/s", "n_gpus": 1, "steps": main_r["steps"],
        "warmup": a.warmup, "ms_per_step": round(main_r["ms"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": f"synthetic (seeded {n}-point {a.workload} workload)",
        "config": {"workload": wl, "points": n, "network": tables.name, "lu": f"fp{a.lu}",
                   "schedule": "densest first" if a.order == "densest" else "index order",
                   "max_steps_per_point": a.max_steps},
        "steps_requested": a.steps, "launches_ms": main_r["launches_ms"],
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
        "roofline": {"bound": "fp64", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": traffic,
                     "note": (f"solve_points_kernel (one launch = the whole grid), algorithmic FP64 FLOPs from the "
                              f"solver counters over the event-timed launch: sparse-plan factorisation "
                              f"({lu_flop:.4g} per Newton-matrix setup: Crout pairs x2, divisions, 2/3 T^3 of the "
                              f"dense tail), solves ({solve_flop:.4g} per Newton iteration) and RHS "
                              f"({rhs_flop:.4g} per evaluation); traffic: {traffic_src}")},
    }
    PARTIAL.update(line)
    if want32:
        est = main_r["ms"] / 1e3
        if elapsed() + 1.15 * est + cpu_reserve + 10.0 <= a.budget:
            torch.cuda.synchronize()
            ms32 = run(32, n, "fp32")
            torch.cuda.synchronize()
            r32 = results[32] = collect([ms32], 32)
            line["fp32_lu"] = {"value": round(n / (r32["ms"] / 1e3), 3), "ms_per_step": round(r32["ms"], 3),
                               "steps": 1, "latency_p50": float(np.percentile(r32["lat"], 50)),
                               "latency_p99": float(np.percentile(r32["lat"], 99)),
                               "status_nonzero": int(np.count_nonzero(r32["status"]))}
            ok = (main_r["status"] == 0) & (r32["status"] == 0)
            y64, y32 = main_r["y"][ok, :tables.nspec], r32["y"][ok, :tables.nspec]
            rel = np.where(y64 > 1e-14, np.abs(y32 - y64) / np.maximum(y64, 1e-300), 0.0).max(axis=1)
            line["fp32_lu"].update({"points_compared": int(ok.sum()),
                                    "points_within_1pct_of_fp64": int((rel <= 0.01).sum()),
                                    "median_max_rel_vs_fp64": float(np.median(rel)) if len(rel) else None})
        else:
            skipped["fp32_lu"] = f"did not fit the {a.budget:.0f} s budget at t={elapsed():.0f} s"
        PARTIAL.update(line)
    cores = min(16, len(os.sched_getaffinity(0)))
    if a.no_cpu:
        line["cpu_baseline"] = None
    elif elapsed() + cpu_reserve <= a.budget + 20.0:
        line["cpu_baseline"] = cpu_baseline(params, cores, a.cpu_cap, a.cpu_sample or cores, a.max_steps)
    else:
        line["cpu_baseline"] = None
        skipped["cpu_baseline"] = f"did not fit the {a.budget:.0f} s budget at t={elapsed():.0f} s"
    if skipped:
        line["skipped"] = skipped
    line["wall_s"] = round(elapsed(), 1)
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
    slots = serving.plan_shared_slots(nets, n_cu, budget, lu_bits=a.lu)   # one queue shared by the networks
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
                               "one request queue shared by the three networks with continuous admission, "
                               "slots by CU count and HBM budget",
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
