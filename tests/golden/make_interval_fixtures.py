"""Write single-interval reference fixtures (tests/golden/ref_intervals.npz) from the reference
Fortran hot path (oracle/_ref/libucref.so).

Run in the build container (needs the reference build):
    python tests/golden/make_interval_fixtures.py [--procs 6] [--budget 2400]

Each row is one updateChemistry interval of the reference's model loop
(src/fortran_src/chemistry.f90:145-239): the point's parameter vector, the state the interval
starts from (abundances; density, gasTemp, dustTemp, Av, cloudSize, phi, abstol_factor, time),
the interval's end, and the reference's result (state, time reached, steps, attempts).  The
states come from the reference's own trajectories walked interval by interval
(ucref_interval + ucref_update_physics, which continue the loop of ucref_model bit for bit):

* "hot": hot-core warm-up points (temp_indx 4, maxTemp 300 K: the hot_core_grid workload)
  at n_H near 1e5, 1e6, 1e7 and 1e8 cm^-3, walked from t = 0 for a wall-time budget;
* "tail": points 594, 380, 1203 and 2549 of the 4096-point static-cloud grid, whose late
  intervals need ~1e5 reference steps.

Rows kept per trajectory: the hardest intervals the reference completes in one attempt;
where its ten attempts end short of the interval's end, the stretch its first attempt covers
(the target is then the time that attempt reached); and, for hot-core points, the last
walked state with its temperature raised to 20, 50, 100 and 300 K ("jump" rows: the
desorption conditions of the warm-up, one interval each, from an icy state the walk reached).

On the first-attempt stretches the reference's DVODE spends its whole step budget, and its
answer for the smallest abundances depends on its tolerance.  For those rows the fixture also
holds y_tight, the stretch's solution at tighter tolerances, so that a test can tell the
reference's own integration error from a difference of method (other rows: NaN): the
reference itself at 100 times tighter reltol and abstol_factor where it finishes within the
per-row cap (tight_src 1), otherwise the BDF core of tests/bdf_host.cpp on the reference's own
F (ucref_rhs, the stretch's rates) with a finite-difference Jacobian and pivoted dense LU at
the same tighter tolerances (tight_src 2).
    python tests/golden/make_interval_fixtures.py --tight-only [--cap 1800] [--only 3,4]
"""
from __future__ import annotations

import argparse
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from uclchem_amd import workloads  # noqa: E402
from uclchem_amd.params import INDEX, SECONDS_PER_YEAR as SPY  # noqa: E402

TAIL_POINTS = (594, 380, 1203, 2549)
HOT_TARGETS = (1e5, 1e6, 1e7, 1e8)
JUMP_TEMPS = (20.0, 50.0, 100.0, 300.0)


def hot_points():
    """Indices of the hot-core workload whose n_H is nearest each of HOT_TARGETS (first 256)."""
    g = workloads.hot_core_grid(256)
    ln = np.log10(g[:, INDEX["initialdens"]])
    return [int(np.argmin(np.abs(ln - np.log10(t)))) for t in HOT_TARGETS]


def walk(job):
    """Walk one trajectory interval by interval; returns its candidate rows."""
    kind, idx, budget = job
    from oracle import RefLib, Oracle
    from uclchem_amd import makerates_hip as mh
    ref = RefLib()
    orc = Oracle(mh.load_network(mh.network_path("default")))
    p = (workloads.hot_core_grid(256) if kind == "hot" else workloads.static_cloud_grid(4096))[idx]
    model = int(p[INDEX["model"]])
    mxstep = float(p[INDEX["mxstep"]])
    t_wall = time.time()
    rc, _, _, y, phys = ref.model_until(p, 1.0)
    rows = []
    while rc == 0 and phys[7] / SPY < 1e6 and time.time() - t_wall < budget:
        t1 = orc.next_target(model, phys[7] / SPY)
        ts = time.time()
        rc, y1, ph1, st = ref.interval(p, y, phys, t1)
        wall = time.time() - ts
        reached = ph1[7] >= t1
        if reached and st[1] == 1:
            rows.append(dict(kind=kind, idx=idx, how="interval", params=p, y_in=y, phys_in=phys, target=t1,
                             y_ref=y1, phys_ref=ph1, rc=rc, steps=st[0], attempts=st[1], wall=wall))
        elif not reached:   # the stretch the first attempt covers
            rc1, y2, ph2, st2 = ref.interval(p, y, phys, t1, maxsteps=mxstep)
            if ph2[7] > phys[7]:
                rows.append(dict(kind=kind, idx=idx, how="first-attempt", params=p, y_in=y, phys_in=phys,
                                 target=ph2[7], y_ref=y2, phys_ref=ph2, rc=0, steps=st2[0], attempts=st2[1],
                                 wall=wall))
        ref.update_physics(ph1[7])
        y, phys = ref.state()
        print(f"{kind} {idx}: t={phys[7] / SPY:.4g} yr T={phys[1]:.1f} K steps {st[0]} attempts {st[1]} "
              f"{wall:.1f}s", flush=True)
    if kind == "hot":   # temperature jumps from the last walked state
        for T in JUMP_TEMPS:
            ph = phys.copy()
            ph[1] = ph[2] = T
            t1 = orc.next_target(model, ph[7] / SPY)
            ts = time.time()
            rc, y1, ph1, st = ref.interval(p, y, ph, t1)
            rows.append(dict(kind=kind, idx=idx, how=f"jump {T:.0f} K", params=p, y_in=y, phys_in=ph, target=t1,
                             y_ref=y1, phys_ref=ph1, rc=rc, steps=st[0], attempts=st[1], wall=time.time() - ts))
            print(f"{kind} {idx}: jump to {T} K: steps {st[0]} attempts {st[1]} reached "
                  f"{ph1[7] >= t1}", flush=True)
    return rows


def select(rows, n_hard=3, n_first=2):
    """Per trajectory: the n_hard single-attempt intervals with the most steps, the n_first
    first-attempt rows with the most steps, every jump row whose interval the reference ends."""
    out = []
    keys = sorted({(r["kind"], r["idx"]) for r in rows})
    for key in keys:
        mine = [r for r in rows if (r["kind"], r["idx"]) == key]
        hard = sorted([r for r in mine if r["how"] == "interval"], key=lambda r: -r["steps"])[:n_hard]
        first = sorted([r for r in mine if r["how"] == "first-attempt"], key=lambda r: -r["steps"])[:n_first]
        jump = [r for r in mine if r["how"].startswith("jump") and r["phys_ref"][7] >= r["target"]]
        out += hard + first + jump
    return out


def _host_core():
    """The host build of the BDF core (tests/bdf_host.cpp), compiled to a temporary file."""
    import ctypes
    import subprocess
    import tempfile
    out = os.path.join(tempfile.mkdtemp(), "libbdfhost.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", out,
                           os.path.join(os.path.dirname(HERE), "bdf_host.cpp")])
    lib = ctypes.CDLL(out)
    lib.bdf_host_solve_fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    return lib


def tight_ref(job):
    """The reference over one first-attempt stretch at 100x tighter tolerances, in a child
    process with a wall-time cap; None when it is exceeded or the stretch is not reached."""
    import subprocess
    import tempfile
    p, y, phys, target, cap = job
    p = p.copy()
    ph = phys.copy()
    p[INDEX["reltol"]] *= 1e-2
    p[INDEX["mxstep"]] = 1e7
    ph[6] *= 1e-2
    with tempfile.NamedTemporaryFile(suffix=".npz") as tmp:
        np.savez(tmp.name + ".in.npz", p=p, y=y, ph=ph, t=target)
        code = ("import sys,numpy as np; sys.path.insert(0,%r); from oracle import RefLib; "
                "z=np.load(sys.argv[1]); rc,y1,ph1,st=RefLib().interval(z['p'],z['y'],z['ph'],float(z['t'])); "
                "np.savez(sys.argv[2],rc=rc,y1=y1,ph1=ph1,st=st)") % os.path.join(ROOT, "oracle")
        try:
            subprocess.run([sys.executable, "-c", code, tmp.name + ".in.npz", tmp.name], timeout=cap, check=True)
        except subprocess.TimeoutExpired:
            return None
        finally:
            os.remove(tmp.name + ".in.npz")
        z = np.load(tmp.name)
        ok = int(z["rc"]) == 0 and z["ph1"][7] >= target * (1 - 1e-12)
        print(f"tight reference: {int(z['st'][0])} steps, rc {int(z['rc'])}, reached {ok}", flush=True)
        return z["y1"].copy() if ok else None


def tight_host(p, y, phys, target):
    """The BDF core on the reference's F over the stretch at 100x tighter tolerances."""
    import ctypes
    from oracle import RefLib
    lib = _host_core()
    ref = RefLib()
    ref.interval(p, y, phys, target, maxsteps=1.0)     # leaves the stretch's rates in the reference
    rhs = ctypes.cast(ref.lib.ucref_rhs, ctypes.c_void_p)
    atol = np.maximum(1e-2 * phys[6] * y, 1e-2 * p[INDEX["abstol_min"]])
    yo = y.copy()
    st = np.zeros(7, dtype=np.int32)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    r = lib.bdf_host_solve_fn(len(yo), rhs, phys[7], target, 1e-2 * p[INDEX["reltol"]], ptr(atol), ptr(yo), 500000,
                              ptr(st))
    print(f"tight host core: {st[0]} steps, return {r}", flush=True)
    return yo if r == 2 else None


def _tight_row(job):
    q, p, y, phys, target, cap = job
    y1 = tight_ref((p, y, phys, target, cap))
    if y1 is not None:
        return q, y1, 1
    y1 = tight_host(p, y, phys, target)
    return q, y1, (2 if y1 is not None else 0)


def add_tight(procs, cap, only=None):
    """Adds y_tight / tight_src row by row, writing the fixture after every finished row (a
    rerun with `only` fills the rows still missing)."""
    from uclchem_amd.params import widen
    path = os.path.join(HERE, "ref_intervals.npz")
    f = dict(np.load(path))
    if "y_tight" not in f:
        f["y_tight"] = np.full_like(f["y_ref"], np.nan)
        f["tight_src"] = np.zeros(len(f["target"]), dtype=np.int32)
    rows = [q for q in range(len(f["target"])) if f["how"][q] == "first-attempt" and (only is None or q in only)]
    params = widen(f["params"])
    jobs = [(q, params[q], f["y_in"][q], f["phys_in"][q], float(f["target"][q]), cap) for q in rows]
    with mp.get_context("fork").Pool(procs) as pool:
        for q, y1, src in pool.imap_unordered(_tight_row, jobs):
            f["y_tight"][q] = np.nan if y1 is None else y1
            f["tight_src"][q] = src
            np.savez_compressed(path, **f)
            print(f"row {q}: tight source {src}", flush=True)


def split_multicall():
    """Rows whose interval the reference ends in more than one integrator call (the first runs
    out of MXSTEP, the next starts from where it stopped with the rates recomputed there,
    chemistry.f90:164-221) are split into those calls: "<how> call 1" (the stretch the first
    call covers, as the first-attempt rows) and "<how> call 2" (from the reference's own state
    at that point to the interval's end).  Each of them is one call with frozen rates; the
    whole row depends on where the step budget ran out and is kept for reference only."""
    from oracle import RefLib
    from uclchem_amd.params import widen
    path = os.path.join(HERE, "ref_intervals.npz")
    f = dict(np.load(path))
    ref = RefLib()
    params = widen(f["params"])
    n0 = len(f["target"])
    add = {k: [] for k in f}
    for q in range(n0):
        how = str(f["how"][q])
        if f["attempts"][q] <= 1 or how == "first-attempt" or " call " in how or f["phys_ref"][q, 7] < f["target"][q]:
            continue
        if any(str(h) == how + " call 1" and f["idx"][r] == f["idx"][q] for r, h in enumerate(f["how"])):
            continue
        p, y, ph, t1 = params[q], f["y_in"][q], f["phys_in"][q], float(f["target"][q])
        ta = time.time()
        rc1, y1, ph1, st1 = ref.interval(p, y, ph, t1, maxsteps=float(p[INDEX["mxstep"]]))
        w1 = time.time() - ta
        tb = time.time()
        rc2, y2, ph2, st2 = ref.interval(p, y1, ph1, t1)
        w2 = time.time() - tb
        # the two calls continue the reference's own interval bit for bit (the first stops at its step
        # bound, -5)
        assert rc1 in (0, -5) and rc2 == 0 and np.array_equal(y2, f["y_ref"][q]), (q, rc1, rc2)
        for part, (yi, phi, tt, yo, pho, st, w) in enumerate([(y, ph, ph1[7], y1, ph1, st1, w1),
                                                              (y1, ph1, t1, y2, ph2, st2, w2)]):
            row = dict(kind=f["kind"][q], idx=f["idx"][q], how=f"{how} call {part + 1}", params=f["params"][q],
                       y_in=yi, phys_in=phi, target=tt, y_ref=yo, phys_ref=pho, rc=0, steps=st[0],
                       attempts=st[1], wall=w, y_tight=np.full_like(yo, np.nan), tight_src=0)
            for k in add:
                add[k].append(row[k])
            print(f"row {q} ({f['kind'][q]} {f['idx'][q]} {how}) call {part + 1}: t {phi[7] / SPY:.6g} -> "
                  f"{tt / SPY:.6g} yr, steps {st[0]}, attempts {st[1]}", flush=True)
    for k in add:
        if add[k]:
            new = np.array(add[k])
            dt = np.result_type(f[k].dtype, new.dtype) if f[k].dtype.kind in "US" else f[k].dtype
            f[k] = np.concatenate([f[k].astype(dt), new.astype(dt)])
    np.savez_compressed(path, **f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--tight-only", action="store_true", help="add y_tight to the existing fixture")
    ap.add_argument("--split-multicall", action="store_true",
                    help="add the single-call parts of rows the reference ends in several calls")
    ap.add_argument("--cap", type=float, default=1800.0, help="wall seconds per tight reference stretch")
    ap.add_argument("--only", default=None, help="comma-separated rows for --tight-only")
    ap.add_argument("--procs", type=int, default=6)
    ap.add_argument("--budget", type=float, default=2400.0, help="wall seconds per hot-core trajectory walk")
    ap.add_argument("--tail-budget", type=float, default=900.0, help="wall seconds per tail-point walk")
    a = ap.parse_args()
    if a.split_multicall:
        split_multicall()
        return
    if a.tight_only:
        add_tight(a.procs, a.cap, None if a.only is None else [int(x) for x in a.only.split(",")])
        return
    jobs = [("hot", i, a.budget) for i in hot_points()] + [("tail", i, a.tail_budget) for i in TAIL_POINTS]
    with mp.get_context("fork").Pool(a.procs) as pool:
        rows = [r for part in pool.map(walk, jobs, chunksize=1) for r in part]
    sel = select(rows)
    cols = dict(kind=np.array([r["kind"] for r in sel]), idx=np.array([r["idx"] for r in sel]),
                how=np.array([r["how"] for r in sel]))
    for k in ("params", "y_in", "phys_in", "target", "y_ref", "phys_ref", "rc", "steps", "attempts", "wall"):
        cols[k] = np.array([r[k] for r in sel])
    np.savez_compressed(os.path.join(HERE, "ref_intervals.npz"), **cols)
    for r in sel:
        print(f"{r['kind']:4s} {r['idx']:5d} {r['how']:14s} T={r['phys_in'][1]:6.1f} K "
              f"t={r['phys_in'][7] / SPY:10.4g} -> {r['target'] / SPY:10.4g} yr steps {r['steps']:6d} "
              f"attempts {r['attempts']} rc {r['rc']} ({r['wall']:.1f}s)")


if __name__ == "__main__":
    main()
