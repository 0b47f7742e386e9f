"""Write golden fixtures from the reference Fortran hot path (oracle/_ref/libucref.so).

Run in the build container (needs /root/reference for the build and the example output):
    python tests/golden/make_ref_fixtures.py [--grid 24] [--procs 8]

Outputs (small .npz files next to this script):
  static_full_t1e6.npz   examples/example-output/static-full.dat row t=1e6 yr (6 digits)
  ref_cloud_point.npz    the reference model for the default static cloud point
  ref_cloud_grid.npz     the reference model for points of the 4096-point grid, one per density
                         stratum (--grid-only), with its answers at tight, half and twice its
                         default tolerances (--grid-tight, --grid-variants)
  example_collapse_hotcore.npz  examples/example-output/phase1-full.dat, startcollapse.dat,
                         phase2-full.dat (the collapse and hot-core example models)
  ref_collapse500.npz    the reference compiled on the 500/6000 synthetic network (written by
                         uclchem_amd/reference_format.py: oracle/_ref/net_synthetic_500_6000)
                         for 16 points of the 8192-point free-fall grid, one per density stratum
  ref_serving.npz        the reference, compiled on each request's network, for the seeded
                         24-request stream over the default and synthetic networks
                         (t_end log-uniform 1e3..1e7 yr)
  ref_coeff.npz          the static cloud to 1e5 yr with reaction coefficients changed (the
                         alpha/beta/gamma keys) and the rate array at five output rows
  ref_start_hotcore.npz  a model from given start abundances and a hot-core model (temp_indx 3,
                         300 K, n_H 1e5 to 2e4 yr), with the final abundances at tight
                         tolerances (--start-hotcore-only)
  ref_collapse_models.npz  the reference's collapse module (collapse.f90: BE1.1, BE4, filament,
                         ambipolar) for two points each: the physics of every output row, the
                         abundances of five rows (collapse_rows), and those at tight tolerances
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

REF_EXAMPLE = "/root/reference/examples/example-output/static-full.dat"


SERVING_NETS = ["synthetic_150_1500", "default", "synthetic_500_6000"]


def _lib_path(net):
    if net is None:
        return None
    return os.path.join(ROOT, "oracle", "_ref", f"net_{net}", "libucref.so")


def _run_one(job):
    p, net = job if isinstance(job, tuple) else (job, None)
    from oracle import RefLib
    ref = RefLib() if net is None else RefLib(_lib_path(net))
    t0 = time.time()
    rc, rows, stats = ref.model(p, maxrec=4000)
    return rc, rows[-1].copy(), rows[:, 0].copy(), stats, time.time() - t0


def run_models(params, procs, nets=None):
    jobs = list(params) if nets is None else list(zip(params, nets))
    with mp.get_context("fork").Pool(procs) as pool:
        out = pool.map(_run_one, jobs, chunksize=1)
    rc = np.array([o[0] for o in out])
    rows = [o[1] for o in out]
    final = np.stack(rows) if len({len(r) for r in rows}) == 1 else rows   # networks of different sizes
    stats = np.stack([o[3] for o in out])
    wall = np.array([o[4] for o in out])
    nrows = np.array([len(o[2]) for o in out])
    return rc, final, stats, wall, nrows


def _read_output(path):
    """Header names and rows of a reference full-output file (io.f90:85-88 format)."""
    lines = open(path).read().splitlines()
    hdr = [h.strip() for h in lines[0].split(",")]
    rows = np.array([[float(x) for x in ln.split(",")] for ln in lines[1:]])
    return hdr, rows


def example_models():
    """The collapse and hot-core example outputs the reference ships
    (scripts/run_uclchem_tests.py:28-58): phase 1 (free-fall 1e2 -> 1e5 cm^-3, stop at the
    final density), its saved final abundances, phase 2 (hot core, temp_indx 3, 300 K, from
    those abundances).  Data only; no reference code runs."""
    d = os.path.dirname(REF_EXAMPLE)
    hdr, r1 = _read_output(os.path.join(d, "phase1-full.dat"))
    start = np.array([float(x) for x in open(os.path.join(d, "startcollapse.dat")).read().replace(",", " ").split()])
    hdr2, r2 = _read_output(os.path.join(d, "phase2-full.dat"))
    assert hdr == hdr2 and len(start) == len(hdr) - 8
    np.savez_compressed(os.path.join(HERE, "example_collapse_hotcore.npz"), species=np.array(hdr[8:], dtype="U16"),
                        phase1_times=r1[:, 0], phase1_physics=r1[:, :8], phase1_final=r1[-1, 8:],
                        start_abundances=start, phase2_times=r2[:, 0], phase2_physics=r2[:, :8],
                        phase2_final=r2[-1, 8:])


def _model_capped(args):
    """One reference model in a child process with a wall-time cap; None when it is exceeded.
    args = (params, cap) or (params, cap, network name of an oracle/_ref/net_<name> build)."""
    import subprocess
    import tempfile
    p, cap = args[:2]
    lib = _lib_path(args[2]) if len(args) > 2 else None
    with tempfile.NamedTemporaryFile(suffix=".npz") as tmp:
        code = ("import sys,time,numpy as np; sys.path.insert(0,%r); from oracle import RefLib; "
                "p=np.array([float(x) for x in sys.argv[1].split(',')]); t=time.time(); "
                "ref=RefLib(sys.argv[3]) if len(sys.argv)>3 else RefLib(); "
                "rc,rows,st=ref.model(p,maxrec=2000); "
                "np.savez(sys.argv[2],rc=rc,final=rows[-1],nrows=len(rows),stats=st,wall=time.time()-t)"
                ) % os.path.join(ROOT, "oracle")
        try:
            subprocess.run([sys.executable, "-c", code, ",".join(repr(float(x)) for x in p), tmp.name] +
                           ([lib] if lib else []), timeout=cap, check=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
        except subprocess.TimeoutExpired:
            return None
        z = np.load(tmp.name)
        return int(z["rc"]), z["final"].copy(), int(z["nrows"]), z["stats"].copy(), float(z["wall"])


def start_and_hotcore_fixture():
    """ref_start_hotcore.npz: (1) a model started from given abundances (abundanceStart,
    wrap.f90:636-640): the static cloud (n_H 1e4, 10 K) to 1e3 yr, its final abundances then the
    start of a denser, warmer cloud (n_H 1e5, 20 K) to 1e4 yr; (2) the hot core of the reference's
    model API (temp_indx 3, maxTemp 300 K, n_H 1e5) to 2e4 yr: every row's time, density, gasTemp
    and Av, and the final abundances, also at tight tolerances (reltol 1e-11, abstol_factor 1e-17,
    abstol_min 1e-30)."""
    from oracle import RefLib
    from uclchem_amd.params import INDEX, make_params
    ref = RefLib()
    p1 = make_params(initialdens=1e4, initialtemp=10.0, finaltime=1e3)
    rc1, rows1, _ = ref.model(p1, maxrec=200)
    y_start = np.append(rows1[-1, 4:], 1e5)          # abundances, then the next model's density
    p2 = make_params(initialdens=1e5, initialtemp=20.0, finaltime=1e4)
    rc2, rows2, _ = ref.model_from(p2, y_start, maxrec=200)
    p3 = make_params(dict(model=1, tempindx=3, maxtemp=300.0), initialdens=1e5, finaltime=2e4)
    rc3, rows3, _ = ref.model(p3, maxrec=400)
    assert rc1 == 0 and rc2 == 0 and rc3 == 0

    def tight(p):
        t = p.copy()
        t[INDEX["reltol"]], t[INDEX["abstol_factor"]], t[INDEX["abstol_min"]] = 1e-11, 1e-17, 1e-30
        return t
    rc2t, rows2t, _ = ref.model_from(tight(p2), y_start, maxrec=200)
    rc3t, rows3t, _ = ref.model(tight(p3), maxrec=400)
    assert rc2t == 0 and rc3t == 0 and len(rows2t) == len(rows2) and len(rows3t) == len(rows3)
    np.savez_compressed(os.path.join(HERE, "ref_start_hotcore.npz"), start_params=p2, y_start=y_start,
                        start_rows=rows2[:, :4], start_final=rows2[-1], start_final_tight=rows2t[-1],
                        hot_params=p3, hot_rows=rows3[:, :4], hot_final=rows3[-1], hot_final_tight=rows3t[-1])
    print(f"ref_start_hotcore: start {len(rows2)} rows, hot core {len(rows3)} rows", flush=True)


COLLAPSE_MODES = {"BE1.1": 1, "BE4": 2, "filament": 3, "ambipolar": 4}


def collapse_points():
    """Two points per collapse model: the reference's defaults, and one with a warmer, more
    strongly irradiated cloud (initialTemp 14 K, zeta 3, radfield 5, baseAv 1, rout 0.08 pc).
    Modes 1 and 2 end at 0.97 maxTime whatever finalTime says; the filament fits hold until
    5.47 time units (about 1.39 Myr), so modes 3 and 4 run to 1 Myr and 1.2 Myr."""
    from uclchem_amd.params import make_params
    out, names = [], []
    for name, mode in COLLAPSE_MODES.items():
        for q, extra in enumerate([dict(), dict(initialtemp=14.0, zeta=3.0, radfield=5.0, baseav=1.0, rout=0.08)]):
            ft = 1.0e6 if mode <= 3 else 1.2e6
            if mode == 3 and q:
                ft = 1.2e6
            out.append(make_params(dict(model=2, collapse=mode, finaltime=ft, **extra)))
            names.append(name)
    return np.stack(out), np.array(names)


def collapse_rows(n):
    """The output rows of a collapse fixture whose abundances are kept: the row after 1e4 yr
    (13th), 1/8, 1/4 and 1/2 of the way, and the last."""
    return np.array([12, n // 8, n // 4, n // 2, n - 1])


def _collapse_capped(args):
    """One reference model in a child process with a wall-time cap: (rc, physics rows [nrec, 4],
    the rows of collapse_rows, stats, wall), or None when the cap is exceeded.
    args = (params, cap, lib)."""
    import subprocess
    import tempfile
    p, cap, lib = args
    with tempfile.NamedTemporaryFile(suffix=".npz") as tmp:
        code = ("import sys,time,numpy as np; sys.path.insert(0,%r); from oracle import RefLib; "
                "p=np.array([float(x) for x in sys.argv[1].split(',')]); t=time.time(); ref=RefLib(sys.argv[3]); "
                "rc,rows,st=ref.model(p,maxrec=4000); n=len(rows); k=np.array([12,n//8,n//4,n//2,n-1]); "
                "np.savez(sys.argv[2],rc=rc,phys=rows[:,:4],keep=rows[k],stats=st,wall=time.time()-t)"
                ) % os.path.join(ROOT, "oracle")
        try:
            subprocess.run([sys.executable, "-c", code, ",".join(repr(float(x)) for x in p), tmp.name, lib],
                           timeout=cap, check=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
        except subprocess.TimeoutExpired:
            return None
        z = np.load(tmp.name)
        return int(z["rc"]), z["phys"].copy(), z["keep"].copy(), z["stats"].copy(), float(z["wall"])


def collapse_fixture(procs, cap, lib):
    """ref_collapse_models.npz: collapse_points through the reference at its default
    tolerances (every row's time, density, gasTemp and Av; the rows of collapse_rows in full)
    and at tight tolerances (rows_tight, NaN where the run exceeds `cap` seconds or writes
    other rows)."""
    from uclchem_amd.params import INDEX
    params, names = collapse_points()
    tight = params.copy()
    tight[:, INDEX["reltol"]] = 1e-11
    tight[:, INDEX["abstol_factor"]] = 1e-17
    tight[:, INDEX["abstol_min"]] = 1e-30
    t0 = time.time()
    jobs = [(p, cap, lib) for p in params] + [(p, cap, lib) for p in tight]
    with mp.get_context("fork").Pool(procs) as pool:
        res = pool.map(_collapse_capped, jobs, chunksize=1)
    n = len(params)
    base, tres = res[:n], res[n:]
    assert all(r is not None for r in base), "a default-tolerance collapse model exceeded the cap"
    nrows = np.array([len(r[1]) for r in base])
    phys = np.full((n, nrows.max(), 4), np.nan)
    for q, r in enumerate(base):
        phys[q, :nrows[q]] = r[1]
    rows = np.stack([r[2] for r in base])               # [n, 5, 4 + nspec]
    rows_tight = np.full_like(rows, np.nan)
    tight_wall = np.full(n, np.nan)
    for q, r in enumerate(tres):
        if r is not None and r[0] == 0 and np.array_equal(r[2][:, 0], rows[q, :, 0]):
            rows_tight[q] = r[2]
            tight_wall[q] = r[4]
    np.savez_compressed(os.path.join(HERE, "ref_collapse_models.npz"), params=params, model=names,
                        rc=np.array([r[0] for r in base]), nrows=nrows, phys=phys,
                        row_index=np.stack([collapse_rows(k) for k in nrows]), rows=rows, rows_tight=rows_tight,
                        stats=np.stack([r[3] for r in base]), wall=np.array([r[4] for r in base]),
                        tight_wall=tight_wall)
    print(f"ref_collapse_models: {n} points in {time.time() - t0:.1f}s, rows {nrows.tolist()}, "
          f"tight {int(np.isfinite(rows_tight[:, 0, 0]).sum())} of {n}", flush=True)


def grid4096_sample(n, procs, cap):
    """n points of the benchmarked 4096-point static-cloud grid, one per initial-density
    stratum (seeded draw), each run through the reference; a point whose reference run
    exceeds `cap` seconds (the grid's pathologically stiff tail, covered interval by interval
    by ref_intervals.npz) is replaced by the next draw from its stratum."""
    from uclchem_amd.params import INDEX
    g = workloads.static_cloud_grid(4096)
    order = np.argsort(g[:, INDEX["initialdens"]], kind="stable")
    rng = np.random.Generator(np.random.PCG64(5))
    strata = [list(rng.permutation(s)) for s in np.array_split(order, n)]
    chosen = [None] * n
    with mp.get_context("fork").Pool(procs) as pool:
        while any(c is None for c in chosen):
            todo = [k for k in range(n) if chosen[k] is None]
            cand = [int(strata[k].pop(0)) for k in todo]
            res = pool.map(_model_capped, [(g[i], cap) for i in cand], chunksize=1)
            for k, i, r in zip(todo, cand, res):
                if r is not None:
                    chosen[k] = (i, r)
                else:
                    print(f"grid point {i} exceeded {cap:.0f} s in the reference; drawing again", flush=True)
    idx = np.array([c[0] for c in chosen])
    res = [c[1] for c in chosen]
    return idx, g[idx], np.array([r[0] for r in res]), np.stack([r[1] for r in res]), \
        np.stack([r[3] for r in res]), np.array([r[4] for r in res]), np.array([r[2] for r in res])


def grid_tight(procs, cap, name="ref_cloud_grid.npz", net=None, only=None):
    """Add to a whole-model fixture (ref_cloud_grid.npz, or ref_alltypes.npz with net
    "alltypes") the reference's answer at tight tolerances (reltol 1e-11, abstol_factor 1e-17,
    abstol_min 1e-30) for each of its points: final_tight (NaN where the run exceeds `cap`
    seconds).  |final - final_tight| is the reference's own integration noise on that point:
    on the grid's stiff tail it includes the different places where its integrator runs out
    of steps and the rates are recomputed mid-interval (chemistry.f90:164-221)."""
    from uclchem_amd.params import INDEX, widen
    path = os.path.join(HERE, name)
    f = dict(np.load(path))
    params = widen(f["params"])
    tight = params.copy()
    tight[:, INDEX["reltol"]] = 1e-11
    tight[:, INDEX["abstol_factor"]] = 1e-17
    tight[:, INDEX["abstol_min"]] = 1e-30
    todo = [q for q in range(len(params)) if only is None or q in only]
    with mp.get_context("fork").Pool(procs) as pool:
        res = pool.map(_model_capped, [(tight[q], cap) if net is None else (tight[q], cap, net) for q in todo],
                       chunksize=1)
    ft = f["final_tight"].copy() if "final_tight" in f else np.full_like(f["final"], np.nan)
    wall = f["tight_wall"].copy() if "tight_wall" in f else np.full(len(params), np.nan)
    for q, r in zip(todo, res):
        ft[q] = np.nan
        wall[q] = np.nan
        # the same end time (at tight tolerances the reference may split more intervals and
        # write more rows before it)
        if r is not None and r[0] == 0 and r[1][0] == f["final"][q, 0]:
            ft[q] = r[1]
            wall[q] = r[4]
        print(f"point {f['index'][q] if 'index' in f else q}: tight "
              f"{'capped' if r is None else (r[0], r[2], r[1][0], round(r[4], 1))}", flush=True)
    f["final_tight"] = ft
    f["tight_wall"] = wall
    np.savez_compressed(path, **f)


def grid_variants(procs, cap, name="ref_cloud_grid.npz"):
    """Add to ref_cloud_grid.npz the reference's answers at half and at twice its relative
    tolerance (final_half, final_double; NaN where a run exceeds `cap` seconds or ends at
    another time).  With final_tight they measure how far the reference's own answer moves
    under a small change of tolerance: on the grid's chaotic points (integrator calls that run
    out of steps, so that the rates are recomputed mid-interval at places that depend on the
    step sequence, chemistry.f90:164-221) it moves by far more than 1%."""
    from uclchem_amd.params import INDEX, widen
    path = os.path.join(HERE, name)
    f = dict(np.load(path))
    params = widen(f["params"])
    jobs = []
    for fac in (0.5, 2.0):
        for q in range(len(params)):
            p = params[q].copy()
            p[INDEX["reltol"]] *= fac
            jobs.append((p, cap))
    with mp.get_context("fork").Pool(procs) as pool:
        res = pool.map(_model_capped, jobs, chunksize=1)
    n = len(params)
    for k, key in enumerate(("final_half", "final_double")):
        out = np.full_like(f["final"], np.nan)
        for q in range(n):
            r = res[k * n + q]
            if r is not None and r[0] == 0 and r[1][0] == f["final"][q, 0]:
                out[q] = r[1]
        f[key] = out
        print(f"{key}: {int(np.isfinite(out[:, 0]).sum())} of {n}", flush=True)
    np.savez_compressed(path, **f)


def coeff_changes(tables):
    """The coefficient changes of ref_coeff.npz: H3+ + CO (alpha, beta), the first CRP
    reaction (alpha), the first LH and LHDES reactions (gamma); 1-based indices."""
    from uclchem_amd import makerates_hip as mh
    names = list(tables.species_names)
    first = lambda kind: next(k for k in range(tables.nreac) if mh.ALL_TYPES[int(tables.rtype[k])] == kind)
    two = next(k for k in range(tables.nreac) if mh.ALL_TYPES[int(tables.rtype[k])] == "TWOBODY"
               and sorted(names[i] for i in tables.re[k][:2] if 0 <= i < tables.nspec) == ["CO", "H3+"])
    return dict(alpha={two + 1: 3.3e-9, first("CRP") + 1: 0.7}, beta={two + 1: -0.25},
                gamma={first("LH") + 1: 600.0, first("LHDES") + 1: 50.0})


RATE_ROWS = (1, 5, 10, 20, -1)


def coeff_fixture():
    """ref_coeff.npz: the static cloud (n_H 1e4, 10 K) to 1e5 yr with the coefficient changes
    of coeff_changes, through the reference with its module arrays changed the same way
    (ucref_set_coefficient), recording the rate array at output rows RATE_ROWS
    (returnRates, io.f90:90-93)."""
    from oracle import RefLib
    from uclchem_amd import makerates_hip as mh
    from uclchem_amd.params import make_params
    t = mh.load_network(mh.network_path("default"))
    ch = coeff_changes(t)
    ref = RefLib()
    old = {w: {k: float(t.arrays["gama" if w == "gamma" else w][k - 1]) for k in c} for w, c in ch.items()}
    for w, c in ch.items():
        for k, v in c.items():
            ref.set_coefficient(w, k, v)
    p = make_params(initialdens=1e4, initialtemp=10.0, finaltime=1e5)
    rc, rows, stats, rates = ref.model_rates(p, maxrec=200)
    for w, c in old.items():
        for k, v in c.items():
            ref.set_coefficient(w, k, v)
    which = np.array([w for w, c in ch.items() for _ in c])
    idx = np.array([k for c in ch.values() for k in c])
    val = np.array([v for c in ch.values() for v in c.values()])
    sel = np.array([r % len(rows) for r in RATE_ROWS])
    np.savez_compressed(os.path.join(HERE, "ref_coeff.npz"), params=p, which=which, index=idx, value=val, rc=rc,
                        rows=rows, stats=stats, rate_rows=sel, rates=rates[sel])
    print(f"ref_coeff: rc {rc}, {len(rows)} rows, rates at rows {sel.tolist()}", flush=True)


def nondefault_params(n=8, seed=9):
    """n static-cloud points to 1e5 yr whose initial-abundance, desorption-constant and switch
    parameters are all away from their defaults (seeded), in the product's parameter layout."""
    from uclchem_amd.params import INDEX, make_params
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n):
        sw = rng.integers(0, 2, 6)
        out.append(make_params(
            initialdens=10 ** rng.uniform(2, 6), initialtemp=rng.uniform(10, 60), zeta=10 ** rng.uniform(0, 1.5),
            radfield=10 ** rng.uniform(0, 2), baseav=rng.uniform(1, 5), finaltime=1e5,
            metallicity=rng.uniform(0.5, 1.5), ion=int(i % 3), fh=rng.uniform(0.2, 0.8), fhe=rng.uniform(0.08, 0.12),
            fc=rng.uniform(1.2e-4, 2.4e-4), fo=rng.uniform(2.5e-4, 4.5e-4), fn=rng.uniform(4e-5, 8e-5),
            fs=rng.uniform(2e-6, 6e-6), ebmaxh2=rng.uniform(800, 2000), ebmaxcr=rng.uniform(800, 2000),
            ebmaxuvcr=rng.uniform(5e3, 2e4), epsilon=rng.uniform(0.005, 0.03), uv_yield=rng.uniform(0.01, 0.06),
            phi=10 ** rng.uniform(4, 6), uvcreff=10 ** rng.uniform(-3.5, -2.5), omega=rng.uniform(0.3, 0.7),
            desorb=float(sw[0] or i % 2), h2desorb=float(sw[1]), crdesorb=float(sw[2]), uvdesorb=float(sw[3]),
            cosmicrayattenuation=float(sw[5]), ionmodel=float(i % 2), improvedh2crpdissociation=float(sw[5]),
            enforcechargeconservation=float(sw[4])))
        out[-1][INDEX["thermdesorb"]] = float(i % 2)       # the slot itself (the dict key sets uvdesorb)
    return np.array(out)


def collapse500_sample(n=16):
    """Indices of n points of the 8192-point free-fall grid, one drawn per initial-density
    stratum (seeded)."""
    from uclchem_amd.params import INDEX
    g = workloads.collapse_grid(8192)
    order = np.argsort(g[:, INDEX["initialdens"]], kind="stable")
    rng = np.random.Generator(np.random.PCG64(11))
    return np.array(sorted(int(rng.choice(s)) for s in np.array_split(order, n)))


def serving_stream():
    """The seeded 24-request stream of the serving parity test."""
    from uclchem_amd import serving
    return serving.poisson_requests(24, rate_per_s=40.0, networks=SERVING_NETS, seed=7, t_end_range=(1e3, 1e7))


def synthetic_fixtures(procs):
    idx = collapse500_sample()
    params = workloads.collapse_grid(8192)[idx]
    t0 = time.time()
    rc, final, stats, wall, nrows = run_models(params, procs, ["synthetic_500_6000"] * len(params))
    np.savez_compressed(os.path.join(HERE, "ref_collapse500.npz"), index=idx, params=params, rc=rc, final=final,
                        stats=stats, wall=wall, nrows=nrows)
    print(f"ref_collapse500: {len(params)} points in {time.time() - t0:.1f}s, per-point wall median "
          f"{np.median(wall):.2f}s max {wall.max():.2f}s, rc {set(rc.tolist())}", flush=True)
    reqs = serving_stream()
    params = np.array([p for _, _, p in reqs])
    nets = [n for _, n, _ in reqs]
    t0 = time.time()
    rc, final, stats, wall, nrows = run_models(params, procs, nets)
    width = max(len(f) for f in final)
    fin = np.zeros((len(final), width))
    for i, f in enumerate(final):
        fin[i, :len(f)] = f
    np.savez_compressed(os.path.join(HERE, "ref_serving.npz"), params=params, network=np.array(nets),
                        arrival=np.array([a for a, _, _ in reqs]), rc=rc, final=fin, stats=stats, wall=wall,
                        nrows=nrows)
    print(f"ref_serving: {len(params)} requests in {time.time() - t0:.1f}s, per-request wall median "
          f"{np.median(wall):.2f}s max {wall.max():.2f}s, rc {set(rc.tolist())}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=24)
    ap.add_argument("--procs", type=int, default=8)
    ap.add_argument("--examples-only", action="store_true", help="only convert the reference's example outputs")
    ap.add_argument("--synthetic-only", action="store_true", help="only the synthetic-network fixtures")
    ap.add_argument("--grid-only", action="store_true", help="only the 4096-grid sample (ref_cloud_grid.npz)")
    ap.add_argument("--params-only", action="store_true", help="only the non-default-parameter points")
    ap.add_argument("--alltypes-only", action="store_true", help="only the all-reaction-types network points")
    ap.add_argument("--grid-tight", action="store_true", help="add tight-tolerance answers to ref_cloud_grid.npz")
    ap.add_argument("--coeff-only", action="store_true", help="only the coefficient-change / rates-output model")
    ap.add_argument("--alltypes-tight", action="store_true", help="add tight-tolerance answers to ref_alltypes.npz")
    ap.add_argument("--only", default=None, help="comma-separated fixture rows to (re)compute with --*-tight")
    ap.add_argument("--cap", type=float, default=900.0, help="reference wall-time cap per grid point, s")
    ap.add_argument("--collapse-only", action="store_true", help="only the collapse-module models")
    ap.add_argument("--start-hotcore-only", action="store_true",
                    help="only the given-start-abundances and hot-core models (ref_start_hotcore.npz)")
    ap.add_argument("--grid-variants", action="store_true",
                    help="add the answers at half and twice the relative tolerance to ref_cloud_grid.npz")
    ap.add_argument("--ref-lib", default=os.path.join(ROOT, "oracle", "_ref", "libucref.so"),
                    help="the reference build to run (--collapse-only)")
    a = ap.parse_args()
    only = None if a.only is None else [int(x) for x in a.only.split(",")]
    if a.collapse_only:
        collapse_fixture(a.procs, a.cap, a.ref_lib)
        return
    if a.start_hotcore_only:
        start_and_hotcore_fixture()
        return
    if a.grid_variants:
        grid_variants(a.procs, a.cap)
        return
    if a.grid_tight:
        grid_tight(a.procs, a.cap, only=only)
        return
    if a.alltypes_tight:
        grid_tight(a.procs, a.cap, "ref_alltypes.npz", "alltypes", only=only)
        return
    if a.coeff_only:
        coeff_fixture()
        return
    if a.synthetic_only:
        synthetic_fixtures(a.procs)
        return
    if a.alltypes_only:
        # static clouds on the all-types check network (net_alltypes.npz, written by
        # make_alltypes_network.py), through the reference compiled on it
        params = workloads.static_cloud_grid(6, seed=31, final_time=1e5)
        t0 = time.time()
        rc, final, stats, wall, nrows = run_models(params, a.procs, ["alltypes"] * len(params))
        np.savez_compressed(os.path.join(HERE, "ref_alltypes.npz"), params=params, rc=rc, final=final, stats=stats,
                            wall=wall, nrows=nrows)
        print(f"ref_alltypes: {len(params)} points in {time.time() - t0:.1f}s, rc {set(rc.tolist())}", flush=True)
        return
    if a.params_only:
        params = nondefault_params()
        t0 = time.time()
        rc, final, stats, wall, nrows = run_models(params, a.procs)
        np.savez_compressed(os.path.join(HERE, "ref_params.npz"), params=params, rc=rc, final=final, stats=stats,
                            wall=wall, nrows=nrows)
        print(f"ref_params: {len(params)} points in {time.time() - t0:.1f}s, rc {set(rc.tolist())}", flush=True)
        return
    if a.grid_only:
        t0 = time.time()
        idx, params, rc, final, stats, wall, nrows = grid4096_sample(a.grid, a.procs, a.cap)
        np.savez_compressed(os.path.join(HERE, "ref_cloud_grid.npz"), index=idx, params=params, rc=rc, final=final,
                            stats=stats, wall=wall, nrows=nrows)
        print(f"ref_cloud_grid: {len(params)} points of the 4096-point grid in {time.time() - t0:.1f}s, per-point "
              f"wall median {np.median(wall):.2f}s max {wall.max():.2f}s, rc {set(rc.tolist())}", flush=True)
        return
    # 1. the reference's own example output at t = 1e6 yr
    lines = open(REF_EXAMPLE).read().splitlines()
    hdr = [h.strip() for h in lines[0].split(",")]
    rows = np.array([[float(x) for x in ln.split(",")] for ln in lines[1:]])
    k = int(np.where(np.isclose(rows[:, 0], 1e6))[0][0])
    np.savez_compressed(os.path.join(HERE, "static_full_t1e6.npz"), species=np.array(hdr[8:], dtype="U16"),
                        abundances=rows[k, 8:], physics=rows[k, :8], times=rows[:k + 1, 0])
    example_models()
    if a.examples_only:
        return
    # 2. the reference model of the default static cloud point (the grid sample and the other
    # fixtures have flags of their own)
    for name, params in [("ref_cloud_point", workloads.static_cloud_point())]:
        t0 = time.time()
        rc, final, stats, wall, nrows = run_models(params, a.procs)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), params=params, rc=rc, final=final, stats=stats,
                            wall=wall, nrows=nrows)
        print(f"{name}: {len(params)} points in {time.time() - t0:.1f}s, per-point wall "
              f"median {np.median(wall):.2f}s max {wall.max():.2f}s, rc {set(rc.tolist())}", flush=True)


if __name__ == "__main__":
    main()
