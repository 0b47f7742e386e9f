"""Write golden fixtures from the reference Fortran hot path (oracle/_ref/libucref.so).

Run in the build container (needs /root/reference for the build and the example output):
    python tests/golden/make_ref_fixtures.py [--grid 24] [--procs 8]

Outputs (small .npz files next to this script):
  static_full_t1e6.npz   examples/example-output/static-full.dat row t=1e6 yr (6 digits)
  ref_cloud_point.npz    the reference model for the default static cloud point
  ref_cloud_grid.npz     the reference model for the first N seeded grid points
  ref_hot_core.npz       the reference model for seeded hot-core points
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


def _run_one(p):
    from oracle import RefLib
    ref = RefLib()
    t0 = time.time()
    rc, rows, stats = ref.model(p, maxrec=2000)
    return rc, rows[-1].copy(), rows[:, 0].copy(), stats, time.time() - t0


def run_models(params, procs):
    with mp.get_context("fork").Pool(procs) as pool:
        out = pool.map(_run_one, list(params))
    rc = np.array([o[0] for o in out])
    final = np.stack([o[1] for o in out])
    stats = np.stack([o[3] for o in out])
    wall = np.array([o[4] for o in out])
    nrows = np.array([len(o[2]) for o in out])
    return rc, final, stats, wall, nrows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=24)
    ap.add_argument("--hot", type=int, default=4)
    ap.add_argument("--procs", type=int, default=8)
    a = ap.parse_args()
    # 1. the reference's own example output at t = 1e6 yr
    lines = open(REF_EXAMPLE).read().splitlines()
    hdr = [h.strip() for h in lines[0].split(",")]
    rows = np.array([[float(x) for x in ln.split(",")] for ln in lines[1:]])
    k = int(np.where(np.isclose(rows[:, 0], 1e6))[0][0])
    np.savez_compressed(os.path.join(HERE, "static_full_t1e6.npz"), species=np.array(hdr[8:], dtype="U16"),
                        abundances=rows[k, 8:], physics=rows[k, :8], times=rows[:k + 1, 0])
    # 2. reference models
    for name, params in [("ref_cloud_point", workloads.static_cloud_point()),
                         ("ref_cloud_grid", workloads.static_cloud_grid(a.grid)),
                         ("ref_hot_core", workloads.hot_core_grid(a.hot))]:
        t0 = time.time()
        rc, final, stats, wall, nrows = run_models(params, a.procs)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), params=params, rc=rc, final=final, stats=stats,
                            wall=wall, nrows=nrows)
        print(f"{name}: {len(params)} points in {time.time() - t0:.1f}s, per-point wall "
              f"median {np.median(wall):.2f}s max {wall.max():.2f}s, rc {set(rc.tolist())}", flush=True)


if __name__ == "__main__":
    main()
