"""Per-phase device timing of the batched solve (ucg_solve_batch_profiled).

    python -m uclchem_amd.phase_profile [--points 8] [--network default]

Solves the first `points` points of the seeded static-cloud grid with the profiled entry
point, which stamps the device wall clock (100 MHz) around every phase inside the solve
kernel, and prints per-step and per-call times: RHS, Jacobian assembly, factorisation
(sparse part / dense tail), Newton solves (sparse part / dense tail), for the FP64 and FP32
factorisations.
"""
from __future__ import annotations

import argparse
import time

from . import gpu, workloads
from . import makerates_hip as mh


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=8)
    ap.add_argument("--network", default="default")
    a = ap.parse_args()
    net = gpu.Network(mh.load_network(mh.network_path(a.network)))
    grid = workloads.static_cloud_grid(a.points)
    for bits in (64, 32):
        t0 = time.time()
        out = net.solve_profiled(grid, lu_bits=bits)
        tm, st = out["timers"], out["stats"]
        S, T = st.sum(0), tm.sum(0)
        print(f"lu{bits}: wall {time.time() - t0:.2f}s, shader MHz {tm[:, 6].mean():.0f}, steps {S[0]}, "
              f"rhs {S[1]}, setups {S[2]}, newton iterations {S[3]}", flush=True)
        print(f"  ms/step {1e3 * T[5] / S[0]:.3f}; per call us: rhs {1e6 * T[1] / S[1]:.1f} "
              f"jacobian {1e6 * T[2] / S[2]:.1f} factorisation {1e6 * T[3] / S[2]:.1f} "
              f"(sparse {1e6 * T[7] / S[2]:.1f}, tail {1e6 * T[8] / S[2]:.1f}) solve {1e6 * T[4] / S[3]:.1f} "
              f"(sparse {1e6 * T[9] / S[3]:.1f}, tail {1e6 * T[10] / S[3]:.1f}); "
              f"other per step us {1e6 * (T[5] - T[:5].sum()) / S[0]:.1f}", flush=True)


if __name__ == "__main__":
    main()
