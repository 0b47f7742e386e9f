"""Seeded grid-point workloads (parameter vectors) for tests and the benchmark.

* ``static_cloud_point``: one static cloud on the default network, n_H=1e4, T=10 K,
  zeta=1, G0=1, default Av (2.9644 mag from baseAv=2 and the 0.05 pc slab), to 1 Myr.
* ``static_cloud_grid``: n points, log n_H ~ U[2,7], T ~ U[10,100] K, log zeta ~ U[0,2],
  log G0 ~ U[0,3], Av ~ U[1,20] mapped through baseAv = Av - rout*pc*n_H/1.6e21 so the
  point's Av is the sampled one (physics-core.f90:50,68-71); all to 1 Myr.
* ``hot_core_grid``: n points, log n_H ~ U[5,8], hot_core(temp_indx=4, max_temp=300 K).
* ``collapse_grid``: free-fall collapse from n_H,0 log-uniform in [1e2,1e6] cm^-3 towards
  finalDens=1e6 (densdot is zero once capped, physics-core.f90:97), T=10 K, to 1 Myr.
Random draws use numpy's PCG64 with the seed given (default 20261202).
"""
from __future__ import annotations

import numpy as np

from .params import NPARAM, base_av_for, make_params

SEED = 20261202


def static_cloud_point(final_time: float = 1e6) -> np.ndarray:
    return make_params(initialdens=1e4, initialtemp=10.0, zeta=1.0, radfield=1.0, finaltime=final_time)[None, :]


def static_cloud_grid(n: int, seed: int = SEED, final_time: float = 1e6) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    lognh = rng.uniform(2.0, 7.0, n)
    temp = rng.uniform(10.0, 100.0, n)
    logzeta = rng.uniform(0.0, 2.0, n)
    logg0 = rng.uniform(0.0, 3.0, n)
    av = rng.uniform(1.0, 20.0, n)
    out = np.zeros((n, NPARAM))
    for i in range(n):
        nh = 10.0 ** lognh[i]
        out[i] = make_params(initialdens=nh, initialtemp=temp[i], zeta=10.0 ** logzeta[i],
                             radfield=10.0 ** logg0[i], baseav=base_av_for(av[i], nh), finaltime=final_time)
    return out


def hot_core_grid(n: int, seed: int = SEED, final_time: float = 1e6) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed + 5))
    lognh = rng.uniform(5.0, 8.0, n)
    out = np.zeros((n, NPARAM))
    for i in range(n):
        out[i] = make_params(initialdens=10.0 ** lognh[i], initialtemp=10.0, model=1, tempindx=4, maxtemp=300.0,
                             finaltime=final_time)
    return out


def collapse_grid(n: int, seed: int = SEED, final_time: float = 1e6) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed + 3))
    lognh = rng.uniform(2.0, 6.0, n)
    out = np.zeros((n, NPARAM))
    for i in range(n):
        out[i] = make_params(initialdens=10.0 ** lognh[i], initialtemp=10.0, freefall=1, finaldens=1e6,
                             finaltime=final_time)
    return out
