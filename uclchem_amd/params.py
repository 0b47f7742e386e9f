"""Point-parameter vectors (the UCG_P_* layout of include/uclchem_gpu.h).

Defaults are UCLCHEM's (src/fortran_src/defaultparameters.f90:18-122) as the compiled
reference stores them: values written there as default-kind literals (rout = 0.05) are
float32-rounded.  Keys follow the reference's parameter-dictionary names
(case-insensitive, src/fortran_src/wrap.f90:699-983).
"""
from __future__ import annotations

from typing import Dict, Iterable

import numpy as np

NPARAM = 24
NAMES = ["initialdens", "initialtemp", "zeta", "radfield", "baseav", "rout", "rin", "finaltime",
         "freefall", "finaldens", "reltol", "abstol_factor", "abstol_min", "model", "tempindx",
         "maxtemp", "freezefactor", "endatfinaldensity", "freefallfactor", "mxstep", "maxsteps"]
INDEX = {n: i for i, n in enumerate(NAMES)}

DEFAULTS: Dict[str, float] = dict(
    initialdens=1.0e2, initialtemp=10.0, zeta=1.0, radfield=1.0, baseav=2.0,
    rout=float(np.float32(0.05)), rin=0.0, finaltime=5.0e6, freefall=0.0, finaldens=1.0e5,
    reltol=1e-8, abstol_factor=1e-14, abstol_min=1e-25, model=0.0, tempindx=1.0, maxtemp=300.0,
    freezefactor=1.0, endatfinaldensity=0.0, freefallfactor=1.0, mxstep=10000.0,
    maxsteps=0.0)   # total-step bound per point, 0 = unbounded (no counterpart in the reference)

ALIASES = {"dens": "initialdens", "temp": "initialtemp", "t_end": "finaltime"}
SECONDS_PER_YEAR = 3.16e7   # constants.f90:14
PC = 3.086e18               # constants.f90:11


def make_params(d: Dict[str, float] | None = None, **kw) -> np.ndarray:
    """One parameter vector from a UCLCHEM-style parameter dict."""
    vals = dict(DEFAULTS)
    merged = {}
    if d:
        merged.update(d)
    merged.update(kw)
    for k, v in merged.items():
        key = ALIASES.get(k.lower(), k.lower())
        if key not in INDEX:
            raise KeyError(f"unknown or unsupported parameter {k!r}")
        if isinstance(v, bool):
            v = 1.0 if v else 0.0
        vals[key] = float(v)
    p = np.zeros(NPARAM)
    for n, i in INDEX.items():
        p[i] = vals[n]
    return p


def stack(dicts: Iterable[Dict[str, float]]) -> np.ndarray:
    return np.stack([make_params(d) for d in dicts])


def base_av_for(av: float, n_h: float, rout: float = DEFAULTS["rout"], rin: float = 0.0) -> float:
    """baseAv giving a point whose initial Av equals `av` (physics-core.f90:50,68-71)."""
    return av - (rout - rin) * PC * n_h / 1.6e21
