"""Point-parameter vectors (the UCG_P_* layout of include/uclchem_gpu.h).

Defaults are UCLCHEM's (src/fortran_src/defaultparameters.f90:18-122) as the compiled
reference stores them: values written there as default-kind literals (rout = 0.05) are
float32-rounded.  Keys follow the reference's parameter-dictionary names
(case-insensitive, src/fortran_src/wrap.f90:699-983).
"""
from __future__ import annotations

from typing import Dict, Iterable

import numpy as np

NPARAM = 64
NAMES = ["initialdens", "initialtemp", "zeta", "radfield", "baseav", "rout", "rin", "finaltime",
         "freefall", "finaldens", "reltol", "abstol_factor", "abstol_min", "model", "tempindx",
         "maxtemp", "freezefactor", "endatfinaldensity", "freefallfactor", "mxstep", "maxsteps",
         # initial elemental abundances (defaultparameters.f90:70-90, applied at chemistry.f90:63-105)
         "metallicity", "ion", "fh", "fhe", "fc", "fo", "fn", "fs", "fmg", "fsi", "fcl", "fp", "ffe", "ff",
         "fd", "fli", "fna", "fpah", "f15n", "f13c", "f18o",
         # desorption and photon constants (defaultparameters.f90:113-120)
         "ebmaxh2", "ebmaxcr", "ebmaxuvcr", "epsilon", "uv_yield", "phi", "uvcreff", "omega",
         # behavioural switches (defaultparameters.f90:40-49)
         "desorb", "h2desorb", "crdesorb", "uvdesorb", "thermdesorb", "cosmicrayattenuation", "ionmodel",
         "improvedh2crpdissociation", "enforcechargeconservation",
         # the collapse model of model 2 (collapse.f90:1-7): 1 BE1.1, 2 BE4, 3 filament, 4 ambipolar
         "collapse"]
INDEX = {n: i for i, n in enumerate(NAMES)}
# the first slots, in the layout files written before the elemental and desorption slots existed
NPARAM_V1 = 24

_f32 = lambda x: float(np.float32(x))   # noqa: E731  (default-real literals of the reference)
DEFAULTS: Dict[str, float] = dict(
    initialdens=1.0e2, initialtemp=10.0, zeta=1.0, radfield=1.0, baseav=2.0,
    rout=_f32(0.05), rin=0.0, finaltime=5.0e6, freefall=0.0, finaldens=1.0e5,
    reltol=1e-8, abstol_factor=1e-14, abstol_min=1e-25, model=0.0, tempindx=1.0, maxtemp=300.0,
    freezefactor=1.0, endatfinaldensity=0.0, freefallfactor=1.0, mxstep=10000.0,
    maxsteps=0.0,   # total-step bound per point, 0 = unbounded (no counterpart in the reference)
    metallicity=1.0, ion=2.0, fh=0.5, fhe=_f32(0.1), fc=1.77e-04, fo=3.34e-04, fn=6.18e-05, fs=3.51e-6,
    fmg=2.256e-06, fsi=1.78e-06, fcl=3.39e-08, fp=7.78e-08, ffe=2.01e-7, ff=3.6e-08, fd=0.0, fli=0.0,
    fna=0.0, fpah=0.0, f15n=0.0, f13c=0.0, f18o=0.0,
    ebmaxh2=1.21e3, ebmaxcr=1.21e3, ebmaxuvcr=1.0e4, epsilon=_f32(0.01), uv_yield=_f32(0.03), phi=1.0e5,
    uvcreff=1.0e-3, omega=0.5,
    desorb=1.0, h2desorb=1.0, crdesorb=1.0, uvdesorb=1.0, thermdesorb=1.0, cosmicrayattenuation=0.0,
    ionmodel=0.0, improvedh2crpdissociation=0.0, enforcechargeconservation=0.0, collapse=0.0)

ALIASES = {"dens": "initialdens", "temp": "initialtemp", "t_end": "finaltime"}
# dictionaryParser quirks kept so a parameter dict means what it means to the reference:
# its "thermdesorb" branch sets uvdesorb (wrap.f90:802-803), and ionModel is a character
# ('L' or 'H'), stored here as 0 or 1
KEY_TARGET = {"thermdesorb": "uvdesorb"}
SECONDS_PER_YEAR = 3.16e7   # constants.f90:14
PC = 3.086e18               # constants.f90:11


def widen(params: np.ndarray) -> np.ndarray:
    """Parameter rows of the earlier 24-slot layout (its first 21 slots are unchanged)
    completed with the defaults of every later slot; rows already NPARAM wide pass through."""
    p = np.atleast_2d(np.asarray(params, dtype=np.float64))
    if p.shape[1] == NPARAM:
        return p.copy()
    if p.shape[1] != NPARAM_V1:
        raise ValueError(f"parameter rows must be {NPARAM} (or {NPARAM_V1}) wide, got {p.shape[1]}")
    out = np.tile(make_params(), (len(p), 1))
    out[:, :INDEX["maxsteps"] + 1] = p[:, :INDEX["maxsteps"] + 1]
    return out


def make_params(d: Dict[str, float] | None = None, **kw) -> np.ndarray:
    """One parameter vector from a UCLCHEM-style parameter dict."""
    vals = dict(DEFAULTS)
    merged = {}
    if d:
        merged.update(d)
    merged.update(kw)
    for k, v in merged.items():
        key = ALIASES.get(k.lower(), k.lower())
        key = KEY_TARGET.get(key, key)
        if key not in INDEX:
            raise KeyError(f"unknown or unsupported parameter {k!r}")
        if key == "ionmodel" and isinstance(v, str):
            if v.upper() not in ("L", "H"):
                raise ValueError("ionModel must be 'L' or 'H'")
            v = 0.0 if v.upper() == "L" else 1.0
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
