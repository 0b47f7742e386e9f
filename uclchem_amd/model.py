"""Batched counterparts of the reference's Python model API on the GPU library.

The reference runs one grid point per call: ``uclchem.model.cloud`` / ``hot_core``
(``src/uclchem/model.py:277-316``) call ``uclchemwrap.cloud`` / ``hot_core``
(``src/fortran_src/wrap.f90:15-217``).  The functions here take a list of parameter
dictionaries (the same case-insensitive keys) and solve all of them in one launch (the
densest points taken first, which shortens heavy-tailed grids), returning results in the
reference's own layouts:

* ``return_array=False``: per point, ``[success_flag] + abundances of out_species``
  (``_format_output``, ``model.py:71-76``);
The physics drivers are the reference's cloud, hot core (``hot_core_grid``) and collapse
(``collapse_grid``: the BE1.1, BE4, filament and ambipolar models of ``collapse.f90``).

* ``return_array=True``: ``(physicsArray, chemicalAbunArray, ratesArray, abundanceStart,
  success_flags)`` with ``physicsArray[timepoints+1, points, 8]`` (Time, Density, gasTemp,
  dustTemp, Av, radfield, zeta, dstep; ``io.f90:74-82``) and
  ``chemicalAbunArray[timepoints+1, points, nspec]``, trimmed like ``_array_clean``
  (``model.py:158-190``) to the last written row; ``ratesArray[rows, points, nreac]`` (the
  rate coefficients as each row is written, ``io.f90:90-93``) with ``return_rates=True``,
  else None, as the reference returns it.

A parameter dictionary may carry the reference's ``alpha``, ``beta`` and ``gamma`` keys
(1-based reaction index -> coefficient, ``wrap.f90:744-761``): points with the same changes
run together on a copy of the network with those coefficients
(``makerates_hip.with_coefficients``).  The reference keeps such a change in its module
arrays for later runs in the process; here it applies to the dictionary that carries it.

``get_rates`` / ``get_odes`` restate ``wrap.f90:446-547``: parse the dictionary, start the
physics, take the given abundances, run the first update (rates at the given state, one
1e-7-year interval), then return the rate coefficients and F at the updated state.
"""
from __future__ import annotations

from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np

from . import gpu
from . import makerates_hip as mh
from .params import INDEX, PC, make_params

TIMEPOINTS = 500          # the reference's default row budget (constants.py:22)
N_PHYSICAL_PARAMETERS = 8

_nets: Dict[str, gpu.Network] = {}


def network(name: str = "default") -> gpu.Network:
    """The resident network of that name (loaded once per process)."""
    if name not in _nets:
        _nets[name] = gpu.Network(mh.load_network(mh.network_path(name)))
    return _nets[name]


_COEFF_KEYS = ("alpha", "beta", "gamma")
_derived: Dict[tuple, gpu.Network] = {}


def _coefficient_network(name: str, changes: tuple) -> gpu.Network:
    """The named network with reaction coefficients replaced (cached per change set)."""
    if not changes:
        return network(name)
    key = (name, changes)
    if key not in _derived:
        ch = {k: dict(v) for k, v in changes}
        t = mh.with_coefficients(network(name).t, **ch)
        _derived[key] = gpu.Network(t)
    return _derived[key]


def _split_dict(d: Optional[Dict]):
    """(parameter dict without coefficient keys, hashable coefficient change set)."""
    rest, changes = {}, []
    for k, v in dict(d or {}).items():
        if str(k).lower() in _COEFF_KEYS:
            items = tuple(sorted((int(i), float(x)) for i, x in dict(v).items()))
            if items:
                changes.append((str(k).lower(), items))
        else:
            rest[k] = v
    return rest, tuple(sorted(changes))


def _arrays(net: gpu.Network, params: np.ndarray, out, timepoints: int):
    n = params.shape[0]
    rows = int(out["n_rec"].max()) if n else 0
    phys = np.zeros((timepoints + 1, n, N_PHYSICAL_PARAMETERS))
    chem = np.zeros((timepoints + 1, n, net.nspec))
    for q in range(n):
        k = int(out["n_rec"][q])
        rec = out["rec"][q, :k]
        phys[:k, q, 0] = rec[:, 0]
        phys[:k, q, 1] = rec[:, 1]
        phys[:k, q, 2] = rec[:, 2]
        phys[:k, q, 3] = rec[:, 2]          # dustTemp = gasTemp for cloud and hot-core models
        phys[:k, q, 4] = rec[:, 3]
        phys[:k, q, 5] = params[q, INDEX["radfield"]]
        phys[:k, q, 6] = params[q, INDEX["zeta"]]
        phys[:k, q, 7] = 1
        chem[:k, q, :] = rec[:, 4:]
    last = np.maximum(out["n_rec"] - 1, 0)
    start = np.stack([chem[last[q], q, :] for q in range(n)]) if n else np.zeros((0, net.nspec))
    rates = None
    if "rates" in out:
        rates = np.zeros((rows, n, net.nreac))
        for q in range(n):
            k = int(out["n_rec"][q])
            rates[:k, q, :] = out["rates"][q, :k]
    return phys[:rows], chem[:rows], start, rates


def _run(param_dicts: Iterable[Dict], out_species, return_array, starting_chemistry, timepoints, network_name,
         lu_bits, extra: Dict, return_rates: bool = False):
    net = network(network_name)
    split = [_split_dict(d) for d in param_dicts]
    params = np.stack([make_params(d, **extra) for d, _ in split])
    y0 = None
    if starting_chemistry is not None:
        sc = np.atleast_2d(np.asarray(starting_chemistry, dtype=np.float64))
        if sc.shape[1] != net.nspec:
            raise ValueError(f"starting_chemistry needs {net.nspec} abundances per point")
        if sc.shape[0] == 1 and params.shape[0] > 1:
            sc = np.repeat(sc, params.shape[0], axis=0)
        y0 = np.concatenate([sc, params[:, INDEX["initialdens"]][:, None]], axis=1)
    want_rates = bool(return_rates and return_array)
    groups: Dict[tuple, List[int]] = {}
    for q, (_, ch) in enumerate(split):
        groups.setdefault(ch, []).append(q)
    out = None
    for ch, members in groups.items():   # one launch per coefficient change set
        m = np.array(members)
        part = _coefficient_network(network_name, ch).solve(
            params[m], y_init=None if y0 is None else y0[m], max_rec=timepoints + 1, lu_bits=lu_bits,
            order="densest", return_rates=want_rates)
        if out is None:
            out = {k: (None if v is None else np.zeros((len(params),) + v.shape[1:], v.dtype)) for k, v in part.items()}
        for k, v in part.items():
            if v is not None:
                out[k][m] = v
    flags = out["status"].astype(int)
    if return_array:
        phys, chem, start, rates = _arrays(net, params, out, timepoints)
        return phys, chem, rates, start, flags
    names = net.t.species_names
    idx = [names.index(s) for s in (out_species or [])]
    res = []
    for q in range(params.shape[0]):
        if flags[q] < 0 or not idx:
            res.append([int(flags[q])])
        else:
            res.append([int(flags[q])] + [float(out["y"][q, i]) for i in idx])
    return res


def cloud_grid(param_dicts: Sequence[Dict], out_species: Optional[List[str]] = None, return_array: bool = False,
               starting_chemistry=None, timepoints: int = TIMEPOINTS, network_name: str = "default",
               lu_bits: int = 64, return_rates: bool = False):
    """``uclchem.model.cloud`` for every dict of `param_dicts` in one GPU launch."""
    return _run(param_dicts, out_species, return_array, starting_chemistry, timepoints, network_name, lu_bits, {},
                return_rates)


def hot_core_grid(temp_indx: int, max_temperature: float, param_dicts: Sequence[Dict],
                  out_species: Optional[List[str]] = None, return_array: bool = False, starting_chemistry=None,
                  timepoints: int = TIMEPOINTS, network_name: str = "default", lu_bits: int = 64,
                  return_rates: bool = False):
    """``uclchem.model.hot_core(temp_indx, max_temperature, ...)`` for every dict in one launch."""
    if not 1 <= int(temp_indx) <= 6:
        raise ValueError("temp_indx must be 1..6 (hotcore.f90:21-22)")
    return _run(param_dicts, out_species, return_array, starting_chemistry, timepoints, network_name, lu_bits,
                dict(model=1, tempindx=int(temp_indx), maxtemp=float(max_temperature)), return_rates)


COLLAPSE_MODELS = {"BE1.1": 1, "BE4": 2, "filament": 3, "ambipolar": 4}   # model.py:355


def collapse_grid(collapse: str, param_dicts: Sequence[Dict], out_species: Optional[List[str]] = None,
                  return_array: bool = False, starting_chemistry=None, timepoints: int = TIMEPOINTS,
                  network_name: str = "default", lu_bits: int = 64, return_rates: bool = False):
    """``uclchem.model.collapse(collapse, physics_output, ...)`` (model.py:319-426) for every dict
    in one launch: `collapse` is 'BE1.1', 'BE4', 'filament' or 'ambipolar', the parameterised
    prestellar-core collapses of collapse.f90.  BE1.1 and BE4 end at 0.97 of their fitted
    collapse time (1.14 Myr and 0.18 Myr, 1144 and 184 output rows) whatever finalTime says, so
    `timepoints` must be at least that to return every row (else status -6, as the
    reference's).  The reference's physics_output file is not written."""
    if collapse not in COLLAPSE_MODELS:
        raise ValueError("collapse must be one of 'BE1.1', 'BE4', 'filament', or 'ambipolar'")
    return _run(param_dicts, out_species, return_array, starting_chemistry, timepoints, network_name, lu_bits,
                dict(model=2, collapse=COLLAPSE_MODELS[collapse]), return_rates)


def _initial_phys(params: np.ndarray) -> np.ndarray:
    """[density, Tgas, Tdust, Av, cloudSize, phi, abstol_factor, 0] at t = 0
    (physics-core.f90:42-73, cloud.f90:20-31)."""
    out = np.zeros((params.shape[0], 8))
    for q, p in enumerate(params):
        cs = (p[INDEX["rout"]] - p[INDEX["rin"]]) * PC
        dens = p[INDEX["initialdens"]] * (float(np.float32(1.001)) if p[INDEX["freefall"]] else 1.0)
        av = p[INDEX["baseav"]] + cs * p[INDEX["initialdens"]] / 1.6e21
        out[q] = [dens, p[INDEX["initialtemp"]], p[INDEX["initialtemp"]], av, cs, 1.0e5, p[INDEX["abstol_factor"]], 0.0]
    return out


def get_rates_and_odes(param_dicts: Sequence[Dict], abundances, network_name: str = "default"):
    """Batched ``get_rates`` + ``get_odes`` (wrap.f90:446-547): returns (rates [n, nreac],
    odes [n, neq]) -- the rate coefficients of the first update at the given abundances and
    F evaluated after that 1e-7-year update."""
    net = network(network_name)
    params = np.stack([make_params(dict(d or {}), finaltime=1e-7) for d in param_dicts])
    ab = np.atleast_2d(np.asarray(abundances, dtype=np.float64))
    if ab.shape != (params.shape[0], net.nspec):
        raise ValueError(f"abundances must be [points, {net.nspec}]")
    dens = params[:, INDEX["initialdens"]][:, None]
    y0 = np.concatenate([ab, dens], axis=1)
    phys = _initial_phys(params)
    rate, _, _ = net.rates(params, phys, y0)
    out = net.solve(params, y_init=y0)
    y1 = out["y"]
    odes, _ = net.rhs(params, phys, rate, y1)
    return rate, odes
