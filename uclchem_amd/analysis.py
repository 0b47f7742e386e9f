"""Batched counterparts of the reference's rate analysis (src/uclchem/analysis.py:691-890).

``rates_to_dy_and_flux`` rebuilds, from a model's output rows (physics, abundances and the
rate coefficients of every row, ``return_rates=True``), the right-hand side the integrator
solved (GETYDOT) and the individual reaction terms that make it up.  The reference does this
on the host, row by row, with pandas; here every row of every point is one workgroup of the
RHS kernel (``ucg_rhs_fluxes``), which returns GETYDOT and its per-reaction terms, and the
reference's bookkeeping of the bulk<->surface transfer (the "SWAP_GEOMETRIC" pseudo-reactions,
analysis.py:813-880) is added on the host from those terms.

Differences kept on purpose: the reference multiplies the rates as written in the output
row; the kernel recomputes the H2 and CO photodissociation rates at the row's abundances, as
F does (chemistry.f90:311-319), and clamps SURFACE and BULK below at 1e-30 as F does.  The
reference applies its geometric swap terms to dy with the opposite sign to its own GETYDOT's
bulk<->surface transfer (odes.f90's transfer terms: with the surface shrinking, the bulk species
lose what their surface partners gain; analysis.py:846-870 moves it the other way), so its dy
on the ice species is not the F it says it rebuilds; ``dy`` here is GETYDOT, and
``reference_signs=True`` returns the reference function's own dy instead (the flux columns are
the same either way).
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

from . import makerates_hip as mh
from .params import INDEX, PC, make_params

# surfacereactions.f90:38,49-51 (the default-real PI, as compiled)
_PI_F = float(np.float32(3.141592654))
NUM_SITES_PER_GRAIN = 1.0e-5 * 1.0e-5 * 1.5e15 * 4.0 * _PI_F
GAS_DUST_DENSITY_RATIO = (4.0 * _PI_F * (1e-5 ** 3) * 3.0 * 100.0) / (3.0 * 1.66053892e-24)


def construct_incidence(tables: mh.NetworkTables) -> np.ndarray:
    """[nreac, nspec] int8: +1 per product, -1 per reactant (analysis.py:711-734)."""
    a = tables.arrays
    inc = np.zeros((tables.nreac, tables.nspec), dtype=np.int8)
    for k in range(tables.nreac):
        for j in a["pr"][k]:
            if 0 <= j < tables.nspec:
                inc[k, j] += 1
        for j in a["re"][k]:
            if 0 <= j < tables.nspec:
                inc[k, j] -= 1
    return inc


def _swap_pairs(tables: mh.NetworkTables):
    """BULKSWAP and SURFSWAP reactions, each list ordered by its reactant's species index,
    paired in that order (analysis.py:815-826: zip, so as many pairs as the shorter list)."""
    a = tables.arrays
    kinds = [mh.ALL_TYPES[int(c)] for c in a["rtype"]]
    bulk = sorted([k for k in range(tables.nreac) if kinds[k] == "BULKSWAP"], key=lambda k: int(a["re"][k][0]))
    surf = sorted([k for k in range(tables.nreac) if kinds[k] == "SURFSWAP"], key=lambda k: int(a["re"][k][0]))
    return list(zip(bulk, surf))


def get_total_swap(rates: np.ndarray, abundances: np.ndarray, tables: mh.NetworkTables) -> np.ndarray:
    """sum over BULKSWAP reactions of rate x abundance of the reactant, per row
    (analysis.py:691-708); rates [..., nreac], abundances [..., nspec]."""
    a = tables.arrays
    out = np.zeros(rates.shape[:-1])
    for k in range(tables.nreac):
        if mh.ALL_TYPES[int(a["rtype"][k])] == "BULKSWAP":
            out = out + rates[..., k] * abundances[..., int(a["re"][k][0])]
    return out


def flux_names(tables: mh.NetworkTables) -> List[str]:
    """Column names of the flux array in the reference's reaction-string form ("A + B + TYPE ->
    C + D", analysis.py's column names): one per reaction, then the geometric swap terms (the
    swap reactions' strings with SWAP replaced by SWAP_GEOMETRIC: every BULKSWAP pair first,
    then every SURFSWAP pair)."""
    a, names = tables.arrays, tables.species_names
    kinds = [mh.ALL_TYPES[int(c)] for c in a["rtype"]]

    def text(k, kind):
        re = [names[j] for j in a["re"][k] if 0 <= j < tables.nspec]
        pr = [names[j] for j in a["pr"][k] if 0 <= j < tables.nspec]
        return " + ".join(re + ([kind] if kind != "TWOBODY" else [])) + " -> " + " + ".join(pr)

    cols = [text(k, kinds[k]) for k in range(tables.nreac)]
    pairs = _swap_pairs(tables)
    cols += [text(b, "BULKSWAP_GEOMETRIC") for b, _ in pairs]
    cols += [text(s, "SURFSWAP_GEOMETRIC") for _, s in pairs]
    return cols


def rates_to_dy_and_flux(physics: np.ndarray, abundances: np.ndarray, rates: np.ndarray, param_dicts=None,
                         network_name: str = "default", net=None,
                         reference_signs: bool = False) -> Tuple[np.ndarray, np.ndarray, List[str]]:
    """The reference's ``rates_to_dy_and_flux`` for output arrays in the layouts the batched model
    API returns: physics [rows, points, 8], abundances [rows, points, nspec], rates [rows, points,
    nreac] (``cloud_grid(..., return_array=True, return_rates=True)``), with the parameter
    dictionaries the points ran with (radfield and the cloud size enter the photo rates; None:
    defaults).  Returns (dy [rows, points, nspec], flux [rows, points, nreac + 2 x swap pairs],
    flux column names): dy is GETYDOT at each row (with reference_signs, the reference
    function's own dy: production minus loss of the terms with its swap corrections), flux its
    reaction terms followed by the geometric swap terms the reference appends
    (analysis.py:826-880)."""
    from . import model
    physics = np.asarray(physics, dtype=np.float64)
    abundances = np.asarray(abundances, dtype=np.float64)
    rates = np.asarray(rates, dtype=np.float64)
    if physics.ndim != 3 or abundances.shape[:2] != physics.shape[:2] or rates.shape[:2] != physics.shape[:2]:
        raise ValueError("physics, abundances and rates must be [rows, points, ...] with the same rows and points")
    rows, npt = physics.shape[:2]
    net = net or model.network(network_name)
    tables = net.t
    if abundances.shape[2] != tables.nspec or rates.shape[2] != tables.nreac:
        raise ValueError(f"abundances need {tables.nspec} species and rates {tables.nreac} reactions per row")
    dicts = list(param_dicts) if param_dicts is not None else [{}] * npt
    if len(dicts) != npt:
        raise ValueError("one parameter dictionary per point")
    base = np.stack([make_params(d) for d in dicts])
    n = rows * npt
    params = np.repeat(base[None], rows, axis=0).reshape(n, -1)
    ph = physics.reshape(n, -1)
    phys = np.zeros((n, 8))
    phys[:, 0] = ph[:, 1]                     # density
    phys[:, 1] = ph[:, 2]                     # gasTemp
    phys[:, 2] = ph[:, 3]                     # dustTemp
    phys[:, 3] = ph[:, 4]                     # Av
    phys[:, 4] = (params[:, INDEX["rout"]] - params[:, INDEX["rin"]]) * PC
    phys[:, 5] = params[:, INDEX["phi"]]
    phys[:, 6] = params[:, INDEX["abstol_factor"]]
    y = np.concatenate([abundances.reshape(n, -1), ph[:, 1:2]], axis=1)
    ydot, flux = net.rhs_fluxes(params, phys, rates.reshape(n, -1), y)
    # the reference's geometric swap terms, from the reaction terms before the transfer
    inc = construct_incidence(tables).astype(np.float64)
    a = tables.arrays
    ab = abundances.reshape(n, -1)
    pre = flux @ inc                                            # dy before the bulk<->surface transfer
    surface = pre[:, np.asarray(a["surface_list"], dtype=int)].sum(axis=1)
    ysurf, ybulk = ab[:, tables.scalars["nsurface"]], ab[:, tables.scalars["nbulk"]]
    pairs = _swap_pairs(tables)
    geo = np.zeros((n, 2 * len(pairs)))
    cover_up = 0.5 * GAS_DUST_DENSITY_RATIO / NUM_SITES_PER_GRAIN
    with np.errstate(divide="ignore", invalid="ignore"):
        for i, (b, s) in enumerate(pairs):
            shrink = surface < 0.0
            cov = np.minimum(1.0, ybulk / ysurf)
            geo[:, i] = np.where(shrink, surface * cov * ab[:, int(a["re"][b][0])] / ybulk, 0.0)
            geo[:, len(pairs) + i] = np.where(shrink, 0.0, surface * cover_up * ab[:, int(a["re"][s][0])])
    if reference_signs:   # analysis.py:846-870 and 884-885, as the reference applies them
        d = pre.copy()
        for i, (b, s) in enumerate(pairs):
            rb, pb = int(a["re"][b][0]), int(a["pr"][b][0])
            rs, ps = int(a["re"][s][0]), int(a["pr"][s][0])
            d[:, rb] -= geo[:, i]
            d[:, pb] += geo[:, i]
            d[:, ps] -= geo[:, len(pairs) + i]
            d[:, rs] += geo[:, len(pairs) + i]
        names = tables.species_names
        d[:, tables.scalars["nsurface"]] = d[:, [j for j in range(tables.nspec) if names[j].startswith("#")]].sum(axis=1)
        d[:, tables.scalars["nbulk"]] = d[:, [j for j in range(tables.nspec) if names[j].startswith("@")]].sum(axis=1)
        dy = d.reshape(rows, npt, tables.nspec)
    else:
        dy = ydot[:, :tables.nspec].reshape(rows, npt, tables.nspec)
    out = np.concatenate([flux, geo], axis=1).reshape(rows, npt, -1)
    return dy, out, flux_names(tables)


def get_production_and_destruction(species: str, flux: np.ndarray, names: Sequence[str]):
    """The columns of `flux` (or of a rate array) whose reaction produces `species` and those
    whose reaction destroys it, read from the column names as the reference does
    (analysis.py:893-902): (production [..., p], destruction [..., d], production names,
    destruction names)."""
    names = [str(r).strip() for r in names]
    dest = [i for i, r in enumerate(names) if species in r.split(" -> ")[0].split(" + ")]
    prod = [i for i, r in enumerate(names) if species in r.split(" -> ")[-1].split(" + ")]
    return flux[..., prod], flux[..., dest], [names[i] for i in prod], [names[i] for i in dest]
