"""The per-reaction terms of the RHS (the oracle's ucl_fluxes, the quantity
uclchem_amd.analysis.rates_to_dy_and_flux returns) pinned on the CPU: against the reference's
own definition of a reaction's flux in its analysis code (src/uclchem/analysis.py:783-806 with
reaction.py:93-103's count of density factors), restated here from the reaction tables, and
against the compiled reference's GETYDOT (gas-phase species: production minus loss of the
terms is ydot), at seeded states; plus the host bookkeeping of the analysis module."""
import numpy as np
import pytest

from helpers import random_state


def _reference_flux(tables, rate, y):
    """analysis.py:783-806: rate x density^body_count x reactant abundances x the type's factor."""
    from uclchem_amd import analysis
    from uclchem_amd import makerates_hip as mh
    a, names = tables.arrays, tables.species_names
    ns = tables.nspec
    D = y[-1]
    surface, bulk = y[tables.scalars["nsurface"]], y[tables.scalars["nbulk"]]
    blr = min(1.0, analysis.NUM_SITES_PER_GRAIN / (analysis.GAS_DUST_DENSITY_RATIO * bulk))
    total_swap = blr * analysis.get_total_swap(rate[None], y[None, :ns], tables)[0]
    out = np.zeros(tables.nreac)
    for k in range(tables.nreac):
        kind = mh.ALL_TYPES[int(a["rtype"][k])]
        sp = [int(j) for j in a["re"][k] if 0 <= j < ns]
        body = -1 + len(sp) + (1 if kind in ("DESOH2", "FREEZE") else 0) - (1 if kind in ("LH", "LHDES") else 0)
        f = rate[k] * D ** body
        for j in sp:
            f *= y[j]
        if kind in ("LH", "LHDES", "BULKSWAP") and any(names[j].startswith("@") for j in sp):
            f *= blr
        elif kind == "SURFSWAP":
            f *= total_swap / surface
        elif kind in ("DESCR", "DEUVCR", "ER", "ERDES"):
            f /= surface
        elif kind == "DESOH2":
            f *= y[tables.scalars["nh"]] / surface
        elif kind == "H2FORM":
            f /= y[tables.scalars["nh"]]
        out[k] = f
    return out


@pytest.mark.parametrize("seed", [3, 4])
def test_fluxes_follow_the_reference_definition_and_its_rhs(reflib, oracle, tables, seed):
    from oracle import make_params
    rng = np.random.default_rng(seed)
    ns = tables.nspec
    for trial in range(4):
        p = make_params(dens=10 ** rng.uniform(3, 7), temp=rng.uniform(10, 100), finaltime=1e6)
        reflib.init(p)
        pt, _ = oracle.init_point(p)
        _, y = random_state(tables, oracle, rng, p)
        y[tables.scalars["nsurface"]] = y[tables.surface_list].sum()   # the ice totals as F sees them
        y[tables.scalars["nbulk"]] = y[tables.bulk_list].sum()
        rr, yr, _ = reflib.rates(y)
        flux, rate = oracle.fluxes(pt, rr, yr)
        ref = _reference_flux(tables, rate, yr)
        scale = np.abs(ref) + 1e-300
        assert np.all(np.abs(flux - ref) <= 1e-12 * scale), np.max(np.abs(flux - ref) / scale)
        # production minus loss of the terms is the compiled reference's GETYDOT on gas species
        ydot = reflib.rhs(yr)
        from uclchem_amd import analysis
        pre = flux @ analysis.construct_incidence(tables).astype(np.float64)
        gas = [i for i in range(ns) if not tables.species_names[i][0] in "#@" and i not in
               (tables.scalars["nsurface"], tables.scalars["nbulk"])]
        mag = np.abs(flux) @ np.abs(analysis.construct_incidence(tables)).astype(np.float64)
        assert np.all(np.abs(pre[gas] - ydot[gas]) <= 1e-12 * mag[gas] + 1e-300)


def test_incidence_names_and_swap_pairs(tables):
    from uclchem_amd import analysis
    inc = analysis.construct_incidence(tables)
    a = tables.arrays
    k = 0
    assert inc[k, a["re"][k][0]] == -1 and inc[k, a["pr"][k][0]] == 1
    names = analysis.flux_names(tables)
    pairs = analysis._swap_pairs(tables)
    from uclchem_amd import makerates_hip as mh
    kinds = [mh.ALL_TYPES[int(c)] for c in a["rtype"]]
    # the reference pairs the two sorted lists with zip (analysis.py:826): as many pairs as the
    # shorter list holds (the default network has one BULKSWAP more than SURFSWAP reactions)
    assert len(pairs) == min(kinds.count("BULKSWAP"), kinds.count("SURFSWAP")) > 0
    assert len(names) == tables.nreac + 2 * len(pairs)
    assert all(int(a["re"][b][0]) < int(a["re"][c][0]) for (b, _), (c, _) in zip(pairs, pairs[1:]))
    assert all("BULKSWAP_GEOMETRIC" in n for n in names[tables.nreac:tables.nreac + len(pairs)])
    # production / destruction columns from the reaction strings, as the reference reads them
    flux = np.arange(len(names), dtype=float)[None]
    prod, dest, pn, dn = analysis.get_production_and_destruction("CO", flux, names)
    assert pn and dn and all("CO" in n.split(" -> ")[-1].split(" + ") for n in pn)
    assert all("CO" in n.split(" -> ")[0].split(" + ") for n in dn)
    assert prod.shape == (1, len(pn)) and dest.shape == (1, len(dn))
