"""GPU parity of the per-reaction RHS terms (ucg_rhs_fluxes) and of the batched analysis built on
them (uclchem_amd.analysis.rates_to_dy_and_flux, the reference's analysis.py:737-890): the
kernel's terms against the oracle's (pinned on the CPU against the reference's own flux
definition and its compiled GETYDOT, tests/test_analysis_host.py) within 1e-12 of each term,
and, on a model's own output rows, that the analysis returns GETYDOT and that the reference's
geometric swap terms, applied to the terms' production minus loss, rebuild GETYDOT's
bulk<->surface transfer."""
import numpy as np
import pytest

from helpers import phys_of, random_state

pytestmark = pytest.mark.gpu


def test_fluxes_match_oracle(gpu_net, tables, oracle):
    from uclchem_amd.params import make_params
    rng = np.random.default_rng(11)
    P, PH, R, Y, FL = [], [], [], [], []
    for trial in range(8):
        p = make_params(initialdens=10 ** rng.uniform(2, 8), initialtemp=rng.uniform(10, 300),
                        radfield=10 ** rng.uniform(0, 3), baseav=rng.uniform(0, 10), finaltime=1e6)
        pt, y = random_state(tables, oracle, rng, p, bare_grains=(trial % 3 == 0))
        r, yr, _ = oracle.rates(pt, y)
        fl, _ = oracle.fluxes(pt, r, yr)
        P.append(p)
        PH.append(phys_of(pt))
        R.append(r)
        Y.append(yr)
        FL.append(fl)
    ydot, flux = gpu_net.rhs_fluxes(np.array(P), np.array(PH), np.array(R), np.array(Y))
    ref = np.array(FL)
    err = np.abs(flux - ref)
    assert np.all(err <= 1e-12 * np.abs(ref) + 1e-300), float((err / (np.abs(ref) + 1e-300)).max())
    ydot2, _ = gpu_net.rhs(np.array(P), np.array(PH), np.array(R), np.array(Y))
    assert np.array_equal(ydot, ydot2)          # the same kernel, the same F


def test_rates_to_dy_and_flux_on_model_output(gpu_net, tables):
    from uclchem_amd import analysis, model
    dicts = [{"initialDens": 1e4, "finalTime": 1e5}, {"initialDens": 3e5, "initialTemp": 15.0, "finalTime": 1e5}]
    phys, chem, rates, _, flags = model.cloud_grid(dicts, return_array=True, return_rates=True)
    assert list(flags) == [0, 0] and rates is not None
    rows = slice(1, None)                        # row 0 is written before any rates exist
    dy, flux, names = analysis.rates_to_dy_and_flux(phys[rows], chem[rows], rates[rows], dicts, net=gpu_net)
    ns, nr = tables.nspec, tables.nreac
    pairs = analysis._swap_pairs(tables)
    assert dy.shape == chem[rows].shape and flux.shape == chem[rows].shape[:2] + (nr + 2 * len(pairs),)
    assert len(names) == flux.shape[2]
    # dy is F at each row: the RHS kernel on the same state
    n = dy.shape[0] * dy.shape[1]
    from uclchem_amd.params import INDEX, PC, make_params
    base = np.stack([make_params(d) for d in dicts])
    params = np.repeat(base[None], dy.shape[0], axis=0).reshape(n, -1)
    ph = phys[rows].reshape(n, -1)
    phv = np.zeros((n, 8))
    phv[:, :4] = ph[:, 1:5]
    phv[:, 4] = (params[:, INDEX["rout"]] - params[:, INDEX["rin"]]) * PC
    phv[:, 5], phv[:, 6] = params[:, INDEX["phi"]], params[:, INDEX["abstol_factor"]]
    y = np.concatenate([chem[rows].reshape(n, -1), ph[:, 1:2]], axis=1)
    ydot, _ = gpu_net.rhs(params, phv, rates[rows].reshape(n, -1), y)
    assert np.array_equal(dy.reshape(n, -1), ydot[:, :ns])
    # the geometric swap terms, applied with the sign of F's transfer (the bulk species lose what
    # their surface partners gain while the surface shrinks, odes.f90's transfer terms), rebuild
    # F on every ice species the reference's pairing covers; the reference's own dy applies
    # them the other way round (analysis.py:846-870)
    fl = flux.reshape(n, -1)
    inc = analysis.construct_incidence(tables).astype(np.float64)
    pre = fl[:, :nr] @ inc
    a = tables.arrays
    corr = np.zeros_like(pre)
    touched = set()
    for i, (b, s) in enumerate(pairs):
        rb, pb, rs, ps = int(a["re"][b][0]), int(a["pr"][b][0]), int(a["re"][s][0]), int(a["pr"][s][0])
        assert not np.any((fl[:, nr + i] != 0) & (fl[:, nr + len(pairs) + i] != 0))
        corr[:, rb] += fl[:, nr + i]
        corr[:, pb] -= fl[:, nr + i]
        corr[:, rs] -= fl[:, nr + len(pairs) + i]
        corr[:, ps] += fl[:, nr + len(pairs) + i]
        touched |= {rb, pb, rs, ps}
    ice = sorted(j for j in touched if j in set(np.flatnonzero(np.abs(corr).sum(axis=0))))
    assert ice
    scale = np.abs(fl[:, :nr]) @ np.abs(inc) + np.abs(corr)
    unpaired = [int(a["re"][k][0]) for k in range(nr) if tables.species_names[int(a["re"][k][0])].startswith("@")
                and k not in {b for b, _ in pairs} and analysis.mh.ALL_TYPES[int(a["rtype"][k])] == "BULKSWAP"]
    check = [j for j in ice if j not in unpaired and j not in
             [int(a["pr"][k][0]) for k in range(nr) if int(a["re"][k][0]) in unpaired]]
    err = np.abs(pre[:, check] + corr[:, check] - ydot[:, check])
    assert np.all(err <= 1e-9 * scale[:, check] + 1e-300), float((err / (scale[:, check] + 1e-300)).max())
    # reference_signs: the reference function's own dy, production minus loss with the
    # corrections the other way round and the ice totals summed over the '#' and '@' species
    dyr, flr, _ = analysis.rates_to_dy_and_flux(phys[rows], chem[rows], rates[rows], dicts, net=gpu_net,
                                                reference_signs=True)
    assert np.array_equal(flr, flux)
    dyr = dyr.reshape(n, -1)
    assert np.allclose(dyr[:, check], pre[:, check] - corr[:, check], rtol=1e-12, atol=0)
