"""GPU tests of the batched model API (uclchem_amd/model.py) in the reference's layouts.

* cloud_grid(return_array=True) on the reference's example model reproduces
  examples/example-output/static-full.dat (times column, physics row and abundances at
  1 Myr), in the (timepoints+1, points, ...) arrays of io.f90:74-82;
* the out_species form is [success_flag] + abundances, as model.py's _format_output;
* hot_core_grid runs the hot-core schedule with the reference's temp_indx check;
* get_rates_and_odes: the rate coefficients equal the oracle's at the given abundances and
  F after the first update equals the oracle's F at the same updated state.
"""
import os

import numpy as np
import pytest

from helpers import term_magnitude

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_cloud_grid_arrays_match_example_output():
    from uclchem_amd import model
    g = np.load(os.path.join(GOLD, "static_full_t1e6.npz"))
    cfg1 = {"initialDens": 1e4, "initialTemp": 10.0, "finalTime": 1e6}
    other = {"initialDens": 1e5, "initialTemp": 20.0, "finalTime": 1e5}
    phys, chem, rates, start, flags = model.cloud_grid([cfg1, other], return_array=True)
    assert rates is None and list(flags) == [0, 0]
    n = len(g["times"])
    assert phys.shape[1:] == (2, 8) and chem.shape[1:] == (2, 335) and phys.shape[0] >= n
    assert np.allclose(phys[:n, 0, 0], g["times"], rtol=1e-12)
    assert np.allclose(phys[n - 1, 0, 1:6], g["physics"][1:6], rtol=1e-4)
    m = g["abundances"] > 1e-14
    rel = np.abs(chem[n - 1, 0, m] - g["abundances"][m]) / g["abundances"][m]
    assert rel.max() < 1e-3
    assert np.array_equal(start[0], chem[n - 1, 0])
    assert phys[0, 1, 1] == 1e5


def test_out_species_format():
    from uclchem_amd import model
    res = model.cloud_grid([{"initialDens": 1e4, "finalTime": 1e4}], out_species=["CO", "H2O", "#H2O"])
    assert len(res) == 1 and len(res[0]) == 4 and res[0][0] == 0
    phys, chem, _, _, _ = model.cloud_grid([{"initialDens": 1e4, "finalTime": 1e4}], return_array=True)
    from uclchem_amd import makerates_hip as mh
    names = mh.load_network(mh.network_path("default")).species_names
    assert res[0][1] == chem[-1, 0, names.index("CO")]


def test_hot_core_grid_matches_reference():
    """model.hot_core_grid (the reference's model.hot_core, hotcore.f90's temperature schedule)
    against the reference (tests/golden/ref_start_hotcore.npz): temp_indx 3, maxTemp 300 K,
    n_H 1e5 to 2e4 yr, every output row's time, density, gasTemp and Av, and the final
    abundances above 1e-14 within 1% plus twice the reference's own integration noise."""
    from helpers import excess_over_reference
    from uclchem_amd import model
    with pytest.raises(ValueError):
        model.hot_core_grid(7, 300.0, [{}])
    path = os.path.join(GOLD, "ref_start_hotcore.npz")
    if not os.path.exists(path):
        pytest.skip("ref_start_hotcore.npz not generated (make_ref_fixtures.py --start-hotcore-only)")
    f = np.load(path)
    n = len(f["hot_rows"])
    phys, chem, _, start, flags = model.hot_core_grid(3, 300.0, [{"initialDens": 1e5, "finalTime": 2e4}],
                                                      return_array=True, timepoints=n + 8)
    assert flags[0] == 0
    np.testing.assert_allclose(phys[:n, 0, 0], f["hot_rows"][:, 0], rtol=1e-12)
    np.testing.assert_allclose(phys[:n, 0, 1:3], f["hot_rows"][:, 1:3], rtol=1e-12)
    np.testing.assert_allclose(phys[:n, 0, 4], f["hot_rows"][:, 3], rtol=1e-12)
    assert phys[n - 1, 0, 2] > 40.0          # the gas warmed along the hot-core schedule
    e = excess_over_reference(start[0], f["hot_final"][4:], f["hot_final_tight"][4:])
    print(f"[hot core] largest excess over the 1% + noise allowance: {e:.3g}", flush=True)
    assert e <= 1.0


def test_get_rates_and_odes_match_oracle(tables, oracle):
    from oracle import make_params as omake
    from uclchem_amd import model
    rng = np.random.default_rng(9)
    dicts, ab = [], []
    for _ in range(3):
        d = {"initialDens": float(10 ** rng.uniform(3, 6)), "initialTemp": float(rng.uniform(10, 60))}
        dicts.append(d)
        y = 10 ** rng.uniform(-20, -4, tables.nspec)
        y[tables.nh2] = 0.4
        ab.append(y)
    rate, odes = model.get_rates_and_odes(dicts, np.array(ab))
    for q, d in enumerate(dicts):
        pt, _ = oracle.init_point(omake(dens=d["initialDens"], temp=d["initialTemp"], finaltime=1e-7))
        y0 = np.append(ab[q], d["initialDens"])
        ro, y0r, _ = oracle.rates(pt, y0)
        rel = np.abs(rate[q] - ro) / np.maximum(np.abs(ro), 1e-300)
        rel[(rate[q] == 0) & (ro == 0)] = 0
        assert rel.max() <= 1e-12
        # F at the updated state: same rates, the GPU solve's y after the 1e-7-yr interval
        out = model.network().solve(np.stack([model.make_params(dict(d), finaltime=1e-7)]), y_init=y0[None])
        fo, r2 = oracle.rhs(pt, ro, out["y"][0])
        assert np.all(np.abs(odes[q] - fo) <= 1e-12 * term_magnitude(tables, r2, out["y"][0]))


def test_coefficient_changes_and_rates_output_match_reference():
    """The alpha/beta/gamma dictionary keys (wrap.f90:744-761) and return_rates
    (ratesarray, io.f90:90-93) against the reference run with its module arrays changed the
    same way (tests/golden/ref_coeff.npz): the static cloud to 1e5 yr, every abundance above
    1e-14 within 1%, and the rate array at five output rows within 1e-3 relative (the
    coefficients follow the state each interval starts from, which the two integrators reach
    within their tolerances).  The H2 and CO photodissociation rates of a row are those of the
    integrator's last RHS evaluation (chemistry.f90:311-319), at a state inside its last step,
    which differs between the two integrators: within 10%."""
    import os
    from uclchem_amd import model
    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_coeff.npz"))
    d = dict(initialdens=1e4, initialtemp=10.0, finaltime=1e5)
    for w in ("alpha", "beta", "gamma"):
        sel = f["which"] == w
        d[w] = {int(k): float(v) for k, v in zip(f["index"][sel], f["value"][sel])}
    phys, chem, rates, start, flags = model.cloud_grid([d], return_array=True, return_rates=True)
    assert flags[0] == 0 and chem.shape[0] == len(f["rows"])
    ref = f["rows"][-1, 4:]
    m = ref > 1e-14
    assert (np.abs(chem[-1, 0][m] - ref[m]) / ref[m]).max() < 0.01
    plain = model.cloud_grid([dict(initialdens=1e4, initialtemp=10.0, finaltime=1e5)], return_array=True)
    assert (np.abs(plain[1][-1, 0][m] - ref[m]) / ref[m]).max() > 0.05    # the changes matter
    assert plain[2] is None
    got = rates[f["rate_rows"], 0, :]
    want = f["rates"]
    photo = [model.network().t.scalars[k] for k in ("nR_H2_hv", "nR_CO_hv")]
    nz = np.abs(want) > 1e-30
    assert not np.any(np.abs(got[~nz]) >= 1e-30)
    rel = np.where(nz, np.abs(got - want) / np.where(nz, np.abs(want), 1.0), 0.0)
    other = np.ones(rel.shape[1], dtype=bool)
    other[photo] = False
    assert rel[:, other].max() < 1e-3, (rel[:, other].max(), np.unravel_index(np.argmax(rel[:, other]), rel[:, other].shape))
    assert rel[:, photo].max() < 0.1, rel[:, photo]
    assert np.all(rates[0, 0] == 0.0)
