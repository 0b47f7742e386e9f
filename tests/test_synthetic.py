"""CPU tests of the seeded synthetic networks (500/6000 collapse network, 150/1500 serving network).

Pinned by what the construction guarantees and the reference's own invariants:
* every reaction conserves atoms, and the oracle's RHS conserves every element on
  seeded states (the reference's tests/test_ode_conservation.py invariant);
* the mandatory species and reactions UCLCHEM looks up by name are present and the
  reaction types are contiguous (the range loops of rates.f90);
* the type mix is the one the 500/6000 workload asks for (10% each of cosmic-ray,
  photo, freeze-out and desorption reactions);
* generation is deterministic for a seed and the saved network equals a rebuild;
* the static-pattern Newton plan of a synthetic network reproduces a dense LU.
"""
import numpy as np
import pytest

from helpers import random_state, term_magnitude


@pytest.fixture(scope="module")
def syn150():
    from uclchem_amd import synthetic as sy
    return sy.load_or_build("synthetic_150_1500")


@pytest.fixture(scope="module")
def syn500():
    from uclchem_amd import synthetic as sy
    return sy.load_or_build("synthetic_500_6000")


def test_conservation_by_construction():
    from uclchem_amd import synthetic as sy
    sp, rx = sy.synthetic_species_and_reactions(150, 1500, seed=7)
    sy.check_conservation(sp, rx)
    bad = [r for r in rx if r.rtype == "TWOBODY"][0]
    broken = type(bad)(list(bad.reactants), [bad.products[0], "NAN", "NAN", "NAN"], 1.0, 0.0, 0.0)
    if broken.products[0] != "NAN" and any(p != "NAN" for p in bad.products[1:]):
        with pytest.raises(ValueError):
            sy.check_conservation(sp, rx + [broken])


def test_mandatory_species_reactions_and_type_order(syn500, syn150):
    from uclchem_amd import makerates_hip as mh
    for t in (syn150, syn500):
        names = t.species_names
        for s in ["H", "H2", "HE", "C", "C+", "CO", "H2O", "E-", "#H", "#H2", "#H2O", "BULK", "SURFACE"]:
            assert s in names, s
        for key in ["nR_H2Form_CT", "nR_H2Form_ERDes", "nR_H2Form_ER", "nR_H2Form_LH", "nR_H2Form_LHDes",
                    "nR_HFreeze", "nR_EFreeze", "nR_H2_hv", "nR_CO_hv", "nR_C_hv", "nR_H2Freeze"]:
            assert t.scalars[key] >= 0, key
        rt = t.arrays["rtype"]
        # each type occupies one contiguous block
        for code in np.unique(rt):
            w = np.where(rt == code)[0]
            assert w[-1] - w[0] + 1 == len(w)
        assert rt[-1] == mh.TYPE_CODE["TWOBODY"]
        for i in t.surface_list:
            assert "@" + names[i][1:] in names


def test_type_mix_500(syn500):
    from uclchem_amd import synthetic as sy
    m = sy.type_mix(syn500)
    assert syn500.nspec == 500 and syn500.nreac == 6000
    assert m["FREEZE"] == 600 and m["PHOTON"] == 600
    assert m["CRP"] + m["CRPHOT"] == 600
    assert m["DESCR"] + m["DESOH2"] + m["DEUVCR"] + m["THERM"] == 600
    be = syn500.binding_energy
    heavy = [k for k, i in enumerate(syn500.ice_list) if syn500.species_names[i].lstrip("#@") not in ("H", "H2")]
    assert be[heavy].min() >= 499.0 and be[heavy].max() <= 5001.0


def test_deterministic_and_saved_equals_rebuild(syn150):
    from uclchem_amd import synthetic as sy
    t = sy.synthetic_network(150, 1500)
    for k in ("alpha", "beta", "gama", "re", "pr", "rtype", "loss_idx", "prod_idx", "nw_perm", "nw_sp_col"):
        assert np.array_equal(t.arrays[k], syn150.arrays[k]), k
    assert t.species_names == syn150.species_names


@pytest.mark.parametrize("which", ["syn150", "syn500"])
def test_oracle_rhs_conserves_elements(which, request):
    from oracle import Oracle, make_params
    from uclchem_amd.makerates_hip import ELEMENTS
    t = request.getfixturevalue(which)
    orc = Oracle(t)
    rng = np.random.default_rng(5)
    ec = np.delete(t.elem_counts[:t.nspec].T.astype(float), ELEMENTS.index("E-"), axis=0)
    for trial in range(4):
        p = make_params(dens=10 ** rng.uniform(2, 7), temp=rng.uniform(10, 200), finaltime=1e6)
        pt, y = random_state(t, orc, rng, p, bare_grains=(trial == 0))
        rate, y, _ = orc.rates(pt, y)
        ydot, rate2 = orc.rhs(pt, rate, y)
        assert np.all(np.isfinite(ydot))
        mag = term_magnitude(t, rate2, y)
        scale = ec @ mag[:t.nspec]
        sums = ec @ ydot[:t.nspec]
        assert np.all(np.abs(sums) <= 1e-13 * scale + 1e-300)


def test_newton_plan_solves_like_lapack_150(syn150):
    """The static plan of a synthetic network, executed as the device executes it, solves
    systems with its pattern like numpy.linalg.solve, and its level schedule only reads
    rows of earlier levels."""
    from test_newton_plan import execute, slot_table
    from uclchem_amd import newton_plan as npl
    a = syn150.arrays
    s = npl.plan_summary(a)
    assert s["N"] == syn150.neq + 2 and s["ns"] + s["T"] == s["N"]
    lev = a["nw_lev_ptr"]
    level_of = np.empty(s["ns"], dtype=int)
    for l in range(len(lev) - 1):
        level_of[a["nw_lev_rows"][lev[l]:lev[l + 1]]] = l
    sp, dpos, col = a["nw_sp_ptr"], a["nw_sp_dpos"], a["nw_sp_col"]
    for r in range(s["ns"]):
        assert np.all(level_of[col[sp[r]:dpos[r]]] < level_of[r])
    rc = slot_table(a)
    struct = set()
    for k in ("nw_jnz_slot", "nw_diag_slot", "nw_ph_slot", "nw_tr_slot", "nw_s1c", "nw_s1r", "nw_s2c", "nw_s2r",
              "nw_misc", "nw_nbrow", "nw_nsrow"):
        struct |= {int(x) for x in a[k] if x >= 0}
    idx = np.array(sorted(struct))
    rng = np.random.default_rng(3)
    v = np.zeros(len(rc))
    v[idx] = rng.uniform(-1, 1, len(idx))
    v[[q for q in idx if rc[q, 0] == rc[q, 1]]] += 40.0
    A = np.zeros((s["N"], s["N"]))
    A[rc[idx, 0], rc[idx, 1]] = v[idx]
    b = rng.standard_normal(s["N"])
    ref = np.linalg.solve(A, b)
    assert np.allclose(execute(a, v, b), ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
