"""CPU checks of the host-side parameter plumbing and seeded workloads."""
import numpy as np

from uclchem_amd import params, workloads


def test_defaults_and_float32_literals():
    p = params.make_params()
    assert p[params.INDEX["rout"]] == float(np.float32(0.05))     # default-real literal
    assert p[params.INDEX["reltol"]] == 1e-8
    q = params.make_params({"initialDens": 1e4, "freefall": True})
    assert q[params.INDEX["initialdens"]] == 1e4 and q[params.INDEX["freefall"]] == 1.0


def test_unknown_key_rejected():
    import pytest
    with pytest.raises(KeyError):
        params.make_params(notaparam=1)


def test_grid_av_mapping():
    g = workloads.static_cloud_grid(16)
    nh = g[:, params.INDEX["initialdens"]]
    av = g[:, params.INDEX["baseav"]] + params.DEFAULTS["rout"] * params.PC * nh / 1.6e21
    assert np.all((av >= 1.0 - 1e-9) & (av <= 20.0 + 1e-9))
    assert np.array_equal(g, workloads.static_cloud_grid(16))    # seeded


def test_oracle_and_product_param_layouts_agree():
    import oracle as orc
    # the trailing slot (maxsteps, a work bound with no reference counterpart) is honoured by
    # the reference driver too, so both layouts carry it
    assert params.NAMES[-1] == "maxsteps" and params.DEFAULTS["maxsteps"] == 0.0
    assert orc.P_NAMES == [
        "dens", "temp", "zeta", "radfield", "baseav", "rout", "rin", "finaltime", "freefall", "finaldens",
        "reltol", "abstol_factor", "abstol_min", "model", "tempindx", "maxtemp", "freezefactor",
        "endatfinaldensity", "freefallfactor", "mxstep", "maxsteps"]
    a = orc.make_params(dens=3e3, temp=20.0, zeta=2.0, maxsteps=5e4)
    b = params.make_params(initialdens=3e3, initialtemp=20.0, zeta=2.0, maxsteps=5e4)
    assert np.array_equal(a, b)


def test_collapse_and_hot_core_workloads():
    """The seeded free-fall grid (log n_H U[2,6] towards 1e6 cm^-3) and hot-core grid
    (log n_H U[5,8], temp_indx 4, 300 K) carry the model switches the kernel reads."""
    I = params.INDEX
    c = workloads.collapse_grid(64)
    assert np.all(c[:, I["freefall"]] == 1.0) and np.all(c[:, I["finaldens"]] == 1e6)
    assert np.all((c[:, I["initialdens"]] >= 1e2) & (c[:, I["initialdens"]] <= 1e6))
    h = workloads.hot_core_grid(64)
    assert np.all(h[:, I["model"]] == 1.0) and np.all(h[:, I["tempindx"]] == 4.0)
    assert np.all(h[:, I["maxtemp"]] == 300.0) and np.all(h[:, I["initialtemp"]] == 10.0)
    assert np.all((h[:, I["initialdens"]] >= 1e5) & (h[:, I["initialdens"]] <= 1e8))
    assert np.array_equal(c, workloads.collapse_grid(64)) and np.array_equal(h, workloads.hot_core_grid(64))
