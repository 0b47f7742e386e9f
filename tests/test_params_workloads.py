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
    # the same 64 slots in the same order (the oracle spells the first two dens/temp); the
    # trailing work bound (maxsteps, no reference counterpart) is honoured by the reference
    # driver too
    assert len(params.NAMES) <= params.NPARAM == orc.P_NPARAM == 64
    assert orc.P_NAMES[:2] == ["dens", "temp"] and orc.P_NAMES[2:] == params.NAMES[2:]
    for name in params.NAMES[2:]:
        assert orc.DEFAULTS[name] == params.DEFAULTS[name], name
    a = orc.make_params(dens=3e3, temp=20.0, zeta=2.0, maxsteps=5e4, metallicity=0.5, ion=1, phi=3e4)
    b = params.make_params(initialdens=3e3, initialtemp=20.0, zeta=2.0, maxsteps=5e4, metallicity=0.5, ion=1,
                           phi=3e4)
    assert np.array_equal(a, b)


def test_parameter_dict_quirks_of_the_reference_parser():
    """dictionaryParser's thermdesorb branch sets uvdesorb (wrap.f90:802-803), and ionModel
    is the character 'L' or 'H'; default-real literals keep their float32 values."""
    p = params.make_params({"thermdesorb": False, "ionModel": "H"})
    assert p[params.INDEX["uvdesorb"]] == 0.0 and p[params.INDEX["thermdesorb"]] == 1.0
    assert p[params.INDEX["ionmodel"]] == 1.0
    d = params.make_params()
    assert d[params.INDEX["epsilon"]] == float(np.float32(0.01))
    assert d[params.INDEX["fhe"]] == float(np.float32(0.1))
    import pytest
    with pytest.raises(ValueError):
        params.make_params(ionModel="X")


def test_widen_keeps_the_early_layout():
    old = params.make_params(initialdens=5e3, maxsteps=7.0)[:params.NPARAM_V1]
    w = params.widen(old)
    assert w.shape == (1, params.NPARAM)
    assert np.array_equal(w[0], params.make_params(initialdens=5e3, maxsteps=7.0))


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
