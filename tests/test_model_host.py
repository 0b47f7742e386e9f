"""Host-side layout logic of the batched model API (uclchem_amd/model.py) on the CPU, with a
stand-in for the resident network whose solve returns outputs that depend only on each
point's parameters: the reference's physicsArray / chemicalAbunArray layouts (io.f90:74-82),
the trimming to the last written row, abundanceStart, the out_species form (model.py:71-76)
and the argument checks."""
import numpy as np
import pytest

from uclchem_amd import model
from uclchem_amd.params import INDEX

NSPEC = 4


class _Tables:
    species_names = ["H", "H2", "CO", "#CO"]


NREAC = 3


class _FakeNet:
    nspec, neq, nreac = NSPEC, NSPEC + 1, NREAC
    t = _Tables()

    def __init__(self, tag=0.0):
        self.calls = []
        self.tag = tag

    def solve(self, params, y_init=None, max_rec=0, lu_bits=64, order=None, return_rates=False):
        self.calls.append(dict(params=params.copy(), y_init=None if y_init is None else y_init.copy(),
                               max_rec=max_rec, order=order, return_rates=return_rates))
        n = params.shape[0]
        dens = params[:, INDEX["initialdens"]]
        n_rec = (2 + np.arange(n) % 3).astype(np.int32)          # 2, 3, 4, 2, ... rows written
        rec = np.zeros((n, max_rec, 4 + NSPEC))
        for q in range(n):
            for r in range(n_rec[q]):
                rec[q, r, 0] = 10.0 ** r                            # time
                rec[q, r, 1] = dens[q]                              # density
                rec[q, r, 2] = 10.0 + q                             # gasTemp
                rec[q, r, 3] = 2.0 + q                              # Av
                rec[q, r, 4:] = dens[q] * (r + 1) + np.arange(NSPEC)
        y = np.zeros((n, NSPEC + 1))
        for q in range(n):
            y[q, :NSPEC] = rec[q, n_rec[q] - 1, 4:]
            y[q, NSPEC] = dens[q]
        status = np.where(dens > 1e6, -5, 0).astype(np.int32)
        out = dict(y=y, rec=rec, n_rec=n_rec, status=status, stats=np.zeros((n, 8), np.int32))
        if return_rates:   # rate k of row r: density + 10 r + k + the network's tag
            out["rates"] = (dens[:, None, None] + 10.0 * np.arange(max_rec)[None, :, None]
                            + np.arange(NREAC)[None, None, :] + self.tag)
        return out


@pytest.fixture
def net(monkeypatch):
    fake = _FakeNet()
    monkeypatch.setattr(model, "network", lambda name="default": fake)
    return fake


def test_array_layout_matches_the_reference(net):
    dicts = [{"initialDens": 1e3, "zeta": 2.0, "radfield": 5.0}, {"initialDens": 1e4}, {"initialDens": 1e5}]
    phys, chem, rates, start, flags = model.cloud_grid(dicts, return_array=True, timepoints=10)
    assert rates is None and list(flags) == [0, 0, 0]
    assert net.calls[0]["max_rec"] == 11 and net.calls[0]["order"] == "densest"
    rows = 4                                            # the most rows any point wrote
    assert phys.shape == (rows, 3, 8) and chem.shape == (rows, 3, NSPEC)
    n_rec = [2, 3, 4]
    for q, d in enumerate(dicts):
        k = n_rec[q]
        assert np.array_equal(phys[:k, q, 0], 10.0 ** np.arange(k))         # time
        assert np.all(phys[:k, q, 1] == d["initialDens"])                    # density
        assert np.all(phys[:k, q, 2] == 10.0 + q) and np.all(phys[:k, q, 3] == 10.0 + q)   # gas = dust T
        assert np.all(phys[:k, q, 4] == 2.0 + q)                             # Av
        assert np.all(phys[:k, q, 5] == d.get("radfield", 1.0)) and np.all(phys[:k, q, 6] == d.get("zeta", 1.0))
        assert np.all(phys[:k, q, 7] == 1)                                   # dstep
        assert np.all(phys[k:, q] == 0) and np.all(chem[k:, q] == 0)         # rows never written stay zero
        assert np.array_equal(start[q], chem[k - 1, q])                      # abundanceStart = last row


def test_out_species_form_and_failed_points(net):
    res = model.cloud_grid([{"initialDens": 1e4}, {"initialDens": 1e7}], out_species=["CO", "H"])
    assert res[0] == [0, 1e4 * 2 + 2, 1e4 * 2 + 0]         # [flag] + final CO, H (point 0 wrote 2 rows)
    assert res[1] == [-5]                                   # a failed point returns its flag only
    assert model.cloud_grid([{"initialDens": 1e4}]) == [[0]]


def test_starting_chemistry_and_hot_core_arguments(net):
    sc = np.array([0.5, 0.25, 1e-4, 1e-6])
    model.cloud_grid([{"initialDens": 1e3}, {"initialDens": 1e4}], starting_chemistry=sc)
    y0 = net.calls[-1]["y_init"]
    assert y0.shape == (2, NSPEC + 1) and np.all(y0[:, :NSPEC] == sc)
    assert list(y0[:, NSPEC]) == [1e3, 1e4]                 # density appended as the last equation
    with pytest.raises(ValueError):
        model.cloud_grid([{}], starting_chemistry=np.zeros(3))
    with pytest.raises(ValueError):
        model.hot_core_grid(0, 300.0, [{}])
    with pytest.raises(ValueError):
        model.hot_core_grid(7, 300.0, [{}])
    model.hot_core_grid(3, 250.0, [{"initialDens": 1e6}])
    p = net.calls[-1]["params"][0]
    assert (p[INDEX["model"]], p[INDEX["tempindx"]], p[INDEX["maxtemp"]]) == (1.0, 3.0, 250.0)


def test_collapse_arguments(net):
    """collapse_grid names the reference's four collapse models (model.py:355) and passes the
    mode in the collapse slot of model 2."""
    for name, mode in [("BE1.1", 1), ("BE4", 2), ("filament", 3), ("ambipolar", 4)]:
        model.collapse_grid(name, [{"initialTemp": 12.0}])
        p = net.calls[-1]["params"][0]
        assert (p[INDEX["model"]], p[INDEX["collapse"]], p[INDEX["initialtemp"]]) == (2.0, float(mode), 12.0)
    for bad in ("BE2", "be1.1", 1):
        with pytest.raises(ValueError):
            model.collapse_grid(bad, [{}])
    assert INDEX["collapse"] == 59   # UCG_P_COLLAPSE


def test_rates_rows_and_coefficient_groups(net, monkeypatch):
    """return_rates gives ratesArray[rows, points, nreac] trimmed like the abundances; points
    whose dictionaries carry the same alpha/beta/gamma changes run together on one derived
    network, and every result lands at its point's index."""
    derived = {}

    def coeff_net(name, changes):
        if not changes:
            return net
        return derived.setdefault(changes, _FakeNet(tag=1000.0 * (len(derived) + 1)))
    monkeypatch.setattr(model, "_coefficient_network", coeff_net)
    dicts = [{"initialDens": 1e3, "alpha": {2: 1e-9}}, {"initialDens": 1e4}, {"initialDens": 1e5, "ALPHA": {"2": 1e-9}},
             {"initialDens": 2e3, "gamma": {1: 10.0}, "beta": {3: 0.5}}]
    phys, chem, rates, start, flags = model.cloud_grid(dicts, return_array=True, return_rates=True, timepoints=10)
    assert len(derived) == 2 and len(net.calls) == 1 and net.calls[0]["return_rates"]
    a = derived[(("alpha", ((2, 1e-9),)),)]
    assert [c["params"].shape[0] for c in a.calls] == [2]
    assert rates.shape == (3, 4, NREAC)
    n_rec = {0: 2, 1: 2, 2: 3, 3: 2}          # the fake writes 2 + (position in its launch) % 3 rows
    tags = [1000.0, 0.0, 1000.0, 2000.0]
    for q, d in enumerate(dicts):
        dens = d["initialDens"]
        assert np.all(phys[:n_rec[q], q, 1] == dens)
        for r in range(n_rec[q]):
            assert np.array_equal(rates[r, q], dens + 10.0 * r + np.arange(NREAC) + tags[q])
        assert np.all(rates[n_rec[q]:, q] == 0)
    assert model.cloud_grid(dicts[1:2], return_array=True)[2] is None
