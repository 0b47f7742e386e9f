"""GPU parity on the all-types check network (tests/golden/net_alltypes.npz, written by
make_alltypes_network.py): the default network plus reactions of every type it lacks —
IONOPOL1/2 (rates.f90:301-313), CRS, EXRELAX, EXSOLID (rates.f90:158-197), GAR
(rates.f90:316-332) and DESORB.

Gates: rates and RHS against the oracle (C restatement, itself bit-exact against the
reference compiled on this network, tests/test_reference_format.py) on seeded states, with
every switch of the reaction types toggled: rates <= 1e-12 relative, RHS <= 1e-12 of each
species' term magnitude; and whole models against the reference Fortran compiled on this
network (oracle/_ref/net_alltypes, from tests/golden/ref_alltypes.npz): every abundance above
1e-14 within 1% plus twice the reference's own integration noise (|final - final_tight|, the
reference at reltol 1e-11, abstol_factor 1e-17, abstol_min 1e-30)."""
import os

import numpy as np
import pytest

from helpers import phys_of, random_state, term_magnitude

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def alltypes():
    from oracle import Oracle
    from uclchem_amd import gpu, makerates_hip as mh
    t = mh.load_network(os.path.join(GOLD, "net_alltypes.npz"))
    return t, gpu.Network(t), Oracle(t)


def test_every_type_present(alltypes):
    from uclchem_amd import makerates_hip as mh
    t, _, _ = alltypes
    have = {mh.ALL_TYPES[int(x)] for x in t.rtype}
    assert {"IONOPOL1", "IONOPOL2", "CRS", "EXRELAX", "EXSOLID", "GAR", "DESORB"} <= have


def test_rates_and_rhs_match_oracle(alltypes):
    t, net, orc = alltypes
    from uclchem_amd.params import make_params
    rng = np.random.default_rng(21)
    P, PH, Y, pts = [], [], [], []
    for trial in range(8):
        sw = rng.integers(0, 2, 4)
        p = make_params(initialdens=10 ** rng.uniform(2, 7), initialtemp=rng.uniform(10, 200),
                        zeta=10 ** rng.uniform(0, 2), radfield=10 ** rng.uniform(0, 3), baseav=rng.uniform(0, 10),
                        finaltime=1e6, ionmodel=float(sw[0]), cosmicrayattenuation=float(sw[1]),
                        improvedh2crpdissociation=float(sw[1] and sw[2]), crdesorb=float(sw[3]),
                        phi=10 ** rng.uniform(3, 6))
        pt, y = random_state(t, orc, rng, p, bare_grains=(trial % 3 == 0))
        P.append(p)
        PH.append(phys_of(pt))
        Y.append(y)
        pts.append(pt)
    P, PH, Y = np.array(P), np.array(PH), np.array(Y)
    rg, yg, _ = net.rates(P, PH, Y)
    R, Yr, YD, MAG = [], [], [], []
    for q in range(len(P)):
        ro, yo, _ = orc.rates(pts[q], Y[q])
        d = np.abs(rg[q] - ro) / np.maximum(np.abs(ro), 1e-300)
        d[(rg[q] == 0) & (ro == 0)] = 0
        assert d.max() <= 1e-12, (q, int(np.argmax(d)))
        yd, r2 = orc.rhs(pts[q], ro, yo)
        R.append(ro)
        Yr.append(yo)
        YD.append(yd)
        MAG.append(term_magnitude(t, r2, yo))
    ydg, _ = net.rhs(P, PH, np.array(R), np.array(Yr))
    assert np.all(np.abs(ydg - np.array(YD)) <= 1e-12 * np.array(MAG))


@pytest.mark.parametrize("lu_bits", [64, 32])
def test_models_match_reference(alltypes, lu_bits):
    path = os.path.join(GOLD, "ref_alltypes.npz")
    if not os.path.exists(path):
        pytest.skip("ref_alltypes.npz not generated")
    t, net, _ = alltypes
    f = np.load(path)
    out = net.solve(f["params"], max_rec=200, lu_bits=lu_bits)
    bad = []
    from helpers import excess_over_reference
    for q in range(len(f["params"])):
        assert f["rc"][q] == 0 and out["status"][q] == 0, (q, f["rc"][q], out["status"][q])
        assert out["n_rec"][q] == f["nrows"][q]
        tight = f["final_tight"][q, 4:] if "final_tight" in f.files else None
        e = excess_over_reference(out["y"][q, :t.nspec], f["final"][q, 4:], tight)
        if e > 1.0:
            bad.append((q, e))
    assert not bad, bad
