"""GPU parity of the collapse models (model 2: the BE1.1, BE4, filament and ambipolar
parameterisations of collapse.f90, Priestley et al. 2018) against the reference's own collapse
module, compiled in place into oracle/_ref (tests/golden/ref_collapse_models.npz, written by
make_ref_fixtures.py --collapse-only): two points per model, every output row's time equal,
its density and Av (the fitted profiles, the 10,000-trapezoid column to the cloud edge and, for
BE1.1 and BE4, the radius that encloses the parcel's mass) within 1e-9 relative, and the
abundances of five rows (the row after 1e4 yr, 1/8, 1/4 and 1/2 of the way, the last) above
1e-14 within 1% plus twice the reference's own integration noise on that row (|ref - tight|,
tight: the reference at reltol 1e-11, abstol_factor 1e-17, abstol_min 1e-30)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture():
    path = os.path.join(GOLD, "ref_collapse_models.npz")
    if not os.path.exists(path):
        pytest.skip("ref_collapse_models.npz not generated (make_ref_fixtures.py --collapse-only)")
    return np.load(path)


@pytest.mark.parametrize("lu_bits", [64, 32])
def test_collapse_models_match_reference(gpu_net, tables, lu_bits):
    from helpers import excess_over_reference
    f = _fixture()
    params = f["params"]
    for q in range(len(params)):
        print(f"point {q}: {f['model'][q]} T={params[q, 1]:.0f} K rout={params[q, 5]:.3g} pc, "
              f"{f['nrows'][q]} rows (reference {f['wall'][q]:.0f} s)", flush=True)
    out = gpu_net.solve(params, max_rec=int(f["nrows"].max()) + 8, lu_bits=lu_bits, order="densest")
    bad, worst = [], 0.0
    for q in range(len(params)):
        n = int(f["nrows"][q])
        assert f["rc"][q] == 0 and out["status"][q] == 0, (q, f["rc"][q], out["status"][q])
        assert out["n_rec"][q] == n, (q, out["n_rec"][q], n)
        rec, ref = out["rec"][q, :n], f["phys"][q, :n]
        np.testing.assert_allclose(rec[:, 0], ref[:, 0], rtol=1e-12, err_msg=f"point {q} times")
        np.testing.assert_allclose(rec[:, 1], ref[:, 1], rtol=1e-9, err_msg=f"point {q} density")
        np.testing.assert_allclose(rec[:, 2], ref[:, 2], rtol=1e-12, err_msg=f"point {q} gasTemp")
        np.testing.assert_allclose(rec[:, 3], ref[:, 3], rtol=1e-9, err_msg=f"point {q} Av")
        for j, k in enumerate(f["row_index"][q]):
            assert rec[k, 0] == f["rows"][q, j, 0]
            e = excess_over_reference(out["rec"][q, k, 4:], f["rows"][q, j, 4:], f["rows_tight"][q, j, 4:])
            worst = max(worst, e)
            if e > 1.0:
                bad.append((q, str(f["model"][q]), int(k), e))
        assert np.array_equal(out["y"][q, :tables.nspec], out["rec"][q, n - 1, 4:])
    print(f"[collapse] largest excess over the 1% + noise allowance: {worst:.3g}", flush=True)
    assert not bad, bad


def test_collapse_grid_entry(gpu_net):
    """model.collapse_grid returns the reference's array layout, the same rows as the solve
    of the same parameter vectors, and refuses an unknown collapse name as the reference's
    model.collapse does."""
    from uclchem_amd import model
    from uclchem_amd.params import make_params
    d = dict(finaltime=2.0e4)
    phys, chem, rates, start, flags = model.collapse_grid("ambipolar", [d, dict(d, initialtemp=12.0)],
                                                          return_array=True, timepoints=200)
    assert flags.tolist() == [0, 0] and rates is None
    p = np.stack([make_params(dict(dd, model=2, collapse=4)) for dd in (d, dict(d, initialtemp=12.0))])
    out = gpu_net.solve(p, max_rec=201)
    for q in range(2):
        k = int(out["n_rec"][q])
        assert np.array_equal(phys[:k, q, 0], out["rec"][q, :k, 0])
        assert np.array_equal(phys[:k, q, 1], out["rec"][q, :k, 1])
        assert np.array_equal(chem[:k, q, :], out["rec"][q, :k, 4:])
        assert np.array_equal(start[q], out["rec"][q, k - 1, 4:])
    with pytest.raises(ValueError):
        model.collapse_grid("BE2", [d])


def test_unacceptable_collapse_mode_is_refused(gpu_net):
    """A collapse mode outside 1..4 is the reference's "unacceptable collapse mode"
    (collapse.f90:43-46): the host API refuses the batch."""
    from uclchem_amd.gpu import GpuError
    from uclchem_amd.params import make_params
    with pytest.raises(GpuError, match="collapse mode"):
        gpu_net.solve(make_params(model=2, collapse=5, finaltime=10.0)[None])
