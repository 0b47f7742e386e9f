"""Reaction-coefficient changes (the reference's alpha/beta/gamma dictionary keys,
wrap.f90:744-761; coefficientParser :985-1019) on the MakeRates backend's tables
(makerates_hip.with_coefficients): unchanged tables come back bit-identical; a change to an
LH or LHDES gamma moves exactly those reactions' derived constants; and the C restatement on
the changed tables gives the compiled reference's rates bit for bit after the reference's
own module arrays are changed the same way."""
import numpy as np
import pytest

from uclchem_amd import makerates_hip as mh


def _first(tables, kind):
    return next(k for k in range(tables.nreac) if mh.ALL_TYPES[int(tables.rtype[k])] == kind)


def test_no_change_is_identity(tables):
    t2 = mh.with_coefficients(tables)
    for k, v in tables.arrays.items():
        assert np.array_equal(v, t2.arrays[k]), k


def test_derived_constants_follow_gamma(tables):
    lh, lhdes = _first(tables, "LH"), _first(tables, "LHDES")
    t2 = mh.with_coefficients(tables, gamma={lh + 1: 500.0, lhdes + 1: 1000.0}, alpha={1: 2.0})
    changed = [k for k in range(tables.nreac) if not np.array_equal(tables.rstat[k], t2.rstat[k])]
    assert changed == sorted([lh, lhdes])
    assert t2.alpha[0] == 2.0 and t2.gama[lh] == 500.0
    assert np.array_equal(tables.alpha[1:], t2.alpha[1:])
    with pytest.raises(ValueError):
        mh.with_coefficients(tables, beta={tables.nreac + 1: 1.0})


def test_changed_rates_bit_exact_vs_reference(reflib, tables):
    from oracle import Oracle, make_params
    from helpers import random_state
    names = list(tables.species_names)
    two = next(k for k in range(tables.nreac) if mh.ALL_TYPES[int(tables.rtype[k])] == "TWOBODY"
               and sorted(names[i] for i in tables.re[k][:2] if 0 <= i < tables.nspec) == ["CO", "H3+"])
    lh, lhdes, crp = _first(tables, "LH"), _first(tables, "LHDES"), _first(tables, "CRP")
    changes = dict(alpha={two + 1: 3.3e-9, crp + 1: 0.7}, beta={two + 1: -0.25}, gamma={lh + 1: 600.0, lhdes + 1: 50.0})
    t2 = mh.with_coefficients(tables, **changes)
    orc, base = Oracle(t2), Oracle(tables)
    old = {w: {k: float(tables.arrays["gama" if w == "gamma" else w][k - 1]) for k in ch} for w, ch in changes.items()}
    try:
        for w, ch in changes.items():
            for k, v in ch.items():
                reflib.set_coefficient(w, k, v)
        rng = np.random.default_rng(12)
        seen = set()
        for trial in range(4):
            p = make_params(dens=10 ** rng.uniform(3, 6), temp=rng.uniform(10, 60) + trial)
            rc, _ = reflib.init(p)
            pt, y0 = orc.init_point(p)
            _, y = random_state(t2, orc, rng, p)
            rr, _, _ = reflib.rates(y)
            ro, _, _ = orc.rates(pt, y)
            assert np.array_equal(rr, ro), (trial, int(np.argmax(np.abs(rr - ro))))
            pb, _ = base.init_point(p)
            rb, _, _ = base.rates(pb, y)
            seen |= {k for k in (two, crp, lh, lhdes) if ro[k] != rb[k]}
        assert seen == {two, crp, lh, lhdes}       # every change reached the rates
    finally:
        for w, ch in old.items():
            for k, v in ch.items():
                reflib.set_coefficient(w, k, v)
