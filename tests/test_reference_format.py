"""CPU checks of the reference-format writer (uclchem_amd/reference_format.py).

Pinned against the reference's own generated sources for its default network: the written
odes.f90 is the shipped src/fortran_src/odes.f90 statement for statement (whitespace, line
breaks and comments aside), every table of the written network.f90 holds the values the
shipped network.f90 compiles to, and the reference hot path compiled against the written
files (oracle/_ref/net_default) runs the reference's models bit-identically to the one
compiled from the shipped files.  For the seeded synthetic networks, which MakeRates never
compiled, the reference built on the written files (oracle/_ref/net_<name>) is checked
against the oracle's C restatement: rates and RHS bit-exact on seeded states.
"""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src/fortran_src"


def _net_lib(name):
    path = os.path.join(ROOT, "oracle", "_ref", f"net_{name}", "libucref.so")
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (make -C oracle netref NET={name})")
    from oracle import RefLib
    return RefLib(path)


def _statements(text):
    text = re.sub(r"&\s*\n\s*&", "", text)
    return [ln.strip().replace(" ", "") for ln in text.splitlines() if ln.strip() and not ln.strip().startswith("!")]


def _declarations(text):
    """name -> (declaration, [values]) for every array constructor and every scalar index
    constant (n* = ...) of a network module."""
    text = re.sub(r"&\s*\n\s*&", "", text)
    out = {}
    for ln in text.splitlines():
        m = re.match(r"\s*(.*?)::\s*(\w+)\s*\((\d+)\)\s*=\s*\(/(.*)/\)\s*$", ln)
        if m:
            out[m.group(2)] = (m.group(1).replace(" ", ""), [v.strip() for v in m.group(4).split(",")])
            continue
        m = re.match(r"\s*INTEGER\(dp\),\s*PARAMETER\s*::\s*(n\w*=\d+(,\s*n\w*=\d+)*)\s*$", ln)
        if m:
            for kv in m.group(1).split(","):
                k, v = kv.strip().split("=")
                out[k] = ("scalar", [v])
    return out


@pytest.fixture(scope="module")
def written(tables):
    from uclchem_amd import reference_format as rf
    return rf.network_f90(tables), rf.odes_f90(tables)


def test_odes_regenerates_the_shipped_file(written):
    path = os.path.join(REF_SRC, "odes.f90")
    if not os.path.exists(path):
        pytest.skip("reference sources not mounted")
    ours, theirs = _statements(written[1]), _statements(open(path).read())
    assert len(ours) == len(theirs)
    bad = [i for i, (a, b) in enumerate(zip(ours, theirs)) if a != b]
    assert not bad, (bad[:3], ours[bad[0]][:120], theirs[bad[0]][:120])


def test_network_tables_match_the_shipped_file(written):
    path = os.path.join(REF_SRC, "network.f90")
    if not os.path.exists(path):
        pytest.skip("reference sources not mounted")
    ours, theirs = _declarations(written[0]), _declarations(open(path).read())
    assert set(theirs) <= set(ours), set(theirs) - set(ours)
    for name, (decl, vals) in theirs.items():
        mine = ours[name][1]
        assert len(mine) == len(vals), name
        if "REAL" in decl:        # the shipped default-real literals vs our exact doubles
            a = np.array([float(v.replace("d", "e")) for v in vals], dtype=np.float32).astype(np.float64)
            b = np.array([float(v.replace("_dp", "")) for v in mine])
            assert np.array_equal(a, b), name
        elif "CHARACTER" in decl:
            assert [v.strip('"').strip() for v in vals] == [v.strip('"').strip() for v in mine], name
        else:
            assert [v.lower() for v in vals] == [v.lower() for v in mine], name


def test_writer_planted_mistake_is_caught(tables):
    """A swapped pair of reactants in the tables changes the written odes.f90 (the statement
    comparison above would fail)."""
    from uclchem_amd import reference_format as rf
    import copy
    t2 = copy.copy(tables)
    t2.arrays = dict(tables.arrays)
    ops = tables.arrays["ops"].copy()
    k = int(np.where((ops[:, 0] >= 0) & (ops[:, 1] >= 0) & (ops[:, 0] != ops[:, 1]))[0][0])
    ops[k, 0], ops[k, 1] = ops[k, 1], ops[k, 0]
    t2.arrays["ops"] = ops
    assert _statements(rf.odes_f90(t2)) != _statements(rf.odes_f90(tables))


def test_reference_built_on_written_files_runs_bit_identically(reflib):
    ref2 = _net_lib("default")
    from uclchem_amd import workloads
    params = np.concatenate([workloads.static_cloud_point(final_time=1e5), workloads.static_cloud_grid(3, final_time=1e5)])
    for p in params:
        rc1, rows1, st1 = reflib.model(p, maxrec=100)
        rc2, rows2, st2 = ref2.model(p, maxrec=100)
        assert rc1 == rc2 == 0
        assert np.array_equal(rows1, rows2) and np.array_equal(st1, st2)


@pytest.mark.parametrize("name", ["synthetic_150_1500", "synthetic_500_6000"])
def test_synthetic_reference_matches_oracle(name):
    """Rates and RHS of the reference compiled on a written synthetic network against the
    oracle's C restatement (bit-exact, as on the default network), on states of a short
    reference free-fall run (the phi carried between calls included)."""
    ref = _net_lib(name)
    from oracle import Oracle
    from uclchem_amd import synthetic as sy, workloads
    t = sy.load_or_build(name)
    orc = Oracle(t)
    assert (ref.neq, ref.nreac) == (t.neq, t.nreac)
    p = workloads.collapse_grid(4)[1].copy()
    p[7] = 1e3
    rc, rows, _ = ref.model(p, maxrec=100)
    assert rc == 0
    rc, y0 = ref.init(p)
    pt, y0o = orc.init_point(p)
    assert np.array_equal(y0, y0o)
    for row in rows[::5]:
        y = np.append(row[4:], row[1])
        r1, y1, s1 = ref.rates(y)
        r2, y2, s2 = orc.rates(pt, y)
        assert np.array_equal(r1, r2) and np.array_equal(y1, y2)
        f1 = ref.rhs(y1)
        f2, _ = orc.rhs(pt, r2, y2)
        assert np.array_equal(f1, f2)
