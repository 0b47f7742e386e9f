"""Reference-format network sources written from this backend's network tables.

UCLCHEM's MakeRates compiles every network into three Fortran files the reference's
hot path USEs: `network.f90` (species/reaction tables, special indices, reaction-type
ranges; written by src/uclchem/makerates/io_functions.py:801-951), `odes.f90` (GETYDOT,
the per-species production/loss sums, the bulk/surface transfer branch;
io_functions.py:533-661) and `f2py_constants.f90` (nspec, nReac).  This module writes the
same three files for any network this backend builds — the seeded synthetic networks
included, which never went through MakeRates — so that the reference Fortran can be
compiled for them (the CPU baseline and the reference runs of those networks).

The module layout, names, kinds and attributes are the reference's (the compiled hot path
depends on them); the code that writes them is this project's own.  Real-valued tables are
written as exact double literals (`..._dp`, 17 significant digits), so the compiled module
holds exactly the values the device tables hold (MakeRates writes `{:.4e}` default-real
literals; the tables here already carry their float32 rounding, so both compile to the same
numbers).  Indices are 1-based; 9999 marks an empty reactant/product slot and
(99999, 99999) an absent reaction type, as in the reference.
"""
from __future__ import annotations

import os
from typing import Iterable, List

import numpy as np

from . import makerates_hip as mh

# reaction-type range names in the backend's type order (T_PHOTON ... T_TWOBODY)
RANGE_NAMES = ["photonReacs", "crpReacs", "crphotReacs", "freezeReacs", "desorbReacs", "thermReacs",
               "desoh2Reacs", "descrReacs", "deuvcrReacs", "h2formReacs", "erReacs", "erdesReacs", "lhReacs",
               "lhdesReacs", "bulkswapReacs", "surfswapReacs", "ionopol1Reacs", "ionopol2Reacs", "crsReacs",
               "exsolidReacs", "exrelaxReacs", "garReacs", "twobodyReacs"]
NONE_SLOT = 9999
NONE_RANGE = 99999
WIDTH = 100


def _wrap(text: str, indent: str = "    ") -> str:
    """Break one Fortran statement into continuation lines of at most WIDTH characters,
    splitting only after a comma or a closing parenthesis (never inside a name or a
    literal)."""
    out: List[str] = []
    line = ""
    tok = ""
    for ch in text:
        tok += ch
        if ch in ",)":
            if len(line) + len(tok) > WIDTH - 2 and line:
                out.append(line)
                line = ""
            line += tok
            tok = ""
    if len(line) + len(tok) > WIDTH - 2 and line:
        out.append(line)
        line = ""
    out.append(line + tok)
    return ("&\n" + indent + "&").join(out) + "\n"


def _real(v: float) -> str:
    """Exact double literal: repr round-trips, the _dp suffix keeps the compiler from
    rounding it to default real."""
    r = repr(float(v))
    if r in ("inf", "-inf", "nan"):
        raise ValueError(f"non-finite table value {r}")
    return r + "_dp" if ("e" in r or "." in r) else r + ".0_dp"


def _ints(vals: Iterable[int]) -> str:
    return ",".join(str(int(v)) for v in vals)


def _reals(vals: Iterable[float]) -> str:
    return ",".join(_real(v) for v in vals)


def _one_based(a: np.ndarray, none: int) -> np.ndarray:
    a = np.asarray(a, dtype=np.int64)
    return np.where(a >= 0, a + 1, none)


def _array(decl: str, name: str, body: str, n: int) -> str:
    return _wrap(f"{decl} :: {name} ({n})=(/{body}/)", indent="    ")


def network_f90(t: mh.NetworkTables) -> str:
    a, s = t.arrays, t.scalars
    nspec, nreac = t.nspec, t.nreac
    keys: List[str] = []
    for n in mh.SPECIAL_SPECIES:
        k = mh.species_key(n)
        if k not in keys:
            keys.append(k)
    out = ["MODULE network\n", "USE constants\n", "IMPLICIT NONE\n"]
    out.append(_wrap("    INTEGER(dp), PARAMETER ::" + ",".join(f"{k}={int(s[k]) + 1}" for k in keys)))
    out.append(f"    LOGICAL, PARAMETER :: THREE_PHASE = .{'TRUE' if len(a['bulk_list']) else 'FALSE'}.\n")
    out.append("    REAL(dp) :: SURFGROWTHUNCORRECTED\n")
    ln = max(len(n) for n in t.species_names)
    names = ",".join('"' + n.ljust(ln) + '"' for n in t.species_names)
    out.append(_array(f"CHARACTER(Len={ln}), PARAMETER", "specname", names, nspec))
    out.append(_array("REAL(dp), PARAMETER", "mass", _reals(a["mass"]), nspec))
    out.append(_array("INTEGER(dp), PARAMETER", "atomCounts", _ints(a["n_atoms"]), nspec))
    for nm, key in (("surfaceList", "surface_list"), ("bulkList", "bulk_list"), ("iceList", "ice_list"),
                    ("gasIceList", "gas_ice_list")):
        out.append(_array("INTEGER(dp), PARAMETER", nm, _ints(np.asarray(a[key]) + 1), len(a[key])))
    nsurf = len(a["surface_list"])
    frac = mh.fortran_real(np.asarray(a["sputter_fractions"]) if "sputter_fractions" in a else np.zeros((nsurf, 3)))
    for col, nm in enumerate(("solidFractions", "monoFractions", "volcanicFractions")):
        out.append(_array("REAL(dp), PARAMETER", nm, _reals(frac.reshape(-1, 3)[:, col]), nsurf))
    nice = len(a["ice_list"])
    out.append(_array("REAL(dp)", "bindingEnergy", _reals(a["binding_energy"]), nice))
    out.append(_array("REAL(dp), PARAMETER", "formationEnthalpy", _reals(a["formation_enthalpy"]), nice))
    refr = np.where(np.asarray(a["refractory"]) != 0)[0]
    out.append(_array("INTEGER(dp), PARAMETER", "refractoryList", _ints(refr + 1) if len(refr) else "-999",
                      max(len(refr), 1)))
    out.append(_wrap("    INTEGER(dp), PARAMETER ::" + ",".join(
        f"{k}={int(s[k]) + 1 if int(s[k]) >= 0 else NONE_SLOT}" for k in mh.SPECIAL_REACTIONS)))
    out.append("    REAL(dp) :: REACTIONRATE(1)\n")
    out.append("    LOGICAL :: ReactionRatesToDisk=.false.\n")
    re, pr = np.asarray(a["re"]), np.asarray(a["pr"])
    for q in range(3):
        out.append(_array("INTEGER(dp), PARAMETER", f"re{q + 1}", _ints(_one_based(re[:, q], NONE_SLOT)), nreac))
    for q in range(4):
        out.append(_array("INTEGER(dp), PARAMETER", f"p{q + 1}", _ints(_one_based(pr[:, q], NONE_SLOT)), nreac))
    for nm, key, attr in (("alpha", "alpha", "REAL(dp)"), ("beta", "beta", "REAL(dp)"), ("gama", "gama", "REAL(dp)"),
                          ("minTemps", "tmin", "REAL(dp), PARAMETER"), ("maxTemps", "tmax", "REAL(dp), PARAMETER"),
                          ("reducedMasses", "reduced_mass", "REAL(dp), PARAMETER")):
        out.append(_array(attr, nm, _reals(a[key]), nreac))
    out.append(_array("LOGICAL(dp), PARAMETER", "ExtrapolateRates",
                      ",".join(".True." if v else ".False." for v in a["extrapolate"]), nreac))
    fp = np.asarray(a["freeze_partners"])
    out.append(_array("INTEGER(dp), PARAMETER", "freezePartners", _ints(fp + 1), len(fp)))
    gar = np.asarray(a["gar_params"], dtype=np.float64).reshape(-1, 7)
    if len(gar) == 0:
        gar = np.zeros((1, 7))
    out.append(_wrap(f"REAL(dp), PARAMETER :: garParams({len(gar)},7) = RESHAPE((/ "
                     f"{_reals(gar.T.reshape(-1))} /), (/ {len(gar)}, 7 /))"))
    tr = np.asarray(a["type_ranges"]).reshape(-1, 2)
    for code, nm in enumerate(RANGE_NAMES):
        lo, hi = int(tr[code, 0]), int(tr[code, 1])
        body = f"{lo + 1},{hi + 1}" if lo >= 0 else f"{NONE_RANGE},{NONE_RANGE}"
        out.append(f"INTEGER(dp), PARAMETER :: {nm} (2)=(/{body}/)\n")
    out.append("END MODULE network\n")
    return "".join(out)


def _term(t: mh.NetworkTables, k: int) -> str:
    """One reaction's flux as the reference's ODE code spells it (reaction.py:779-819)."""
    s = f"RATE({k + 1})" + "*D" * int(t.arrays["n_d"][k])
    for op in t.arrays["ops"][k]:
        op = int(op)
        if op == mh.OP_NONE:
            continue
        s += {mh.OP_BLR: "*bulkLayersReciprocal", mh.OP_SWAP: "*totalSwap/safeMantle",
              mh.OP_DIVSM: "/safeMantle"}.get(op, f"*Y({op + 1})")
    return s


def odes_f90(t: mh.NetworkTables) -> str:
    a = t.arrays
    nspec = t.nspec
    nb, ns = int(t.scalars["nbulk"]), int(t.scalars["nsurface"])
    surf, bulk = list(np.asarray(a["surface_list"])), list(np.asarray(a["bulk_list"]))
    refractory = np.asarray(a["refractory"])
    out = ["MODULE ODES\n", "USE constants\n", "USE network\n", "IMPLICIT NONE\n", "CONTAINS\n",
           "SUBROUTINE GETYDOT(RATE, Y, bulkLayersReciprocal, surfaceCoverage, safeMantle, safebulk, D, YDOT)\n",
           "REAL(dp), INTENT(IN) :: RATE(:), Y(:), bulkLayersReciprocal, safeMantle, safebulk, D\n",
           "REAL(dp), INTENT(INOUT) :: YDOT(:), surfaceCoverage\n",
           "REAL(dp) :: totalSwap, LOSS, PROD\n"]
    swap = "+".join(_term(t, int(k)) for k in a["swap_list"])
    out.append(_wrap("    totalSwap=" + (swap if swap else "0.0_dp")))
    sums = {nb: "+".join(f"YDOT({i + 1})" for i in bulk), ns: "+".join(f"YDOT({i + 1})" for i in surf)}

    def species(i: int) -> str:
        if i in sums:
            loss, prod = "", sums[i]
        else:
            loss = "+".join(_term(t, int(k)) for k in a["loss_idx"][a["loss_ptr"][i]:a["loss_ptr"][i + 1]])
            prod = "+".join(_term(t, int(k)) for k in a["prod_idx"][a["prod_ptr"][i]:a["prod_ptr"][i + 1]])
        txt = ""
        if loss:
            txt += _wrap("    LOSS = " + loss)
        if prod:
            txt += _wrap("    PROD = " + prod)
        rhs = ("PROD" if prod else "") + ("-LOSS" if loss else "")
        return txt + f"    YDOT({i + 1}) = {rhs if rhs else '0.0_dp'}\n"

    for i in range(nspec):
        out.append("\n" + species(i))
    S = f"YDOT({ns + 1})"
    out.append(f"    SURFGROWTHUNCORRECTED = {S}\n")
    out.append(f"IF ({S} .lt. 0) THEN\n    surfaceCoverage = MIN(1.0,safeBulk/safeMantle)\n")
    bps = dict(zip(surf, np.asarray(a["bulk_partner_of_surface"])))
    spb = dict(zip(bulk, np.asarray(a["surface_partner_of_bulk"])))
    for i in range(nspec):   # bulk lost, surface formed, in proportion to each bulk species
        if i in bps and bps[i] >= 0 and not refractory[bps[i]]:
            out.append(f"    YDOT({i + 1})=YDOT({i + 1})-{S}*surfaceCoverage*Y({bps[i] + 1})/safeBulk\n")
        if i in spb and not refractory[i]:
            out.append(f"    YDOT({i + 1})=YDOT({i + 1})+{S}*surfaceCoverage*Y({i + 1})/safeBulk\n")
    out.append("ELSE\n")
    for i in range(nspec):   # surface buried into the bulk
        if i in spb:
            out.append(f"    YDOT({i + 1})=YDOT({i + 1})+{S}*surfaceCoverage*Y({spb[i] + 1})\n")
        if i in bps:
            out.append(f"    YDOT({i + 1})=YDOT({i + 1})-{S}*surfaceCoverage*Y({i + 1})\n")
    out.append("ENDIF\n")
    out.append(species(nb))
    out.append(species(ns))
    out.append("    END SUBROUTINE GETYDOT\nEND MODULE ODES\n")
    return "".join(out)


def f2py_constants_f90(t: mh.NetworkTables) -> str:
    return ("MODULE F2PY_CONSTANTS\n    INTEGER, PARAMETER :: N_PHYSICS_PARAMS=8\n"
            f"    INTEGER, PARAMETER :: nspec={t.nspec}, nReac={t.nreac}\nEND MODULE F2PY_CONSTANTS\n")


def write_reference_sources(t: mh.NetworkTables, outdir: str) -> List[str]:
    """network.f90, odes.f90 and f2py_constants.f90 for `t` into outdir; returns the paths."""
    os.makedirs(outdir, exist_ok=True)
    paths = []
    for fname, text in (("network.f90", network_f90(t)), ("odes.f90", odes_f90(t)),
                        ("f2py_constants.f90", f2py_constants_f90(t))):
        p = os.path.join(outdir, fname)
        with open(p, "w") as f:
            f.write(text)
        paths.append(p)
    return paths


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="write reference-format network sources for a network")
    ap.add_argument("network", help="network name (uclchem_amd/networks/<name>.npz) or path of a .npz")
    ap.add_argument("outdir")
    args = ap.parse_args()
    path = args.network if args.network.endswith(".npz") else mh.network_path(args.network)
    print("\n".join(write_reference_sources(mh.load_network(path), args.outdir)))
