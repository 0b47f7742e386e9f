"""MakeRates backend for the MI355X solver: species/reaction lists -> device network tables.

MakeRates serialises a finished network as ``species.csv`` / ``reactions.csv``
(``src/uclchem/makerates/io_functions.py:231-282``) and as generated Fortran
(``network.f90``, ``odes.f90``).  This module consumes the CSV pair (or the same
lists built in memory, as the synthetic-network generator does) and emits the
tables the HIP library runs on:

* species tables: mass, atom counts, element counts, ice lists, binding energies
  and enthalpies, with MakeRates' ``{:.4e}``-then-default-real rounding applied
  (``io_functions.py:994-1061``; the compiled value is double(float32(x)));
* reaction tables: reactants/products (0-based, -1 = none), coefficients,
  T-ranges, reduced masses, extrapolation flags, contiguous type ranges
  (``io_functions.py:936-948``, including the 1-reaction-range quirk that makes
  ``idx1 == idx2``), freeze partners (``io_functions.py:970-991``), special
  reaction and species indices (``network.py:1165-1276``);
* the RHS "term programs": one multiplicative program per reaction restating
  ``_generate_reaction_ode_bit`` (``reaction.py:779-819``), and species-major
  loss/production lists in the order ``build_ode_string`` accumulates them
  (``io_functions.py:533-661``), so the device RHS sums terms in the same order as
  the generated Fortran;
* the analytic-Jacobian structure: for every reaction the list of state
  variables its term depends on (species, density, SURFACE via safeMantle, BULK
  via bulkLayersReciprocal, bulk species via totalSwap).

Everything is plain numpy; ``save_network``/``load_network`` use ``.npz`` with no
pickles.
"""
from __future__ import annotations

import csv
import dataclasses
import os
from collections import Counter
from typing import Dict, List, Sequence

import numpy as np

# ---------------------------------------------------------------- vocabulary

# reaction.py:5-28 (order matters: it is the order of the *Reacs range arrays)
REACTION_TYPES = [
    "PHOTON", "CRP", "CRPHOT", "FREEZE", "DESORB", "THERM", "DESOH2", "DESCR", "DEUVCR",
    "H2FORM", "ER", "ERDES", "LH", "LHDES", "BULKSWAP", "SURFSWAP", "IONOPOL1", "IONOPOL2",
    "CRS", "EXSOLID", "EXRELAX", "GAR",
]
ALL_TYPES = REACTION_TYPES + ["TWOBODY"]
TYPE_CODE = {t: i for i, t in enumerate(ALL_TYPES)}

# species.py:5-26 element list and masses
ELEMENTS = ["H", "D", "HE", "C", "N", "O", "F", "P", "S", "CL", "LI", "NA", "MG", "SI",
            "PAH", "15N", "13C", "18O", "E-", "FE"]
ELEMENT_MASS = [1, 2, 4, 12, 14, 16, 19, 31, 32, 35, 7, 23, 24, 28, 420, 15, 13, 18, 0, 56]

# Term-program opcodes (value >= 0 means "multiply by y[value]").
OP_NONE = -1
OP_BLR = -2        # * bulkLayersReciprocal
OP_SWAP = -3       # * totalSwap / safeMantle
OP_DIVSM = -4      # / safeMantle
MAX_OPS = 6
FAC_D = -5         # Jacobian factor code for one power of the density
MAX_FAC = 10

# special species looked up by name (network.py:1238-1276); missing ones map to the
# density slot nspec (0-based), exactly like MakeRates' dummy index nspec+1.
SPECIAL_SPECIES = ["C+", "H+", "H2", "HE", "HE+", "N", "N+", "O", "O+", "C", "SI+", "S+",
                   "H2O", "CH3OH", "CL", "CL+", "CO", "MG", "MG+", "#H", "#H2", "#N", "#O",
                   "#OH", "SURFACE", "BULK"] + ELEMENTS

SPECIAL_REACTIONS = ["nR_H2Form_CT", "nR_H2Form_ERDes", "nR_H2Form_ER", "nR_H2Form_LH",
                     "nR_H2Form_LHDes", "nR_HFreeze", "nR_EFreeze", "nR_H2_hv", "nR_H2_crp",
                     "nR_H2Freeze", "nR_C_hv", "nR_CO_hv"]


def species_key(name: str) -> str:
    """MakeRates' Fortran index name for a species (network.py:1270-1274)."""
    return "n" + name.lower().replace("+", "x").replace("e-", "elec").replace("#", "g")


def fortran_real(values) -> np.ndarray:
    """Value the Fortran compiler stores for a MakeRates float table entry.

    MakeRates prints ``{:.4e}`` (io_functions.py:1047) and the literal has default
    (single) kind, so the stored double is float32(decimal) widened to double."""
    arr = np.asarray(values, dtype=np.float64).ravel()
    out = np.empty(arr.shape, dtype=np.float64)
    for i, v in enumerate(arr):
        out[i] = np.float64(np.float32("{0:.4e}".format(float(v))))
    return out.reshape(np.shape(values))


def parse_atoms(name: str) -> List[str]:
    """Constituent atoms of a species name (restates species.py:244-345).

    Ice prefixes (#, @) and charges are ignored; brackets with multipliers are
    expanded.  Returns the list of element symbols, one entry per atom."""
    if name in ("BULK", "SURFACE", "E-"):
        return ["E-"] if name == "E-" else []
    s = name
    atoms: List[str] = []
    i = 0
    bracket = False
    content: List[str] = []
    while i < len(s):
        c = s[i]
        if c in "#@+-":
            i += 1
            continue
        if c == "(":
            bracket = True
            content = []
            i += 1
            continue
        if c == ")":
            bracket = False
            if i + 1 < len(s) and s[i + 1].isdigit():
                atoms.extend(content * int(s[i + 1]))
                i += 2
            else:
                atoms.extend(content)
                i += 1
            continue
        j = i
        if i + 2 < len(s) + 1 and s[i:i + 3] in ELEMENTS:
            j = i + 3
        elif s[i:i + 2] in ELEMENTS:
            j = i + 2
        elif s[i] in ELEMENTS:
            j = i + 1
        if j == i:
            raise ValueError(f"species {name!r}: unknown element at {s[i:]!r}")
        el = s[i:j]
        mult = 1
        if j < len(s) and s[j].isdigit():
            mult = int(s[j])
            j += 1
        (content if bracket else atoms).extend([el] * mult)
        i = j
    return atoms


# ---------------------------------------------------------------- data classes

@dataclasses.dataclass
class Species:
    name: str
    mass: float
    binding_energy: float = 0.0
    solid_fraction: float = 0.0
    mono_fraction: float = 0.0
    volc_fraction: float = 0.0
    enthalpy: float = 0.0


@dataclasses.dataclass
class Reaction:
    reactants: List[str]   # length 3, "NAN" for none
    products: List[str]    # length 4, "NAN" for none
    alpha: float
    beta: float
    gamma: float
    tmin: float = 0.0
    tmax: float = 10000.0
    reduced_mass: float = 0.0
    extrapolate: bool = False

    @property
    def rtype(self) -> str:
        # reaction.py:406-419
        if self.reactants[2] in REACTION_TYPES:
            return self.reactants[2]
        if self.reactants[1] in REACTION_TYPES:
            return self.reactants[1]
        return "TWOBODY"

    @property
    def body_count(self) -> int:
        # reaction.py:96-103
        n = -1
        for r in self.reactants:
            if r not in REACTION_TYPES and r != "NAN":
                n += 1
            if r in ("DESOH2", "FREEZE"):
                n += 1
            if r in ("LH", "LHDES"):
                n -= 1
        return n


# ---------------------------------------------------------------- CSV readers

def read_species_csv(path: str) -> List[Species]:
    out = []
    with open(path) as fh:
        rd = csv.reader(fh)
        next(rd)
        for row in rd:
            if not row:
                continue
            out.append(Species(row[0].strip().upper(), float(row[1]), float(row[2]),
                               float(row[3]), float(row[4]), float(row[5]), float(row[6])))
    return out


def read_reactions_csv(path: str) -> List[Reaction]:
    out = []
    with open(path) as fh:
        rd = csv.reader(fh)
        next(rd)
        for row in rd:
            if not row:
                continue
            reac = [x.strip().upper() for x in row[0:3]]
            prod = [x.strip().upper() for x in row[3:7]]
            extrap = len(row) > 13 and row[13].strip().lower() == "true"
            out.append(Reaction(reac, prod, float(row[7]), float(row[8]), float(row[9]),
                                float(row[10]), float(row[11]),
                                float(row[12]) if len(row) > 12 else 0.0, extrap))
    return out


# ---------------------------------------------------------------- the network tables

@dataclasses.dataclass
class NetworkTables:
    """All device-side tables of one network (0-based indices, -1 = none)."""
    name: str
    species_names: List[str]
    nspec: int
    neq: int
    nreac: int
    arrays: Dict[str, np.ndarray]
    scalars: Dict[str, int]

    def __getattr__(self, item):
        a = self.__dict__.get("arrays", {})
        if item in a:
            return a[item]
        s = self.__dict__.get("scalars", {})
        if item in s:
            return s[item]
        raise AttributeError(item)


def build_tables(species: Sequence[Species], reactions: Sequence[Reaction],
                 name: str = "network", gar_database: Dict[str, Sequence[float]] | None = None) -> NetworkTables:
    """gar_database: the grain-assisted recombination parameters (Weingartner & Draine 2001,
    seven per ion) keyed by ion name, needed when the network has GAR reactions; each GAR
    reaction takes its first reactant's row, in reaction order (makerates.py:107-123)."""
    names = [s.name for s in species]
    for r in reactions:
        if r.rtype in UNSUPPORTED_TYPES:
            raise ValueError(f"reaction type {r.rtype} is not supported by the device rate evaluator")
    gar_rows = []
    for r in reactions:
        if r.rtype == "GAR":
            ion = r.reactants[0]
            if not gar_database or ion not in gar_database:
                raise ValueError(f"GAR reaction of {ion} has no grain-assisted recombination parameters")
            gar_rows.append([float(v) for v in gar_database[ion]])
    gar_params = fortran_real(np.array(gar_rows, dtype=np.float64).reshape(-1, 7))
    nspec = len(names)
    neq = nspec + 1
    nreac = len(reactions)
    idx = {n: i for i, n in enumerate(names)}
    dummy = nspec  # the density slot (MakeRates' nspec+1 in 1-based Fortran)

    def sidx(n: str) -> int:
        return idx.get(n, -1)

    # --- species tables
    atoms = [parse_atoms(n) for n in names]
    n_atoms = np.array([len(a) if n not in ("E-",) else 1 for a, n in zip(atoms, names)],
                       dtype=np.int64)
    # MakeRates counts E- as one "atom"; BULK/SURFACE have 0
    elem_counts = np.zeros((nspec, len(ELEMENTS)), dtype=np.int32)
    for i, a in enumerate(atoms):
        c = Counter(a)
        for e, k in c.items():
            elem_counts[i, ELEMENTS.index(e)] = k
    charge = np.array([(1 if n.endswith("+") else (-1 if n == "E-" else 0)) for n in names],
                      dtype=np.int32)
    mass_raw = np.array([s.mass for s in species], dtype=np.float64)

    surface = [i for i, n in enumerate(names) if n.startswith("#")]
    bulk = [i for i, n in enumerate(names) if n.startswith("@")]
    ice = [i for i, n in enumerate(names) if n[0] in "#@"]
    # desorption product of each ice species: MakeRates sets it from the user's DESORB
    # entries (network.py:658-676), which survive in the serialised network as the
    # product of the species' generated desorption reactions; default: strip the prefix
    # (species.py:96-102).  Bulk species share their surface partner's product.
    desorb_product = {}
    for r in reactions:
        if r.rtype in ("DESCR", "DEUVCR", "DESOH2", "THERM") and r.reactants[0].startswith("#"):
            desorb_product.setdefault(r.reactants[0], r.products[0])
    gas_of_ice = []
    for i in ice:
        surf_name = "#" + names[i][1:]
        g = desorb_product.get(surf_name, names[i][1:])
        if g not in idx:
            raise ValueError(f"ice species {names[i]} desorbs to {g}, not in the species list")
        gas_of_ice.append(idx[g])
    bulk_partner_of_surface = []
    for i in surface:
        b = "@" + names[i][1:]
        bulk_partner_of_surface.append(idx.get(b, -1))
    surface_partner_of_bulk = [idx["#" + names[i][1:]] for i in bulk]
    ice_pos = np.full(neq, -1, dtype=np.int32)        # species -> position in iceList
    for p, i in enumerate(ice):
        ice_pos[i] = p
    gas_ice_pos = np.full(neq, -1, dtype=np.int32)    # gas species -> last position in gasIceList
    for p, g in enumerate(gas_of_ice):
        gas_ice_pos[g] = p
    is_bulk = np.zeros(neq, dtype=np.int8)
    is_bulk[bulk] = 1
    is_surface = np.zeros(neq, dtype=np.int8)
    is_surface[surface] = 1

    # --- reaction tables
    def ri(n):
        return idx.get(n, -1)
    re = np.array([[ri(r.reactants[k]) for k in range(3)] for r in reactions], dtype=np.int32).reshape(nreac, 3)
    pr = np.array([[ri(r.products[k]) for k in range(4)] for r in reactions], dtype=np.int32).reshape(nreac, 4)
    rtype = np.array([TYPE_CODE[r.rtype] for r in reactions], dtype=np.int32)
    alpha = fortran_real([r.alpha for r in reactions])
    # the serialised network already holds the final coefficients (the compiled
    # network.f90 carries beta=0 for ionic freeze-out, as reactions.csv does)
    beta = fortran_real([r.beta for r in reactions])
    gama = fortran_real([r.gamma for r in reactions])
    tmin = fortran_real([r.tmin for r in reactions])
    tmax = fortran_real([r.tmax for r in reactions])
    redmass = fortran_real([r.reduced_mass for r in reactions])
    extrap = np.array([1 if r.extrapolate else 0 for r in reactions], dtype=np.int8)

    # type ranges (io_functions.py:936-948); 0-based inclusive, (-1,-1) when absent
    ranges = np.full((len(ALL_TYPES), 2), -1, dtype=np.int32)
    for t, code in TYPE_CODE.items():
        w = np.where(rtype == code)[0]
        if len(w) > 0:
            ranges[code] = (w[0], w[-1])
    # freeze partners (io_functions.py:970-991)
    partners = []
    for r in reactions:
        if r.reactants[1] == "DESCR":
            spec = r.products[0]
            for k, r2 in enumerate(reactions):
                if r2.rtype == "FREEZE" and r2.reactants[0] == spec:
                    partners.append(k)
                    break
    freeze_partners = np.array(partners, dtype=np.int32)

    # special reactions (network.py:1165-1236)
    special = {k: -1 for k in SPECIAL_REACTIONS}
    for k, r in enumerate(reactions):
        rs, ps = r.reactants, r.products
        filt = {
            "nR_CO_hv": ("CO" in rs) and ("PHOTON" in rs) and ("O" in ps) and ("C" in ps),
            "nR_C_hv": ("C" in rs) and ("PHOTON" in rs),
            "nR_H2Form_CT": "H2FORM" in rs,
            "nR_H2Form_ERDes": ("H" in rs) and ("#H" in rs) and ("H2" in ps),
            "nR_H2Form_ER": ("H" in rs) and ("#H" in rs) and ("#H2" in ps),
            "nR_H2Form_LH": (rs.count("#H") == 2) and ("LH" in rs),
            "nR_H2Form_LHDes": (rs.count("#H") == 2) and ("LHDES" in rs),
            "nR_HFreeze": ("H" in rs) and ("FREEZE" in rs),
            "nR_H2Freeze": ("H2" in rs) and ("FREEZE" in rs),
            "nR_EFreeze": ("E-" in rs) and ("FREEZE" in rs),
            "nR_H2_hv": ("H2" in rs) and ("PHOTON" in rs),
            "nR_H2_crp": ("H2" in rs) and ("CRP" in rs) and (ps.count("H") == 2),
        }
        for key, hit in filt.items():
            if hit:
                if special[key] != -1:
                    raise ValueError(f"duplicate special reaction {key}")
                special[key] = k
    for key in ("nR_H2Form_CT", "nR_H2Form_ERDes", "nR_H2Form_ER", "nR_H2Form_LH",
                "nR_H2Form_LHDes", "nR_HFreeze", "nR_EFreeze", "nR_H2_hv"):
        if special[key] < 0:
            raise ValueError(f"network lacks mandatory reaction {key}")

    spec_idx = {}
    for n in SPECIAL_SPECIES:
        spec_idx[species_key(n)] = idx.get(n, dummy)

    # --- term programs (reaction.py:779-819)
    h_index = idx.get("H", dummy)
    n_d = np.zeros(nreac, dtype=np.int32)
    ops = np.full((nreac, MAX_OPS), OP_NONE, dtype=np.int32)
    for k, r in enumerate(reactions):
        nd = max(r.body_count, 0)
        if r.rtype == "GAR":
            nd += 1
        n_d[k] = nd
        prog = []
        for s in r.reactants:
            if s in idx:
                prog.append(idx[s])
            elif s == "BULKSWAP":
                prog.append(OP_BLR)
            elif s == "SURFSWAP":
                prog.append(OP_SWAP)
            elif s in ("DEUVCR", "DESCR", "DESOH2", "ER", "ERDES"):
                prog.append(OP_DIVSM)
                if s == "DESOH2":
                    prog.append(h_index)
            if "H2FORM" in r.reactants:
                break
        if "LH" in r.reactants[2] and "@" in r.reactants[0]:
            prog.append(OP_BLR)
        if len(prog) > MAX_OPS:
            raise ValueError("term program too long")
        ops[k, :len(prog)] = prog
        if r.body_count < 0:
            raise ValueError(f"reaction {k} has negative body count")

    # --- species-major loss / production lists (io_functions.py:563-581)
    loss: List[List[int]] = [[] for _ in range(nspec)]
    prod: List[List[int]] = [[] for _ in range(nspec)]
    swap_list: List[int] = []
    for k, r in enumerate(reactions):
        for s in r.reactants:
            if s in idx:
                if r.rtype == "ER" and not s.startswith("#"):
                    loss[idx["#" + s]].append(k)
                else:
                    loss[idx[s]].append(k)
                if r.rtype == "BULKSWAP":
                    swap_list.append(k)
        for s in r.products:
            if s in idx:
                prod[idx[s]].append(k)

    def csr(lists):
        ptr = np.zeros(len(lists) + 1, dtype=np.int32)
        for i, l in enumerate(lists):
            ptr[i + 1] = ptr[i] + len(l)
        ix = np.array([k for l in lists for k in l], dtype=np.int32)
        return ptr, ix

    loss_ptr, loss_idx = csr(loss)
    prod_ptr, prod_idx = csr(prod)

    nsurface = idx.get("SURFACE", dummy)
    nbulk = idx.get("BULK", dummy)
    if nsurface == dummy or nbulk == dummy:
        raise ValueError("three-phase networks need SURFACE and BULK species")
    # refractory bulk species are skipped in the transfer branch (io_functions.py:624-629)
    refractory = np.zeros(nspec, dtype=np.int8)

    # --- analytic-Jacobian structure
    # factor list of every term: FAC_D per density power, then the ops in emission order
    fac = np.full((nreac, MAX_FAC), OP_NONE, dtype=np.int32)
    nfac = np.zeros(nreac, dtype=np.int32)
    for k in range(nreac):
        f = [FAC_D] * int(n_d[k]) + [int(o) for o in ops[k] if o != OP_NONE]
        if len(f) > MAX_FAC:
            raise ValueError("term has too many factors")
        fac[k, :len(f)] = f
        nfac[k] = len(f)
    # species affected by each reaction with stoichiometric sign (the ER rule moves the
    # gas reactant's loss to its surface twin, io_functions.py:566-574)
    aff: List[Dict[int, int]] = [dict() for _ in range(nreac)]
    for k, r in enumerate(reactions):
        for sname in r.reactants:
            if sname in idx:
                tgt = idx["#" + sname] if (r.rtype == "ER" and not sname.startswith("#")) else idx[sname]
                aff[k][tgt] = aff[k].get(tgt, 0) - 1
        for sname in r.products:
            if sname in idx:
                aff[k][idx[sname]] = aff[k].get(idx[sname], 0) + 1
    aff_ptr = np.zeros(nreac + 1, dtype=np.int32)
    aff_sp, aff_cf = [], []
    for k in range(nreac):
        items = [(i, c) for i, c in sorted(aff[k].items()) if c != 0]
        aff_ptr[k + 1] = aff_ptr[k] + len(items)
        aff_sp += [i for i, _ in items]
        aff_cf += [c for _, c in items]

    def fac_var(code):
        if code >= 0:
            return code
        return {FAC_D: neq - 1, OP_BLR: nbulk_i, OP_DIVSM: nsurf_i, OP_SWAP: nsurf_i}[code]

    nsurf_i = idx.get("SURFACE", dummy)
    nbulk_i = idx.get("BULK", dummy)
    contrib: Dict[tuple, Dict[tuple, int]] = {}
    for k in range(nreac):
        for m in range(int(nfac[k])):
            v = fac_var(int(fac[k, m]))
            for i, c in aff[k].items():
                if c == 0:
                    continue
                d = contrib.setdefault((v, i), {})
                d[(k, m)] = d.get((k, m), 0) + c
    keys = sorted(contrib.keys())                  # column-major order (col v, then row i)
    jnz_row = np.array([i for v, i in keys], dtype=np.int32)
    jnz_col = np.array([v for v, i in keys], dtype=np.int32)
    jnz_ptr = np.zeros(len(keys) + 1, dtype=np.int32)
    jc_k, jc_m, jc_c = [], [], []
    for n_, key in enumerate(keys):
        items = sorted((km, c) for km, c in contrib[key].items() if c != 0)
        jnz_ptr[n_ + 1] = jnz_ptr[n_] + len(items)
        jc_k += [km[0] for km, _ in items]
        jc_m += [km[1] for km, _ in items]
        jc_c += [c for _, c in items]
    # SURFSWAP terms depend on every bulk species through totalSwap: per-species
    # coefficient lists for the rank-1 correction u_i = sum_k c_ik * w_k
    sw_lists: List[List[tuple]] = [[] for _ in range(neq)]
    for k in range(nreac):
        if any(int(o) == OP_SWAP for o in ops[k]):
            for i, c in aff[k].items():
                if c != 0:
                    sw_lists[i].append((k, c))
    sw_ptr = np.zeros(neq + 1, dtype=np.int32)
    for i in range(neq):
        sw_ptr[i + 1] = sw_ptr[i] + len(sw_lists[i])
    sw_k = np.array([k for l in sw_lists for k, _ in l], dtype=np.int32)
    sw_c = np.array([c for l in sw_lists for _, c in l], dtype=np.int32)
    swap_bulk = np.array([int(ops[k][0]) for k in swap_list], dtype=np.int32)  # bulk reactant of each BULKSWAP

    # --- static per-reaction constants for the device rate evaluator (network-only
    # parts of rates.f90 / surfacereactions.f90; temperature-dependent parts stay on device)
    rstat, ice_kind = _static_rate_constants(names, species, reactions, re, pr, rtype, alpha, gama,
                                             redmass, fortran_real(mass_raw), ice, gas_of_ice,
                                             fortran_real([species[i].binding_energy for i in ice]),
                                             fortran_real([species[i].enthalpy for i in ice]),
                                             n_atoms, bulk, idx, dummy)

    arrays = dict(
        rstat=rstat, ice_kind=ice_kind,
        fac=fac, nfac=nfac, aff_ptr=aff_ptr, aff_sp=np.array(aff_sp, dtype=np.int32),
        aff_cf=np.array(aff_cf, dtype=np.int32), jnz_row=jnz_row, jnz_col=jnz_col, jnz_ptr=jnz_ptr,
        jc_k=np.array(jc_k, dtype=np.int32), jc_m=np.array(jc_m, dtype=np.int32),
        jc_c=np.array(jc_c, dtype=np.int32), sw_ptr=sw_ptr, sw_k=sw_k, sw_c=sw_c,
        swap_bulk=swap_bulk,
        mass=fortran_real(mass_raw), mass_raw=mass_raw, n_atoms=n_atoms, elem_counts=elem_counts,
        charge=charge,
        surface_list=np.array(surface, dtype=np.int32), bulk_list=np.array(bulk, dtype=np.int32),
        ice_list=np.array(ice, dtype=np.int32), gas_ice_list=np.array(gas_of_ice, dtype=np.int32),
        bulk_partner_of_surface=np.array(bulk_partner_of_surface, dtype=np.int32),
        surface_partner_of_bulk=np.array(surface_partner_of_bulk, dtype=np.int32),
        ice_pos=ice_pos, gas_ice_pos=gas_ice_pos, is_bulk=is_bulk, is_surface=is_surface,
        binding_energy=fortran_real([species[i].binding_energy for i in ice]),
        formation_enthalpy=fortran_real([species[i].enthalpy for i in ice]),
        # per surface species: solid, mono and volcanic fractions (sputtering inputs that
        # network.f90 carries; the chemistry solve does not read them)
        sputter_fractions=np.array([[species[i].solid_fraction, species[i].mono_fraction, species[i].volc_fraction]
                                    for i in surface], dtype=np.float64).reshape(-1, 3),
        refractory=refractory,
        re=re, pr=pr, rtype=rtype, alpha=alpha, beta=beta, gama=gama, tmin=tmin, tmax=tmax,
        reduced_mass=redmass, extrapolate=extrap, type_ranges=ranges,
        freeze_partners=freeze_partners, gar_params=gar_params,
        n_d=n_d, ops=ops, loss_ptr=loss_ptr, loss_idx=loss_idx, prod_ptr=prod_ptr,
        prod_idx=prod_idx, swap_list=np.array(swap_list, dtype=np.int32),
    )
    scalars = dict(nspec=nspec, neq=neq, nreac=nreac, h_index=h_index, nsurface=nsurface,
                   nbulk=nbulk, n_elements=len(ELEMENTS))
    for k, v in special.items():
        scalars[k] = int(v)
    for k, v in spec_idx.items():
        scalars[k] = int(v)
    t = NetworkTables(name, names, nspec, neq, nreac, arrays, scalars)
    from .newton_plan import build_plan
    t.arrays.update(build_plan(t))      # static-pattern sparse Newton factorisation
    return t


# ---------------------------------------------------------------- static rate constants

K_BOLTZ = 1.38065040e-16
AMU = 1.66053892e-24
REDUCED_PLANCK = 1.054571628e-27
PI_F = float(np.float32(3.141592654))           # constants.f90:9 is a default-real literal
SURFACE_SITE_DENSITY = 1.5e15
CHEMICAL_BARRIER_THICKNESS = 1.40e-8
VDIFF_PREFACTOR = 2.0 * K_BOLTZ * SURFACE_SITE_DENSITY / PI_F / PI_F / AMU


def _static_rate_constants(names, species, reactions, re, pr, rtype, alpha, gama, redmass, mass,
                           ice, gas_of_ice, be, fe, n_atoms, bulk, idx, dummy):
    """Per-reaction constants that depend only on the network.

    rstat[k] holds, by type:
      THERM     [vdiff(re1), valid]                     (rates.f90:205-214)
      BULKSWAP  [vdiff(re1), E_bind(re1), valid]        (surfacereactions.f90:131-150)
      LH        [vdiff1, vdiff2, E1, E2, tunnel]        (surfacereactions.f90:158-211)
      LHDES/ERDES [desorption fraction, re1 is bulk]    (surfacereactions.f90:218-297)
    ice_kind[k]: 1 surface, 2 bulk first reactant (PHOTON/CRPHOT ice factors, rates.f90:47-56)."""
    import math
    nreac = len(reactions)
    nice = len(ice)
    ice_arr = np.array(ice)
    gas_arr = np.array(gas_of_ice)
    vdiff = np.array([math.sqrt(VDIFF_PREFACTOR * be[p] / mass[ice[p]]) for p in range(nice)])
    bulk_set = set(bulk)
    surf_set = set(i for i, n in enumerate(names) if n.startswith("#"))

    def ice_pos(sp):
        hits = np.where(ice_arr == sp)[0]
        return int(hits[-1]) if len(hits) else -1

    def desorption_fraction(k):
        r1, r2 = int(re[k, 0]), int(re[k, 1])
        ri1 = ri2 = 0
        pidx = [-1, -1, -1, -1]
        for i in range(nice):
            if ice_arr[i] == r1: ri1 = i
            if gas_arr[i] == r1: ri1 = i
            if ice_arr[i] == r2: ri2 = i
            if gas_arr[i] == r2: ri2 = i
            for q in range(4):
                pq = int(pr[k, q])
                if pq < 0:
                    continue
                if ice_arr[i] == pq: pidx[q] = i
                if gas_arr[i] == pq: pidx[q] = i
        maxbe = prodh = eps = 0.0
        for q in range(4):
            if pidx[q] >= 0:
                maxbe = max(maxbe, be[pidx[q]])
                prodh = prodh + fe[pidx[q]]
                eps = eps + mass[pidx[q]]        # indexed by ice position, as the reference does
        eps = ((eps - 120.0) / (eps + 120.0)) ** 2
        dh = fe[ri1] + fe[ri2] - prodh
        dh = dh * 4.184e03 / (1.38054e-23 * 6.02214129e23)
        dh = dh + gama[k]
        if dh == 0.0:
            dh = float(np.float32(1e-30))
        dof = int(n_atoms[pidx[0]]) if pidx[0] >= 0 else 0
        for q in range(1, 4):
            if pidx[q] >= 0:
                dof = max(dof, int(n_atoms[pidx[q]]))
        dof = 3 * dof
        f = math.exp((-maxbe * float(np.float32(dof))) / (eps * dh))
        if dh < 0.0:
            f = 0.0
        f = f / 10
        ngn, ngo, ngoh, nh = idx.get("#N", dummy), idx.get("#O", dummy), idx.get("#OH", dummy), idx.get("H", dummy)
        if r1 == ngn and r2 == ngn:
            f = 0.5
        if (r1 == ngo and r2 == nh) or (r1 == nh and r2 == ngo):
            f = float(np.float32(0.3))
        if (r1 == ngoh and r2 == nh) or (r1 == nh and r2 == ngoh):
            f = 0.25
        return f

    rstat = np.zeros((nreac, 5))
    ice_kind = np.zeros(nreac, dtype=np.int32)
    for k in range(nreac):
        t = ALL_TYPES[int(rtype[k])]
        r1 = int(re[k, 0])
        if t in ("PHOTON", "CRPHOT"):
            ice_kind[k] = 2 if r1 in bulk_set else (1 if r1 in surf_set else 0)
        elif t == "THERM":
            p = ice_pos(r1)
            if p >= 0:
                rstat[k, 0], rstat[k, 1] = vdiff[p], 1.0
        elif t == "BULKSWAP":
            p = ice_pos(r1)
            if p >= 0:
                rstat[k, 0], rstat[k, 1], rstat[k, 2] = vdiff[p], be[p], 1.0
        elif t == "LH":
            i1, i2 = int(re[k, 0]), int(re[k, 1])
            p1, p2 = ice_pos(i1), ice_pos(i2)
            i1 = p1 if p1 >= 0 else i1
            i2 = p2 if p2 >= 0 else i2
            rm = redmass[k]
            if rm == 0.0:
                m1, m2 = mass[ice[i1]], mass[ice[i2]]
                rm = m1 * m2 / (m1 + m2)
            tunnel = 2.0 * CHEMICAL_BARRIER_THICKNESS / REDUCED_PLANCK * math.sqrt(2.0 * AMU * rm * K_BOLTZ * gama[k])
            rstat[k] = (vdiff[i1], vdiff[i2], be[i1], be[i2], tunnel)
        elif t in ("LHDES", "ERDES"):
            rstat[k, 0] = desorption_fraction(k)
            rstat[k, 1] = 1.0 if r1 in bulk_set else 0.0
        elif t in ("EXRELAX", "EXSOLID"):
            # diffusion frequencies of the ice reactants (rates.f90:170-197 take the last
            # iceList match; a reactant outside the ice list would leave them undefined)
            p1 = ice_pos(r1)
            p2 = ice_pos(int(re[k, 1])) if t == "EXSOLID" else 0
            if p1 < 0 or p2 < 0:
                raise ValueError(f"{t} reaction {k} has a reactant outside the ice list")
            if t == "EXRELAX":
                rstat[k, 0] = vdiff[p1]
            else:
                rstat[k, 0] = (vdiff[p2] + vdiff[p1]) / (SURFACE_SITE_DENSITY * 1.8e-8)
    return rstat, ice_kind


UNSUPPORTED_TYPES: tuple = ()   # every reaction type of rates.f90 has a device rate law


def with_coefficients(t: NetworkTables, alpha=None, beta=None, gamma=None) -> NetworkTables:
    """A copy of the tables with reaction coefficients replaced: the reference's ``alpha``,
    ``beta`` and ``gamma`` parameter-dictionary keys, each a dictionary from the 1-based
    reaction index to its new value (wrap.f90:744-761, coefficientParser :985-1019).  The
    network constants derived from them (LH tunnelling widths and LHDES/ERDES desorption
    fractions, which the reference computes from gama at every rate evaluation) are
    recomputed, so the copy behaves as the reference does after the same assignment."""
    a = {k: v.copy() for k, v in t.arrays.items()}
    for key, changes in (("alpha", alpha), ("beta", beta), ("gama", gamma)):
        for i, v in (changes or {}).items():
            k = int(i)
            if not 1 <= k <= t.nreac:
                raise ValueError(f"{key if key != 'gama' else 'gamma'} index {i} outside 1..{t.nreac}")
            a[key][k - 1] = float(v)
    names = list(t.species_names)
    idx = {n: i for i, n in enumerate(names)}
    a["rstat"], a["ice_kind"] = _static_rate_constants(
        names, None, range(t.nreac), a["re"], a["pr"], a["rtype"], a["alpha"], a["gama"], a["reduced_mass"],
        a["mass"], [int(i) for i in a["ice_list"]], [int(i) for i in a["gas_ice_list"]], a["binding_energy"],
        a["formation_enthalpy"], a["n_atoms"], [int(i) for i in a["bulk_list"]], idx, t.nspec)
    return NetworkTables(name=t.name, species_names=list(t.species_names), nspec=t.nspec, neq=t.neq, nreac=t.nreac,
                         arrays=a, scalars=dict(t.scalars))


# ---------------------------------------------------------------- persistence

def save_network(t: NetworkTables, path: str) -> None:
    payload = {f"a_{k}": v for k, v in t.arrays.items()}
    payload["species_names"] = np.array(t.species_names, dtype="U16")
    payload["scalar_keys"] = np.array(list(t.scalars.keys()), dtype="U32")
    payload["scalar_vals"] = np.array(list(t.scalars.values()), dtype=np.int64)
    payload["name"] = np.array(t.name, dtype="U64")
    np.savez_compressed(path, **payload)


def load_network(path: str) -> NetworkTables:
    with np.load(path, allow_pickle=False) as z:
        arrays = {k[2:]: z[k] for k in z.files if k.startswith("a_")}
        names = [str(s) for s in z["species_names"]]
        scalars = {str(k): int(v) for k, v in zip(z["scalar_keys"], z["scalar_vals"])}
        name = str(z["name"])
    return NetworkTables(name, names, len(names), len(names) + 1, len(arrays["alpha"]),
                         arrays, scalars)


# ------------------------------------------------------------------ device descriptor
SPECIAL_SPECIES_ORDER = ["nh", "nh2", "nhe", "nc", "ncx", "nn", "no", "nf", "np", "nmg", "nmgx", "nna",
                         "nli", "npah", "nsx", "nsix", "nclx", "nd", "n18o", "n15n", "n13c", "nelec", "nco",
                         "nsurface", "nbulk"]
SPECIAL_REACTION_ORDER = ["nR_H2Form_CT", "nR_H2Form_ER", "nR_H2Form_ERDes", "nR_HFreeze", "nR_EFreeze",
                          "nR_H2Freeze", "nR_H2_hv", "nR_CO_hv", "nR_C_hv", "nR_H2_crp"]
DESC_SIZES = ["nspec", "neq", "nreac", "nice", "nsurfl", "nbulkl", "nfp", "nswap",
              "nloss", "nprod", "nnz", "ncontrib", "naff", "nsw"]
DESC_F64 = ["alpha", "beta", "gama", "tmin", "tmax", "mass", "rstat"]
DESC_I32 = ["rtype", "re", "pr", "ice_kind", "n_d", "ops", "fac", "nfac", "type_ranges", "loss_ptr",
            "loss_idx", "prod_ptr", "prod_idx", "swap_list", "surface_list", "bulk_list",
            "bulk_partner_of_surface", "surface_partner_of_bulk", "freeze_partners", "is_surface", "is_bulk",
            "jnz_row", "jnz_col", "jnz_ptr", "jc_k", "jc_m", "jc_c", "aff_ptr", "aff_sp", "aff_cf",
            "sw_ptr", "sw_k", "sw_c"]
DESC_I8 = ["extrapolate"]
DESC_NW = ["nw_perm", "nw_inv", "nw_sp_ptr", "nw_sp_col", "nw_sp_dpos", "nw_lev_ptr", "nw_lev_rows", "nw_ph_ptr",
           "nw_tg_slot", "nw_tg_ptr", "nw_tg_div", "nw_pr_l", "nw_pr_u", "nw_jnz_slot", "nw_diag_slot",
           "nw_ph_slot", "nw_tr_slot", "nw_s1c", "nw_s1r", "nw_s2c", "nw_s2r", "nw_gs_ptr", "nw_gs_slot",
           "nw_misc", "nw_nbrow", "nw_nsrow"]


def device_descriptor(t: NetworkTables):
    """Sizes, typed arrays and special indices of ucg_network_desc (include/uclchem_gpu.h).

    Partner arrays are expanded to length neq (-1 where a species has no partner) and
    the surface/bulk class flags are added, so the device indexes them directly."""
    a = t.arrays
    neq = t.neq
    bps = np.full(neq, -1, dtype=np.int32)
    bps[a["surface_list"]] = a["bulk_partner_of_surface"]
    spb = np.full(neq, -1, dtype=np.int32)
    spb[a["bulk_list"]] = a["surface_partner_of_bulk"]
    is_s = np.zeros(neq, dtype=np.int32)
    is_s[a["surface_list"]] = 1
    is_b = np.zeros(neq, dtype=np.int32)
    is_b[a["bulk_list"]] = 1
    arrays = {}
    for n in DESC_F64:
        arrays[n] = np.ascontiguousarray(a[n], dtype=np.float64)
    for n in DESC_I32:
        if n == "bulk_partner_of_surface":
            arrays[n] = bps
        elif n == "surface_partner_of_bulk":
            arrays[n] = spb
        elif n == "is_surface":
            arrays[n] = is_s
        elif n == "is_bulk":
            arrays[n] = is_b
        else:
            arrays[n] = np.ascontiguousarray(a[n], dtype=np.int32)
    arrays["extrapolate"] = np.ascontiguousarray(a["extrapolate"], dtype=np.int8)
    for n in DESC_NW:
        arrays[n] = np.ascontiguousarray(a[n], dtype=np.int32)
    sizes = dict(nspec=t.nspec, neq=neq, nreac=t.nreac, nice=len(a["ice_list"]),
                 nsurfl=len(a["surface_list"]), nbulkl=len(a["bulk_list"]), nfp=len(a["freeze_partners"]),
                 nswap=len(a["swap_list"]), nloss=len(a["loss_idx"]), nprod=len(a["prod_idx"]),
                 nnz=len(a["jnz_row"]), ncontrib=len(a["jc_k"]), naff=len(a["aff_sp"]), nsw=len(a["sw_k"]))
    special_species = np.array([t.scalars[k] for k in SPECIAL_SPECIES_ORDER], dtype=np.int32)
    special_reactions = np.array([t.scalars.get(k, -1) for k in SPECIAL_REACTION_ORDER], dtype=np.int32)
    sizes["nw_sizes"] = np.ascontiguousarray(a["nw_sizes"], dtype=np.int32)
    arrays["ion_list"] = ion_list(t)
    sizes["nion"] = len(arrays["ion_list"])
    for n in ("nw_erow", "nw_eion"):
        arrays[n] = np.ascontiguousarray(a[n], dtype=np.int32)
    sizes["nw_nerow"] = len(arrays["nw_erow"])
    gar = np.ascontiguousarray(np.asarray(a.get("gar_params", np.zeros((0, 7))), dtype=np.float64).reshape(-1, 7))
    arrays["gar_params"] = gar.reshape(-1)
    sizes["ngar"] = len(gar)
    return sizes, arrays, special_species, special_reactions


def ion_list(t: "NetworkTables") -> np.ndarray:
    """0-based indices of the positive ions: species whose name holds '+'
    (initializeChemistry's ionlist, chemistry.f90:114-120)."""
    return np.array([i for i, n in enumerate(t.species_names) if "+" in n], dtype=np.int32)


BIN_MAGIC = b"UCGNET01"
_BIN_DTYPES = {np.dtype(np.float64): 0, np.dtype(np.int32): 1, np.dtype(np.int8): 2}


def save_network_bin(t: NetworkTables, path: str) -> None:
    """Write the device descriptor as a flat little-endian file that the C ABI loads
    directly (ucg_network_load_file), for callers without Python (the Fortran core).

    Layout: magic "UCGNET01", int32 record count, then per record a 24-byte
    NUL-padded name, int32 dtype (0 f64, 1 i32, 2 i8), int64 element count, the data,
    and zero padding to a multiple of 8 bytes."""
    sizes, arrays, ssp, srx = device_descriptor(t)
    recs = [("sizes", np.array([sizes[k] for k in DESC_SIZES], dtype=np.int32)),
            ("special_species", ssp), ("special_reactions", srx), ("nw_sizes", sizes["nw_sizes"]),
            ("nion", np.array([sizes["nion"]], dtype=np.int32)), ("ngar", np.array([sizes["ngar"]], dtype=np.int32)),
            ("nw_nerow", np.array([sizes["nw_nerow"]], dtype=np.int32))]
    recs += [(n, arrays[n]) for n in DESC_F64 + DESC_I32 + DESC_I8 + DESC_NW +
             ["ion_list", "gar_params", "nw_erow", "nw_eion"]]
    import struct
    with open(path, "wb") as f:
        f.write(BIN_MAGIC)
        f.write(struct.pack("<i", len(recs)))
        for name, arr in recs:
            arr = np.ascontiguousarray(arr)
            f.write(name.encode().ljust(24, b"\0"))
            f.write(struct.pack("<iq", _BIN_DTYPES[arr.dtype], arr.size))
            raw = arr.astype(arr.dtype.newbyteorder("<"), copy=False).tobytes()
            f.write(raw)
            f.write(b"\0" * ((-len(raw)) % 8))


NETWORK_DIR = os.path.join(os.path.dirname(__file__), "networks")


def network_path(name: str) -> str:
    return os.path.join(NETWORK_DIR, f"{name}.npz")


def build_from_csv(species_csv: str, reactions_csv: str, name: str) -> NetworkTables:
    return build_tables(read_species_csv(species_csv), read_reactions_csv(reactions_csv), name)


if __name__ == "__main__":  # regenerate the default network from MakeRates' CSV output
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--species", default="/root/reference/src/uclchem/species.csv")
    ap.add_argument("--reactions", default="/root/reference/src/uclchem/reactions.csv")
    ap.add_argument("--name", default="default")
    a = ap.parse_args()
    tb = build_from_csv(a.species, a.reactions, a.name)
    save_network(tb, network_path(a.name))
    save_network_bin(tb, network_path(a.name)[:-4] + ".ucgnet")
    print(f"wrote {network_path(a.name)} (+ .ucgnet): {tb.nspec} species, {tb.nreac} reactions")
