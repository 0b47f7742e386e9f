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
from typing import Dict, List, Optional, Sequence

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
                 name: str = "network") -> NetworkTables:
    names = [s.name for s in species]
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

    arrays = dict(
        mass=fortran_real(mass_raw), mass_raw=mass_raw, n_atoms=n_atoms, elem_counts=elem_counts,
        charge=charge,
        surface_list=np.array(surface, dtype=np.int32), bulk_list=np.array(bulk, dtype=np.int32),
        ice_list=np.array(ice, dtype=np.int32), gas_ice_list=np.array(gas_of_ice, dtype=np.int32),
        bulk_partner_of_surface=np.array(bulk_partner_of_surface, dtype=np.int32),
        surface_partner_of_bulk=np.array(surface_partner_of_bulk, dtype=np.int32),
        ice_pos=ice_pos, gas_ice_pos=gas_ice_pos, is_bulk=is_bulk, is_surface=is_surface,
        binding_energy=fortran_real([species[i].binding_energy for i in ice]),
        formation_enthalpy=fortran_real([species[i].enthalpy for i in ice]),
        refractory=refractory,
        re=re, pr=pr, rtype=rtype, alpha=alpha, beta=beta, gama=gama, tmin=tmin, tmax=tmax,
        reduced_mass=redmass, extrapolate=extrap, type_ranges=ranges,
        freeze_partners=freeze_partners, gar_params=np.zeros((0, 7)),
        n_d=n_d, ops=ops, loss_ptr=loss_ptr, loss_idx=loss_idx, prod_ptr=prod_ptr,
        prod_idx=prod_idx, swap_list=np.array(swap_list, dtype=np.int32),
    )
    scalars = dict(nspec=nspec, neq=neq, nreac=nreac, h_index=h_index, nsurface=nsurface,
                   nbulk=nbulk, n_elements=len(ELEMENTS))
    for k, v in special.items():
        scalars[k] = int(v)
    for k, v in spec_idx.items():
        scalars[k] = int(v)
    return NetworkTables(name, names, nspec, neq, nreac, arrays, scalars)


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
    print(f"wrote {network_path(a.name)}: {tb.nspec} species, {tb.nreac} reactions")
