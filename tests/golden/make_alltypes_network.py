"""Write the all-types check network (tests/golden/net_alltypes.npz): UCLCHEM's default network
(MakeRates' species.csv / reactions.csv of the reference) plus reactions of every type the
default network lacks, so that each rate law of rates.f90 runs on the device and in the
reference Fortran compiled on this network (oracle/_ref/net_alltypes, its network.f90 and
odes.f90 written by uclchem_amd/reference_format.py):

* IONOPOL1 / IONOPOL2 (rates.f90:301-313): ion-dipole captures between species of the
  network, alpha a branching ratio, beta a Langevin rate, gamma a dipole moment;
* CRS, EXRELAX, EXSOLID (rates.f90:158-197): cosmic-ray processing of ices, relaxation and a
  solid-phase reaction, on ice species of the network;
* GAR (rates.f90:316-332): grain-assisted recombination of C+, S+, Si+ and Mg+ with the
  Weingartner & Draine (2001) parameters the reference ships
  (Makerates/data/databases/weingartner01_grain_assisted_recombination.yaml values);
* DESORB: a user desorption route, whose rate the reference leaves at zero.

Every added reaction conserves atoms and charge.  The network has no excited species, so the
excited-state laws act on ground-state ices: EXRELAX (rate = the reactant's diffusion
frequency, ~1e12 s^-1, rates.f90:164-175) breaks two trace radicals up, and EXSOLID carries a
small branching ratio (alpha) so that its diffusion-limited rate does not consume the major
ices at once; the rate laws themselves are checked for any values.  Reactions are ordered by type as MakeRates
orders them (alphabetically, TWOBODY last), so the reference's range loops see each type as
one contiguous range.

Run in the build container:  python tests/golden/make_alltypes_network.py
"""
from __future__ import annotations

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from uclchem_amd import makerates_hip as mh  # noqa: E402

SPECIES = "/root/reference/src/uclchem/species.csv"
REACTIONS = "/root/reference/src/uclchem/reactions.csv"
OUT = os.path.join(HERE, "net_alltypes.npz")

# Weingartner & Draine (2001) parameters (the reference's grain-assisted recombination data)
GAR = {"C+": [45.58e-14, 6.089e-3, 1.128, 4.331e+2, 4.845e-2, 0.8120, 1.333e-4],
       "MG+": [2.510e-14, 8.116e-8, 1.864, 6.170e+4, 2.169e-6, 0.9605, 7.232e-5],
       "SI+": [2.166e-14, 5.678e-8, 1.874, 4.375e+4, 1.635e-6, 0.8964, 7.538e-5],
       "S+": [3.064e-14, 7.769e-5, 1.319, 1.087e+2, 3.475e-1, 0.4790, 4.689e-2]}

N = "NAN"
EXTRA = [
    # reactants (3), products (4), alpha, beta, gamma
    (("H3+", "H2O", N), ("H3O+", "H2", N, N), 1.0, 4.5e-9, 1.85),        # IONOPOL1
    (("HCO+", "H2O", N), ("H3O+", "CO", N, N), 1.0, 2.1e-9, 1.85),
    (("H3+", "NH3", N), ("NH4+", "H2", N, N), 1.0, 4.4e-9, 1.47),        # IONOPOL2 (set below)
    (("C+", "NH3", N), ("HCN+", "H2", N, N), 0.5, 2.3e-9, 1.47),
    (("#H2O", "CRS", N), ("#OH", "#H", N, N), 1.0, 1.747, 1.287e-15),    # CRS
    (("#CO", "CRS", N), ("#C", "#O", N, N), 1.0, 1.269, 1.287e-15),
    (("#HCO", "EXRELAX", N), ("#H", "#CO", N, N), 1.0, 0.0, 0.0),        # EXRELAX
    (("#HNO", "EXRELAX", N), ("#H", "#NO", N, N), 1.0, 0.0, 0.0),
    (("#O", "#CO", "EXSOLID"), ("#CO2", N, N, N), 1e-15, 0.0, 0.0),      # EXSOLID
    (("#H", "#OH", "EXSOLID"), ("#H2O", N, N, N), 5e-16, 0.0, 0.0),
    (("C+", "GAR", N), ("C", N, N, N), 1.0, 0.0, 0.0),                   # GAR
    (("MG+", "GAR", N), ("MG", N, N, N), 1.0, 0.0, 0.0),
    (("SI+", "GAR", N), ("SI", N, N, N), 1.0, 0.0, 0.0),
    (("S+", "GAR", N), ("S", N, N, N), 1.0, 0.0, 0.0),
    (("#CH4", "DESORB", N), ("CH4", N, N, N), 1.0, 0.0, 0.0),            # DESORB
]


def build():
    species = mh.read_species_csv(SPECIES)
    reactions = mh.read_reactions_csv(REACTIONS)
    names = {s.name for s in species}
    extra = []
    for q, (re, pr, a, b, g) in enumerate(EXTRA):
        for n in re + pr:
            assert n == N or n in names or n in mh.REACTION_TYPES, n
        re = list(re)
        if q in (2, 3):            # the IONOPOL2 pair
            re[2] = "IONOPOL2"
        elif q in (0, 1):
            re[2] = "IONOPOL1"
        extra.append(mh.Reaction(re, list(pr), a, b, g, 10.0, 41000.0))
    order = {t: i for i, t in enumerate(sorted(mh.REACTION_TYPES))}
    allr = reactions + extra
    allr.sort(key=lambda r: (1, 0) if r.rtype == "TWOBODY" else (0, order[r.rtype]))
    t = mh.build_tables(species, allr, name="alltypes", gar_database=GAR)
    mh.save_network(t, OUT)
    return t


if __name__ == "__main__":
    t = build()
    print(f"wrote {OUT}: {t.nspec} species, {t.nreac} reactions")
