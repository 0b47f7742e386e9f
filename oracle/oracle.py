"""ORACLE / TEST INFRASTRUCTURE ONLY -- ctypes front end of the two CPU checkers.

* ``Oracle``: our C restatement (oracle/liboracle.so) of rates, RHS, physics and schedule.
* ``RefLib``: the UCLCHEM reference Fortran hot path compiled in place
  (oracle/_ref/libucref.so, built by oracle/Makefile from the reference sources).

Only tests/, the smoke check and bench.py's CPU-baseline leg import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
# UCREF_LIB selects another build of the reference library (e.g. one compiled on a written network)
REF_SO = os.environ.get("UCREF_LIB", os.path.join(HERE, "_ref", "libucref.so"))

# parameter vector layout (oracle/ucl_oracle.h, oracle/ref_driver.f90)
P_NAMES = ["dens", "temp", "zeta", "radfield", "baseav", "rout", "rin", "finaltime", "freefall",
           "finaldens", "reltol", "abstol_factor", "abstol_min", "model", "tempindx", "maxtemp",
           "freezefactor", "endatfinaldensity", "freefallfactor", "mxstep", "maxsteps",
           "metallicity", "ion", "fh", "fhe", "fc", "fo", "fn", "fs", "fmg", "fsi", "fcl", "fp", "ffe", "ff",
           "fd", "fli", "fna", "fpah", "f15n", "f13c", "f18o",
           "ebmaxh2", "ebmaxcr", "ebmaxuvcr", "epsilon", "uv_yield", "phi", "uvcreff", "omega",
           "desorb", "h2desorb", "crdesorb", "uvdesorb", "thermdesorb", "cosmicrayattenuation", "ionmodel",
           "improvedh2crpdissociation", "enforcechargeconservation", "collapse"]
P_NPARAM = 64
_F32 = lambda x: float(np.float32(x))   # noqa: E731  default-real literals
# Defaults of defaultparameters.f90:18-122 as the compiled reference holds them.
DEFAULTS = dict(dens=1.0e2, temp=10.0, zeta=1.0, radfield=1.0, baseav=2.0,
                rout=_F32(0.05), rin=0.0, finaltime=5.0e6, freefall=0.0,
                finaldens=1.0e5, reltol=1e-8, abstol_factor=1e-14, abstol_min=1e-25, model=0.0,
                tempindx=1.0, maxtemp=300.0, freezefactor=1.0, endatfinaldensity=0.0,
                freefallfactor=1.0, mxstep=10000.0, maxsteps=0.0,
                metallicity=1.0, ion=2.0, fh=0.5, fhe=_F32(0.1), fc=1.77e-04, fo=3.34e-04, fn=6.18e-05,
                fs=3.51e-6, fmg=2.256e-06, fsi=1.78e-06, fcl=3.39e-08, fp=7.78e-08, ffe=2.01e-7, ff=3.6e-08,
                fd=0.0, fli=0.0, fna=0.0, fpah=0.0, f15n=0.0, f13c=0.0, f18o=0.0,
                ebmaxh2=1.21e3, ebmaxcr=1.21e3, ebmaxuvcr=1.0e4, epsilon=_F32(0.01), uv_yield=_F32(0.03),
                phi=1.0e5, uvcreff=1.0e-3, omega=0.5,
                desorb=1.0, h2desorb=1.0, crdesorb=1.0, uvdesorb=1.0, thermdesorb=1.0, cosmicrayattenuation=0.0,
                ionmodel=0.0, improvedh2crpdissociation=0.0, enforcechargeconservation=0.0, collapse=0.0)


def make_params(**kw) -> np.ndarray:
    p = np.zeros(P_NPARAM)
    d = dict(DEFAULTS)
    for k, v in kw.items():
        if k not in d:
            raise KeyError(k)
        d[k] = v
    for i, n in enumerate(P_NAMES):
        p[i] = d[n]
    return p


def build(ref: bool = False) -> None:
    """Compile the C restatement (and, when the reference tree exists, the reference library)."""
    target = "all" if ref else "oracle"
    subprocess.check_call(["make", "-s", "-C", HERE, target])


class _Net(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int32) for n in
                 ["nspec", "neq", "nreac", "nice", "nsurfl", "nbulkl", "nfp", "nswap"]] +
                [(n, ctypes.c_void_p) for n in
                 ["mass", "alpha", "beta", "gama", "tmin", "tmax", "redmass", "binding_energy",
                  "formation_enthalpy", "n_atoms", "re", "pr", "rtype", "extrapolate",
                  "type_ranges", "ice_list", "gas_ice_list", "surface_list", "bulk_list",
                  "bulk_partner_of_surface", "freeze_partners", "n_d", "ops", "loss_ptr",
                  "loss_idx", "prod_ptr", "prod_idx", "swap_list"]] +
                [(n, ctypes.c_int32) for n in
                 ["nh", "nh2", "nhe", "nc", "ncx", "nn", "no", "nf", "np", "nmg", "nmgx", "nna",
                  "nli", "npah", "nsx", "nsix", "nclx", "nd", "n18o", "n15n", "n13c", "nelec",
                  "nco", "nsurface", "nbulk", "ngn", "ngo", "ngoh",
                  "r_h2form_ct", "r_h2form_erdes", "r_h2form_er", "r_hfreeze", "r_efreeze",
                  "r_h2freeze", "r_h2_hv", "r_co_hv", "r_c_hv", "r_h2_crp"]] +
                [("gar_params", ctypes.c_void_p), ("ngar", ctypes.c_int32),
                 ("ion_list", ctypes.c_void_p), ("nion", ctypes.c_int32)])


class _Point(ctypes.Structure):
    _fields_ = [("p", ctypes.c_double * P_NPARAM)] + [
        (n, ctypes.c_double) for n in ["density", "gas_temp", "dust_temp", "av", "cloud_size",
                                       "coldens", "phi", "h2col", "cocol", "ccol", "safe_mantle",
                                       "safe_bulk", "blr", "h2crp"]]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """C restatement bound to one network."""

    def __init__(self, tables):
        if not os.path.exists(ORACLE_SO):
            build()
        self.lib = ctypes.CDLL(ORACLE_SO)
        self.t = tables
        self._keep = {}
        a = tables.arrays
        net = _Net()
        net.nspec, net.neq, net.nreac = tables.nspec, tables.neq, tables.nreac
        net.nice = len(a["ice_list"])
        net.nsurfl = len(a["surface_list"])
        net.nbulkl = len(a["bulk_list"])
        net.nfp = len(a["freeze_partners"])
        net.nswap = len(a["swap_list"])
        conv = dict(mass=np.float64, alpha=np.float64, beta=np.float64, gama=np.float64,
                    tmin=np.float64, tmax=np.float64, reduced_mass=np.float64,
                    binding_energy=np.float64, formation_enthalpy=np.float64, n_atoms=np.int64,
                    re=np.int32, pr=np.int32, rtype=np.int32, extrapolate=np.int8,
                    type_ranges=np.int32, ice_list=np.int32, gas_ice_list=np.int32,
                    surface_list=np.int32, bulk_list=np.int32, bulk_partner_of_surface=np.int32,
                    freeze_partners=np.int32, n_d=np.int32, ops=np.int32, loss_ptr=np.int32,
                    loss_idx=np.int32, prod_ptr=np.int32, prod_idx=np.int32, swap_list=np.int32)
        for k, dt in conv.items():
            arr = np.ascontiguousarray(a[k], dtype=dt)
            self._keep[k] = arr
            setattr(net, "redmass" if k == "reduced_mass" else k, arr.ctypes.data)
        s = tables.scalars
        for n in ["nh", "nh2", "nhe", "nc", "ncx", "nn", "no", "nf", "np", "nmg", "nmgx", "nna",
                  "nli", "npah", "nsx", "nsix", "nclx", "nd", "n18o", "n15n", "n13c", "nelec",
                  "nco", "nsurface", "nbulk", "ngn", "ngo", "ngoh"]:
            setattr(net, n, s[n])
        for fld, key in [("r_h2form_ct", "nR_H2Form_CT"), ("r_h2form_erdes", "nR_H2Form_ERDes"),
                         ("r_h2form_er", "nR_H2Form_ER"), ("r_hfreeze", "nR_HFreeze"),
                         ("r_efreeze", "nR_EFreeze"), ("r_h2freeze", "nR_H2Freeze"),
                         ("r_h2_hv", "nR_H2_hv"), ("r_co_hv", "nR_CO_hv"), ("r_c_hv", "nR_C_hv"),
                         ("r_h2_crp", "nR_H2_crp")]:
            setattr(net, fld, s[key])
        gar = np.ascontiguousarray(np.asarray(a.get("gar_params", np.zeros((0, 7))), dtype=np.float64).reshape(-1))
        self._keep["gar_params"] = gar
        net.gar_params = gar.ctypes.data if len(gar) else None
        net.ngar = len(gar) // 7
        # positive ions: every species whose name holds '+' (chemistry.f90:114-120)
        ions = np.array([i for i, nm in enumerate(tables.species_names) if "+" in nm], dtype=np.int32)
        self._keep["ion_list"] = ions
        net.ion_list = ions.ctypes.data if len(ions) else None
        net.nion = len(ions)
        self.net = net
        L = self.lib
        L.ucl_next_target.restype = ctypes.c_double
        L.ucl_next_target.argtypes = [ctypes.c_int, ctypes.c_double]
        for fn in ["ucl_h2_photodiss", "ucl_co_photodiss", "ucl_co_self_shielding", "ucl_scatter",
                   "ucl_h2_form_efficiency", "ucl_c_ionization"]:
            getattr(L, fn).restype = ctypes.c_double
        L.ucl_update_physics.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]

    def init_point(self, params):
        if len(params) != P_NPARAM:
            raise ValueError(f"parameter vector must have {P_NPARAM} entries")
        pt = _Point()
        y = np.zeros(self.t.neq)
        self.lib.ucl_point_init(ctypes.byref(self.net), _ptr(np.ascontiguousarray(params, dtype=np.float64)),
                                ctypes.byref(pt), _ptr(y))
        return pt, y

    def rates(self, pt, y):
        """Returns (rate, y_after_ice_reset, scal[8]); pt is updated (phi leak)."""
        y = np.array(y, dtype=np.float64)
        rate = np.zeros(self.t.nreac)
        scal = np.zeros(8)
        self.lib.ucl_rates(ctypes.byref(self.net), ctypes.byref(pt), _ptr(y), _ptr(rate), _ptr(scal))
        return rate, y, scal

    def rhs(self, pt, rate, y):
        rate = np.array(rate, dtype=np.float64)
        ydot = np.zeros(self.t.neq)
        self.lib.ucl_rhs(ctypes.byref(self.net), ctypes.byref(pt), _ptr(rate),
                         _ptr(np.ascontiguousarray(y, dtype=np.float64)), _ptr(ydot))
        return ydot, rate

    def fluxes(self, pt, rate, y):
        """GETYDOT's per-reaction terms at y (ucl_fluxes); returns (flux, rate)."""
        rate = np.array(rate, dtype=np.float64)
        flux = np.zeros(self.t.nreac)
        self.lib.ucl_fluxes(ctypes.byref(self.net), ctypes.byref(pt), _ptr(rate),
                            _ptr(np.ascontiguousarray(y, dtype=np.float64)), _ptr(flux))
        return flux, rate

    def next_target(self, model, t_years):
        return self.lib.ucl_next_target(int(model), float(t_years))

    def update_physics(self, pt, t_s):
        self.lib.ucl_update_physics(ctypes.byref(self.net), ctypes.byref(pt), ctypes.c_double(t_s))


class RefLib:
    """The reference Fortran hot path (module globals => one point at a time per process)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: build it with `make -C oracle ref`")
        self.lib = ctypes.CDLL(path)
        neq, nr = ctypes.c_int(), ctypes.c_int()
        self.lib.ucref_sizes(ctypes.byref(neq), ctypes.byref(nr))
        self.neq, self.nreac = neq.value, nr.value
        self.nspec = self.neq - 1
        self.lib.ucref_set_phi.argtypes = [ctypes.c_double]

    def init(self, params):
        y0 = np.zeros(self.neq)
        rc = self.lib.ucref_init(_ptr(np.ascontiguousarray(params, dtype=np.float64)), _ptr(y0))
        return rc, y0

    def rates(self, y):
        y = np.array(y, dtype=np.float64)
        rate = np.zeros(self.nreac)
        scal = np.zeros(8)
        self.lib.ucref_rates(_ptr(y), _ptr(rate), _ptr(scal))
        return rate, y, scal

    def set_rates(self, rate):
        self.lib.ucref_set_rates(_ptr(np.ascontiguousarray(rate, dtype=np.float64)))

    def set_phi(self, v):
        self.lib.ucref_set_phi(ctypes.c_double(v))

    def rhs(self, y):
        ydot = np.zeros(self.neq)
        self.lib.ucref_rhs(_ptr(np.ascontiguousarray(y, dtype=np.float64)), _ptr(ydot))
        return ydot

    def model(self, params, maxrec: int = 1000):
        """Run one full single-point model; returns (rc, rows[nrec, 4+nspec], stats[8]).

        Row columns: time [yr], density, gasTemp, Av, abundances."""
        rec = np.zeros((maxrec, 4 + self.nspec))
        nrec = ctypes.c_int()
        stats = np.zeros(8, dtype=np.int32)
        rc = self.lib.ucref_model(_ptr(np.ascontiguousarray(params, dtype=np.float64)),
                                  ctypes.c_int(maxrec), _ptr(rec), ctypes.byref(nrec), _ptr(stats))
        return rc, rec[:nrec.value].copy(), stats

    def model_rates(self, params, maxrec: int = 1000):
        """ucref_model with the rate array of every output row: (rc, rows, stats, rates[nrec, nreac])."""
        rec = np.zeros((maxrec, 4 + self.nspec))
        rates = np.zeros((maxrec, self.nreac))
        nrec = ctypes.c_int()
        stats = np.zeros(8, dtype=np.int32)
        rc = self.lib.ucref_model_rates(_ptr(np.ascontiguousarray(params, dtype=np.float64)), ctypes.c_int(maxrec),
                                        _ptr(rec), ctypes.byref(nrec), _ptr(stats), _ptr(rates))
        return rc, rec[:nrec.value].copy(), stats, rates[:nrec.value].copy()

    def set_coefficient(self, which: str, k: int, v: float) -> None:
        """alpha/beta/gamma of reaction k (1-based) in the reference's module arrays."""
        code = {"alpha": 0, "beta": 1, "gamma": 2}[which]
        self.lib.ucref_set_coefficient.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double]
        if self.lib.ucref_set_coefficient(code, int(k), float(v)) != 0:
            raise ValueError(f"ucref_set_coefficient({which}, {k}) failed")

    def model_from(self, params, y0, maxrec: int = 1000):
        """The model from given start abundances y0 [neq] (the reference's abundanceStart)."""
        rec = np.zeros((maxrec, 4 + self.nspec))
        nrec = ctypes.c_int()
        stats = np.zeros(8, dtype=np.int32)
        rc = self.lib.ucref_model_from(_ptr(np.ascontiguousarray(params, dtype=np.float64)),
                                       _ptr(np.ascontiguousarray(y0, dtype=np.float64)), ctypes.c_int(maxrec),
                                       _ptr(rec), ctypes.byref(nrec), _ptr(stats))
        return rc, rec[:nrec.value].copy(), stats

    def model_until(self, params, t_stop_years: float, maxrec: int = 1000):
        """Run the model loop until the first interval ending at or beyond t_stop_years; returns
        (rc, rows, stats, y, phys) with phys = (density, gasTemp, dustTemp, Av, cloudSize, phi,
        abstol_factor, time [s]), the state the next interval starts from."""
        rec = np.zeros((maxrec, 4 + self.nspec))
        nrec = ctypes.c_int()
        stats = np.zeros(8, dtype=np.int32)
        y, phys = np.zeros(self.neq), np.zeros(8)
        rc = self.lib.ucref_model_until(_ptr(np.ascontiguousarray(params, dtype=np.float64)),
                                        ctypes.c_double(t_stop_years), ctypes.c_int(maxrec), _ptr(rec),
                                        ctypes.byref(nrec), _ptr(stats), _ptr(y), _ptr(phys))
        return rc, rec[:nrec.value].copy(), stats, y, phys

    def interval(self, params, y, phys, t_target: float, maxsteps: float = 0.0):
        """One updateChemistry interval from the state (y, phys) to t_target seconds;
        returns (rc, y_out, phys_out, stats) with stats = (steps, attempts, ...)."""
        y_out, phys_out = np.zeros(self.neq), np.zeros(8)
        stats = np.zeros(8, dtype=np.int32)
        rc = self.lib.ucref_interval(_ptr(np.ascontiguousarray(params, dtype=np.float64)),
                                     _ptr(np.ascontiguousarray(y, dtype=np.float64)),
                                     _ptr(np.ascontiguousarray(phys, dtype=np.float64)),
                                     ctypes.c_double(t_target), ctypes.c_double(maxsteps),
                                     _ptr(y_out), _ptr(phys_out), _ptr(stats))
        return rc, y_out, phys_out, stats

    def update_physics(self, t_s: float) -> float:
        """The loop's physics update after an interval ending at t_s seconds; returns gasTemp."""
        temp = ctypes.c_double()
        self.lib.ucref_update_physics(ctypes.c_double(t_s), ctypes.byref(temp))
        return temp.value

    def state(self):
        """(y, phys) of the point as the next interval would start from it."""
        y, phys = np.zeros(self.neq), np.zeros(8)
        self.lib.ucref_get_state(_ptr(y), _ptr(phys))
        return y, phys

    def network_array(self, sym, n, ctype=ctypes.c_double):
        return np.ctypeslib.as_array((ctype * n).in_dll(self.lib, sym)).copy()
