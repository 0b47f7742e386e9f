"""ctypes binding of libuclchem_gpu.so (the C ABI declared in include/uclchem_gpu.h).

The HIP library is the only compute path: if it is missing or no GPU is visible every
call raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from . import makerates_hip as mh
from .params import INDEX, NPARAM

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libuclchem_gpu.so")
# torch device type of the buffers solve_device hands to the kernel (host tensors would be
# dereferenced by the GPU as device addresses)
DEVICE_TYPE = "cuda"

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)


class _Desc(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int) for n in
                 ["nspec", "neq", "nreac", "nice", "nsurfl", "nbulkl", "nfp", "nswap",
                  "nloss", "nprod", "nnz", "ncontrib", "naff", "nsw"]] +
                [(n, ctypes.c_void_p) for n in
                 ["alpha", "beta", "gama", "tmin", "tmax", "mass", "rstat", "rtype", "re", "pr",
                  "ice_kind", "n_d", "ops", "fac", "nfac", "extrapolate", "type_ranges", "loss_ptr",
                  "loss_idx", "prod_ptr", "prod_idx", "swap_list", "surface_list", "bulk_list",
                  "bulk_partner_of_surface", "surface_partner_of_bulk", "freeze_partners",
                  "is_surface", "is_bulk", "jnz_row", "jnz_col", "jnz_ptr", "jc_k", "jc_m", "jc_c",
                  "aff_ptr", "aff_sp", "aff_cf", "sw_ptr", "sw_k", "sw_c"]] +
                [("special_species", ctypes.c_int * 25), ("special_reactions", ctypes.c_int * 10)] +
                [("nw_sizes", ctypes.c_int * 10)] + [(n, ctypes.c_void_p) for n in mh.DESC_NW] +
                [("nion", ctypes.c_int), ("ion_list", ctypes.c_void_p), ("ngar", ctypes.c_int),
                 ("gar_params", ctypes.c_void_p), ("nw_nerow", ctypes.c_int), ("nw_erow", ctypes.c_void_p),
                 ("nw_eion", ctypes.c_void_p)])


class GpuError(RuntimeError):
    pass


_lib = None


def _init_torch_runtime_first():
    """torch ships its own HIP runtime; when both are used in one process, torch's must
    initialise the device before this library's (/opt/rocm) runtime does, or torch finds
    no GPU afterwards.  Device buffers are shared between the two (one GPU address space
    per process)."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def lib():
    """Load the HIP library once; raise loudly when it is not built."""
    global _lib
    if _lib is None:
        _init_torch_runtime_first()
        if not os.path.exists(LIB_PATH):
            raise GpuError(f"{LIB_PATH} is not built; run `python -m uclchem_amd.build` "
                           "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.ucg_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise GpuError(f"{what} failed ({rc}): {lib().ucg_last_error().decode()}")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _rows(a, n: int, width: int, name: str) -> np.ndarray:
    """[n, width] float64 C-contiguous copy of `a`; the library reads exactly n*width values."""
    a = np.ascontiguousarray(np.atleast_2d(a), dtype=np.float64)
    if a.shape != (n, width):
        raise ValueError(f"{name} must be [{n}, {width}], got {list(a.shape)}")
    return a


class Network:
    """A network resident on the GPU (tables uploaded once, immutable)."""

    def __init__(self, tables: mh.NetworkTables, device: int = 0, from_file: Optional[str] = None):
        """Upload `tables`; with `from_file` the library reads the same tables from a
        .ucgnet file itself (ucg_network_load_file), as a Fortran caller would."""
        self.t = tables
        self.nspec, self.neq, self.nreac = tables.nspec, tables.neq, tables.nreac
        _check(lib().ucg_init(ctypes.c_int(device)), "ucg_init")
        if from_file is not None:
            h = ctypes.c_int(-1)
            _check(lib().ucg_network_load_file(from_file.encode(), ctypes.byref(h)), "ucg_network_load_file")
            self.handle = h.value
            return
        sizes, arrays, ssp, srx = mh.device_descriptor(tables)
        self._keep = arrays
        d = _Desc()
        for k, v in sizes.items():
            if k == "nw_sizes":
                for i in range(10):
                    d.nw_sizes[i] = int(v[i])
            else:
                setattr(d, k, v)
        for k, v in arrays.items():
            setattr(d, k, _p(v).value)
        for i in range(25):
            d.special_species[i] = int(ssp[i])
        for i in range(10):
            d.special_reactions[i] = int(srx[i])
        h = ctypes.c_int(-1)
        _check(lib().ucg_network_load(ctypes.byref(d), ctypes.byref(h)), "ucg_network_load")
        self.handle = h.value

    SPARSE, DENSE = 0, 1
    CANCELLED = -10

    def cancel(self) -> None:
        """Stop running device solves on this network (unfinished points: CANCELLED)."""
        _check(lib().ucg_cancel(ctypes.c_int(self.handle)), "ucg_cancel")

    def set_linear_solver(self, kind: int) -> None:
        """Network.SPARSE (static-pattern plan, default) or Network.DENSE (full LU)."""
        _check(lib().ucg_set_linear_solver(ctypes.c_int(self.handle), ctypes.c_int(kind)), "ucg_set_linear_solver")

    # ------------------------------------------------------------ full solves
    def solve(self, params: np.ndarray, y_init: Optional[np.ndarray] = None, max_rec: int = 0,
              lu_bits: int = 64, order=None, return_rates: bool = False):
        """Host-array solve (ucg_solve_batch).  order: None (index order) or "densest"
        (points taken highest initial density first, as solve_device); results are returned
        in index order and are the same either way.  return_rates (needs max_rec > 0): also
        the rate coefficients of every output row, "rates" [n, max_rec, nreac]
        (ucg_solve_batch_rates)."""
        params = np.atleast_2d(params)
        n = params.shape[0]
        params = _rows(params, n, NPARAM, "params")
        y0 = None if y_init is None else _rows(y_init, n, self.neq, "y_init")
        if return_rates and max_rec <= 0:
            raise ValueError("return_rates needs max_rec > 0")
        if order is not None:
            if order != "densest":
                raise ValueError(f"unknown order {order!r} (expected 'densest')")
            perm = np.argsort(-params[:, INDEX["initialdens"]], kind="stable")
            out = self.solve(params[perm], None if y0 is None else y0[perm], max_rec=max_rec, lu_bits=lu_bits,
                             return_rates=return_rates)
            back = np.empty_like(perm)
            back[perm] = np.arange(n)
            return {k: (None if v is None else v[back]) for k, v in out.items()}
        y_out = np.zeros((n, self.neq))
        rec = np.zeros((n, max_rec, 4 + self.nspec)) if max_rec > 0 else None
        n_rec = np.zeros(n, dtype=np.int32)
        status = np.zeros(n, dtype=np.int32)
        stats = np.zeros((n, 8), dtype=np.int32)
        out = dict(y=y_out, rec=rec, n_rec=n_rec, status=status, stats=stats)
        if return_rates:
            rates = np.zeros((n, max_rec, self.nreac))
            rc = lib().ucg_solve_batch_rates(ctypes.c_int(self.handle), ctypes.c_int(n), _p(params), _p(y0),
                                             _p(y_out), ctypes.c_int(max_rec), _p(rec), _p(rates), _p(n_rec),
                                             _p(status), _p(stats), ctypes.c_int(lu_bits))
            _check(rc, "ucg_solve_batch_rates")
            out["rates"] = rates
            return out
        rc = lib().ucg_solve_batch(ctypes.c_int(self.handle), ctypes.c_int(n), _p(params), _p(y0), _p(y_out),
                                   ctypes.c_int(max_rec), _p(rec), _p(n_rec), _p(status), _p(stats),
                                   ctypes.c_int(lu_bits))
        _check(rc, "ucg_solve_batch")
        return out

    def interval(self, params, phys, y, t_target, lu_bits: int = 64):
        """One updateChemistry interval per point (ucg_interval): from the state (phys, y) —
        phys rows (density, gasTemp, dustTemp, Av, cloudSize, phi, abstol_factor, time [s]) —
        to t_target seconds.  Returns dict(y, phys, status, stats)."""
        params = np.atleast_2d(params)
        n = params.shape[0]
        params = _rows(params, n, NPARAM, "params")
        phys = _rows(phys, n, 8, "phys")
        y = _rows(y, n, self.neq, "y")
        tt = np.ascontiguousarray(np.broadcast_to(np.asarray(t_target, dtype=np.float64), (n,)))
        y_out = np.zeros((n, self.neq))
        phys_out = np.zeros((n, 8))
        status = np.zeros(n, dtype=np.int32)
        stats = np.zeros((n, 8), dtype=np.int32)
        _check(lib().ucg_interval(ctypes.c_int(self.handle), ctypes.c_int(n), _p(params), _p(phys), _p(y), _p(tt),
                                  _p(y_out), _p(phys_out), _p(status), _p(stats), ctypes.c_int(lu_bits)),
               "ucg_interval")
        return dict(y=y_out, phys=phys_out, status=status, stats=stats)

    def solve_profiled(self, params, lu_bits=64, n_slots=0):
        params = np.atleast_2d(params)
        n = params.shape[0]
        params = _rows(params, n, NPARAM, "params")
        y_out = np.zeros((n, self.neq))
        status = np.zeros(n, dtype=np.int32)
        stats = np.zeros((n, 8), dtype=np.int32)
        timers = np.zeros((n, 16))
        _check(lib().ucg_solve_batch_profiled(ctypes.c_int(self.handle), ctypes.c_int(n), _p(params), _p(y_out),
                                              _p(status), _p(stats), _p(timers), ctypes.c_int(lu_bits),
                                              ctypes.c_int(n_slots)), "ucg_solve_batch_profiled")
        return dict(y=y_out, status=status, stats=stats, timers=timers)

    def solve_device(self, params, y_out, status, stats, y_init=None, rec=None, n_rec=None, max_rec=0,
                     t_start=None, t_done=None, lu_bits=64, n_slots=0, stream=None, order=None):
        """Device-resident solve; arguments are torch CUDA tensors (or None).

        order: the sequence in which the workgroups take the points — None (index order),
        "densest" (schedule_order: highest initial density first) or a device tensor holding
        a permutation of 0..n-1.  Outputs stay in index order and each point's results are
        the same whatever the order; the inputs are gathered into and the outputs scattered
        from order-sorted copies on `stream`."""
        def ptr(t):
            return ctypes.c_void_p(t.data_ptr()) if t is not None else None
        n = params.shape[0]
        self._check_device_args(params, y_out, status, stats, y_init, rec, n_rec, max_rec, t_start, t_done)
        if order is not None:
            import contextlib
            import torch
            s = None
            if params.device.type == "cuda":
                s = torch.cuda.ExternalStream(stream) if stream else torch.cuda.current_stream(params.device)
            with torch.cuda.stream(s) if s is not None else contextlib.nullcontext():
                perm = schedule_order(params, order)
                outs = [y_out, status, stats, rec, n_rec, t_start, t_done]
                tmp = [t.index_select(0, perm) if t is not None else None for t in outs]
                gi = y_init.index_select(0, perm) if y_init is not None else None
                self.solve_device(params.index_select(0, perm), tmp[0], tmp[1], tmp[2], y_init=gi, rec=tmp[3],
                                  n_rec=tmp[4], max_rec=max_rec, t_start=tmp[5], t_done=tmp[6], lu_bits=lu_bits,
                                  n_slots=n_slots, stream=s.cuda_stream if s is not None else stream)
                for t, u in zip(outs, tmp):
                    if t is not None:
                        t.index_copy_(0, perm, u)
            return
        rc = lib().ucg_solve_batch_device(ctypes.c_int(self.handle), ctypes.c_int(n), ptr(params), ptr(y_init),
                                          ptr(y_out), ctypes.c_int(max_rec), ptr(rec), ptr(n_rec), ptr(status),
                                          ptr(stats), ptr(t_start), ptr(t_done), ctypes.c_int(lu_bits),
                                          ctypes.c_int(n_slots), ctypes.c_void_p(stream or 0))
        _check(rc, "ucg_solve_batch_device")

    def _check_device_args(self, params, y_out, status, stats, y_init, rec, n_rec, max_rec, t_start, t_done):
        """The kernel indexes these buffers by point with the sizes below; a wrong shape,
        dtype or layout would make it read or write outside them, so refuse it here."""
        import torch
        if params.device.type != DEVICE_TYPE:
            raise ValueError(f"solve_device takes {DEVICE_TYPE} tensors, got params on {params.device}")
        if params.dim() != 2 or params.shape[1] != NPARAM:
            raise ValueError(f"params must be [n, {NPARAM}]")
        n = params.shape[0]
        want = [("params", params, (n, NPARAM), torch.float64), ("y_out", y_out, (n, self.neq), torch.float64),
                ("status", status, (n,), torch.int32), ("stats", stats, (n, 8), torch.int32),
                ("y_init", y_init, (n, self.neq), torch.float64), ("t_start", t_start, (n,), torch.float64),
                ("t_done", t_done, (n,), torch.float64)]
        if rec is not None:
            want += [("rec", rec, (n, max_rec, 4 + self.nspec), torch.float64), ("n_rec", n_rec, (n,), torch.int32)]
        for name, t, shape, dtype in want:
            if t is None:
                if name in ("params", "y_out", "status", "stats", "n_rec"):
                    raise ValueError(f"{name} is required")
                continue
            if tuple(t.shape) != shape or t.dtype != dtype or not t.is_contiguous() or t.device != params.device:
                raise ValueError(f"{name} must be a contiguous {dtype} tensor of shape {shape} on {params.device}, "
                                 f"got {t.dtype} {tuple(t.shape)} on {t.device}")

    # ------------------------------------------------------------ pieces
    def rates(self, params, phys, y):
        params = np.atleast_2d(params)
        n = params.shape[0]
        params = _rows(params, n, NPARAM, "params")
        phys = _rows(phys, n, 8, "phys")
        y = _rows(y, n, self.neq, "y")
        rate = np.zeros((n, self.nreac))
        y_out = np.zeros((n, self.neq))
        scal = np.zeros((n, 8))
        _check(lib().ucg_rates(ctypes.c_int(self.handle), ctypes.c_int(n), _p(params), _p(phys), _p(y),
                               _p(rate), _p(y_out), _p(scal)), "ucg_rates")
        return rate, y_out, scal

    def rhs(self, params, phys, rate, y):
        params = np.atleast_2d(params)
        n = params.shape[0]
        params = _rows(params, n, NPARAM, "params")
        phys = _rows(phys, n, 8, "phys")
        rate = _rows(rate, n, self.nreac, "rate")
        y = _rows(y, n, self.neq, "y")
        ydot = np.zeros((n, self.neq))
        rate_out = np.zeros((n, self.nreac))
        _check(lib().ucg_rhs(ctypes.c_int(self.handle), ctypes.c_int(n), _p(params), _p(phys), _p(rate), _p(y),
                             _p(ydot), _p(rate_out)), "ucg_rhs")
        return ydot, rate_out

    def rhs_fluxes(self, params, phys, rate, y):
        """(ydot [n, neq], flux [n, nreac]): F and its per-reaction terms at y (ucg_rhs_fluxes)."""
        params = np.atleast_2d(params)
        n = params.shape[0]
        params = _rows(params, n, NPARAM, "params")
        phys = _rows(phys, n, 8, "phys")
        rate = _rows(rate, n, self.nreac, "rate")
        y = _rows(y, n, self.neq, "y")
        ydot = np.zeros((n, self.neq))
        flux = np.zeros((n, self.nreac))
        _check(lib().ucg_rhs_fluxes(ctypes.c_int(self.handle), ctypes.c_int(n), _p(params), _p(phys), _p(rate), _p(y),
                                    _p(ydot), _p(flux)), "ucg_rhs_fluxes")
        return ydot, flux

    def jacobian(self, params, phys, rate, y):
        """Returns J[n, neq, neq] with J[p, i, j] = d ydot_i / d y_j."""
        params = np.atleast_2d(params)
        n = params.shape[0]
        params = _rows(params, n, NPARAM, "params")
        phys = _rows(phys, n, 8, "phys")
        rate = _rows(rate, n, self.nreac, "rate")
        y = _rows(y, n, self.neq, "y")
        jac = np.zeros((n, self.neq, self.neq))   # column-major per point: [p, j, i]
        _check(lib().ucg_jacobian(ctypes.c_int(self.handle), ctypes.c_int(n), _p(params), _p(phys), _p(rate),
                                  _p(y), _p(jac)), "ucg_jacobian")
        return np.ascontiguousarray(np.transpose(jac, (0, 2, 1)))

    def newton_solve(self, params, phys, y, gamma, b, lu_bits=64):
        """x = (I - gamma J(y))^-1 b with the library's LU (rates evaluated at y)."""
        params = np.atleast_2d(params)
        n = params.shape[0]
        params = _rows(params, n, NPARAM, "params")
        phys = _rows(phys, n, 8, "phys")
        y = _rows(y, n, self.neq, "y")
        b = _rows(b, n, self.neq, "b")
        x = np.zeros((n, self.neq))
        _check(lib().ucg_newton_solve(ctypes.c_int(self.handle), ctypes.c_int(n), _p(params), _p(phys), _p(y),
                                      ctypes.c_double(gamma), _p(b), _p(x), ctypes.c_int(lu_bits)),
               "ucg_newton_solve")
        return x


def schedule_order(params, order):
    """The permutation solve_device(order=...) takes the points in, as an int64 tensor on
    params' device.

    "densest": initial density descending, ties in index order.  A point's integration cost
    grows with its density (freeze-out and grain-surface chemistry get stiffer), and a batch
    ends when its last point ends, so the costliest points are best started first (on the
    seeded 4096-point static-cloud grid the slowest points are all denser than 8e5 cm^-3).
    A tensor is checked to be a permutation of 0..n-1."""
    import torch
    n = params.shape[0]
    if isinstance(order, str):
        if order != "densest":
            raise ValueError(f"unknown order {order!r} (expected 'densest' or a permutation tensor)")
        return torch.sort(params[:, INDEX["initialdens"]], descending=True, stable=True).indices
    perm = torch.as_tensor(order).to(device=params.device, dtype=torch.int64)
    if perm.shape != (n,) or not torch.equal(torch.sort(perm).values, torch.arange(n, device=params.device)):
        raise ValueError("order must be a permutation of 0..n-1")
    return perm


def load_network(name: str = "default", device: int = 0) -> Network:
    return Network(mh.load_network(mh.network_path(name)), device=device)
