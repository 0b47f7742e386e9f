"""Serving loop: continuous batching of single-point solves over several resident networks.

The reference runs one model per call (``uclchem.model.cloud`` -> ``uclchemwrap.cloud``,
``src/uclchem/model.py:277-289``, ``src/fortran_src/wrap.f90:15-80``).  Here every network
of the mix keeps a request queue open on the GPU (``ucg_queue_*`` in
include/uclchem_gpu.h): a persistent solve kernel with ``n_slots`` resident workgroups
that takes each posted request as soon as one of its workgroups is free.  Requests that
arrive while others are running therefore start without waiting for a batch boundary,
and the work of all networks overlaps on the device.

By default the networks share one queue (``ucg_queue_open_shared``): its resident workgroups
take the next posted request whatever its network, so capacity follows the load -- a burst
on one network runs on every slot while the others are quiet.  With ``shared=False`` each
network keeps a queue of its own with a fixed share of the CUs.

Admission by HBM budget: the slot count is cut so that the workspaces (slots x per-slot
bytes, ``ucg_queue_slot_bytes``) stay within the budget; requests beyond the free slots wait
inside the device queue (posted, not yet taken).  Network tables and workspaces stay
resident for the life of the server.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Union

import numpy as np

from .gpu import GpuError, Network, _check, _p, lib
from .params import NPARAM, make_params


class RequestQueue:
    """ctypes wrapper of one ucg_queue on a loaded network."""

    def __init__(self, net: Network, capacity: int, n_slots: int, lu_bits: int = 64, idle_timeout_s: float = 600.0):
        self.net = net
        q = ctypes.c_int(-1)
        _check(lib().ucg_queue_open(ctypes.c_int(net.handle), ctypes.c_int(capacity), ctypes.c_int(n_slots),
                                    ctypes.c_int(lu_bits), ctypes.c_double(idle_timeout_s), ctypes.byref(q)),
               "ucg_queue_open")
        self.id = q.value
        self.open = True

    @staticmethod
    def slot_bytes(net: Network, lu_bits: int = 64) -> int:
        b = ctypes.c_size_t(0)
        _check(lib().ucg_queue_slot_bytes(ctypes.c_int(net.handle), ctypes.c_int(lu_bits), ctypes.byref(b)),
               "ucg_queue_slot_bytes")
        return int(b.value)

    def submit(self, params: np.ndarray) -> int:
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(NPARAM)
        r = ctypes.c_int(-1)
        _check(lib().ucg_queue_submit(ctypes.c_int(self.id), _p(p), ctypes.byref(r)), "ucg_queue_submit")
        return r.value

    def done(self, request: int) -> bool:
        d = ctypes.c_int(0)
        _check(lib().ucg_queue_test(ctypes.c_int(self.id), ctypes.c_int(request), ctypes.byref(d)), "ucg_queue_test")
        return bool(d.value)

    def result(self, request: int):
        y = np.zeros(self.net.neq)
        st = ctypes.c_int(0)
        stats = np.zeros(8, dtype=np.int32)
        _check(lib().ucg_queue_result(ctypes.c_int(self.id), ctypes.c_int(request), _p(y), ctypes.byref(st),
                                      _p(stats)), "ucg_queue_result")
        return y, st.value, stats

    def times(self, request: int):
        """(taken, finished) of a finished request, seconds of the device wall clock."""
        return _queue_times(self.id, request)

    def close(self) -> None:
        if self.open:
            self.open = False
            _check(lib().ucg_queue_close(ctypes.c_int(self.id)), "ucg_queue_close")


class SharedQueue:
    """ctypes wrapper of one ucg_queue shared by several loaded networks: requests name their
    network by its index in `nets`."""

    def __init__(self, nets: List[Network], capacity: int, n_slots: int, lu_bits: int = 64,
                 idle_timeout_s: float = 600.0):
        self.nets = list(nets)
        handles = (ctypes.c_int * len(self.nets))(*[n.handle for n in self.nets])
        q = ctypes.c_int(-1)
        _check(lib().ucg_queue_open_shared(handles, ctypes.c_int(len(self.nets)), ctypes.c_int(capacity),
                                           ctypes.c_int(n_slots), ctypes.c_int(lu_bits),
                                           ctypes.c_double(idle_timeout_s), ctypes.byref(q)),
               "ucg_queue_open_shared")
        self.id = q.value
        self.open = True
        self._neq_max = max(n.neq for n in self.nets)

    def submit(self, params: np.ndarray, network: int) -> int:
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(NPARAM)
        r = ctypes.c_int(-1)
        _check(lib().ucg_queue_submit_to(ctypes.c_int(self.id), ctypes.c_int(network), _p(p), ctypes.byref(r)),
               "ucg_queue_submit_to")
        return r.value

    def done(self, request: int) -> bool:
        d = ctypes.c_int(0)
        _check(lib().ucg_queue_test(ctypes.c_int(self.id), ctypes.c_int(request), ctypes.byref(d)), "ucg_queue_test")
        return bool(d.value)

    def result(self, request: int, network: int):
        y = np.zeros(self._neq_max)          # the library writes the request's own network's NEQ
        st = ctypes.c_int(0)
        stats = np.zeros(8, dtype=np.int32)
        _check(lib().ucg_queue_result(ctypes.c_int(self.id), ctypes.c_int(request), _p(y), ctypes.byref(st),
                                      _p(stats)), "ucg_queue_result")
        return y[:self.nets[network].neq].copy(), st.value, stats

    def times(self, request: int):
        """(taken, finished) of a finished request, seconds of the device wall clock."""
        return _queue_times(self.id, request)

    def close(self) -> None:
        if self.open:
            self.open = False
            _check(lib().ucg_queue_close(ctypes.c_int(self.id)), "ucg_queue_close")


def _queue_times(qid: int, request: int):
    a, b = ctypes.c_double(0.0), ctypes.c_double(0.0)
    _check(lib().ucg_queue_times(ctypes.c_int(qid), ctypes.c_int(request), ctypes.byref(a), ctypes.byref(b)),
           "ucg_queue_times")
    return a.value, b.value


@dataclass
class Ticket:
    network: str
    request: int
    t_arrival: float
    t_done: float = 0.0
    y: Optional[np.ndarray] = None
    status: int = 0
    stats: Optional[np.ndarray] = None


@dataclass
class Server:
    """Resident networks served by one shared request queue (default) or one queue each.

    slots: resident workgroups -- an int, or {network: slots} (summed for a shared queue, the
    per-network shares otherwise), e.g. from :func:`plan_shared_slots` / :func:`plan_slots`;
    capacity: requests the server will take (per network without sharing)."""
    networks: Dict[str, Network]
    slots: Union[Dict[str, int], int]
    capacity: int = 4096
    lu_bits: int = 64
    shared: bool = True
    queues: Dict[str, RequestQueue] = field(default_factory=dict)
    pending: List[Ticket] = field(default_factory=list)
    shared_queue: Optional[SharedQueue] = None

    def __post_init__(self):
        names = list(self.networks)
        if self.shared:
            n = self.slots if isinstance(self.slots, int) else sum(self.slots.values())
            self._index = {k: i for i, k in enumerate(names)}
            self.shared_queue = SharedQueue([self.networks[k] for k in names], self.capacity, int(n), self.lu_bits)
        else:
            for name in names:
                self.queues[name] = RequestQueue(self.networks[name], self.capacity, self.slots[name], self.lu_bits)

    def submit(self, network: str, params: np.ndarray, t_arrival: Optional[float] = None) -> Ticket:
        if self.shared_queue is not None:
            r = self.shared_queue.submit(params, self._index[network])
        else:
            r = self.queues[network].submit(params)
        t = Ticket(network, r, time.perf_counter() if t_arrival is None else t_arrival)
        self.pending.append(t)
        return t

    def poll(self) -> List[Ticket]:
        """Tickets that finished since the last poll (results fetched)."""
        out, keep = [], []
        now = time.perf_counter()
        for t in self.pending:
            q = self.shared_queue if self.shared_queue is not None else self.queues[t.network]
            if q.done(t.request):
                t.t_done = now
                if self.shared_queue is not None:
                    t.y, t.status, t.stats = q.result(t.request, self._index[t.network])
                else:
                    t.y, t.status, t.stats = q.result(t.request)
                out.append(t)
            else:
                keep.append(t)
        self.pending = keep
        return out

    def cancel(self) -> None:
        """Stops the server's queues (ucg_cancel on the network whose cancel word each queue
        watches): running requests end with UCG_STATUS_CANCELLED, waiting ones are not run."""
        nets = [self.shared_queue.nets[0]] if self.shared_queue is not None else [q.net for q in self.queues.values()]
        for net in nets:
            _check(lib().ucg_cancel(ctypes.c_int(net.handle)), "ucg_cancel")

    def close(self) -> None:
        if self.shared_queue is not None:
            self.shared_queue.close()
        for q in self.queues.values():
            q.close()


def plan_shared_slots(networks: Dict[str, Network], n_cu: int, hbm_budget_bytes: float, lu_bits: int = 64) -> int:
    """Resident workgroups of a shared queue: the CUs less four (left to other work), cut so
    that the workspaces fit the HBM budget.  A shared slot holds the largest network's
    Newton matrix, J-form values and rates; the sum of the networks' slot sizes bounds it."""
    per = sum(RequestQueue.slot_bytes(net, lu_bits) for net in networks.values())
    n = min(max(1, n_cu - 4), int(hbm_budget_bytes // per))
    if n < 1:
        raise GpuError("HBM budget too small for one shared slot")
    return n


def plan_slots(networks: Dict[str, Network], n_cu: int, hbm_budget_bytes: float, lu_bits: int = 64,
               weights: Optional[Dict[str, float]] = None) -> Dict[str, int]:
    """Resident workgroups per network: a share of the CUs in proportion to `weights`
    (default equal), reduced until the workspaces fit the HBM budget.  The total stays
    within the CU count so that every queue's workgroups are resident together."""
    names = list(networks)
    w = np.array([(weights or {}).get(n, 1.0) for n in names], dtype=float)
    share = np.maximum(1, np.floor((n_cu - 4) * w / w.sum())).astype(int)
    per = np.array([RequestQueue.slot_bytes(networks[n], lu_bits) for n in names], dtype=float)
    while (share * per).sum() > hbm_budget_bytes and share.max() > 1:
        share[int(np.argmax(np.where(share > 1, share * per, -1.0)))] -= 1   # never below one slot
    if (share * per).sum() > hbm_budget_bytes:
        raise GpuError("HBM budget too small for one slot per network")
    return {n: int(s) for n, s in zip(names, share)}


def poisson_requests(n: int, rate_per_s: float, networks: List[str], seed: int = 20261202,
                     t_end_range=(1e3, 1e7)):
    """Seeded request stream: exponential inter-arrival gaps (rate lambda), a network chosen
    uniformly from `networks`, t_end log-uniform in t_end_range years, physics of the
    single static cloud point (n_H=1e4, T=10 K, zeta=1, G0=1, default Av)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gaps = rng.exponential(1.0 / rate_per_s, n)
    arrive = np.cumsum(gaps) - gaps[0]
    which = rng.integers(0, len(networks), n)
    t_end = 10 ** rng.uniform(np.log10(t_end_range[0]), np.log10(t_end_range[1]), n)
    reqs = []
    for i in range(n):
        p = make_params(initialdens=1e4, initialtemp=10.0, zeta=1.0, radfield=1.0, finaltime=float(t_end[i]))
        reqs.append((float(arrive[i]), networks[int(which[i])], p))
    return reqs


def run_stream(server: Server, requests, poll_interval_s: float = 2e-4, timeout_s: float = 600.0,
               progress=None) -> List[Ticket]:
    """Replay `requests` (arrival offset s, network, params) in real time against the
    server; returns every finished ticket (latency = t_done - t_arrival)."""
    done: List[Ticket] = []
    t0 = time.perf_counter()
    i = 0
    last = t0
    while len(done) < len(requests):
        now = time.perf_counter()
        while i < len(requests) and now - t0 >= requests[i][0]:
            server.submit(requests[i][1], requests[i][2], t_arrival=t0 + requests[i][0])
            i += 1
        done += server.poll()
        if progress and now - last > 10.0:
            progress(len(done), i, now - t0)
            last = now
        if now - t0 > timeout_s:
            server.cancel()   # so that closing the queues does not wait for the unfinished requests
            raise GpuError(f"stream not finished after {timeout_s} s ({len(done)}/{len(requests)} done)")
        time.sleep(poll_interval_s)
    return done
