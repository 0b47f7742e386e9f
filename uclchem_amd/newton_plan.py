"""Build-time plan of the static-pattern sparse Newton factorisation (MakeRates HIP backend).

DVODE factors the dense Newton matrix M = I - gamma*J with LINPACK (partial pivoting,
src/fortran_src/dvode.f90:8372-8387, 11982-12206).  Here the pattern of M is known when the
network is generated, so the ordering, the fill and the whole elimination schedule are
fixed once per network and the device only does arithmetic:

* Augmented (bordered) system.  Two couplings of UCLCHEM's RHS make J dense in blocks
  but are rank one: totalSwap (odes.f90:10, every SURFSWAP term depends on every bulk
  species) and the bulk<->surface transfer (odes.f90:4818-5153, every surface/bulk row
  gets t_s times the column sums of the surface rows).  Each becomes one extra unknown:
      sigma1 = -gamma * g1.x        (g1 = d totalSwap / dy)
      sigma2 = -gamma * gsum0.x     (gsum0 = sum of the surface rows of the sparse part)
  so that M x = b  <=>  [M0  u1 w2] [x ]   [b]
                        [r1  -1  0] [s1] = [0]
                        [r2  c1 -1] [s2]   [0]
  with M0 = I - gamma*J0 sparse, u1/w2 the rank-one column factors, r1 = -gamma*g1,
  r2 = -gamma*gsum0 and c1 = sum of u1 over the surface species (gsum of the full J
  includes the totalSwap part of the surface rows).
* Sum rows.  ydot(BULK) and ydot(SURFACE) are the sums of the bulk and surface rows
  (odes.f90:4801,4814), so their rows of M are replaced exactly by
  row(BULK) - sum_b row(b) = e_BULK - sum_b e_b (right-hand side transformed alike).
* Charge row.  With enforceChargeConservation (a per-point switch) ydot(e-) is the sum of
  the positive-ion rows (chemistry.f90:333-335), so the electron row of M is replaced the
  same way, by e_e - sum_ion e_ion; the pattern holds those entries beside the electron's
  own chemistry row, and the point's switch picks one of the two when M is formed.
* Ordering: minimum degree on the symmetrised pattern (ties by index), no pivoting.
* Split: pivots [0, ns) are eliminated sparsely (Crout dot products, one thread per
  entry, grouped in dependency levels); the trailing T = N - ns block, nearly full after
  fill, is stored dense (column-major) and factored with the blocked dense LU.

Slot numbering of the value array: the sparse part (entries with min(row, col) < ns) in
CSR order by row, then the dense tail at nsp + (col - ns) * T + (row - ns).
"""
from __future__ import annotations

from typing import Dict

import numpy as np


def _augmented_pattern(t) -> tuple:
    a = t.arrays
    n = t.neq
    N = n + 2
    S1, S2 = n, n + 1
    NB, NS = t.scalars["nbulk"], t.scalars["nsurface"]
    surf = [int(s) for s in a["surface_list"]]
    bulk = [int(b) for b in a["bulk_list"]]
    bps = [int(b) for b in a["bulk_partner_of_surface"]]
    if any(b < 0 for b in bps):
        raise ValueError("every surface species needs its bulk partner for the transfer terms")
    pat = set()
    for i, j in zip(a["jnz_row"], a["jnz_col"]):
        if int(i) in (NB, NS):
            raise ValueError("BULK/SURFACE rows are sums of rows and take no generated entries")
        pat.add((int(i), int(j)))
    for i in range(N):
        pat.add((i, i))
    aff = {}
    for k in (t.scalars["nR_H2_hv"], t.scalars["nR_CO_hv"]):
        aff[k] = [int(i) for i in a["aff_sp"][a["aff_ptr"][k]:a["aff_ptr"][k + 1]]]
        for i in aff[k]:
            for j in (t.scalars["nh2"], t.scalars["nco"], n - 1):
                pat.add((i, int(j)))
    for s, b in zip(surf, bps):
        for r in (s, b):
            for c in (s, b, NS, NB):
                pat.add((r, c))
    sw_rows = [i for i in range(n) if a["sw_ptr"][i + 1] > a["sw_ptr"][i]]
    for i in sw_rows:
        pat.add((i, S1))
    g1_cols = sorted(set(int(x) for x in a["swap_bulk"]) | {NB})
    for j in g1_cols:
        pat.add((S1, j))
    for s, b in zip(surf, bps):
        pat.add((s, S2))
        pat.add((b, S2))
    # gsum0 columns: generated entries of the surface rows
    surf_set = set(surf)
    gs_cols: Dict[int, list] = {}
    for q, (i, j) in enumerate(zip(a["jnz_row"], a["jnz_col"])):
        if int(i) in surf_set:
            gs_cols.setdefault(int(j), []).append(q)
    for j in gs_cols:
        pat.add((S2, j))
    pat.add((S2, S1))
    # the electron row's replacement under charge enforcement
    NE = t.scalars["nelec"]
    from .makerates_hip import ion_list
    ions = [int(i) for i in ion_list(t)]
    if ions and NE in (NB, NS):
        raise ValueError("the electron cannot be a BULK/SURFACE sum species")
    for i in ions:
        pat.add((NE, i))
    # replaced sum rows
    pat = {(i, j) for (i, j) in pat if i not in (NB, NS)}
    pat.add((NB, NB))
    for b in bulk:
        pat.add((NB, b))
    pat.add((NS, NS))
    for s in surf:
        pat.add((NS, s))
    return N, pat, dict(aff=aff, sw_rows=sw_rows, g1_cols=g1_cols, gs_cols=gs_cols, surf=surf, bulk=bulk,
                        bps=bps, NB=NB, NS=NS, S1=S1, S2=S2, NE=NE, ions=ions)


def minimum_degree(N: int, pat) -> np.ndarray:
    """Greedy minimum-degree elimination order of the symmetrised pattern (ties by index)."""
    adj = [set() for _ in range(N)]
    for (i, j) in pat:
        if i != j:
            adj[i].add(j)
            adj[j].add(i)
    alive = set(range(N))
    order = []
    while alive:
        v = min(alive, key=lambda x: (len(adj[x]), x))
        order.append(v)
        alive.remove(v)
        nb = adj[v]
        for u in nb:
            adj[u] |= nb
            adj[u].discard(u)
            adj[u].discard(v)
        adj[v] = set()
    return np.array(order, dtype=np.int64)


def symbolic_lu(N: int, rows_new) -> list:
    """Row structures of L+U (no pivoting) in the given order."""
    R = [set(r) for r in rows_new]
    col_rows = [set() for _ in range(N)]
    for i in range(N):
        for j in R[i]:
            col_rows[j].add(i)
    for k in range(N):
        upper = {j for j in R[k] if j > k}
        for i in sorted(x for x in col_rows[k] if x > k):
            new = upper - R[i]
            if new:
                R[i] |= new
                for j in new:
                    col_rows[j].add(i)
    return R


def choose_split(N: int, R, pair_cost: float = 4.0) -> int:
    """ns minimising (sparse dot-product work) * pair_cost + tail dense LU work."""
    cols = [set() for _ in range(N)]
    for i in range(N):
        for j in R[i]:
            cols[j].add(i)
    lower = [sorted(x for x in cols[k] if x > k) for k in range(N)]
    upper = [sorted(x for x in R[k] if x > k) for k in range(N)]
    best, best_ns = None, N
    for ns in range(N // 2, N + 1):
        T = N - ns
        ops = sum(len(lower[k]) * len(upper[k]) for k in range(ns))
        cost = pair_cost * ops + T ** 3 / 3.0
        if best is None or cost < best:
            best, best_ns = cost, ns
    return best_ns


def build_plan(t, split: int | None = None) -> Dict[str, np.ndarray]:
    N, pat, info = _augmented_pattern(t)
    n = t.neq
    perm = minimum_degree(N, pat)            # new -> old
    inv = np.empty(N, dtype=np.int64)
    inv[perm] = np.arange(N)
    rows_new = [set() for _ in range(N)]
    for (i, j) in pat:
        rows_new[inv[i]].add(int(inv[j]))
    R = symbolic_lu(N, rows_new)
    ns = choose_split(N, R) if split is None else split
    T = N - ns
    # slots
    sp_ptr = np.zeros(N + 1, dtype=np.int64)
    sp_col = []
    for r in range(N):
        cs = sorted(c for c in R[r] if min(r, c) < ns)
        sp_col += cs
        sp_ptr[r + 1] = sp_ptr[r] + len(cs)
    nsp = len(sp_col)
    sp_col = np.array(sp_col, dtype=np.int64)
    pos = {}
    for r in range(N):
        for q in range(sp_ptr[r], sp_ptr[r + 1]):
            pos[(r, int(sp_col[q]))] = q

    def slot(r, c):
        if min(r, c) < ns:
            return pos[(r, c)]
        return nsp + (c - ns) * T + (r - ns)

    def oslot(i, j):   # original indices
        return slot(int(inv[i]), int(inv[j]))

    sp_dpos = np.array([pos[(r, r)] for r in range(ns)], dtype=np.int64)
    # levels of the sparse pivots
    Lrow = [[c for c in R[r] if c < r] for r in range(N)]            # L(r, c), c < r
    cols = [set() for _ in range(N)]
    for i in range(N):
        for j in R[i]:
            cols[j].add(i)
    Ucol = [[i for i in cols[c] if i < c] for c in range(N)]         # U(i, c), i < c
    level = np.zeros(ns, dtype=np.int64)
    for k in range(ns):
        deps = [p for p in Lrow[k]] + [p for p in Ucol[k]]
        level[k] = 1 + max((level[p] for p in deps), default=-1)
    nlev = int(level.max()) + 1 if ns else 0
    lev_rows = [[k for k in range(ns) if level[k] == l] for l in range(nlev)]
    # Crout targets
    Rset = [set(r) for r in R]
    tg_slot, tg_ptr, tg_div, pr_l, pr_u = [], [0], [], [], []
    ph_ptr = [0]

    def add_target(r, c, div):
        pairs = [(slot(r, p), slot(p, c)) for p in sorted(Rset[r]) if p < min(r, c, ns) and c in Rset[p]]
        tg_slot.append(slot(r, c))
        tg_div.append(div)
        for a_, b_ in pairs:
            pr_l.append(a_)
            pr_u.append(b_)
        tg_ptr.append(len(pr_l))

    for l in range(nlev):
        for k in lev_rows[l]:                       # U row k, diagonal included
            for c in sorted(c for c in R[k] if c >= k):
                add_target(k, c, -1)
        ph_ptr.append(len(tg_slot))
        for k in lev_rows[l]:                       # L column k
            for r in sorted(cols[k]):
                if r > k:
                    add_target(r, k, slot(k, k))
        ph_ptr.append(len(tg_slot))
    for c in range(ns, N):                          # Schur complement of the tail
        for r in range(ns, N):
            if c in Rset[r] and any(p < ns and c in Rset[p] for p in Rset[r]):
                add_target(r, c, -1)
    ph_ptr.append(len(tg_slot))
    # ---------------- assembly slot maps (original indices)
    a = t.arrays
    NB, NS, S1, S2 = info["NB"], info["NS"], info["S1"], info["S2"]
    jnz_slot = np.array([oslot(int(i), int(j)) for i, j in zip(a["jnz_row"], a["jnz_col"])], dtype=np.int64)
    diag_slot = np.array([-1 if i in (NB, NS) else oslot(i, i) for i in range(n)], dtype=np.int64)
    ph_slot = []
    for k in (t.scalars["nR_H2_hv"], t.scalars["nR_CO_hv"]):
        for i in info["aff"][k]:
            ph_slot += [oslot(i, t.scalars["nh2"]), oslot(i, t.scalars["nco"]), oslot(i, n - 1)]
    tr_slot = []
    for s, b in zip(info["surf"], info["bps"]):
        for r in (s, b):
            for c in (s, b, NS, NB):
                tr_slot.append(oslot(r, c))
    s1c = np.full(n, -1, dtype=np.int64)
    for i in info["sw_rows"]:
        s1c[i] = oslot(i, S1)
    s1r = np.full(n, -1, dtype=np.int64)
    for j in info["g1_cols"]:
        s1r[j] = oslot(S1, j)
    s2c = np.full(n, -1, dtype=np.int64)
    for s, b in zip(info["surf"], info["bps"]):
        s2c[s] = oslot(s, S2)
        s2c[b] = oslot(b, S2)
    s2r = np.full(n, -1, dtype=np.int64)
    gs_ptr = np.zeros(n + 1, dtype=np.int64)
    gs_slot = []
    for j in range(n):
        if j in info["gs_cols"]:
            s2r[j] = oslot(S2, j)
            gs_slot += [int(jnz_slot[q]) for q in info["gs_cols"][j]]
        gs_ptr[j + 1] = len(gs_slot)
    misc = np.array([oslot(S1, S1), oslot(S2, S2), oslot(S2, S1), oslot(NB, NB), oslot(NS, NS)], dtype=np.int64)
    nbrow = np.array([oslot(NB, b) for b in info["bulk"]], dtype=np.int64)
    nsrow = np.array([oslot(NS, s) for s in info["surf"]], dtype=np.int64)
    # charge row: every pattern slot of the electron row (cleared), then its ion entries
    NE = info["NE"]
    erow = np.array(sorted(oslot(NE, j) for (i, j) in pat if i == NE) if info["ions"] else [], dtype=np.int64)
    eion = np.array([oslot(NE, i) for i in info["ions"]], dtype=np.int64)
    i32 = lambda x: np.ascontiguousarray(x, dtype=np.int32)
    lev_ptr = np.zeros(nlev + 1, dtype=np.int64)
    for l in range(nlev):
        lev_ptr[l + 1] = lev_ptr[l] + len(lev_rows[l])
    plan = dict(
        nw_sizes=i32([N, ns, T, nsp, nlev, len(tg_slot), len(pr_l), len(ph_ptr) - 1, len(ph_slot),
                      len(gs_slot)]),
        nw_perm=i32(perm), nw_inv=i32(inv), nw_sp_ptr=i32(sp_ptr), nw_sp_col=i32(sp_col), nw_sp_dpos=i32(sp_dpos),
        nw_lev_ptr=i32(lev_ptr), nw_lev_rows=i32([k for l in lev_rows for k in l]),
        nw_ph_ptr=i32(ph_ptr), nw_tg_slot=i32(tg_slot), nw_tg_ptr=i32(tg_ptr), nw_tg_div=i32(tg_div),
        nw_pr_l=i32(pr_l), nw_pr_u=i32(pr_u),
        nw_jnz_slot=i32(jnz_slot), nw_diag_slot=i32(diag_slot), nw_ph_slot=i32(ph_slot), nw_tr_slot=i32(tr_slot),
        nw_s1c=i32(s1c), nw_s1r=i32(s1r), nw_s2c=i32(s2c), nw_s2r=i32(s2r), nw_gs_ptr=i32(gs_ptr),
        nw_gs_slot=i32(gs_slot), nw_misc=i32(misc), nw_nbrow=i32(nbrow), nw_nsrow=i32(nsrow),
        nw_erow=i32(erow), nw_eion=i32(eion))
    return plan


NW_SIZE_NAMES = ["N", "ns", "T", "nsp", "nlev", "ntg", "npair", "nph", "nphs", "ngs"]


def plan_summary(plan) -> Dict[str, int]:
    s = dict(zip(NW_SIZE_NAMES, [int(v) for v in plan["nw_sizes"]]))
    s["values"] = s["nsp"] + s["T"] * s["T"]
    s["tail_lu_mads"] = s["T"] ** 3 // 3
    return s


LDS_BYTES_PER_CU = 160 * 1024


def lds_bytes(tables, staged: bool = True) -> int:
    """LDS one solve workgroup needs (mirror of lds_bytes/flux_region_len in the HIP
    library): the shared region (fluxes and gather partials, Jacobian term partials, and
    for a staged plan the staged factors, solve program and dense tail) plus the NEQ-long
    vectors, reduction slots, scalars and parameters."""
    a = tables.arrays
    z = a["nw_sizes"]
    N, T, nsp, nlev = int(z[0]), int(z[2]), int(z[3]), int(z[4])
    npad = (tables.neq + 7) & ~7
    Npad = (N + 7) & ~7
    npart = int(a["nfac"].sum())
    ngt = sum(int(np.sum((np.diff(a[k]) + 23) // 24)) for k in ("prod_ptr", "loss_ptr"))
    b = max(tables.nreac, ((npart + 7) & ~7) + 2 * npad, ((tables.nreac + 7) & ~7) + ngt)
    if staged:
        al = lambda x: (x + 15) & ~15
        stage = lambda nst, lut: al(nst * lut) + al(nsp * 4) + N * 16 + al((N + nlev + 1) * 4) + N * 8 + T * 8
        c = max(T * T + ((schur_sources(tables) + 7) & ~7) + Npad, (stage(nsp + T * T, 4) + 7) // 8,
                (stage(nsp, 8) + 7) // 8)
    else:
        c = Npad
    from .params import NPARAM
    return 8 * (max(b, c) + 4 * npad + 16 + 24 + NPARAM)   # + reduction slots, scalars, parameters


def schur_sources(tables) -> int:
    """Distinct sparse factor entries the last elimination phase reads when that phase is
    the pure Schur-complement update of the dense tail (0 otherwise); the library gathers
    them into LDS beside the staged tail."""
    a = tables.arrays
    nsp, nph = int(a["nw_sizes"][3]), int(a["nw_sizes"][7])
    ph, tp = a["nw_ph_ptr"], a["nw_tg_ptr"]
    if nph < 2:
        return 0
    t0, t1 = int(ph[nph - 1]), int(ph[nph])
    q0, q1 = int(tp[t0]), int(tp[t1])
    if q1 <= q0 or (a["nw_tg_slot"][t0:t1] < nsp).any() or (a["nw_tg_div"][t0:t1] >= 0).any():
        return 0
    refs = np.concatenate([a["nw_pr_l"][q0:q1], a["nw_pr_u"][q0:q1]])
    if (refs >= nsp).any():
        return 0
    return int(len(np.unique(refs)))


def plan_is_staged(tables) -> bool:
    """Whether the library stages this network's plan in LDS (else it runs it from the workspace)."""
    return lds_bytes(tables, True) <= LDS_BYTES_PER_CU
