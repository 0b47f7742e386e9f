"""CPU checks of the build-time sparse Newton plan (uclchem_amd/newton_plan.py).

The plan's schedule is executed here in numpy exactly as the device executes it (Crout
phases, dense tail LU, level-ordered triangular solves) on matrices with the plan's
pattern, and compared with LAPACK (numpy.linalg.solve); every assembly slot map is
checked to address the (row, col) it claims."""
import numpy as np
import pytest

from uclchem_amd import newton_plan as npl


@pytest.fixture(scope="module")
def plan(tables):
    return npl.build_plan(tables)


def slot_table(plan):
    """(new row, new col) of every slot."""
    s = npl.plan_summary(plan)
    N, ns, T, nsp = s["N"], s["ns"], s["T"], s["nsp"]
    rc = np.empty((nsp + T * T, 2), dtype=np.int64)
    ptr, col = plan["nw_sp_ptr"], plan["nw_sp_col"]
    for r in range(N):
        for q in range(ptr[r], ptr[r + 1]):
            rc[q] = (r, col[q])
    for c in range(T):
        for r in range(T):
            rc[nsp + c * T + r] = (ns + r, ns + c)
    return rc


def execute(plan, v, b_new):
    """The device algorithm, sequentially: factor v in place, solve for b (new order)."""
    s = npl.plan_summary(plan)
    N, ns, T, nsp = s["N"], s["ns"], s["T"], s["nsp"]
    v = v.copy()
    ph, tg, tp, dv, pl, pu = (plan[k] for k in ("nw_ph_ptr", "nw_tg_slot", "nw_tg_ptr", "nw_tg_div", "nw_pr_l", "nw_pr_u"))
    for p in range(len(ph) - 1):
        new = {}
        for t in range(ph[p], ph[p + 1]):       # targets of one phase read only earlier phases
            acc = v[tg[t]] - np.dot(v[pl[tp[t]:tp[t + 1]]], v[pu[tp[t]:tp[t + 1]]])
            if dv[t] >= 0:
                acc /= v[dv[t]]
            new[tg[t]] = acc
        for k, x in new.items():
            v[k] = x
    A = v[nsp:].reshape(T, T).T.copy()          # tail, [row][col]
    for k in range(T):
        A[k + 1:, k] /= A[k, k]
        A[k + 1:, k + 1:] -= np.outer(A[k + 1:, k], A[k, k + 1:])
    ptr, col, dpos = plan["nw_sp_ptr"], plan["nw_sp_col"], plan["nw_sp_dpos"]
    lp, lr = plan["nw_lev_ptr"], plan["nw_lev_rows"]
    y = b_new.astype(float).copy()
    for l in range(len(lp) - 1):
        for r in lr[lp[l]:lp[l + 1]]:
            q = np.arange(ptr[r], dpos[r])
            y[r] -= np.dot(v[q], y[col[q]])
    for r in range(ns, N):
        q = np.arange(ptr[r], ptr[r + 1])
        y[r] -= np.dot(v[q], y[col[q]])
    yt = y[ns:].copy()
    for k in range(T):
        yt[k + 1:] -= A[k + 1:, k] * yt[k]
    for k in range(T - 1, -1, -1):
        yt[k] /= A[k, k]
        yt[:k] -= A[:k, k] * yt[k]
    y[ns:] = yt
    for l in range(len(lp) - 2, -1, -1):
        for r in lr[lp[l]:lp[l + 1]]:
            q = np.arange(dpos[r] + 1, ptr[r + 1])
            y[r] = (y[r] - np.dot(v[q], y[col[q]])) / v[dpos[r]]
    return y


def test_plan_sizes(plan, tables):
    s = npl.plan_summary(plan)
    assert s["N"] == tables.neq + 2
    assert 0 < s["ns"] < s["N"] and s["T"] == s["N"] - s["ns"]
    # the point of the plan: far less work than the dense factorisation DVODE does
    dense_mads = tables.neq ** 3 / 3
    assert s["npair"] + s["tail_lu_mads"] < dense_mads / 10
    perm, inv = plan["nw_perm"], plan["nw_inv"]
    assert np.array_equal(np.sort(perm), np.arange(s["N"])) and np.array_equal(perm[inv], np.arange(s["N"]))


def test_schedule_solves_random_systems(plan):
    rc = slot_table(plan)
    s = npl.plan_summary(plan)
    N, nsp = s["N"], s["nsp"]
    rng = np.random.default_rng(11)
    # structural entries of the original pattern are the slots that assembly writes;
    # fill slots start at zero.  Use every slot the assembly maps name.
    struct = set()
    for k in ("nw_jnz_slot", "nw_diag_slot", "nw_ph_slot", "nw_tr_slot", "nw_s1c", "nw_s1r", "nw_s2c", "nw_s2r",
              "nw_misc", "nw_nbrow", "nw_nsrow"):
        struct |= {int(x) for x in plan[k] if x >= 0}
    for trial in range(3):
        v = np.zeros(len(rc))
        idx = np.array(sorted(struct))
        v[idx] = rng.uniform(-1, 1, len(idx))
        diag = [q for q in idx if rc[q, 0] == rc[q, 1]]
        v[diag] += 30.0 * (1 + trial)             # diagonally dominant: no pivoting needed
        A = np.zeros((N, N))
        A[rc[idx, 0], rc[idx, 1]] = v[idx]
        b = rng.standard_normal(N)
        x = execute(plan, v, b)
        ref = np.linalg.solve(A, b)
        assert np.allclose(x, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())


def test_assembly_maps_address_their_entries(plan, tables):
    rc = slot_table(plan)
    inv = plan["nw_inv"]
    a = tables.arrays
    n = tables.neq
    S1, S2 = n, n + 1
    NB, NS = tables.scalars["nbulk"], tables.scalars["nsurface"]

    def at(q, i, j):
        return tuple(rc[q]) == (inv[i], inv[j])

    assert all(at(q, i, j) for q, i, j in zip(plan["nw_jnz_slot"], a["jnz_row"], a["jnz_col"]))
    assert all(at(q, i, i) for i, q in enumerate(plan["nw_diag_slot"]) if q >= 0)
    assert plan["nw_diag_slot"][NB] == -1 and plan["nw_diag_slot"][NS] == -1
    assert all(at(q, i, S1) for i, q in enumerate(plan["nw_s1c"]) if q >= 0)
    assert all(at(q, S1, j) for j, q in enumerate(plan["nw_s1r"]) if q >= 0)
    assert all(at(q, i, S2) for i, q in enumerate(plan["nw_s2c"]) if q >= 0)
    assert all(at(q, S2, j) for j, q in enumerate(plan["nw_s2r"]) if q >= 0)
    m = plan["nw_misc"]
    assert at(m[0], S1, S1) and at(m[1], S2, S2) and at(m[2], S2, S1) and at(m[3], NB, NB) and at(m[4], NS, NS)
    assert all(at(q, NB, b) for q, b in zip(plan["nw_nbrow"], a["bulk_list"]))
    assert all(at(q, NS, s) for q, s in zip(plan["nw_nsrow"], a["surface_list"]))
    tr = plan["nw_tr_slot"].reshape(-1, 8)
    for q, (s, b) in enumerate(zip(a["surface_list"], a["bulk_partner_of_surface"])):
        want = [(r, c) for r in (s, b) for c in (s, b, NS, NB)]
        assert all(at(tr[q, m_], r, c) for m_, (r, c) in enumerate(want))
    # gsum0 lists: the generated entries of the surface rows, by column
    surf = set(int(s) for s in a["surface_list"])
    jslot = plan["nw_jnz_slot"]
    for j in range(n):
        got = sorted(plan["nw_gs_slot"][plan["nw_gs_ptr"][j]:plan["nw_gs_ptr"][j + 1]])
        want = sorted(int(jslot[q]) for q in range(len(jslot)) if a["jnz_col"][q] == j and int(a["jnz_row"][q]) in surf)
        assert got == want


def test_charge_row_slots(plan, tables):
    """The charge row lists every assembly slot of the electron row and its positive-ion
    entries (ion_list order)."""
    from uclchem_amd.makerates_hip import ion_list
    rc = slot_table(plan)
    inv = plan["nw_inv"]
    a = tables.arrays
    NE = tables.scalars["nelec"]
    ions = ion_list(tables)
    assert len(ions) > 0 and len(plan["nw_eion"]) == len(ions)
    erow = set(int(q) for q in plan["nw_erow"])
    assert all(rc[q, 0] == inv[NE] for q in erow)
    assert all(tuple(rc[q]) == (inv[NE], inv[i]) for q, i in zip(plan["nw_eion"], ions))
    assert set(int(q) for q in plan["nw_eion"]) <= erow
    assert int(plan["nw_diag_slot"][NE]) in erow
    assert {int(q) for q, i in zip(plan["nw_jnz_slot"], a["jnz_row"]) if i == NE} <= erow


def test_charge_row_replacement_is_exact(plan, tables):
    """With charge enforcement the Newton matrix's electron row is e_e - gamma sum_ion J_ion.
    The device clears the electron row's slots, writes 1 and -1 (e_e - sum_ion e_ion) and
    subtracts the ion entries of the right-hand side from the electron's: the plan's
    schedule on that system gives the solution of the true matrix (numpy.linalg.solve)."""
    from uclchem_amd.makerates_hip import ion_list
    rc = slot_table(plan)
    s = npl.plan_summary(plan)
    N = s["N"]
    perm, inv = plan["nw_perm"], plan["nw_inv"]
    NE = tables.scalars["nelec"]
    ions = ion_list(tables)
    struct = set()
    for k in ("nw_jnz_slot", "nw_diag_slot", "nw_ph_slot", "nw_tr_slot", "nw_s1c", "nw_s1r", "nw_s2c", "nw_s2r",
              "nw_misc", "nw_nbrow", "nw_nsrow"):
        struct |= {int(x) for x in plan[k] if x >= 0}
    idx = np.array(sorted(struct))
    rng = np.random.default_rng(5)
    for trial in range(2):
        v = np.zeros(len(rc))
        v[idx] = rng.uniform(-1, 1, len(idx))
        v[[q for q in idx if rc[q, 0] == rc[q, 1]]] += 30.0
        A = np.zeros((N, N))                       # original indices
        A[perm[rc[idx, 0]], perm[rc[idx, 1]]] = v[idx]
        Mt = A.copy()
        Mt[NE] = np.eye(N)[NE] + (A[ions] - np.eye(N)[ions]).sum(axis=0)
        b = rng.standard_normal(N)
        v[plan["nw_erow"]] = 0.0
        v[plan["nw_eion"]] = -1.0
        v[plan["nw_diag_slot"][NE]] = 1.0
        bt = b.copy()
        bt[NE] -= b[ions].sum()
        x = np.empty(N)
        x[perm] = execute(plan, v, bt[perm])
        ref = np.linalg.solve(Mt, b)
        assert np.allclose(x, ref, rtol=1e-10, atol=1e-11 * np.abs(ref).max())
