"""Shared test helpers (seeded states, tolerance metrics)."""
import numpy as np

REFERENCE = "/root/reference"


def random_state(tables, oracle, rng, params, bare_grains=False):
    """A seeded abundance vector for rate/RHS/Jacobian checks, plus the oracle point."""
    pt, y0 = oracle.init_point(params)
    y = 10 ** rng.uniform(-30, -3, tables.neq)
    y[-1] = y0[-1]
    y[tables.nh2] = 0.4
    y[tables.nh] = 0.1
    if bare_grains:
        y[tables.surface_list] = 1e-32
    return pt, y


def phys_of(pt, absfac=1e-14):
    return np.array([pt.density, pt.gas_temp, pt.dust_temp, pt.av, pt.cloud_size, pt.phi, absfac, 0.0])


def term_magnitude(tables, rate, y):
    """Per species sum of |ODE terms| (scale for RHS comparisons with cancellation)."""
    from uclchem_amd.makerates_hip import OP_BLR, OP_DIVSM, OP_NONE, OP_SWAP
    neq = tables.neq
    D = y[-1]
    sm = max(1e-30, y[tables.nsurface])
    sbk = max(1e-30, y[tables.nbulk])
    pi_f = float(np.float32(3.141592654))
    nspg = 1.0e-5 * 1.0e-5 * 1.5e15 * 4.0 * pi_f
    gdr = (4.0 * pi_f * (1e-5 ** 3) * 3.0 * 100.0) / (3.0 * 1.66053892e-24)
    blr = min(1.0, nspg / (gdr * sbk))

    def tv(k, ts):
        v = rate[k] * D ** tables.n_d[k]
        for o in tables.ops[k]:
            if o == OP_NONE:
                continue
            if o >= 0:
                v *= y[o]
            elif o == OP_BLR:
                v *= blr
            elif o == OP_SWAP:
                v = v * ts / sm
            elif o == OP_DIVSM:
                v = v / sm
        return v
    ts = sum(tv(k, 0.0) for k in tables.swap_list)
    flux = np.array([tv(k, ts) for k in range(tables.nreac)])
    mag = np.zeros(neq)
    for i in range(tables.nspec):
        mag[i] = np.abs(flux[tables.loss_idx[tables.loss_ptr[i]:tables.loss_ptr[i + 1]]]).sum() + \
            np.abs(flux[tables.prod_idx[tables.prod_ptr[i]:tables.prod_ptr[i + 1]]]).sum()
    # bulk<->surface transfer terms (odes.f90:4818-5153) move S*t_s between partners
    pre = np.zeros(neq)
    for i in tables.surface_list:
        pre[i] = flux[tables.prod_idx[tables.prod_ptr[i]:tables.prod_ptr[i + 1]]].sum() - \
            flux[tables.loss_idx[tables.loss_ptr[i]:tables.loss_ptr[i + 1]]].sum()
    S = pre[tables.surface_list].sum()
    sc0 = 0.5 * gdr / nspg
    for q, sidx in enumerate(tables.surface_list):
        b = tables.bulk_partner_of_surface[q]
        t = (min(1.0, sbk / sm) * y[b] / sbk) if S < 0 else sc0 * y[sidx]
        mag[sidx] += abs(S * t) + np.abs(pre[tables.surface_list]).sum() * t
        mag[b] += abs(S * t) + np.abs(pre[tables.surface_list]).sum() * t
    mag[tables.nbulk] = mag[tables.bulk_list].sum()
    mag[tables.nsurface] = mag[tables.surface_list].sum()
    return mag + 1e-300


def element_totals(tables, y):
    """Total abundance of each element (rows of elem_counts) over the species."""
    return tables.elem_counts[:tables.nspec].T.astype(np.float64) @ y[:tables.nspec]


def conserved_elements(tables):
    """Elements every reaction conserves.  Four freeze-outs of the default network
    hydrogenate on the grain (e.g. C2N -> #CH3CNH), so H is not an invariant of the
    reference equations; electrons follow charge, which ionic freeze-out breaks."""
    from uclchem_amd.makerates_hip import ELEMENTS
    a = tables.arrays
    ok = np.ones(len(ELEMENTS), dtype=bool)
    for k in range(tables.nreac):
        bal = np.zeros(len(ELEMENTS))
        for i, c in zip(a["aff_sp"][a["aff_ptr"][k]:a["aff_ptr"][k + 1]], a["aff_cf"][a["aff_ptr"][k]:a["aff_ptr"][k + 1]]):
            if i < tables.nspec:
                bal += c * tables.elem_counts[i]
        ok &= bal == 0
    ok[ELEMENTS.index("E-")] = False
    return np.where(ok)[0]


def excess_over_reference(y, y_ref, y_tight=None, floor=1e-14):
    """Largest |y - y_ref| / allowance over species with y_ref > floor, where the allowance is
    1% of y_ref plus, when the reference's tight-tolerance answer y_tight is given (finite),
    twice the reference's own integration noise |y_ref - y_tight|.  <= 1 passes."""
    import numpy as np
    m = y_ref > floor
    allow = 0.01 * y_ref[m]
    if y_tight is not None and np.all(np.isfinite(y_tight)):
        allow = allow + 2.0 * np.abs(y_ref[m] - y_tight[m])
    return float((np.abs(y[m] - y_ref[m]) / allow).max())
