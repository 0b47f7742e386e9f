// Variable-order, variable-step BDF integrator in fixed-leading-coefficient
// Nordsieck form with modified (chord) Newton iteration.
//
// This is the method family of the reference's DVODE_F90 with METHOD_FLAG=22
// (src/fortran_src/dvode.f90: DVODE 5654, DVHIN 6768, DVSTEP 7180, DVSET 7616,
// DVJUST 7790, DVNLSD 7926, DVNORM 9021) written from the published algorithm
// (Brown, Byrne & Hindmarsh 1989; Hindmarsh et al. 2005, ACM TOMS 31:363, whose
// CVODE formulation of the same method is followed for coefficient updates).  It
// is not a transliteration: the Jacobian is analytic and re-evaluated at every
// Newton-matrix setup instead of a saved finite-difference copy.
//
// The control logic below is scalar and identical for every thread that runs it.
// All vector work goes through the Ops policy, whose methods are collective: on
// the GPU every thread of the workgroup calls each method (each owns a slice of
// the state vectors; block reductions and the RHS/LU/solve synchronise inside).
//
// Ops::setup(gamma, new_jacobian) forms and factors M = I - gamma J, evaluating J first
// when new_jacobian (else from the J saved at the last evaluation, as DVODE's JSV=1 path
// does); Ops::can_reuse_jacobian() false forces an evaluation at every setup.
//
// Call contract of one integrate() call mirrors DVODE_F90 with ITASK=1, ISTATE=1
// (a cold start per call, src/fortran_src/chemistry.f90:241-292): integrate from
// t0 past tout, interpolate back to tout, report DVODE-style states:
//   2 success, -1 step limit, -2 tolerance too small, -4 repeated error-test
//   failures, -5 repeated convergence failures, -3 illegal input,
//   -6 singular/invalid Newton matrix or non-finite state, -7 cancelled by the caller
//   (Ops::cancelled(), checked once per step).
// On a negative return the state is the last accepted step (t = tn, y = zn[0]).
#pragma once

#if defined(__HIPCC__)
#define BDF_HD __host__ __device__ __forceinline__
#else
#define BDF_HD inline
#include <cmath>
#endif

namespace bdf {

constexpr int QMAX = 5;
constexpr double UROUND = 2.220446049250313e-16;
// step/order control constants (VODE/CVODE published defaults; the error-test failure
// limits are DVODE_F90's: order crash from the third consecutive failure (KFC = -3), give up
// at the fifteenth (KFH = -15, raised from VODE's 7), dvode.f90:1865-1876)
constexpr double ADDON = 1.0e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0, THRESH = 1.5;
constexpr double ETAMX1 = 1.0e4, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2, ETAMIN = 0.1;
#ifndef BDF_DGMAX
#define BDF_DGMAX 0.3
#endif
constexpr double ETACF = 0.25, CRDOWN = 0.3, DGMAX = BDF_DGMAX, RDIV = 2.0, NLSCOEF = 0.1;
constexpr int SMALL_NST = 10, MXNEF = 15, MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10;
constexpr int MSBP = 20, MAXCOR = 3, MXNCF = 10, HIN_ITERS = 4;
// saved-Jacobian reuse (DVODE's JSV=1 / CVODE's lsetup): re-evaluate J at most every
// MSBJ steps, after a convergence failure, or when a stale-J failure is not explained by
// a change of gamma of more than DGMAX_J
constexpr int MSBJ = 50;
constexpr double DGMAX_J = 0.2;
constexpr double HLB_FACTOR = 100.0, HUB_FACTOR = 0.1, H_BIAS = 0.5;

// vector slots the Ops policy provides
enum Vec { Z0 = 0, Z1, Z2, Z3, Z4, Z5, ACOR, EWT, Y, F, TMP, F0 };

struct Stats {
    int nst = 0, nfe = 0, nje = 0, nsetups = 0, nni = 0, ncfn = 0, netf = 0;
};

struct State {
    double tn, h, hscale, hprime, eta, etamax, etaq, etaqm1, etaqp1;
    double tau[QMAX + 2];
    double l[QMAX + 1];
    double tq[6];
    double gamma, gammap, gamrat, rl1, crate, acnrm, saved_tq5;
    int q, qprime, qwait, L, nstlp, nstlj, nst_call;
    bool jcur;
};

enum { FIRST_CALL = 0, PREV_CONV_FAIL, PREV_ERR_FAIL };

BDF_HD double dmax(double a, double b) { return a > b ? a : b; }
BDF_HD double dmin(double a, double b) { return a < b ? a : b; }
BDF_HD double dabs(double a) { return a < 0 ? -a : a; }
BDF_HD double powi(double x, int k)
{
    double r = 1.0;
    for (int i = 0; i < k; i++) r *= x;
    return r;
}

// BDF coefficients l[] and test constants tq[] for order q with step history tau[].
BDF_HD void set_bdf(State &s)
{
    const int q = s.q;
    double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
    for (int i = 0; i <= QMAX; i++) s.l[i] = 0.0;
    s.l[0] = s.l[1] = xi_inv = xistar_inv = 1.0;
    alpha0 = alpha0_hat = -1.0;
    hsum = s.h;
    if (q > 1) {
        for (int j = 2; j < q; j++) {
            hsum += s.tau[j - 1];
            xi_inv = s.h / hsum;
            alpha0 -= 1.0 / j;
            for (int i = j; i >= 1; i--) s.l[i] += s.l[i - 1] * xi_inv;
        }
        alpha0 -= 1.0 / q;
        xistar_inv = -s.l[1] - alpha0;
        hsum += s.tau[q - 1];
        xi_inv = s.h / hsum;
        alpha0_hat = -s.l[1] - xi_inv;
        for (int i = q; i >= 1; i--) s.l[i] += s.l[i - 1] * xistar_inv;
    }
    // error and convergence test constants
    double A1 = 1.0 - alpha0_hat + alpha0;
    double A2 = 1.0 + q * A1;
    s.tq[2] = dabs(A1 / (alpha0 * A2));
    s.tq[5] = dabs(A2 * xistar_inv / (s.l[q] * xi_inv));
    if (q > 1) {
        double C = xistar_inv / s.l[q];
        double A3 = alpha0 + 1.0 / q;
        double A4 = alpha0_hat + xi_inv;
        double Cpinv = (1.0 - A4 + A3) / A3;
        s.tq[1] = dabs(C * Cpinv);
        hsum += s.tau[q];
        xi_inv = s.h / hsum;
        double A5 = alpha0 - 1.0 / (q + 1);
        double A6 = alpha0_hat - xi_inv;
        double Cppinv = (1.0 - A6 + A5) / A2;
        s.tq[3] = dabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    s.tq[4] = NLSCOEF / s.tq[2];
}

// Nordsieck-array correction coefficients for an order change by +1 / -1.
BDF_HD void increase_coeffs(const State &s, double *lc, double &A1)
{
    const int q = s.q;
    for (int i = 0; i <= QMAX; i++) lc[i] = 0.0;
    lc[2] = 1.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0, hsum = s.hscale;
    if (q > 1) {
        for (int j = 1; j < q; j++) {
            hsum += s.tau[j + 1];
            double xi = hsum / s.hscale;
            prod *= xi;
            alpha0 -= 1.0 / (j + 1);
            alpha1 += 1.0 / xi;
            for (int i = j + 2; i >= 2; i--) lc[i] = lc[i] * xiold + lc[i - 1];
            xiold = xi;
        }
    }
    A1 = (-alpha0 - alpha1) / prod;
}

BDF_HD void decrease_coeffs(const State &s, double *lc)
{
    const int q = s.q;
    for (int i = 0; i <= QMAX; i++) lc[i] = 0.0;
    lc[2] = 1.0;
    double hsum = 0.0;
    for (int j = 1; j <= q - 2; j++) {
        hsum += s.tau[j];
        double xi = hsum / s.hscale;
        for (int i = j + 2; i >= 2; i--) lc[i] = lc[i] * xi + lc[i - 1];
    }
}

template <class Ops>
BDF_HD void adjust_order(Ops &o, State &s, int dq)
{
    double lc[QMAX + 1];
    if (dq == 1) {
        double A1;
        increase_coeffs(s, lc, A1);
        // zn[q+1] = A1 * saved acor (kept in zn[QMAX]); zn[j] += lc[j] * zn[q+1]
        o.scale_into_z(s.q + 1, A1, QMAX);
        for (int j = 2; j <= s.q; j++) o.axpy_zz(j, lc[j], s.q + 1);
    } else if (dq == -1) {
        decrease_coeffs(s, lc);
        for (int j = 2; j < s.q; j++) o.axpy_zz(j, -lc[j], s.q);
    }
}

template <class Ops>
BDF_HD void rescale(Ops &o, State &s)
{
    o.rescale_z(s.q, s.eta);
    s.h = s.hscale * s.eta;
    s.hscale = s.h;
}

template <class Ops>
BDF_HD void restore(Ops &o, State &s, double saved_t)
{
    s.tn = saved_t;
    o.unpredict(s.q);
}

// Newton iteration for the corrector; returns 0 converged, 1 recoverable
// convergence failure, -6 setup failure.
template <class Ops>
BDF_HD int nls(Ops &o, State &s, Stats &st, int nflag)
{
    bool call_setup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (st.nst == 0) ||
                      (st.nst >= s.nstlp + MSBP) || (dabs(s.gamrat - 1.0) > DGMAX);
    bool stale_retry = false;
    for (;;) {
        o.start_corrector();          // acor = 0, y = zn[0]
        o.rhs();                      // F = f(tn, y)
        st.nfe++;
        if (call_setup) {
            const bool jbad = (st.nst == 0) || (st.nst >= s.nstlj + MSBJ) || (nflag == PREV_CONV_FAIL) ||
                              (stale_retry && dabs(s.gamrat - 1.0) < DGMAX_J) || !o.can_reuse_jacobian();
            int ierr = o.setup(s.gamma, jbad);   // new J at y = zn[0] if jbad; M = I - gamma J; factor
            st.nsetups++;
            if (jbad) {
                st.nje++;
                s.nstlj = st.nst;
            }
            s.jcur = jbad;
            s.gamrat = 1.0;
            s.gammap = s.gamma;
            s.crate = 1.0;
            s.nstlp = st.nst;
            if (ierr) {
                if (!jbad) {          // singular with a saved J: evaluate J and try once more
                    stale_retry = false;
                    s.nstlj = -MSBJ;
                    continue;
                }
                return 1;             // singular Newton matrix: retry with a smaller step
            }
        }
        double del = 0.0, delp = 0.0;
        int m = 0;
        bool retry = false;
        for (;;) {
            st.nni++;
            o.form_residual(s.gamma, s.rl1);             // TMP = gamma F - (rl1 zn1 + acor)
            o.solve();                                    // TMP = M^-1 TMP
            double scale = (s.gamrat != 1.0) ? 2.0 / (1.0 + s.gamrat) : 1.0;
            del = o.apply_correction(scale);              // acor += scale*TMP; y = zn0 + acor
            if (!(del == del)) {                          // non-finite update
                if (!s.jcur) { retry = true; break; }
                return 1;
            }
            if (m > 0) s.crate = dmax(CRDOWN * s.crate, del / delp);
            double dcon = del * dmin(1.0, s.crate) / s.tq[4];
            if (dcon <= 1.0) {
                s.acnrm = (m == 0) ? del : o.wrms(ACOR);
                s.jcur = false;
                return 0;
            }
            m++;
            if (m == MAXCOR || (m >= 2 && del > RDIV * delp)) {
                if (!s.jcur) { retry = true; break; }
                return 1;
            }
            delp = del;
            o.rhs();
            st.nfe++;
        }
        if (retry) {
            call_setup = true;        // stale Jacobian: refresh and start the iteration again
            stale_retry = true;
            continue;
        }
    }
}

// One accepted step (or a failure code -4/-5/-6).
template <class Ops>
BDF_HD int step(Ops &o, State &s, Stats &st)
{
    double saved_t = s.tn;
    int ncf = 0, nef = 0;
    int nflag = FIRST_CALL;
    double dsm = 0.0;
    if (st.nst > 0 && s.hprime != s.h) {
        if (s.qprime != s.q) {
            adjust_order(o, s, s.qprime - s.q);
            s.q = s.qprime;
            s.L = s.q + 1;
            s.qwait = s.L;
        }
        rescale(o, s);
    }
    for (;;) {
        // predict
        s.tn += s.h;
        o.predict(s.q);
        set_bdf(s);
        s.rl1 = 1.0 / s.l[1];
        s.gamma = s.h * s.rl1;
        if (st.nst == 0) s.gammap = s.gamma;
        s.gamrat = (st.nst > 0) ? s.gamma / s.gammap : 1.0;

        int r = nls(o, s, st, nflag);
        if (r < 0) {
            restore(o, s, saved_t);
            return -6;
        }
        if (r == 1) {
            st.ncfn++;
            ncf++;
            s.etamax = 1.0;
            restore(o, s, saved_t);
            if (ncf == MXNCF) return -5;
            s.eta = ETACF;
            nflag = PREV_CONV_FAIL;
            rescale(o, s);
            continue;
        }
        // local error test
        dsm = s.acnrm * s.tq[2];
        if (dsm <= 1.0) break;
        nef++;
        st.netf++;
        nflag = PREV_ERR_FAIL;
        restore(o, s, saved_t);
        if (nef == MXNEF) return -4;
        s.etamax = 1.0;
        if (nef < MXNEF1) {
            s.eta = 1.0 / (pow(BIAS2 * dsm, 1.0 / s.L) + ADDON);
            s.eta = dmax(ETAMIN, s.eta);
            if (nef >= SMALL_NEF) s.eta = dmin(s.eta, ETAMXF);
            rescale(o, s);
            continue;
        }
        if (s.q > 1) {
            s.eta = ETAMIN;
            adjust_order(o, s, -1);
            s.L = s.q;
            s.q--;
            s.qwait = s.L;
            rescale(o, s);
            continue;
        }
        // order 1 after repeated failures: shrink and restart the history from f
        s.eta = ETAMIN;
        s.h *= s.eta;
        s.hscale = s.h;
        s.qwait = LONG_WAIT;
        o.start_corrector();
        o.rhs();
        st.nfe++;
        o.scale_into_z_from(1, s.h, F);   // zn[1] = h f(tn, zn[0])
    }
    // complete the step
    st.nst++;
    s.nst_call++;
    for (int i = s.q; i >= 2; i--) s.tau[i] = s.tau[i - 1];
    if (s.q == 1 && st.nst > 1) s.tau[2] = s.tau[1];
    s.tau[1] = s.h;
    o.complete(s.q, s.l);
    s.qwait--;
    if (s.qwait == 1 && s.q != QMAX) {
        o.copy_acor_to_z(QMAX);
        s.saved_tq5 = s.tq[5];
    }
    // choose the next step size and order; after a failure in this step keep h and q
    if (s.etamax == 1.0) {
        if (s.qwait < 2) s.qwait = 2;
        s.qprime = s.q;
        s.hprime = s.h;
        s.eta = 1.0;
        s.etamax = (st.nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
        return 0;
    }
    s.etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / s.L) + ADDON);
    if (s.qwait != 0) {
        s.eta = s.etaq;
        s.qprime = s.q;
    } else {
        s.qwait = 2;
        s.etaqm1 = 0.0;
        if (s.q > 1) {
            double ddn = o.wrms(s.q) * s.tq[1];
            s.etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / s.q) + ADDON);
        }
        s.etaqp1 = 0.0;
        if (s.q != QMAX) {
            double cquot = (s.tq[5] / s.saved_tq5) * powi(s.h / s.tau[2], s.L);
            double dup = o.wrms_acor_minus(cquot, QMAX) * s.tq[3];
            s.etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (s.L + 1)) + ADDON);
        }
        double etam = dmax(s.etaqm1, dmax(s.etaq, s.etaqp1));
        if (etam < THRESH) {
            s.eta = 1.0;
            s.qprime = s.q;
        } else if (etam == s.etaq) {
            s.eta = s.etaq;
            s.qprime = s.q;
        } else if (etam == s.etaqm1) {
            s.eta = s.etaqm1;
            s.qprime = s.q - 1;
        } else {
            s.eta = s.etaqp1;
            s.qprime = s.q + 1;
            o.copy_acor_to_z(QMAX);
        }
    }
    if (s.eta < THRESH) {
        s.eta = 1.0;
        s.hprime = s.h;
        s.qprime = s.q;
    } else {
        s.eta = dmin(s.eta, s.etamax);
        s.hprime = s.h * s.eta;
    }
    s.etamax = (st.nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
    return 0;
}

// Initial step size from an estimate of y'' (same approach as DVHIN).
template <class Ops>
BDF_HD double initial_step(Ops &o, double t0, double tout, Stats &st)
{
    double tdist = dabs(tout - t0);
    double tround = UROUND * dmax(dabs(t0), dabs(tout));
    double hlb = HLB_FACTOR * tround;
    double hub = HUB_FACTOR * tdist;
    double hub_inv = o.hub_inverse(HUB_FACTOR);    // max_i |f0_i| / (0.1|y0_i| + atol_i)
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    double hg = sqrt(hlb * hub);
    if (hub < hlb) return hg;
    double hnew = hg;
    for (int count = 1; count <= HIN_ITERS; count++) {
        double yddnrm = o.ydd_norm(t0, hg);        // ||(f(t0+hg, y0+hg f0) - f0)/hg||
        st.nfe++;
        hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
        if (!(hnew == hnew)) { hnew = hg; break; }
        if (count == HIN_ITERS) break;
        double hrat = hnew / hg;
        if (hrat > 0.5 && hrat < 2.0) break;
        if (count > 1 && hrat > 2.0) { hnew = hg; break; }
        hg = hnew;
    }
    double h0 = H_BIAS * hnew;
    if (h0 < hlb) h0 = hlb;
    if (h0 > hub) h0 = hub;
    return h0;
}

// One cold-start integration t0 -> tout (ITASK=1 semantics).  Before the call the
// Ops hold y0 in Z0 and the absolute tolerances; on return Z0 holds y(t).
template <class Ops>
BDF_HD int integrate(Ops &o, double &t, double tout, double rtol, int mxstep, Stats &st)
{
    State s;
    if (!(tout > t)) return -3;
    s.tn = t;
    if (!o.set_ewt(rtol)) return -3;
    o.start_corrector();                 // y = zn[0]
    o.rhs();                             // F = f(t0, y0)
    st.nfe++;
    o.save_f0();
    double h0 = initial_step(o, t, tout, st);
    if (!(h0 > 0.0)) return -3;
    o.scale_into_z_from(1, h0, F0);      // zn[1] = h0 f0
    s.h = s.hscale = s.hprime = h0;
    s.q = s.qprime = 1;
    s.L = 2;
    s.qwait = 2;
    s.etamax = ETAMX1;
    s.eta = 1.0;
    for (int i = 0; i < QMAX + 2; i++) s.tau[i] = 0.0;
    for (int i = 0; i < 6; i++) s.tq[i] = 0.0;
    s.gamma = s.gammap = s.gamrat = 1.0;
    s.crate = 1.0;
    s.acnrm = 0.0;
    s.saved_tq5 = 1.0;
    s.nstlp = 0;
    s.nstlj = 0;
    s.nst_call = 0;
    s.jcur = false;
    s.etaq = s.etaqm1 = s.etaqp1 = 0.0;
    Stats local;                         // per-call counters drive the step logic
    int ret = 2;
    for (;;) {
        if (s.nst_call > 0) {
            if (!o.set_ewt(rtol)) { ret = -3; break; }
        }
        if (s.nst_call >= mxstep) { ret = -1; break; }
        if (o.cancelled()) { ret = -7; break; }   // the caller asked the whole batch to stop
        double tolsf = UROUND * o.wrms(Z0);
        if (tolsf > 1.0) { ret = -2; break; }
        int k = step(o, s, local);
        if (k < 0) { ret = k; break; }
        if (s.tn >= tout) {
            o.interpolate((tout - s.tn) / s.h, s.q);   // Z0 <- y(tout)
            break;
        }
    }
    // every exit accounts the call's work into the caller's counters
    st.nst += local.nst; st.nfe += local.nfe; st.nje += local.nje; st.nsetups += local.nsetups;
    st.nni += local.nni; st.ncfn += local.ncfn; st.netf += local.netf;
    if (ret == 2) {
        t = tout;
    } else {
        t = s.tn;
        o.finish_at_tn();
    }
    return ret;
}

}  // namespace bdf
