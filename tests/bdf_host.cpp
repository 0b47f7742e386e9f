// Host build of the solver's BDF core (uclchem_amd/csrc/bdf_core.h, the header the GPU
// kernel compiles) driven by a plain dense Ops, for known-answer tests of the integrator
// on small stiff problems.  Test infrastructure: built by tests/test_bdf_core.py with g++.
//
// HostOps mirrors the GPU policy's vector semantics one for one (Nordsieck history z0..z5,
// acor, ewt, WRMS norms over all equations, residual gamma f - (rl1 z1 + acor), correction
// y = z0 + acor) with a dense Newton matrix I - gamma J factored by LU with partial pivoting.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../uclchem_amd/csrc/bdf_core.h"

namespace {

struct Problem {
    int n;
    void (*f)(const double *y, double *dy);
    void (*jac)(const double *y, double *J);   // J[i * n + j] = d f_i / d y_j
};

// y_i' = -k_i y_i
const double DECAY_K[3] = {1.0, 1.0e3, 1.0e6};
void decay_f(const double *y, double *dy)
{
    for (int i = 0; i < 3; i++) dy[i] = -DECAY_K[i] * y[i];
}
void decay_j(const double *, double *J)
{
    std::memset(J, 0, 9 * sizeof(double));
    for (int i = 0; i < 3; i++) J[i * 3 + i] = -DECAY_K[i];
}

// Robertson's chemical kinetics problem (Robertson 1966; Hairer & Wanner, Solving ODEs II)
void rober_f(const double *y, double *dy)
{
    dy[0] = -0.04 * y[0] + 1.0e4 * y[1] * y[2];
    dy[2] = 3.0e7 * y[1] * y[1];
    dy[1] = -dy[0] - dy[2];
}
void rober_j(const double *y, double *J)
{
    J[0] = -0.04; J[1] = 1.0e4 * y[2]; J[2] = 1.0e4 * y[1];
    J[6] = 0.0; J[7] = 6.0e7 * y[1]; J[8] = 0.0;
    J[3] = 0.04; J[4] = -1.0e4 * y[2] - 6.0e7 * y[1]; J[5] = -1.0e4 * y[1];
}

// HIRES, plant physiology (Schaefer 1975; Hairer & Wanner, Solving ODEs II)
void hires_f(const double *y, double *dy)
{
    dy[0] = -1.71 * y[0] + 0.43 * y[1] + 8.32 * y[2] + 0.0007;
    dy[1] = 1.71 * y[0] - 8.75 * y[1];
    dy[2] = -10.03 * y[2] + 0.43 * y[3] + 0.035 * y[4];
    dy[3] = 8.32 * y[1] + 1.71 * y[2] - 1.12 * y[3];
    dy[4] = -1.745 * y[4] + 0.43 * y[5] + 0.43 * y[6];
    dy[5] = -280.0 * y[5] * y[7] + 0.69 * y[3] + 1.71 * y[4] - 0.43 * y[5] + 0.69 * y[6];
    dy[6] = 280.0 * y[5] * y[7] - 1.81 * y[6];
    dy[7] = -dy[6];
}
void hires_j(const double *y, double *J)
{
    std::memset(J, 0, 64 * sizeof(double));
    J[0 * 8 + 0] = -1.71; J[0 * 8 + 1] = 0.43; J[0 * 8 + 2] = 8.32;
    J[1 * 8 + 0] = 1.71; J[1 * 8 + 1] = -8.75;
    J[2 * 8 + 2] = -10.03; J[2 * 8 + 3] = 0.43; J[2 * 8 + 4] = 0.035;
    J[3 * 8 + 1] = 8.32; J[3 * 8 + 2] = 1.71; J[3 * 8 + 3] = -1.12;
    J[4 * 8 + 4] = -1.745; J[4 * 8 + 5] = 0.43; J[4 * 8 + 6] = 0.43;
    J[5 * 8 + 3] = 0.69; J[5 * 8 + 4] = 1.71; J[5 * 8 + 5] = -0.43 - 280.0 * y[7]; J[5 * 8 + 6] = 0.69;
    J[5 * 8 + 7] = -280.0 * y[5];
    J[6 * 8 + 5] = 280.0 * y[7]; J[6 * 8 + 6] = -1.81; J[6 * 8 + 7] = 280.0 * y[5];
    for (int j = 0; j < 8; j++) J[7 * 8 + j] = -J[6 * 8 + j];
}

const Problem PROBLEMS[] = {{3, decay_f, decay_j}, {3, rober_f, rober_j}, {8, hires_f, hires_j}};

struct HostOps {
    int n;
    const Problem *P = nullptr;
    int (*cb)(double *y, double *ydot) = nullptr;   // external RHS; its Jacobian by differences
    std::vector<double> zs, acor, ewt, atol, f0, y, f, b, M, J, yp, fp;
    std::vector<int> piv;
    double *z[6];

    explicit HostOps(int n_) : n(n_), zs(6 * n_), acor(n_), ewt(n_), atol(n_), f0(n_), y(n_), f(n_), b(n_),
                               M((size_t)n_ * n_), J((size_t)n_ * n_), yp(n_), fp(n_), piv(n_)
    {
        for (int j = 0; j < 6; j++) z[j] = zs.data() + (size_t)j * n_;
    }
    void eval(const double *yy, double *ff)
    {
        if (cb) cb(const_cast<double *>(yy), ff);
        else P->f(yy, ff);
    }
    double zget(int v, int i) const { return v == bdf::ACOR ? acor[i] : z[v][i]; }
    double wrms(int v) const
    {
        double a = 0.0;
        for (int i = 0; i < n; i++) {
            const double t = zget(v, i) * ewt[i];
            a += t * t;
        }
        return std::sqrt(a / n);
    }
    double wrms_acor_minus(double c, int zj) const
    {
        double a = 0.0;
        for (int i = 0; i < n; i++) {
            const double t = (acor[i] - c * z[zj][i]) * ewt[i];
            a += t * t;
        }
        return std::sqrt(a / n);
    }
    bool set_ewt(double rtol)
    {
        bool ok = true;
        for (int i = 0; i < n; i++) {
            const double d = rtol * std::fabs(z[0][i]) + atol[i];
            if (!(d > 0.0)) ok = false;
            ewt[i] = 1.0 / d;
        }
        return ok;
    }
    void start_corrector()
    {
        for (int i = 0; i < n; i++) {
            acor[i] = 0.0;
            y[i] = z[0][i];
        }
    }
    void rhs() { eval(y.data(), f.data()); }
    void save_f0() { f0 = f; }
    void scale_into_z_from(int j, double a, int src)
    {
        for (int i = 0; i < n; i++) z[j][i] = a * (src == bdf::F0 ? f0[i] : f[i]);
    }
    void scale_into_z(int j, double a, int src)
    {
        for (int i = 0; i < n; i++) z[j][i] = a * z[src][i];
    }
    void axpy_zz(int j, double a, int src)
    {
        for (int i = 0; i < n; i++) z[j][i] += a * z[src][i];
    }
    void rescale_z(int q, double eta)
    {
        double s = eta;
        for (int j = 1; j <= q; j++) {
            for (int i = 0; i < n; i++) z[j][i] *= s;
            s *= eta;
        }
    }
    void predict(int q)
    {
        for (int k = 1; k <= q; k++)
            for (int j = q; j >= k; j--)
                for (int i = 0; i < n; i++) z[j - 1][i] += z[j][i];
    }
    void unpredict(int q)
    {
        for (int k = 1; k <= q; k++)
            for (int j = q; j >= k; j--)
                for (int i = 0; i < n; i++) z[j - 1][i] -= z[j][i];
    }
    void complete(int q, const double *l)
    {
        for (int j = 0; j <= q; j++)
            for (int i = 0; i < n; i++) z[j][i] += l[j] * acor[i];
    }
    void copy_acor_to_z(int j) { std::copy(acor.begin(), acor.end(), z[j]); }
    void form_residual(double gamma, double rl1)
    {
        for (int i = 0; i < n; i++) b[i] = gamma * f[i] - (rl1 * z[1][i] + acor[i]);
    }
    // Jacobian by forward differences with DVODE's increments (dvode.f90:8336-8356), the
    // gamma of the setup standing in for the step size in the r0 floor
    void fd_jacobian(double gamma)
    {
        const double srur = std::sqrt(bdf::UROUND);
        double fac = 0.0;
        for (int i = 0; i < n; i++) fac += (f[i] * ewt[i]) * (f[i] * ewt[i]);
        fac = std::sqrt(fac / n);
        double r0 = 1000.0 * std::fabs(gamma) * bdf::UROUND * n * fac;
        if (r0 == 0.0) r0 = 1.0;
        yp = y;
        for (int j = 0; j < n; j++) {
            const double r = std::fmax(srur * std::fabs(y[j]), r0 / ewt[j]);
            yp[j] = y[j] + r;
            eval(yp.data(), fp.data());
            for (int i = 0; i < n; i++) J[(size_t)i * n + j] = (fp[i] - f[i]) / r;
            yp[j] = y[j];
        }
    }
    int setup(double gamma, bool)
    {
        if (cb) fd_jacobian(gamma);
        else P->jac(y.data(), J.data());
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) M[i * n + j] = (i == j ? 1.0 : 0.0) - gamma * J[i * n + j];
        for (int k = 0; k < n; k++) {   // LU with partial pivoting, row-major
            int p = k;
            for (int i = k + 1; i < n; i++)
                if (std::fabs(M[i * n + k]) > std::fabs(M[p * n + k])) p = i;
            piv[k] = p;
            if (M[p * n + k] == 0.0 || !std::isfinite(M[p * n + k])) return 1;
            if (p != k)
                for (int j = 0; j < n; j++) std::swap(M[k * n + j], M[p * n + j]);
            for (int i = k + 1; i < n; i++) {
                const double m = M[i * n + k] / M[k * n + k];
                M[i * n + k] = m;
                for (int j = k + 1; j < n; j++) M[i * n + j] -= m * M[k * n + j];
            }
        }
        return 0;
    }
    bool can_reuse_jacobian() const { return false; }
    void solve()
    {
        // the factorisation swapped whole rows (multipliers included), so every interchange
        // applies to b before the forward substitution
        for (int k = 0; k < n; k++) std::swap(b[k], b[piv[k]]);
        for (int k = 0; k < n; k++)
            for (int i = k + 1; i < n; i++) b[i] -= M[i * n + k] * b[k];
        for (int k = n - 1; k >= 0; k--) {
            for (int j = k + 1; j < n; j++) b[k] -= M[k * n + j] * b[j];
            b[k] /= M[k * n + k];
        }
    }
    double apply_correction(double scale)
    {
        double a = 0.0;
        for (int i = 0; i < n; i++) {
            const double d = scale * b[i];
            const double dw = d * ewt[i];
            a += dw * dw;
            acor[i] += d;
            y[i] = z[0][i] + acor[i];
        }
        return std::sqrt(a / n);
    }
    double hub_inverse(double fac) const
    {
        double m = 0.0;
        for (int i = 0; i < n; i++) m = std::fmax(m, std::fabs(f0[i]) / (fac * std::fabs(z[0][i]) + atol[i]));
        return m;
    }
    double ydd_norm(double, double hg)
    {
        for (int i = 0; i < n; i++) y[i] = z[0][i] + hg * f0[i];
        eval(y.data(), f.data());
        double a = 0.0;
        for (int i = 0; i < n; i++) {
            const double t = (f[i] - f0[i]) / hg * ewt[i];
            a += t * t;
        }
        return std::sqrt(a / n);
    }
    void finish_at_tn() {}
    void interpolate(double s, int q)
    {
        for (int i = 0; i < n; i++) {
            double a = z[q][i];
            for (int j = q - 1; j >= 0; j--) a = a * s + z[j][i];
            z[0][i] = a;
        }
    }
    bool cancelled() const { return false; }
};

int run(HostOps &o, double t0, double tout, double rtol, const double *atol, double *y, int mxstep, int *stats)
{
    for (int i = 0; i < o.n; i++) {
        o.z[0][i] = y[i];
        o.atol[i] = atol[i];
    }
    bdf::Stats st;
    double t = t0;
    const int r = bdf::integrate(o, t, tout, rtol, mxstep, st);
    for (int i = 0; i < o.n; i++) y[i] = o.z[0][i];
    const int s[7] = {st.nst, st.nfe, st.nje, st.nsetups, st.nni, st.ncfn, st.netf};
    std::memcpy(stats, s, sizeof(s));
    return r;
}

}  // namespace

extern "C" {

// Integrates problem `problem` (0 decay, 1 Robertson, 2 HIRES) from t0 to tout with one
// cold-start integrate() call; y [n] in/out, atol [n]; stats [7] = nst nfe nje nsetups nni
// ncfn netf.  Returns integrate()'s DVODE-style state (2 = success).
int bdf_host_solve(int problem, double t0, double tout, double rtol, const double *atol, double *y, int mxstep,
                   int *stats)
{
    if (problem < 0 || problem > 2) return -100;
    HostOps o(PROBLEMS[problem].n);
    o.P = &PROBLEMS[problem];
    return run(o, t0, tout, rtol, atol, y, mxstep, stats);
}

// Same for an external RHS f(y, ydot) of n equations (autonomous), Jacobian by differences.
int bdf_host_solve_fn(int n, int (*rhs)(double *, double *), double t0, double tout, double rtol,
                      const double *atol, double *y, int mxstep, int *stats)
{
    if (n <= 0 || !rhs) return -100;
    HostOps o(n);
    o.cb = rhs;
    return run(o, t0, tout, rtol, atol, y, mxstep, stats);
}

int bdf_host_n(int problem) { return (problem >= 0 && problem <= 2) ? PROBLEMS[problem].n : -1; }
}
