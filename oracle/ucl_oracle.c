/* ORACLE / TEST INFRASTRUCTURE ONLY.  Never linked into the product library.
 *
 * Plain-C FP64 restatement of UCLCHEM's per-point chemistry hot path
 * (src/fortran_src/{chemistry,rates,surfacereactions,photoreactions,physics-core,
 * cloud,hotcore,constants,defaultparameters}.f90).  Default-kind (float32) literals of
 * the Fortran source are written here as float literals widened to double, so the
 * arithmetic matches what the reference compiler produces.
 *
 * Pinning (see tests/test_oracle_*.py): rates and RHS against the compiled reference
 * (oracle/_ref/libucref.so) on seeded states; RHS element/charge conservation; the
 * photo helpers against the reference procedures called directly; the schedule
 * against examples/example-output time columns.
 */
#include "ucl_oracle.h"
#include <math.h>
#include <string.h>

/* ---------------------------------------------------------------- constants.f90:1-19 */
#define K_BOLTZ 1.38065040e-16
#define AMU 1.66053892e-24
#define REDUCED_PLANCK 1.054571628e-27
#define PI ((double)3.141592654f)          /* constants.f90:9, default-real literal */
#define PC 3.086e18
#define SECONDS_PER_YEAR 3.16e7

/* ---------------------------------------------------------------- surfacereactions.f90:33-51 */
#define GAS_DUST_MASS_RATIO 100.0
#define GRAIN_RADIUS 1.0e-5
#define GRAIN_DENSITY 3.0
#define SURFACE_SITE_DENSITY 1.5e15
#define CHEMICAL_BARRIER_THICKNESS 1.40e-8
#define DIFFUSION_BIND_RATIO 0.5
#define MAX_GRAIN_TEMP 150.0
#define MIN_SURFACE_ABUND 1.0e-20
static double thermal_vel(void) { return sqrt(8.0 * K_BOLTZ / (PI * AMU)); }
static double gas_dust_density_ratio(void)
{
    return (4.0 * PI * (GRAIN_RADIUS * GRAIN_RADIUS * GRAIN_RADIUS) * GRAIN_DENSITY * GAS_DUST_MASS_RATIO) / (3.0 * AMU);
}
#define GRAIN_CROSSSECTION_PER_H (0.5 * (7.908e-22 + 8.473e-22))
#define GRAIN_SURFACEAREA_PER_H (4.0 * GRAIN_CROSSSECTION_PER_H)
static double vdiff_prefactor(void) { return 2.0 * K_BOLTZ * SURFACE_SITE_DENSITY / PI / PI / AMU; }
static double num_sites_per_grain(void) { return GRAIN_RADIUS * GRAIN_RADIUS * SURFACE_SITE_DENSITY * 4.0 * PI; }

/* defaultparameters.f90 "here be dragons" values; epsilon and uv_yield are default-real */
#define ICE_GAS_RATIO ((double)0.3f)      /* photoreactions.f90:41 */
#define TURB_VEL 1.0                      /* rates.f90:17 */

/* ---------------------------------------------------------------- small helpers */
static int ice_index(const ucl_net *n, int species)
{   /* position of a species in iceList; the reference loops keep the last match */
    int r = -1;
    for (int i = 0; i < n->nice; i++)
        if (n->ice_list[i] == species) r = i;
    return r;
}
static int in_list(const int32_t *l, int len, int v)
{
    for (int i = 0; i < len; i++)
        if (l[i] == v) return 1;
    return 0;
}
static double vdiff_of(const ucl_net *n, int icepos)
{   /* chemistry.f90:107-111 */
    double v = vdiff_prefactor() * n->binding_energy[icepos] / n->mass[n->ice_list[icepos]];
    return sqrt(v);
}
static void range_of(const ucl_net *n, int type, int *a, int *b)
{
    *a = n->type_ranges[2 * type];
    *b = n->type_ranges[2 * type + 1];
}

/* ---------------------------------------------------------------- photoreactions.f90 */
static const double LAMBDA_GRID[30] = {910.0, 950.0, 1000.0, 1050.0, 1110.0, 1180.0, 1250.0, 1390.0, 1490.0, 1600.0,
    1700.0, 1800.0, 1900.0, 2000.0, 2100.0, 2190.0, 2300.0, 2400.0, 2500.0, 2740.0, 3440.0, 4000.0, 4400.0,
    5500.0, 7000.0, 9000.0, 12500.0, 22000.0, 34000.0, 1.0e9};
static const double XLAMBDA_GRID[30] = {5.76, 5.18, 4.65, 4.16, 3.73, 3.40, 3.11, 2.74, 2.63, 2.62, 2.54, 2.50,
    2.58, 2.78, 3.01, 3.12, 2.86, 2.58, 2.35, 2.00, 1.58, 1.42, 1.32, 1.00, 0.75, 0.48, 0.28, 0.12, 0.05, 0.00};
static const double NCO_GRID[8] = {12.0, 13.0, 14.0, 15.0, 16.0, 17.0, 18.0, 19.0};
static const double NH2_GRID[6] = {18.0, 19.0, 20.0, 21.0, 22.0, 23.0};
/* SCO_GRID(8,6) in Fortran column-major storage order (photoreactions.f90:31-39) */
static const double SCO_FLAT[48] = {
    0.000e+00, -1.408e-02, -1.099e-01, -4.400e-01, -1.154e+00, -1.888e+00, -2.760e+00, -4.001e+00,
    -8.539e-02, -1.015e-01, -2.104e-01, -5.608e-01, -1.272e+00, -1.973e+00, -2.818e+00, -4.055e+00,
    -1.451e-01, -1.612e-01, -2.708e-01, -6.273e-01, -1.355e+00, -2.057e+00, -2.902e+00, -4.122e+00,
    -4.559e-01, -4.666e-01, -5.432e-01, -8.665e-01, -1.602e+00, -2.303e+00, -3.146e+00, -4.421e+00,
    -1.303e+00, -1.312e+00, -1.367e+00, -1.676e+00, -2.305e+00, -3.034e+00, -3.758e+00, -5.077e+00,
    -3.883e+00, -3.888e+00, -3.936e+00, -4.197e+00, -4.739e+00, -5.165e+00, -5.441e+00, -6.446e+00};

/* natural cubic spline second derivatives (photoreactions.f90:340-418, NR "spline") */
static void spline_natural(const double *x, const double *y, int n, double *y2)
{
    double u[100];
    y2[0] = 0.0;
    u[0] = 0.0;
    for (int i = 1; i < n - 1; i++) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    double qn = 0.0, un = 0.0;
    y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
    for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
}
/* cubic spline evaluation (photoreactions.f90:445-560; bracket search by bisection,
 * which gives the same bracket as the reference hunt for interior points) */
static double splint(const double *xa, const double *ya, const double *y2a, int n, double x)
{
    int jlo = 0, jhi = n + 1;   /* 1-based as in the reference */
    while (jhi - jlo != 1) {
        int jm = (jhi + jlo) / 2;
        if (x > xa[jm - 1]) jlo = jm; else jhi = jm;
    }
    if (jlo == 0) { jlo = 1; jhi = 2; }
    if (jlo == n) { jlo = n - 1; jhi = n; }
    double h = xa[jhi - 1] - xa[jlo - 1];
    double a = (xa[jhi - 1] - x) / h;
    double b = (x - xa[jlo - 1]) / h;
    return a * ya[jlo - 1] + b * ya[jhi - 1] + ((a * a * a - a) * y2a[jlo - 1] + (b * b * b - b) * y2a[jhi - 1]) * (h * h) / 6.0;
}

static double xlambda(double lambda)
{   /* photoreactions.f90:211-228 */
    double d2[30];
    spline_natural(LAMBDA_GRID, XLAMBDA_GRID, 30, d2);
    double v = lambda;
    if (lambda < LAMBDA_GRID[0]) v = LAMBDA_GRID[0];
    if (lambda > LAMBDA_GRID[29]) v = LAMBDA_GRID[29];
    double r = splint(LAMBDA_GRID, XLAMBDA_GRID, d2, 30, v);
    return r < 0.0 ? 0.0 : r;
}

double ucl_scatter(double x1, double av)
{   /* photoreactions.f90:138-196 */
    static const double c[6] = {1.0, 2.006, -1.438, 7.364e-01, -5.076e-01, -5.920e-02};
    static const double k1[6] = {7.514e-01, 8.490e-01, 1.013, 1.282, 2.005, 5.832};
    double tv = av / 1.086;
    double tl = tv * xlambda(x1);
    double s = 0.0;
    if (tl < 1.0) {
        double expo = k1[0] * tl;
        if (expo < 100.0) s = c[0] * exp(-expo);
    } else {
        for (int i = 1; i < 6; i++) {
            double expo = k1[i] * tl;
            if (expo < 100.0) s = s + c[i] * exp(-expo);
        }
    }
    return s;
}

static double h2_self_shielding(double nh2, double doppler, double radwidth)
{   /* photoreactions.f90:95-133 */
    const double FPARA = 0.5, FOSC = 1.0e-2;
    double taud = FPARA * nh2 * (double)1.5e-2f * FOSC / doppler;
    double sj, sr;
    if (taud == 0.0) sj = 1.0;
    else if (taud < 2.0) sj = exp(-0.6666667 * taud);
    else if (taud < 10.0) sj = 0.638 * pow(taud, -1.25);
    else if (taud < 100.0) sj = 0.505 * pow(taud, -1.15);
    else sj = 0.344 * pow(taud, -1.0667);
    if (radwidth == 0.0) sr = 0.0;
    else {
        double r = radwidth / (1.7724539 * doppler);
        double t = 3.02 * pow(r * 1.0e3, -0.064);
        double u = sqrt(taud * r) / t;
        sr = r / (t * sqrt(0.78539816 + u * u));
    }
    return sj + sr;
}

double ucl_h2_photodiss(double nh2, double radfield, double av, double turbvel)
{   /* photoreactions.f90:46-55 */
    const double base = 5.18e-11, xl = 1000.0, radwidth = 8.0e7;
    double doppler = turbvel / (xl * 1.0e-8);
    return base * (radfield / (double)1.7f) * ucl_scatter(xl, av) * h2_self_shielding(nh2, doppler, radwidth);
}

double ucl_co_self_shielding(double nh2, double nco)
{   /* photoreactions.f90:238-257 with splie2/splin2 (:289-338).  The reference passes the
     * 8x6 table with DIMCO=7 rows, so the routines see the first 42 stored values as a
     * 7x6 column-major array; that view is reproduced here. */
    const int m = 7, n = 6;
    double ya[7][6], y2a[7][6];
    for (int j = 0; j < m; j++)
        for (int k = 0; k < n; k++) ya[j][k] = SCO_FLAT[j + m * k];
    for (int j = 0; j < m; j++) spline_natural(NH2_GRID, ya[j], n, y2a[j]);
    double lognco = log10(nco + 1.0), lognh2 = log10(nh2 + 1.0);
    if (lognco < NCO_GRID[0]) lognco = NCO_GRID[0];
    if (lognh2 < NH2_GRID[0]) lognh2 = NH2_GRID[0];
    if (lognco > NCO_GRID[m - 1]) lognco = NCO_GRID[m - 1];
    if (lognh2 > NH2_GRID[n - 1]) lognh2 = NH2_GRID[n - 1];
    double yy[7], y2[7];
    for (int j = 0; j < m; j++) yy[j] = splint(NH2_GRID, ya[j], y2a[j], n, lognh2);
    spline_natural(NCO_GRID, yy, m, y2);
    double v = splint(NCO_GRID, yy, y2, m, lognco);
    return pow(10.0, v);
}

static double lbar(double u, double w)
{   /* photoreactions.f90:262-287 */
    double lu = log10(fabs(u) + 1.0), lw = log10(fabs(w) + 1.0);
    double l = (5675.0 - 200.6 * lw) - (571.6 - 24.09 * lw) * lu + (18.22 - 0.7664 * lw) * (lu * lu);
    if (l > 1076.1) l = 1076.1;
    if (l < 913.6) l = 913.6;
    return l;
}

double ucl_co_photodiss(double nh2, double nco, double radfield, double av)
{   /* photoreactions.f90:57-72 */
    double ssf = ucl_co_self_shielding(nh2, nco);
    double lba = lbar(nco, nh2);
    double sca = ucl_scatter(lba, av);
    return 2.0e-10 * (radfield / (double)1.7f) * ssf * sca;
}

double ucl_c_ionization(double alpha, double gamma, double t, double nc, double nh2, double av, double radfield)
{   /* photoreactions.f90:74-87 */
    double tauc = gamma * av + 1.1e-17 * nc + (0.9 * pow(t, 0.27) * pow(nh2 / 1.59e21, 0.45));
    return alpha * (radfield / (double)1.7f) * exp(-tauc);
}

/* ---------------------------------------------------------------- surfacereactions.f90 */
double ucl_h2_form_efficiency(double gas_temp, double dust_temp)
{   /* surfacereactions.f90:63-107 (Cazaux & Tielens 2002/2004) */
    const double S_MU = 0.005, S_ES = 110.0, S_EH2 = 320.0, S_EHP = 450.0, S_EHC = 3.0e4, S_NUH2 = 3.0e12,
                 S_NUHC = 1.3e13, S_CS = 8.473e-22;
    const double G_MU = 0.005, G_ES = 260.0, G_EH2 = 520.0, G_EHP = 800.0, G_EHC = 3.0e4, G_NUH2 = 3.0e12,
                 G_NUHC = 1.3e13, G_CS = 7.908e-22;
    double tv = 1.45e5 * sqrt(gas_temp / 1.0e2);
    double stick = 1.0 / (1.0 + 0.04 * sqrt(gas_temp + dust_temp) + 0.2 * (gas_temp / 1.0e2) + 0.08 * pow(gas_temp / 1.0e2, 2));
    double hflux = 1.0e-10;
    double f1 = S_MU * hflux / (2 * S_NUH2 * exp(-S_EH2 / dust_temp));
    double q = 1.0 + sqrt((S_EHC - S_ES) / (S_EHP - S_ES));
    double f2 = 1.0 * (q * q) / 4.0 * exp(-S_ES / dust_temp);
    double eps = 1.0 / (1.0 + S_NUHC / (2 * hflux) * exp(-1.5 * S_EHC / dust_temp) * (q * q));
    double sil = 1.0 / (1.0 + f1 + f2) * eps;
    f1 = G_MU * hflux / (2 * G_NUH2 * exp(-G_EH2 / dust_temp));
    q = 1.0 + sqrt((G_EHC - G_ES) / (G_EHP - G_ES));
    f2 = 1.0 * (q * q) / 4.0 * exp(-G_ES / dust_temp);
    eps = 1.0 / (1.0 + G_NUHC / (2 * hflux) * exp(-1.5 * G_EHC / dust_temp) * (q * q));
    double gra = 1.0 / (1.0 + f1 + f2) * eps;
    return 0.5 * tv * (S_CS * sil + G_CS * gra) * stick;
}

static double diffusion_reaction_rate(const ucl_net *n, int j, double dust_temp)
{   /* surfacereactions.f90:158-211 */
    int i1 = n->re[3 * j], i2 = n->re[3 * j + 1];
    int p1 = ice_index(n, i1), p2 = ice_index(n, i2);
    if (p1 >= 0) i1 = p1;
    if (p2 >= 0) i2 = p2;
    double v1 = vdiff_of(n, i1), v2 = vdiff_of(n, i2);
    double be1 = n->binding_energy[i1], be2 = n->binding_energy[i2];
    double diffuse = v1 * exp(-DIFFUSION_BIND_RATIO * be1 / dust_temp);
    diffuse = diffuse + v2 * exp(-DIFFUSION_BIND_RATIO * be2 / dust_temp);
    double desorb = v1 * exp(-be1 / dust_temp);
    desorb = desorb + v2 * exp(-be2 / dust_temp);
    double reac = n->gama[j] / dust_temp;
    double rm = n->redmass[j];
    if (rm == 0.0) {
        double m1 = n->mass[n->ice_list[i1]], m2 = n->mass[n->ice_list[i2]];
        rm = m1 * m2 / (m1 + m2);
    }
    double tunnel = 2.0 * CHEMICAL_BARRIER_THICKNESS / REDUCED_PLANCK * sqrt(2.0 * AMU * rm * K_BOLTZ * n->gama[j]);
    if (reac > tunnel) reac = tunnel;
    reac = fmax(v1, v2) * exp(-reac);
    reac = reac / (reac + desorb + diffuse);
    return n->alpha[j] * reac * diffuse * gas_dust_density_ratio() / num_sites_per_grain();
}

static double desorption_fraction(const ucl_net *n, int j)
{   /* surfacereactions.f90:218-297, including its indexing of mass() and atomCounts()
     * by iceList position and "last match wins" searches */
    int r1 = n->re[3 * j], r2 = n->re[3 * j + 1];
    int ri1 = 0, ri2 = 0, pidx[4] = {-1, -1, -1, -1};
    for (int i = 0; i < n->nice; i++) {
        if (n->ice_list[i] == r1) ri1 = i;
        if (n->gas_ice_list[i] == r1) ri1 = i;
        if (n->ice_list[i] == r2) ri2 = i;
        if (n->gas_ice_list[i] == r2) ri2 = i;
        for (int k = 0; k < 4; k++) {
            int pk = n->pr[4 * j + k];
            if (pk < 0) continue;
            if (n->ice_list[i] == pk) pidx[k] = i;
            if (n->gas_ice_list[i] == pk) pidx[k] = i;
        }
    }
    double maxbe = 0.0, prodh = 0.0, eps = 0.0;
    for (int k = 0; k < 4; k++) {
        if (pidx[k] >= 0) {
            maxbe = fmax(maxbe, n->binding_energy[pidx[k]]);
            prodh = prodh + n->formation_enthalpy[pidx[k]];
            eps = eps + n->mass[pidx[k]];          /* mass() indexed by ice position */
        }
    }
    const double EFF = 120.0;
    eps = pow((eps - EFF) / (eps + EFF), 2);
    double dh = n->formation_enthalpy[ri1] + n->formation_enthalpy[ri2] - prodh;
    dh = dh * 4.184e03 / (1.38054e-23 * 6.02214129e23);
    dh = dh + n->gama[j];
    if (dh == 0.0) dh = (double)1e-30f;
    /* atomCounts() indexed by ice position; pidx[0] is never absent for these types */
    long dof = pidx[0] >= 0 ? (long)n->n_atoms[pidx[0]] : 0;
    for (int k = 1; k < 4; k++)
        if (pidx[k] >= 0 && n->n_atoms[pidx[k]] > dof) dof = (long)n->n_atoms[pidx[k]];
    dof = 3 * dof;
    double f = exp((-maxbe * (double)(float)dof) / (eps * dh));
    if (dh < 0.0) f = 0.0;
    f = f / 10;
    if (r1 == n->ngn && r2 == n->ngn) f = 0.5;
    if ((r1 == n->ngo && r2 == n->nh) || (r1 == n->nh && r2 == n->ngo)) f = (double)0.3f;
    if ((r1 == n->ngoh && r2 == n->nh) || (r1 == n->nh && r2 == n->ngoh)) f = 0.25;
    return f;
}

static double freeze_out_rate(const ucl_net *n, const ucl_point *pt, int j)
{   /* rates.f90:354-367 */
    double f = 1.0 + n->beta[j] * 16.71e-4 / (GRAIN_RADIUS * pt->gas_temp);
    if (pt->p[P_FREEZEFAC] == 0.0 || pt->dust_temp > MAX_GRAIN_TEMP) return 0.0;
    return f * pt->p[P_FREEZEFAC] * n->alpha[j] * thermal_vel() * sqrt(pt->gas_temp / n->mass[n->re[3 * j]]) *
           GRAIN_CROSSSECTION_PER_H;
}

static double sticking(double s0, double tcrit, double t)
{   /* rates.f90:370-378 */
    double r = t / tcrit;
    return s0 * (1.0 + 2.5 * r) / pow(1.0 + r, 2.5);
}

/* ---------------------------------------------------------------- chemistry.f90:46-141 */
int ucl_point_init(const ucl_net *n, const double *params, ucl_point *pt, double *y)
{
    memset(pt, 0, sizeof(*pt));
    memcpy(pt->p, params, sizeof(double) * P_NPARAM);
    /* physics-core.f90:42-73 + cloud.f90:20-31 */
    pt->cloud_size = (pt->p[P_ROUT] - pt->p[P_RIN]) * PC;
    pt->gas_temp = pt->p[P_TEMP];
    pt->dust_temp = pt->gas_temp;
    pt->density = pt->p[P_DENS];
    pt->coldens = 1.0 * pt->cloud_size / 1.0 * pt->p[P_DENS];   /* real(points-dstep+1)*cloudSize/real(points)*initialDens */
    pt->av = pt->p[P_BASEAV] + pt->coldens / 1.6e21;
    if (pt->p[P_FREEFALL] != 0.0) pt->density = (double)1.001f * pt->p[P_DENS];
    pt->phi = pt->p[P_PHI];
    pt->h2crp = 0.0;
    /* elemental abundances (chemistry.f90:58-105) from the parameter vector */
    const double *q = pt->p;
    int neq = n->neq;
    for (int i = 0; i < neq; i++) y[i] = 1.0e-30;
    y[n->no] = q[P_FO];
    y[n->nn] = q[P_FN];
    y[n->nmg] = q[P_FMG];
    y[n->np] = q[P_FP];
    y[n->nf] = q[P_FF];
    y[n->nna] = q[P_FNA];
    y[n->nli] = q[P_FLI];
    y[n->npah] = q[P_FPAH];
    y[n->nsx] = q[P_FS];
    y[n->nsix] = q[P_FSI];
    y[n->nclx] = q[P_FCL];
    switch ((int)q[P_ION]) {   /* how much carbon starts ionised */
    case 0: y[n->nc] = q[P_FC]; y[n->ncx] = 1.0e-10; break;
    case 1: y[n->nc] = q[P_FC] * 0.5; y[n->ncx] = q[P_FC] * 0.5; break;
    case 2: y[n->nc] = 1.0e-10; y[n->ncx] = q[P_FC]; break;
    default: break;
    }
    y[n->n18o] = q[P_F18O];
    y[n->n15n] = q[P_F15N];
    y[n->n13c] = q[P_F13C];
    y[n->nelec] = y[n->ncx] + y[n->nsix] + y[n->nsx] + y[n->nclx] + y[n->nmgx];
    for (int i = 0; i < neq; i++) y[i] = y[i] * q[P_METALLICITY];
    y[n->nh] = q[P_FH];
    y[n->nh2] = 0.5 * (1.0 - q[P_FH]);
    y[n->nd] = q[P_FD];
    y[n->nhe] = q[P_FHE];
    y[neq - 1] = pt->density;
    return 0;
}

/* ---------------------------------------------------------------- rates.f90:21-343 */
int ucl_rates(const ucl_net *n, ucl_point *pt, double *y, double *rate, double *scal)
{
    const int neq = n->neq, ns = n->nspec;
    const double T = pt->gas_temp, Td = pt->dust_temp, av = pt->av;
    const double zeta = pt->p[P_ZETA], radfield = pt->p[P_RADFIELD];
    int a, b;
    /* chemistry.f90:167-200: density, columns, ice totals, ice scalars */
    if (pt->p[P_FREEFALL] == 0.0) y[neq - 1] = pt->density;
    double half = pt->cloud_size / 1.0;
    pt->h2col = 0.5 * y[n->nh2] * pt->density * half;
    pt->cocol = 0.5 * y[n->nco] * pt->density * half;
    pt->ccol = 0.5 * y[n->nc] * pt->density * half;
    double sb = 0.0, ss = 0.0;
    for (int i = 0; i < n->nbulkl; i++) sb += y[n->bulk_list[i]];
    for (int i = 0; i < n->nsurfl; i++) ss += y[n->surface_list[i]];
    y[n->nbulk] = sb;
    y[n->nsurface] = ss;
    double sm = fmax(1e-30, y[n->nsurface]);
    double sbk = fmax(1e-30, y[n->nbulk]);
    double blr = fmin(1.0, num_sites_per_grain() / (gas_dust_density_ratio() * sbk));
    pt->safe_mantle = sm;
    pt->safe_bulk = sbk;
    pt->blr = blr;
    const double dens = pt->density;

    for (int k = 0; k < n->nreac; k++) rate[k] = 0.0;

    const double *q = pt->p;
    const int desorb = q[P_DESORB] != 0.0;
    range_of(n, T_CRP, &a, &b);
    if (a != b)
        for (int j = a; j <= b; j++) rate[j] = n->alpha[j] * zeta;
    if (q[P_IMPROVEDH2CRP] != 0.0 && n->r_h2_crp >= 0) rate[n->r_h2_crp] = pt->h2crp;

    const double photo_bulk_pow = pow((double)(1.0f - 0.007f), 1.0 + 0.5 / blr);
    range_of(n, T_PHOTON, &a, &b);
    if (a != b) {
        for (int j = a; j <= b; j++) {
            rate[j] = n->alpha[j] * exp(-n->gama[j] * av) * radfield / (double)1.7f;
            int r1 = n->re[3 * j];
            if (in_list(n->bulk_list, n->nbulkl, r1)) rate[j] = rate[j] * ICE_GAS_RATIO * photo_bulk_pow;
            else if (in_list(n->surface_list, n->nsurfl, r1)) rate[j] = rate[j] * ICE_GAS_RATIO;
        }
    }
    range_of(n, T_CRPHOT, &a, &b);
    if (a != b) {
        for (int j = a; j <= b; j++) {
            rate[j] = n->alpha[j] * n->gama[j] * 1.0 / (1.0 - q[P_OMEGA]) * zeta * pow(T / 300, n->beta[j]);
            int r1 = n->re[3 * j];
            if (in_list(n->bulk_list, n->nbulkl, r1)) rate[j] = rate[j] * ICE_GAS_RATIO * photo_bulk_pow;
            else if (in_list(n->surface_list, n->nsurfl, r1)) rate[j] = rate[j] * ICE_GAS_RATIO;
        }
    }
    range_of(n, T_FREEZE, &a, &b);
    if (a != b) {
        for (int j = a; j <= b; j++) rate[j] = freeze_out_rate(n, pt, j);
        rate[n->r_efreeze] = rate[n->r_hfreeze];
        rate[n->r_h2freeze] = sticking(0.87, 87.0, T) * rate[n->r_h2freeze];
        rate[n->r_hfreeze] = sticking(1.0, 52.0, T) * rate[n->r_hfreeze];
    }
    /* non-thermal desorption (rates.f90:91-155) with freeze-out switch-off */
    range_of(n, T_DESOH2, &a, &b);
    if (a != b) {
        if (desorb && q[P_H2DESORB] != 0.0 && sm > MIN_SURFACE_ABUND) {
            double v = q[P_EPSILON] * ucl_h2_form_efficiency(T, Td);
            for (int j = a; j <= b; j++) rate[j] = n->gama[j] > q[P_EBMAXH2] ? 0.0 : v;
        }
        for (int q = 0; q < n->nfp; q++) {
            int f = n->freeze_partners[q];
            if (rate[f] * y[n->re[3 * f]] < MIN_SURFACE_ABUND * rate[a + q]) rate[f] = 0.0;
        }
    }
    range_of(n, T_DESCR, &a, &b);
    if (a != b) {
        if (desorb && q[P_CRDESORB] != 0.0 && sm > MIN_SURFACE_ABUND) {
            double v = 4.0 * PI * zeta * 1.64e-4 * (GRAIN_SURFACEAREA_PER_H) * pt->phi;
            for (int j = a; j <= b; j++) rate[j] = n->gama[j] > q[P_EBMAXCR] ? 0.0 : v;
        }
        for (int q = 0; q < n->nfp; q++) {
            int f = n->freeze_partners[q];
            if (rate[f] * y[n->re[3 * f]] * dens < MIN_SURFACE_ABUND * rate[a + q]) rate[f] = 0.0;
        }
    }
    range_of(n, T_DEUVCR, &a, &b);
    if (a != b) {
        if (desorb && q[P_UVDESORB] != 0.0 && sm > MIN_SURFACE_ABUND && zeta > 0) {
            double v = GRAIN_CROSSSECTION_PER_H * q[P_UV_YIELD] * 4.875e3 * zeta;
            v = v * (1 + (radfield / q[P_UVCREFF]) * (1.0 / zeta) * exp((double)-1.8f * av));
            for (int j = a; j <= b; j++) rate[j] = n->gama[j] > q[P_EBMAXUVCR] ? 0.0 : v;
        }
        for (int q = 0; q < n->nfp; q++) {
            int f = n->freeze_partners[q];
            if (rate[f] * y[n->re[3 * f]] * dens < MIN_SURFACE_ABUND * rate[a + q]) rate[f] = 0.0;
        }
    }
    /* excited-species reactions (rates.f90:158-197): CR excitation, relaxation at the
     * reactant's diffusion frequency, and solid-phase reactions of excited species */
    range_of(n, T_CRS, &a, &b);
    if (a != b)
        for (int j = a; j <= b; j++)
            rate[j] = n->alpha[j] * (n->beta[j] * (n->gama[j] / 100) * ((double)8.6f * zeta * (double)1.3f));
    range_of(n, T_EXRELAX, &a, &b);
    if (a != b)
        for (int j = a; j <= b; j++) rate[j] = vdiff_of(n, ice_index(n, n->re[3 * j]));
    range_of(n, T_EXSOLID, &a, &b);
    if (a != b)
        for (int j = a; j <= b; j++) {
            const double va = vdiff_of(n, ice_index(n, n->re[3 * j])), vb = vdiff_of(n, ice_index(n, n->re[3 * j + 1]));
            rate[j] = (vb + va) / (SURFACE_SITE_DENSITY * 1.8e-8);
            rate[j] = n->alpha[j] * rate[j];
        }
    /* continuous thermal desorption (rates.f90:200-229) */
    range_of(n, T_THERM, &a, &b);
    if (a != b && q[P_THERMDESORB] != 0.0) {   /* off: the range stays zero, no switch-off */
        for (int j = a; j <= b; j++) {
            int p = ice_index(n, n->re[3 * j]);
            if (p >= 0) rate[j] = vdiff_of(n, p) * exp(-n->gama[j] / Td);
        }
        /* non-conforming WHERE: the freeze-partner list is compared element-wise with the
         * first entries of the thermal range (surface species come first) */
        for (int q = 0; q < n->nfp; q++) {
            int f = n->freeze_partners[q];
            if (rate[f] * y[n->re[3 * f]] * dens < MIN_SURFACE_ABUND * rate[a + q]) rate[f] = 0.0;
        }
        if (sm < MIN_SURFACE_ABUND)
            for (int j = a; j <= b; j++) rate[j] = 0.0;
    }
    /* Langmuir-Hinshelwood with chemical desorption (rates.f90:236-263) */
    range_of(n, T_LH, &a, &b);
    if (a != b) {
        int c, d;
        range_of(n, T_LHDES, &c, &d);
        if (Td < MAX_GRAIN_TEMP && sm > MIN_SURFACE_ABUND) {
            for (int j = a; j <= b; j++) rate[j] = diffusion_reaction_rate(n, j, Td);
            for (int j = 0; j <= b - a; j++) rate[c + j] = rate[a + j];
            for (int j = c; j <= d; j++) {
                rate[j] = desorption_fraction(n, j) * rate[j];
                if (in_list(n->bulk_list, n->nbulkl, n->re[3 * j])) rate[j] = 0.0;
            }
            for (int j = 0; j <= b - a; j++) rate[a + j] = rate[a + j] - rate[c + j];
        } else {
            for (int j = a; j <= b; j++) rate[j] = 0.0;
            for (int j = c; j <= d; j++) rate[j] = 0.0;
        }
    }
    /* Eley-Rideal (rates.f90:267-279) */
    range_of(n, T_ER, &a, &b);
    if (a != b) {
        int c, d;
        range_of(n, T_ERDES, &c, &d);
        for (int j = a; j <= b; j++) rate[j] = freeze_out_rate(n, pt, j) * exp(-n->gama[j] / Td);
        for (int j = 0; j <= b - a; j++) rate[c + j] = rate[a + j];
        for (int j = c; j <= d; j++) {
            rate[j] = desorption_fraction(n, j) * rate[j];
            if (in_list(n->bulk_list, n->nbulkl, n->re[3 * j])) rate[j] = 0.0;
        }
        for (int j = 0; j <= b - a; j++) rate[a + j] = rate[a + j] - rate[c + j];
    }
    /* parameterised H2 formation (rates.f90:281-289) */
    rate[n->r_h2form_ct] = ucl_h2_form_efficiency(Td, Td);
    rate[n->r_h2form_er] = 0.0;
    rate[n->r_h2form_erdes] = 0.0;
    /* bulk <-> surface exchange (surfacereactions.f90:109-150) */
    range_of(n, T_BULKSWAP, &a, &b);
    if (a >= 0) {
        if (Td > MAX_GRAIN_TEMP || sm < MIN_SURFACE_ABUND) {
            for (int j = a; j <= b; j++) rate[j] = 0.0;
        } else {
            for (int j = a; j <= b; j++) {
                int p = ice_index(n, n->re[3 * j]);
                if (p >= 0) rate[j] = vdiff_of(n, p) * exp(-n->binding_energy[p] / Td);
            }
        }
    }
    range_of(n, T_SURFSWAP, &a, &b);
    if (a >= 0) {
        double v = (Td > MAX_GRAIN_TEMP || sm < MIN_SURFACE_ABUND) ? 0.0 : 1.0;
        for (int j = a; j <= b; j++) rate[j] = v;
    }
    /* gas-phase two-body (rates.f90:293-299) */
    range_of(n, T_TWOBODY, &a, &b);
    if (a >= 0)
        for (int j = a; j <= b; j++)
            rate[j] = n->alpha[j] * pow(T / 300.0, n->beta[j]) * exp(-n->gama[j] / T);
    /* ion-polar reactions (rates.f90:301-313, KIDA formulas) */
    range_of(n, T_IONOPOL1, &a, &b);
    if (a != b)
        for (int j = a; j <= b; j++)
            rate[j] = n->alpha[j] * n->beta[j] * (0.62 + 0.4767 * n->gama[j] * sqrt(300.0 / T));
    range_of(n, T_IONOPOL2, &a, &b);
    if (a != b)
        for (int j = a; j <= b; j++)
            rate[j] = n->alpha[j] * n->beta[j] *
                      (1.0 + 0.0967 * n->gama[j] * sqrt(300.0 / T) + n->gama[j] * n->gama[j] * 300.0 / ((double)10.526f * T));
    /* phi: used by GAR now and by the next call's CR desorption (rates.f90:320-326) */
    double phi = radfield * exp(-2.5 * av) * sqrt(T) / (y[neq - 1] * y[n->nelec]);
    pt->phi = fmin(fmax(phi, 1e2), 1e6);
    range_of(n, T_GAR, &a, &b);
    if (a != b)
        for (int j = a; j <= b; j++) {
            const double *g = n->gar_params + 7 * (j - a), ph = pt->phi;
            rate[j] = (double)0.6f * n->alpha[j] * g[0] /
                      (1. + g[1] * pow(ph, g[2]) * (1. + g[3] * pow(T, g[4]) * pow(ph, -g[5] - g[6] * log(T))));
        }
    /* temperature-range masks (rates.f90:335-337) */
    for (int k = 0; k < n->nreac; k++) {
        if (!n->extrapolate[k] && T < n->tmin[k]) rate[k] = 0.0;
        if (!n->extrapolate[k] && T > n->tmax[k]) rate[k] = 0.0;
    }
    /* photo overrides (rates.f90:340-342) */
    rate[n->r_h2_hv] = ucl_h2_photodiss(pt->h2col, radfield, av, TURB_VEL);
    rate[n->r_co_hv] = ucl_co_photodiss(pt->h2col, pt->cocol, radfield, av);
    rate[n->r_c_hv] = ucl_c_ionization(n->alpha[n->r_c_hv], n->gama[n->r_c_hv], T, pt->ccol, pt->h2col, av, radfield);
    if (scal) {
        scal[0] = pt->h2col; scal[1] = pt->cocol; scal[2] = pt->ccol; scal[3] = sm; scal[4] = sbk;
        scal[5] = blr; scal[6] = pt->phi; scal[7] = av;
    }
    (void)ns;
    return 0;
}

/* ---------------------------------------------------------------- chemistry.f90:294-352 + odes.f90 */
static double term_value(const ucl_net *n, int k, const double *rate, const double *y, double D, double blr,
                         double total_swap, double sm)
{   /* one reaction's ODE term, tokens in the order reaction.py:779-819 emits them */
    double v = rate[k];
    for (int q = 0; q < n->n_d[k]; q++) v = v * D;
    for (int q = 0; q < UCL_MAXOPS; q++) {
        int op = n->ops[UCL_MAXOPS * k + q];
        if (op == -1) continue;
        if (op >= 0) v = v * y[op];
        else if (op == -2) v = v * blr;
        else if (op == -3) v = v * total_swap / sm;
        else if (op == -4) v = v / sm;
    }
    return v;
}

/* The per-reaction terms of GETYDOT (odes.f90) at y: flux[k] = RATE(k) times its density and
 * abundance factors (and the bulk-layer, totalSwap and 1/SURFACE factors of its type), the
 * quantity the reference's analysis.rates_to_dy_and_flux (analysis.py:737-800) rebuilds from
 * output rows; H2 and CO photodissociation recomputed at y as in F (chemistry.f90:311-319). */
int ucl_fluxes(const ucl_net *n, const ucl_point *pt, double *rate, const double *y, double *flux)
{
    const int neq = n->neq;
    const double D = y[neq - 1];
    const double radfield = pt->p[P_RADFIELD];
    double half = pt->cloud_size / 1.0;
    double cocol = 0.5 * y[n->nco] * D * half;
    double h2col = 0.5 * y[n->nh2] * D * half;
    rate[n->r_h2_hv] = ucl_h2_photodiss(h2col, radfield, pt->av, TURB_VEL);
    rate[n->r_co_hv] = ucl_co_photodiss(h2col, cocol, radfield, pt->av);
    double sm = fmax(1e-30, y[n->nsurface]);
    double sbk = fmax(1e-30, y[n->nbulk]);
    double blr = fmin(1.0, num_sites_per_grain() / (gas_dust_density_ratio() * sbk));
    double total_swap = 0.0;
    for (int q = 0; q < n->nswap; q++) total_swap += term_value(n, n->swap_list[q], rate, y, D, blr, 0.0, sm);
    for (int k = 0; k < n->nreac; k++) flux[k] = term_value(n, k, rate, y, D, blr, total_swap, sm);
    return 0;
}

int ucl_rhs(const ucl_net *n, const ucl_point *pt, double *rate, const double *y, double *ydot)
{
    const int neq = n->neq, ns = n->nspec;
    const double D = y[neq - 1];
    const double radfield = pt->p[P_RADFIELD];
    /* H2 and CO photodissociation follow the current abundances (chemistry.f90:311-319) */
    double half = pt->cloud_size / 1.0;
    double cocol = 0.5 * y[n->nco] * D * half;
    double h2col = 0.5 * y[n->nh2] * D * half;
    rate[n->r_h2_hv] = ucl_h2_photodiss(h2col, radfield, pt->av, TURB_VEL);
    rate[n->r_co_hv] = ucl_co_photodiss(h2col, cocol, radfield, pt->av);
    double sm = fmax(1e-30, y[n->nsurface]);
    double sbk = fmax(1e-30, y[n->nbulk]);
    double blr = fmin(1.0, num_sites_per_grain() / (gas_dust_density_ratio() * sbk));
    double sc = 0.5 * gas_dust_density_ratio() / num_sites_per_grain();   /* bulkGainFromMantleBuildUp */
    double total_swap = 0.0;
    for (int q = 0; q < n->nswap; q++) total_swap += term_value(n, n->swap_list[q], rate, y, D, blr, 0.0, sm);
    for (int i = 0; i < ns; i++) {
        double loss = 0.0, prod = 0.0;
        int nl = n->loss_ptr[i + 1] - n->loss_ptr[i], np = n->prod_ptr[i + 1] - n->prod_ptr[i];
        for (int q = n->loss_ptr[i]; q < n->loss_ptr[i + 1]; q++)
            loss += term_value(n, n->loss_idx[q], rate, y, D, blr, total_swap, sm);
        for (int q = n->prod_ptr[i]; q < n->prod_ptr[i + 1]; q++)
            prod += term_value(n, n->prod_idx[q], rate, y, D, blr, total_swap, sm);
        if (nl && np) ydot[i] = prod - loss;
        else if (np) ydot[i] = prod;
        else if (nl) ydot[i] = -loss;
        else ydot[i] = 0.0;
    }
    /* ice totals, then the bulk<->surface transfer branch (odes.f90:4789-5153) */
    double sum_b = 0.0, sum_s = 0.0;
    for (int i = 0; i < n->nbulkl; i++) sum_b += ydot[n->bulk_list[i]];
    for (int i = 0; i < n->nsurfl; i++) sum_s += ydot[n->surface_list[i]];
    ydot[n->nbulk] = sum_b;
    ydot[n->nsurface] = sum_s;
    double ds = ydot[n->nsurface];
    if (ds < 0) {
        sc = fmin(1.0, sbk / sm);
        for (int q = 0; q < n->nsurfl; q++) {
            int s = n->surface_list[q], bp = n->bulk_partner_of_surface[q];
            if (bp < 0) continue;
            double t = ds * sc * y[bp] / sbk;
            ydot[s] = ydot[s] - t;
            ydot[bp] = ydot[bp] + t;
        }
    } else {
        for (int q = 0; q < n->nsurfl; q++) {
            int s = n->surface_list[q], bp = n->bulk_partner_of_surface[q];
            if (bp < 0) continue;
            double t = ds * sc * y[s];
            ydot[bp] = ydot[bp] + t;
            ydot[s] = ydot[s] - t;
        }
    }
    sum_b = 0.0;
    sum_s = 0.0;
    for (int i = 0; i < n->nbulkl; i++) sum_b += ydot[n->bulk_list[i]];
    for (int i = 0; i < n->nsurfl; i++) sum_s += ydot[n->surface_list[i]];
    ydot[n->nbulk] = sum_b;
    ydot[n->nsurface] = sum_s;
    /* enforceChargeConservation: the electron row is the sum of the positive-ion rows
       (chemistry.f90:333-335) */
    if (pt->p[P_ENFORCECHARGE] != 0.0) {
        double e = 0.0;
        for (int q = 0; q < n->nion; q++) e += ydot[n->ion_list[q]];
        ydot[n->nelec] = e;
    }
    /* density (physics-core.f90:90-103) */
    double dd = 0.0;
    double d = y[neq - 1];
    const double n0 = pt->p[P_DENS];
    if (d < pt->p[P_FINALDENS] && pt->p[P_FREEFALL] != 0.0)
        dd = pt->p[P_FREEFALLFAC] * pow(pow(d, 4.) / n0, (double)0.33f) *
             pow(8.4e-30 * n0 * (pow(d / n0, (double)0.33f) - 1.), 0.5);
    ydot[neq - 1] = dd;
    return 0;
}

/* ---------------------------------------------------------------- schedules */
double ucl_next_target(int model, double t)
{
    if (model == 1) {   /* hotcore.f90:56-72 */
        if (t > 1.0e6) return (t + 1.0e5) * SECONDS_PER_YEAR;
        if (t > 1.0e5) return (t + 1.0e4) * SECONDS_PER_YEAR;
        if (t > 1.0e4) return (t + 1000.0) * SECONDS_PER_YEAR;
        if (t > 1000) return (t + 100.0) * SECONDS_PER_YEAR;
        if (t > 100) return (t + 10.0) * SECONDS_PER_YEAR;
        if (t > 0.0) return (t * 10.0) * SECONDS_PER_YEAR;
        return SECONDS_PER_YEAR * 1.0e-7;
    }
    /* cloud.f90:39-51 */
    if (t >= 1.0e6) return (t + 1.0e5) * SECONDS_PER_YEAR;
    if (t > 10.0) {
        double om = pow(10.0, floor(log10(t)));
        return ((floor(t / om) + 1.0) * om) * SECONDS_PER_YEAR;
    }
    if (t > 0.0) return 10 * t * SECONDS_PER_YEAR;
    return SECONDS_PER_YEAR * 1.0e-7;
}

void ucl_update_physics(const ucl_net *n, ucl_point *pt, double current_time_s)
{   /* physics-core.f90:75-88 (points=1), hotcore.f90:78-90 */
    pt->coldens = pt->cloud_size / 1.0 * pt->density;
    pt->av = pt->p[P_BASEAV] + pt->coldens / 1.6e21;
    pt->dust_temp = pt->gas_temp;
    if (pt->p[P_CRATTEN] != 0.0 && pt->p[P_IMPROVEDH2CRP] != 0.0) {
        /* physics-core.f90:117-155: H2 dissociation by attenuated cosmic rays, polynomial in
         * log10 of the column density (Padovani et al. 2018 fits, L or H model) */
        static const double ck_l[10] = {1.582911005330e7, -6.465722684896e6, 1.172189025424e6, -1.237950798073e5,
                                        8.393404654312e3, -3.788811358130e2, 1.138688455029e1, -2.197136304567e-1,
                                        2.469841278950e-3, -1.232393620924e-5};
        static const double ck_h[10] = {1.217227462831e7, -4.989649250304e6, 9.079152156645e5, -9.624890825395e4,
                                        6.551161486120e3, -2.968976216187e2, 8.959037875226e0, -1.735757324445e-1,
                                        1.959267277734e-3, -9.816996707980e-6};
        const double *ck = pt->p[P_IONMODEL] != 0.0 ? ck_h : ck_l;
        const double lc = log10(pt->coldens);
        double sum = 0.0, pw = 1.0;
        for (int k = 0; k < 10; k++) {
            sum += ck[k] * pw;
            pw *= lc;
        }
        pt->h2crp = pow(10.0, sum) * pt->p[P_ZETA];
    }
    if ((int)pt->p[P_MODEL] == 1) {
        static const double tempa[6] = {1.927e-1, 4.8560e-2, 7.8470e-3, 9.6966e-4, 1.706e-4, 4.74e-7};
        static const float tempb[6] = {0.5339f, 0.6255f, 0.8395f, 1.085f, 1.289f, 1.98f};
        int ti = (int)pt->p[P_TEMPINDX] - 1;
        double maxt = pt->p[P_MAXTEMP];
        if (pt->gas_temp < maxt) {
            double g = (pt->cloud_size / (pt->p[P_ROUT] * PC)) * ((double)1.0f / (double)1.0f);
            g = pow(g, (double)-0.5f);
            g = pt->p[P_TEMP] + ((tempa[ti] * pow(current_time_s / SECONDS_PER_YEAR, (double)tempb[ti])) * g);
            if (g > maxt) g = maxt;
            pt->gas_temp = g;
        }
        pt->dust_temp = pt->gas_temp;
    }
    (void)n;
}
