// Device-side data structures and per-point physics for the UCLCHEM chemistry solve.
//
// One network descriptor (DevNet) holds device pointers to the tables emitted by the
// MakeRates backend (uclchem_amd/makerates_hip.py).  The physics restated here follows
// src/fortran_src/{rates,surfacereactions,photoreactions,physics-core,cloud,hotcore}.f90;
// default-kind (single precision) literals of the reference are written as float
// literals widened to double so the arithmetic matches the compiled reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ucg {

// pointer into workgroup-local memory (LDS); keeps the compiler from falling back to
// flat addressing, whose loads wait on both the LDS and the vector-memory counters
#define LDS __attribute__((address_space(3)))
// global-memory views of the table and workspace pointers: loads through them are global_load
// (counted by vmcnt only), so waiting for an LDS value does not also wait for them, as it
// must for flat (generic) loads
#define GLB __attribute__((address_space(1)))
template <class T> __device__ __forceinline__ const GLB T *glb(const T *p) { return (const GLB T *)p; }
template <class T> __device__ __forceinline__ GLB T *glbw(T *p) { return (GLB T *)p; }
__device__ __forceinline__ int2 ld_int2(const GLB int *p, int i)
{
    int2 r;
    r.x = p[2 * i];
    r.y = p[2 * i + 1];
    return r;
}
__device__ __forceinline__ int4 ld_int4(const GLB int *p, int i)
{
    int4 r;
    r.x = p[4 * i];
    r.y = p[4 * i + 1];
    r.z = p[4 * i + 2];
    r.w = p[4 * i + 3];
    return r;
}

constexpr int NT = 256;          // threads per workgroup (one workgroup solves one grid point)
constexpr int EPT = 2;           // state elements owned per thread => NEQ <= 512
constexpr int JZ_LONG = 16;      // Jacobian entries with more contributions than this...
constexpr int JZ_GROUP = 16;     // ...are summed by a group of this many lanes
constexpr int MAXOPS = 6;
constexpr int MAXFAC = 10;
constexpr int NTYPES = 23;

// reaction type codes (reaction.py:5-28 order, TWOBODY last)
enum { T_PHOTON, T_CRP, T_CRPHOT, T_FREEZE, T_DESORB, T_THERM, T_DESOH2, T_DESCR, T_DEUVCR, T_H2FORM,
       T_ER, T_ERDES, T_LH, T_LHDES, T_BULKSWAP, T_SURFSWAP, T_IONOPOL1, T_IONOPOL2, T_CRS, T_EXSOLID,
       T_EXRELAX, T_GAR, T_TWOBODY };
// term-program / Jacobian factor codes (>= 0: species index)
enum { OP_NONE = -1, OP_BLR = -2, OP_SWAP = -3, OP_DIVSM = -4, FAC_D = -5 };

// point-parameter vector layout (shared with include/uclchem_gpu.h)
enum { P_DENS, P_TEMP, P_ZETA, P_RADFIELD, P_BASEAV, P_ROUT, P_RIN, P_FINALTIME, P_FREEFALL,
       P_FINALDENS, P_RELTOL, P_ABSFAC, P_ABSMIN, P_MODEL, P_TEMPINDX, P_MAXTEMP, P_FREEZEFAC,
       P_ENDATFINALDENS, P_FREEFALLFAC, P_MXSTEP, P_MAXSTEPS,
       P_METALLICITY, P_ION, P_FH, P_FHE, P_FC, P_FO, P_FN, P_FS, P_FMG, P_FSI, P_FCL, P_FP, P_FFE, P_FF,
       P_FD, P_FLI, P_FNA, P_FPAH, P_F15N, P_F13C, P_F18O,
       P_EBMAXH2, P_EBMAXCR, P_EBMAXUVCR, P_EPSILON, P_UV_YIELD, P_PHI, P_UVCREFF, P_OMEGA,
       P_DESORB, P_H2DESORB, P_CRDESORB, P_UVDESORB, P_THERMDESORB, P_CRATTEN, P_IONMODEL, P_IMPROVEDH2CRP,
       P_ENFORCECHARGE, P_COLLAPSE, P_NPARAM = 64 };

// physical constants (constants.f90, surfacereactions.f90:33-51)
constexpr double K_BOLTZ = 1.38065040e-16;
constexpr double AMU = 1.66053892e-24;
constexpr double PI_F = (double)3.141592654f;
constexpr double PC = 3.086e18;
constexpr double SECONDS_PER_YEAR = 3.16e7;
constexpr double GRAIN_RADIUS = 1.0e-5;
constexpr double MAX_GRAIN_TEMP = 150.0;
constexpr double MIN_SURFACE_ABUND = 1.0e-20;
constexpr double GRAIN_CROSSSECTION_PER_H = 0.5 * (7.908e-22 + 8.473e-22);
constexpr double GRAIN_SURFACEAREA_PER_H = 4.0 * GRAIN_CROSSSECTION_PER_H;
constexpr double GAS_DUST_DENSITY_RATIO =
    (4.0 * PI_F * (GRAIN_RADIUS * GRAIN_RADIUS * GRAIN_RADIUS) * 3.0 * 100.0) / (3.0 * AMU);
constexpr double NUM_SITES_PER_GRAIN = GRAIN_RADIUS * GRAIN_RADIUS * 1.5e15 * 4.0 * PI_F;
constexpr double SURFACE_COVERAGE0 = 0.5 * GAS_DUST_DENSITY_RATIO / NUM_SITES_PER_GRAIN;  // bulkGainFromMantleBuildUp
constexpr double MIN_ABUND = 1.0e-30;

struct DevNet {
    int32_t nspec, neq, nreac, nice, nsurfl, nbulkl, nfp, nswap, nnz, ld;
    const double *alpha, *beta, *gama, *tmin, *tmax, *mass, *rstat;  // rstat [nreac*5]
    const int32_t *rtype, *re, *pr, *ice_kind, *n_d, *ops, *fac, *nfac;
    const int32_t *prog;   // [nreac*8] {n_d, ops[6], nfac}, 32-byte aligned records
    const int32_t *facp;   // [nreac*4] factor codes as int16 x 8 (density factors first), -1 padded
    const int32_t *fac_off, *jc_pk;   // partial offsets [nreac+1]; contributions (partial << 4 | coef + 8)
    int32_t npart, npart_pad;         // number of partials; LDS offset of the scratch after them
    // balanced production/loss sums: tasks {start, end, species, 0 prod | 1 loss} (longest
    // first), per species the task lists [2i, 2i+1) prod and [2i+1, 2i+2) loss
    const int32_t *gt_rec, *gt_sp_ptr, *gt_sp_task;
    int32_t ngt, gt_off;              // task count; LDS offset (after the fluxes) of the partial sums
    const int8_t *extrapolate;
    const int32_t *type_ranges;                                     // [NTYPES*2]
    const int32_t *loss_ptr, *loss_idx, *prod_ptr, *prod_idx, *swap_list;
    const int32_t *surface_list, *bulk_list, *bulk_partner_of_surface, *surface_partner_of_bulk;
    const int32_t *freeze_partners, *is_surface, *is_bulk;          // is_*: [neq]
    const int32_t *jnz_row, *jnz_col, *jnz_ptr, *jc_k, *jc_m, *jc_c;
    // Jacobian entries in evaluation order: nj_short with at most JZ_LONG contributions (one
    // lane each, dealt so that every lane gets a near-equal share), nj_long with more (one
    // group of JZ_GROUP lanes each)
    int32_t nj_short, nj_long;
    const int32_t *jz_short, *jz_long;
    const int32_t *aff_ptr, *aff_sp, *aff_cf, *sw_ptr, *sw_k, *sw_c;
    // special species (0-based; missing -> nspec, the density slot)
    int32_t nh, nh2, nhe, nc, ncx, nn, no, nf, np, nmg, nmgx, nna, nli, npah, nsx, nsix, nclx, nd,
        n18o, n15n, n13c, nelec, nco, nsurface, nbulk;
    int32_t r_h2form_ct, r_h2form_er, r_h2form_erdes, r_hfreeze, r_efreeze, r_h2freeze, r_h2_hv, r_co_hv,
        r_c_hv, r_h2_crp;   // r_h2_crp: H2 + CRP -> H + H, or -1
    int32_t nion;             // positive ions (names holding '+'), for enforceChargeConservation
    const int32_t *ion_list;
    int32_t ngar;             // GAR reactions; gar [ngar*7] their Weingartner & Draine parameters
    const double *gar;
    // static-pattern sparse Newton plan (augmented order of size nw_N = neq + 2)
    int32_t nw_N, nw_ns, nw_T, nw_nsp, nw_nlev, nw_nph, nw_nval;
    int32_t nw_stage;            // 1: factors and solve program are staged in LDS; 0: they stay in the workspace
    const int32_t *nw_perm, *nw_inv, *nw_sp_ptr, *nw_sp_col, *nw_sp_dpos, *nw_lev_ptr, *nw_lev_rows, *nw_ph_ptr;
    const int32_t *nw_tg_slot, *nw_tg_ptr, *nw_tg_div, *nw_pr_l, *nw_pr_u, *nw_jnz_slot, *nw_diag_slot;
    const int32_t *nw_ph_slot, *nw_tr_slot, *nw_s1c, *nw_s1r, *nw_s2c, *nw_s2r, *nw_gs_ptr, *nw_gs_slot;
    const int32_t *nw_misc, *nw_nbrow, *nw_nsrow;
    int32_t nw_nerow;                   // charge row (enforceChargeConservation): slots of the electron row,
    const int32_t *nw_erow, *nw_eion;   // then of its entries in the positive-ion columns
    const int32_t *sv_rec;   // solve rows {row, first, diagonal, end}: sparse rows in level order, then tail rows
    const int32_t *nw_pr;    // Crout pairs interleaved {L slot, U slot}
    // the last elimination phase is the Schur-complement update of the dense tail; its
    // sources (sparse factor entries) are gathered into LDS beside the staged tail
    int32_t nlp_src;          // number of distinct source entries (0: the phase runs in place)
    const int32_t *lp_src;    // [nlp_src] factor slots of the sources
    const int32_t *lp_pr;     // the phase's pairs as {source index, source index}
    const int8_t *nw_unscaled;   // [nval] 1 where the Newton matrix takes the J-form value without gamma
};

// LDS bytes the sparse solve stages: factors (lut bytes each; FP32 all of them, FP64 the
// sparse part), column indices, row records, permutation, work vector and the dense-tail
// product's scratch (16-byte aligned parts)
__host__ __device__ inline int solve_stage_bytes(const DevNet &N, int lut)
{
    const int nst = lut == 4 ? N.nw_nval : N.nw_nsp;
    auto al = [](int x) { return (x + 15) & ~15; };
    return al(nst * lut) + al(N.nw_nsp * 4) + N.nw_N * 16 + al((N.nw_N + N.nw_nlev + 1) * 4) + N.nw_N * 8 +
           N.nw_T * 8;
}

// length (in doubles) of the LDS region shared by the reaction fluxes (RHS), the term
// partials plus two NEQ-long scratch vectors (Jacobian assembly), and the staged factors
// (all of them in FP32, the dense tail in FP64) plus the permuted work vector.  Plans too
// large to stage (strongly coupled networks with a wide dense tail) keep their factors in
// the workspace and need only the work vector here.
__host__ __device__ inline int flux_region_len(const DevNet &N)
{
    const int npad = (N.neq + 7) & ~7, Npad = (N.nw_N + 7) & ~7;
    const int a = N.npart_pad + 2 * npad;
    const int g = N.gt_off + N.ngt;                  // fluxes + chunk partial sums
    const int b0 = N.nreac > a ? N.nreac : a;
    const int b1 = b0 > g ? b0 : g;
    const int b = b1 > 256 ? b1 : 256;               // >= one value per thread (ordered_sum)
    if (!N.nw_stage) return b > Npad ? b : Npad;               // unstaged plan: the work vector only
    const int t64 = N.nw_T * N.nw_T + ((N.nlp_src + 7) & ~7) + Npad;   // FP64: staged tail + Schur sources
    const int s32 = (solve_stage_bytes(N, 4) + 7) / 8, s64 = (solve_stage_bytes(N, 8) + 7) / 8;
    const int f = s32 > s64 ? s32 : s64;
    const int c = t64 > f ? t64 : f;
    return b > c ? b : c;
}

// static spline tables of the photo-rate functions (photoreactions.f90:14-39),
// second derivatives precomputed once on the host
struct PhotoTables {
    double lam[30], xl[30], xl_d2[30];
    double nco[8], nh2[6];
    double ya[7][6], y2a[7][6];   // the 7x6 view of SCO_GRID the reference's splie2 call sees
};

extern __constant__ PhotoTables c_photo;

// ---------------------------------------------------------------- photo-rate helpers

__device__ __forceinline__ double splint_dev(const double *xa, const double *ya, const double *y2a, int n,
                                             double x)
{
    int jlo = 0, jhi = n + 1;
    while (jhi - jlo != 1) {
        int jm = (jhi + jlo) >> 1;
        if (x > xa[jm - 1]) jlo = jm; else jhi = jm;
    }
    if (jlo == 0) { jlo = 1; jhi = 2; }
    if (jlo == n) { jlo = n - 1; jhi = n; }
    double h = xa[jhi - 1] - xa[jlo - 1];
    double a = (xa[jhi - 1] - x) / h;
    double b = (x - xa[jlo - 1]) / h;
    return a * ya[jlo - 1] + b * ya[jhi - 1] +
           ((a * a * a - a) * y2a[jlo - 1] + (b * b * b - b) * y2a[jhi - 1]) * (h * h) / 6.0;
}

__device__ __forceinline__ void spline_dev(const double *x, const double *y, int n, double *y2)
{   // natural cubic spline (photoreactions.f90:340-418)
    double u[8];
    y2[0] = 0.0;
    u[0] = 0.0;
    for (int i = 1; i < n - 1; i++) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) /
                    (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    y2[n - 1] = 0.0;
    for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
}

__device__ inline double scatter_dev(double x1, double av)
{   // photoreactions.f90:138-196
    const double c[6] = {1.0, 2.006, -1.438, 7.364e-01, -5.076e-01, -5.920e-02};
    const double k1[6] = {7.514e-01, 8.490e-01, 1.013, 1.282, 2.005, 5.832};
    double lv = x1;
    if (x1 < c_photo.lam[0]) lv = c_photo.lam[0];
    if (x1 > c_photo.lam[29]) lv = c_photo.lam[29];
    double xl = splint_dev(c_photo.lam, c_photo.xl, c_photo.xl_d2, 30, lv);
    if (xl < 0.0) xl = 0.0;
    double tl = (av / 1.086) * xl;
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

__device__ inline double h2_self_shielding_dev(double nh2)
{   // photoreactions.f90:95-133 with dopplerWidth = turbVel/(xl*1e-8) = 1e5, radWidth = 8e7
    const double doppler = 1.0 / (1000.0 * 1.0e-8), radwidth = 8.0e7;
    double taud = 0.5 * nh2 * (double)1.5e-2f * 1.0e-2 / doppler;
    double sj;
    if (taud == 0.0) sj = 1.0;
    else if (taud < 2.0) sj = exp(-0.6666667 * taud);
    else if (taud < 10.0) sj = 0.638 * pow(taud, -1.25);
    else if (taud < 100.0) sj = 0.505 * pow(taud, -1.15);
    else sj = 0.344 * pow(taud, -1.0667);
    double r = radwidth / (1.7724539 * doppler);
    double t = 3.02 * pow(r * 1.0e3, -0.064);
    double u = sqrt(taud * r) / t;
    double sr = r / (t * sqrt(0.78539816 + u * u));
    return sj + sr;
}

// H2 photodissociation given scatter(1000 A, Av) (photoreactions.f90:46-55)
__device__ inline double h2_photodiss_dev(double nh2col, double radfield, double scat1000)
{
    return 5.18e-11 * (radfield / (double)1.7f) * scat1000 * h2_self_shielding_dev(nh2col);
}

// CO photodissociation (photoreactions.f90:57-72, 238-287) in parts, so that a wave can
// evaluate the seven row splines and the scattering term in parallel lanes.
__device__ __forceinline__ void co_log_columns(double nh2col, double ncocol, double &lognh2, double &lognco)
{
    lognco = log10(ncocol + 1.0);
    lognh2 = log10(nh2col + 1.0);
    if (lognco < c_photo.nco[0]) lognco = c_photo.nco[0];
    if (lognh2 < c_photo.nh2[0]) lognh2 = c_photo.nh2[0];
    if (lognco > c_photo.nco[6]) lognco = c_photo.nco[6];
    if (lognh2 > c_photo.nh2[5]) lognh2 = c_photo.nh2[5];
}
__device__ __forceinline__ double co_row_spline(int j, double lognh2)
{   // row j of the 7x6 shielding table interpolated in log N(H2) (splin2's first pass)
    return splint_dev(c_photo.nh2, c_photo.ya[j], c_photo.y2a[j], 6, lognh2);
}
__device__ __forceinline__ double co_shielding(const double yy[7], double lognco)
{   // splin2's second pass along log N(CO)
    double y2[7];
    spline_dev(c_photo.nco, yy, 7, y2);
    return pow(10.0, splint_dev(c_photo.nco, yy, y2, 7, lognco));
}
__device__ __forceinline__ double co_scatter(double nh2col, double ncocol, double av)
{   // dust scattering at the mean wavelength lbar (photoreactions.f90:289-300)
    double lu = log10(fabs(ncocol) + 1.0), lw = log10(fabs(nh2col) + 1.0);
    double lb = (5675.0 - 200.6 * lw) - (571.6 - 24.09 * lw) * lu + (18.22 - 0.7664 * lw) * (lu * lu);
    if (lb > 1076.1) lb = 1076.1;
    if (lb < 913.6) lb = 913.6;
    return scatter_dev(lb, av);
}
__device__ __forceinline__ double co_rate_from_parts(double radfield, double ssf, double scat)
{
    return 2.0e-10 * (radfield / (double)1.7f) * ssf * scat;
}
__device__ inline double co_photodiss_dev(double nh2col, double ncocol, double radfield, double av)
{
    double lognh2, lognco;
    co_log_columns(nh2col, ncocol, lognh2, lognco);
    double yy[7];
    for (int j = 0; j < 7; j++) yy[j] = co_row_spline(j, lognh2);
    const double ssf = co_shielding(yy, lognco);
    return co_rate_from_parts(radfield, ssf, co_scatter(nh2col, ncocol, av));
}

__device__ inline double c_ionization_dev(double alpha, double gamma, double t, double ncol, double nh2col,
                                          double av, double radfield)
{   // photoreactions.f90:74-87
    double tauc = gamma * av + 1.1e-17 * ncol + (0.9 * pow(t, 0.27) * pow(nh2col / 1.59e21, 0.45));
    return alpha * (radfield / (double)1.7f) * exp(-tauc);
}

__device__ inline double h2_form_efficiency_dev(double gas_temp, double dust_temp)
{   // surfacereactions.f90:63-107
    double tv = 1.45e5 * sqrt(gas_temp / 1.0e2);
    double g2 = gas_temp / 1.0e2;
    double stick = 1.0 / (1.0 + 0.04 * sqrt(gas_temp + dust_temp) + 0.2 * g2 + 0.08 * (g2 * g2));
    const double hflux = 1.0e-10;
    double q = 1.0 + sqrt((3.0e4 - 110.0) / (450.0 - 110.0));
    double f1 = 0.005 * hflux / (2 * 3.0e12 * exp(-320.0 / dust_temp));
    double f2 = 1.0 * (q * q) / 4.0 * exp(-110.0 / dust_temp);
    double eps = 1.0 / (1.0 + 1.3e13 / (2 * hflux) * exp(-1.5 * 3.0e4 / dust_temp) * (q * q));
    double sil = 1.0 / (1.0 + f1 + f2) * eps;
    q = 1.0 + sqrt((3.0e4 - 260.0) / (800.0 - 260.0));
    f1 = 0.005 * hflux / (2 * 3.0e12 * exp(-520.0 / dust_temp));
    f2 = 1.0 * (q * q) / 4.0 * exp(-260.0 / dust_temp);
    eps = 1.0 / (1.0 + 1.3e13 / (2 * hflux) * exp(-1.5 * 3.0e4 / dust_temp) * (q * q));
    double gra = 1.0 / (1.0 + f1 + f2) * eps;
    return 0.5 * tv * (8.473e-22 * sil + 7.908e-22 * gra) * stick;
}

__device__ __forceinline__ double sticking_dev(double s0, double tcrit, double t)
{
    double r = t / tcrit;
    return s0 * (1.0 + 2.5 * r) / pow(1.0 + r, 2.5);
}

template <typename PP>
__device__ inline double densdot_dev(const PP p, double d)
{   // physics-core.f90:90-103
    if (d < p[P_FINALDENS] && p[P_FREEFALL] != 0.0) {
        double n0 = p[P_DENS];
        return p[P_FREEFALLFAC] * pow(pow(d, 4.0) / n0, (double)0.33f) *
               pow(8.4e-30 * n0 * (pow(d / n0, (double)0.33f) - 1.0), 0.5);
    }
    return 0.0;
}

__device__ inline double next_target_dev(int model, double t)
{
    if (model == 1) {   // hotcore.f90:56-72
        if (t > 1.0e6) return (t + 1.0e5) * SECONDS_PER_YEAR;
        if (t > 1.0e5) return (t + 1.0e4) * SECONDS_PER_YEAR;
        if (t > 1.0e4) return (t + 1000.0) * SECONDS_PER_YEAR;
        if (t > 1000) return (t + 100.0) * SECONDS_PER_YEAR;
        if (t > 100) return (t + 10.0) * SECONDS_PER_YEAR;
        if (t > 0.0) return (t * 10.0) * SECONDS_PER_YEAR;
        return SECONDS_PER_YEAR * 1.0e-7;
    }
    // cloud.f90:39-51
    if (t >= 1.0e6) return (t + 1.0e5) * SECONDS_PER_YEAR;
    if (t > 10.0) {
        double om = 1.0;
        int e = (int)floor(log10(t));
        for (int i = 0; i < e; i++) om *= 10.0;
        return ((floor(t / om) + 1.0) * om) * SECONDS_PER_YEAR;
    }
    if (t > 0.0) return 10 * t * SECONDS_PER_YEAR;
    return SECONDS_PER_YEAR * 1.0e-7;
}

// ---------------------------------------------------------------- collapse models
// The parameterised prestellar-core collapses of collapse.f90 (Priestley et al. 2018, AJ 156,
// 51): mode 1 Bonnor-Ebert sphere with overdensity 1.1, 2 with overdensity 4, 3 magnetised
// filament, 4 ambipolar-diffusion cloud.  Literals the reference writes as default-kind reals
// are float32 values (F4), as its compiled code holds them.
#define F4(x) ((double)(x##f))
constexpr double MH = 1.67262164e-24;   // constants.f90:7
constexpr double AU_PER_PC = 2.063e5;   // constants.f90:12 ("au": parsecs to AU)

__device__ inline double collapse_next_target(double t)
{   // collapse.f90:65-77 (years in, seconds out)
    if (t > 10000) return (t + 1000.0) * SECONDS_PER_YEAR;
    if (t > 1000) return (t + 100.0) * SECONDS_PER_YEAR;
    if (t > 0.0) return (t * 10) * SECONDS_PER_YEAR;
    return 3.16e7 * 10.e-8;
}

struct CollapseFit {
    int mode;        // 1..4
    double maxt;     // maxTime [yr] of modes 1 and 2 (collapse.f90:35-40)

    __device__ static double unit_t()
    {   // (2 pi G rho0)^-1/2 in years (collapse.f90:210-211)
        return pow(2 * PI_F * 6.67e-8 * 2.2e4 * MH, -0.5) / SECONDS_PER_YEAR;
    }
    __device__ static double sound_speed() { return sqrt(1.38e-16 * 10 / 2 / MH); }
    __device__ static double unit_r()
    {   // c_s (2 pi G rho0)^-1/2 in pc (collapse.f90:189-190)
        return sound_speed() * pow(2 * PI_F * 6.67e-8 * 2.2e4 * MH, -0.5) / PC;
    }
    // central density, radius parameter and slope against time in years (collapse.f90:201-264)
    __device__ double rho0(double t) const
    {
        double l;
        switch (mode) {
        case 1: l = F4(61.8) * pow(maxt - t, F4(-0.01)) - F4(49.4); break;
        case 2: l = F4(68.4) * pow(maxt - t, F4(-0.01)) - F4(55.7); break;
        case 3: l = F4(3.54) * pow(F4(5.47) - t / unit_t(), F4(-0.15)) - F4(2.73); break;
        default:
            if (t <= 6.0) return 2.0e3 + 1.7e3 * (t / 6.0 - 1.0);
            l = F4(5.3) * pow(F4(16.138) - 1e-6 * t, F4(-0.1)) - 1.0;
        }
        return pow(10.0, l);
    }
    __device__ double r0(double t) const
    {
        double l;
        switch (mode) {
        case 1: l = F4(-28.5) * pow(maxt - t, F4(-0.01)) + F4(28.93); break;
        case 2: l = F4(-39.0) * pow(maxt - t, F4(-0.01)) + F4(38.7); break;
        case 3: l = F4(-1.34) * pow(F4(5.47) - t / unit_t(), F4(-0.15)) + F4(1.47); break;
        default: l = F4(-2.57) * pow(F4(16.138) - 1e-6 * t, F4(-0.1)) + F4(1.85);
        }
        return pow(10.0, l);
    }
    __device__ double slope(double t) const
    {
        switch (mode) {
        case 1: return 2.4;
        case 2: return F4(1.9) + 0.5 * exp(-t / 1e5);
        case 3: {   // x**9 by repeated squaring
            const double x = t / unit_t() / F4(5.47), x2 = x * x, x4 = x2 * x2;
            return 2.0 - 0.5 * (x4 * x4 * x);
        }
        default: {  // x**40 = x^32 x^8
            const double x = 1e-6 * t / F4(16.138), x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
            return F4(2.4) - F4(0.2) * (x16 * x16 * x8);
        }
        }
    }
    // density at radius r [pc] for the profile (rho0, r0, a) (collapse.f90:178-198)
    __device__ double rho(double r, double rho0v, double r0v, double a) const
    {
        switch (mode) {
        case 1:
        case 2: return rho0v / (1 + pow(r * AU_PER_PC / r0v, a));
        case 3: {
            const double x = r / unit_r() / r0v;
            return 2.2e4 * rho0v / pow(1 + x * x, a);
        }
        default: return rho0v / (1 + pow(r / 7.5e-1 / r0v, a));
        }
    }
    // radial velocity [cm/s] of modes 3 and 4 and its time fits (collapse.f90:266-383)
    __device__ double vr(double r, double rmin, double vmin, double a) const
    {
        if (mode == 3) {
            const double nr = r / unit_r() - rmin;
            double v;
            if (nr < 0.0) {
                const double x = nr / rmin;
                v = vmin * (x * x - 1);
            } else {
                v = vmin * (exp(-2.0 * a * nr) - 2 * exp(-a * nr));
            }
            return sound_speed() * v;
        }
        const double rmid = 0.5, r75 = r / 7.5e-1, nr = r75 - rmin;
        double v;
        if (r75 < rmin) {
            const double x = nr / rmin;
            v = vmin * (x * x - 1);
        } else if (r75 <= rmid) {
            v = (vmin - a) * pow(nr / (rmid - rmin), F4(0.3)) - vmin;
        } else {
            v = a / (1.0 - rmid) * (r75 - rmid) - a;
        }
        return 1e3 * v;
    }
    __device__ double rmin(double t) const
    {
        if (mode == 3) {
            const double x = t / unit_t();
            if (x == 0.0) return 7.2;
            const double lx = log(x);
            if (lx < 1.6) return F4(-1.149) * x + F4(7.2);
            if (lx < 1.674) return F4(-9.2) * lx + F4(16.25);
            return -22.0 * lx + F4(37.65);
        }
        const double t6 = 1e-6 * t;
        if (t6 <= F4(10.2)) return F4(-0.0039) * t6 + F4(0.49);
        if (t6 <= F4(15.1)) return F4(-0.0306) * (t6 - F4(10.2)) + F4(0.45);
        return F4(-0.282) * (t6 - F4(15.1)) + F4(0.3);
    }
    __device__ double vmin(double t) const
    {
        if (mode == 3) {
            const double x = t / unit_t();
            if (x == 0.0) return 0.0;
            const double lx = log(x);
            if (lx < 1.6) return F4(0.0891) * x;
            if (lx < 1.674) return 5.5 * lx - F4(8.37);
            return F4(18.9) * lx - F4(30.8);
        }
        const double t6 = 1e-6 * t;
        return F4(3.44) * pow(F4(16.138) - t6, F4(-0.35)) - F4(0.7);
    }
    __device__ double vslope(double t) const
    {
        if (mode == 3) {
            const double x = t / unit_t();
            if (x == 0.0) return 0.4;
            const double lx = log(x);
            if (lx < 1.6) return F4(0.0101) * x + F4(0.4);
            if (lx < 1.674) return F4(0.695) * lx - F4(0.663);
            return F4(2.69) * lx - 4.0;
        }
        const double t6 = 1e-6 * t;
        if (t6 <= F4(10.2)) return F4(0.143) * t6;
        return F4(0.217) * (t6 - F4(10.2)) + F4(1.46);
    }
};

__device__ inline CollapseFit collapse_fit(int mode)
{   // collapse.f90:33-46 (modes 3 and 4 have no maxTime)
    CollapseFit f;
    f.mode = mode;
    f.maxt = mode == 1 ? 1.175e6 : mode == 2 ? 1.855e5 : 0.0;
    return f;
}

}  // namespace ucg
