// Batched stiff chemistry solve for MI355X (gfx950).
//
// One workgroup (256 threads, 4 wave64) integrates one grid point at a time and then
// takes the next point from a device work counter, so uneven points (step counts vary
// by >30x) never stall a whole batch: this is the persistent "retire and refill"
// kernel.  Per point it reproduces UCLCHEM's per-interval semantics
// (src/fortran_src/wrap.f90:598-690, chemistry.f90:145-292): the output schedule,
// rates frozen per interval attempt, a cold-start BDF integration per interval with
// DVODE's ITASK=1 contract, the retry policy, the 1e-30 clamp and the physics update.
//
// Memory layout per workgroup slot: the dense Newton matrix (column-major, leading
// dimension ld) in HBM (FP64 or FP32, template LUT), everything else in LDS or
// registers: rate and flux arrays [NREAC], the RHS input/output and solve vectors
// [NEQ], and each thread's slice of the Nordsieck history (registers).
#include "ucg_device.h"
#include "bdf_core.h"

namespace ucg {

__constant__ PhotoTables c_photo;

// LDS scalar slots
enum { S_SM, S_SBK, S_BLR, S_TS, S_SPRE, S_SCAT, S_H2COL, S_COCOL, S_CCOL, S_PHI, S_DENS, S_TGAS, S_TDUST,
       S_AV, S_CS, S_COLDENS, S_ABSFAC, S_FLAG, S_H2CRP, S_RAD, S_MASS, S_NSCAL = 24 };

// Large collective routines can be kept out of line to bound the kernel's register
// allocation (UCG_BIG=__attribute__((noinline))); inlined by default.
#ifndef UCG_BIG
#define UCG_BIG
#endif
#ifndef UCG_WAVES_PER_EU
#define UCG_WAVES_PER_EU 1
#endif
#ifndef UCG_LDS_SCHUR
#define UCG_LDS_SCHUR 1     // FP64: the tail's Schur update from sources gathered into LDS
#endif

struct Lds {
    double *rate;                                   // workspace (global memory)
    LDS double *flux, *y, *f, *b, *g, *red, *sc, *prm;
};

__device__ __forceinline__ int tid() { return (int)threadIdx.x; }

__device__ inline double block_sum(double v, LDS double *red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();
    if ((tid() & 63) == 0) red[tid() >> 6] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) r += red[w];
    return r;
}

__device__ inline void block_sum2(double &a, double &b, LDS double *red)
{
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    __syncthreads();
    if ((tid() & 63) == 0) {
        red[tid() >> 6] = a;
        red[NT / 64 + (tid() >> 6)] = b;
    }
    __syncthreads();
    double ra = 0.0, rb = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) {
        ra += red[w];
        rb += red[NT / 64 + w];
    }
    a = ra;
    b = rb;
}

__device__ inline double block_max(double v, LDS double *red)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    __syncthreads();
    if ((tid() & 63) == 0) red[tid() >> 6] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < NT / 64; w++) r = fmax(r, red[w]);
    return r;
}

__device__ __forceinline__ bool type_active(const DevNet &N, int t)
{
    int a = N.type_ranges[2 * t], b = N.type_ranges[2 * t + 1];
    if (t == T_TWOBODY || t == T_BULKSWAP || t == T_SURFSWAP) return a >= 0;
    return a >= 0 && a != b;   // rates.f90 range blocks are skipped when idx1 == idx2
}

__device__ inline double freeze_out(const DevNet &N, const LDS double *prm, int j, double T, double Td)
{   // rates.f90:354-367
    double f = 1.0 + N.beta[j] * 16.71e-4 / (GRAIN_RADIUS * T);
    if (prm[P_FREEZEFAC] == 0.0 || Td > MAX_GRAIN_TEMP) return 0.0;
    const double thermal_vel = sqrt(8.0 * K_BOLTZ / (PI_F * AMU));
    return f * prm[P_FREEZEFAC] * N.alpha[j] * thermal_vel * sqrt(T / N.mass[N.re[3 * j]]) *
           GRAIN_CROSSSECTION_PER_H;
}

// Rate coefficients for the current interval attempt (rates.f90:21-343 preceded by the
// ice bookkeeping of chemistry.f90:167-200).  Reads y (LDS), writes rate (LDS), updates
// y[BULK], y[SURFACE], the column densities, the ice scalars and phi.
UCG_BIG __device__ void block_rates(const DevNet &N, const Lds L)
{
    const int neq = N.neq;
    const LDS double *prm = L.prm;
    LDS double *y = L.y, *sc = L.sc;
    double *rate = L.rate;
    const double T = sc[S_TGAS], Td = sc[S_TDUST], av = sc[S_AV], dens = sc[S_DENS];
    const double zeta = prm[P_ZETA], radfield = prm[P_RADFIELD];
    if (tid() == 0 && prm[P_FREEFALL] == 0.0) y[neq - 1] = dens;
    double sb = 0.0, ss = 0.0;
    for (int i = tid(); i < N.nbulkl; i += NT) sb += y[N.bulk_list[i]];
    for (int i = tid(); i < N.nsurfl; i += NT) ss += y[N.surface_list[i]];
    block_sum2(sb, ss, L.red);
    if (tid() == 0) {
        y[N.nbulk] = sb;
        y[N.nsurface] = ss;
        double sm = fmax(1e-30, ss), sbk = fmax(1e-30, sb);
        sc[S_SM] = sm;
        sc[S_SBK] = sbk;
        sc[S_BLR] = fmin(1.0, NUM_SITES_PER_GRAIN / (GAS_DUST_DENSITY_RATIO * sbk));
        double half = sc[S_CS] / 1.0;
        sc[S_H2COL] = 0.5 * y[N.nh2] * dens * half;
        sc[S_COCOL] = 0.5 * y[N.nco] * dens * half;
        sc[S_CCOL] = 0.5 * y[N.nc] * dens * half;
    }
    if (tid() == 64) sc[S_SCAT] = scatter_dev(1000.0, av);
    __syncthreads();
    const double sm = sc[S_SM], blr = sc[S_BLR], phi = sc[S_PHI];   // phi: the previous call's (rates.f90:118)
    // this call's grain-assisted-recombination phi = G0 e^(-2.5 Av) T^0.5 / n_e (rates.f90:321-325):
    // the GAR rates use it, CR desorption above still the previous one; stored for the next call
    const double phi_new = fmin(fmax(radfield * exp(-2.5 * av) * sqrt(T) / (y[neq - 1] * y[N.nelec]), 1e2), 1e6);
    const double photo_bulk_pow = pow((double)(1.0f - 0.007f), 1.0 + 0.5 / blr);
    const double ice_ratio = (double)0.3f;
    const bool act_crp = type_active(N, T_CRP), act_ph = type_active(N, T_PHOTON),
               act_crph = type_active(N, T_CRPHOT), act_fr = type_active(N, T_FREEZE),
               act_h2d = type_active(N, T_DESOH2), act_cr = type_active(N, T_DESCR),
               act_uv = type_active(N, T_DEUVCR),
               act_th = type_active(N, T_THERM) && prm[P_THERMDESORB] != 0.0,   // off: zero, no switch-off
               act_lh = type_active(N, T_LH), act_er = type_active(N, T_ER),
               act_io1 = type_active(N, T_IONOPOL1), act_io2 = type_active(N, T_IONOPOL2),
               act_crs = type_active(N, T_CRS), act_exr = type_active(N, T_EXRELAX),
               act_exs = type_active(N, T_EXSOLID), act_gar = type_active(N, T_GAR);
    const int gar0 = N.type_ranges[2 * T_GAR];
    const int lh0 = N.type_ranges[2 * T_LH], lhd0 = N.type_ranges[2 * T_LHDES];
    const int er0 = N.type_ranges[2 * T_ER], erd0 = N.type_ranges[2 * T_ERDES];
    const double h2eff = h2_form_efficiency_dev(T, Td);
    const bool desorb = prm[P_DESORB] != 0.0;
    // phase 1: base rates (every entry written)
    for (int k = tid(); k < N.nreac; k += NT) {
        const int t = N.rtype[k];
        double r = 0.0;
        switch (t) {
        case T_CRP:
            if (act_crp) r = N.alpha[k] * zeta;
            if (k == N.r_h2_crp && prm[P_IMPROVEDH2CRP] != 0.0) r = sc[S_H2CRP];   // rates.f90:39-41
            break;
        case T_PHOTON:
        case T_CRPHOT:
            if ((t == T_PHOTON && act_ph) || (t == T_CRPHOT && act_crph)) {
                if (t == T_PHOTON) r = N.alpha[k] * exp(-N.gama[k] * av) * radfield / (double)1.7f;
                else r = N.alpha[k] * N.gama[k] * 1.0 / (1.0 - prm[P_OMEGA]) * zeta * pow(T / 300, N.beta[k]);
                if (N.ice_kind[k] == 2) r = r * ice_ratio * photo_bulk_pow;
                else if (N.ice_kind[k] == 1) r = r * ice_ratio;
            }
            break;
        case T_FREEZE:
            if (act_fr) r = freeze_out(N, prm, k, T, Td);
            break;
        case T_DESOH2:
            if (act_h2d && desorb && prm[P_H2DESORB] != 0.0 && sm > MIN_SURFACE_ABUND)
                r = N.gama[k] > prm[P_EBMAXH2] ? 0.0 : prm[P_EPSILON] * h2eff;
            break;
        case T_DESCR:
            if (act_cr && desorb && prm[P_CRDESORB] != 0.0 && sm > MIN_SURFACE_ABUND)
                r = N.gama[k] > prm[P_EBMAXCR] ? 0.0 : 4.0 * PI_F * zeta * 1.64e-4 * (GRAIN_SURFACEAREA_PER_H) * phi;
            break;
        case T_DEUVCR:
            if (act_uv && desorb && prm[P_UVDESORB] != 0.0 && sm > MIN_SURFACE_ABUND && zeta > 0) {
                double v = GRAIN_CROSSSECTION_PER_H * prm[P_UV_YIELD] * 4.875e3 * zeta;
                v = v * (1 + (radfield / prm[P_UVCREFF]) * (1.0 / zeta) * exp((double)-1.8f * av));
                r = N.gama[k] > prm[P_EBMAXUVCR] ? 0.0 : v;
            }
            break;
        case T_THERM:
            if (act_th && N.rstat[5 * k + 1] != 0.0) r = N.rstat[5 * k] * exp(-N.gama[k] / Td);
            break;
        case T_LH:
            if (act_lh) {
                const int c = lhd0 + (k - lh0);
                double rl = 0.0, rd = 0.0;
                if (Td < MAX_GRAIN_TEMP && sm > MIN_SURFACE_ABUND) {
                    const double *st = N.rstat + 5 * k;
                    double v1 = st[0], v2 = st[1], e1 = st[2], e2 = st[3], tunnel = st[4];
                    double diffuse = v1 * exp(-0.5 * e1 / Td);
                    diffuse = diffuse + v2 * exp(-0.5 * e2 / Td);
                    double desorb = v1 * exp(-e1 / Td);
                    desorb = desorb + v2 * exp(-e2 / Td);
                    double reac = N.gama[k] / Td;
                    if (reac > tunnel) reac = tunnel;
                    reac = fmax(v1, v2) * exp(-reac);
                    reac = reac / (reac + desorb + diffuse);
                    rl = N.alpha[k] * reac * diffuse * GAS_DUST_DENSITY_RATIO / NUM_SITES_PER_GRAIN;
                    rd = N.rstat[5 * c + 1] != 0.0 ? 0.0 : N.rstat[5 * c] * rl;
                    rl = rl - rd;
                }
                rate[c] = rd;
                r = rl;
            }
            break;
        case T_LHDES:
            if (act_lh) continue;   // written together with its LH partner
            break;
        case T_ER:
            if (act_er) {
                const int c = erd0 + (k - er0);
                double rl = freeze_out(N, prm, k, T, Td) * exp(-N.gama[k] / Td);
                double rd = N.rstat[5 * c + 1] != 0.0 ? 0.0 : N.rstat[5 * c] * rl;
                rate[c] = rd;
                r = rl - rd;
            }
            break;
        case T_ERDES:
            if (act_er) continue;
            break;
        case T_BULKSWAP:
            if (!(Td > MAX_GRAIN_TEMP || sm < MIN_SURFACE_ABUND) && N.rstat[5 * k + 2] != 0.0)
                r = N.rstat[5 * k] * exp(-N.rstat[5 * k + 1] / Td);
            break;
        case T_SURFSWAP:
            r = (Td > MAX_GRAIN_TEMP || sm < MIN_SURFACE_ABUND) ? 0.0 : 1.0;
            break;
        case T_TWOBODY:
            r = N.alpha[k] * pow(T / 300.0, N.beta[k]) * exp(-N.gama[k] / T);
            break;
        case T_IONOPOL1:   // rates.f90:301-305 (KIDA ion-polar formula 1)
            if (act_io1) r = N.alpha[k] * N.beta[k] * (0.62 + 0.4767 * N.gama[k] * sqrt(300.0 / T));
            break;
        case T_IONOPOL2:   // rates.f90:307-313 (formula 2; 300.0 and 10.526 are default-real literals)
            if (act_io2)
                r = N.alpha[k] * N.beta[k] *
                    (1.0 + 0.0967 * N.gama[k] * sqrt(300.0 / T) + N.gama[k] * N.gama[k] * 300.0 / ((double)10.526f * T));
            break;
        case T_CRS:        // rates.f90:158-165: excitation by cosmic rays (Shingledecker et al. 2018)
            if (act_crs) r = N.alpha[k] * (N.beta[k] * (N.gama[k] / 100) * ((double)8.6f * zeta * (double)1.3f));
            break;
        case T_EXRELAX:    // rates.f90:168-179: relaxation at the reactant's diffusion frequency
            if (act_exr) r = N.rstat[5 * k];
            break;
        case T_EXSOLID:    // rates.f90:182-197: alpha (vA + vB) / (site density 1.8e-8)
            if (act_exs) r = N.alpha[k] * N.rstat[5 * k];
            break;
        case T_GAR:        // rates.f90:316-332: grain-assisted recombination (Weingartner & Draine 2001)
            if (act_gar) {
                const double *g = N.gar + 7 * (k - gar0);
                r = (double)0.6f * N.alpha[k] * g[0] /
                    (1. + g[1] * pow(phi_new, g[2]) * (1. + g[3] * pow(T, g[4]) * pow(phi_new, -g[5] - g[6] * log(T))));
            }
            break;
        default:
            break;
        }
        rate[k] = r;
    }
    __syncthreads();
    // freeze-out specials (rates.f90:78-83)
    if (tid() == 0 && act_fr) {
        rate[N.r_efreeze] = rate[N.r_hfreeze];
        rate[N.r_h2freeze] = sticking_dev(0.87, 87.0, T) * rate[N.r_h2freeze];
        rate[N.r_hfreeze] = sticking_dev(1.0, 52.0, T) * rate[N.r_hfreeze];
    }
    __syncthreads();
    // phase 2: freeze-out switch-off where desorption dominates (rates.f90:104-107,125-128,148-151,221-224);
    // the thermal list is compared with its first nfp entries (non-conforming WHERE)
    for (int q = tid(); q < N.nfp; q += NT) {
        const int f = N.freeze_partners[q];
        double rf = rate[f];
        double yf = y[N.re[3 * f]];
        bool off = false;
        if (act_h2d) off |= (rf * yf < MIN_SURFACE_ABUND * rate[N.type_ranges[2 * T_DESOH2] + q]);
        if (act_cr) off |= (rf * yf * dens < MIN_SURFACE_ABUND * rate[N.type_ranges[2 * T_DESCR] + q]);
        if (act_uv) off |= (rf * yf * dens < MIN_SURFACE_ABUND * rate[N.type_ranges[2 * T_DEUVCR] + q]);
        if (act_th) off |= (rf * yf * dens < MIN_SURFACE_ABUND * rate[N.type_ranges[2 * T_THERM] + q]);
        if (off) rate[f] = 0.0;
    }
    __syncthreads();
    // phase 3: thermal zeroing on bare grains, H2 formation specials, temperature masks
    for (int k = tid(); k < N.nreac; k += NT) {
        double r = rate[k];
        if (N.rtype[k] == T_THERM && act_th && sm < MIN_SURFACE_ABUND) r = 0.0;
        if (k == N.r_h2form_ct) r = h2_form_efficiency_dev(Td, Td);
        if (k == N.r_h2form_er || k == N.r_h2form_erdes) r = 0.0;
        if (!N.extrapolate[k] && T < N.tmin[k]) r = 0.0;
        if (!N.extrapolate[k] && T > N.tmax[k]) r = 0.0;
        rate[k] = r;
    }
    __syncthreads();
    // photo overrides and the phi update (rates.f90:316-342)
    if (tid() == 0) {
        rate[N.r_h2_hv] = h2_photodiss_dev(sc[S_H2COL], radfield, sc[S_SCAT]);
        sc[S_PHI] = phi_new;
    }
    if (tid() == 64) rate[N.r_co_hv] = co_photodiss_dev(sc[S_H2COL], sc[S_COCOL], radfield, av);
    if (tid() == 128)
        rate[N.r_c_hv] = c_ionization_dev(N.alpha[N.r_c_hv], N.gama[N.r_c_hv], T, sc[S_CCOL], sc[S_H2COL], av,
                                          radfield);
    __syncthreads();
}

// value of one reaction's ODE term, tokens in the order MakeRates emits them.
// prog[k] packs {n_d, op0..op5, nfac} into two 16-byte words.
// Factor codes of one reaction from its 16-byte record (int16 x 8, density factors first,
// then the MakeRates tokens in order, -1 padded).
__device__ __forceinline__ void decode_codes(const int4 w, int c[8])
{
    const int wd[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int h = 0; h < 4; h++) {
        c[2 * h] = (int)(short)(wd[h] & 0xffff);
        c[2 * h + 1] = wd[h] >> 16;   // arithmetic shift keeps the sign of the high half
    }
}

// Value of one reaction's ODE term, factors applied in the order MakeRates emits them
// (rate * D^b * y... with /safeMantle, *bulkLayersReciprocal, *totalSwap/safeMantle).
__device__ __forceinline__ double term_from_codes(const int c[8], double r, const LDS double *y, double D, double blr,
                                                  double TS, double sm)
{
    double v = r;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int o = c[m];
        if (o == OP_NONE) continue;
        if (o == FAC_D) v = v * D;
        else if (o >= 0) v = v * y[o];
        else if (o == OP_BLR) v = v * blr;
        else if (o == OP_SWAP) v = v * TS / sm;
        else v = v / sm;
    }
    return v;
}

__device__ __forceinline__ double term_value(const DevNet &N, int k, const double *rate, const LDS double *y,
                                             double D, double blr, double TS, double sm)
{
    int c[8];
    decode_codes(reinterpret_cast<const int4 *>(N.facp)[k], c);
    return term_from_codes(c, rate[k], y, D, blr, TS, sm);
}

// RHS f(y) (chemistry.f90:294-352 + the generated GETYDOT, odes.f90).  Input L.y, output
// L.f; also leaves the pre-transfer surface total in sc[S_SPRE] for the Jacobian.
UCG_BIG __device__ void block_rhs(const DevNet &N, const Lds L, uint64_t *split = nullptr)
{
    const uint64_t c0 = wall_clock64();
    const int neq = N.neq;
    LDS double *y = L.y, *f = L.f, *flux = L.flux, *sc = L.sc;
    double *rate = L.rate;
    const LDS double *prm = L.prm;
    const double D = y[neq - 1];
    const double radfield = prm[P_RADFIELD];
    const double sm = fmax(1e-30, y[N.nsurface]);
    const double sbk = fmax(1e-30, y[N.nbulk]);
    const double blr = fmin(1.0, NUM_SITES_PER_GRAIN / (GAS_DUST_DENSITY_RATIO * sbk));
    double ts = 0.0;
    for (int q = tid(); q < N.nswap; q += NT) ts += term_value(N, N.swap_list[q], rate, y, D, blr, 0.0, sm);
    const double TS = block_sum(ts, L.red);
    const uint64_t c1 = wall_clock64();
    // H2 and CO photodissociation depend on the current column densities (chemistry.f90:311-319):
    // the last wave evaluates them (the seven CO row splines, the scattering term and the H2
    // rate in separate lanes) and their two fluxes while the other waves take every other flux.
    if (tid() >= NT - 64) {
        const int ln = tid() - (NT - 64);
        const double h2col = 0.5 * y[N.nh2] * D * (sc[S_CS] / 1.0);
        const double cocol = 0.5 * y[N.nco] * D * (sc[S_CS] / 1.0);
        double lognh2, lognco;
        co_log_columns(h2col, cocol, lognh2, lognco);
        double v = 0.0;
        if (ln < 7) v = co_row_spline(ln, lognh2);
        else if (ln == 7) v = co_scatter(h2col, cocol, sc[S_AV]);
        else if (ln == 8) v = h2_photodiss_dev(h2col, radfield, sc[S_SCAT]);
        double yy[7];
#pragma unroll
        for (int j = 0; j < 7; j++) yy[j] = __shfl(v, j);
        const double scat = __shfl(v, 7), kh2 = __shfl(v, 8);
        if (ln == 0) {
            rate[N.r_co_hv] = co_rate_from_parts(radfield, co_shielding(yy, lognco), scat);
            rate[N.r_h2_hv] = kh2;
            flux[N.r_co_hv] = term_value(N, N.r_co_hv, rate, y, D, blr, TS, sm);
            flux[N.r_h2_hv] = term_value(N, N.r_h2_hv, rate, y, D, blr, TS, sm);
        }
    } else {
        const GLB int *fp = glb(N.facp);
        const GLB double *rg = glb(rate);
        const int nr = N.nreac;
        const int w = tid(), nw = NT - 64;
        const int kh = N.r_h2_hv, kc = N.r_co_hv;
        for (int k0 = w; k0 < nr; k0 += 4 * nw) {   // four reactions per pass, records loaded up front
            int4 wr[4];
            double r[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int k = k0 + u * nw;
                if (k < nr) {
                    wr[u] = ld_int4(fp, k);
                    r[u] = rg[k];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int k = k0 + u * nw;
                if (k < nr && k != kh && k != kc) {
                    int c[8];
                    decode_codes(wr[u], c);
                    flux[k] = term_from_codes(c, r[u], y, D, blr, TS, sm);
                }
            }
        }
    }
    __syncthreads();
    const uint64_t c2 = wall_clock64();
    // production and loss sums in balanced chunks (partial sums in LDS after the fluxes),
    // then per species in chunk order
    LDS double *part = flux + N.gt_off;
    {
        const GLB int *gr = glb(N.gt_rec);
        const GLB int *lidx = glb(N.loss_idx), *pidx = glb(N.prod_idx);
        for (int t = tid(); t < N.ngt; t += NT) {
            const int4 rc = ld_int4(gr, t);
            const GLB int *src = rc.w ? lidx : pidx;
            double acc = 0.0;
#pragma unroll 8
            for (int q = rc.x; q < rc.y; q++) acc += flux[src[q]];
            part[t] = acc;
        }
    }
    __syncthreads();
    const uint64_t c3 = wall_clock64();
    double pre[EPT];
    double spre = 0.0;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const int i = tid() + e * NT;
        pre[e] = 0.0;
        if (i < N.nspec) {
            double prod = 0.0, loss = 0.0;
            const GLB int *spp = glb(N.gt_sp_ptr), *spt = glb(N.gt_sp_task);
            const int q0 = spp[2 * i], q1 = spp[2 * i + 1], q2 = spp[2 * i + 2];
            for (int q = q0; q < q1; q++) prod += part[spt[q]];
            for (int q = q1; q < q2; q++) loss += part[spt[q]];
            pre[e] = prod - loss;
            if (N.is_surface[i]) spre += pre[e];
        }
    }
    const double S = block_sum(spre, L.red);
    double tb = 0.0, tsu = 0.0;
    double out[EPT];
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const int i = tid() + e * NT;
        out[e] = pre[e];
        if (i < N.nspec) {
            // bulk<->surface transfer (odes.f90:4818-5153)
            if (N.is_surface[i]) {
                const int b = N.bulk_partner_of_surface[i];
                if (b >= 0) {
                    if (S < 0) out[e] = out[e] - S * fmin(1.0, sbk / sm) * y[b] / sbk;
                    else out[e] = out[e] - S * SURFACE_COVERAGE0 * y[i];
                }
                tsu += out[e];
            } else if (N.is_bulk[i]) {
                const int s = N.surface_partner_of_bulk[i];
                if (S < 0) out[e] = out[e] + S * fmin(1.0, sbk / sm) * y[i] / sbk;
                else out[e] = out[e] + S * SURFACE_COVERAGE0 * y[s];
                tb += out[e];
            }
        }
    }
    block_sum2(tb, tsu, L.red);
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const int i = tid() + e * NT;
        if (i < neq) {
            double v = out[e];
            if (i == N.nbulk) v = tb;
            if (i == N.nsurface) v = tsu;
            if (i == neq - 1) v = densdot_dev(prm, D);
            f[i] = v;
        }
    }
    if (tid() == 0) sc[S_SPRE] = S;
    __syncthreads();
    if (prm[P_ENFORCECHARGE] != 0.0 && N.nelec < N.nspec) {
        // enforceChargeConservation: ydot(e-) is the sum of the positive-ion rows (chemistry.f90:333-335)
        double e = 0.0;
        for (int q = tid(); q < N.nion; q += NT) e += f[N.ion_list[q]];
        e = block_sum(e, L.red);
        if (tid() == 0) f[N.nelec] = e;
        __syncthreads();
    }
    if (split) {   // totalSwap, photo rates + fluxes, chunked sums, species totals + transfer
        split[0] += c1 - c0;
        split[1] += c2 - c1;
        split[2] += c3 - c2;
        split[3] += wall_clock64() - c3;
    }
}

// ---------------------------------------------------------------- analytic Jacobian

__device__ __forceinline__ double fac_value(int c, const LDS double *y, double D, double blr, double ism, double tsm)
{
    if (c >= 0) return y[c];
    if (c == FAC_D) return D;
    if (c == OP_BLR) return blr;
    if (c == OP_DIVSM) return ism;
    return tsm;   // OP_SWAP
}

// dJ/dy for the factor's own variable
__device__ __forceinline__ double fac_deriv(int c, double dblr, double dism, double dtsm)
{
    if (c >= 0 || c == FAC_D) return 1.0;
    if (c == OP_BLR) return dblr;
    if (c == OP_DIVSM) return dism;
    return dtsm;
}

// Partial derivatives of every reaction term with respect to each of its factors:
// P[fac_off[k] + m] = rate_k * d(factor m) * prod_{m' != m} factor m'.  One 16-byte
// record per reaction holds its (up to 8) factor codes as int16, -1 padded.
__device__ void block_term_partials(const DevNet &N, const double *rate, const LDS double *y, LDS double *P, double D,
                                    double blr, double ism, double tsm, double dblr, double dism, double dtsm)
{
    const GLB int *fp = glb(N.facp);
    const GLB double *rg = glb(rate);
    for (int k = tid(); k < N.nreac; k += NT) {
        int c[8];
        decode_codes(ld_int4(fp, k), c);
        double v[8], pre[8];
        double acc = 1.0;
#pragma unroll
        for (int m = 0; m < 8; m++) {
            v[m] = (c[m] == OP_NONE) ? 1.0 : fac_value(c[m], y, D, blr, ism, tsm);
            pre[m] = acc;
            acc *= v[m];
        }
        const double r = rg[k];
        const int off = glb(N.fac_off)[k];
        double suf = 1.0;
#pragma unroll
        for (int m = 7; m >= 0; m--) {
            if (c[m] != OP_NONE) P[off + m] = r * fac_deriv(c[m], dblr, dism, dtsm) * pre[m] * suf;
            suf *= v[m];
        }
    }
}

// Value of structural nonzero n of dJ (generated part): sum of coefficient * partial.
__device__ __forceinline__ double jac_entry(const DevNet &N, const LDS double *P, int n)
{
    double acc = 0.0;
    const GLB int *jp = glb(N.jnz_ptr), *pk_ = glb(N.jc_pk);
    const int q1 = jp[n + 1];
#pragma unroll 4
    for (int q = jp[n]; q < q1; q++) {
        const int pk = pk_[q];
        acc += (double)((pk & 15) - 8) * P[pk >> 4];
    }
    return acc;
}

// Every generated entry of dJ, handed to store(n, value): the short contribution lists one per
// lane in the dealt order, the long ones split over JZ_GROUP lanes (strided partial sums, then a
// shuffle tree inside the group; all lanes of a group take the same trip count, so the shuffles
// read active lanes only).  The longest list (hundreds of terms) no longer serialises one lane.
template <typename Store>
__device__ __forceinline__ void jac_entries(const DevNet &N, const LDS double *P, Store store)
{
    const GLB int *js = glb(N.jz_short);
    for (int q = tid(); q < N.nj_short; q += NT) {
        const int n = js[q];
        store(n, jac_entry(N, P, n));
    }
    const GLB int *jl = glb(N.jz_long), *jp = glb(N.jnz_ptr), *pk_ = glb(N.jc_pk);
    const int grp = tid() / JZ_GROUP, ln = tid() % JZ_GROUP;
    for (int q = grp; q < N.nj_long; q += NT / JZ_GROUP) {
        const int n = jl[q];
        double acc = 0.0;
        const int q1 = jp[n + 1];
        for (int c = jp[n] + ln; c < q1; c += JZ_GROUP) {
            const int pk = pk_[c];
            acc += (double)((pk & 15) - 8) * P[pk >> 4];
        }
#pragma unroll
        for (int off = JZ_GROUP / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (ln == 0) store(n, acc);
    }
}

// Newton matrix M = I - gamma J(y) in column-major storage, J from the generated
// structure plus the structured couplings (photo rates, totalSwap, ice transfer and the
// BULK/SURFACE totals).  With scale != 0 the matrix is equilibrated by the error weights:
// M' = W M W^-1, W = diag(ewt) (held in L.g).
template <typename LUT>
UCG_BIG __device__ void block_newton_matrix(const DevNet &N, const Lds L, LUT *M, double gamma, bool convert = true,
                                    bool scaled = false)
{
    const int neq = N.neq, ld = N.ld;
    LDS double *y = L.y, *sc = L.sc, *g = L.g;
    double *rate = L.rate;
    LDS double *scr = L.flux + N.npart_pad;   // scratch [>= 2*neq] beside the partials
    const double D = y[neq - 1];
    const double ysur = y[N.nsurface], ybulk = y[N.nbulk];
    const double sm = fmax(1e-30, ysur), sbk = fmax(1e-30, ybulk);
    const double blr = fmin(1.0, NUM_SITES_PER_GRAIN / (GAS_DUST_DENSITY_RATIO * sbk));
    const double ism = 1.0 / sm;
    const double dblr = (blr < 1.0 && ybulk > 1e-30) ? -blr / sbk : 0.0;
    const double dism = (ysur > 1e-30) ? -ism * ism : 0.0;
    // totalSwap
    double ts = 0.0;
    for (int q = tid(); q < N.nswap; q += NT) ts += term_value(N, N.swap_list[q], rate, y, D, blr, 0.0, sm);
    const double TS = block_sum(ts, L.red);
    const double tsm = TS / sm;
    const double dtsm = (ysur > 1e-30) ? -TS * ism * ism : 0.0;
    // 1. zero J; partial derivatives of every term
    for (int idx = tid(); idx < neq * ld; idx += NT) M[idx] = (LUT)0;
    LDS double *P = L.flux;
    block_term_partials(N, rate, y, P, D, blr, ism, tsm, dblr, dism, dtsm);
    __syncthreads();
    // 2. generated sparse part
    jac_entries(N, P, [&](int n, double v) { M[(size_t)N.jnz_col[n] * ld + N.jnz_row[n]] = (LUT)v; });
    __syncthreads();
    // 3a. totalSwap dependence of SURFSWAP terms: J[i][b] += u_i * dTS/dy_b
    for (int i = tid(); i < neq; i += NT) {
        double u = 0.0;
        for (int q = N.sw_ptr[i]; q < N.sw_ptr[i + 1]; q++) {
            const int k = N.sw_k[q];
            const int *fc = N.fac + MAXFAC * k;
            double w = rate[k];
            for (int mm = 0; mm < N.nfac[k]; mm++) {
                const int c = fc[mm];
                w *= (c == OP_SWAP) ? ism : fac_value(c, y, D, blr, ism, tsm);
            }
            u += N.sw_c[q] * w;
        }
        scr[i] = u;
        scr[neq + i] = 0.0;
    }
    __syncthreads();
    if (tid() == 0) {
        for (int q = 0; q < N.nswap; q++) {
            const int k = N.swap_list[q];
            scr[neq + N.ops[MAXOPS * k]] += rate[k] * blr;
        }
        scr[neq + N.nbulk] = (blr < 1.0 && ybulk > 1e-30) ? -TS / sbk : 0.0;
    }
    __syncthreads();
    for (int idx = tid(); idx < neq * (N.nbulkl + 1); idx += NT) {
        const int i = idx % neq, q = idx / neq;
        const double u = scr[i];
        if (u == 0.0) continue;
        const int b = (q < N.nbulkl) ? N.bulk_list[q] : N.nbulk;
        const double gb = scr[neq + b];
        if (gb != 0.0) M[(size_t)b * ld + i] = (LUT)((double)M[(size_t)b * ld + i] + u * gb);
    }
    __syncthreads();
    // 3b. photo-rate dependence on y(H2), y(CO) and density through the column densities,
    //     and the free-fall density row
    {   // the six photo-rate evaluations run in parallel lanes of wave 0
        const double cs = sc[S_CS], G0 = L.prm[P_RADFIELD], av = sc[S_AV];
        const double h2col = 0.5 * y[N.nh2] * D * cs, cocol = 0.5 * y[N.nco] * D * cs;
        const double eps = 1e-6;
        const int t = tid();
        if (t < 6) {
            const double h = (t == 1 || t == 4) ? h2col * (1 + eps) : h2col;
            const double c = (t == 5) ? cocol * (1 + eps) : cocol;
            scr[t] = (t < 2) ? h2_photodiss_dev(h, G0, sc[S_SCAT]) : (t >= 3 ? co_photodiss_dev(h, c, G0, av) : 0.0);
        }
    }
    __syncthreads();
    if (tid() == 0) {
        const double cs = sc[S_CS];
        const double h2col = 0.5 * y[N.nh2] * D * cs, cocol = 0.5 * y[N.nco] * D * cs;
        const double eps = 1e-6;
        const double kh = scr[0];
        const double dkh = h2col > 0 ? (scr[1] - kh) / (h2col * eps) : 0.0;
        const double kc = scr[3];
        const double dkch = h2col > 0 ? (scr[4] - kc) / (h2col * eps) : 0.0;
        const double dkcc = cocol > 0 ? (scr[5] - kc) / (cocol * eps) : 0.0;
        for (int which = 0; which < 2; which++) {
            const int k = which ? N.r_co_hv : N.r_h2_hv;
            const int *fc = N.fac + MAXFAC * k;
            double pf = 1.0;
            for (int mm = 0; mm < N.nfac[k]; mm++) pf *= fac_value(fc[mm], y, D, blr, ism, tsm);
            double d_h2 = which ? dkch * 0.5 * D * cs : dkh * 0.5 * D * cs;
            double d_co = which ? dkcc * 0.5 * D * cs : 0.0;
            double d_d = which ? (dkch * 0.5 * y[N.nh2] * cs + dkcc * 0.5 * y[N.nco] * cs) : dkh * 0.5 * y[N.nh2] * cs;
            for (int q = N.aff_ptr[k]; q < N.aff_ptr[k + 1]; q++) {
                const int i = N.aff_sp[q];
                const double c = N.aff_cf[q] * pf;
                M[(size_t)N.nh2 * ld + i] = (LUT)((double)M[(size_t)N.nh2 * ld + i] + c * d_h2);
                if (d_co != 0.0) M[(size_t)N.nco * ld + i] = (LUT)((double)M[(size_t)N.nco * ld + i] + c * d_co);
                M[(size_t)(neq - 1) * ld + i] = (LUT)((double)M[(size_t)(neq - 1) * ld + i] + c * d_d);
            }
        }
        const double dn = D * 1e-7;
        M[(size_t)(neq - 1) * ld + neq - 1] =
            (LUT)((densdot_dev(L.prm, D + dn) - densdot_dev(L.prm, D)) / dn);
    }
    __syncthreads();
    // 4. ice transfer: rows of surface/bulk species get -/+ (t_s * grad S + S * grad t_s)
    const double S = sc[S_SPRE];
    for (int j = tid(); j < neq; j += NT) {
        double gsum = 0.0;
#pragma unroll 8
        for (int q = 0; q < N.nsurfl; q++) gsum += (double)M[(size_t)j * ld + N.surface_list[q]];
        scr[j] = gsum;
    }
    __syncthreads();
    const bool neg = S < 0;
    const bool ratio_lt1 = (sbk / sm) < 1.0;
    for (int idx = tid(); idx < neq * N.nsurfl; idx += NT) {
        const int q = idx % N.nsurfl, j = idx / N.nsurfl;
        const int s = N.surface_list[q], b = N.bulk_partner_of_surface[s];
        if (b < 0) continue;
        double t = neg ? (ratio_lt1 ? y[b] / sm : y[b] / sbk) : SURFACE_COVERAGE0 * y[s];
        double v = t * scr[j];
        // derivative of t_s
        if (neg) {
            if (j == b) v += S * (ratio_lt1 ? 1.0 / sm : 1.0 / sbk);
            if (ratio_lt1 && j == N.nsurface && ysur > 1e-30) v += S * (-y[b] / (sm * sm));
            if (!ratio_lt1 && j == N.nbulk && ybulk > 1e-30) v += S * (-y[b] / (sbk * sbk));
        } else if (j == s) {
            v += S * SURFACE_COVERAGE0;
        }
        if (v == 0.0) continue;
        M[(size_t)j * ld + s] = (LUT)((double)M[(size_t)j * ld + s] - v);
        M[(size_t)j * ld + b] = (LUT)((double)M[(size_t)j * ld + b] + v);
    }
    __syncthreads();
    // 5. BULK / SURFACE rows are the sums of their member rows
    for (int j = tid(); j < neq; j += NT) {
        double sb = 0.0, ss = 0.0;
#pragma unroll 8
        for (int q = 0; q < N.nbulkl; q++) sb += (double)M[(size_t)j * ld + N.bulk_list[q]];
#pragma unroll 8
        for (int q = 0; q < N.nsurfl; q++) ss += (double)M[(size_t)j * ld + N.surface_list[q]];
        M[(size_t)j * ld + N.nbulk] = (LUT)sb;
        M[(size_t)j * ld + N.nsurface] = (LUT)ss;
    }
    __syncthreads();
    // 5b. with enforceChargeConservation the electron row is the sum of the positive-ion rows
    if (L.prm[P_ENFORCECHARGE] != 0.0 && N.nelec < N.nspec) {
        for (int j = tid(); j < neq; j += NT) {
            double se = 0.0;
            for (int q = 0; q < N.nion; q++) se += (double)M[(size_t)j * ld + N.ion_list[q]];
            M[(size_t)j * ld + N.nelec] = (LUT)se;
        }
        __syncthreads();
    }
    if (!convert) return;
    // 6. M = I - gamma W J W^-1 (W = I unless the caller equilibrates)
    if (!scaled) {
        for (int idx = tid(); idx < neq * neq; idx += NT) {
            const int j = idx / neq, i = idx - j * neq;
            double v = -gamma * (double)M[(size_t)j * ld + i];
            if (i == j) v += 1.0;
            M[(size_t)j * ld + i] = (LUT)v;
        }
    } else {
        for (int idx = tid(); idx < neq * neq; idx += NT) {
            const int j = idx / neq, i = idx - j * neq;
            double v = -gamma * (double)M[(size_t)j * ld + i] * (g[i] / g[j]);
            if (i == j) v += 1.0;
            M[(size_t)j * ld + i] = (LUT)v;
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------- dense LU (no pivoting)

#ifndef UCG_LU_NB
#define UCG_LU_NB 16
#endif
constexpr int LU_NB = UCG_LU_NB;

// Right-looking blocked LU of the column-major n x n matrix M (unit lower L, upper U
// stored in place).  Per block column: the 16x16 diagonal block is factored by wave 0 in
// registers (lane r holds row r, pivot rows exchanged with wave shuffles), then
// L21 = A21 U11^-1 (one row per thread), U12 = L11^-1 A12 (one column per thread) and
// the trailing update A22 -= L21 U12 (threads own rows, four columns per pass).
// `stage` is LDS scratch of >= 2*16*16 doubles.  Returns nonzero (uniformly) on a zero
// or non-finite pivot.
template <typename LUT, typename MP>
UCG_BIG __device__ int block_lu(MP M, int n, int ld, LDS double *red, LDS double *stage)
{
    LDS double *U11 = stage;          // U11[p*16+q] = U(kb+p, kb+q)
    LDS double *L11 = stage + LU_NB * LU_NB;
    for (int kb = 0; kb < n; kb += LU_NB) {
        const int nb = min(LU_NB, n - kb);
        __syncthreads();
        // 1. diagonal block in registers of wave 0
        if (tid() < 64) {
            const int r = tid();
            double a[LU_NB];
#pragma unroll
            for (int q = 0; q < LU_NB; q++) a[q] = (r < nb && q < nb) ? (double)M[(size_t)(kb + q) * ld + kb + r] : 0.0;
            double bad = 0.0;
#pragma unroll
            for (int k = 0; k < LU_NB; k++) {
                if (k < nb) {
                    const double piv = __shfl(a[k], k);
                    if (!(piv == piv) || piv == 0.0 || isinf(piv)) bad = 1.0;
                    const double l = (r > k && r < nb) ? a[k] / piv : 0.0;
                    if (r > k && r < nb) a[k] = l;
#pragma unroll
                    for (int q = k + 1; q < LU_NB; q++) {
                        const double uq = __shfl(a[q], k);
                        if (r > k && r < nb) a[q] -= l * uq;
                    }
                }
            }
            if (r < nb) {
#pragma unroll
                for (int q = 0; q < LU_NB; q++) {
                    if (q < nb) {
                        M[(size_t)(kb + q) * ld + kb + r] = (LUT)a[q];
                        U11[r * LU_NB + q] = (q >= r) ? a[q] : 0.0;
                        L11[r * LU_NB + q] = (q < r) ? a[q] : 0.0;
                    }
                }
            }
            if (r == 0) red[8] = bad;
        }
        __syncthreads();
        if (red[8] != 0.0) return 1;
        const int c0 = kb + nb;
        if (c0 >= n) break;
        // 2. L21 = A21 U11^-1 (rows) and U12 = L11^-1 A12 (columns)
        for (int w = tid(); w < 2 * (n - c0); w += NT) {
            double a[LU_NB];
            if (w < n - c0) {
                const int i = c0 + w;
#pragma unroll
                for (int p = 0; p < LU_NB; p++) a[p] = (p < nb) ? (double)M[(size_t)(kb + p) * ld + i] : 0.0;
#pragma unroll
                for (int p = 0; p < LU_NB; p++) {
                    if (p < nb) {
                        double acc = a[p];
#pragma unroll
                        for (int q = 0; q < p; q++) acc -= a[q] * U11[q * LU_NB + p];
                        a[p] = acc / U11[p * LU_NB + p];
                    }
                }
#pragma unroll
                for (int p = 0; p < LU_NB; p++)
                    if (p < nb) M[(size_t)(kb + p) * ld + i] = (LUT)a[p];
            } else {
                const int c = c0 + (w - (n - c0));
#pragma unroll
                for (int r = 0; r < LU_NB; r++) a[r] = (r < nb) ? (double)M[(size_t)c * ld + kb + r] : 0.0;
#pragma unroll
                for (int r = 1; r < LU_NB; r++) {
                    if (r < nb) {
                        double acc = a[r];
#pragma unroll
                        for (int p = 0; p < r; p++) acc -= L11[r * LU_NB + p] * a[p];
                        a[r] = acc;
                    }
                }
#pragma unroll
                for (int r = 0; r < LU_NB; r++)
                    if (r < nb) M[(size_t)c * ld + kb + r] = (LUT)a[r];
            }
        }
        __syncthreads();
        // 3. A22 -= L21 U12: threads own rows, U12 columns read uniformly, 4 columns a pass
        for (int i = c0 + tid(); i < n; i += NT) {
            LUT l[LU_NB];
#pragma unroll
            for (int p = 0; p < LU_NB; p++) l[p] = (p < nb) ? M[(size_t)(kb + p) * ld + i] : (LUT)0;
            int c = c0;
            for (; c + 4 <= n; c += 4) {
                const LUT *u0 = M + (size_t)c * ld + kb;
                const LUT *u1 = u0 + ld;
                const LUT *u2 = u1 + ld;
                const LUT *u3 = u2 + ld;
                LUT x0 = 0, x1 = 0, x2 = 0, x3 = 0;
#pragma unroll
                for (int p = 0; p < LU_NB; p++) {
                    if (p < nb) {
                        x0 += l[p] * u0[p];
                        x1 += l[p] * u1[p];
                        x2 += l[p] * u2[p];
                        x3 += l[p] * u3[p];
                    }
                }
                M[(size_t)c * ld + i] -= x0;
                M[(size_t)(c + 1) * ld + i] -= x1;
                M[(size_t)(c + 2) * ld + i] -= x2;
                M[(size_t)(c + 3) * ld + i] -= x3;
            }
            for (; c < n; c++) {
                const LUT *u = M + (size_t)c * ld + kb;
                LUT x = 0;
#pragma unroll
                for (int p = 0; p < LU_NB; p++)
                    if (p < nb) x += l[p] * u[p];
                M[(size_t)c * ld + i] -= x;
            }
        }
    }
    __syncthreads();
    return 0;
}

// Broadcast of lane k's double to the whole wave without an LDS round trip.
__device__ __forceinline__ double wave_bcast(double v, int k)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
    return __hiloint2double(hi, lo);
}

// Blocked right-looking LU without pivoting for a dense block (the trailing block of the
// sparse plan, in LDS).  Panels of LU_NB columns: the diagonal block is factored in the
// registers of wave 0 (row r in lane r, pivots broadcast lane to lane), the panel's L21
// rows and U12 columns are solved one per thread against reciprocal pivots, and the
// rank-LU_NB trailing update runs over every thread in 4x4 output tiles.  ld should be
// odd for column accesses to spread over the LDS banks.
template <typename LUT, typename MP>
UCG_BIG __device__ int block_lu_tiled(MP M, int n, int ld, LDS double *red, LDS double *stage)
{
    LDS double *U11 = stage;                  // U11[p*16+q] = U(kb+p, kb+q), 1/U(p,p) on the diagonal
    LDS double *L11 = stage + LU_NB * LU_NB;
    for (int kb = 0; kb < n; kb += LU_NB) {
        const int nb = min(LU_NB, n - kb);
        __syncthreads();
        if (tid() < 64) {
            const int r = tid();
            double a[LU_NB];
#pragma unroll
            for (int q = 0; q < LU_NB; q++) a[q] = (r < nb && q < nb) ? (double)M[(size_t)(kb + q) * ld + kb + r] : 0.0;
            double bad = 0.0;
#pragma unroll
            for (int k = 0; k < LU_NB; k++) {
                if (k < nb) {
                    const double piv = wave_bcast(a[k], k);
                    if (!(piv == piv) || piv == 0.0 || isinf(piv)) bad = 1.0;
                    const double l = (r > k && r < nb) ? a[k] / piv : 0.0;
                    if (r > k && r < nb) a[k] = l;
#pragma unroll
                    for (int q = k + 1; q < LU_NB; q++) {
                        const double uq = wave_bcast(a[q], k);
                        if (r > k && r < nb) a[q] -= l * uq;
                    }
                }
            }
            if (r < nb) {
#pragma unroll
                for (int q = 0; q < LU_NB; q++) {
                    if (q < nb) {
                        M[(size_t)(kb + q) * ld + kb + r] = (LUT)a[q];
                        U11[r * LU_NB + q] = (q > r) ? a[q] : (q == r ? 1.0 / a[q] : 0.0);
                        L11[r * LU_NB + q] = (q < r) ? a[q] : 0.0;
                    }
                }
            }
            if (r == 0) red[8] = bad;
        }
        __syncthreads();
        if (red[8] != 0.0) return 1;
        const int c0 = kb + nb;
        if (c0 >= n) break;
        for (int w = tid(); w < 2 * (n - c0); w += NT) {   // L21 rows and U12 columns
            double a[LU_NB];
            if (w < n - c0) {
                const int i = c0 + w;
#pragma unroll
                for (int p = 0; p < LU_NB; p++) a[p] = (p < nb) ? (double)M[(size_t)(kb + p) * ld + i] : 0.0;
#pragma unroll
                for (int p = 0; p < LU_NB; p++) {
                    if (p < nb) {
                        double acc = a[p];
#pragma unroll
                        for (int q = 0; q < p; q++) acc -= a[q] * U11[q * LU_NB + p];
                        a[p] = acc * U11[p * LU_NB + p];
                    }
                }
#pragma unroll
                for (int p = 0; p < LU_NB; p++)
                    if (p < nb) M[(size_t)(kb + p) * ld + i] = (LUT)a[p];
            } else {
                const int c = c0 + (w - (n - c0));
#pragma unroll
                for (int r = 0; r < LU_NB; r++) a[r] = (r < nb) ? (double)M[(size_t)c * ld + kb + r] : 0.0;
#pragma unroll
                for (int r = 1; r < LU_NB; r++) {
                    if (r < nb) {
                        double acc = a[r];
#pragma unroll
                        for (int p = 0; p < r; p++) acc -= L11[r * LU_NB + p] * a[p];
                        a[r] = acc;
                    }
                }
#pragma unroll
                for (int r = 0; r < LU_NB; r++)
                    if (r < nb) M[(size_t)c * ld + kb + r] = (LUT)a[r];
            }
        }
        __syncthreads();
        const int m = n - c0;
        const int tr = (m + 3) >> 2;
        for (int tt = tid(); tt < tr * tr; tt += NT) {   // A22 -= L21 U12, 4x4 tiles
            const int ti = tt % tr, tj = tt / tr;
            const int i0 = c0 + 4 * ti, j0 = c0 + 4 * tj;
            LUT acc[4][4];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int c = 0; c < 4; c++) acc[r][c] = (LUT)0;
#pragma unroll
            for (int p = 0; p < LU_NB; p++) {
                if (p < nb) {
                    LUT a[4], b[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) a[r] = (i0 + r < n) ? M[(size_t)(kb + p) * ld + i0 + r] : (LUT)0;
#pragma unroll
                    for (int c = 0; c < 4; c++) b[c] = (j0 + c < n) ? M[(size_t)(j0 + c) * ld + kb + p] : (LUT)0;
#pragma unroll
                    for (int r = 0; r < 4; r++)
#pragma unroll
                        for (int c = 0; c < 4; c++) acc[r][c] += a[r] * b[c];
                }
            }
#pragma unroll
            for (int c = 0; c < 4; c++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (i0 + r < n && j0 + c < n) M[(size_t)(j0 + c) * ld + i0 + r] -= acc[r][c];
        }
    }
    __syncthreads();
    return 0;
}

// In-place inverse of the dense n x n block M (column-major, leading dimension ld) by
// blocked Gauss-Jordan elimination without pivoting, GJ_B = 8 pivots per sweep: the 8x8 pivot
// block is inverted by wave 0, one element per lane (eight elimination steps whose pivot row
// and column travel by shuffles), its row panel is multiplied by that inverse (one column per
// thread), every entry outside the pivot rows and columns takes the rank-8 update in 4x4
// register tiles, then the pivot column panel becomes -A_IK inv(A_KK).  Eight pivots per sweep
// halve the sweeps and barriers of a four-pivot sweep and load each 4x4 tile once per eight
// rank-one terms.  The pivots are those of the unpivoted LU and are checked the same way; it
// is used for the trailing block of the sparse plan so that each Newton solve applies the
// block as one matrix-vector product instead of two dependent triangular sweeps.  stage: >= 64
// doubles of LDS.  Returns nonzero (uniformly) on a zero or non-finite pivot.
constexpr int GJ_B = 8;
template <typename LUT, typename MP>
__device__ __attribute__((always_inline)) int block_gj_inverse(MP M, int n, int ld, LDS double *red, LDS double *stage)
{
    using AT = LUT;          // arithmetic in the storage precision
    constexpr int B = GJ_B;
    LDS AT *P = (LDS AT *)stage;   // P[l*B+m] = inv(A_KK)(l, m)
    for (int kb = 0; kb < n; kb += B) {
        const int nb = min(B, n - kb);
        __syncthreads();
        if (tid() < 64) {   // wave 0: lane r*B + c holds element (r, c) of the pivot block
            const int r = tid() / B, c = tid() % B;
            AT a = (r < nb && c < nb) ? (AT)M[(size_t)(kb + c) * ld + kb + r] : (r == c ? (AT)1 : (AT)0);
            bool bad = false;
#pragma unroll
            for (int k = 0; k < B; k++) {
                const AT piv = __shfl(a, k * B + k);
                const AT rowk = __shfl(a, k * B + c);   // a(k, c)
                const AT colk = __shfl(a, r * B + k);   // a(r, k)
                if (!(piv == piv) || piv == (AT)0 || isinf(piv)) bad = true;
                const AT q = (AT)1 / piv;
                if (r == k) a = (c == k) ? q : a * q;
                else a = (c == k) ? -colk * q : a - colk * (rowk * q);
            }
            P[r * B + c] = a;
            if (r < nb && c < nb) M[(size_t)(kb + c) * ld + kb + r] = (LUT)a;
            if (tid() == 0) red[8] = bad ? 1.0 : 0.0;
        }
        __syncthreads();
        if (red[8] != 0.0) return 1;
        for (int j = tid(); j < n; j += NT) {   // row panel: A_KJ <- inv(A_KK) A_KJ
            if (j >= kb && j < kb + nb) continue;
            AT v[B];
#pragma unroll
            for (int m = 0; m < B; m++) v[m] = m < nb ? (AT)M[(size_t)j * ld + kb + m] : (AT)0;
#pragma unroll
            for (int l = 0; l < B; l++) {
                AT acc = 0;
#pragma unroll
                for (int m = 0; m < B; m++) acc += P[l * B + m] * v[m];
                if (l < nb) M[(size_t)j * ld + kb + l] = (LUT)acc;
            }
        }
        __syncthreads();
        const int nt = (n + 3) >> 2;
        for (int tt = tid(); tt < nt * nt; tt += NT) {   // A_IJ -= A_IK A_KJ outside the pivot block
            const int ti = tt % nt, tj = tt / nt;
            const int i0 = 4 * ti, j0 = 4 * tj;
            if ((i0 < kb + nb && i0 + 4 > kb) || (j0 < kb + nb && j0 + 4 > kb)) continue;
            AT acc[4][4];
#pragma unroll
            for (int c = 0; c < 4; c++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    acc[r][c] = (i0 + r < n && j0 + c < n) ? (AT)M[(size_t)(j0 + c) * ld + i0 + r] : (AT)0;
#pragma unroll
            for (int m = 0; m < B; m++) {
                if (m < nb) {
                    AT a[4], b[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) a[r] = (i0 + r < n) ? (AT)M[(size_t)(kb + m) * ld + i0 + r] : (AT)0;
#pragma unroll
                    for (int c = 0; c < 4; c++) b[c] = (j0 + c < n) ? (AT)M[(size_t)(j0 + c) * ld + kb + m] : (AT)0;
#pragma unroll
                    for (int r = 0; r < 4; r++)
#pragma unroll
                        for (int c = 0; c < 4; c++) acc[r][c] -= a[r] * b[c];
                }
            }
#pragma unroll
            for (int c = 0; c < 4; c++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (i0 + r < n && j0 + c < n) M[(size_t)(j0 + c) * ld + i0 + r] = (LUT)acc[r][c];
        }
        __syncthreads();
        for (int i = tid(); i < n; i += NT) {   // column panel: A_IK <- -A_IK inv(A_KK)
            if (i >= kb && i < kb + nb) continue;
            AT c[B];
#pragma unroll
            for (int l = 0; l < B; l++) c[l] = l < nb ? (AT)M[(size_t)(kb + l) * ld + i] : (AT)0;
#pragma unroll
            for (int m = 0; m < B; m++) {
                AT acc = 0;
#pragma unroll
                for (int l = 0; l < B; l++) acc += c[l] * P[l * B + m];
                if (m < nb) M[(size_t)(kb + m) * ld + i] = (LUT)(-acc);
            }
        }
    }
    __syncthreads();
    return 0;
}

// x <- Ainv x for the n x n inverse Ainv (column-major, leading dimension n) and the LDS
// vector x: two threads per row (column halves) when 2n fits the workgroup, the second
// half's partial passed through `scratch` (n doubles of LDS); x is overwritten only after
// every thread has read it.
template <typename TP>
__device__ __forceinline__ void tail_matvec(TP Ainv, int n, LDS double *x, LDS double *scratch)
{
    const int t = tid();
    if (2 * n <= NT) {
        const int half = (n + 1) >> 1;
        double acc = 0.0;
        int r = -1;
        if (t < 2 * n) {
            r = t % n;
            const int h = t / n;
            const int c1 = h ? n : half;
            int c = h ? half : 0;
            double a1 = 0.0, a2 = 0.0, a3 = 0.0;
            for (; c + 4 <= c1; c += 4) {
                acc += (double)Ainv[(size_t)c * n + r] * x[c];
                a1 += (double)Ainv[(size_t)(c + 1) * n + r] * x[c + 1];
                a2 += (double)Ainv[(size_t)(c + 2) * n + r] * x[c + 2];
                a3 += (double)Ainv[(size_t)(c + 3) * n + r] * x[c + 3];
            }
            for (; c < c1; c++) acc += (double)Ainv[(size_t)c * n + r] * x[c];
            acc = (acc + a1) + (a2 + a3);
            if (h) scratch[r] = acc;
        }
        __syncthreads();
        if (t < n) x[r] = acc + scratch[r];
    } else {
        for (int r = t; r < n; r += NT) {
            double a0 = 0.0, a1 = 0.0;
            int c = 0;
            for (; c + 2 <= n; c += 2) {
                a0 += (double)Ainv[(size_t)c * n + r] * x[c];
                a1 += (double)Ainv[(size_t)(c + 1) * n + r] * x[c + 1];
            }
            if (c < n) a0 += (double)Ainv[(size_t)c * n + r] * x[c];
            scratch[r] = a0 + a1;
        }
        __syncthreads();
        for (int r = t; r < n; r += NT) x[r] = scratch[r];
    }
    __syncthreads();
}

#ifndef UCG_SB
#define UCG_SB 32
#endif
constexpr int SB = UCG_SB;   // triangular-solve block: wave 0 holds the diagonal block rows

// Solve (LU) x = b in place on the LDS vector b (FP64 accumulation).
template <typename LUT, typename MP>
UCG_BIG __device__ void block_lu_solve(MP M, int n, int ld, LDS double *b)
{
    const int lane = tid() & 63, wave = tid() >> 6;
    // forward, unit lower
    for (int kb = 0; kb < n; kb += SB) {
        const int nb = min(SB, n - kb);
        if (wave == 0) {
            double row[SB];
#pragma unroll
            for (int c = 0; c < SB; c++)
                row[c] = (lane < nb && c < lane) ? (double)M[(size_t)(kb + c) * ld + kb + lane] : 0.0;
            double x = (lane < nb) ? b[kb + lane] : 0.0;
#pragma unroll
            for (int c = 0; c < SB - 1; c++) {
                const double xc = __shfl(x, c);
                if (c < lane) x -= row[c] * xc;
            }
            if (lane < nb) b[kb + lane] = x;
        }
        __syncthreads();
        for (int i = kb + nb + tid(); i < n; i += NT) {
            double acc = b[i];
#pragma unroll 8
            for (int p = 0; p < nb; p++) acc -= (double)M[(size_t)(kb + p) * ld + i] * b[kb + p];
            b[i] = acc;
        }
        __syncthreads();
    }
    // backward, upper
    const int nblk = (n + SB - 1) / SB;
    for (int blk = nblk - 1; blk >= 0; blk--) {
        const int kb = blk * SB;
        const int nb = min(SB, n - kb);
        if (wave == 0) {
            double row[SB];
#pragma unroll
            for (int c = 0; c < SB; c++)
                row[c] = (lane < nb && c < nb && c >= lane) ? (double)M[(size_t)(kb + c) * ld + kb + lane] : 1.0;
            double x = (lane < nb) ? b[kb + lane] : 0.0;
#pragma unroll
            for (int c = SB - 1; c >= 0; c--) {
                if (c < nb && lane == c) x = x / row[c];
                const double xc = __shfl(x, c);
                if (c < nb && lane < c) x -= row[c] * xc;
            }
            if (lane < nb) b[kb + lane] = x;
        }
        __syncthreads();
        for (int i = tid(); i < kb; i += NT) {
            double acc = b[i];
#pragma unroll 8
            for (int p = 0; p < nb; p++) acc -= (double)M[(size_t)(kb + p) * ld + i] * b[kb + p];
            b[i] = acc;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- sparse Newton system
// J-form values (gamma = 1, no identity) of the augmented Newton matrix of the build-time
// plan (newton_plan.py): the sparse part -J0 of the generated reaction terms, photo
// columns and transfer S-terms; two bordered unknowns carrying the rank-one totalSwap and
// transfer couplings (columns u1 and w2, rows -g1 and -gsum0, and c1); the BULK and
// SURFACE rows are replaced by constant rows when the Newton matrix is formed
// (sparse_form_newton).  Values are written in plan slot order; fill slots start at zero.
UCG_BIG __device__ void sparse_jacobian(const DevNet &N, const Lds L, double *V)
{
    const int neq = N.neq;
    LDS double *y = L.y, *sc = L.sc;
    double *rate = L.rate;
    LDS double *scr = L.flux + N.npart_pad;   // scratch [>= 2*neq] beside the partials
    const double D = y[neq - 1];
    const double ysur = y[N.nsurface], ybulk = y[N.nbulk];
    const double sm = fmax(1e-30, ysur), sbk = fmax(1e-30, ybulk);
    const double blr = fmin(1.0, NUM_SITES_PER_GRAIN / (GAS_DUST_DENSITY_RATIO * sbk));
    const double ism = 1.0 / sm;
    const double dblr = (blr < 1.0 && ybulk > 1e-30) ? -blr / sbk : 0.0;
    const double dism = (ysur > 1e-30) ? -ism * ism : 0.0;
    double ts = 0.0;
    for (int q = tid(); q < N.nswap; q += NT) ts += term_value(N, N.swap_list[q], rate, y, D, blr, 0.0, sm);
    const double TS = block_sum(ts, L.red);
    const double tsm = TS / sm;
    const double dtsm = (ysur > 1e-30) ? -TS * ism * ism : 0.0;
    {
        GLB double *Vw = glbw(V);
        for (int q = tid(); q < N.nw_nval; q += NT) Vw[q] = 0;
    }
    LDS double *P = L.flux;
    block_term_partials(N, rate, y, P, D, blr, ism, tsm, dblr, dism, dtsm);
    __syncthreads();
    // 1. generated entries of J0, stored as -dJ; u1 (totalSwap column factor)
    {
        GLB double *Vw = glbw(V);
        const GLB int *slot = glb(N.nw_jnz_slot);
        jac_entries(N, P, [&](int n, double v) { Vw[slot[n]] = -v; });
    }
    for (int i = tid(); i < neq; i += NT) {
        double u = 0.0;
        for (int q = N.sw_ptr[i]; q < N.sw_ptr[i + 1]; q++) {
            const int k = N.sw_k[q];
            const int *fc = N.fac + MAXFAC * k;
            double w = rate[k];
            for (int mm = 0; mm < N.nfac[k]; mm++) {
                const int c = fc[mm];
                w *= (c == OP_SWAP) ? ism : fac_value(c, y, D, blr, ism, tsm);
            }
            u += N.sw_c[q] * w;
        }
        scr[i] = u;
        scr[neq + i] = 0.0;
    }
    __syncthreads();
    // g1 = d totalSwap / dy: one BULKSWAP per bulk species, so the targets are distinct
    // (checked when the network is loaded) and the sums need no ordering
    for (int q = tid(); q < N.nswap; q += NT) {
        const int k = glb(N.swap_list)[q];
        scr[neq + glb(N.ops)[MAXOPS * k]] += glb(rate)[k] * blr;
    }
    __syncthreads();
    if (tid() == 0) scr[neq + N.nbulk] = (blr < 1.0 && ybulk > 1e-30) ? -TS / sbk : 0.0;
    __syncthreads();
    // 2. border rows and columns, constant sum rows
    for (int j = tid(); j < neq; j += NT) {
        const int r2 = N.nw_s2r[j];
        if (r2 >= 0) {
            double acc = 0.0;
            for (int q = N.nw_gs_ptr[j]; q < N.nw_gs_ptr[j + 1]; q++) acc += (double)V[N.nw_gs_slot[q]];
            V[r2] = acc;
        }
        const int r1 = N.nw_s1r[j];
        if (r1 >= 0) V[r1] = -scr[neq + j];
        const int cc = N.nw_s1c[j];
        if (cc >= 0) V[cc] = scr[j];
    }
    double cs = 0.0;
    for (int q = tid(); q < N.nsurfl; q += NT) cs += scr[N.surface_list[q]];
    const double c1 = block_sum(cs, L.red);
    if (tid() == 0) {
        V[N.nw_misc[2]] = c1;
    }
    __syncthreads();
    // 3. identity; transfer column w2 (-t_s on the surface row, +t_s on its bulk partner)
    const double S = sc[S_SPRE];
    const bool neg = S < 0;
    const bool ratio_lt1 = (sbk / sm) < 1.0;
    for (int q = tid(); q < N.nsurfl; q += NT) {
        const int s_ = N.surface_list[q], b = N.bulk_partner_of_surface[s_];
        const double t = neg ? (ratio_lt1 ? y[b] / sm : y[b] / sbk) : SURFACE_COVERAGE0 * y[s_];
        V[N.nw_s2c[s_]] = (-t);
        V[N.nw_s2c[b]] = t;
    }
    __syncthreads();
    // 4. photo-rate columns (H2, CO, density) and the free-fall density row; transfer S-terms
    {
        const double cs_ = sc[S_CS], G0 = L.prm[P_RADFIELD], av = sc[S_AV];
        const double h2col = 0.5 * y[N.nh2] * D * cs_, cocol = 0.5 * y[N.nco] * D * cs_;
        const double eps = 1e-6;
        const int t = tid();
        if (t < 6) {
            const double h = (t == 1 || t == 4) ? h2col * (1 + eps) : h2col;
            const double c = (t == 5) ? cocol * (1 + eps) : cocol;
            scr[t] = (t < 2) ? h2_photodiss_dev(h, G0, sc[S_SCAT]) : (t >= 3 ? co_photodiss_dev(h, c, G0, av) : 0.0);
        }
    }
    __syncthreads();
    if (tid() == 0) {
        const double cs_ = sc[S_CS];
        const double h2col = 0.5 * y[N.nh2] * D * cs_, cocol = 0.5 * y[N.nco] * D * cs_;
        const double eps = 1e-6;
        const double kh = scr[0];
        const double dkh = h2col > 0 ? (scr[1] - kh) / (h2col * eps) : 0.0;
        const double kc = scr[3];
        const double dkch = h2col > 0 ? (scr[4] - kc) / (h2col * eps) : 0.0;
        const double dkcc = cocol > 0 ? (scr[5] - kc) / (cocol * eps) : 0.0;
        int m3 = 0;
        for (int which = 0; which < 2; which++) {
            const int k = which ? N.r_co_hv : N.r_h2_hv;
            const int *fc = N.fac + MAXFAC * k;
            double pf = 1.0;
            for (int mm = 0; mm < N.nfac[k]; mm++) pf *= fac_value(fc[mm], y, D, blr, ism, tsm);
            const double d_h2 = which ? dkch * 0.5 * D * cs_ : dkh * 0.5 * D * cs_;
            const double d_co = which ? dkcc * 0.5 * D * cs_ : 0.0;
            const double d_d = which ? (dkch * 0.5 * y[N.nh2] * cs_ + dkcc * 0.5 * y[N.nco] * cs_)
                                     : dkh * 0.5 * y[N.nh2] * cs_;
            for (int q = N.aff_ptr[k]; q < N.aff_ptr[k + 1]; q++, m3 += 3) {
                const double c = N.aff_cf[q] * pf;
                const int *ps = N.nw_ph_slot + m3;
                V[ps[0]] = ((double)V[ps[0]] - c * d_h2);
                V[ps[1]] = ((double)V[ps[1]] - c * d_co);
                V[ps[2]] = ((double)V[ps[2]] - c * d_d);
            }
        }
        const double dn = D * 1e-7;
        const int dd = N.nw_diag_slot[neq - 1];
        V[dd] = ((double)V[dd] - (densdot_dev(L.prm, D + dn) - densdot_dev(L.prm, D)) / dn);
    }
    for (int q = tid(); q < N.nsurfl; q += NT) {
        const int s_ = N.surface_list[q], b = N.bulk_partner_of_surface[s_];
        double d[4] = {0.0, 0.0, 0.0, 0.0};   // S * dt_s/dy over columns (s, b, SURFACE, BULK)
        if (neg) {
            d[1] = S * (ratio_lt1 ? 1.0 / sm : 1.0 / sbk);
            if (ratio_lt1 && ysur > 1e-30) d[2] = S * (-y[b] / (sm * sm));
            if (!ratio_lt1 && ybulk > 1e-30) d[3] = S * (-y[b] / (sbk * sbk));
        } else {
            d[0] = S * SURFACE_COVERAGE0;
        }
        const int *ts_ = N.nw_tr_slot + 8 * q;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            if (d[m] != 0.0) {   // J[s][c] -= d, J[b][c] += d, stored as -dJ
                V[ts_[m]] = ((double)V[ts_[m]] + d[m]);
                V[ts_[4 + m]] = ((double)V[ts_[4 + m]] - d[m]);
            }
        }
    }
    __syncthreads();
}

// M = I - gamma J in plan slots from the saved J-form values (sparse_jacobian): entries
// that carry gamma are scaled, the rank-one column factors and c1 are not; then the
// identity, the border diagonals and the replaced BULK/SURFACE rows.
template <typename LUT>
UCG_BIG __device__ void sparse_form_newton(const DevNet &N, const double *J, LUT *V, double gamma, bool enforce)
{
    for (int q = tid(); q < N.nw_nval; q += NT) V[q] = (LUT)(N.nw_unscaled[q] ? J[q] : gamma * J[q]);
    __syncthreads();
    for (int i = tid(); i < N.neq; i += NT) {
        const int d = N.nw_diag_slot[i];
        if (d >= 0) V[d] = (LUT)(gamma * J[d] + 1.0);
    }
    for (int q = tid(); q < N.nbulkl; q += NT) V[N.nw_nbrow[q]] = (LUT)(-1.0);
    for (int q = tid(); q < N.nsurfl; q += NT) V[N.nw_nsrow[q]] = (LUT)(-1.0);
    if (tid() == 0) {
        V[N.nw_misc[0]] = (LUT)(-1.0);
        V[N.nw_misc[1]] = (LUT)(-1.0);
        V[N.nw_misc[3]] = (LUT)1.0;
        V[N.nw_misc[4]] = (LUT)1.0;
    }
    __syncthreads();
    if (enforce && N.nelec < N.nspec) {   // charge row: row(e) - sum_ion row(ion) = e_e - sum_ion e_ion
        for (int q = tid(); q < N.nw_nerow; q += NT) V[N.nw_erow[q]] = (LUT)0.0;
        __syncthreads();
        for (int q = tid(); q < N.nion; q += NT) V[N.nw_eion[q]] = (LUT)(-1.0);
        if (tid() == 0) V[N.nw_diag_slot[N.nelec]] = (LUT)1.0;
        __syncthreads();
    }
}

// Crout elimination of the sparse pivots in plan phases (one thread per target entry,
// dependencies only on earlier phases), then the trailing block (its Schur complement):
// inverted in place by blocked Gauss-Jordan when the plan is staged, LU-factored when it
// is not.  The sparse factors stay in the workspace in their precision (LUT); a staged
// plan's tail, with the Schur-update sources, is worked on in LDS (region) in the same
// precision (the FP32 form does its whole factorisation in FP32: the Newton iteration's FP64
// residuals absorb the difference) and stored back.  stage: 512 doubles of LDS.
template <typename LUT, typename WP>
__device__ __forceinline__ void crout_phases(const DevNet &N, WP W, LDS double *red, int nph)
{
    const GLB int *ph = glb(N.nw_ph_ptr), *tgs = glb(N.nw_tg_slot), *tgp = glb(N.nw_tg_ptr), *tgd = glb(N.nw_tg_div);
    const GLB int *pr = glb(N.nw_pr);
    for (int p = 0; p < nph; p++) {
        const int t1 = ph[p + 1];
        for (int t = ph[p] + tid(); t < t1; t += NT) {
            const int tg = tgs[t];
            double a0 = (double)W[tg], a1 = 0.0, a2 = 0.0, a3 = 0.0;
            const int q1 = tgp[t + 1];
            int q = tgp[t];
            for (; q + 4 <= q1; q += 4) {
                const int2 p0 = ld_int2(pr, q), p1 = ld_int2(pr, q + 1), p2 = ld_int2(pr, q + 2), p3 = ld_int2(pr, q + 3);
                a0 -= (double)W[p0.x] * (double)W[p0.y];
                a1 -= (double)W[p1.x] * (double)W[p1.y];
                a2 -= (double)W[p2.x] * (double)W[p2.y];
                a3 -= (double)W[p3.x] * (double)W[p3.y];
            }
            for (; q < q1; q++) {
                const int2 p0 = ld_int2(pr, q);
                a0 -= (double)W[p0.x] * (double)W[p0.y];
            }
            double acc = (a0 + a1) + (a2 + a3);
            const int dv = tgd[t];
            if (dv >= 0) {
                const double piv = (double)W[dv];
                if (!(piv == piv) || piv == 0.0 || isinf(piv)) red[9] = 1.0;
                acc = acc / piv;
            }
            W[tg] = (LUT)acc;
        }
        __syncthreads();
    }
    for (int r = tid(); r < N.nw_ns; r += NT) {
        const double piv = (double)W[N.nw_sp_dpos[r]];
        if (!(piv == piv) || piv == 0.0 || isinf(piv)) red[9] = 1.0;
    }
    __syncthreads();
}

template <typename LUT>
UCG_BIG __device__ int sparse_lu(const DevNet &N, LUT *V, LDS double *region, LDS double *red, LDS double *stage,
                                 uint64_t *split = nullptr)
{
    const int nval = N.nw_nval, nsp = N.nw_nsp, T = N.nw_T;
    if (tid() == 0) red[9] = 0.0;
    int r = 0;
    const uint64_t c0 = wall_clock64();
    uint64_t c1;
    if (!N.nw_stage) {                    // factors too large to stage: everything in the workspace
        __syncthreads();
        crout_phases<LUT>(N, glbw(V), red, N.nw_nph);
        c1 = wall_clock64();
        if (red[9] != 0.0) return 1;
        if (T > 0) r = block_lu_tiled<LUT>(V + nsp, T, T, red, stage);
    } else {
        __syncthreads();
        const bool lp = UCG_LDS_SCHUR && N.nlp_src > 0;   // the last phase (Schur update of the tail) in LDS
        if (lp) {
            // the phases before it touch only the sparse part's nsp values (a staged region
            // holds them): eliminate on an LDS copy, so that every dependent access of the
            // phase chain is an LDS access, then store the factors back for the solve
            LDS LUT *Ws = (LDS LUT *)region;
            const GLB LUT *Vg = glb(V);
            for (int q = tid(); q < nsp; q += NT) Ws[q] = Vg[q];
            __syncthreads();
            crout_phases<LUT>(N, Ws, red, N.nw_nph - 1);
            if (red[9] != 0.0) return 1;
            for (int q = tid(); q < nsp; q += NT) V[q] = Ws[q];
            __syncthreads();
        } else {
            crout_phases<LUT>(N, glbw(V), red, N.nw_nph);
            if (red[9] != 0.0) return 1;
        }
        LDS LUT *tl = (LDS LUT *)region;   // the tail, worked on in the factor precision
        for (int q = tid(); q < T * T; q += NT) tl[q] = V[nsp + q];
        if (lp) {
            LDS LUT *src = tl + T * T;
            const GLB LUT *Vg = glb(V);
            const GLB int *lsrc = glb(N.lp_src), *tgs = glb(N.nw_tg_slot), *tgp = glb(N.nw_tg_ptr);
            for (int k = tid(); k < N.nlp_src; k += NT) src[k] = Vg[lsrc[k]];
            __syncthreads();
            const int p = N.nw_nph - 1, t0 = N.nw_ph_ptr[p], t1 = N.nw_ph_ptr[p + 1], qb = N.nw_tg_ptr[t0];
            const GLB int *pr = glb(N.lp_pr);
            for (int t = t0 + tid(); t < t1; t += NT) {
                const int e = tgs[t] - nsp;
                LUT a0 = tl[e], a1 = 0, a2 = 0, a3 = 0;
                const int q1 = tgp[t + 1] - qb;
                int q = tgp[t] - qb;
                for (; q + 4 <= q1; q += 4) {
                    const int2 p0 = ld_int2(pr, q), p1 = ld_int2(pr, q + 1), p2 = ld_int2(pr, q + 2), p3 = ld_int2(pr, q + 3);
                    a0 -= src[p0.x] * src[p0.y];
                    a1 -= src[p1.x] * src[p1.y];
                    a2 -= src[p2.x] * src[p2.y];
                    a3 -= src[p3.x] * src[p3.y];
                }
                for (; q < q1; q++) {
                    const int2 p0 = ld_int2(pr, q);
                    a0 -= src[p0.x] * src[p0.y];
                }
                tl[e] = (a0 + a1) + (a2 + a3);
            }
        }
        c1 = wall_clock64();
        __syncthreads();
        if (T > 0) r = block_gj_inverse<LUT>(tl, T, T, red, stage);
        for (int q = tid(); q < T * T; q += NT) V[nsp + q] = tl[q];
    }
    __syncthreads();
    if (split) {
        split[0] += c1 - c0;
        split[1] += wall_clock64() - c1;
    }
    return r;
}

// Lanes per row for a level of R independent rows: the largest power of two with R groups
// fitting the workgroup, at most a wave (groups never straddle a wave).
__device__ __forceinline__ int row_group(int R)
{
    int G = 64;
    while (G > 1 && G * R > NT) G >>= 1;
    return G;
}

// Partial dot product of one row's entries e0, e0+G, ... < e1 with the work vector.
template <typename VP, typename IP>
__device__ __forceinline__ double row_dot(VP V, IP cols, const LDS double *w, int e0, int e1, int G)
{
    double a0 = 0.0, a1 = 0.0;
    int e = e0;
    for (; e + G < e1; e += 2 * G) {
        a0 += (double)V[e] * w[cols[e]];
        a1 += (double)V[e + G] * w[cols[e + G]];
    }
    if (e < e1) a0 += (double)V[e] * w[cols[e]];
    return a0 + a1;
}

// Sum over an aligned group of G lanes (G a power of two <= 64); every lane of the wave
// takes part, each lane ends with its group's sum.
__device__ __forceinline__ double group_sum(double v, int G)
{
    for (int off = G >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Level-scheduled substitutions of the static plan on the permuted work vector w (LDS):
// forward over the sparse rows by level, the tail rows' L21 part, the dense tail, then
// backward by level.  V, the column indices and the row records are LDS or workspace
// pointers (staged or unstaged plan); b (length neq, LDS) is overwritten by x.
template <typename LUT, typename VP, typename IP, typename TP>
__device__ __forceinline__ void solve_program(const DevNet &N, VP V, IP cols, IP recs, IP inv, IP lev, TP tail,
                                              LDS double *w, LDS double *b, double sb, double ss, uint64_t *split,
                                              uint64_t c0, uint64_t cs)
{
    const int neq = N.neq, NN = N.nw_N, ns = N.nw_ns, T = N.nw_T;
    for (int i = tid(); i < NN; i += NT) {
        double v = 0.0;
        if (i < neq) {
            v = b[i];
            if (i == N.nbulk) v -= sb;
            else if (i == N.nsurface) v -= ss;
        }
        w[inv[i]] = v;
    }
    __syncthreads();
    const uint64_t cf0 = wall_clock64();
    for (int l = 0; l < N.nw_nlev; l++) {   // forward, sparse rows by level
        const int q0 = lev[l];
        const int R = lev[l + 1] - q0;
        const int G = row_group(R);
        for (int base = 0; base < R; base += NT / G) {
            const int g = base + tid() / G, ln = tid() % G;
            double acc = 0.0;
            int r = -1;
            if (g < R) {
                const int q = q0 + g;
                r = recs[4 * q];
                acc = row_dot(V, cols, w, recs[4 * q + 1] + ln, recs[4 * q + 2], G);
            }
            acc = group_sum(acc, G);
            if (r >= 0 && ln == 0) w[r] -= acc;
        }
        __syncthreads();
    }
    {   // tail rows: L21
        const int R = NN - ns, G = row_group(R);
        for (int base = 0; base < R; base += NT / G) {
            const int g = base + tid() / G, ln = tid() % G;
            double acc = 0.0;
            int r = -1;
            if (g < R) {
                const int q = ns + g;
                r = recs[4 * q];
                acc = row_dot(V, cols, w, recs[4 * q + 1] + ln, recs[4 * q + 3], G);
            }
            acc = group_sum(acc, G);
            if (r >= 0 && ln == 0) w[r] -= acc;
        }
    }
    __syncthreads();
    const uint64_t c1 = wall_clock64();
    if (T > 0) {
        if (N.nw_stage) tail_matvec(tail, T, w + ns, w + NN);   // staged plans hold the block's inverse
        else block_lu_solve<LUT>(tail, T, T, w + ns);          // unstaged plans hold its LU factors
    }
    const uint64_t c2 = wall_clock64();
    for (int l = N.nw_nlev - 1; l >= 0; l--) {   // backward, sparse rows by level
        const int q0 = lev[l];
        const int R = lev[l + 1] - q0;
        const int G = row_group(R);
        for (int base = 0; base < R; base += NT / G) {
            const int g = base + tid() / G, ln = tid() % G;
            double acc = 0.0;
            int r = -1, ed = 0;
            if (g < R) {
                const int q = q0 + g;
                r = recs[4 * q];
                ed = recs[4 * q + 2];
                acc = row_dot(V, cols, w, ed + 1 + ln, recs[4 * q + 3], G);
            }
            acc = group_sum(acc, G);
            if (r >= 0 && ln == 0) w[r] = (w[r] - acc) / (double)V[ed];
        }
        __syncthreads();
    }
    for (int i = tid(); i < neq; i += NT) b[i] = w[inv[i]];
    __syncthreads();
    if (split) {
        split[0] += (c1 - c0) + (wall_clock64() - c2);   // staging, sparse rows, permutations
        split[1] += c2 - c1;                             // dense tail
        split[2] += cs - c0;                             // staging + right-hand-side sums
        split[3] += c1 - cf0;                            // forward levels + tail rows
    }
}

// Solve M x = b with the factors of sparse_lu: b (length neq, in LDS) is overwritten by x.
// A staged plan copies the solve program (row records in level order, column indices,
// permutation) and the sparse factors into LDS (region) so that every dependent access of
// the level-scheduled substitutions is an LDS access; FP32 factors bring their dense tail
// too, FP64 tails are read from the workspace by the dense triangular solves.  An unstaged
// plan reads all of it from the workspace and keeps only the work vector in region.
template <typename LUT>
UCG_BIG __device__ void sparse_lu_solve(const DevNet &N, const Lds L, const LUT *Vg, LDS double *b, LDS double *region,
                                       uint64_t *split = nullptr)
{
    const uint64_t c0 = wall_clock64();
    const int NN = N.nw_N, nsp = N.nw_nsp;
    if (L.prm[P_ENFORCECHARGE] != 0.0 && N.nelec < N.nspec) {   // right-hand side of the charge row
        double se = 0.0;
        for (int q = tid(); q < N.nion; q += NT) se += b[N.ion_list[q]];
        se = block_sum(se, L.red);
        if (tid() == 0) b[N.nelec] -= se;
        __syncthreads();
    }
    double sb = 0.0, ss = 0.0;
    for (int q = tid(); q < N.nbulkl; q += NT) sb += b[N.bulk_list[q]];
    for (int q = tid(); q < N.nsurfl; q += NT) ss += b[N.surface_list[q]];
    if (!N.nw_stage) {
        block_sum2(sb, ss, L.red);
        solve_program<LUT>(N, Vg, N.nw_sp_col, N.sv_rec, N.nw_inv, N.nw_lev_ptr, Vg + nsp, region, b, sb, ss, split, c0,
                           wall_clock64());
        return;
    }
    constexpr bool STAGE_TAIL = sizeof(LUT) == 4;
    const int nst = STAGE_TAIL ? N.nw_nval : nsp;                       // staged factor entries
    // layout in bytes, 16-byte aligned parts (solve_stage_bytes sizes the same)
    LDS char *base = (LDS char *)region;
    int off = 0;
    LDS LUT *Vs = (LDS LUT *)base;
    off = (nst * (int)sizeof(LUT) + 15) & ~15;
    LDS int *cols = (LDS int *)(base + off);
    off += (nsp * 4 + 15) & ~15;
    LDS int *recs = (LDS int *)(base + off);   // 4 ints per row: {row, first, diagonal, end}
    off += NN * 16;
    LDS int *inv = (LDS int *)(base + off);
    off += ((NN + N.nw_nlev + 1) * 4 + 15) & ~15;   // permutation, then the level pointers
    LDS double *w = (LDS double *)(base + off);
    {
        const GLB LUT *vg = glb(Vg);
        const GLB int *cg = glb(N.nw_sp_col);
        for (int q = tid(); q < nst; q += NT) Vs[q] = vg[q];
        for (int q = tid(); q < nsp; q += NT) cols[q] = cg[q];
    }
    {
        const GLB int *gr = glb(N.sv_rec);
        for (int q = tid(); q < NN; q += NT) {
            const int4 g = ld_int4(gr, q);
            recs[4 * q] = g.x;
            recs[4 * q + 1] = g.y;
            recs[4 * q + 2] = g.z;
            recs[4 * q + 3] = g.w;
        }
    }
    for (int q = tid(); q < NN; q += NT) inv[q] = glb(N.nw_inv)[q];
    LDS int *lev = inv + NN;   // level pointers (nlev + 1), after the permutation
    for (int q = tid(); q <= N.nw_nlev; q += NT) lev[q] = glb(N.nw_lev_ptr)[q];
    block_sum2(sb, ss, L.red);   // its barriers also publish the staged program
    const uint64_t cs = wall_clock64();
    const LDS LUT *V = Vs;
    if constexpr (STAGE_TAIL)
        solve_program<LUT>(N, V, (const LDS int *)cols, (const LDS int *)recs, (const LDS int *)inv,
                           (const LDS int *)lev, V + nsp, w, b, sb, ss, split, c0, cs);
    else
        solve_program<LUT>(N, V, (const LDS int *)cols, (const LDS int *)recs, (const LDS int *)inv,
                           (const LDS int *)lev, Vg + nsp, w, b, sb, ss, split, c0, cs);
}

// ---------------------------------------------------------------- BDF ops policy

// A launch of generation gen stops once the host has raised the handle's cancel word to gen
// or beyond (launches are numbered in issue order), so a cancel also reaches launches that
// were issued before it but have not started yet, and no reset is needed between launches.
__device__ __forceinline__ bool cancel_requested(const int *cancel, int gen)
{
    return cancel && __hip_atomic_load(cancel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) >= gen;
}

template <typename LUT>
struct GpuOps {
    const DevNet &N;
    const Lds L;   // by value: the LDS and workspace pointers live in registers
    LUT *M;
    int neq;
    double z[6][EPT], acor[EPT], ewt[EPT], atol[EPT], f0[EPT];
    uint64_t t_rhs = 0, t_jac = 0, t_lu = 0, t_sol = 0;   // device wall-clock ticks per phase
    uint64_t t_lu_split[2] = {0, 0}, t_sol_split[4] = {0, 0, 0, 0};   // sparse part, dense tail, ...
    uint64_t t_rhs_split[4] = {0, 0, 0, 0};
    bool equilibrate = false;   // factor W M W^-1 (W = diag(ewt)) instead of M (dense form only)
    bool sparse = true;         // static-pattern sparse plan (else dense LU of the full matrix)
    double *Jsave = nullptr;    // saved J-form values of the sparse plan (workspace)
    const int *cancel = nullptr;  // highest cancelled launch generation, raised by the host while kernels run
    int gen = 0;                  // this launch's generation
    __device__ bool cancelled() const { return cancel_requested(cancel, gen); }

    __device__ GpuOps(const DevNet &n_, Lds &l_, LUT *m_) : N(n_), L(l_), M(m_), neq(n_.neq) {}

    __device__ __forceinline__ bool own(int e) const { return tid() + e * NT < neq; }
    __device__ __forceinline__ int idx(int e) const { return tid() + e * NT; }

    // z[j][e] with a run-time j but compile-time storage indices (keeps z in VGPRs)
    __device__ __forceinline__ double zget(int j, int e) const
    {
        switch (j) {
        case 0: return z[0][e];
        case 1: return z[1][e];
        case 2: return z[2][e];
        case 3: return z[3][e];
        case 4: return z[4][e];
        default: return z[5][e];
        }
    }
    __device__ __forceinline__ void zset(int j, int e, double v)
    {
        switch (j) {
        case 0: z[0][e] = v; break;
        case 1: z[1][e] = v; break;
        case 2: z[2][e] = v; break;
        case 3: z[3][e] = v; break;
        case 4: z[4][e] = v; break;
        default: z[5][e] = v; break;
        }
    }
    __device__ double wrms(int v)
    {
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < EPT; e++)
            if (own(e)) {
                double x = (v == bdf::ACOR) ? acor[e] : zget(v, e);
                double t = x * ewt[e];
                acc += t * t;
            }
        return sqrt(block_sum(acc, L.red) / neq);
    }
    __device__ double wrms_acor_minus(double c, int zj)
    {
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < EPT; e++)
            if (own(e)) {
                double t = (acor[e] - c * zget(zj, e)) * ewt[e];
                acc += t * t;
            }
        return sqrt(block_sum(acc, L.red) / neq);
    }
    __device__ bool set_ewt(double rtol)
    {
        double bad = 0.0;
#pragma unroll
        for (int e = 0; e < EPT; e++)
            if (own(e)) {
                double d = rtol * fabs(z[0][e]) + atol[e];
                if (!(d > 0.0)) bad = 1.0;
                ewt[e] = 1.0 / d;
            }
        return block_max(bad, L.red) == 0.0;
    }
    __device__ void start_corrector()
    {
#pragma unroll
        for (int e = 0; e < EPT; e++)
            if (own(e)) {
                acor[e] = 0.0;
                L.y[idx(e)] = z[0][e];
            }
        __syncthreads();
    }
    __device__ void rhs()
    {
        uint64_t t0 = wall_clock64();
        block_rhs(N, L, t_rhs_split);
        t_rhs += wall_clock64() - t0;
    }
    __device__ void save_f0()
    {
#pragma unroll
        for (int e = 0; e < EPT; e++)
            if (own(e)) f0[e] = L.f[idx(e)];
    }
    __device__ void scale_into_z_from(int j, double a, int src)
    {
#pragma unroll
        for (int e = 0; e < EPT; e++)
            if (own(e)) zset(j, e, a * ((src == bdf::F0) ? f0[e] : L.f[idx(e)]));
    }
    __device__ void scale_into_z(int j, double a, int src)
    {
#pragma unroll
        for (int e = 0; e < EPT; e++) zset(j, e, a * zget(src, e));
    }
    __device__ void axpy_zz(int j, double a, int src)
    {
#pragma unroll
        for (int e = 0; e < EPT; e++) zset(j, e, zget(j, e) + a * zget(src, e));
    }
    __device__ void rescale_z(int q, double eta)
    {
        double f = eta;
#pragma unroll
        for (int j = 1; j <= bdf::QMAX; j++) {
            if (j <= q) {
#pragma unroll
                for (int e = 0; e < EPT; e++) z[j][e] *= f;
                f *= eta;
            }
        }
    }
    __device__ void predict(int q)
    {
#pragma unroll
        for (int k = 1; k <= bdf::QMAX; k++)
#pragma unroll
            for (int j = bdf::QMAX; j >= k; j--)
                if (k <= q && j <= q)
#pragma unroll
                    for (int e = 0; e < EPT; e++) z[j - 1][e] += z[j][e];
    }
    __device__ void unpredict(int q)
    {
#pragma unroll
        for (int k = 1; k <= bdf::QMAX; k++)
#pragma unroll
            for (int j = bdf::QMAX; j >= k; j--)
                if (k <= q && j <= q)
#pragma unroll
                    for (int e = 0; e < EPT; e++) z[j - 1][e] -= z[j][e];
    }
    __device__ void complete(int q, const double *l)
    {
#pragma unroll
        for (int j = 0; j <= bdf::QMAX; j++)
            if (j <= q)
#pragma unroll
                for (int e = 0; e < EPT; e++) z[j][e] += l[j] * acor[e];
    }
    __device__ void copy_acor_to_z(int j)
    {
#pragma unroll
        for (int e = 0; e < EPT; e++) zset(j, e, acor[e]);
    }
    __device__ void form_residual(double gamma, double rl1)
    {
#pragma unroll
        for (int e = 0; e < EPT; e++)
            if (own(e)) {
                const int i = idx(e);
                // equilibrated system: b' = W b
                L.b[i] = (gamma * L.f[i] - (rl1 * z[1][e] + acor[e])) * (equilibrate ? ewt[e] : 1.0);
            }
        __syncthreads();
    }
    __device__ void solve()
    {
        uint64_t t0 = wall_clock64();
        if (sparse) sparse_lu_solve<LUT>(N, L, M, L.b, L.flux, t_sol_split);
        else block_lu_solve<LUT>(M, neq, N.ld, L.b);
        t_sol += wall_clock64() - t0;
    }
    __device__ double apply_correction(double scale)
    {
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < EPT; e++)
            if (own(e)) {
                const int i = idx(e);
                double d = scale * L.b[i];
                if (equilibrate) d = d / ewt[e];  // the equilibrated solve returns W d
                double dw = d * ewt[e];
                acc += dw * dw;
                acor[e] += d;
                L.y[i] = z[0][e] + acor[e];
            }
        double r = sqrt(block_sum(acc, L.red) / neq);
        return r;
    }
    // Reusing a saved J between Newton-matrix setups (DVODE's JSV=1) was measured on the
    // seeded grid: 31% more steps (slower Newton convergence) for 7x fewer J evaluations,
    // a net loss since an analytic J costs under half a factorisation here; kept off.
    bool reuse_jacobian = false;
    __device__ bool can_reuse_jacobian() const { return sparse && reuse_jacobian; }
    __device__ int setup(double gamma, bool new_jacobian)
    {
#pragma unroll
        for (int e = 0; e < EPT; e++)
            if (own(e)) L.g[idx(e)] = equilibrate ? ewt[e] : 1.0;
        __syncthreads();
        uint64_t t0 = wall_clock64();
        int r;
        uint64_t t1;
        if (sparse) {
            if (new_jacobian) sparse_jacobian(N, L, Jsave);
            sparse_form_newton<LUT>(N, Jsave, M, gamma, L.prm[P_ENFORCECHARGE] != 0.0);
            t1 = wall_clock64();
            r = sparse_lu<LUT>(N, M, L.flux, L.red, L.b, t_lu_split);
        } else {
            block_newton_matrix<LUT>(N, L, M, gamma, true, equilibrate);
            t1 = wall_clock64();
            r = block_lu<LUT>(M, neq, N.ld, L.red, L.flux);
        }
        uint64_t t2 = wall_clock64();
        t_jac += t1 - t0;
        t_lu += t2 - t1;
        return r;
    }
    __device__ double hub_inverse(double fac)
    {
        double m = 0.0;
#pragma unroll
        for (int e = 0; e < EPT; e++)
            if (own(e)) m = fmax(m, fabs(f0[e]) / (fac * fabs(z[0][e]) + atol[e]));
        return block_max(m, L.red);
    }
    __device__ double ydd_norm(double t0, double hg)
    {
#pragma unroll
        for (int e = 0; e < EPT; e++)
            if (own(e)) L.y[idx(e)] = z[0][e] + hg * f0[e];
        __syncthreads();
        block_rhs(N, L);
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < EPT; e++)
            if (own(e)) {
                double t = (L.f[idx(e)] - f0[e]) / hg * ewt[e];
                acc += t * t;
            }
        return sqrt(block_sum(acc, L.red) / neq);
    }
    __device__ void finish_at_tn() {}
    __device__ void interpolate(double s, int q)
    {
#pragma unroll
        for (int e = 0; e < EPT; e++) {
            double acc = 0.0;
#pragma unroll
            for (int j = bdf::QMAX; j >= 0; j--)
                if (j <= q) acc = (j == q) ? z[j][e] : acc * s + z[j][e];
            z[0][e] = acc;
        }
    }
};

// ---------------------------------------------------------------- persistent solve kernel

constexpr int STATUS_CANCELLED = -10;   // UCG_E_CANCELLED
constexpr int MAX_INTERVALS = 100000;

struct SolveArgs {
    int n_points;
    const double *params;     // [n][P_NPARAM]
    const double *y_init;     // [n][neq] or null (elemental start, chemistry.f90:46-141)
    double *y_out;            // [n][neq]
    int max_rec;              // rows per point in rec (0: none)
    double *rec;              // [n][max_rec][4 + nspec]
    double *rates_rec;        // [n][max_rec][nreac] rate coefficients per output row, or null
    int *n_rec;               // [n]
    int *status;              // [n]
    int *stats;               // [n][8]: nst nfe nje nni ncfn netf n_intervals n_retries
    double *timers;           // [n][8] seconds: rates, rhs, jacobian, lu, solve, total, -, - (or null)
    double *t_start;          // [n] device wall clock when the point was taken, or null
    double *t_done;           // [n] device wall clock at retirement, or null
    void *ws;                 // slot workspaces
    size_t ws_stride;         // bytes per slot: Newton matrix (ws_mat_bytes), saved J-form values
    size_t ws_mat_bytes;      // (ws_jac_bytes, sparse plan only), then rate[nreac]
    size_t ws_jac_bytes;
    int *counter;             // work queue head
    const int *cancel;        // nullable: the highest cancelled launch generation (points this launch
    int gen;                  // has not finished get STATUS_CANCELLED once it reaches gen)
    // single-interval mode (phys_in non-null): each point is one updateChemistry call from the
    // state phys_in/y_init to t_target, with no physics update and no output rows
    const double *phys_in;    // [n][8]: density, Tgas, Tdust, Av, cloudSize, phi, abstol_factor, time [s]
    double *phys_out;         // [n][8]: the same after the interval
    const double *t_target;   // [n] seconds
    int equilibrate;          // factor the error-weight-equilibrated Newton matrix (dense form)
    int sparse;               // 1: static-pattern sparse plan, 0: dense LU
    // request-queue mode (all null/0 for a batch): n_points is the queue capacity and the
    // host posts requests while the kernel runs
    const int *q_avail;       // requests posted so far (host-written, read at system scope)
    const int *q_closed;      // nonzero once the host admits no more requests
    int *q_done;              // [n]: set to 1 after request i's outputs are written
    long long q_idle_ticks;   // a workgroup that waits this long (100 MHz ticks) for a request leaves
    // shared queue over several networks (null for one network): request i runs on network
    // q_net[i] (host-written with the request), whose descriptor is q_nets[q_net[i]]; outputs
    // are y_stride apart (the largest NEQ of the networks)
    const DevNet *const *q_nets;
    const int *q_net;
    int y_stride;             // y_out row stride, 0: the network's NEQ
};

__device__ void elemental_start(const DevNet &N, const LDS double *prm, LDS double *y, double density)
{   // chemistry.f90:58-105: elemental abundances from the point's parameters, in the
    // reference's order (species a network lacks share the density slot, restored last)
    const int neq = N.neq;
    for (int i = tid(); i < neq; i += NT) y[i] = MIN_ABUND;
    __syncthreads();
    if (tid() == 0) {
        y[N.no] = prm[P_FO];
        y[N.nn] = prm[P_FN];
        y[N.nmg] = prm[P_FMG];
        y[N.np] = prm[P_FP];
        y[N.nf] = prm[P_FF];
        y[N.nna] = prm[P_FNA];
        y[N.nli] = prm[P_FLI];
        y[N.npah] = prm[P_FPAH];
        y[N.nsx] = prm[P_FS];
        y[N.nsix] = prm[P_FSI];
        y[N.nclx] = prm[P_FCL];
        const double ion = prm[P_ION];      // how much carbon starts ionised
        if (ion == 0.0) {
            y[N.nc] = prm[P_FC];
            y[N.ncx] = 1.0e-10;
        } else if (ion == 1.0) {
            y[N.nc] = prm[P_FC] * 0.5;
            y[N.ncx] = prm[P_FC] * 0.5;
        } else if (ion == 2.0) {
            y[N.nc] = 1.0e-10;
            y[N.ncx] = prm[P_FC];
        }
        y[N.n18o] = prm[P_F18O];
        y[N.n15n] = prm[P_F15N];
        y[N.n13c] = prm[P_F13C];
        y[N.nelec] = y[N.ncx] + y[N.nsix] + y[N.nsx] + y[N.nclx] + y[N.nmgx];
    }
    __syncthreads();
    const double met = prm[P_METALLICITY];
    for (int i = tid(); i < neq; i += NT) y[i] = y[i] * met;
    __syncthreads();
    if (tid() == 0) {
        y[N.nh] = prm[P_FH];
        y[N.nh2] = 0.5 * (1.0 - prm[P_FH]);
        y[N.nd] = prm[P_FD];
        y[N.nhe] = prm[P_FHE];
        y[neq - 1] = density;
    }
    __syncthreads();
}

// H2 dissociation by cosmic rays attenuated by the column (physics-core.f90:117-155, the
// improvedH2CRPDissociation rate the physics update sets): a degree-9 polynomial in
// log10(coldens), coefficients of the L or H model
__device__ inline double h2_crp_rate_dev(double coldens, double zeta, bool high)
{
    const double ck_l[10] = {1.582911005330e7, -6.465722684896e6, 1.172189025424e6, -1.237950798073e5,
                             8.393404654312e3, -3.788811358130e2, 1.138688455029e1, -2.197136304567e-1,
                             2.469841278950e-3, -1.232393620924e-5};
    const double ck_h[10] = {1.217227462831e7, -4.989649250304e6, 9.079152156645e5, -9.624890825395e4,
                             6.551161486120e3, -2.968976216187e2, 8.959037875226e0, -1.735757324445e-1,
                             1.959267277734e-3, -9.816996707980e-6};
    const double lc = log10(coldens);
    double sum = 0.0, pw = 1.0;
    for (int k = 0; k < 10; k++) {
        sum += (high ? ck_h[k] : ck_l[k]) * pw;
        pw *= lc;
    }
    return pow(10.0, sum) * zeta;
}

__device__ void record_row(const DevNet &N, const Lds L, const SolveArgs &A, int pidx, int row, double tyears)
{
    if (A.max_rec <= 0 || row >= A.max_rec) return;
    const int ns = N.nspec;
    double *r = A.rec + ((size_t)pidx * A.max_rec + row) * (4 + ns);
    if (tid() == 0) {
        r[0] = tyears;
        r[1] = L.sc[S_DENS];
        r[2] = L.sc[S_TGAS];
        r[3] = L.sc[S_AV];
    }
    for (int i = tid(); i < ns; i += NT) r[4 + i] = L.y[i];
    if (A.rates_rec) {
        // ratesarray (io.f90:90-93): the rate array as the row is written, i.e. the last
        // attempt's coefficients with the H2/CO photo rates of the last RHS evaluation; before
        // the first interval the point has no rates yet (zeros)
        double *rr = A.rates_rec + ((size_t)pidx * A.max_rec + row) * N.nreac;
        for (int k = tid(); k < N.nreac; k += NT) rr[k] = row ? L.rate[k] : 0.0;
    }
}

// LDS layout shared by all kernels; the rate coefficients live in the workspace slot
// (read coalesced, written once per interval attempt).
__device__ void lds_layout(const DevNet &N, Lds &L, double *lds_raw, double *rate)
{
    const int npad = (N.neq + 7) & ~7;
    L.rate = rate;
    L.flux = (LDS double *)lds_raw;
    L.y = L.flux + flux_region_len(N);
    L.f = L.y + npad;
    L.b = L.f + npad;
    L.g = L.b + npad;
    L.red = L.g + npad;
    L.sc = L.red + 16;
    L.prm = L.sc + S_NSCAL;
}

// Resets the work counter and the cancel flag and marks every point "not run" before a
// solve; a kernel rather than memsets so that the launch sequence can be graph-captured.
__global__ void solve_init_kernel(int *counter, int *status, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *counter = 0;
    if (i < n) status[i] = STATUS_CANCELLED;
}

// Sum of term(i), i = 1..n, added in increasing i as the reference's sequential loops add
// them (collapse.f90:140-176).  With `stop_at`, the sum ends after the first i at which it
// is no longer below `stop` and *hit = i (0 if it never gets there).  The terms are
// evaluated NT at a time into `buf` (NT doubles of LDS); every thread adds them in the same
// order, so every thread returns the same sum and hit.
template <class Term>
__device__ double ordered_sum(const Term &term, int n, double stop, bool stop_at, int *hit, LDS double *buf)
{
    double s = 0.0;
    int h = 0;
    for (int base = 0; base < n && !h; base += NT) {
        const int i = base + tid() + 1;
        __syncthreads();   // the previous chunk has been read
        buf[tid()] = i <= n ? term(i) : 0.0;
        __syncthreads();
        const int m = n - base < NT ? n - base : NT;
        for (int j = 0; j < m; j++) {
            s = s + buf[j];
            if (stop_at && !(s < stop)) {
                h = base + j + 1;
                break;
            }
        }
    }
    __syncthreads();
    *hit = h;
    return s;
}

// Steps of findNewRadius's outward search (dr = rout/1e4): the reference searches without a
// bound; a parcel this far out (400 rout) has no meaning in these profiles.
constexpr int COLLAPSE_MAX_STEPS = 4000000;

// collapse.f90:27-59 initializePhysics for one parcel (points = 1): the parcel sits at rout,
// the density is the profile's at rin, and modes 1 and 2 record the mass inside the parcel
// (findMassInRadius, 1000 trapezoids).  Called by every thread.
__device__ __attribute__((noinline)) void collapse_setup(const LDS double *prm, Lds &L)
{
    const CollapseFit F = collapse_fit((int)prm[P_COLLAPSE]);
    const double rho0 = F.rho0(0.0), r0 = F.r0(0.0), a = F.slope(0.0);
    const double rad = 1 * prm[P_ROUT] / (double)1.0f;
    double mass = 0.0;
    if (F.mode <= 2) {
        const double dr = rad / 1000;
        int hit;
        mass = ordered_sum(
            [&](int i) {
                const double x = i * dr;
                return 0.5 * (F.rho(x, rho0, r0, a) + F.rho((i - 1) * dr, rho0, r0, a)) * dr * (x * x);
            },
            1000, 0.0, false, &hit, L.flux);
    }
    if (tid() == 0) {
        L.sc[S_RAD] = rad;
        L.sc[S_MASS] = mass;
        L.sc[S_DENS] = F.rho(prm[P_RIN], rho0, r0, a);
    }
    __syncthreads();
}

// collapse.f90:81-107 updatePhysics at `ty` years, after the physics core's update: the column
// to the cloud edge through the current profile (findcoldens, 10000 trapezoids) sets Av; modes
// 1 and 2 move the parcel to the radius that encloses its mass (findNewRadius), modes 3 and 4
// by the fitted radial velocity over `dt` seconds (the part of the interval the integrator did
// not cover: the reference takes targetTime - currentTime after the interval); the density
// is the profile's at the new radius.  Called by every thread; both collapse routines stay out
// of line so that they add no register pressure to the integration loop.
__device__ __attribute__((noinline)) void collapse_update(const LDS double *prm, Lds &L, double ty, double dt)
{
    const CollapseFit F = collapse_fit((int)prm[P_COLLAPSE]);
    const double rho0 = F.rho0(ty), r0 = F.r0(ty), a = F.slope(ty);
    const double rin = prm[P_RIN], rout = prm[P_ROUT];
    int hit;
    double cd = 0.0;
    const double size = rout - rin;
    if (size > 0.0) {
        const double dr = size / 10000;
        cd = ordered_sum(
            [&](int i) {
                const double r1 = rin + (i - 1) * dr, r2 = rin + i * dr;
                return 0.5 * (F.rho(r2, rho0, r0, a) + F.rho(r1, rho0, r0, a)) * dr * PC;
            },
            10000, 0.0, false, &hit, L.flux);
    }
    double rad = L.sc[S_RAD];
    if (F.mode <= 2) {
        const double mass = L.sc[S_MASS];
        if (0.0 < mass) {   // else the reference's search loop does not run and the radius stays
            const double dr = rout / 1.0e4;
            ordered_sum(
                [&](int i) {
                    const double x = i * dr;
                    return 0.5 * (F.rho(x, rho0, r0, a) + F.rho((i - 1) * dr, rho0, r0, a)) * dr * (x * x);
                },
                COLLAPSE_MAX_STEPS, mass, true, &hit, L.flux);
            rad = (hit ? hit : COLLAPSE_MAX_STEPS) * dr;
        }
    } else {
        rad = rad + F.vr(rad, F.rmin(ty), F.vmin(ty), F.vslope(ty)) * dt / PC;
    }
    if (tid() == 0) {
        L.sc[S_COLDENS] = cd;
        L.sc[S_AV] = prm[P_BASEAV] + cd / 1.6e21;
        L.sc[S_RAD] = rad;
        L.sc[S_DENS] = F.rho(rad, rho0, r0, a);
    }
    __syncthreads();
}

// The checks the host API makes of a parameter vector, repeated on the device for callers of
// the device-pointer entry points: 0, or the status of a point that is not integrated:
// -1 (PARAMETER_READ_ERROR: a value the kernel cannot use, e.g. a tempindx that would index
// the hot-core tables out of bounds) or -2 (PHYSICS_INIT_ERROR: improvedH2CRPDissociation
// without cosmicRayAttenuation, physics-core.f90:60-64).
__device__ inline int params_status(const LDS double *q)
{
    const double model = q[P_MODEL], ti = q[P_TEMPINDX], im = q[P_IONMODEL], cm = q[P_COLLAPSE];
    bool ok = q[P_DENS] > 0.0 && q[P_TEMP] > 0.0 && q[P_RELTOL] > 0.0 && q[P_ABSMIN] >= 0.0 &&
              q[P_MXSTEP] >= 1.0 && q[P_ROUT] > q[P_RIN] && q[P_MAXSTEPS] >= 0.0 &&
              (model == 0.0 || model == 1.0 || model == 2.0) && (im == 0.0 || im == 1.0) && q[P_OMEGA] != 1.0 &&
              q[P_UVCREFF] != 0.0;
    if (model == 1.0) ok = ok && ti >= 1.0 && ti < 7.0;
    if (model == 2.0) ok = ok && (cm == 1.0 || cm == 2.0 || cm == 3.0 || cm == 4.0);
    if (!ok) return -1;
    if (q[P_IMPROVEDH2CRP] != 0.0 && q[P_CRATTEN] == 0.0) return -2;
    return 0;
}

template <typename LUT>
__global__ __launch_bounds__(NT, UCG_WAVES_PER_EU) void solve_points_kernel(const DevNet *__restrict__ Np, SolveArgs A)
{
    extern __shared__ double lds_raw[];
    char *slot = (char *)A.ws + (size_t)blockIdx.x * A.ws_stride;
    __shared__ int s_point, s_net;
    LUT *M = (LUT *)slot;

    for (;;) {
        if (tid() == 0) {
            const bool stop = cancel_requested(A.cancel, A.gen);
            int p = stop ? A.n_points : atomicAdd(A.counter, 1);
            if (A.q_avail && p < A.n_points) {   // request queue: wait until request p is posted
                const uint64_t w0 = wall_clock64();
                for (;;) {
                    if (__hip_atomic_load(A.q_avail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) > p) break;
                    const bool closed = __hip_atomic_load(A.q_closed, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) != 0;
                    if (closed && __hip_atomic_load(A.q_avail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) > p) break;
                    const bool cancelled = cancel_requested(A.cancel, A.gen);
                    if (closed || cancelled || (long long)(wall_clock64() - w0) > A.q_idle_ticks) {
                        p = A.n_points;   // every path of this loop ends: post, close, cancel or idle limit
                        break;
                    }
                    __builtin_amdgcn_s_sleep(127);
                }
            }
            s_point = p;
            // a shared queue's request names its network (written before the request was posted)
            s_net = (A.q_net && p < A.n_points)
                        ? __hip_atomic_load(A.q_net + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)
                        : 0;
        }
        __syncthreads();
        const int pidx = s_point, inet = s_net;
        __syncthreads();
        if (pidx >= A.n_points) break;   // every wave of the workgroup leaves here
        // the point's network: tables in global memory (field loads stay scalar, nothing copied
        // to scratch); its LDS layout fits the launch's LDS, sized for the largest network
        const DevNet &N = A.q_nets ? *A.q_nets[inet] : *Np;
        const int neq = N.neq;
        const size_t ystride = A.y_stride ? (size_t)A.y_stride : (size_t)neq;
        Lds L;
        lds_layout(N, L, lds_raw, (double *)(slot + A.ws_mat_bytes + A.ws_jac_bytes));

        if (A.t_start && tid() == 0) A.t_start[pidx] = (double)wall_clock64();
        // ---- point set-up: parameters, physics (physics-core.f90:42-73, cloud.f90:20-31)
        const double *pin = A.params + (size_t)pidx * P_NPARAM;
        for (int i = tid(); i < P_NPARAM; i += NT) L.prm[i] = pin[i];
        __syncthreads();
        const LDS double *prm = L.prm;
        const bool interval_mode = A.phys_in != nullptr;
        const int pstatus = params_status(prm);
        if (interval_mode) {   // the caller's state at the interval start
            if (tid() == 0) {
                const double *ph = A.phys_in + (size_t)pidx * 8;
                L.sc[S_DENS] = ph[0];
                L.sc[S_TGAS] = ph[1];
                L.sc[S_TDUST] = ph[2];
                L.sc[S_AV] = ph[3];
                L.sc[S_CS] = ph[4];
                L.sc[S_PHI] = ph[5];
                L.sc[S_ABSFAC] = ph[6];
                // the rate the last physics update set from the column at this density
                L.sc[S_H2CRP] = (ph[7] > 0.0 && prm[P_CRATTEN] != 0.0 && prm[P_IMPROVEDH2CRP] != 0.0)
                                    ? h2_crp_rate_dev(ph[4] * ph[0], prm[P_ZETA], prm[P_IONMODEL] != 0.0) : 0.0;
            }
        } else if (tid() == 0) {
            double cs = (prm[P_ROUT] - prm[P_RIN]) * PC;
            L.sc[S_CS] = cs;
            L.sc[S_TGAS] = prm[P_TEMP];
            L.sc[S_TDUST] = prm[P_TEMP];
            double dens = prm[P_DENS];
            L.sc[S_COLDENS] = 1.0 * cs / 1.0 * prm[P_DENS];
            L.sc[S_AV] = prm[P_BASEAV] + L.sc[S_COLDENS] / 1.6e21;
            if (prm[P_FREEFALL] != 0.0) dens = (double)1.001f * prm[P_DENS];
            L.sc[S_DENS] = dens;
            L.sc[S_PHI] = prm[P_PHI];
            L.sc[S_ABSFAC] = prm[P_ABSFAC];
            L.sc[S_H2CRP] = 0.0;   // set by the physics update (physics-core.f90:87,143-155)
        }
        __syncthreads();
        // the collapse models replace the cloud's initializePhysics (wrap.f90:145-147)
        if (!interval_mode && pstatus == 0 && (int)prm[P_MODEL] == 2) collapse_setup(prm, L);
        if (A.y_init) {
            for (int i = tid(); i < neq; i += NT) L.y[i] = A.y_init[(size_t)pidx * neq + i];
            __syncthreads();
        } else {
            elemental_start(N, prm, L.y, L.sc[S_DENS]);
        }
        GpuOps<LUT> o(N, L, M);
        o.sparse = A.sparse != 0;
        o.Jsave = (double *)(slot + A.ws_mat_bytes);
        o.cancel = A.cancel;
        o.equilibrate = !o.sparse && A.equilibrate != 0;
        o.gen = A.gen;
        bdf::Stats st;
        int row = 0, nint = 0, nretry = 0, flag = pstatus;
        if (!interval_mode) record_row(N, L, A, pidx, row++, 0.0);
        double t = interval_mode ? A.phys_in[(size_t)pidx * 8 + 7] : 0.0;
        double tyears = t / SECONDS_PER_YEAR;
        const int model = (int)prm[P_MODEL];
        const double reltol = prm[P_RELTOL], absmin = prm[P_ABSMIN];
        const int mxstep = (int)prm[P_MXSTEP];
        const double maxsteps = prm[P_MAXSTEPS];   // 0: unbounded
        const bool end_dens = prm[P_ENDATFINALDENS] != 0.0;
        const bool freefall = prm[P_FREEFALL] != 0.0;
        // collapse modes 1 and 2 end at 0.97 maxTime whatever finalTime says (collapse.f90:35-40)
        const double final_time = (model == 2 && pstatus == 0 && prm[P_COLLAPSE] <= 2.0)
                                      ? (double)0.97f * collapse_fit((int)prm[P_COLLAPSE]).maxt
                                      : prm[P_FINALTIME];
        // ---- time loop (wrap.f90:657-690).  The reference loops forever when the
        // integrator cannot advance t at all; here such a point stops with
        // INT_TOO_MANY_FAILS_ERROR after 10 intervals without progress.
        int stuck = 0;
        uint64_t t_rates = 0;
        const uint64_t t_point0 = wall_clock64();
        const uint64_t c_point0 = clock64();   // shader clock, for the effective frequency
        while (!flag && (interval_mode ? nint == 0
                                       : ((end_dens && L.sc[S_DENS] < prm[P_FINALDENS]) ||
                                          (!end_dens && tyears < final_time)))) {
            if (nint >= MAX_INTERVALS || stuck > 10) {
                flag = -5;
                break;
            }
            const double t_before = t;
            tyears = t / SECONDS_PER_YEAR;
            double target = interval_mode ? A.t_target[pidx]
                            : model == 2  ? collapse_next_target(tyears)
                                          : next_target_dev(model, tyears);
            const double orig = target;
            int loops = 0;
            while (t < target && loops < 10) {   // updateChemistry retry loop (chemistry.f90:164-223)
                if (maxsteps > 0.0 && (double)st.nst >= maxsteps) {
                    flag = -5;   // the caller's work bound (INT_TOO_MANY_FAILS_ERROR)
                    break;
                }
                // the density the physics set, unless it is integrated (chemistry.f90:166)
                if (!freefall && tid() == 0) L.y[neq - 1] = L.sc[S_DENS];
                __syncthreads();
                const uint64_t tr0 = wall_clock64();
                block_rates(N, L);
                t_rates += wall_clock64() - tr0;
                const double absfac = L.sc[S_ABSFAC];
#pragma unroll
                for (int e = 0; e < EPT; e++)
                    if (o.own(e)) {
                        double yi = L.y[o.idx(e)];
                        o.z[0][e] = yi;
                        double a = absfac * yi;
                        o.atol[e] = a < absmin ? absmin : a;
                    }
                __syncthreads();
                int is = bdf::integrate(o, t, target, reltol, mxstep, st);
                if (is == -2 && tid() == 0) L.sc[S_ABSFAC] = absfac * 10.0;
                if (is < 0 && is != -7) nretry++;
                if (is == -3) {
                    flag = -4;   // INT_UNRECOVERABLE_ERROR (constants.f90:25)
                }
                if (is == -7) flag = STATUS_CANCELLED;
                // enforceChargeConservation (chemistry.f90:288-291: electrons = sum of the
                // positive ions after the integrator call), then the clamp (:215) and the
                // density bookkeeping (:216)
                if (prm[P_ENFORCECHARGE] != 0.0 && is != -3 && is != -7) {
#pragma unroll
                    for (int e = 0; e < EPT; e++)
                        if (o.own(e)) L.y[o.idx(e)] = o.z[0][e];
                    __syncthreads();
                    double ions = 0.0;
                    for (int q = tid(); q < N.nion; q += NT) ions += L.y[N.ion_list[q]];
                    ions = block_sum(ions, L.red);
#pragma unroll
                    for (int e = 0; e < EPT; e++)
                        if (o.own(e) && o.idx(e) == N.nelec) o.z[0][e] = ions;
                    __syncthreads();
                }
#pragma unroll
                for (int e = 0; e < EPT; e++)
                    if (o.own(e)) {
                        double yi = o.z[0][e];
                        if (!(yi >= MIN_ABUND)) yi = (yi == yi) ? MIN_ABUND : yi;
                        L.y[o.idx(e)] = yi;
                    }
                __syncthreads();
                if (tid() == 0) L.sc[S_DENS] = L.y[neq - 1];
                __syncthreads();
                if (flag) break;
                loops++;
                target = orig;
            }
            if (flag) break;
            nint++;
            if (interval_mode) break;   // updateChemistry ends here; the physics update is the caller's
            stuck = (t > t_before) ? 0 : stuck + 1;
            tyears = target / SECONDS_PER_YEAR;
            // physics update (physics-core.f90:75-88, then hotcore.f90:78-90 or collapse.f90:81-107)
            if (tid() == 0) {
                const double cs = L.sc[S_CS];
                L.sc[S_COLDENS] = cs / 1.0 * L.sc[S_DENS];
                L.sc[S_AV] = prm[P_BASEAV] + L.sc[S_COLDENS] / 1.6e21;
                L.sc[S_TDUST] = L.sc[S_TGAS];
                if (prm[P_CRATTEN] != 0.0 && prm[P_IMPROVEDH2CRP] != 0.0)
                    L.sc[S_H2CRP] = h2_crp_rate_dev(L.sc[S_COLDENS], prm[P_ZETA], prm[P_IONMODEL] != 0.0);
                if (model == 1) {
                    const double tempa[6] = {1.927e-1, 4.8560e-2, 7.8470e-3, 9.6966e-4, 1.706e-4, 4.74e-7};
                    const float tempb[6] = {0.5339f, 0.6255f, 0.8395f, 1.085f, 1.289f, 1.98f};
                    const int ti = (int)prm[P_TEMPINDX] - 1;
                    const double maxt = prm[P_MAXTEMP];
                    if (L.sc[S_TGAS] < maxt) {
                        double gt = (cs / (prm[P_ROUT] * PC)) * ((double)1.0f / (double)1.0f);
                        gt = pow(gt, (double)-0.5f);
                        gt = prm[P_TEMP] + ((tempa[ti] * pow(t / SECONDS_PER_YEAR, (double)tempb[ti])) * gt);
                        if (gt > maxt) gt = maxt;
                        L.sc[S_TGAS] = gt;
                    }
                    L.sc[S_TDUST] = L.sc[S_TGAS];
                }
            }
            __syncthreads();
            if (model == 2) collapse_update(prm, L, tyears, target - t);
            if (row < A.max_rec) record_row(N, L, A, pidx, row, tyears);
            row++;
        }
        if (A.max_rec > 0 && row > A.max_rec && flag == 0) flag = -6;   // NOT_ENOUGH_TIMEPOINTS_ERROR
        // ---- retire the point
        for (int i = tid(); i < neq; i += NT) A.y_out[(size_t)pidx * ystride + i] = L.y[i];
        if (tid() == 0) {
            A.status[pidx] = flag;
            if (A.phys_out) {
                double *ph = A.phys_out + (size_t)pidx * 8;
                ph[0] = L.sc[S_DENS]; ph[1] = L.sc[S_TGAS]; ph[2] = L.sc[S_TDUST]; ph[3] = L.sc[S_AV];
                ph[4] = L.sc[S_CS]; ph[5] = L.sc[S_PHI]; ph[6] = L.sc[S_ABSFAC]; ph[7] = t;
            }
            if (A.n_rec) A.n_rec[pidx] = row < A.max_rec ? row : A.max_rec;
            int *s = A.stats + (size_t)pidx * 8;
            s[0] = st.nst; s[1] = st.nfe; s[2] = st.nje; s[3] = st.nni;
            s[4] = st.ncfn; s[5] = st.netf; s[6] = nint; s[7] = nretry;
            if (A.t_done) A.t_done[pidx] = (double)wall_clock64();
            if (A.timers) {
                const double tick = 1.0e-8;   // wall_clock64 runs at 100 MHz
                double *tm = A.timers + (size_t)pidx * 16;
                tm[0] = t_rates * tick; tm[1] = o.t_rhs * tick; tm[2] = o.t_jac * tick; tm[3] = o.t_lu * tick;
                const uint64_t dw = wall_clock64() - t_point0, dc = clock64() - c_point0;
                tm[4] = o.t_sol * tick; tm[5] = dw * tick;
                tm[6] = dw ? (double)dc / ((double)dw * tick) * 1e-6 : 0.0;   // mean shader clock, MHz
                tm[7] = o.t_lu_split[0] * tick; tm[8] = o.t_lu_split[1] * tick;
                tm[9] = o.t_sol_split[0] * tick; tm[10] = o.t_sol_split[1] * tick;
                tm[11] = o.t_sol_split[2] * tick;   // solve: staging + right-hand-side sums
                for (int k = 1; k < 4; k++) tm[11 + k] = o.t_rhs_split[k] * tick;
                tm[15] = o.t_sol_split[3] * tick;   // solve: forward levels (staging in [9])
            }
        }
        __syncthreads();
        if (A.q_done) {   // request queue: publish the outputs, then the completion flag
            __threadfence_system();
            __syncthreads();
            if (tid() == 0) __hip_atomic_store(A.q_done + pidx, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// ---------------------------------------------------------------- fine-grained kernels (tests)

struct PieceArgs {
    int n;
    const double *params;   // [n][P_NPARAM]
    const double *y;        // [n][neq]
    const double *rate_in;  // [n][nreac] (rhs/jac) or null
    const double *phys;     // [n][8]: density, Tgas, Tdust, Av, cloudSize, phi, abstol_factor, unused
    double *rate_out;       // [n][nreac]
    double *y_out;          // [n][neq] (rates: y after the ice reset; rhs: ydot)
    double *scal_out;       // [n][8]
    double *jac_out;        // [n][neq*neq] column-major, J (not M)
    double *flux_out;       // [n][nreac] per-reaction terms of the RHS (rhs_kernel), or null
    double gamma;
    int sparse;
    void *ws;
    size_t ws_stride;      // bytes per point: Newton matrix (ws_mat_bytes), J-form values
    size_t ws_mat_bytes;   // (ws_jac_bytes), then rate[nreac]
    size_t ws_jac_bytes;
};

__device__ void piece_setup(const DevNet &N, Lds &L, const PieceArgs &A, int p)
{
    for (int i = tid(); i < P_NPARAM; i += NT) L.prm[i] = A.params[(size_t)p * P_NPARAM + i];
    for (int i = tid(); i < N.neq; i += NT) L.y[i] = A.y[(size_t)p * N.neq + i];
    __syncthreads();
    if (tid() == 0) {
        const double *ph = A.phys + (size_t)p * 8;
        L.sc[S_DENS] = ph[0];
        L.sc[S_TGAS] = ph[1];
        L.sc[S_TDUST] = ph[2];
        L.sc[S_AV] = ph[3];
        L.sc[S_CS] = ph[4];
        L.sc[S_PHI] = ph[5];
        L.sc[S_ABSFAC] = ph[6];
        const LDS double *prm = L.prm;
        L.sc[S_H2CRP] = (ph[7] > 0.0 && prm[P_CRATTEN] != 0.0 && prm[P_IMPROVEDH2CRP] != 0.0)
                            ? h2_crp_rate_dev(ph[4] * ph[0], prm[P_ZETA], prm[P_IONMODEL] != 0.0) : 0.0;
    }
    if (A.rate_in)
        for (int k = tid(); k < N.nreac; k += NT) L.rate[k] = A.rate_in[(size_t)p * N.nreac + k];
    __syncthreads();
    if (tid() == 64) L.sc[S_SCAT] = scatter_dev(1000.0, L.sc[S_AV]);
    __syncthreads();
}


__global__ __launch_bounds__(NT) void rates_kernel(const DevNet *__restrict__ Np, PieceArgs A)
{
    const DevNet &N = *Np;   // tables in global memory: field loads stay scalar, nothing copied to scratch
    extern __shared__ double lds_raw[];
    Lds L;
    lds_layout(N, L, lds_raw, (double *)((char *)A.ws + (size_t)blockIdx.x * A.ws_stride + A.ws_mat_bytes + A.ws_jac_bytes));
    const int p = blockIdx.x;
    if (p >= A.n) return;
    piece_setup(N, L, A, p);
    block_rates(N, L);
    for (int k = tid(); k < N.nreac; k += NT) A.rate_out[(size_t)p * N.nreac + k] = L.rate[k];
    for (int i = tid(); i < N.neq; i += NT) A.y_out[(size_t)p * N.neq + i] = L.y[i];
    if (tid() == 0) {
        double *s = A.scal_out + (size_t)p * 8;
        s[0] = L.sc[S_H2COL]; s[1] = L.sc[S_COCOL]; s[2] = L.sc[S_CCOL]; s[3] = L.sc[S_SM];
        s[4] = L.sc[S_SBK]; s[5] = L.sc[S_BLR]; s[6] = L.sc[S_PHI]; s[7] = L.sc[S_AV];
    }
}

__global__ __launch_bounds__(NT) void rhs_kernel(const DevNet *__restrict__ Np, PieceArgs A)
{
    const DevNet &N = *Np;   // tables in global memory: field loads stay scalar, nothing copied to scratch
    extern __shared__ double lds_raw[];
    Lds L;
    lds_layout(N, L, lds_raw, (double *)((char *)A.ws + (size_t)blockIdx.x * A.ws_stride + A.ws_mat_bytes + A.ws_jac_bytes));
    const int p = blockIdx.x;
    if (p >= A.n) return;
    piece_setup(N, L, A, p);
    block_rhs(N, L);
    for (int i = tid(); i < N.neq; i += NT) A.y_out[(size_t)p * N.neq + i] = L.f[i];
    if (A.rate_out)
        for (int k = tid(); k < N.nreac; k += NT) A.rate_out[(size_t)p * N.nreac + k] = L.rate[k];
    if (A.flux_out)   // the fluxes stay in the first nreac slots of the region after block_rhs
        for (int k = tid(); k < N.nreac; k += NT) A.flux_out[(size_t)p * N.nreac + k] = L.flux[k];
}

__global__ __launch_bounds__(NT) void jac_kernel(const DevNet *__restrict__ Np, PieceArgs A)
{
    const DevNet &N = *Np;   // tables in global memory: field loads stay scalar, nothing copied to scratch
    extern __shared__ double lds_raw[];
    Lds L;
    lds_layout(N, L, lds_raw, (double *)((char *)A.ws + (size_t)blockIdx.x * A.ws_stride + A.ws_mat_bytes + A.ws_jac_bytes));
    const int p = blockIdx.x;
    if (p >= A.n) return;
    piece_setup(N, L, A, p);
    block_rhs(N, L);   // leaves the pre-transfer surface total and the photo rates at y
    double *M = (double *)((char *)A.ws + (size_t)blockIdx.x * A.ws_stride);
    block_newton_matrix<double>(N, L, M, 0.0, false);   // J itself, before forming I - gamma J
    const int neq = N.neq;
    for (int idx = tid(); idx < neq * neq; idx += NT) {
        const int j = idx / neq, i = idx % neq;
        A.jac_out[(size_t)p * neq * neq + idx] = M[(size_t)j * N.ld + i];
    }
}

// LU + solve of M = I - gamma J: returns x = M^-1 b (b passed in y, x returned in y_out)
template <typename LUT>
__global__ __launch_bounds__(NT) void newton_solve_kernel(const DevNet *__restrict__ Np, PieceArgs A)
{
    const DevNet &N = *Np;   // tables in global memory: field loads stay scalar, nothing copied to scratch
    extern __shared__ double lds_raw[];
    Lds L;
    lds_layout(N, L, lds_raw, (double *)((char *)A.ws + (size_t)blockIdx.x * A.ws_stride + A.ws_mat_bytes + A.ws_jac_bytes));
    const int p = blockIdx.x;
    if (p >= A.n) return;
    // y holds the state for J; the right-hand side is rate_in reinterpreted as [n][neq]
    piece_setup(N, L, A, p);
    block_rates(N, L);
    block_rhs(N, L);
    for (int i = tid(); i < N.neq; i += NT) L.g[i] = 1.0;
    __syncthreads();
    LUT *M = (LUT *)((char *)A.ws + (size_t)blockIdx.x * A.ws_stride);
    int bad;
    if (A.sparse) {
        double *J = (double *)((char *)A.ws + (size_t)blockIdx.x * A.ws_stride + A.ws_mat_bytes);
        sparse_jacobian(N, L, J);
        sparse_form_newton<LUT>(N, J, M, A.gamma, L.prm[P_ENFORCECHARGE] != 0.0);
        bad = sparse_lu<LUT>(N, M, L.flux, L.red, L.b);
    } else {
        block_newton_matrix<LUT>(N, L, M, A.gamma);
        bad = block_lu<LUT>(M, N.neq, N.ld, L.red, L.flux);
    }
    for (int i = tid(); i < N.neq; i += NT) L.b[i] = A.scal_out[(size_t)p * N.neq + i];
    __syncthreads();
    if (!bad) {
        if (A.sparse) sparse_lu_solve<LUT>(N, L, M, L.b, L.flux);
        else block_lu_solve<LUT>(M, N.neq, N.ld, L.b);
    }
    for (int i = tid(); i < N.neq; i += NT) A.y_out[(size_t)p * N.neq + i] = bad ? NAN : L.b[i];
}

template __global__ void solve_points_kernel<double>(const DevNet *, SolveArgs);
template __global__ void solve_points_kernel<float>(const DevNet *, SolveArgs);
template __global__ void newton_solve_kernel<double>(const DevNet *, PieceArgs);
template __global__ void newton_solve_kernel<float>(const DevNet *, PieceArgs);

}  // namespace ucg
