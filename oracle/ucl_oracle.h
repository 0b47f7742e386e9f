/* ORACLE / TEST INFRASTRUCTURE ONLY.  Never linked into the product library.
 *
 * CPU restatement (plain C, FP64) of the UCLCHEM per-point chemistry hot path:
 * initial abundances, per-interval rate coefficients, the RHS, the photo-rate
 * self-shielding functions and the output schedules.  Each function cites the
 * reference file:line it restates.  Network tables come from the MakeRates
 * backend (uclchem_amd/makerates_hip.py) through ucl_net; they are pinned
 * bit-for-bit against the compiled reference network by the tests.
 */
#ifndef UCL_ORACLE_H
#define UCL_ORACLE_H
#include <stdint.h>

#define UCL_NTYPES 23
#define UCL_MAXOPS 6

/* reaction type codes, reaction.py:5-28 order, TWOBODY last */
enum { T_PHOTON, T_CRP, T_CRPHOT, T_FREEZE, T_DESORB, T_THERM, T_DESOH2, T_DESCR, T_DEUVCR,
       T_H2FORM, T_ER, T_ERDES, T_LH, T_LHDES, T_BULKSWAP, T_SURFSWAP, T_IONOPOL1, T_IONOPOL2,
       T_CRS, T_EXSOLID, T_EXRELAX, T_GAR, T_TWOBODY };

/* point-parameter vector (same layout as oracle/ref_driver.f90) */
enum { P_DENS, P_TEMP, P_ZETA, P_RADFIELD, P_BASEAV, P_ROUT, P_RIN, P_FINALTIME, P_FREEFALL,
       P_FINALDENS, P_RELTOL, P_ABSFAC, P_ABSMIN, P_MODEL, P_TEMPINDX, P_MAXTEMP, P_FREEZEFAC,
       P_ENDATFINALDENS, P_FREEFALLFAC, P_MXSTEP, P_MAXSTEPS,
       /* initial elemental abundances (defaultparameters.f90:70-90) */
       P_METALLICITY, P_ION, P_FH, P_FHE, P_FC, P_FO, P_FN, P_FS, P_FMG, P_FSI, P_FCL, P_FP, P_FFE, P_FF,
       P_FD, P_FLI, P_FNA, P_FPAH, P_F15N, P_F13C, P_F18O,
       /* desorption and photon constants (defaultparameters.f90:113-120) */
       P_EBMAXH2, P_EBMAXCR, P_EBMAXUVCR, P_EPSILON, P_UV_YIELD, P_PHI, P_UVCREFF, P_OMEGA,
       /* switches (defaultparameters.f90:40-49) */
       P_DESORB, P_H2DESORB, P_CRDESORB, P_UVDESORB, P_THERMDESORB, P_CRATTEN, P_IONMODEL, P_IMPROVEDH2CRP,
       P_ENFORCECHARGE, P_NPARAM = 64 };

typedef struct {
    int32_t nspec, neq, nreac, nice, nsurfl, nbulkl, nfp, nswap;   /* list lengths */
    const double *mass, *alpha, *beta, *gama, *tmin, *tmax, *redmass;
    const double *binding_energy, *formation_enthalpy;       /* [nice] */
    const int64_t *n_atoms;                                   /* [nspec] */
    const int32_t *re, *pr;                                   /* [nreac*3], [nreac*4], -1 = none */
    const int32_t *rtype;
    const int8_t *extrapolate;
    const int32_t *type_ranges;                               /* [UCL_NTYPES*2], -1 = absent */
    const int32_t *ice_list, *gas_ice_list, *surface_list, *bulk_list;
    const int32_t *bulk_partner_of_surface;                   /* [nsurf] */
    const int32_t *freeze_partners;                           /* [nfp] */
    const int32_t *n_d, *ops;                                 /* [nreac], [nreac*UCL_MAXOPS] */
    const int32_t *loss_ptr, *loss_idx, *prod_ptr, *prod_idx, *swap_list;
    /* special species (0-based; missing species map to nspec, the density slot) */
    int32_t nh, nh2, nhe, nc, ncx, nn, no, nf, np, nmg, nmgx, nna, nli, npah, nsx, nsix, nclx,
        nd, n18o, n15n, n13c, nelec, nco, nsurface, nbulk, ngn, ngo, ngoh;
    /* special reactions (0-based) */
    int32_t r_h2form_ct, r_h2form_erdes, r_h2form_er, r_hfreeze, r_efreeze, r_h2freeze,
        r_h2_hv, r_co_hv, r_c_hv;
    int32_t r_h2_crp;           /* H2 + CRP -> H + H, or -1 */
    const double *gar_params;   /* [ngar*7], one row per GAR reaction */
    int32_t ngar;
    const int32_t *ion_list;    /* [nion] positive ions, names holding '+' (chemistry.f90:114-120) */
    int32_t nion;
} ucl_net;

/* per-point physical state carried between intervals */
typedef struct {
    double p[P_NPARAM];
    double density, gas_temp, dust_temp, av, cloud_size, coldens;
    double phi;                 /* CR desorption yield; leaks between calls (rates.f90:118,322-326) */
    double h2col, cocol, ccol;  /* set by ucl_rates, reused by ucl_rhs */
    double safe_mantle, safe_bulk, blr;
    double h2crp;               /* H2 CR dissociation rate the physics update sets (physics-core.f90:143-155) */
} ucl_point;

int ucl_point_init(const ucl_net *net, const double *params, ucl_point *pt, double *y0);
int ucl_rates(const ucl_net *net, ucl_point *pt, double *y, double *rate, double *scal);
int ucl_rhs(const ucl_net *net, const ucl_point *pt, double *rate, const double *y, double *ydot);
int ucl_fluxes(const ucl_net *net, const ucl_point *pt, double *rate, const double *y, double *flux);
double ucl_next_target(int model, double time_in_years);           /* seconds */
void ucl_update_physics(const ucl_net *net, ucl_point *pt, double current_time_s);

/* photo-rate helpers exposed for direct tests */
double ucl_h2_photodiss(double nh2, double radfield, double av, double turbvel);
double ucl_co_photodiss(double nh2, double nco, double radfield, double av);
double ucl_c_ionization(double alpha, double gamma, double t, double nc, double nh2, double av, double radfield);
double ucl_co_self_shielding(double nh2, double nco);
double ucl_scatter(double x1, double av);
double ucl_h2_form_efficiency(double gas_temp, double dust_temp);

#endif
