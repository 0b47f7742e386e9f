/*
 * uclchem_gpu.h -- C ABI of the MI355X batched chemistry solver (libuclchem_gpu.so).
 *
 * Drop-in boundary for UCLCHEM's per-point chemistry solve.  Each entry point replaces
 * one reference interface for a whole batch of independent grid points:
 *
 *   ucg_solve_batch      <- the solveAbundances time loop with points as independent grid
 *                           points: updateChemistry + DVODE_F90 per output interval
 *                           (src/fortran_src/wrap.f90:549-697, chemistry.f90:145-292,
 *                            dvode.f90:2084 DVODE_F90(F,NEQ,Y,T,TOUT,ITASK=1,ISTATE=1,OPTS))
 *   ucg_rates            <- calculateReactionRates (src/fortran_src/rates.f90:21) with the
 *                           ice bookkeeping that precedes it (chemistry.f90:167-200)
 *   ucg_rhs              <- chemistry::F -> odes::GETYDOT (chemistry.f90:294, odes.f90:6)
 *   ucg_jacobian         <- the Newton-matrix Jacobian that DVODE builds by finite
 *                           differences in DVJAC (dvode.f90:8326-8360); analytic here
 *   ucg_newton_solve     <- DVJAC's M = I - gamma J factorisation + DGESL solve
 *                           (dvode.f90:8372-8387, 11982-12206)
 *
 * Conventions: plain C types only; arrays are row-major in C (point-major: point p's
 * vector starts at p*len).  Host pointers unless the function name ends in _device.
 * Species/reaction indices are 0-based.  The parameter vector of one point has
 * UCG_NPARAM doubles in the order of the UCG_P_* constants.  Every function returns
 * 0 on success or a negative UCG_E_* code; ucg_last_error() describes the failure.
 * The library is re-entrant per network handle and keeps no per-point global state.
 */
#ifndef UCLCHEM_GPU_H
#define UCLCHEM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define UCG_NPARAM 64
/* point parameters (defaultparameters.f90 names) */
#define UCG_P_DENS 0          /* initialDens [cm^-3] */
#define UCG_P_TEMP 1          /* initialTemp [K] */
#define UCG_P_ZETA 2          /* zeta [1.3e-17 s^-1] */
#define UCG_P_RADFIELD 3      /* radfield [Habing] */
#define UCG_P_BASEAV 4        /* baseAv [mag] */
#define UCG_P_ROUT 5          /* rout [pc] */
#define UCG_P_RIN 6           /* rin [pc] */
#define UCG_P_FINALTIME 7     /* finalTime [yr] */
#define UCG_P_FREEFALL 8      /* freefall (0/1) */
#define UCG_P_FINALDENS 9     /* finalDens [cm^-3] */
#define UCG_P_RELTOL 10       /* reltol */
#define UCG_P_ABSFAC 11       /* abstol_factor */
#define UCG_P_ABSMIN 12       /* abstol_min */
#define UCG_P_MODEL 13        /* 0 cloud, 1 hot core, 2 collapse (collapse.f90) */
#define UCG_P_TEMPINDX 14     /* hot-core mass index 1..6 */
#define UCG_P_MAXTEMP 15      /* hot-core maxTemp [K] */
#define UCG_P_FREEZEFAC 16    /* freezeFactor */
#define UCG_P_ENDATFINALDENS 17
#define UCG_P_FREEFALLFAC 18  /* freefallFactor */
#define UCG_P_MXSTEP 19       /* MXSTEP per integrator call */
#define UCG_P_MAXSTEPS 20     /* total BDF steps a point may take (0: unbounded, the reference's
                                 semantics); a point that reaches it stops with UCG_E_TOO_MANY_FAILS,
                                 the code UCLCHEM reserves for integrations that cannot advance.
                                 No counterpart in the reference: a work bound for batch and serving
                                 callers whose grids hold a few pathologically stiff points. */
/* initial elemental abundances (defaultparameters.f90:70-90, applied at chemistry.f90:58-105) */
#define UCG_P_METALLICITY 21  /* scales every element heavier than He */
#define UCG_P_ION 22          /* 0: C atomic, 1: half ionised, 2: C+ (default) */
#define UCG_P_FH 23           /* fraction of H atomic; the rest in H2 */
#define UCG_P_FHE 24
#define UCG_P_FC 25
#define UCG_P_FO 26
#define UCG_P_FN 27
#define UCG_P_FS 28
#define UCG_P_FMG 29
#define UCG_P_FSI 30
#define UCG_P_FCL 31
#define UCG_P_FP 32
#define UCG_P_FFE 33          /* read by the reference's parser, not used by its start */
#define UCG_P_FF 34
#define UCG_P_FD 35
#define UCG_P_FLI 36
#define UCG_P_FNA 37
#define UCG_P_FPAH 38
#define UCG_P_F15N 39
#define UCG_P_F13C 40
#define UCG_P_F18O 41
/* desorption and photon constants (defaultparameters.f90:113-120) */
#define UCG_P_EBMAXH2 42
#define UCG_P_EBMAXCR 43
#define UCG_P_EBMAXUVCR 44
#define UCG_P_EPSILON 45
#define UCG_P_UV_YIELD 46
#define UCG_P_PHI 47          /* CR-desorption yield of the first rates call */
#define UCG_P_UVCREFF 48
#define UCG_P_OMEGA 49
/* switches, 0 or 1 (defaultparameters.f90:40-49) */
#define UCG_P_DESORB 50
#define UCG_P_H2DESORB 51
#define UCG_P_CRDESORB 52
#define UCG_P_UVDESORB 53
#define UCG_P_THERMDESORB 54
#define UCG_P_CRATTENUATION 55    /* cosmicRayAttenuation */
#define UCG_P_IONMODEL 56         /* 0: 'L', 1: 'H' */
#define UCG_P_IMPROVEDH2CRP 57    /* improvedH2CRPDissociation (needs UCG_P_CRATTENUATION) */
#define UCG_P_ENFORCECHARGE 58    /* enforceChargeConservation */
#define UCG_P_COLLAPSE 59         /* collapse mode for UCG_P_MODEL 2: 1 BE1.1, 2 BE4, 3 filament,
                                     4 ambipolar (collapse_mode, collapse.f90:1-7; model.py collapse) */
/* slots 60..63 are reserved (0) */

/* status codes (UCLCHEM's own, src/fortran_src/constants.f90:22-27) */
#define UCG_OK 0
#define UCG_E_PARAM -1
#define UCG_E_PHYSICS_INIT -2
#define UCG_E_UNRECOVERABLE -4
#define UCG_E_TOO_MANY_FAILS -5
#define UCG_E_TIMEPOINTS -6
/* library errors */
#define UCG_E_NODEVICE -100
#define UCG_E_HIP -101
#define UCG_E_ARG -102
#define UCG_E_HANDLE -103
/* per-point status of a point a cancelled batch did not finish (ucg_cancel) */
#define UCG_E_CANCELLED -10

/* Network tables as emitted by the MakeRates backend (uclchem_amd/makerates_hip.py).
 * All pointers are host pointers, copied to the device by ucg_network_load. */
typedef struct ucg_network_desc {
    int nspec, neq, nreac, nice, nsurfl, nbulkl, nfp, nswap;
    int nloss, nprod, nnz, ncontrib, naff, nsw;
    const double *alpha, *beta, *gama, *tmin, *tmax, *mass, *rstat;     /* rstat [nreac*5] */
    const int *rtype, *re, *pr, *ice_kind, *n_d, *ops, *fac, *nfac;       /* re [nreac*3], pr [nreac*4], ops [nreac*6], fac [nreac*10] */
    const signed char *extrapolate;
    const int *type_ranges;                                               /* [23*2] */
    const int *loss_ptr, *loss_idx, *prod_ptr, *prod_idx, *swap_list;
    const int *surface_list, *bulk_list, *bulk_partner_of_surface, *surface_partner_of_bulk; /* partner arrays [neq] */
    const int *freeze_partners, *is_surface, *is_bulk;                    /* is_* [neq] */
    const int *jnz_row, *jnz_col, *jnz_ptr, *jc_k, *jc_m, *jc_c;
    const int *aff_ptr, *aff_sp, *aff_cf, *sw_ptr, *sw_k, *sw_c;
    int special_species[25];   /* nh nh2 nhe nc ncx nn no nf np nmg nmgx nna nli npah nsx nsix nclx nd n18o n15n n13c nelec nco nsurface nbulk */
    int special_reactions[10]; /* H2Form_CT H2Form_ER H2Form_ERDes HFreeze EFreeze H2Freeze H2_hv CO_hv C_hv,
                                  H2_crp (-1 when the network has none) */
    /* static-pattern sparse Newton plan (uclchem_amd/newton_plan.py): sizes are
     * N ns T nsp nlev ntg npair nph nphs ngs; arrays as named there */
    int nw_sizes[10];
    const int *nw_perm, *nw_inv, *nw_sp_ptr, *nw_sp_col, *nw_sp_dpos, *nw_lev_ptr, *nw_lev_rows, *nw_ph_ptr;
    const int *nw_tg_slot, *nw_tg_ptr, *nw_tg_div, *nw_pr_l, *nw_pr_u, *nw_jnz_slot, *nw_diag_slot;
    const int *nw_ph_slot, *nw_tr_slot, *nw_s1c, *nw_s1r, *nw_s2c, *nw_s2r, *nw_gs_ptr, *nw_gs_slot;
    const int *nw_misc, *nw_nbrow, *nw_nsrow;
    /* positive ions (species names holding '+', chemistry.f90:114-120) */
    int nion;
    const int *ion_list;
    /* grain-assisted recombination parameters, 7 per GAR reaction in reaction order
       (network.f90 garParams, rates.f90:316-332) */
    int ngar;
    const double *gar_params;
    /* charge row of the sparse plan, used by points with enforceChargeConservation
       (chemistry.f90:333-335): the nw_nerow plan slots of the electron row, cleared, and
       the nion slots of its positive-ion entries (ion_list order) */
    int nw_nerow;
    const int *nw_erow, *nw_eion;
} ucg_network_desc;

/* Select the GPU and upload the constant photo-rate tables.  Called implicitly. */
int ucg_init(int device);
const char *ucg_last_error(void);

int ucg_network_load(const ucg_network_desc *desc, int *handle);
/* Same from the flat descriptor file MakeRates' HIP backend writes beside the network
 * (uclchem_amd/networks/<name>.ucgnet), for callers without Python: the Fortran core's
 * counterpart of USE network (src/fortran_src/network.f90:1-5169). */
int ucg_network_load_file(const char *path, int *handle);
int ucg_network_free(int handle);

/* Full single-point models for a batch of independent grid points.
 *   params  [n][UCG_NPARAM]; y_init [n][neq] or NULL for UCLCHEM's elemental start
 *   y_out   [n][neq] final state (abundances + density)
 *   rec     [n][max_rec][4+nspec] output rows (time yr, density, gasTemp, Av, abundances)
 *           like the reference's physicsarray/chemicalabunarray, or NULL with max_rec=0
 *   n_rec   [n] rows written; status [n] UCLCHEM success flag per point
 *   stats   [n][8] steps, f-evals, Newton-matrix setups, Newton iterations,
 *           convergence failures, error-test failures, intervals, integrator retries
 *   lu_bits 64: FP64 Newton-matrix factorisation; 32: FP32 factorisation with the FP64
 *           residual (every Newton iterate is re-evaluated in FP64).
 * Returns when all points are done.  It waits for the whole device, so while a request
 * queue's kernel runs (ucg_queue_open) use ucg_solve_batch_device on a stream instead. */
int ucg_solve_batch(int handle, int n, const double *params, const double *y_init, double *y_out,
                    int max_rec, double *rec, int *n_rec, int *status, int *stats, int lu_bits);

/* ucg_solve_batch that also returns the rate coefficients of every output row: the
 * reference's returnRates with returnArray (ratesarray(dtime, dstep, :) = rate(:nreac),
 * src/fortran_src/io.f90:90-93; wrap.f90:15-80).  rates_rec [n][max_rec][nreac]: row r
 * holds the rates of the interval that ended at row r (the last attempt's coefficients,
 * with the H2 and CO photodissociation rates of its last RHS evaluation); row 0, written
 * before any interval, is zero.  rec and max_rec > 0 are required. */
int ucg_solve_batch_rates(int handle, int n, const double *params, const double *y_init, double *y_out,
                          int max_rec, double *rec, double *rates_rec, int *n_rec, int *status, int *stats,
                          int lu_bits);

/* Same on device-resident buffers (all pointers are device pointers, stream is a
 * hipStream_t or NULL).  t_start/t_done [n] (nullable) receive the device wall clock
 * (100 MHz ticks) when each point was taken and retired.  n_slots = resident
 * workgroups (0: two per compute unit).  Asynchronous with respect to the host.
 * Every buffer is checked to be device memory of the library's GPU, large enough for n
 * points (UCG_E_ARG otherwise); each point's parameter vector is checked on the device
 * and a point that fails the checks of ucg_solve_batch reports UCG_E_PARAM unintegrated. */
int ucg_solve_batch_device(int handle, int n, const double *params, const double *y_init, double *y_out,
                           int max_rec, double *rec, int *n_rec, int *status, int *stats,
                           double *t_start, double *t_done, int lu_bits, int n_slots, void *stream);

/* ucg_solve_batch_device that allocates nothing: the slot workspaces and the work counter
 * come from a fixed device array inside the library (80 MiB per GPU, set up by ucg_init and
 * held until the process ends), so the call makes no hipMalloc/hipFree and no allocating
 * library call.  No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * Arguments and results as ucg_solve_batch_device (results are identical); n_slots = 0 takes
 * as many resident workgroups as fit the array, up to two per compute unit; an explicit
 * n_slots beyond what fits, or a network one slot of which does not fit, is refused with
 * UCG_E_ARG.  Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (on different streams or threads) race, so order them on one stream. */
int ucg_solve_batch_device_noalloc(int handle, int n, const double *params, const double *y_init, double *y_out,
                                   int max_rec, double *rec, int *n_rec, int *status, int *stats,
                                   double *t_start, double *t_done, int lu_bits, int n_slots, void *stream);

/* Stop every solve launched on a handle before this call (ucg_solve_batch_device launches,
 * including ones queued behind other work, and the handle's request queues): workgroups
 * finish their current integrator step and take no further points; unfinished points
 * report UCG_E_CANCELLED and a stopped queue's unrun requests fail ucg_queue_test.
 * Launches issued after the call are not affected.  Callable from another host thread
 * while the solves run (not concurrently with ucg_network_free of the same handle).  No
 * counterpart in the reference (its grids run until every point ends); for serving loops
 * and time-bounded batches. */
int ucg_cancel(int handle);

/* ucg_solve_batch with per-point device timers [n][16]: seconds in rates, RHS, Jacobian
 * assembly, LU, triangular solves, total; the mean shader clock over the point's solve
 * in MHz; then LU sparse part, LU dense tail, solve sparse part, solve dense tail
 * (sparse plan only); RHS photo rates, totalSwap, fluxes, gather and totals; 1 reserved
 * (no output rows).  n_slots as above. */
int ucg_solve_batch_profiled(int handle, int n, const double *params, double *y_out, int *status, int *stats,
                             double *timers, int lu_bits, int n_slots);

/* One output interval of UCLCHEM's updateChemistry for a batch of points, each from its
 * own state: replaces one pass of the solveAbundances loop body's chemistry call
 * (src/fortran_src/chemistry.f90:145-239: ice bookkeeping and calculateReactionRates at
 * the interval start, integrateODESystem -> DVODE_F90 up to 10 attempts with the
 * reference's retry handling, chemistry.f90:241-292, dvode.f90:2084; the 1e-30 clamp).
 * No physics update afterwards (wrap.f90:672-676 is the caller's).  Host arrays:
 *   params [n][UCG_NPARAM] (tolerances, MXSTEP, model switches);
 *   phys   [n][8] = density, gasTemp, dustTemp, Av, cloudSize [cm], phi (the CR-desorption
 *          yield left by the previous rates call), abstol_factor, current time [s];
 *   y      [n][neq] state at the interval start; t_target [n] end of the interval [s];
 *   y_out  [n][neq]; phys_out [n][8] the same fields after the interval (phi and
 *          abstol_factor as the attempts left them, the time reached);
 *   status [n] (0, or UCLCHEM's error codes); stats [n][8] as for ucg_solve_batch.
 * lu_bits as for ucg_solve_batch. */
int ucg_interval(int handle, int n, const double *params, const double *phys, const double *y,
                 const double *t_target, double *y_out, double *phys_out, int *status, int *stats, int lu_bits);

/* Fine-grained kernels (host arrays).  phys [n][8] = density, gasTemp, dustTemp, Av,
 * cloudSize [cm], phi (the CR-desorption yield carried from the previous call),
 * abstol_factor, unused. */
int ucg_rates(int handle, int n, const double *params, const double *phys, const double *y,
              double *rate_out, double *y_out, double *scal_out);
int ucg_rhs(int handle, int n, const double *params, const double *phys, const double *rate,
            const double *y, double *ydot, double *rate_out);
int ucg_jacobian(int handle, int n, const double *params, const double *phys, const double *rate,
                 const double *y, double *jac);
/* ucg_rhs that also returns flux [n][nreac], the per-reaction terms of GETYDOT at y (RATE(k)
 * times its density and abundance factors and its type's bulk-layer, totalSwap or 1/SURFACE
 * factor; the H2 and CO photo rates recomputed at y as in F): the quantity the reference's
 * analysis.rates_to_dy_and_flux rebuilds from output rows (src/uclchem/analysis.py:737-800). */
int ucg_rhs_fluxes(int handle, int n, const double *params, const double *phys, const double *rate,
                   const double *y, double *ydot, double *flux);
int ucg_newton_solve(int handle, int n, const double *params, const double *phys, const double *y,
                     double gamma, const double *b, double *x, int lu_bits);

/* Linear solver of the Newton iterations for a handle (DVJAC/DGEFA/DGESL's job,
 * dvode.f90:8182-8390, 11982-12206):
 *   UCG_LINSOL_SPARSE (default): the static-pattern plan built with the network
 *     (fill-reducing order fixed at build time, bordered rank-one couplings, dense tail);
 *   UCG_LINSOL_DENSE: blocked LU of the full NEQ x NEQ matrix. */
#define UCG_LINSOL_SPARSE 0
#define UCG_LINSOL_DENSE 1
int ucg_set_linear_solver(int handle, int kind);

/* Network sizes of a loaded handle. */
int ucg_network_info(int handle, int *nspec, int *neq, int *nreac);

/* Request queues (serving).  The reference solves one model per call
 * (uclchemwrap.cloud, wrap.f90:15-80, driven from uclchem.model.cloud, model.py:277-289);
 * a queue keeps a persistent solve kernel resident on one network and admits single-point
 * requests while it runs: a workgroup that finishes a request takes the next posted one,
 * so requests arriving while others are on the GPU start as soon as a slot frees up.
 * Each request is a full model to params[UCG_P_FINALTIME] with the batch semantics
 * (status codes, counters, results bit-identical to ucg_solve_batch of the same point).
 *   ucg_queue_slot_bytes: workspace of one slot (for admission by HBM budget);
 *   ucg_queue_open: capacity = requests the queue will ever take, n_slots resident
 *     workgroups (<= CUs; queues open together should not exceed the CU count in total),
 *     idle_timeout_s: a workgroup waiting this long for a request leaves (bounds the kernel
 *     if the caller never closes); the queue's workspace and pinned buffers are its own;
 *   ucg_queue_submit: posts one request (params [UCG_NPARAM]), returns its index;
 *   ucg_queue_test: *done = 1 once the request's results are available (error if the
 *     kernel left before running it);
 *   ucg_queue_result: y [neq], status, stats [8] of a finished request;
 *   ucg_queue_close: admits nothing more, waits for posted requests, frees the queue. */
int ucg_queue_slot_bytes(int handle, int lu_bits, size_t *bytes);
int ucg_queue_open(int handle, int capacity, int n_slots, int lu_bits, double idle_timeout_s, int *queue);
int ucg_queue_submit(int queue, const double *params, int *request);
int ucg_queue_test(int queue, int request, int *done);
int ucg_queue_result(int queue, int request, double *y, int *status, int *stats);
int ucg_queue_close(int queue);

/* Shared request queue over several resident networks (no reference counterpart: the
 * reference serves one model per call, wrap.f90:15-80).  Its n_slots resident workgroups take
 * the next posted request whatever its network, so the capacity follows the load instead of
 * being split per network: a burst on one network runs on every slot.  Each slot's workspace
 * and the launch's LDS are sized for the largest network of the queue; the networks must use
 * the same linear solver; ucg_cancel on handles[0] stops the queue.
 *   ucg_queue_open_shared: handles[n_networks] (at most 64), then as ucg_queue_open;
 *   ucg_queue_submit_to: posts a request for network handles[network] (ucg_queue_submit
 *   takes requests only on a one-network queue);
 *   ucg_queue_test / _result / _close as for a one-network queue (y [neq of the request's
 *   network]). */
int ucg_queue_open_shared(const int *handles, int n_networks, int capacity, int n_slots, int lu_bits,
                          double idle_timeout_s, int *queue);
int ucg_queue_submit_to(int queue, int network, const double *params, int *request);

/* When a finished request was taken by a workgroup and when it finished, in seconds of the
 * device's constant-rate wall clock (differences between requests of one queue are
 * meaningful; no reference counterpart: serving latency and occupancy measurements). */
int ucg_queue_times(int queue, int request, double *t_taken, double *t_done);

#ifdef __cplusplus
}
#endif
#endif
