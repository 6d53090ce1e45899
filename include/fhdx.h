/*
 * fhdx.h -- C ABI of the MI355X (gfx950) compressible fluctuating-hydrodynamics step.
 *
 * Every entry point takes plain pointers and sizes (no AMReX, HIP or torch types), returns an
 * int error code (FHDX_OK == 0) and launches asynchronously on the HIP stream passed as
 * `void* stream` (NULL = the null stream).  All arrays are FP64 device pointers laid out like
 * an AMReX FArrayBox / Array4: x fastest, then y, then z, then component.
 *
 * Index conventions (reference: src_compressible/main_driver.cpp:160-297), for a valid box
 * [lo_d, hi_d] of the domain [0, n_d):
 *   cell data  : indices [lo_d - ngc, hi_d + ngc] per direction (ngc ghost layers);
 *   face data  : direction d nodal, indices [lo_d, hi_d + 1] in d and [lo_t, hi_t] across;
 *   node data  : [lo_d, hi_d + 1] in every direction, no ghosts.
 * fhdx_cons_to_prim, fhdx_transport, fhdx_set_bc, fhdx_calculate_flux, fhdx_fill_noise,
 * fhdx_chemistry_source, fhdx_fill_chem_noise and fhdx_rk3_stage_ref take any box of the domain
 * (one FAB of a MultiFab split into boxes, main_driver.cpp:90-117); the other entry points take
 * the whole domain as one box (lo = 0, hi = n - 1) and fill its periodic ghosts themselves.
 * Component maps (reference: main_driver.cpp:172-190, 248-260; transCoeffs.cpp:76-80):
 *   cons  : 0 rho, 1-3 momentum, 4 rho*E, 5.. rho*Y_k                         (nvars = 5+NS)
 *   prim  : 0 rho, 1-3 velocity, 4 T, 5 p, 6.. Y_k, 6+NS.. X_k                 (nprimvars = 6+2NS)
 *   flux  : 0..nvars-1 as cons, nvars heat, +1 viscous heating (diag), +2 (shear), +3 Dufour
 *   D     : component ll*NS+kk, multiplied by rho
 */
#ifndef FHDX_H
#define FHDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FHDX_MAX_SPECIES 8
#define FHDX_MAX_REACTION 8
#define FHDX_ABI_VERSION 2

enum fhdx_status {
    FHDX_OK = 0,
    FHDX_ERR_ARG = 1,         /* bad argument or array shape */
    FHDX_ERR_UNSUPPORTED = 2, /* parameter combination outside the implemented scope */
    FHDX_ERR_HIP = 3,         /* HIP runtime failure (message in fhdx_last_error) */
    FHDX_ERR_MASSFRAC = 4,    /* Y_k <= 0 or >= 1 met in the transport step: the reference aborts
                                 there (src_compressible/transCoeffs.cpp:45-46) */
    FHDX_ERR_NOMEM = 5,       /* device memory budget exhausted */
    FHDX_ERR_NONFINITE = 6,   /* non-finite state detected */
    FHDX_ERR_BUSY = 7,        /* serving engine: no free slot for the request */
    FHDX_ERR_SINGULAR = 8     /* transport_type 2: a singular matrix in the Valk-Waldmann solves
                                 (the reference aborts, compressible_functions.H:251-253) */
};

/* Run parameters: the values the reference reads from its AMREX_GPU_MANAGED namespaces
 * (src_common/common_namespace.H:24-115, src_compressible/compressible_namespace.H:1-15). */
typedef struct fhdx_params {
    int32_t nspecies;          /* NS, 2..FHDX_MAX_SPECIES (one gas: two identical species)  */
    int32_t n_cells[3];        /* domain cells per direction                                 */
    int32_t ngc;               /* ghost layers of cell data (>= 2; <= n_cells/2 at walls)    */
    int32_t algorithm_type;    /* 2: multispecies (species fluxes on), else species off      */
    int32_t advection_type;    /* 0: none, 1: interpolate prim, 2: interpolate cons          */
    int32_t visc_type;         /* |3| adds bulk viscosity zeta = zeta_ratio*eta              */
    int32_t transport_type;    /* 1: Giovangigli, 2: Valk-Waldmann, 3: Hirschfelder-Curtiss-
                                  Bird (binary)                                              */
    int32_t stoch_stress_form; /* 1: stochastic fluxes on, 0: deterministic                  */
    int32_t nreaction;         /* 0, or the reaction count given to fhdx_ctx_set_chemistry   */
    /* Boundary codes per direction (common_namespace.H, boundary.cpp:172-2001).  bc_vel -1:
     * periodic (then every code of that direction is -1, SetupBC boundary.cpp:3-12); 1: slip
     * wall, 2: no-slip wall.  bc_mass 1: wall (no species flux), 2: fixed composition bc_Yk,
     * 3: reservoir (x faces only; the reference's y/z reservoir ghosts keep stale velocity and
     * energy, boundary.cpp:648-649, 806-807).  bc_therm 1: adiabatic, 2: isothermal at t_lo/t_hi.
     * Walls are served by the reference-shaped entry points; the fused step, the ensemble
     * engine and the server take periodic boxes only. */
    int32_t bc_vel_lo[3], bc_vel_hi[3];
    int32_t bc_mass_lo[3], bc_mass_hi[3];
    int32_t bc_therm_lo[3], bc_therm_hi[3];
    int32_t pad0;
    double dx[3];              /* cell sizes [cm]                                            */
    double molmass[FHDX_MAX_SPECIES];  /* g/mol                                              */
    double diameter[FHDX_MAX_SPECIES]; /* cm                                                 */
    double dof[FHDX_MAX_SPECIES];
    double hcv[FHDX_MAX_SPECIES];      /* < 0: derived from dof (compressible_functions.cpp:91-100) */
    double hcp[FHDX_MAX_SPECIES];
    double e0[FHDX_MAX_SPECIES];
    double k_B;                /* 1.38064852e-16 by default                                  */
    double Runiv;              /* 8.314462175e7 by default                                   */
    double variance_coef_mom, variance_coef_ener, variance_coef_mass;
    double grav[3];
    double zeta_ratio;
    double t_lo[3], t_hi[3];   /* isothermal wall temperatures [K]                           */
    /* Wall composition of bc_mass 2 faces, face f = 2*direction + (0 lo, 1 hi): the reference's
     * bc_Yk_x_lo .. bc_Yk_z_hi and bc_Xk_*.  fhdx_params_finalize fills whichever of the two sums
     * to 0 from the other (SetupCWall, boundary.cpp:14-170). */
    double bc_Yk[6][FHDX_MAX_SPECIES];
    double bc_Xk[6][FHDX_MAX_SPECIES];
    /* Reservoirs (bc_mass 3, x faces only): pressure, density and temperature (t_lo/t_hi) of the
     * reservoir gas.  p must be given (the reference's default, the ambient pressure of the
     * initial state, is computed by the AMReX adapter); fhdx_params_finalize fills rho <= 0
     * from p, T and the wall composition, or T <= 0 from p and rho (SetupCWall,
     * boundary.cpp:38-58, 76-96). */
    double p_lo[3], p_hi[3];
    double rho_lo[3], rho_hi[3];
} fhdx_params;

/* One FAB: data pointer plus the inclusive index box it covers (ghosts included). */
typedef struct fhdx_fab {
    double* p;
    int32_t lo[3];
    int32_t hi[3];
    int32_t ncomp;
    int32_t pad;
} fhdx_fab;

typedef struct fhdx_ctx fhdx_ctx;

/* ---- library / parameters -------------------------------------------------------------- */
int fhdx_abi_version(void);
const char* fhdx_status_string(int status);
/* Text of the last error raised on this host thread ("" if none). */
const char* fhdx_last_error(void);

/* Reference defaults of src_common/common_functions.cpp:320-490 with the compressible
 * periodic settings every reference input uses (ngc 2, algorithm/visc/advection type 2). */
void fhdx_params_default(fhdx_params* p);
/* GetHcGas (src_compressible/compressible_functions.cpp:91-100): hcv/hcp < 0 -> from dof;
 * SetupBC (boundary.cpp:3-12) and the wall compositions of SetupCWall (boundary.cpp:14-170). */
void fhdx_params_finalize(fhdx_params* p);
/* 0 if the parameters are inside the implemented scope, else an fhdx_status.  Cell counts:
 * at least 1 per periodic direction and 2 ngc per wall direction; the fused step, the
 * ensemble engine and the server further need max(4, 2 ngc) per direction and no walls. */
int fhdx_params_check(const fhdx_params* p);

/* A context owns per-run device workspace; one per host thread.  Creating it selects nothing:
 * the current HIP device is used. */
int fhdx_ctx_create(const fhdx_params* p, fhdx_ctx** out);
int fhdx_ctx_destroy(fhdx_ctx* ctx);
/* Waits for the work queued on `stream`, then reports (and clears) the device-side error
 * flags raised by earlier calls on this context (e.g. FHDX_ERR_MASSFRAC). */
int fhdx_ctx_check(fhdx_ctx* ctx, void* stream);

/* ---- reference-shaped entry points ------------------------------------------------------ */

/* conservedToPrimitive (src_compressible/conservedPrimitiveConversions.cpp:4-92): valid cells. */
int fhdx_cons_to_prim(fhdx_ctx* ctx, fhdx_fab prim, fhdx_fab cons, void* stream);

/* calculateTransportCoeffs (src_compressible/transCoeffs.cpp:7-90), transport_type 1, 2 or 3, on
 * the box grown by ngc.  A mass fraction outside (0,1) raises FHDX_ERR_MASSFRAC, a singular
 * matrix in the type-2 solves FHDX_ERR_SINGULAR (see ctx_check). */
int fhdx_transport(fhdx_ctx* ctx, fhdx_fab prim, fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa,
                   fhdx_fab chi, fhdx_fab D, void* stream);

/* MultiFab::FillBoundary(periodicity) for a single-box domain (all components): every ghost cell
 * whose periodic image is a valid cell gets that image; ghost cells outside a wall are left
 * alone.  A domain split into several FABs exchanges ghosts between them (AMReX's FillBoundary). */
int fhdx_fill_boundary(fhdx_ctx* ctx, fhdx_fab mf, void* stream);

/* setBC (src_compressible/boundary.cpp:172-1209): the physical ghost cells of prim and cons at
 * walls, face by face in the order lo x, hi x, lo y, hi y, lo z, hi z.  No-op when periodic. */
int fhdx_set_bc(fhdx_ctx* ctx, fhdx_fab prim, fhdx_fab cons, void* stream);

/* calculateFlux (src_compressible/flux.cpp:4-2376; signature compressible_functions.H:23-35):
 * stochastic, diffusive, corner and hyperbolic face fluxes into flux[3] (nvars+4 comps),
 * the corner arrays and visccorn.  stochFlux/rancorn hold the already weighted noise.  At walls
 * the one-sided face branches of flux.cpp, StochFlux (boundary.cpp:1392-1970) and
 * BCWallSpeciesFlux (boundary.cpp:1212-1390) apply. */
int fhdx_calculate_flux(fhdx_ctx* ctx, fhdx_fab cons, fhdx_fab prim, fhdx_fab eta,
                        fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi, fhdx_fab D,
                        const fhdx_fab flux[3], const fhdx_fab stochFlux[3],
                        const fhdx_fab cornx[3], const fhdx_fab corny[3],
                        const fhdx_fab cornz[3], fhdx_fab visccorn, fhdx_fab rancorn,
                        double dt, void* stream);

/* The noise of one step (replaces the MultiFabFillRandom calls of timeStep.cpp:75-91):
 * N(0,1)*|variance_coef| from a counter-based Philox4x32-10 keyed by (seed, step, instance),
 * set A and set B, face components 1..nvars-1 and the nodal field.  Duplicate periodic
 * faces/nodes (index n_d == index 0) carry the same value, as after OverrideSync; in a wall
 * direction the faces and nodes at index 0 and n_d are distinct and draw their own values. */
int fhdx_fill_noise(fhdx_ctx* ctx, uint64_t seed, uint64_t step, uint32_t instance,
                    const fhdx_fab stochA[3], const fhdx_fab stochB[3], fhdx_fab rancornA,
                    fhdx_fab rancornB, void* stream);

/* RK3step (src_compressible/timeStep.cpp:11-427; signature compressible_functions.H:41-51)
 * with the reference's own arrays: cu is advanced in place, cup/cup2 are stage scratch, prim,
 * eta..D, flux, stochFlux, corners, visccorn and rancorn end as the reference leaves them.
 * Noise: if stochA is non-NULL the caller's A/B arrays are used (injected noise), otherwise
 * Philox noise for (seed, step, instance 0) is generated into ctx workspace.  At walls the
 * ghost cells after each stage come from setBC (timeStep.cpp:207-209, 312-314, 420-425); the
 * caller's cu/prim enter with their ghosts set (main_driver.cpp:535-542). */
int fhdx_rk3_step_ref(fhdx_ctx* ctx, fhdx_fab cu, fhdx_fab cup, fhdx_fab cup2, fhdx_fab prim,
                      fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi, fhdx_fab D,
                      const fhdx_fab flux[3], const fhdx_fab stochFlux[3],
                      const fhdx_fab cornx[3], const fhdx_fab corny[3], const fhdx_fab cornz[3],
                      fhdx_fab visccorn, fhdx_fab rancorn,
                      const fhdx_fab* stochA, const fhdx_fab* stochB,
                      const fhdx_fab* rancornA, const fhdx_fab* rancornB,
                      double dt, uint64_t seed, uint64_t step, void* stream);

/* One stage of RK3step (src_compressible/timeStep.cpp:99-202 stage 0, 211-307 stage 1, 316-415
 * stage 2) on ONE FAB of MultiFabs that may be split into several boxes (main_driver.cpp:90-117;
 * e.g. inputs_equil_3d's max_grid_size 32 32 64): the transport coefficients of the stage input
 * on the grown box, stochFlux/rancorn = A + w_s B (LinComb, timeStep.cpp:118-128), calculateFlux,
 * with chemistry ranchem = A + w_s B and the source (timeStep.cpp:135-143), the stage combination
 * (stage 0: cu -> cup, 1: cup -> cup2, 2: cup2 -> cu) and conservedToPrimitive of the output on
 * the valid cells.  It fills no ghost cell: between stages the caller runs
 * FillBoundary(periodicity) on the output state and prim across all FABs, then setBC
 * (fhdx_set_bc per FAB), as timeStep.cpp:204-209, 309-314 and 417-425 do; before stage 0 cu and
 * prim enter with their ghosts set (main_driver.cpp:535-542).  The A/B sets are filled once per
 * step per FAB by fhdx_fill_noise (and fhdx_fill_chem_noise): the noise is keyed by the global
 * face / node index, so a face two FABs share draws one value in both, which is what
 * OverrideSync leaves in the reference (MultiFabFillRandom.cpp:25-28).  Each FAB is the
 * reference's: cell arrays on the box grown by ngc, face arrays nodal in their direction, node
 * arrays nodal in all, no ghosts.  A box must lie in the domain and hold >= ngc cells across a
 * wall it touches.  source (nvars comps), ranchem and ranchemA/B (nreaction comps, cell FABs
 * with ngc ghosts) are read only when nreaction > 0 (pass zeroed fabs otherwise); the reference
 * fills source and ranchem inside RK3step (timeStep.cpp:135-143). */
int fhdx_rk3_stage_ref(fhdx_ctx* ctx, int stage, fhdx_fab cu, fhdx_fab cup, fhdx_fab cup2,
                       fhdx_fab prim, fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi,
                       fhdx_fab D, const fhdx_fab flux[3], const fhdx_fab stochFlux[3],
                       const fhdx_fab cornx[3], const fhdx_fab corny[3], const fhdx_fab cornz[3],
                       fhdx_fab visccorn, fhdx_fab rancorn,
                       const fhdx_fab stochA[3], const fhdx_fab stochB[3],
                       fhdx_fab rancornA, fhdx_fab rancornB,
                       fhdx_fab source, fhdx_fab ranchem, fhdx_fab ranchemA, fhdx_fab ranchemB,
                       double dt, void* stream);

/* evaluateStats (src_compressible/stats.cpp:5-363; called at main_driver.cpp:613): running
 * means consMean (nvars) and primMean (nprimvars), variances consVar (nvars) and primVar
 * (nprimvars+5), the cross-correlations against the x = cross_cell slice in spatialCross (8)
 * and miscStats (>= 6 comps), all cell FABs updated in place on the valid box; miscVals[20]
 * (host memory) receives the cross_cell slice averages.  steps is the running sample count
 * (statsCount, >= 1).  Per-cell arithmetic follows the reference's order; the yz-slice sums are
 * added in a fixed parallel order, so they match the reference's serial sums to rounding.
 * Waits for `stream` before returning (miscVals is on the host). */
int fhdx_evaluate_stats(fhdx_ctx* ctx, fhdx_fab cons, fhdx_fab consMean, fhdx_fab consVar,
                        fhdx_fab prim, fhdx_fab primMean, fhdx_fab primVar,
                        fhdx_fab spatialCross, fhdx_fab miscStats, double* miscVals, int steps,
                        int cross_cell, void* stream);

/* ---- fused fast path --------------------------------------------------------------------- */

/* One RK3 step of RK3step's arithmetic with fused stage kernels: cu (nvars comps, ngc ghosts)
 * advanced in place, prim (nprimvars comps) recomputed from the new cu; ghost layers of both
 * are filled periodically, as after timeStep.cpp:417-425.  Noise is generated in-kernel from
 * Philox (seed, step, instance 0), bit-identical to fhdx_fill_noise for the same key.
 * Periodic boxes of at least max(4, 2 ngc) cells per direction only (walls, and thinner
 * boxes such as the quasi-2-D n_cells = nx ny 1 decks: FHDX_ERR_UNSUPPORTED; use
 * fhdx_rk3_step_ref, whose periodic ghosts are images any number of periods away). */
int fhdx_rk3_step(fhdx_ctx* ctx, fhdx_fab cu, fhdx_fab prim, double dt, uint64_t seed,
                  uint64_t step, void* stream);

/* As fhdx_rk3_step but the A/B noise is read from caller arrays (parity runs).  Not with
 * chemistry (its noise has no array here): FHDX_ERR_UNSUPPORTED. */
int fhdx_rk3_step_injected(fhdx_ctx* ctx, fhdx_fab cu, fhdx_fab prim,
                           const fhdx_fab stochA[3], const fhdx_fab stochB[3],
                           fhdx_fab rancornA, fhdx_fab rancornB, double dt, void* stream);

/* No counterpart in the reference: fhdx_rk3_step for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * arithmetic and bits as fhdx_rk3_step; its workspace (two stage states, cell records, node
 * terms, face fluxes: (2 nvars + (8 + 4 NS + NS^2) + 14 + 3 nvars) x (n_0+2 ngc)(n_1+2 ngc)
 * (n_2+2 ngc) doubles) lives in a file-scope device array of 64 MiB per GPU, held until the
 * process ends and looked up per device by fhdx_ctx_create, so the call allocates nothing.  A
 * box whose workspace exceeds 64 MiB (45^3 cells hold with two species, 39^3 with four) returns
 * FHDX_ERR_UNSUPPORTED.  The context's device must be current.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (two streams, two threads) race and give
 * wrong results; serialise them.  Nothing needs clearing between calls (each step writes every
 * workspace value before reading it). */
int fhdx_rk3_step_noalloc(fhdx_ctx* ctx, fhdx_fab cu, fhdx_fab prim, double dt, uint64_t seed,
                          uint64_t step, void* stream);

/* ---- chemistry: the +dt*S source of the RK3 stages ------------------------------------ */
/* Reaction data of src_chemistry/chemistry_namespace.H:1-39 as compute_compressible_chemistry_
 * source_CLE (chemistry_functions.cpp:128-193) uses it: law-of-mass-action rates
 * k_m exp(-alpha_m/Runiv (1/T - 1/T0)) (T/T0)^beta_m prod_n c_n^R(m,n) with molar
 * concentrations c_n = X_n p/(Runiv T), and for reaction_type 1 the chemical-Langevin noise
 * sqrt(rate/N_A) W, W = ranchem/sqrt(dt dV), ranchem = A + w_s B per stage like the flux noise. */
typedef struct fhdx_chemistry {
    int32_t nreaction;       /* 1..FHDX_MAX_REACTION; must equal fhdx_params.nreaction       */
    int32_t reaction_type;   /* 0: deterministic rates, 1: chemical Langevin (CLE)           */
    int32_t stoich_R[FHDX_MAX_REACTION][FHDX_MAX_SPECIES]; /* reactants (stoich_<m>R), >= 0   */
    int32_t stoich_P[FHDX_MAX_REACTION][FHDX_MAX_SPECIES]; /* products  (stoich_<m>P), >= 0   */
    double rate_const[FHDX_MAX_REACTION];
    double alpha_param[FHDX_MAX_REACTION];
    double beta_param[FHDX_MAX_REACTION];
    double T0_chem;          /* > 0                                                          */
} fhdx_chemistry;

/* Sets the reactions of a context whose fhdx_params.nreaction > 0 (required before its first
 * step).  The steps then add dt * S to the species equations of every stage, S evaluated from
 * the stage's input state, and draw the reaction noise from the same Philox key as the flux
 * noise (field 64 + m).  exp and pow are evaluated with IEEE operations only (identical on host
 * and device); integer powers of concentrations by repeated multiplication. */
int fhdx_ctx_set_chemistry(fhdx_ctx* ctx, const fhdx_chemistry* chem);

/* compute_compressible_chemistry_source_CLE (chemistry_functions.cpp:128-193): source
 * components 5..5+NS-1 of the valid cells from prim (nprimvars) and ranchem (nreaction comps,
 * the already weighted noise); other components are not touched. */
int fhdx_chemistry_source(fhdx_ctx* ctx, fhdx_fab prim, fhdx_fab source, fhdx_fab ranchem,
                          double dt, void* stream);

/* The reaction noise sets A and B (nreaction comps, valid cells) of Philox key
 * (seed, step, instance): the values the steps evaluate in place. */
int fhdx_fill_chem_noise(fhdx_ctx* ctx, uint64_t seed, uint64_t step, uint32_t instance,
                         fhdx_fab ranchemA, fhdx_fab ranchemB, void* stream);

/* ---- launch-shape tuning ---------------------------------------------------------------- */

/* Launch shape of the fast path: tile width in x (16, 32 or 64 cells; the tile has 256 cells),
 * tiles per region of the region-major block order, how many z levels a tile block marches
 * through, whether the Philox noise is evaluated in place by every stage (1) or written once per
 * step to arrays (0; always so for more than five species), and an optional cap on resident face-kernel blocks per CU (fewer blocks
 * in flight keep more of their neighbours' data in L2). */
typedef struct fhdx_tune {
    int32_t tile_width;
    int32_t region_tiles;     /* 1..64                                                       */
    int32_t z_levels;         /* 1..64                                                       */
    int32_t noise_inline;     /* 0 or 1                                                      */
    int32_t face_blocks_per_cu; /* 0: as many as registers allow; 2 or 3: cap (reserves LDS) */
    int32_t records_in_update;  /* 1: stages 1-2 write the next stage's cell records in the
                                   update kernel (one cell-kernel launch per step, not 3)    */
    double ms_per_step;       /* measured by fhdx_ctx_autotune (0 when set by hand)          */
} fhdx_tune;

/* Times candidate launch shapes on trial_steps fast-path steps of a scratch copy of cu (cu is
 * not modified; HIP events on `stream`): one coordinate at a time (tile width, region height,
 * z levels, noise, face blocks per CU, records in update), each swept with the others at
 * their best so far.  Keeps the
 * fastest for this context and returns it in *best; trials (may be NULL) receives up to max_trials timed candidates, n_trials their
 * count.  The choice is cached per (species count, box shape) for the process, so later
 * contexts of the same shape reuse it without timing (then n_trials = 0). */
int fhdx_ctx_autotune(fhdx_ctx* ctx, fhdx_fab cu, double dt, int trial_steps, fhdx_tune* best,
                      fhdx_tune* trials, int max_trials, int* n_trials, void* stream);
int fhdx_ctx_set_tune(fhdx_ctx* ctx, const fhdx_tune* cfg);
int fhdx_ctx_get_tune(fhdx_ctx* ctx, fhdx_tune* cfg);

/* ---- measurement ----------------------------------------------------------------------- */

/* Per-kernel timing of the fast path.  While enabled, fhdx_rk3_step brackets each of its kernel
 * groups (noise, cells, nodes, faces, update, ghosts) with HIP events recorded on the launch
 * stream; nothing synchronises inside the step.  Enabling (re)sets the totals. */
int fhdx_ctx_timing_enable(fhdx_ctx* ctx, int enable);
/* Number of kernel groups that fhdx_ctx_timing_read reports. */
int fhdx_timing_kinds(void);
/* Waits for the recorded events, then returns group `kind`'s name, timed launches and summed
 * milliseconds. */
int fhdx_ctx_timing_read(fhdx_ctx* ctx, int kind, char* name, int name_len, int64_t* launches,
                         double* total_ms);

/* ---- ensembles: many independent boxes advanced together -------------------------------- */

/* The reference advances one box per process (main_driver.cpp:574 calls RK3step once per
 * step); these calls have no reference counterpart.  They batch independent instances of the
 * same step (any box size, 2-4 species each) into shared launches on one stream: per kernel
 * kind and species count, one launch covers every admitted instance.  Instance results are
 * bit-identical to fhdx_rk3_step on the same (state, dt, seed, step): the noise of an instance
 * is keyed by its own (seed, step, instance 0). */
typedef struct fhdx_engine fhdx_engine;

/* HBM bytes an instance with these parameters holds while admitted. */
uint64_t fhdx_engine_bytes_needed(const fhdx_params* p);
/* At most max_instances slots; admission fails once the admitted instances would hold more
 * than mem_budget_bytes of HBM (FHDX_ERR_NOMEM) or no slot is free (FHDX_ERR_BUSY). */
int fhdx_engine_create(int max_instances, uint64_t mem_budget_bytes, fhdx_engine** out);
int fhdx_engine_destroy(fhdx_engine* e);
/* Admit an instance: cu_init (host or device pointer) holds its conserved state on the box
 * grown by ngc, component-major like a FAB.  The instance takes steps step0, step0+1, ... and
 * idles after nsteps of them (nsteps = 0: no limit).  Returns its slot. */
int fhdx_engine_add(fhdx_engine* e, const fhdx_params* p, double dt, uint64_t seed,
                    uint64_t step0, uint64_t nsteps, const double* cu_init, int* slot);
/* As fhdx_engine_add, with a uniform start: component c of every cell (ghosts included) is
 * values[c], c = 0..nvars-1. */
int fhdx_engine_add_uniform(fhdx_engine* e, const fhdx_params* p, double dt, uint64_t seed,
                            uint64_t step0, uint64_t nsteps, const double* values, int* slot);
/* Release a slot and its memory. */
int fhdx_engine_remove(fhdx_engine* e, int slot);
/* Queue nsteps steps of every admitted instance on the engine's stream.  use_graph != 0
 * captures one step of the current membership in a hipGraph (again after each add/remove)
 * and replays it. Returns without waiting. */
int fhdx_engine_step(fhdx_engine* e, int nsteps, int use_graph);
/* Wait for the queued steps. */
int fhdx_engine_sync(fhdx_engine* e);
/* Device FABs of a slot's conserved and primitive state (valid after fhdx_engine_sync), its
 * next step index and whether it has taken all its steps.  Any output may be NULL.  The engine
 * keeps only the conserved state between steps; asking for prim computes it from that state
 * (conservedToPrimitive, ghost cells included) before returning. */
int fhdx_engine_state(fhdx_engine* e, int slot, fhdx_fab* cu, fhdx_fab* prim, uint64_t* step,
                      int* done);
/* Wait for the queued steps, then copy a slot's grown conserved / primitive state to cu_out /
 * prim_out (host or device pointers; either may be NULL). */
int fhdx_engine_read(fhdx_engine* e, int slot, double* cu_out, double* prim_out);
/* Wait, then copy (and clear) the per-slot device error flags (max_instances ints; bit 0:
 * mass fraction outside (0,1), as FHDX_ERR_MASSFRAC). */
int fhdx_engine_errors(fhdx_engine* e, int* flags);
/* HBM bytes held by the admitted instances. */
uint64_t fhdx_engine_bytes_used(fhdx_engine* e);

/* ---- serving: continuous batching of arriving ensemble runs ----------------------------- */

/* No reference counterpart.  A worker thread owns an engine: submitted requests are admitted
 * in arrival order as the memory budget and a slot allow, every admitted instance advances in
 * the same launches (new arrivals join at the next chunk of at most chunk_steps steps), and a
 * finished instance's state is copied to the request's cu_out / prim_out before its completion
 * is posted.  A
 * request whose instance raises an error flag is evicted alone.  Results are bit-identical to a
 * solo fhdx_rk3_step run with the same (state, dt, seed), steps 0..nsteps-1. */
typedef struct fhdx_server fhdx_server;

typedef struct fhdx_request {
    fhdx_params params;       /* box, species, physics (as for fhdx_ctx_create)            */
    double dt;
    uint64_t seed;            /* Philox key; steps 0..nsteps-1                              */
    uint64_t nsteps;          /* >= 1                                                       */
    const double* cu_init;    /* grown conserved start state (host or device), kept alive by
                                 the caller until completion; NULL: uniform start below     */
    double* cu_out;           /* host/device buffer for the grown final cu, or NULL; with
                                 cu_out and prim_out both NULL the final state is dropped   */
    double* prim_out;         /* grown final prim, or NULL                                  */
    double rho0, T0;          /* uniform start: density, temperature ...                    */
    double Y0[FHDX_MAX_SPECIES]; /* ... and mass fractions (InitConsVar prob_type 1)       */
} fhdx_request;

typedef struct fhdx_completion {
    int64_t id;               /* as returned by fhdx_server_submit                          */
    int32_t status;           /* FHDX_OK, or why the request failed / was evicted           */
    int32_t pad;
    int64_t arrival_ns;       /* steady-clock times: submission, admission (0: never) ...   */
    int64_t admit_ns;
    int64_t done_ns;          /* ... and result copied to cu_out / prim_out                 */
    uint64_t steps;           /* steps taken                                                */
    char message[96];
} fhdx_completion;

/* max_instances slots, mem_budget_bytes of HBM for admitted instances, at most chunk_steps
 * steps between admission points, use_graph != 0 to replay hipGraph-captured steps. */
int fhdx_server_create(int max_instances, uint64_t mem_budget_bytes, int chunk_steps,
                       int use_graph, fhdx_server** out);
/* Thread-safe.  Records the arrival time and queues the request. */
int fhdx_server_submit(fhdx_server* s, const fhdx_request* r, int64_t* id);
/* Admission order of waiting requests: arrival (FIFO, the default) or smallest work
 * (cells x nsteps) first; max_active_cells > 0 also keeps new admissions waiting while the
 * admitted instances already hold that many cells (one is always admitted onto an idle GPU).
 * Smallest-first can delay a large request for as long as smaller ones keep arriving. */
#define FHDX_ADMIT_ARRIVAL 0
#define FHDX_ADMIT_SMALLEST_FIRST 1
int fhdx_server_set_policy(fhdx_server* s, int policy, uint64_t max_active_cells);
/* Moves up to max completions out (n = how many). */
int fhdx_server_poll(fhdx_server* s, fhdx_completion* out, int max, int* n);
/* Waits until every submitted request has completed or timeout_s passes (then FHDX_ERR_BUSY);
 * open = requests not yet completed. */
int fhdx_server_wait(fhdx_server* s, double timeout_s, int64_t* open);
/* Steps taken by the worker, cell-updates done (summed over instances), HBM bytes held. */
int fhdx_server_stats(fhdx_server* s, uint64_t* steps, uint64_t* cell_updates,
                      uint64_t* bytes_used);
/* Stops the worker; requests not finished complete as FHDX_ERR_BUSY ("server stopped"). */
int fhdx_server_destroy(fhdx_server* s);

/* ---- on-disk formats: AMReX plotfile and checkpoint ------------------------------------ */
/* The MultiFabs WritePlotFile / WriteCheckPoint / ReadCheckPoint take
 * (src_compressible/writePlotFile.cpp:7-18, checkpoint.cpp:22-36 and 135-149), each a cell
 * array of the run's box with ngc ghosts.  Component counts: cu, cuMeans, cuVars nvars; prim,
 * primMeans nprimvars; primVars nprimvars+5; spatialCross 8; miscStats 10 (checkpoint only);
 * eta, kappa 1. */
typedef struct fhdx_fields {
    fhdx_fab cu, cuMeans, cuVars, prim, primMeans, primVars, spatialCross, miscStats, eta, kappa;
} fhdx_fields;

/* flags: the fabs are host memory (default: device memory, copied on `stream`) */
#define FHDX_IO_HOST_ARRAYS 1

/* WritePlotFile (writePlotFile.cpp:7-190): directory <plot_base_name><step, 9 digits> holding an
 * AMReX plotfile (Header, Level_0/Cell_H, Level_0/Cell_D_00000) of cu, prim, [cuMeans 0-4,
 * primMeans 0-5 if plot_means == 1], [cuVars 0-4, primVars 0-4 if plot_vars == 1],
 * spatialCross, eta, kappa with the reference's variable names; valid cells only.  prob_lo is
 * the domain's lower corner (Geometry::ProbLo); the upper one is prob_lo + n_cells * dx. */
int fhdx_write_plotfile(const fhdx_params* p, const char* plot_base_name, int step, double time,
                        const double prob_lo[3], const fhdx_fields* f, int plot_means,
                        int plot_vars, int flags, void* stream);

/* WriteCheckPoint (checkpoint.cpp:22-133): directory <chk_base_name><step, 9 digits> with the
 * reference's Header (title, step, time, statsCount, BoxArray) plus one trailing line with the
 * Philox seed, and one VisMF MultiFab per field (ghost cells included) under Level_0/. */
int fhdx_write_checkpoint(const fhdx_params* p, const char* chk_base_name, int step, double time,
                          int statsCount, uint64_t seed, const fhdx_fields* f, int flags,
                          void* stream);

/* ReadCheckPoint (checkpoint.cpp:135-255) of checkpoint <chk_base_name><restart, 9 digits>:
 * cu, prim, eta, kappa always; the statistics fields if reset_stats == 0, zeroed with
 * statsCount = 1 if reset_stats == 1.  *step is the checkpoint's step + 1 (the next step to
 * take).  *has_seed (may be NULL) is 1 and *seed the recorded Philox seed if the checkpoint
 * records one, else 0 with *seed unchanged. */
int fhdx_read_checkpoint(const fhdx_params* p, const char* chk_base_name, int restart,
                         int reset_stats, int* step, double* time, int* statsCount,
                         uint64_t* seed, int* has_seed, const fhdx_fields* f, int flags,
                         void* stream);

/* ---- staggered scheme (src_compressible_stag) -------------------------------------------
 * Momentum and velocity on faces, the rest in cells.  Arrays, as main_driver.cpp:616-631 and
 * 1023-1050 of src_compressible_stag define them: cu (nvars) and prim (nprimvars) cell FABs
 * with ngc ghosts; cumom[d] and vel[d] face FABs nodal in d with ngc ghosts; eta, zeta, kappa,
 * chi (nspecies), D (nspecies^2) cell FABs with ngc ghosts; faceflux[d] faces nodal in d, no
 * ghosts, nvars+4 comps; edgeflux[6] = edgeflux_x[0] (xy), edgeflux_x[1] (xz), edgeflux_y[0]
 * (xy), edgeflux_y[1] (yz), edgeflux_z[0] (xz), edgeflux_z[1] (yz), edges nodal along both of
 * their axes, no ghosts; cenflux[d] cells with one ghost layer.  Scope: the 3-D branch on a box
 * periodic in every direction (walls: FHDX_ERR_UNSUPPORTED), chemistry included.  The
 * noise replaces the MultiFabFillRandomNormal calls of timeStepStag.cpp:171-223 with a
 * counter-based Philox4x32-10 + Box-Muller keyed by (seed, step, instance 0) and the wrapped
 * global index of each face, edge and cell, so duplicate periodic faces and edges share a
 * value (OverrideSync's result); the stage noise is A + w_s B of the step's two sets. */

/* conservedToPrimitiveStag (conservedPrimitiveConversionsStag.cpp:4-144) with the
 * FillBoundary calls around it (timeStepStag.cpp:526-535): face velocities, prim and cu's
 * cell-centred momentum components on the valid box, then the periodic ghosts of cu, cumom,
 * vel and prim.  Also the start of a run (main_driver.cpp:766-794). */
int fhdx_cons_to_prim_stag(fhdx_ctx* ctx, fhdx_fab prim, const fhdx_fab vel[3], fhdx_fab cu,
                           const fhdx_fab cumom[3], void* stream);

/* calculateFluxStag (fluxStag.cpp:5-1868; signature compressible_functions_stag.H:127-143)
 * with the noise A + weight B of the (seed, step) sets; eta..D are inputs (the caller's
 * calculateTransportCoeffs, fhdx_transport). */
int fhdx_calculate_flux_stag(fhdx_ctx* ctx, fhdx_fab cu, const fhdx_fab cumom[3], fhdx_fab prim,
                             const fhdx_fab vel[3], fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa,
                             fhdx_fab chi, fhdx_fab D, const fhdx_fab faceflux[3],
                             const fhdx_fab edgeflux[6], const fhdx_fab cenflux[3], double weight,
                             double dt, uint64_t seed, uint64_t step, void* stream);

/* RK3stepStag (timeStepStag.cpp:13-1245; signature compressible_functions_stag.H:113-125):
 * cu, cumom advanced in place; prim, vel, the transport coefficients and the fluxes end as the
 * reference leaves them.  cu, cumom, prim and vel enter consistent, ghosts included (as after
 * fhdx_cons_to_prim_stag).  The stage arrays (cup, cup2 and their face momenta) are the
 * context's, allocated on the first call.  With nreaction > 0 (fhdx_ctx_set_chemistry first)
 * every stage adds dt times compute_compressible_chemistry_source_CLE of its prim, with the
 * reaction noise A + w_s B keyed as fhdx_fill_chem_noise keys it (timeStepStag.cpp:389-396 and
 * the stage-2/3 forms); source (>= nvars comps; the staggered driver allocates nprimvars) and
 * ranchem (nreaction comps), cell FABs with ngc ghosts, then end as the reference leaves them on
 * the valid cells (source components 5.. written; 0-4, the drivers' zero "external source", are
 * taken as zero); either may be passed with p = NULL (not written, and ignored without
 * chemistry). */
int fhdx_rk3_step_stag(fhdx_ctx* ctx, fhdx_fab cu, const fhdx_fab cumom[3], fhdx_fab prim,
                       const fhdx_fab vel[3], fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa,
                       fhdx_fab chi, fhdx_fab D, const fhdx_fab faceflux[3],
                       const fhdx_fab edgeflux[6], const fhdx_fab cenflux[3], fhdx_fab source,
                       fhdx_fab ranchem, double dt, uint64_t seed, uint64_t step, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FHDX_H */
