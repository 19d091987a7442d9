/*
 * ice_gpu.h -- C-ABI of the MI355X explicit residual-and-update path.
 *
 * The path replaces the body of ICE's `Explicit_Step(grid)`
 * (src/lib/numerics/time/explicit/Mod_Explicit.f90:11-245), called once per
 * iteration from `ICE_solve` (src/lib/driver/Wrap_Solve.f90:26).  The host keeps
 * ICE's block / boundary-condition types (src/lib/base/Advanced_Types_m.f90:13-153)
 * and hands their arrays over here by plain pointers in their Fortran layouts:
 *
 *   prim     (nc, 1-gc:ni+gc, 1-gc:nj+gc, 1-gc:nk+gc)   gc = 2   (Mod_Allocate_Data.f90:216-238)
 *   vol      (ni, nj, nk)
 *   dl       (3, 0:ni+1, 0:nj+1, 0:nk+1)
 *   M        (3, 3, ni, nj, nk)                  metric tensor, row d = grad xi_d
 *   face_i   (4, 0:ni, 1:nj, 1:nk)  {N(3), A}    ICE_f_metrics_type (Base_Types_m.f90:22-25)
 *   face_j   (4, 1:ni, 0:nj, 1:nk)
 *   face_k   (4, 1:ni, 1:nj, 0:nk)
 *   gas prim (5, ni, nj, nk) = rho, u, v, w, T;  R, gam, k, mu (ni, nj, nk)
 *
 * Every call is synchronous at return, from one host thread.  Status 0 = ok;
 * a non-zero status leaves a message for ice_gpu_last_error.  No torch types,
 * no HIP types: a Fortran host binds these with ISO_C_BINDING (ice_gpu_m.f90),
 * Python with ctypes.
 */
#ifndef ICE_GPU_H
#define ICE_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ICE_GC        2    /* ghost layers, ICE_Global_m gc (Global_m.f90:9) */
#define ICE_NCOND_MAX 14   /* Global_m.f90:24 */
#define ICE_NEP       8    /* packed vapour properties, Lib_Evaporation.f90:15-19 */
#define ICE_NFROZEN_MAX 6  /* frozen slots: v, w and AG's P12, P13, P23 (Mod_Freeze.f90:117-128) */

/* Closures (Lib_Model.f90:105-316).  nc = 6 / 7 / 12. */
enum { ICE_MK = 0, ICE_IG = 1, ICE_AG = 2 };
/* Riemann solvers (Lib_Riemann.f90:30-57) */
enum { ICE_SAUREL = 0, ICE_RUSANOV = 1, ICE_HLLE = 2 };
/* Limiters, in the order of assign_limiter (Lib_Limiters.f90:35-84) */
enum { ICE_LIM_IORD = 0, ICE_LIM_MINMOD, ICE_LIM_VANALBADA, ICE_LIM_VANLEER, ICE_LIM_OSPRE,
       ICE_LIM_UMIST, ICE_LIM_OSHER, ICE_LIM_SWEBY, ICE_LIM_MC, ICE_LIM_KOREN, ICE_LIM_SUPERBEE };
/* Execution modes of ice_gpu_step */
enum { ICE_MODE_PLAIN = 0, ICE_MODE_BATCHED = 1 };

/* One particle family: its closure, its Riemann solver and its material
 * (condensed_phase_t, Config_Types_m.f90:195-293, constant properties). */
typedef struct {
  int32_t closure;          /* ICE_MK / ICE_IG / ICE_AG                                  */
  int32_t riemann;          /* ICE_SAUREL / ICE_RUSANOV / ICE_HLLE                       */
  int32_t nfrozen;          /* slots not integrated (Mod_Freeze.f90:37-141)              */
  int32_t frozen[ICE_NFROZEN_MAX]; /* their 1-based slot numbers                         */
  int32_t evap_select;      /* 0 none, 1 d2-law, 2 CEM, 3 CEM-B, 4 ASM, 5 TC             */
  int32_t intf_select;      /* 0 VLE, 1 LK                                               */
  int32_t surrogate;        /* 1: seed the LK interface solve with the learned model     */
  double  rho_mat;          /* rho_al  [kg/m^3]                                          */
  double  cs;               /* cs_al   [J/(kg K)]                                        */
  double  lv;               /* lv_al   [J/kg]                                            */
  double  emiss;            /* surface emissivity                                        */
  double  ep[ICE_NEP];      /* Mv, Lv, cpv, Le, Yinf, Lv*Mv/Ru, 1/Tboil, alphaE          */
} ice_family_t;

/* time_scheme_t / space_scheme_t / irs_t subset (Config_Types_m.f90:119-180) */
typedef struct {
  int32_t ngroups;          /* particle families                                         */
  int32_t nrk;              /* 1 euler, 2 RK2, 3 RK3                                     */
  int32_t time_accurate;    /* 1: one global dt (min over cells, families)               */
  int32_t cfl_ramp_iter;    /* cfl-rampa-iter                                            */
  int32_t ndir;             /* resolved directions, 1..3 (setup_metrics)                 */
  int32_t limiter;          /* ICE_LIM_*                                                 */
  int32_t shock_detector;   /* 1: Jameson sensor weight on the slopes                    */
  int32_t irs;              /* 1: implicit residual smoothing                            */
  int32_t coupled;          /* one-way coupled to a frozen gas                           */
  int32_t drag_select;      /* Lib_Drag selector 1..13                                   */
  int32_t heat_select;      /* Lib_Heat selector 1..8                                    */
  int32_t blow_select;      /* 0 none, 1 LK blowing factor                               */
  int32_t mode;             /* ICE_MODE_PLAIN / ICE_MODE_BATCHED                         */
  int32_t var_yinf;         /* 1: Yinf per cell from the gas (ice_gpu_put_gas_yinf)      */
  int32_t reserved[2];
  double  cfl, dt_max, tau_factor, irs_beta;
} ice_config_t;

/* Boundary records, one per (boundary cell, family), in the order of ICE's bc
 * array (IO_BC.f90:299, idx = (n-1)*ngroups + p).  Integer fields, 1-based: */
enum { ICE_R_B = 0, ICE_R_I, ICE_R_J, ICE_R_K, ICE_R_F, ICE_R_TYPE, ICE_R_P,
       ICE_R_BS, ICE_R_IS, ICE_R_JS, ICE_R_KS, ICE_R_FS,
       ICE_R_NORMAL_DIR, ICE_R_NCHI1, ICE_R_NCHI2, ICE_R_DONOR0, ICE_REC_NI };
/* Real fields of an inlet record (ICE_bc_type, Advanced_Types_m.f90:111-118) */
enum { ICE_RD_MASSFLUX = 0, ICE_RD_VELOCITY, ICE_RD_TEMPERATURE, ICE_RD_RADIUS,
       ICE_RD_ALPHA, ICE_RD_BETA, ICE_RD_GFIXED, ICE_REC_ND };

typedef struct {
  int64_t n_floored;        /* update events with a negative density (Lib_Newstate.f90:178-193) */
  int32_t bad_state;        /* a state failed check_prim after the update                        */
  int32_t bad_block, bad_i, bad_j, bad_k, bad_p;   /* the first failing cell (1-based)           */
  int32_t bad_recon;        /* an unphysical stencil at first order (Mod_Fluxes.f90:238-241)     */
  int32_t iter;             /* grid%iter after the call                                          */
  double  dtglobal;         /* last global dt (time-accurate)                                    */
  double  time;             /* grid%time                                                         */
} ice_status_t;

typedef struct ice_gpu_ctx ice_gpu_ctx;

/* Context on HIP device `device`: the run's configuration, its families and the
 * cell counts of its nb blocks (dims = 3*nb ints).  Replaces the set-up that
 * Wrap_Setup.f90:48-121 does for the hot path (allocate_data, setup_freeze). */
int  ice_gpu_create(int device, const ice_config_t* cfg, const ice_family_t* fams,
                    int nb, const int32_t* dims, ice_gpu_ctx** ctx);

/* Geometry of block b (1-based), from setup_metrics (Mod_Metrics.f90:19-59). */
int  ice_gpu_set_geometry(ice_gpu_ctx* ctx, int b, const double* vol, const double* dl,
                          const double* M, const double* face_i, const double* face_j,
                          const double* face_k);

/* The boundary table (Setup_BC, IO_BC.f90:65-344): nrec records of ICE_REC_NI
 * ints and ICE_REC_ND reals; chimera donors (b,i,j,k) and weights, record r's
 * first donor at reci[r*ICE_REC_NI+ICE_R_DONOR0] (0-based into the donor list).
 * Pairs the connections (Pair_Connections, IO_BC.f90:360-510) and builds the
 * per-family record lists and boundary cells (Mod_GhostExchange.f90:103-360). */
int  ice_gpu_set_boundaries(ice_gpu_ctx* ctx, int nrec, const int32_t* reci, const double* recd,
                            int ndonor, const int32_t* donor, const double* weight);

/* The pairing it computed: partner record (1-based, 0 = none), owner flag and
 * face number cf of every record; and family p's boundary cells, cell c holding
 * records rec[first[c]-1 .. first[c+1]-2] (1-based, collect_boundary_cells). */
int  ice_gpu_get_bc_maps(ice_gpu_ctx* ctx, int32_t* partner, int32_t* owner, int32_t* cf,
                         int p, int32_t* ncell, int32_t* first, int32_t* rec, int cap);

/* State of family p in block b, primitive variables with ghosts. */
int  ice_gpu_put_state(ice_gpu_ctx* ctx, int b, int p, const double* prim);
int  ice_gpu_get_state(ice_gpu_ctx* ctx, int b, int p, double* prim);

/* Frozen gas of block b (set_gas, Mod_Exchange.f90:63-81; Mod_Phase.f90:11-63). */
int  ice_gpu_put_gas(ice_gpu_ctx* ctx, int b, const double* gprim, const double* R,
                     const double* gam, const double* k, const double* mu);
/* Far-field vapour mass fraction per cell, replacing ep(iYinf) (cfg.var_yinf = 1). */
int  ice_gpu_put_gas_yinf(ice_gpu_ctx* ctx, int b, const double* yinf);

/* Relaxation time of family p (ni,nj,nk) before the first step; the stages write it,
 * and the next step's dt pass reads it (Mod_dt.f90:94-98): a restart that puts it
 * back goes on as the uninterrupted run. */
int  ice_gpu_put_tau(ice_gpu_ctx* ctx, int b, int p, const double* tau);
int  ice_gpu_get_tau(ice_gpu_ctx* ctx, int b, int p, double* tau);

/* nsteps calls of Explicit_Step, all on the device; status after the last one.
 * Returns non-zero when a step hit a fatal state (st->bad_state / bad_recon). */
int  ice_gpu_step(ice_gpu_ctx* ctx, int nsteps, ice_status_t* st);

/* 1: the last stage of every step keeps its source for ice_gpu_get_source.
 * 2: the step's source, each stage's weighted as the update weighs its residual
 *    (add_stage_source / hand_step_source, Mod_Explicit.f90:257-300): what a
 *    two-way coupled run hands the gas (exchange_rate, Mod_Exchange.f90:83-103). */
int  ice_gpu_set_keep_source(ice_gpu_ctx* ctx, int on);

/* res-threshold of a steady run (Mod_Explicit.f90:196-225): each ice_gpu_step then
 * computes the residual norm after every step and ends early, after the step whose
 * density residual is at most threshold times the first non-zero one of the run.
 * 0 (the default) never stops on the residual; time-accurate runs never do.  Setting
 * it restarts the reference. */
int  ice_gpu_set_res_threshold(ice_gpu_ctx* ctx, double threshold);
/* Whether the last ice_gpu_step ended on the residual, and the reference residual. */
int  ice_gpu_res_stop(ice_gpu_ctx* ctx, int32_t* stopped, double* reference);

/* Last stage's source (nc,ni,nj,nk) and relaxation time (ni,nj,nk) of family p. */
int  ice_gpu_get_source(ice_gpu_ctx* ctx, int b, int p, double* source, double* tau);

/* residual_norm (Mod_Diagnostic.f90:22-69) of the last step: total[5], fam[5*ngroups]. */
int  ice_gpu_residual_norm(ice_gpu_ctx* ctx, double* total, double* fam);

/* residual_norm of the last step as ice_gpu_residual_norm, with no allocation: no
 * counterpart in the reference; for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its device memory is a file-scope array of 1.5 MiB per GPU, which
 * holds 604 block-family pairs up to 8 families (larger grids get an error: use
 * ice_gpu_residual_norm); calls that overlap on one GPU race. */
int  ice_gpu_residual_norm_noalloc(ice_gpu_ctx* ctx, double* total, double* fam);

/* Time (ms, HIP events on the context's stream) of the last ice_gpu_step, and
 * the summed duration and launch count of the dominant kernel in it (the fused
 * stage kernel; with ice_gpu_set_kernel_timing). */
int  ice_gpu_timing(ice_gpu_ctx* ctx, double* step_ms, double* kernel_ms, int64_t* kernel_launches);
/* The summed duration and launch count of the deferred interface-solve kernel
 * (k_lk_finish: the cells that run the Langmuir-Knudsen solve, their source,
 * residual and update) in the last ice_gpu_step run with kernel timing on.  No
 * counterpart in the reference: it measures the solve the learned seed speeds up. */
int  ice_gpu_lk_timing(ice_gpu_ctx* ctx, double* ms, int64_t* launches);
/* Cells per workgroup tile of the stage kernel (tx*ty*tz, and the faces of the
 * tile in each direction, at most 256; the tile widened by one cell along each
 * resolved direction, at most 352); tx = -1 restores the default shapes,
 * tx = 0 selects the untiled one-thread-per-cell kernel. */
int  ice_gpu_set_tiles(ice_gpu_ctx* ctx, int tx, int ty, int tz);
/* Dispatch of the batched stage kernel: 0 persistent workgroups over per-XCD tile
 * queues, 1 one workgroup per tile with the families of a tile side by side, 2 one
 * workgroup per tile with each block's tiles family by family (same tiles, same
 * bits). */
int  ice_gpu_set_dispatch(ice_gpu_ctx* ctx, int dispatch);
/* The autotuner: times `steps` steps of the run (plain or batched, as configured)
 * with each candidate stage tile for its blocks' dimensionality (and, batched, each
 * dispatch), restores the state after every trial, and keeps the fastest.
 * best[4] receives the tile and the dispatch, ms[k] the milliseconds per step of
 * trial k (at most 24; infinity when it did not fit).  Needs a device copy of the
 * state (three buffers, tau and dt per family). */
int  ice_gpu_autotune(ice_gpu_ctx* ctx, int steps, int32_t* best, double* ms);
/* 1: bracket every launch of the dominant kernel with HIP events (bench only). */
int  ice_gpu_set_kernel_timing(ice_gpu_ctx* ctx, int on);

/* ---- learned seed of the Langmuir-Knudsen interface solve (no counterpart in the
 * reference: the north star's surrogate).  A 6-64-64-2 ReLU network in BF16 on the
 * MFMA units, FP32 accumulation; parameters are ICE_MLP_PARAMS floats:
 * W1t[64][16] b1[64] W2t[64][64] b2[64] W3t[16][64] b3[16] (W[n][k], input k to
 * output n; inputs padded 6 -> 16 and outputs 2 -> 16 with zeros).  Inputs are the
 * six scaled features of a drop (gas T, pressure, far-field vapour fraction, drop T,
 * diameter, slip Reynolds number); outputs the interface state: the rate over the
 * equilibrium rate and the surface vapour molar fraction over its equilibrium value.
 * ice_family_t.surrogate = 1 seeds each cell's interface solve with the first output
 * (secant iteration, the reference's Picard solve as fallback); 2 starts the same
 * secant iteration from the equilibrium rate (the method without the network). */
#define ICE_MLP_PARAMS 6288
/* Loss (mean squared error) and its gradient for n samples x[n][6] -> y[n][2]. */
int  ice_gpu_mlp_grad(ice_gpu_ctx* ctx, int n, const float* x, const float* y, const float* params, float* grad,
                      float* loss);
/* Network outputs out[n][2] for x[n][6]. */
int  ice_gpu_mlp_forward(ice_gpu_ctx* ctx, int n, const float* x, const float* params, float* out);
/* Samples family p's cells that run the interface solve in the current state, solves
 * them with the plain solver, and trains the family's network on them on the GPU
 * (Adam, mini-batches of `batch`, `epochs` passes, learning rate lr, initial weights
 * from `seed`).  report[6]: samples, initial loss, final loss, mean and max absolute
 * error of the rate ratio, milliseconds. */
int  ice_gpu_surrogate_train(ice_gpu_ctx* ctx, int p, int epochs, int batch, double lr, uint64_t seed,
                             float* report);
/* Online retraining: after every `every` steps of ice_gpu_step (iteration count a
 * multiple of it), each trained network is trained again, `epochs` passes from its
 * current weights, on samples of the current state (0 turns it off). */
int  ice_gpu_surrogate_retrain_every(ice_gpu_ctx* ctx, int every, int epochs);
/* The same retraining of family p's network, now: `epochs` passes from its current
 * weights on samples of the current state, with the settings of its last training. */
int  ice_gpu_surrogate_retrain(ice_gpu_ctx* ctx, int p, int epochs);
int  ice_gpu_surrogate_params(ice_gpu_ctx* ctx, int p, float* params);
int  ice_gpu_surrogate_set_params(ice_gpu_ctx* ctx, int p, const float* params);
/* 1: count the interface solves of every stage; ice_gpu_lk_stats: family p's solves,
 * map evaluations, fallbacks to the plain solve and the largest evaluation count. */
int  ice_gpu_set_lk_stats(ice_gpu_ctx* ctx, int on);
int  ice_gpu_lk_stats(ice_gpu_ctx* ctx, int p, int64_t* counts);

/* ---- grid sequencing (Mod_Multigrid.f90:93-131, Lib_Multigrid.f90:40-191): two
 * contexts of one run one level apart on the same device, with the same families
 * and blocks; each coarse block has half the fine block's cells in every direction
 * (a flat block keeps nk = 1).  Restriction sets the coarse state to the fine
 * one's conservative state averaged over the coarse cell's fine cells, weighted by
 * volume; prolongation sets the fine state by trilinear interpolation of the
 * coarse primitives (27/64, 9/64, 3/64, 1/64; flat: 9/16, 3/16, 1/16), neighbours
 * clamped to the block.  The interiors only: the next stage fills the ghosts. */
int  ice_gpu_restrict(ice_gpu_ctx* fine, ice_gpu_ctx* coarse);
int  ice_gpu_prolong(ice_gpu_ctx* coarse, ice_gpu_ctx* fine);

/* ---- tabulated material properties (Lib_Properties.f90:18-250, the table that
 * Load_Table.f90:131-147 loads).  Family p's material gets n rows on the integer
 * kelvins T_min .. T_min+n-1; cols holds ICE_PROP_NCOL columns of n values each:
 * density, cp, energy e = h - hOff, Psat, Lv, Cpv, emissivity.  Each property is
 * linear between the nodes and takes its end value outside them; e is extended
 * linearly beyond both ends and inverted by bisection for T.  flags says which
 * columns the path takes, as Load_Table decides: density varies (else the first
 * row's), cp varies (the closures' energy is then e(T)), and Psat, Lv, Cpv and
 * emissivity replace Clausius-Clapeyron and the constant values.  The table's first
 * rows of density and cp replace the family's rho_mat and cs.  Call before the
 * first step; it replaces an earlier table of the family. */
enum { ICE_PROP_RHO = 0, ICE_PROP_CP, ICE_PROP_E, ICE_PROP_PSAT, ICE_PROP_LV, ICE_PROP_CPV, ICE_PROP_EMISS,
       ICE_PROP_NCOL };
enum { ICE_PROP_RHO_VARIES = 1, ICE_PROP_CS_VARIES = 2, ICE_PROP_USE_PSAT = 4, ICE_PROP_USE_LV = 8,
       ICE_PROP_USE_CPV = 16, ICE_PROP_USE_EMISS = 32 };
int  ice_gpu_set_properties(ice_gpu_ctx* ctx, int p, int T_min, int n, const double* cols, int32_t flags);

/* ---- particles entering through extrapolation faces (400): count_400_inflow and
 * report_400_inflow (src/lib/diagnostic/Mod_Inflow400.f90:52-171).  Every stage
 * counts, per boundary (the 400 records of one face of one block, for one family),
 * the records whose cell carries the phase (rho > rho_empty) and moves into the
 * domain faster than round-off, and measures it (|w| for MK, |w|/sqrt(P/rho) for IG,
 * |w|/sqrt(Pnn/rho) for AG).  A line is logged at a boundary's first such stage and
 * each time a stage's largest measure exceeds 10 times that of its last line.  The
 * device logs lines in the order the stages run, boundaries by block then face: the
 * reference's order on the plain path (family by family, each family's stages in
 * turn); the batched path runs the families of one closure group stage by stage, so
 * their lines interleave by stage.  The host prints them (the reference's text: ICE
 * prints and flushes them at the stage). */
typedef struct {
  int32_t iter;             /* iter_general of the stage                                 */
  int32_t block, face, family;
  int32_t faces;            /* records above 10 times the boundary's last line           */
  int32_t relative;         /* 1: the boundary had a line before ("faces above 10 times") */
  double  largest;          /* their largest measure                                     */
} ice_inflow400_line_t;
/* A boundary's totals over the run (report_400_inflow). */
typedef struct {
  int32_t block, face, family, stages, first_iter, first_faces;
  double  first_largest, largest;
} ice_inflow400_total_t;
#define ICE_INFLOW400_LOG 4096   /* lines the device keeps between two reads */
/* The lines logged since the last call, oldest first: *n = min(logged, max,
 * ICE_INFLOW400_LOG) lines into out, *dropped = the lines logged beyond them.
 * Clears the log. */
int  ice_gpu_inflow400_lines(ice_gpu_ctx* ctx, ice_inflow400_line_t* out, int max, int32_t* n, int64_t* dropped);
/* The totals of every boundary that let particles in, in the reference's report
 * order (family, block, face): *n of them, at most max. */
int  ice_gpu_inflow400_totals(ice_gpu_ctx* ctx, ice_inflow400_total_t* out, int max, int32_t* n);

/* compute_ghost and fill_second_ghost (Lib_Ghost.f90:29-72, 270-308): both ghost
 * layers of every family on its current state, as the next stage fills them
 * first; ice_gpu_get_state then returns them. */
int  ice_gpu_fill_ghosts(ice_gpu_ctx* ctx);

/* ---- coupler births (Mod_Exchange.f90:225-324).  The rates in force on family p
 * in block b, in the conservative slots (nc, ni, nj, nk) per unit volume and time,
 * as clear_import_rates + add_birth_rate / add_absorption_rate / add_removal_rate
 * leave them (the packets are the coupler's: birth_packet, absorption_packet);
 * every stage adds them to the residual over the cell's dt (add_births,
 * Lib_Residual.f90:164-210).  NULL clears the block's. */
int  ice_gpu_put_imports(ice_gpu_ctx* ctx, int b, int p, const double* imports);
/* add_birth (Mod_Exchange.f90:283-300): n packets of family p at once, packet m
 * born in cell cells[4m..4m+3] = (b, i, j, k) with pk[6m..6m+5] = [mass, momentum
 * (3), energy, number] per unit volume, added to the cell's conservative state;
 * the packets of one cell in their order. */
int  ice_gpu_add_births(ice_gpu_ctx* ctx, int p, int n, const int32_t* cells, const double* pk);

/* ---- run control (no counterpart in the reference, which fixes these at set-up:
 * Assign_Setup.f90:60-117).  Execution mode of the following steps, as cfg.mode
 * at create; the interface solve of family p, as ice_family_t.surrogate. */
int  ice_gpu_set_mode(ice_gpu_ctx* ctx, int mode);
int  ice_gpu_set_surrogate(ice_gpu_ctx* ctx, int p, int mode);

/* ---- restart (Wrap_Setup.f90:158-214): the counters a solution stores for its
 * restart (IO_Solution.f90:20-30) -- grid%iter (ICE_ITERATION, which drives the
 * CFL ramp of Mod_dt.f90:83-85), grid%time, and the residual reference of
 * res-threshold (ICE_RES_REFERENCE).  Set them after put_state (and put_tau) to
 * resume a run as the uninterrupted run goes on. */
int  ice_gpu_set_counters(ice_gpu_ctx* ctx, int32_t iter, double time, double res_reference);
int  ice_gpu_get_counters(ice_gpu_ctx* ctx, int32_t* iter, double* time, double* res_reference);

int  ice_gpu_last_error(ice_gpu_ctx* ctx, char* buf, int len);
void ice_gpu_destroy(ice_gpu_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* ICE_GPU_H */
