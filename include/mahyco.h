/*
 * mahyco.h -- C ABI of the MI355X-native VNR Lagrange + directional remap
 * hot path (libmahyco_hip.so).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns
 * an int status (MHY_OK == 0) and never throws.  The last error of a context
 * is available through mhy_last_error().  One context owns one HIP stream and
 * all device memory of one problem (or of one ensemble of problems); a
 * context is not thread-safe, distinct contexts may be used concurrently.
 *
 * Reference interfaces replaced (file:line in the MaHyCo sources):
 *   mhy_read_donnees      LectureDonneesClass::LectureDonnees
 *                         lecture_donnees/LectureDonnees.cc:12-226
 *   mhy_create            Vnr::Vnr + Vnr::simulate prologue
 *                         vnr_remap/Vnr.h:184-295, vnr_remap/VnrRemap.cc:330-341
 *   mhy_lagrange          Vnr::executeTimeLoopN Lagrange part
 *                         vnr_remap/VnrRemap.cc:139-171
 *   mhy_prepare_remap     computeVariablesForRemap / computeCellQuantitesForRemap /
 *                         computeFaceQuantitesForRemap  vnr_remap/PrepareRemap.cc:14-211
 *   mhy_remap_sweep       Remap::computeGradPhiFace{1,2} .. computeDualUremap{1,2}
 *                         remap/Remap.h:161-171
 *   mhy_remap_final       Vnr::remapVariables + computeNodeMass + EOS
 *                         vnr_remap/PhaseRemapFinal.cc:18-267, vnr_remap/VnrRemap.cc:186-190
 *   mhy_compute_dt        Vnr::computeDeltaT  vnr_remap/VnrRemap.cc:14-50
 *   mhy_step              one iteration of Vnr::executeTimeLoopN
 *                         vnr_remap/VnrRemap.cc:127-253
 */
#ifndef MAHYCO_H
#define MAHYCO_H

#ifdef __cplusplus
extern "C" {
#endif

#define MHY_NBMATMAX 3

/* status codes */
#define MHY_OK 0
#define MHY_ERR_ARG -1         /* invalid argument / unsupported option     */
#define MHY_ERR_IO -2          /* file could not be opened or parsed        */
#define MHY_ERR_DEVICE -3      /* HIP runtime error                         */
#define MHY_ERR_ENERGY -4      /* |e_k| > 5 (PhaseLagrange.cc:358-369)       */
#define MHY_ERR_NAN -5         /* NaN after remap (PhaseRemapFinal.cc:174-196) */
#define MHY_ERR_UNKNOWN_CASE -6 /* Init.cc:368-371                           */
#define MHY_ERR_FINISHED -7    /* time loop already ended                   */
#define MHY_ERR_DT_MIN -8      /* Eucclhyd time step below deltat_min
                                  (eucclhyd_remap/EucclhydRemap.cc:170-176) */

/* option values (includes/Limiteurs.h, Eos.h, ConditionsLimites.h,
 * SchemaLagrange.h, CasTest.h) */
#define MHY_LIM_MINMOD 300
#define MHY_LIM_SUPERBEE 301
#define MHY_LIM_VANLEER 302
#define MHY_LIM_MINMODG 1300
#define MHY_LIM_SUPERBEEG 1301
#define MHY_LIM_VANLEERG 1302
#define MHY_LIM_ARITHMETICG 1303
#define MHY_LIM_ULTRABEEG 1304

#define MHY_EOS_VOID 100
#define MHY_EOS_PERFECT_GAS 101
#define MHY_EOS_STIFFENED_GAS 102
#define MHY_EOS_MURNHAGAN 103
#define MHY_EOS_SOLID_LINEAR 104

#define MHY_BC_SYMMETRY 200
#define MHY_BC_IMPOSED_VELOCITY 201
#define MHY_BC_FREE_SURFACE 202
#define MHY_BC_PERIODIC 203

#define MHY_SCHEMA_EUCCLHYD 2000
#define MHY_SCHEMA_VNR 2001

#define MHY_CASE_UNIT_TEST 0
#define MHY_CASE_SEDOV 1
#define MHY_CASE_TRIPLE_POINT 2
#define MHY_CASE_SOD_X 4
#define MHY_CASE_SOD_Y 5
#define MHY_CASE_NOH 6
#define MHY_CASE_ADVECTION_X 7
#define MHY_CASE_ADVECTION_Y 8
#define MHY_CASE_ADVECTION_VIT_X 9
#define MHY_CASE_ADVECTION_VIT_Y 10
#define MHY_CASE_BI_UNIT_TEST 11
#define MHY_CASE_BI_SEDOV 12
#define MHY_CASE_BI_TRIPLE_POINT 13
#define MHY_CASE_BI_SHOCK_BUBBLE 14
#define MHY_CASE_BI_SOD_X 15
#define MHY_CASE_BI_SOD_Y 16
#define MHY_CASE_BI_NOH 17
#define MHY_CASE_BI_ADVECTION_X 18
#define MHY_CASE_BI_ADVECTION_Y 19
#define MHY_CASE_BI_ADVECTION_VIT_X 20
#define MHY_CASE_BI_ADVECTION_VIT_Y 21
/* Initialisers that are not in the reference (all-symmetry boundaries,
 * per-material gamma taken from gammap[]): */
#define MHY_CASE_RIEMANN_PERTURBED 100 /* 2 materials, sinusoidal interface x_I(y) */
#define MHY_CASE_POINT_EXPLOSION 101   /* 1 material, energy in the 2x2 origin cells */
#define MHY_CASE_TRIPLE_LAYOUT 102     /* 3 materials on [0,3]x[0,1] */
#define MHY_CASE_RIEMANN 103           /* 2 materials, straight interface x = x0 */

/* Problem description.  The first block is exactly what LectureDonnees
 * reads (same order, same defaults as includes/ headers); the second block holds
 * keys the reference has no input for. */
typedef struct mhy_config {
  int test_case;            /* CasTest id or MHY_CASE_* extension      */
  int nx, ny;               /* X_EDGE_ELEMS, Y_EDGE_ELEMS               */
  double dx, dy;            /* X_EDGE_LENGTH, Y_EDGE_LENGTH             */
  double final_time;
  double output_time;
  double deltat_init;
  double deltat_min;
  int max_time_iterations;  /* default 500000000 (GestionTemps.h:12)    */
  int nbmat;
  int eos[MHY_NBMATMAX];
  double gammap[MHY_NBMATMAX]; /* default 1.4 (Eos.h:16)                */
  double gamma;                /* eos->gamma, default 1.4 (Eos.h:17)    */
  int schema;
  int space_order;
  int avec_projection;
  int projection_conservative;
  double threshold;
  int projection_order;
  int limiter;              /* projectionLimiterId                      */
  int pp;                   /* projectionAvecPlateauPente               */
  int pp_mixte;             /* projectionLimiteurMixte                  */
  int limiter_pure;         /* projectionLimiterIdPure                  */
  int pp_complet;           /* projectionPlateauPenteComplet            */
  int pseudo_centree;
  int methode_flux_masse;   /* 0 A1, 1 A2, 2 PB                         */
  int sans_lagrange;
  /* extension keys */
  int abort_on_energy;      /* 1: stop with MHY_ERR_ENERGY when |e_k|>5 */
  int precision;            /* 0: FP64 everywhere; 1: mixed form: FP32
                               storage of the bulk cell / node fields and
                               FP32 slope / limiter / PenteBorne arithmetic,
                               FP64 volume fractions, masses, remap vectors,
                               compensated flux accumulation, dt          */
  unsigned long long seed;  /* initialiser randomisation               */
  /* initialiser parameters:
   *  RIEMANN_PERTURBED: [0] mean interface x, [1] amplitude, [2] modes,
   *                     [3] unused, [4] rho_L, [5] p_L, [6] rho_R, [7] p_R
   *  POINT_EXPLOSION:   [0] rho, [1] background p, [2] deposited energy
   *  TRIPLE_LAYOUT:     [0] relative density jitter amplitude
   *  RIEMANN:           [0] x0, [1] rho_L, [2] p_L, [3] rho_R, [4] p_R  */
  double param[8];
} mhy_config;

/* cell fields (state "n" of the reference, i.e. after the last swap) */
#define MHY_CELL_DENSITY 0      /* m_density_n              */
#define MHY_CELL_PRESSURE 1     /* m_pressure_n             */
#define MHY_CELL_ENERGY 2       /* m_internal_energy_n      */
#define MHY_CELL_PSEUDO 3       /* m_pseudo_viscosity_n     */
#define MHY_CELL_SOUND_SPEED 4  /* m_speed_velocity_n       */
#define MHY_CELL_DIVU 5         /* m_divu_n                 */
#define MHY_CELL_MASS 6         /* m_cell_mass              */
#define MHY_CELL_MIXTE 7        /* varlp->mixte (as double) */
#define MHY_CELL_PURE 8         /* varlp->pure  (as double) */
#define MHY_CELL_VELOCITY_X 9   /* Eucclhyd m_cell_velocity_n[0] */
#define MHY_CELL_VELOCITY_Y 10  /* Eucclhyd m_cell_velocity_n[1] */
#define MHY_CELL_OUT_VX 11      /* Eucclhyd m_x_cell_velocity (output) */
#define MHY_CELL_OUT_VY 12      /* Eucclhyd m_y_cell_velocity (output) */
/* per-material cell fields: add MHY_MAT_STRIDE*k to the id */
#define MHY_MAT_FRACVOL 16      /* m_fracvol_env            */
#define MHY_MAT_MASS_FRACTION 17 /* m_mass_fraction_env     */
#define MHY_MAT_DENSITY 18      /* m_density_env_n          */
#define MHY_MAT_ENERGY 19       /* m_internal_energy_env_n  */
#define MHY_MAT_PRESSURE 20     /* m_pressure_env_n         */
#define MHY_MAT_PSEUDO 21       /* m_pseudo_viscosity_env_n */
#define MHY_MAT_MASS 22         /* m_cell_mass_env          */
#define MHY_MAT_SOUND_SPEED 23  /* m_speed_velocity_env_n   */
#define MHY_MAT_STRIDE 16
/* node fields */
#define MHY_NODE_VELOCITY_X 64  /* m_node_velocity_n[0]     */
#define MHY_NODE_VELOCITY_Y 65  /* m_node_velocity_n[1]     */
#define MHY_NODE_MASS 66        /* m_node_mass              */
#define MHY_NODE_COORD_X 67     /* m_node_coord_n[0]        */
#define MHY_NODE_COORD_Y 68     /* m_node_coord_n[1]        */
#define MHY_NODE_OUT_VX 69      /* m_x_velocity (output)    */
#define MHY_NODE_OUT_VY 70      /* m_y_velocity (output)    */

/* scalars of the time loop */
typedef struct mhy_time_state {
  double t_n, t_nplus1, deltat_n, deltat_nplus1;
  int iteration;            /* n of executeTimeLoopN                   */
  int x_then_y_n;           /* direction of the first sweep            */
  int finished;             /* continueLoop became false               */
  int status;               /* last status of a step                   */
} mhy_time_state;

typedef struct mhy_ctx mhy_ctx;

/* LectureDonnees.cc:12-226: positional donnees.txt reader. */
int mhy_read_donnees(const char* path, mhy_config* cfg);
/* Defaults of includes/ headers (the structs main.cc:17-25 news). */
void mhy_config_default(mhy_config* cfg);
/* JSON input: the keys of donnees.txt (LectureDonnees.cc:12-226) by name,
 * the keys the positional format has no place for (per-material gammap,
 * the MHY_CASE_* extension initialisers by name or id, param, seed,
 * precision "fp64" / "mixed", abort_on_energy), an optional base file
 * "donnees" (path relative to the JSON file) and an optional "problems"
 * array, one configuration per entry, each starting from the keys around
 * it (an ensemble for mhy_batch_create).  Writes *count configurations into
 * cfgs; capacity 0 (cfgs may be NULL) only returns the count.  Unknown keys
 * or malformed JSON: MHY_ERR_IO; capacity too small: MHY_ERR_ARG. */
int mhy_read_json(const char* path, mhy_config* cfgs, int capacity, int* count);

/* Vnr::Vnr + simulate() prologue (VnrRemap.cc:330-341): builds the mesh,
 * initial state (Init.cc), uploads it and runs setUpTimeLoopN. */
int mhy_create(const mhy_config* cfg, mhy_ctx** ctx);
int mhy_destroy(mhy_ctx* ctx);
const char* mhy_last_error(const mhy_ctx* ctx);

/* One iteration of executeTimeLoopN (VnrRemap.cc:130-222) including the
 * n/n+1 swap; no output.  Enqueued on the context stream; scalars stay on
 * the device unless sync != 0.  Neither mhy_step nor mhy_run applies the
 * loop's stopping test (final_time, max_time_iterations): the caller
 * decides how many steps to run; after mhy_simulate has ended the loop both
 * return MHY_ERR_FINISHED. */
int mhy_step(mhy_ctx* ctx, int sync);
/* nsteps iterations without host synchronisation in between.  The error
 * status (|e| > 5 abort, NaN after the final remap) is read once, after the
 * nsteps: when a step fails, the later steps of the call still run, so the
 * fields, time and iteration count returned are those after nsteps steps,
 * not those of the failing step (the reference stops at that step).  A
 * caller that needs the state of the failing step steps with
 * mhy_step(ctx, 1) or uses mhy_simulate, which test the status every step. */
int mhy_run(mhy_ctx* ctx, int nsteps);
/* Whole reference time loop (VnrRemap.cc:127-253) with VTP/PVD output in
 * output_dir (NULL or "none": no output) -- the drop-in for Vnr::simulate. */
int mhy_simulate(mhy_ctx* ctx, const char* output_dir, int verbose);

/* Output of Vnr::dumpVariables (vnr_remap/VnrRemap.cc:255-299) through the
 * reference's PvdFileWriter2D format (cppnabla/mesh/PvdFileWriter2D.cc:26-114):
 * mhy_output_open selects the directory ("none" disables; the directory is
 * created if missing), mhy_output_write writes <dir>/VnrRemap.<index>.vtp for
 * the current state at time t and rewrites <dir>/VnrRemap.pvd. */
int mhy_output_open(mhy_ctx* ctx, const char* output_dir);
int mhy_output_write(mhy_ctx* ctx, int index, double t);

/* Phase-level entry points, 1:1 with the reference call sequence. */
int mhy_compute_dt(mhy_ctx* ctx);     /* computeCornerNormal+computeDeltaT+computeTime */
int mhy_lagrange(mhy_ctx* ctx);       /* updateVelocity .. computePressionMoyenne */
int mhy_prepare_remap(mhy_ctx* ctx);  /* PrepareRemap.cc                          */
int mhy_remap_sweep(mhy_ctx* ctx, int sweep); /* sweep 1 or 2: cells + dual      */
/* One stage of a sweep.  MHY_STAGE_CELLS replaces Remap::computeGradPhiFace{1,2},
 * computeGradPhi{1,2}, computeUpwindFaceQuantitiesForProjection{1,2} and
 * computeUremap{1,2} (remap/Remap.h:161-168, remap/PhaseRemap1.cc,
 * remap/PhaseRemap2.cc), fused in one kernel; MHY_STAGE_NODES replaces
 * Remap::computeDualUremap{1,2} (remap/Remap.h:169-170,
 * remap/DualPhaseRemap1.cc, remap/DualPhaseRemap2.cc).  Stage 2 of a sweep
 * reads the face fluxes of stage 1 of the same sweep. */
#define MHY_STAGE_CELLS 1
#define MHY_STAGE_NODES 2
int mhy_remap_stage(mhy_ctx* ctx, int sweep, int stage);
int mhy_remap_final(mhy_ctx* ctx);    /* remapVariables + node mass + EOS         */
int mhy_end_iteration(mhy_ctx* ctx);  /* loop test + swaps (VnrRemap.cc:193-222)  */

int mhy_synchronize(mhy_ctx* ctx);
int mhy_get_time_state(mhy_ctx* ctx, mhy_time_state* ts);
int mhy_get_mesh_size(const mhy_ctx* ctx, int* nx, int* ny);
/* Copies one field to host memory (n = number of cells or nodes). */
int mhy_get_field(mhy_ctx* ctx, int field, double* out, long long n);
/* Overwrites one field from host memory (test / restart use). */
int mhy_set_field(mhy_ctx* ctx, int field, const double* in, long long n);
/* sums[0] = total mass, sums[1] = total internal energy (sum of M*e),
 * sums[2..4] = mass of material k; FP64 compensated device reductions. */
int mhy_conservation(mhy_ctx* ctx, double* sums);
/* The same sums, allocating nothing: no counterpart in the reference.  For
 * callers that cannot allocate device memory (tight serving loops, code
 * around HIP graph capture, harnesses that forbid hipMalloc).  The block
 * partials live in a static device array of 5120 doubles (40 KiB) per GPU,
 * held until the process ends; every mesh fits it (at most 256 blocks x 5
 * quantities x 4 partials), and a larger request would return MHY_ERR_ARG.
 * Calls on the same GPU share that array: calls that overlap in time (several
 * host threads or streams) race.  Results are bitwise those of
 * mhy_conservation. */
int mhy_conservation_noalloc(mhy_ctx* ctx, double* sums);

/* Ensembles (no counterpart in the reference, which runs one problem per
 * process): n independent VNR Lagrange + remap problems of any mesh sizes,
 * sharing the material count and the PenteBorne option, stepped together.
 * Each kernel of a step is one launch over all problems; every problem keeps
 * its own arrays, dt and time, and its results are bitwise those of the same
 * problem run alone.  mhy_batch_run(b, n) advances every problem by n steps
 * (no per-problem stopping test); mhy_batch_context gives problem i's context
 * for field access (owned by the batch: do not destroy it). */
typedef struct mhy_batch mhy_batch;
int mhy_batch_create(const mhy_config* cfgs, int n, mhy_batch** batch);
int mhy_batch_destroy(mhy_batch* batch);
const char* mhy_batch_last_error(const mhy_batch* batch);
int mhy_batch_run(mhy_batch* batch, int nsteps);
int mhy_batch_context(mhy_batch* batch, int i, mhy_ctx** ctx);
long long mhy_batch_cells(const mhy_batch* batch);

/* Timing of the kernels launched by the last mhy_run (HIP events on the
 * context stream): per-kernel average milliseconds, by kernel index.
 * mhy_set_kernel_timing_mask limits the events to the kernels whose bit
 * (1 << k) is set (default: all), so that timing one kernel perturbs the
 * step as little as possible.  No counterpart in the reference, which has
 * no per-kernel timers. */
int mhy_kernel_count(const mhy_ctx* ctx);
/* The launch-shape tuner: on meshes of at least 2^20 cells the first steps
 * of mhy_step / mhy_run / mhy_simulate time the interchangeable y-direction
 * kernels (nodal remap on 16x48-node LDS tiles, 4x16-node wave strips or
 * 64-column walking bands; cell sweep on 4x16-cell wave strips or 64-column
 * walking bands; the Lagrange update of y-first steps fused into the
 * first sweep or run as its own kernel; results bitwise the same) and keep
 * the fastest before any graph is captured.  Returns the kept variants
 * ("dual_y=tile sweep_y=wave lagrange_y=fused") and a Lagrange variant forced
 * through the environment (MAHYCO_LAG_Y), or "" while tuning or when the
 * tuner is off (MAHYCO_AUTOTUNE=0).  No counterpart in the reference. */
const char* mhy_tuned_kernels(const mhy_ctx* ctx);
const char* mhy_kernel_name(const mhy_ctx* ctx, int k);
int mhy_kernel_bytes(const mhy_ctx* ctx, int k, double* bytes_per_launch);
int mhy_enable_kernel_timing(mhy_ctx* ctx, int on);
int mhy_set_kernel_timing_mask(mhy_ctx* ctx, unsigned mask);
int mhy_kernel_time(mhy_ctx* ctx, int k, double* avg_ms, int* launches);

#ifdef __cplusplus
}
#endif
#endif /* MAHYCO_H */
