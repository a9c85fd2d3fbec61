/*
 * frehg_gw.h — C-ABI of the MI355X-native Richards groundwater step.
 *
 * Two seams of the reference are replaced:
 *
 *  1. The step seam, frehg::gw::RichardsSolver (src/gw/RichardsSolver.hpp:166-423):
 *     construction from a grid/soil/IC/BC description, step(t, dtg), nextDt(),
 *     head(), waterContent(), audit(), ownedVolume(), lastSolve(), the flux and
 *     face-conductivity accessors, attachCoupling().
 *  2. The solve seam, frehg::LinearSystem (src/core/LinearSystem.hpp:109-151):
 *     setPattern / setValues / solve on a COO system, replacing the PETSc
 *     KSPCG call (src/core/LinearSystem.cpp:149,308,320).
 *
 * Conventions
 *  - Every function returns 0 on success and a nonzero FREHG_E* code on
 *    failure; frehg_last_error() then returns a message (thread-local).
 *    This replaces the reference's throwing log::fatal (src/core/Logger.hpp:53-57).
 *  - Host arrays are plain C pointers.  Arrays documented "device" are HIP
 *    device pointers; work on them is ordered on the handle's stream.
 *  - 3D interior arrays are flattened (j*nx + i)*nz + k with k = 0 the top
 *    layer (the reference's flat-file order, src/gw/RichardsSolver.cpp:224-228).
 *    2D column arrays are j*nx + i.
 *  - Halo-layout fields (the getters) are (ny+2) x (nx+2) x nz with one ghost
 *    cell in j and i and k unit stride, exactly the reference's Field3
 *    (src/core/Types.hpp:35-43); z-face fields carry nz+1 planes, plane k the
 *    face above cell k (src/gw/RichardsSolver.hpp:584-587).
 */
#ifndef FREHG_GW_H
#define FREHG_GW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FREHG_GW_ABI_VERSION 1

/* Error codes. */
#define FREHG_OK 0
#define FREHG_EINVAL 1      /* invalid argument or descriptor */
#define FREHG_EDEVICE 2     /* HIP runtime error */
#define FREHG_EDIVERGED 3   /* linear solve did not converge (fatal in the reference) */
#define FREHG_EUNSUPPORTED 4

/* Boundary-condition codes: GwBcCode (src/gw/RichardsSolver.hpp:66-72). */
#define FREHG_BC_NOFLUX 0
#define FREHG_BC_HEAD 1
#define FREHG_BC_FLUX 2
#define FREHG_BC_GRAVITY 3
#define FREHG_BC_HEAD_HYDROSTATIC 4

/* initial_conditions.groundwater forms (src/gw/RichardsSolver.cpp:393-413). */
#define FREHG_IC_MOISTURE 0
#define FREHG_IC_HEAD 1
#define FREHG_IC_WATER_TABLE 2

/* Preconditioners of the PCG that replaces KSPCG. */
#define FREHG_PC_JACOBI 0   /* point Jacobi */
#define FREHG_PC_ZLINE 1    /* z-line block Jacobi: one tridiagonal solve per column */
#define FREHG_PC_ICC0 2     /* incomplete Cholesky (CPU oracle only; the reference default) */
#define FREHG_PC_CHEBYSHEV 3 /* mixed form: the FP32 inner PCG preconditioned by a Jacobi-
                              * Chebyshev polynomial (frehg_gw_set_chebyshev); even nz */

/* Arithmetic of the linear solve. */
#define FREHG_PREC_FP64 0   /* plain FP64 PCG */
#define FREHG_PREC_MIXED 1  /* FP32 inner PCG inside FP64 iterative refinement */

/*
 * Everything RichardsSolver's constructor reads from FrehgConfig, Grid,
 * TerrainMetric and BoundarySet, as plain arrays
 * (src/gw/RichardsSolver.cpp:61-205; YAML keys src/core/Config.cpp:305-420).
 * Zero-initialise and fill; NULL array pointers select the *_const value.
 */
typedef struct frehg_gw_desc {
  /* domain.* (src/core/Config.hpp DomainConfig; Grid, TerrainMetric) */
  int nx, ny, nz;
  double dx, dy, dz;
  double dz_stretch;               /* 0 is read as 1 */
  int follow_terrain;              /* domain.follow_terrain */
  int terrain_layers;              /* 0 scaled (legacy), 1 uniform */
  const double* bottom_elevation;  /* ny*nx raster or NULL */
  double bottom_elevation_const;

  /* soil: per-cell fields, interior cells (j*nx+i)*nz+k.  The reference's
   * soil.types x soil.map staging (src/gw/RichardsSolver.cpp:207-329)
   * expands to exactly these arrays; continuous per-cell fields are the
   * documented extension. */
  const double *ksx, *ksy, *ksz, *vg_alpha, *vg_n, *theta_s, *theta_r, *aev;

  /* groundwater.* */
  double specific_storage;
  int use_full3d;
  int predictor_corrector;         /* 1 consistent (default), 0 legacy */
  int face_conductivity;           /* 0 arithmetic (default), 1 harmonic */
  int saturated_contact;           /* 1 head (default); 0 legacy is rejected */
  int reallocation;                /* 0 legacy (default), 1 disabled */
  int reallocation_surplus;        /* 0 drop (default), 1 redistribute */
  double dt_init, dt_min, dt_max;  /* groundwater.timestep */
  double dq_grow, dq_shrink, courant_max;

  /* initial_conditions.groundwater */
  int ic_form;                     /* FREHG_IC_* */
  const double* ic_values;         /* per cell (moisture/head) or per column (water_table) */
  double ic_const;

  /* Rasterized boundary markers (GwBcCode per column / edge cell, ny*nx each).
   * Side arrays are read at their edge cells only (x-: i = 0, x+: i = nx-1,
   * y-: j = 0, y+: j = ny-1).  NULL code arrays mean no-flux. */
  const int *top_code, *bot_code, *xm_code, *xp_code, *ym_code, *yp_code;
  const double *top_value, *bot_value, *xm_value, *xp_value, *ym_value, *yp_value;

  /* solver.groundwater.* */
  double rtol, atol;
  int max_iterations;
  int preconditioner;              /* FREHG_PC_* */
  int precision;                   /* FREHG_PREC_*: solver.groundwater.precision (extension) */
  int initial_guess_previous;      /* 1: start the PCG from h^n; 0: from zero (PETSc default) */
  double inner_rtol;               /* mixed only: inner FP32 solve tolerance (0 -> 1e-4) */
} frehg_gw_desc;

/* GwStepAudit (src/gw/RichardsSolver.hpp:86-130), the terms this path books. */
typedef struct frehg_gw_audit {
  double boundary_in;
  double ss_storage;
  double realloc_adjust;
  double realloc_dropped;
  double vloss;
  double cpl_exchanged;
  double cpl_bounce;
  double cpl_vent;
  double cpl_evap;
  double cpl_vent_thin;
  double pc_clamp;                 /* mode-3 clamp volume of this step */
  double tr_applied;               /* root-uptake sink volume S V dtg of this step
                                      (GwStepAudit::trApplied, RichardsSolver.hpp:113) */
} frehg_gw_audit;

/* SolveStats (src/core/LinearSystem.hpp:49-53) plus refinement detail. */
typedef struct frehg_gw_solve_stats {
  int iterations;                  /* total PCG iterations (inner ones in mixed mode) */
  int refinements;                 /* FP64 residual sweeps (mixed mode), else 0 */
  int converged;                   /* 1 converged, 0 not */
  double residual_norm;            /* final true residual 2-norm */
  double rhs_norm;                 /* ||b||_2 */
  int fp64_iterations;             /* mixed mode: iterations the solve finished in FP64
                                      after an FP32 cycle stopped reducing the FP64
                                      residual (0 when it did not) */
} frehg_gw_solve_stats;

/* Field identifiers for frehg_gw_get_field / frehg_gw_set_field. */
#define FREHG_F_HEAD 0             /* h     (H)  RichardsSolver::head() */
#define FREHG_F_WC 1               /* theta (H)  waterContent() */
#define FREHG_F_HEAD_N 2           /* h^n   (H) */
#define FREHG_F_WC_N 3             /* theta^n (H) waterContentStart() */
#define FREHG_F_KX 4               /* face K x+ (H)  faceConductivityX() */
#define FREHG_F_KY 5               /* face K y+ (H) */
#define FREHG_F_KZF 6              /* face K z (F)   faceConductivityZFace() */
#define FREHG_F_QX 7               /* Darcy volume flux x (H) fluxXVolumetric() */
#define FREHG_F_QY 8
#define FREHG_F_QZF 9              /* (F) fluxZFaceVolumetric() */
#define FREHG_F_VLOSS 10           /* cumulative clamp volume (H) vlossField() */
#define FREHG_F_DZ3D 11            /* geometry (H) */
#define FREHG_F_BOT3D 12
#define FREHG_F_AREA_X 13
#define FREHG_F_AREA_Y 14
#define FREHG_F_SIN_X 15
#define FREHG_F_COS_X 16
#define FREHG_F_SIN_Y 17
#define FREHG_F_COS_Y 18
#define FREHG_F_CAPACITY 19        /* C(h^n) of the last assembly (H) */
#define FREHG_F_SINK 20            /* root-uptake sink S [1/s] (H) rootUptakeSink(),
                                      RichardsSolver.hpp:295,656 */
/* The transport window (frehg_gw_enable_transport_window; RichardsSolver.hpp:
 * 226-249): face volumes [m^3] summed over the steps since the last reset. */
#define FREHG_F_WIN_X 21           /* through the x+ faces (H) windowVolumeX() */
#define FREHG_F_WIN_Y 22           /* through the y+ faces (H) windowVolumeY() */
#define FREHG_F_WIN_Z 23           /* through the z faces, incl. the post-allocation
                                      transfers on every face crossed (F) windowVolumeZ() */
#define FREHG_F_WIN_SINK 24        /* root-uptake volume (H) windowSinkVolume() */
#define FREHG_F_QZ_PATH_CORR 25    /* the last step's path corrections [m^3/s] (F) qzPathCorr_ */
#define FREHG_F_WIN_VENT 26        /* vented volume, 2D (ny+2)(nx+2) windowVentVolume() */
#define FREHG_F_WIN_VAP 27         /* ET vapour volume, 2D windowVapourVolume() */

/* Opaque handles. */
typedef struct frehg_gw_solver frehg_gw_solver;
typedef struct frehg_linear_system frehg_linear_system;

/* Library identity and errors. */
int frehg_abi_version(void);
const char* frehg_last_error(void);

/* ---- step seam (replaces frehg::gw::RichardsSolver) ------------------- */

/* RichardsSolver::RichardsSolver (src/gw/RichardsSolver.cpp:61-205) together
 * with TerrainMetric (src/gw/TerrainMetric.cpp:69-99) and
 * Grid::buildGlobalIds (src/core/Grid.cpp:160-245): stages geometry, soil,
 * boundary markers and the initial state on device. */
int frehg_gw_create(const frehg_gw_desc* desc, frehg_gw_solver** out);
int frehg_gw_destroy(frehg_gw_solver* s);

/* RichardsSolver::step(t, dtg) (src/gw/RichardsSolver.cpp:729-794). */
int frehg_gw_step(frehg_gw_solver* s, double t, double dtg);
/* RichardsSolver::nextDt() (src/gw/RichardsSolver.hpp:261). */
int frehg_gw_next_dt(const frehg_gw_solver* s, double* dt);
/* RichardsSolver::audit() (src/gw/RichardsSolver.hpp:419). */
int frehg_gw_get_audit(const frehg_gw_solver* s, frehg_gw_audit* out);
/* RichardsSolver::ownedVolume() (src/gw/RichardsSolver.cpp:892-912). */
int frehg_gw_owned_volume(frehg_gw_solver* s, double* volume);
/* RichardsSolver::lastSolve() (src/gw/RichardsSolver.hpp:402). */
int frehg_gw_last_solve(const frehg_gw_solver* s, frehg_gw_solve_stats* out);
/* Active cell count (Grid::activeCount3Local, src/core/Grid.hpp:137). */
int frehg_gw_active_cells(const frehg_gw_solver* s, int64_t* n);
/* Element count of a field in halo layout (H or F). */
int frehg_gw_field_size(const frehg_gw_solver* s, int field, int64_t* n);
/* Copy a halo-layout field to host memory (the head()/waterContent()/flux
 * accessors, src/gw/RichardsSolver.hpp:265-358). */
int frehg_gw_get_field(frehg_gw_solver* s, int field, double* host_out);
/* Overwrite h or theta from host memory and rebuild the derived state
 * (checkpoint restore + RichardsSolver::refreshDerivedState,
 * src/gw/RichardsSolver.cpp:880-890): FREHG_F_HEAD / FREHG_F_WC.
 * FREHG_F_VLOSS restores the cumulative clamp volume, the checkpoint's
 * vloss_cum (src/driver/Simulation.cpp:720-723), with nothing rebuilt.
 * FREHG_F_SINK sets the root-uptake sink S [1/s] that the following steps
 * apply (the vegetation module's sink_, src/gw/Evapotranspiration.cpp:
 * 560-634, consumed at src/gw/Predictor.cpp:700-726 and
 * src/gw/Corrector.cpp:674-705); host_in = NULL detaches it.  The step
 * applies S as given: the reference's vegetation module recomputes it before
 * every step with the supply cap min(want, (theta^n - theta_r) / dtg)
 * (Evapotranspiration.cpp:606-611), so a caller refreshes S, capped at that
 * step's dtg, before each step (a sink computed for a smaller dtg can draw
 * theta below theta_r, which the final clamp then books as vloss). */
int frehg_gw_set_field(frehg_gw_solver* s, int field, const double* host_in);
/* The evapotranspiration soil channel (RichardsSolver::qVap_ / etColumn_,
 * src/gw/RichardsSolver.hpp:620-636, computed by the ET module each
 * substep): columns = 2D (ny+2)(nx+2) int mask of channel columns, qvap =
 * 2D applied top flux [m/s].  A marked column's top face takes qvap
 * guard-free in the assembly and the Darcy flux (Predictor.cpp:660-679,
 * Corrector.cpp:458-472), and in coupled runs the channel's volume leaves
 * the seepage accumulator as evaporation (Corrector.cpp:609-615).  NULL
 * detaches the channel. */
int frehg_gw_set_et(frehg_gw_solver* s, const int* columns, const double* qvap);
/* Density / viscosity ratios from the transport module's scalars
 * (RichardsSolver::updateBaroclinicFaces, src/gw/Baroclinic.cpp:46-187, with
 * groundwater.density_coupling, src/core/Config.cpp:345-356): r_rho = 1 +
 * beta_s s - beta_T (T - T0) and r_visc = 1 / (1 + beta_sv s) per cell, halo
 * layout including ghosts (the caller evaluates them from its scalars).
 * r_rho_top / r_visc_top: the ratios of the ghost above each column's top cell
 * (2D (ny+2)(nx+2); the surface scalar's in coupled runs), NULL for the top
 * cell's own.  r_visc NULL = 1.  The face means follow Baroclinic.cpp's edge
 * rules; the ratios then enter the assembly, the Darcy fluxes, the theta
 * update, the hydrostatic ghosts and the pressurized-top test as in
 * Predictor.cpp / Corrector.cpp.  reallocation_gradient = 1 is
 * density_coupling.reallocation_gradient: density (Reallocate.cpp:221-233).
 * r_rho NULL resets every ratio to 1. */
int frehg_gw_set_density(frehg_gw_solver* s, const double* r_rho, const double* r_visc,
                         const double* r_rho_top, const double* r_visc_top,
                         int reallocation_gradient);
/* RichardsSolver::enableTransportWindow (src/gw/RichardsSolver.cpp:796-813,
 * transport.subsurface_update: consistent): from now on every step adds
 * q dtg of each face into the window fields (FREHG_F_WIN_*) until
 * frehg_gw_reset_transport_window, and the post-allocation walks record
 * each transfer on every face it crosses (FREHG_F_QZ_PATH_CORR, added to
 * the z window; the legacy mirror in the Darcy flux qzF is unchanged).
 * Idempotent. */
int frehg_gw_enable_transport_window(frehg_gw_solver* s);
/* RichardsSolver::resetTransportWindow (src/gw/RichardsSolver.cpp:815-824):
 * zero the window volumes (the transport calls it after consuming them). */
int frehg_gw_reset_transport_window(frehg_gw_solver* s);
/* Refresh per-column boundary values (updateBoundaryValues,
 * src/gw/RichardsSolver.cpp:526-566); which = 0 top, 1 bottom, 2 x-, 3 x+,
 * 4 y-, 5 y+; values is ny*nx. */
int frehg_gw_set_bc_values(frehg_gw_solver* s, int which, const double* host_values);

/* GwCoupling (src/gw/RichardsSolver.hpp:138-163): device arrays in 2D halo
 * layout (ny+2)*(nx+2), aliased and mutated by the step exactly where the
 * reference mutates them.  bounce may be NULL.  RichardsSolver::attachCoupling. */
int frehg_gw_attach_coupling(frehg_gw_solver* s, double* eta, double* dept, double* seep_accum,
                             double* avail, double* gain, double* bounce, double min_depth,
                             int vent_all);

/* Per-phase and per-kernel timing (HIP events on the solver's stream). */
typedef struct frehg_gw_timing {
  double step_ms;          /* accumulated wall of whole steps (events) */
  double solve_ms;         /* accumulated PCG time */
  double spmv_ms;          /* accumulated duration of the SpMV kernel launches */
  int64_t spmv_launches;
  double spmv_bytes;       /* algorithmic bytes per SpMV launch */
  int64_t steps;
  int64_t pcg_iterations;
} frehg_gw_timing;
/* Event timing and reset of the accumulators: 0 off; n > 0 times every step
 * and every solve, and brackets one SpMV launch in n (an event pair around a
 * launch delays the next dispatch by a few microseconds). */
int frehg_gw_set_timing(frehg_gw_solver* s, int enable);
int frehg_gw_get_timing(frehg_gw_solver* s, frehg_gw_timing* out);
/* Block until all work on the solver's stream is done. */
int frehg_gw_synchronize(frehg_gw_solver* s);

/* Step modes.  The reference's coupled substep (src/coupling/Coupler.cpp:
 * 106-119) calls RichardsSolver::step, whose KSP solve
 * (src/core/LinearSystem.cpp:320) iterates under host control.  HOST keeps
 * that shape: the PCG's convergence is polled from the host every 8
 * iterations.  DEVICE takes the same decisions on the device (one-block
 * control kernels between rounds of launches), so a step is a fixed launch
 * sequence with one host synchronization at its end.  GRAPH captures that
 * sequence once per buffer parity in a hipGraph and replays it (the step's
 * dt travels through a pinned host word, so any dt replays the same graph).
 * A device-controlled solve runs `rounds` rounds (FP64: restarted
 * recursions, each confirmed on the true residual; mixed: refinement
 * cycles) of `iterations` direction/update launch pairs; launches past the
 * end of the recursion return at once.  A solve that does not converge
 * inside that budget closes a gate: the step's later kernels return at
 * once, the host finishes the solve host-polled from where it stopped and
 * then runs the rest of the step (the budget then grows, unless the caller
 * fixed it).  DEVICE and GRAPH give bitwise the same results for the same
 * budget; with a budget the solves fit in they give HOST's (FP64) or agree
 * with it to rounding (mixed: the FP32 scale is computed on the device).
 * No counterpart in the reference. */
#define FREHG_STEP_HOST 0
#define FREHG_STEP_DEVICE 1
#define FREHG_STEP_GRAPH 2
typedef struct frehg_gw_graph_stats {
  int mode;
  int budget_rounds;
  int budget_iterations;
  int64_t captures;        /* graphs captured (one per parity and budget) */
  int64_t replays;         /* graph launches */
  int64_t device_steps;    /* steps in DEVICE or GRAPH mode */
  int64_t host_finishes;   /* of those, steps whose solve the host finished */
  int budget_later;        /* launch pairs of each round / cycle after the first */
} frehg_gw_graph_stats;
/* rounds = iterations = 0: the first step runs host-polled and sizes the
 * budget, the first round (cycle) from its own iteration count and the later
 * ones from theirs (and it grows when a solve outruns it); else the budget
 * is fixed, `iterations` for every round (rounds <= 64, iterations <= 4096;
 * mixed cycles <= 400).  Timing mode (frehg_gw_set_timing) runs host-polled
 * steps. */
int frehg_gw_set_step_mode(frehg_gw_solver* s, int mode, int rounds, int iterations);
int frehg_gw_get_graph_stats(const frehg_gw_solver* s, frehg_gw_graph_stats* out);

/* Batched column solve.  When no cell has lateral conductivity (ksx = ksy
 * = 0 everywhere) the operator is a batch of independent tridiagonal
 * columns, and with the z-line preconditioner the step solves every column
 * directly in one launch (one wave a column, active rows only, in buckets of
 * equal depth R = ceil(active rows / 64), deepest first), with one FP64
 * refinement; the solve is checked on its true residual against the PCG
 * tolerance and the PCG finishes any solve that misses it.  The reference
 * runs the same systems through KSPCG (src/core/LinearSystem.cpp:320).
 * Tiles: columns per wave of each bucket, in [1, 64]; the tuner
 * frehg2_amd.autotune.tune_columns picks them by timing steps.
 * get: the bucket count (0: no column solve) and, up to capacity, each
 * bucket's R, column count and tile. */
int frehg_gw_get_column_tiles(const frehg_gw_solver* s, int* rows_per_lane, int* columns,
                              int* columns_per_wave, int capacity, int* buckets);
int frehg_gw_set_column_tiles(frehg_gw_solver* s, const int* columns_per_wave, int buckets);
/* FREHG_PC_CHEBYSHEV: the polynomial's odd degree (3 .. 15, default 3; each
 * inner iteration then runs degree - 1 extra stencil sweeps) and its interval
 * [lower_ratio * 2, 2] on the spectrum of D^-1 A (default lower_ratio 0.05).
 * The reference offers point Jacobi, block Jacobi / ICC and (hypre) AMG
 * preconditioners for KSPCG (src/core/LinearSystem.cpp:158-221); a
 * polynomial preconditioner is an extension. */
int frehg_gw_set_chebyshev(frehg_gw_solver* s, int degree, double lower_ratio);
/* 0: solve laterally decoupled systems with the z-line PCG instead (1, the
 * default, restores the column solve; EUNSUPPORTED where there is none). */
int frehg_gw_enable_column_solve(frehg_gw_solver* s, int enable);

/* Launch shapes of the PCG (no counterpart in the reference, where PETSc
 * makes these choices): waves per CU of the FP64 and FP32 direction kernels
 * and of the column (update) kernels, each in [1, 256]; 0 keeps a value.
 * Defaults 4 / 4 / 16 where the direction kernels run over column pairs
 * (nx even, nz even and above 64; FP64 also nz <= 256), else 16 / 32 / 16.
 * Results depend on the shapes at rounding level (the order of the partial
 * sums) and are deterministic for given shapes.  The tuner
 * frehg2_amd.autotune picks them by timing solves. */
int frehg_gw_set_pcg_shapes(frehg_gw_solver* s, int march64_waves_per_cu,
                            int march32_waves_per_cu, int column_waves_per_cu);
int frehg_gw_get_pcg_shapes(const frehg_gw_solver* s, int* march64_waves_per_cu,
                            int* march32_waves_per_cu, int* column_waves_per_cu);

/* ---- device-memory plumbing -------------------------------------------
 * No counterpart in the reference (its Kokkos views own device memory).
 * They let an FFI caller without its own HIP binding stage the device
 * arrays the seams take (coupling state, COO values, rhs, solution) in the
 * library's own HIP runtime. */
int frehg_device_malloc(int64_t bytes, void** ptr);
int frehg_device_free(void* ptr);
int frehg_memcpy_htod(void* dst_device, const void* src_host, int64_t bytes);
int frehg_memcpy_dtoh(void* dst_host, const void* src_device, int64_t bytes);

/* ---- solve seam (replaces frehg::LinearSystem / PETSc KSPCG) ---------- */

/* LinearSystem(comm, prefix, nLocalRows, nGlobalRows, settings)
 * (src/core/LinearSystem.hpp:109-110; defaults :72-97).  preconditioner is
 * FREHG_PC_JACOBI or FREHG_PC_ZLINE (the z-line block Jacobi, the seam's
 * counterpart of the reference's default bjacobi + icc(0),
 * LinearSystem.cpp:206-221; it needs the 7-point pattern below). */
int frehg_ls_create(int64_t n_rows, double rtol, double atol, int max_iterations,
                    int preconditioner, frehg_linear_system** out);
int frehg_ls_destroy(frehg_linear_system* ls);
/* LinearSystem::setPattern (src/core/LinearSystem.cpp:273-296): host COO
 * index arrays; entries with row or column -1 are ignored, duplicates sum.
 * The 7-point pattern RichardsSolver::buildCooPattern emits
 * (src/gw/RichardsSolver.cpp:680-727: per row [diag, xm, xp, ym, yp, zm, zp],
 * rows in gid order) is recognised and solved on a structured embedding by
 * the step's stencil PCG; any other pattern runs a generic CSR PCG. */
int frehg_ls_set_pattern(frehg_linear_system* ls, int64_t n_coo, const int64_t* rows,
                         const int64_t* cols);
/* LinearSystem::setValues (src/core/LinearSystem.cpp:298-309): device array
 * of n_coo values in pattern order. */
int frehg_ls_set_values(frehg_linear_system* ls, const double* values_device);
/* LinearSystem::solve (src/core/LinearSystem.cpp:332-484): device rhs and
 * solution arrays of n_rows; zero initial guess, CG, true-residual test
 * ||b - A x||_2 <= max(rtol ||b||_2, atol).  Non-convergence returns
 * FREHG_EDIVERGED (fatal in the reference, :426-431). */
int frehg_ls_solve(frehg_linear_system* ls, const double* rhs_device, double* sol_device,
                   frehg_gw_solve_stats* stats);
/* LinearSystem::isSymmetric (src/core/LinearSystem.hpp:137). */
int frehg_ls_is_symmetric(frehg_linear_system* ls, double tol, int* symmetric);
/* Arithmetic of the seam's solve (no counterpart in the reference's
 * SolverSettings, LinearSystem.hpp:72-97; the YAML extension
 * solver.groundwater.precision): FREHG_PREC_FP64 (default) or
 * FREHG_PREC_MIXED (FP32 inner PCG in FP64 refinement, structured path
 * only); inner_rtol as in frehg_gw_desc (0 -> 1e-4). */
int frehg_ls_set_precision(frehg_linear_system* ls, int precision, double inner_rtol);
/* Which path the seam runs (no counterpart in the reference): *structured =
 * 2 when the pattern is the 7-point pattern and the current values are
 * symmetric (the stencil PCG runs), 1 when the pattern is but values are not
 * set or not symmetric, 0 for the generic CSR path.  *asymmetry (may be
 * NULL) is the largest |minus leg - mirror plus leg| / |diagonal| of the last
 * values. */
int frehg_ls_structured(frehg_linear_system* ls, int* structured, double* asymmetry);

/* Largest grid of frehg_ls_solve_stencil_noalloc, in halo cells
 * (nx+2)(ny+2)nz: 2^20, about 100 x 100 x 100 cells.  Its scratch array is
 * 7 vectors of this many doubles plus partial sums, ~59 MB per GPU. */
#define FREHG_NOALLOC_MAX_CELLS 1048576

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the allocation-free variant of frehg_ls_solve's
 * structured path.  It runs the same FP64 stencil PCG kernels (point Jacobi
 * or z-line, zero initial guess, the same convergence test and
 * FREHG_EDIVERGED on failure) on a system given directly in the symmetric
 * 7-point stencil storage: device arrays in halo layout (ny+2)(nx+2)nz, Ad
 * the diagonal, Axp / Ayp / Azp the coupling of a cell with its i+1 / j+1 /
 * k+1 neighbour (zero toward missing neighbours and in the ghost cells),
 * inactive cells as identity rows; rhs and sol likewise (sol is zeroed
 * first).  Every work vector lives in a static device array of the library
 * (FREHG_NOALLOC_MAX_CELLS halo cells; larger grids return FREHG_EINVAL
 * before any device work); no hipMalloc, hipFree or other allocating call is
 * made.  The call is synchronous on the null stream.  All calls on one GPU
 * share that array and it keeps the last call's contents: calls that overlap
 * on one GPU race. */
int frehg_ls_solve_stencil_noalloc(int nx, int ny, int nz, const double* Ad, const double* Axp,
                                   const double* Ayp, const double* Azp, const double* rhs,
                                   double* sol, double rtol, double atol, int max_iterations,
                                   int preconditioner, frehg_gw_solve_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* FREHG_GW_H */
