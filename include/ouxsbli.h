/*
 * ouxsbli.h -- C ABI of the MI355X (gfx950) RHS + Runge-Kutta library.
 *
 * The reference solver is a monolithic CUDA-Fortran executable: its hot path is
 * `RungeKutta` (3D_solver/src/calc_time_dev.f90.fypp:30, :224) driving
 * `calc_EFG` (3D_solver/src/calc_flux_base.f90.fypp:109), the `calc_step*`
 * kernels (3D_solver/src/calc_steps.f90:42-184) and `set_bc`
 * (3D_solver/NSTGV/set.f90:50-57). It has no library boundary; the Python layer
 * (ouxsbli/case.py:180-201) shells out to `mpirun a.out`. Every entry point
 * below replaces one piece of that executable, as cited on each declaration.
 *
 * Conventions
 *   - Every function returns OX_OK (0) on success and a negative OX_E* code on
 *     failure; ox_last_error() returns a thread-local message for the last
 *     failure on the calling thread. Nothing here calls exit() or abort().
 *   - Host pointers are borrowed for the duration of the call only (copies are
 *     synchronous). The handle owns all device memory, its stream and graphs.
 *   - A handle is not thread-safe; distinct handles may be driven from
 *     distinct threads.
 *   - State arrays use the reference layout Q(nx,5,ny,nz) in Fortran order
 *     (x fastest, then the variable, then y, then z), ghost layers included:
 *     element (i,l,j,k) (1-based) lives at
 *     (i-1) + nx*((l-1) + 5*((j-1) + ny*(k-1))).
 *     Variables are (rho, rho*u, rho*v, rho*w, E) -- the conserved set.
 */
#ifndef OUXSBLI_H
#define OUXSBLI_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OX_ABI_VERSION 2

/* Error codes. */
#define OX_OK            0
#define OX_EINVAL       -1   /* bad argument / configuration              */
#define OX_ENOMEM       -2   /* device or host allocation failed          */
#define OX_EDEVICE      -3   /* HIP runtime error                         */
#define OX_EUNSUPPORTED -4   /* configuration not built into the library  */
#define OX_ESTATE       -5   /* non-physical state (rho<=0, p<=0, NaN)    */
#define OX_EBUSY        -6   /* admission refused: footprint over budget  */
#define OX_ENOTFOUND    -7   /* unknown case id                           */

/* config.fypp SCHEME (src/mod_constant.f90.fypp:85-93) */
#define OX_SCHEME_KEEP   0
#define OX_SCHEME_SLAU   1
#define OX_SCHEME_ROE    2
#define OX_SCHEME_HYBRID 3
/* config.fypp VISC (src/mod_constant.f90.fypp:37-43) */
#define OX_VISC_EULER 0
#define OX_VISC_NS    1
#define OX_VISC_LES   2
/* config.fypp TVD (src/mod_constant.f90.fypp:51-57) */
#define OX_TVD_NONE   0
#define OX_TVD_TVD    1
#define OX_TVD_HYBRID 2
/* config.fypp SLAU_VARIANT (src/mod_constant.f90.fypp:58-62) */
#define OX_SLAU_SLAU1   0
#define OX_SLAU_HRSLAU2 1
/* VISC_STENCIL (3D_solver/src/calc_visc_high_internal.f90.fypp:5-7) */
#define OX_STENCIL_CENTRAL 0
#define OX_STENCIL_ME4BASE 1
/* Boundary kind of one direction (both ends), the ghost fill that the case's
 * set_bc applies after every stage.  PERIODIC: the periodic image
 * (set_bc_cyclic*, 3D_solver/src/set_bc_common.f90:101-273).  EXTRAPOLATE:
 * every ghost layer takes the adjacent interior cell (zero gradient,
 * 3D_solver/STZ/set.f90:113-138).  SLIP: mirror images about the boundary face
 * with the normal momentum negated (slip wall, 3D_solver/STZ/set.f90:77-112).
 * The faces of interior cells read only full-order stencils inside the ghost
 * layers, so the interior result equals the reference's BC_X/Y/Z=True flux
 * kernels (calc_*_kernel.f90.fypp) followed by that ghost fill. */
#define OX_BC_PERIODIC    0
#define OX_BC_EXTRAPOLATE 1
#define OX_BC_SLIP        2

/* Arithmetic of the RHS. FP64: everything FP64 (bit-identical to the FP64
 * CPU transcription). FP32_BULK: flux/stencil arithmetic in FP32, conserved
 * state, RK accumulation and primitive recovery in FP64. */
#define OX_PREC_FP64      0
#define OX_PREC_FP32_BULK 1

/* Compile-time switches of one case (config.fypp) plus the physical
 * constants of mod_globals.f90 that the hot path reads.  Replaces the
 * fypp/`parameter` specialisation of the reference
 * (src/mod_constant.f90.fypp:14-100, 3D_solver/NSTGV/mod_globals.f90). */
typedef struct ox_config {
  int32_t abi_version;   /* must be OX_ABI_VERSION                              */
  int32_t nx, ny, nz;    /* extents incl. order/2 ghost layers on each side     */
  int32_t order;         /* ORDER: 2, 4 or 6                                    */
  int32_t scheme;        /* OX_SCHEME_*                                         */
  int32_t visc;          /* OX_VISC_*                                           */
  int32_t visc_order;    /* VISC_ORDER: 4 or 6 (0 = same as order)              */
  int32_t visc_stencil;  /* OX_STENCIL_*                                        */
  int32_t tvd;           /* OX_TVD_*                                            */
  int32_t slau_variant;  /* OX_SLAU_*                                           */
  int32_t rk;            /* RK: 3 (TVD, Shu-Osher) or 4 (low-storage)           */
  int32_t precision;     /* OX_PREC_*                                           */
  int32_t bc_x, bc_y, bc_z; /* OX_BC_*; non-periodic: Euler/NS, not Roe or LES */
  int32_t device;        /* HIP device ordinal                                  */
  double gamma, R, Pr, Prt;
  double threshold;      /* hybrid switch on the Ducros sensor                  */
  double dt;             /* fixed time step (mod_globals.f90 `dt`)             */
  double Lx, Ly, Lz;     /* periodic box lengths                                */
} ox_config;

/* Interior diagnostics, FP64 with compensated summation.  ke and entropy are
 * the quantities of src/print.f90:42-122 (interior means); enstrophy is the
 * mean of 0.5*rho*|curl u|^2 (2nd-order central differences);
 * max_wave_speed = max(|u|+c, |v|+c, |w|+c) for CFL. */
typedef struct ox_diag {
  double ke;          /* mean 0.5*rho*|u|^2                      */
  double entropy;     /* mean rho*ln(p*rho^-gamma)               */
  double enstrophy;   /* mean 0.5*rho*|omega|^2                  */
  double max_wave_speed;
  double mass;        /* sum rho*vol                             */
  double energy;      /* sum E*vol                               */
} ox_diag;

typedef struct ox_solver ox_solver;

/* Creates a solver for one case: allocates the device state, the grid metrics
 * and scratch.  Replaces allocate_device_mem + pre_calc's metric setup
 * (3D_solver/src/preprocess.f90.fypp:21-46, :48-103) and set_grid_cyclic /
 * set_Jacobian_xy3 (src/set_coordinate.f90:99-117, :197-339).  Grids up to
 * nx*ny*nz < 2^31 cells (state indices are 64-bit; about 1100^3 fits the HBM
 * of one MI355X in FP64 NS RK3) with 15*nx*ny < 2^31 and nz <= 65535. */
int ox_create(const ox_config* cfg, ox_solver** out);

/* Frees everything the handle owns.  Replaces the deallocate at
 * 3D_solver/src/calc_time_dev.f90.fypp:110. */
void ox_destroy(ox_solver* s);

/* Uploads the physical conserved state Q(nx,5,ny,nz) and divides by the
 * Jacobian (QJ = Q / J), as pre_calc does before its H2D copy
 * (3D_solver/src/preprocess.f90.fypp:65-71, :95). */
int ox_set_state(ox_solver* s, const double* Q);

/* Downloads J*QJ, the physical conserved state written to recal/Q*.dat at the
 * end of the run (3D_solver/src/main.f90:82-93). */
int ox_get_state(ox_solver* s, double* Q);

/* Raw volume-scaled state QJ(nx,5,ny,nz) (the device array QJ of
 * calc_time_dev.f90.fypp:47), without the Jacobian multiply/divide. */
int ox_set_qj(ox_solver* s, const double* QJ);
int ox_get_qj(ox_solver* s, double* QJ);

/* Advances nsteps full RK steps (RK3: 3 stages, RK4: 4 stages).  Each stage
 * keeps the primitives of the ghost cells equal to those of their periodic
 * images, or of the boundary fill of a non-periodic direction (OX_BC_*),
 * which is what the reference's per-stage set_bc achieves; the state's own
 * ghost layers are filled once at the end of the call.  Replaces
 * the inner `t1` loop of RungeKutta (3D_solver/src/calc_time_dev.f90.fypp:84-101
 * and :266-286). */
int ox_step(ox_solver* s, int32_t nsteps);

/* One right-hand-side evaluation on the handle's current state: writes
 * R(nx,5,ny,nz) = dt*A*(div of face fluxes) at interior cells (zero in the
 * ghost layers), i.e. calc_EFG followed by calc_R
 * (3D_solver/src/calc_flux_base.f90.fypp:109-198, calc_steps.f90:8-38).
 * Parity hook; not used by ox_step. */
int ox_rhs(ox_solver* s, double* R);

/* Host helper of the reference's .d output: the sum, in order, of n REAL(8)
 * terms into a REAL(4) accumulator (each addition in double precision, the
 * running sum rounded to single), as print_KE / print_entropy accumulate
 * (src/print.f90:60-67, :101-109, with io = 4).  No device work. */
int ox_real4_sum(const double* terms, int64_t n, float* out);

/* Interior diagnostics of the current state on the device.  Replaces the
 * host-side FP32 print_KE / print_entropy (src/print.f90:42-122). */
int ox_diagnostics(ox_solver* s, ox_diag* out);

/* The same diagnostics without any allocation: no reference counterpart.  For
 * callers that cannot allocate (serving loops with a no-allocation rule, code
 * that must not call hipMalloc/hipFree or allocate host memory per call).
 * The reduction partials and the merged result live in a file-scope device
 * array of the library (352 KiB per GPU, held until the process ends; a grid
 * of any size fits, since the reduction never uses more than 4096 blocks), and
 * the final merge runs on the device in the order ox_diagnostics merges on the
 * host, so both return identical values.  Calls that overlap on one GPU (from
 * several threads, or handles on different streams) share that array and race;
 * serialise them. */
int ox_diagnostics_noalloc(ox_solver* s, ox_diag* out);

/* Scans the current state for non-physical cells (rho<=0, p<=0, non-finite).
 * Returns OX_OK, or OX_ESTATE with *bad_index set to the first offending
 * linear cell index (x fastest over nx,ny,nz) and *count to how many.
 * The reference has no such watchdog (preprocess.f90.fypp:143-148 checks one
 * rescaling array only). */
int ox_check_state(ox_solver* s, int64_t* bad_index, int64_t* count);

/* Blocks until all work queued on the handle's stream has finished. */
int ox_synchronize(ox_solver* s);

/* Device time of the last ox_step call in milliseconds (HIP events on the
 * handle's stream). */
int ox_last_step_ms(ox_solver* s, double* ms);

/* Per-kernel device time, accumulated while profiling is enabled. */
typedef struct ox_kernel_time {
  char name[32];      /* "prim", "sensor", "sweep_x", "sweep_y", "sweep_z_rk", "ghost_fill" */
  double total_ms;    /* sum of launch durations (HIP events on the handle's stream) */
  int64_t launches;
} ox_kernel_time;

/* Instrumentation (no counterpart in the reference, which relies on external
 * nsys/ncu runs, 3D_solver/NSTGV/profile_ncu.sh): while enabled, ox_step
 * records a HIP event after every kernel on the handle's stream and
 * accumulates per-kernel durations.  Enabling clears the accumulators. */
int ox_set_profiling(ox_solver* s, int32_t enable);
int ox_kernel_times(ox_solver* s, ox_kernel_time* out, int32_t cap, int32_t* n);

/* Device bytes ox_create allocates for this configuration (state, stage
 * buffers, primitives, scratch): the admission measure of the batch engine.
 * No reference counterpart (the reference runs one case per executable). */
int ox_footprint(const ox_config* cfg, uint64_t* bytes);

/* Picks the y- and z-sweep segment lengths for this handle's grid by timing
 * candidate values (16, 32, 64, 128 cells) on right-hand-side evaluations of
 * the current state, keeps the fastest in the handle and returns them.  The
 * state is not changed.  No reference counterpart (the reference's launch
 * shapes are fixed at build time, calc_flux_base.f90.fypp:44-63). */
int ox_autotune(ox_solver* s, int32_t* lseg_y, int32_t* lseg_z);

/* Sets the y- and z-sweep segment lengths (cells one thread marches along the
 * sweep direction) explicitly, as ox_autotune would after timing them: any
 * value in 1..4096, lowered where the state offsets within one segment would
 * reach 2^31 ((lseg_z + 2) * 5 * nx * ny for z; for the default 64, planes of
 * over 6.5e6 cells).  Results do not depend on the choice; the setter exists so
 * that a caller can fix a tuned value, and a test can cover segment seams
 * deterministically.  No reference counterpart (fixed launch shapes,
 * calc_flux_base.f90.fypp:44-63). */
int ox_set_segments(ox_solver* s, int32_t lseg_y, int32_t lseg_z);

/* ---------------------------------------------------------------------------
 * Batch engine: many cases resident on one GPU, advanced together.  Cases
 * that share a kernel specialisation and RK scheme form a group; one RK stage
 * of a group is one persistent launch per sweep over the tiles of all its
 * cases (LES groups first run the SGS viscosity, calc_les.f90:96-125, and
 * its ghost copy, set_bc_common.f90:296-335, as persistent launches over the
 * same cases), and one step of the whole batch is replayed from a captured HIP
 * graph while the set of cases does not change.  Cases join (ox_batch_admit)
 * and leave (when their step count is reached) between steps.  The reference
 * has no counterpart: it runs one case per `mpirun` (ouxsbli/case.py:180-201).
 * Each case's result is bit-identical to advancing it alone with ox_step.
 * A batch is driven from one host thread.
 * ------------------------------------------------------------------------- */
typedef struct ox_batch ox_batch;

typedef struct ox_batch_config {
  int32_t abi_version;   /* OX_ABI_VERSION                                    */
  int32_t device;
  int32_t max_cases;     /* resident cases at most (slots)                    */
  int32_t use_graph;     /* replay each step from a captured HIP graph        */
  int32_t autotune;      /* tune segment lengths per (specialisation, grid)   */
  int32_t pad;
  uint64_t hbm_budget;   /* device bytes the resident cases may hold          */
} ox_batch_config;

typedef struct ox_batch_stats {
  int32_t active;        /* cases still stepping                               */
  int32_t done;          /* finished cases not yet retired                     */
  int32_t groups;        /* kernel-specialisation groups among active cases    */
  int32_t graph_captures;
  int64_t steps;         /* batch steps run                                    */
  int64_t graph_replays;
  uint64_t bytes_in_use; /* footprint of resident (active + done) cases        */
  double cell_stages;    /* interior cell updates x RK stages, all cases       */
  double device_ms;      /* device time of all batch steps (HIP events)        */
} ox_batch_stats;

int ox_batch_create(const ox_batch_config* cfg, ox_batch** out);
void ox_batch_destroy(ox_batch* b);

/* Admits one case: physical state Q(nx,5,ny,nz) (ghosts included), advanced
 * by nsteps RK steps in later ox_batch_step calls.  Returns OX_EBUSY, with
 * nothing allocated, when the case does not fit in the free slots or the
 * free HBM budget now, and OX_EINVAL when its footprint exceeds the engine's
 * whole HBM budget (it could never be admitted). */
int ox_batch_admit(ox_batch* b, const ox_config* cfg, const double* Q, int32_t nsteps, int64_t* id);

/* Fills the engine's segment-length cache for cases of this configuration
 * (grid and specialisation) by tuning on the state Q, without admitting
 * anything: lets a server tune ahead of a request stream. */
int ox_batch_tune(ox_batch* b, const ox_config* cfg, const double* Q);

/* One RK step of every active case (all groups), then blocks until it has
 * finished.  Cases whose step count is reached move to the done list.
 * *active receives the number of cases still stepping. */
int ox_batch_step(ox_batch* b, int32_t* active);

/* Like ox_batch_step for up to max_steps steps, enqueued back to back on the
 * device with one synchronization at the end.  Stops early after the first
 * step in which a case reaches its step count (so that it can be retired and
 * others admitted); *steps_done receives the number of steps run. */
int ox_batch_run(ox_batch* b, int32_t max_steps, int32_t* steps_done, int32_t* active);

/* Ids of finished cases not yet retired (at most cap, oldest first). */
int ox_batch_done(ox_batch* b, int64_t* ids, int32_t cap, int32_t* n);

/* Copies a finished case's physical state Q(nx,5,ny,nz) out (ghost layers
 * filled) and frees the case.  OX_ENOTFOUND for an unknown id, OX_ESTATE for
 * a case still stepping. */
int ox_batch_retire(ox_batch* b, int64_t id, double* Q);

int ox_batch_query(ox_batch* b, ox_batch_stats* out);

/* ---------------------------------------------------------------------------
 * Two-dimensional cases with boundaries.  The reference builds these from
 * 2D_solver/src (one executable per case directory); the case implemented
 * here is its inviscid oblique-shock reflection, 2D_solver/OS.
 *
 *   - State Q(nx,4,ny) in Fortran order (x fastest, then the variable, then
 *     y): element (i,l,j) (1-based) at (i-1) + nx*((l-1) + 4*(j-1)).
 *     Variables are (rho, rho*u, rho*v, E).
 *   - nx, ny count one boundary layer on each side.  Interior points are
 *     2..nx-1 and 2..ny-1, on the uniform grid x = (i-1)*Lx/(nx-1),
 *     y = (j-1)*Ly/(ny-1) (2D_solver/OS/set.f90:8-26).
 *   - Faces next to a boundary fall back to lower order: ORDER 6 uses the
 *     6-point MUSCL for faces 3..n-3, the 4-point one for faces 2 and n-2, and
 *     first order for faces 1 and n-1 (2D_solver/src/calc_slau_kernel.f90.fypp:17-200).
 *   - After every RK stage the boundary points take the oblique-shock
 *     conditions of 2D_solver/OS/set.f90:48-86:
 *       left column      the inflow state;
 *       top row          inflow for i < top_split, the post-shock state otherwise;
 *       bottom row       slip wall (copy of row 2 with v negated);
 *       right column     copy of column nx-1.
 * ------------------------------------------------------------------------- */
#define OX2D_BC_OBLIQUE_SHOCK 1

typedef struct ox2d_config {
  int32_t abi_version;   /* OX_ABI_VERSION                                          */
  int32_t nx, ny;        /* points incl. the boundary layer on each side            */
  int32_t order;         /* ORDER: 2, 4 or 6                                        */
  int32_t scheme;        /* OX_SCHEME_SLAU or OX_SCHEME_ROE                         */
  int32_t tvd;           /* OX_TVD_*                                                */
  int32_t slau_variant;  /* OX_SLAU_*                                               */
  int32_t rk;            /* 3 (TVD, Shu-Osher)                                      */
  int32_t precision;     /* OX_PREC_*                                               */
  int32_t bc;            /* OX2D_BC_OBLIQUE_SHOCK                                   */
  int32_t top_split;     /* 1-based; int(0.1*nx) in OS/set.f90:55                   */
  int32_t device;        /* HIP device ordinal                                      */
  double gamma, R;
  double threshold;      /* hybrid MUSCL switch on the Ducros sensor                */
  double dt;
  double Lx, Ly;
  double inflow[4];      /* rho, u, v, p of the left boundary and the top for i < top_split */
  double post[4];        /* rho, u, v, p of the top for i >= top_split              */
} ox2d_config;

typedef struct ox2d_solver ox2d_solver;

/* Creates a 2-D case: device state, grid metrics.  Replaces
 * allocate_device_mem + pre_calc (2D_solver/src/preprocess.f90.fypp:25-111),
 * set_grid (2D_solver/OS/set.f90:8-26) and set_Jacobian_xy2
 * (src/set_coordinate.f90:78-96). */
int ox2d_create(const ox2d_config* cfg, ox2d_solver** out);
void ox2d_destroy(ox2d_solver* s);

/* Physical conserved state Q(nx,4,ny) in and out (QJ = Q / J on the way in,
 * as pre_calc, preprocess.f90.fypp:86-92; J * QJ on the way out, as
 * 2D_solver/src/main.f90:80-85). */
int ox2d_set_state(ox2d_solver* s, const double* Q);
int ox2d_get_state(ox2d_solver* s, double* Q);

/* Advances nsteps RK3 steps, each stage followed by the boundary conditions.
 * Replaces the `t1` loop of RungeKutta (2D_solver/src/calc_time_dev.f90.fypp:77-94). */
int ox2d_step(ox2d_solver* s, int32_t nsteps);

/* Device time of the last ox2d_step call in milliseconds. */
int ox2d_last_step_ms(ox2d_solver* s, double* ms);

/* Device bytes ox2d_create allocates for this configuration. */
int ox2d_footprint(const ox2d_config* cfg, uint64_t* bytes);

/* Admits a 2-D case into a batch; like ox_batch_admit.  ox_batch_retire of its
 * id copies out Q(nx,4,ny).  Q == NULL builds the oblique-shock initial state
 * on the device instead (ox2d_init_state). */
int ox_batch_admit2d(ox_batch* b, const ox2d_config* cfg, const double* Q, int32_t nsteps, int64_t* id);

/* ---------------------------------------------------------------------------
 * Initial states built on the device.  The case set_init routines sample the
 * initial field at the cell centres of the interior on the host, then the
 * host fills the ghosts and pre_calc divides by the Jacobian
 * (3D_solver/NSTGV/set.f90:16-47, 3D_solver/src/set_bc_common.f90:11-97,
 * 3D_solver/src/preprocess.f90.fypp:65-71; 2D_solver/OS/set.f90:28-46).
 * These entry points do all of it in one kernel plus the ghost fill, with no
 * host array and no upload, so a server can admit a request from its
 * parameters alone.
 * ------------------------------------------------------------------------- */
#define OX_IC_TAYLOR_GREEN 1 /* NSTGV set_init: p, rho, u, v (w = 0) at the cell centres            */
#define OX_IC_SHEAR_LAYER 2  /* double tanh shear layer in y with seeded velocity noise (no reference
                                counterpart: the Kelvin-Helmholtz request stream's field)           */
#define OX_IC_SPECTRAL 3     /* velocity synthesised from spectral modes (replaces the inverse FFT of
                                3D_solver/DHIT/set_init_dhit.f90:170-179): on a cubic n^3 interior,
                                u(x) = sum over kx = 0..K, ky, kz = -K..K of w_kx Re(H(k) e^{i k.x}) / n^3
                                at x = 2 pi (interior index) / n, w_0 = 1, w_kx = 2 otherwise (the
                                inverse real transform of the half spectrum), then scaled so that
                                sqrt(<|u|^2>) = V0; uniform density rho0 and pressure p0            */
#define OX_IC_ETGV 4         /* ETGV set_init (3D_solver/ETGV/set.f90:27-40): rho = rho0,
                                rho u = rho0 M0 sin x cos y cos z, rho v = -rho0 M0 cos x sin y cos z,
                                w = 0, p = 1/gamma + rho0 M0^2 (cos 2x + cos 2y)(cos 2z + 2)/16,
                                with M0 = V0, at the cell centres of the [0, 2 pi) box               */
#define OX_IC_KHI 5          /* KHI set_init (3D_solver/KHI/set.f90:31-47): stream 1 (V0, rho0) where
                                y > 0.75 Lx or y < 0.25 Lx (Lx, as the reference writes it), stream 2
                                (u2, rho2) between, pressure p0, perturbations
                                v = amp V0 sin(2 pi x / Lx), w = amp V0 sin(2 pi z / Lz)             */

typedef struct ox_ic {
  int32_t kind;     /* OX_IC_*                                                                  */
  int32_t kmax;     /* spectral: K of the mode box (K < n/2); 0 for the other kinds             */
  uint64_t seed;    /* shear layer: noise seed                                                  */
  double L0;        /* Taylor-Green: length scale of sin(x/L0), cos(2x/L0), ...                 */
  double V0;        /* Taylor-Green: V0; shear layer: U, the velocity difference across a layer;
                       ETGV: M0; KHI: u1                                                        */
  double rho0;      /* Taylor-Green: RHO0 of the pressure amplitude; shear layer: density;
                       ETGV: RHO0; KHI: rho1                                                    */
  double p0;        /* Taylor-Green: mean pressure; shear layer, KHI: pressure                  */
  double T0;        /* Taylor-Green: temperature, rho = p / (R T0)                              */
  double delta;     /* shear layer: thickness d of tanh((y - Ly/4) / d)                         */
  double noise;     /* shear layer: amplitude of the uniform [-1, 1) velocity noise             */
  const double* modes; /* spectral: host array H[3][K+1][2K+1][2K+1] of complex coefficients
                          (re, im interleaved; component, kx, ky + K, kz + K); copied during the
                          call                                                                     */
  double u2;        /* KHI: velocity of stream 2 (V0, rho0 are stream 1's; p0 the pressure)     */
  double rho2;      /* KHI: density of stream 2                                                 */
  double amp;       /* KHI: perturbation amplitude                                              */
} ox_ic;

/* Builds the initial state of `ic` in the handle (interior at the cell
 * centres of ox_create's grid, then the periodic ghost fill), replacing
 * ox_set_state of a host array. */
int ox_init_state(ox_solver* s, const ox_ic* ic);

/* ox_batch_admit with the initial state built on the device from `ic`. */
int ox_batch_admit_ic(ox_batch* b, const ox_config* cfg, const ox_ic* ic, int32_t nsteps, int64_t* id);

/* The oblique-shock initial state of the configuration (its inflow state
 * everywhere, its post-shock state on the top row from top_split on),
 * replacing ox2d_set_state of a host array (2D_solver/OS/set.f90:28-46). */
int ox2d_init_state(ox2d_solver* s);

/* Page-locked host memory (hipHostMalloc): a device-to-host copy of a result
 * into it runs at the host link's full rate, where a copy into pageable
 * memory is staged.  No counterpart in the reference; a serving helper for
 * callers of ox_batch_retire / ox_get_state that keep result buffers. */
int ox_host_alloc(uint64_t bytes, void** ptr);
int ox_host_free(void* ptr);

/* Thread-local text of the last error on this thread ("" if none). */
const char* ox_last_error(void);

/* Library version string ("ouxsbli-mi355x <abi> <arch>"). */
const char* ox_version(void);

#ifdef __cplusplus
}
#endif
#endif /* OUXSBLI_H */
