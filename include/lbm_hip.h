/*
 * lbm_hip.h -- C ABI of the MI355X lattice-update library (liblbm_hip.so).
 *
 * The library replaces the reference's per-timestep operators (geirev/LBM, called from the
 * time loop src/main.F90:290-329).  One context holds one domain: runtime grid dimensions,
 * lattice, collision options, boundary codes, one HIP stream and all device memory.
 *
 * Conventions
 *   - Every function returns 0 on success and a non-zero LBM_E* code on failure; the message
 *     is available from lbm_last_error().  (The reference's operators return nothing and
 *     `stop` on fatal paths; a Fortran binding stops on a non-zero status.)
 *   - Host arrays are borrowed for the duration of the call and use the reference's layouts:
 *       populations f(nl,0:nx+1,0:ny+1,0:nz+1), cell fields rho(0:nx+1,0:ny+1,0:nz+1),
 *       external_forcing(3,0:nx+1,0:ny+1,0:nz+1), uu(ny,nz,0:nrturb), all column major.
 *     `realkind` is the byte size of the host reals (4 or 8), as the reference's REAL kind.
 *   - Calls on one context are not thread safe; all device work is ordered on its stream.
 *     Every call that returns host data synchronises that stream.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_EINVAL 1   /* invalid argument */
#define LBM_EHIP 2     /* HIP runtime error (message holds hipGetErrorString) */
#define LBM_ESTATE 3   /* call out of order (e.g. step before a state was set) */
#define LBM_EUNSTABLE 4 /* non-finite or non-positive density found */

/* Storage of the populations in HBM; the collision always computes in FP32 (FP64 for F64). */
#define LBM_STORE_F32 0
#define LBM_STORE_F64 1
#define LBM_STORE_F16S 2  /* (f - w_l) * 2^10 as IEEE half */
#define LBM_STORE_BF16S 3 /* (f - w_l) as bfloat16 */
#define LBM_STORE_F64T 4  /* FP64 with a tangent: forward-mode derivative along lbm_set_tangent */

/* Boundary code for jbnd only: the j ghost planes are halos that the caller fills between steps
 * with lbm_halo_pack / lbm_halo_unpack (a domain tiled along j, one tile per GPU). */
#define LBM_BND_HALO 3

typedef struct lbm_ctx lbm_ctx;

/* Run-time description of one domain.  Field meanings follow the reference's infile.in
 * (src/m_readinfile.F90:86-160) and mod_dimensions (src/mod_dimensions.F90:21-26). */
typedef struct {
    int32_t nx, ny, nz;        /* interior cells (>= 2 each)                                  */
    int32_t q;                 /* 19 or 27 (the reference's -DD3Q19 switch)                   */
    int32_t storage;           /* LBM_STORE_*                                                 */
    int32_t ibgk;              /* 2 or 3: order of the Hermite equilibrium                    */
    int32_t ihrr;              /* 1: regularized non-equilibrium (Malaspinas recursion)       */
    int32_t ivreman;           /* 1: Vreman eddy viscosity                                    */
    int32_t ibnd, jbnd, kbnd;  /* 0 periodic, 1 inflow/outflow (i only), 11/12/21/22 closed,
                                  LBM_BND_HALO (j only)                                       */
    int32_t iturb_pos;         /* inflow-turbulence slice (reference: 10)                     */
    int32_t store_tau;         /* 1: keep the per-cell tau of the Vreman closure              */
    int32_t block_x, block_y;  /* step-kernel block shape; 0 = default                        */
    double kinevisc;           /* non-dimensional kinematic viscosity: tau0 = 3*kinevisc+0.5  */
    double smagorinsky;        /* Vreman constant c = 2.5*smagorinsky^2                       */
    double rho0;               /* reference density of the inflow condition (1.0)             */
} lbm_desc;

/* Context lifetime.  Replaces the reference's allocations (src/main.F90:150-177) and
 * hermite_polynomials() (src/mod_D3Q27setup.F90:71-169). */
int lbm_create(const lbm_desc* desc, lbm_ctx** out);
int lbm_destroy(lbm_ctx* ctx);
/* Message of the last failure on ctx (or of the last failed lbm_create when ctx is NULL). */
const char* lbm_last_error(const lbm_ctx* ctx);

/* State in: f in the restart layout including ghosts.  Mirrors the restart path
 * readrestart -> macrovars -> fB = fA (src/main.F90:255-258, src/m_readrestart.F90:163-185):
 * both device buffers receive f and the next step collides it in place. */
int lbm_set_state(lbm_ctx* ctx, const void* f, int realkind);

/* Equilibrium start from moments (src/main.F90:243-247): fequil3, fillghosts, fA = fB,
 * boundarycond(fB,fA), boundarycond(fA,fB).  The first step then collides with these moments,
 * as the reference's first postcoll uses inipert's rho,u,v,w.  it0 selects uvel/udir. */
int lbm_init_equilibrium(lbm_ctx* ctx, const void* rho, const void* u, const void* v, const void* w,
                         int realkind, int it0);

/* State out: the reference's f1 after the last step (post-streaming populations and ghost
 * layers), as saverestart/testing write it (src/m_saverestart.F90:125-136). */
int lbm_get_state(lbm_ctx* ctx, void* f, int realkind);

/* Moments of that state for every cell including ghosts: macrovars (src/m_macrovars.F90:4).
 * Returns LBM_EUNSTABLE (after filling the arrays) when a moment is not finite or a cell's density
 * is <= 0: the test diag makes before it stops the run (src/m_diag.F90:169-173). */
int lbm_get_macros(lbm_ctx* ctx, void* rho, void* u, void* v, void* w, int realkind);

/* Per-cell relaxation time of the last step (needs store_tau and ivreman; interior cells,
 * others hold tau0).  The reference's tau array (src/m_postcoll_kernel.F90:304). */
int lbm_get_tau(lbm_ctx* ctx, void* tau, int realkind);

/* Inputs.  Solid mask lblanking(0:nx+1,0:ny+1,0:nz+1) as bytes (non-zero = solid), NULL to
 * remove; replaces solids(f,lblanking) (src/m_solids.F90:3). */
int lbm_set_blanking(lbm_ctx* ctx, const uint8_t* mask);
/* ablvisc(nz), added to the Vreman viscosity (src/m_postcoll_kernel.F90:304). */
int lbm_set_ablvisc(lbm_ctx* ctx, const void* ablvisc, int realkind);
/* Inflow: uvel = uvel_time(it)*uvel_shear(k), direction udir_time(it) in degrees
 * (src/main.F90:313-314, src/m_inflow.F90).  shear may be NULL (all ones).  Steps outside
 * [it_lo, it_hi] (or with series == NULL) use uini and udir. */
int lbm_set_inflow(lbm_ctx* ctx, const void* shear, double uini, double udir, int realkind);
int lbm_set_inflow_series(lbm_ctx* ctx, int it_lo, int it_hi, const void* uvel_time,
                          const void* udir_time, int realkind);
/* Inflow-turbulence fields uu,vv,ww(ny,nz,0:nrturb) and amplitude: the df of
 * inflow_turbulence_forcing (src/m_inflow_turbulence_forcing.F90:5) added to slice iturb_pos
 * after the collision (inflow_turbulence_apply, src/m_inflow_turbulence_apply.F90:4).
 * uu == NULL switches it off. */
int lbm_set_turbulence(lbm_ctx* ctx, const void* uu, const void* vv, const void* ww, int nrturb,
                       double ampl, int realkind);
/* Inflow-turbulence fields generated on the device, replacing the reference's
 * inflow_turbulence_compute (src/m_inflow_turbulence_compute.F90:3) over pseudo2D
 * (src/m_pseudo2D.F90:3), which the reference runs at start-up and every nrturb steps
 * (src/main.F90:252,381-383): per variable and slice a Gaussian random field with the analytic
 * spectrum of pseudo2D (decorrelation length 10/sqrt(3) cells, r2 formed from rx as there,
 * mean mode zeroed), an unnormalized inverse real 2-D transform of size ny x nz rounded up to
 * even, then slices 1..nrturb mixed as AR(1) with 0.98 and scaled to zero mean and unit sample
 * deviation.  The random phases come from SplitMix64(seed) where the reference calls the
 * compiler's random_number.  The fields become the context's turbulence forcing (as
 * lbm_set_turbulence with ampl); uu, vv, ww, rr (ny,nz,0:nrturb) receive copies when not NULL
 * (the turbulence restart record).  FP32/FP64 compute only. */
int lbm_generate_turbulence(lbm_ctx* ctx, int nrturb, double ampl, uint64_t seed, void* uu, void* vv,
                            void* ww, void* rr, int realkind);
/* external_forcing(3,0:nx+1,0:ny+1,0:nz+1) applied every step by forcings_apply
 * (src/m_forcings_apply.F90:3); NULL switches it off. */
int lbm_set_external_forcing(lbm_ctx* ctx, const void* force, int realkind);

/* Buoyancy: the external-force producer of src/main.F90:293 for iablvisc = 2, replacing
 * buoyancy_forcing (src/m_buoyancy_forcing.F90:3-46, kernel src/m_buoyancy_forcing_kernel.F90:37):
 * every following step adds  F_z -= scaling * (pottemp - theta0) / theta0  to the forcing of each
 * interior cell (on top of the caller's external_forcing or the turbines' deposit) before the
 * forcings_apply increment.  pottemp(0:nx+1,0:ny+1,0:nz+1) is the potential temperature of the
 * step; theta0 the reference's pottemp0 (300 K, src/m_abl_initialize.F90:2); scaling
 * = g * p2l%time^2 / p2l%length with g = 9.81 (src/m_buoyancy_forcing.F90:33-34).  Call it again
 * with the new field whenever the caller's transport of pottemp (heatflux + advection,
 * src/main.F90:332-339, not in this library) has advanced it; NULL switches the buoyancy off.
 * Not in ensembles or adjoint runs. */
int lbm_set_buoyancy(lbm_ctx* ctx, const void* pottemp, double theta0, double scaling, int realkind);

/* Actuator-line turbines: the external-force producer of src/main.F90:291-292, replacing
 * turbine_forcing (src/m_turbine_forcing.F90:8-140) with turbine_distribute_points
 * (src/m_turbine_distribute_points.F90:6-64), turbine_point_forces (src/m_turbine_point_forces.F90:20-82,
 * src/m_turbine_interpolation.F90:25-114, src/m_nrelliftdrag.F90:7-35,
 * src/m_turbine_compute_bladeforce.F90:20-65) and turbine_deposit (src/m_turbine_deposit.F90:7-91).
 * Every step, before the collision, the rotor azimuths advance by omegand, each blade section's point
 * takes the moments of the step's input (rho, u, v, w as macrovars leaves them) by trilinear
 * interpolation, its angle of attack selects cl, cd from its airfoil table, and its lift and drag
 * are spread with a Gaussian stencil (sigma = 2/sqrt(2), at most 5 cells out) into the external
 * forcing, which forcings_apply then adds (src/main.F90:306).  The values are those of
 * turbine_initialize (src/m_turbine_initialize.F90:29-70): positions in grid-index units, lengths in
 * cells (m / p2l%length), angles of the rotor in radians, pitch, twist and table angles in degrees,
 * omegand = 2 pi rpm / 60 * p2l%time.  Exclusive with lbm_set_external_forcing; not in ensembles,
 * adjoint runs or halo domains.  t == NULL removes the turbines.  Host arrays are copied. */
typedef struct {
    int32_t nturbines;       /* rotors                                                        */
    int32_t nblades;         /* blades per rotor (the reference: 3)                           */
    int32_t nchords;         /* sections per blade (17 for the NREL 5-MW blade)               */
    int32_t nfoils;          /* airfoil tables                                                */
    int32_t nang_max;        /* row stride of the airfoil tables                              */
    int32_t reserved;
    const double* hub;       /* [3*nturbines] hub x, y, z (ipos, jpos, kpos)                  */
    const double* yaw;       /* [nturbines] rad                                               */
    const double* tilt;      /* [nturbines] rad                                               */
    const double* omegand;   /* [nturbines] azimuth advance per step, rad                     */
    const double* pitch;     /* [nturbines] collective pitch, deg                             */
    const double* theta;     /* [nturbines] azimuth of the state's step, rad (the restart's)  */
    const double* relm;      /* [nchords] radial position of each section, cells              */
    const double* dc;        /* [nchords] radial length of each section, cells                */
    const double* chord;     /* [nchords] chord, cells                                        */
    const double* twist;     /* [nchords] deg                                                 */
    const int32_t* nfoil;    /* [nchords] airfoil table of each section, 1-based              */
    const int32_t* nang;     /* [nfoils] rows of each table                                   */
    const double* foil;      /* [nfoils][nang_max][3]: angle of attack (deg), cl, cd          */
} lbm_turbines;
int lbm_set_turbines(lbm_ctx* ctx, const lbm_turbines* t);
/* The rotors' azimuths after the last step (the restart record theta_<tile>_<it>.uf,
 * src/m_saverestart.F90:118-122). */
int lbm_get_turbine_theta(lbm_ctx* ctx, double* theta);

/* Time steps it0+1 .. it0+nsteps: the sequence of src/main.F90:291-329 (turbine forcing,
 * turbulence df, postcoll, turbulence apply, forcings_apply, solids, boundarycond, drift, swap,
 * macrovars). */
int lbm_step(lbm_ctx* ctx, int it0, int nsteps);

/* Wait for the context's stream.  The stream itself (a hipStream_t) for callers that order
 * their own work after the library's. */
int lbm_synchronize(lbm_ctx* ctx);
void* lbm_stream(lbm_ctx* ctx);

/* Halo exchange of a domain tiled along j (jbnd = LBM_BND_HALO).  Replaces
 * mpi_halo_exchange_j (src/m_mpi_halo_exchange_j.F90:55-120) with its pack and unpack kernels
 * (src/m_mpi_pack_jplane.F90, src/m_mpi_unpack_jplane.F90); the transfer between them is the
 * caller's (RCCL send/recv, peer copies).  After each lbm_step of one step:
 *   lbm_halo_pack(ctx, 0, s) copies plane j = 1 (for the neighbour below), side 1 plane j = ny
 *   (for the neighbour above), into the device buffer s;
 *   lbm_halo_unpack(ctx, 0, r) fills ghost plane j = 0 from the lower neighbour's side-1 pack,
 *   side 1 ghost plane j = ny+1 from the upper neighbour's side-0 pack.
 * side = 2 packs (unpacks) both in one launch: the buffer then holds side 0's halo followed by
 * side 1's.  Both calls are enqueued on lbm_stream(ctx).  A halo holds only the populations
 * crossing the plane (9 of 27, 5 of 19) over i = 0..nx+1, k = 0..nz+1, in the storage encoding:
 * lbm_halo_bytes gives the size of one side's halo.  Unlike the reference's MPI path, the unpack also completes the closed face's
 * copies whose source lies on the halo plane, so a tiled run reproduces the untiled periodic run
 * bit for bit. */
int lbm_halo_bytes(lbm_ctx* ctx, size_t* bytes);
int lbm_halo_pack(lbm_ctx* ctx, int side, void* dev_buf);
int lbm_halo_unpack(lbm_ctx* ctx, int side, const void* dev_buf);

/* Step it of a halo domain in two launches, so that the halo transfer can run during the
 * second: LBM_PART_EDGE steps the cells with j = 1, 2, ny-1, ny (j = 1, ny when no closed face is
 * folded into the step; all of them when no inner cell is left), after which lbm_halo_pack packs
 * this step's boundary planes; LBM_PART_INNER steps the other cells and completes the step.  The inner cells read no halo and write nothing the pack reads, so the
 * caller's transfer (on its own stream, after an event recorded on lbm_stream following the pack)
 * overlaps the inner launch; lbm_halo_unpack then follows the inner part on lbm_stream.  The pair
 * gives lbm_step(ctx, it - 1, 1)'s result bit for bit.  Between the parts the calls that read or
 * change the state or the step's inputs (lbm_step, lbm_step_part(EDGE), lbm_set_state,
 * lbm_init_equilibrium, lbm_get_state, lbm_get_macros, lbm_get_tau, lbm_get_totals, lbm_set_* of
 * inputs, lbm_generate_turbulence, averaging) fail with LBM_ESTATE: the inner part uses the
 * inputs the edge part saw.  The halo calls, lbm_stream, lbm_synchronize, the timing calls and
 * lbm_get_info remain available.  Needs the inflow/outflow
 * planes folded into the step kernel (the default) and j off the contiguous axis.  The
 * reference's MPI path exchanges between whole steps (src/main.F90:318,
 * src/m_mpi_halo_exchange_j.F90:55-120); the split has no counterpart there. */
#define LBM_PART_EDGE 1
#define LBM_PART_INNER 2
int lbm_step_part(lbm_ctx* ctx, int it, int part);

/* Kernel timing with HIP events on the context's stream.  on = 0: off; on = n >= 1: the step
 * kernel of every n-th step (the first after the call, then every n) and that step's group of
 * boundary kernels, if it launches any, are bracketed by events (each event record costs the
 * stream a few microseconds, so n > 1 keeps the overhead off most steps).  lbm_timing_read
 * synchronises and returns the summed milliseconds and bracketed launch counts since the last
 * read. */
int lbm_set_timing(lbm_ctx* ctx, int on);
int lbm_timing_read(lbm_ctx* ctx, double* step_ms, int* n_step, double* bc_ms, int* n_bc);

/* Time averaging, replacing averaging_full (src/m_averaging_full.F90:14-97 with its kernels
 * src/m_averaging_full_kernel.F90:6-45 and src/m_averaging_full_kernelfin.F90:6-63): after
 * lbm_set_averaging(ctx, 1) every step adds u, v, w, u^2, v^2, w^2 of the moments after it to
 * per-cell sums, accumulated inside the next step's kernel, which computes those moments anyway.
 * lbm_get_averages returns the means and Ti = sqrt((sum_c max(<c^2> - <c>^2, 0)) / 3) for the
 * interior cells (layout of rho, ghosts 0) and the number of states averaged; averaging stays on
 * until lbm_set_averaging(ctx, 0).  Not available inside ensembles or adjoint runs. */
int lbm_set_averaging(lbm_ctx* ctx, int on);
int lbm_get_averages(lbm_ctx* ctx, void* uave, void* vave, void* wave, void* ti, int* count, int realkind);

/* Global mass and momentum of the interior state in FP64 (accumulated on the device).  Returns
 * LBM_EUNSTABLE (after setting the outputs) when a population is not finite or an interior cell's
 * density is <= 0 (src/m_diag.F90:169-173); the message gives the count and the smallest density. */
int lbm_get_totals(lbm_ctx* ctx, double* mass, double* momentum3);

/* lbm_get_totals that allocates nothing.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid hipMalloc).  The
 * totals kernel's per-block partial results go to a fixed array of the library on each GPU
 * (at most 1024 blocks x 6 doubles = 48 KiB, held until the process ends; the kernel
 * grid-strides, so every grid fits) instead of a context buffer allocated on first use.
 * Outputs and status as lbm_get_totals.  Every call on one GPU shares that GPU's array: calls
 * that overlap in time on one GPU (from several host threads, on any contexts) race. */
int lbm_get_totals_noalloc(lbm_ctx* ctx, double* mass, double* momentum3);

/* Bytes of HBM the two population buffers hold, the device layout pitch (elements per row) and
 * the elements of one per-cell array (the padded grid in rows of that pitch). */
int lbm_get_info(lbm_ctx* ctx, int64_t* device_bytes, int32_t* row_pitch, int64_t* plane_elems);

/*
 * Ensembles.  No counterpart in the reference, which runs one domain per process: many contexts
 * of one lattice, storage and collision model (sizes, boundary codes, solids, inflow and
 * turbulence inputs free per member) advance together.  Each step is one launch of a persistent
 * batched step kernel over every member's cells, one launch of a batched inflow/outflow kernel
 * and a device clock tick, captured in HIP graphs.  Members keep behaving as their own contexts:
 * lbm_get_state / lbm_get_macros / lbm_set_* on a member work between ensemble steps (a member
 * changed with lbm_set_* is picked up by the next lbm_ensemble_step), and the result is exactly
 * that of lbm_step on each member.  A member may have at most one closed (11/12/21/22) axis,
 * whose face pass the step kernel performs; the members stay owned by the caller and must outlive
 * the ensemble.
 */
typedef struct lbm_ens lbm_ens;
/* On failure *out may still be set: read lbm_ensemble_last_error(*out), then destroy it. */
int lbm_ensemble_create(lbm_ctx* const* members, int n, lbm_ens** out);
int lbm_ensemble_destroy(lbm_ens* ens);
const char* lbm_ensemble_last_error(const lbm_ens* ens);
/* Steps it0+1 .. it0+nsteps of every member. */
int lbm_ensemble_step(lbm_ens* ens, int it0, int nsteps);
int lbm_ensemble_synchronize(lbm_ens* ens);
/* Launch shape: block threads per tile (multiple of 64, <= 1024) and a persistent grid of
 * blocks_per_cu blocks per compute unit (0: one block per tile).  Default 256, 0. */
int lbm_ensemble_set_launch(lbm_ens* ens, int block, int blocks_per_cu);
int lbm_ensemble_get_launch(lbm_ens* ens, int* block, int* blocks_per_cu);
/* Autotuner: times trial_steps batched steps (after 2 untimed ones) for each candidate launch
 * shape from the same saved state, keeps the fastest and restores every member's state.
 * best_ms: milliseconds per step of the chosen shape. */
int lbm_ensemble_tune(lbm_ens* ens, int it0, int trial_steps, int* best_block, int* best_blocks_per_cu,
                      double* best_ms);
/* HIP-event timing of the batched steps (summed milliseconds and steps since the last read). */
int lbm_ensemble_set_timing(lbm_ens* ens, int on);
int lbm_ensemble_timing_read(lbm_ens* ens, double* ms, int* nsteps);

/*
 * Objective and adjoint.  No counterpart in the reference, which has no derivatives: they
 * differentiate its time loop (src/main.F90:290-329) as this library performs it.
 *
 * Objective: J = (1/nsteps) sum_{n=1..nsteps} (1/|B|) sum_{x in B} 0.5 rho |u|^2 (x, n), the
 * time-mean kinetic energy over box B of the moments after each step (macrovars).  box =
 * {i_lo, i_hi, j_lo, j_hi, k_lo, k_hi}, 1-based, inclusive.
 *
 * lbm_objective_run advances steps it0+1 .. it0+nsteps and returns J.
 *
 * lbm_adjoint_run does the same and also returns, by reverse-mode differentiation with a
 * checkpoint every `every` steps and recomputation in between:
 *   grad_force  dJ/d external_forcing, layout (3,0:nx+1,0:ny+1,0:nz+1) (ghost cells 0); needs
 *               lbm_set_external_forcing (else zeros); may be NULL;
 *   grad_shear  dJ/d uvel_shear(k), k = 1..nz (inflow/outflow i only, else zeros); may be NULL.
 * The context then holds the state after the nsteps steps, as after lbm_step.  Needs FP32 or
 * FP64 storage, a stepped state (not the first step after set_state / init_equilibrium), no
 * inflow turbulence, no turbines or buoyancy, at most one closed axis and at most 2^24 cells in a
 * plane of the two fastest storage axes.  Device memory beside the context's own: (ceil(nsteps/every)
 * + 1) state pairs + (every - 1) single states, kept by the context for the next run (a gradient
 * loop allocates once) and released by lbm_destroy.  Adjoint arithmetic in the storage precision, gradient accumulation
 * in FP64.
 */
int lbm_objective_run(lbm_ctx* ctx, int it0, int nsteps, const int32_t* box, double* J);

/* Tangent run (LBM_STORE_F64T): every value carries its derivative along one input direction.
 * lbm_set_tangent sets the direction: d external_forcing (layout of lbm_set_external_forcing;
 * NULL: zero) and d uvel_shear (nz; NULL: zero); lbm_set_external_forcing and lbm_set_inflow
 * reset it to zero, so call it after them.  The state's own tangent is zero after lbm_set_state.
 * lbm_objective_tangent returns dJ of the last lbm_objective_run. */
int lbm_set_tangent(lbm_ctx* ctx, const double* dforce, const double* dshear);
int lbm_objective_tangent(lbm_ctx* ctx, double* dJ);
int lbm_adjoint_run(lbm_ctx* ctx, int it0, int nsteps, int every, const int32_t* box, double* J,
                    double* grad_force, double* grad_shear);

#ifdef __cplusplus
}
#endif
#endif
