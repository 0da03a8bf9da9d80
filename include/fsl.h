/*
 * fsl.h - C ABI of the MI355X FLIP projection core.
 *
 * One context holds one simulation domain on one GPU: sparse grids stored as
 * 8^3 leaf blocks with a device list of active leaves, particles in SoA
 * arrays sorted by cell, and the Jacobi-PCG pressure solver. The reference's
 * Canary partitions become tiles of the one domain (per-tile statistics are
 * reported; the physics is the one-partition physics).
 *
 * Conventions
 *  - Every call returns FSL_OK (0) or a negative FSL_ERR_* code and never
 *    aborts; fsl_last_error() returns a message for the last failure.
 *  - Host arrays are copied, never aliased. Calls on one context are not
 *    thread-safe; contexts are independent. One HIP stream per context.
 *  - "Box" arrays are dense over the index box [-1, n] on every axis, x-major:
 *    index(i,j,k) = ((i+1)*(ny+2) + (j+1))*(nz+2) + (k+1). Vector grids store
 *    3 floats per voxel; component d of voxel (i,j,k) is the velocity on the
 *    lower d-face of cell (i,j,k) (vector_grid.cc:20-22).
 *  - Particle arrays are float[n][3].
 *
 * Each entry point names the reference symbol it replaces (file:line in the
 * reference repository schinmayee/fluid-sim-lb).
 */
#ifndef FSL_H_
#define FSL_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSL_OK 0
#define FSL_ERR_ARG (-1)            /* bad argument or size */
#define FSL_ERR_HIP (-2)            /* HIP runtime failure */
#define FSL_ERR_STATE (-3)          /* call out of order */
#define FSL_ERR_CAPACITY (-4)       /* particle capacity exceeded */
#define FSL_ERR_ZERO_DIAGONAL (-10) /* FLUID cell whose 6 neighbours are SOLID */
#define FSL_ERR_NAN (-11)           /* NaN residual (solver_driver.cc:327) */
#define FSL_ERR_ALPHA (-12)         /* alpha > 1e5 (solver.cc:333) */
#define FSL_ERR_BETA (-13)          /* beta > 1e5 (solver.cc:361) */
#define FSL_ERR_VELOCITY_UPDATE (-14) /* face active in v, not saved (flip_simulation.cc:788-792) */
#define FSL_ERR_PRECONDITIONER (-15) /* IC(0) pivot below 1e-5: the incomplete Cholesky
                                        build failed (OpenVDB's compatibility test) */

/* PCG arithmetic. */
#define FSL_PCG_FP32 0  /* reference precision: FP32 vectors and dots, ordered stencil, no FMA */
#define FSL_PCG_FP64 1  /* FP64 vectors, stencil and dots */
#define FSL_PCG_MIXED 2 /* FP32 stencil/axpys, FP64 dots, FP64 true-residual refresh */

/* PCG iteration structure. */
#define FSL_PCG_AUTO 0     /* FP32 or IC: two-pass; FP64 and MIXED: single-reduction */
#define FSL_PCG_TWO_PASS 1 /* the reference's stage order (solver_driver.cc:318-343): Ap and
                              p.Ap, then the axpys with r.r and r.z; two reductions */
#define FSL_PCG_SINGLE_REDUCTION 2 /* one pass per iteration: p, Ap, x, r and A(M^-1 r) from
                              a two-cell halo, alpha from the Chronopoulos-Gear recurrence;
                              one reduction */

/* Preconditioner (BuildPreconditioner, solver.cc:101-112). The reference's
 * driver always ends up with Jacobi (solver_driver.cc:250 default-constructs
 * the solver, so the type it was given is never used); IC selects the branch
 * it would otherwise take: block IC(0) per tile (OpenVDB 4.0.1
 * IncompleteCholeskyPreconditioner with openvdb.patch:209-224), two-pass only. */
#define FSL_PRECOND_JACOBI 0
#define FSL_PRECOND_IC 1

/* Summation order of the PCG dot products (two-pass FP32 and FP64 solves). */
#define FSL_DOT_TREE 0      /* per-thread sums, wave and workgroup trees, fixed order */
#define FSL_DOT_ROW_ORDER 1 /* sequential over the rows in the row-index order of
                               CreateIndexAndMaskTree (solver.cc:114-151), one rounding
                               per product and per sum: the oracle's order. For stage
                               parity on small systems (one workgroup sums; slow) */

/* PCG stopping rule. */
#define FSL_STOP_ABS_RR 0 /* stop when r.r < threshold (solver_driver.cc:330) */
#define FSL_STOP_REL 1    /* stop when ||r|| < threshold * ||b|| */

/* Grid selectors. */
#define FSL_GRID_VELOCITY 0        /* grid_velocity_ */
#define FSL_GRID_VELOCITY_UPDATE 1 /* grid_velocity_update_ (saved v, then dv) */
#define FSL_GRID_DIVERGENCE 0      /* grid_divergence_ */
#define FSL_GRID_PRESSURE 1        /* grid_pressure_ */
#define FSL_GRID_MARKER 0          /* grid_marker_: 0 AIR, 1 FLUID, 2 SOLID */
#define FSL_GRID_SOLID 1           /* grid_solid_: 0/1 */
#define FSL_GRID_ROW_INDEX 2       /* solver row index, -1 where not a row */
#define FSL_GRID_CELL_COUNT 3      /* particles per cell */

typedef struct fsl_ctx fsl_ctx;

typedef struct fsl_solve_params {
  int precision;      /* FSL_PCG_* */
  int stop_mode;      /* FSL_STOP_* */
  double threshold;   /* absolute r.r bound, or relative ||r||/||b|| bound */
  int max_iterations; /* reference default 1500 (flip_app.h:52) */
  int refresh_every;  /* MIXED: FP64 residual refresh period (0 = 32) */
  int batch;          /* iterations launched between host checks; 0: sized from
                         the observed residual decay (first batch 16) */
  int warm_start;     /* 0: x0 = 0 (the cleared pressure grid, solver.cc:86-88);
                         1: x0 = the current pressure grid */
  int variant;        /* FSL_PCG_AUTO / TWO_PASS / SINGLE_REDUCTION */
  int preconditioner; /* FSL_PRECOND_JACOBI (0, the reference's effective one) / IC */
  int dot_order;      /* FSL_DOT_TREE (0) / FSL_DOT_ROW_ORDER */
} fsl_solve_params;

typedef struct fsl_solve_stats {
  int iterations;
  int n_rows;          /* FLUID cells in the domain (solver.cc:80) */
  int n_cols;          /* FLUID cells in the domain + 1 ring (solver.cc:81) */
  int n_fluid_leaves;  /* 8^3 leaves holding at least one FLUID cell */
  double rr;           /* final recursive r.r */
  double true_rel_residual; /* ||b - A x|| / ||b||, FP64, after the solve */
  double bnorm;        /* ||b|| */
  double seconds;      /* wall time of the iteration loop on the stream */
} fsl_solve_stats;

/* ---- context ---- */
/* FlipSimulation::InitializeMetadata + InitializeVariables
 * (flip_simulation.cc:83-122); dims must be multiples of 8; tiles are the
 * reference partitions (metagrid.cc:25-35). */
int fsl_create(fsl_ctx **out, const int dims[3], const int tiles[3], int device,
               int64_t particle_capacity);
int fsl_destroy(fsl_ctx *ctx);
const char *fsl_last_error(const fsl_ctx *ctx);
/* Waits until all work enqueued on the context's stream has finished. */
int fsl_synchronize(fsl_ctx *ctx);
/* Scaled gravity (flip_app_sim.cc:74), force = config force * scaled gravity
 * (flip_app_sim.cc:87, SetForces flip_simulation.cc:248-253), flip factor. */
int fsl_set_params(fsl_ctx *ctx, float gravity_scaled, const float force[3], float flip);
/* SetBoxBoundary / SetTankBoundary (flip_simulation.cc:255-377). */
int fsl_set_box_boundary(fsl_ctx *ctx, int tank);
/* AddSolidCuboid (flip_simulation.cc:315-324); inclusive integer box. */
int fsl_add_solid_box(fsl_ctx *ctx, const int lo[3], const int hi[3]);

/* ---- particles ---- */
/* Particles::AddParticleToData for a batch (particles.h); replaces all
 * particles. gv may be NULL (then gv = vel). */
int fsl_set_particles(fsl_ctx *ctx, int64_t n, const float *pos, const float *vel,
                      const float *gv);
/* InitializePhiAndParticles for a given fluid mask (flip_simulation.cc:138-188):
 * ppc Philox-jittered particles per mask cell, seeded on the device; mask is
 * uint8 over [0, n) x-major; velocity and grid velocity are `velocity`. */
int fsl_seed_particles(fsl_ctx *ctx, const uint8_t *mask, int ppc, uint32_t seed,
                       const float velocity[3]);
/* AddContributionFromSources (flip_simulation.cc:1206-1266) for one source
 * box already evaluated at the step's time (Source::get_box, sources.cc:64-82,
 * i.e. with its end reduced by 1): cells from (int)lo to (int)hi, clipped to
 * the domain, each get up to ppc particles at Philox points (key seed, stream)
 * inside the box (bounds inclusive), out of at most 32 attempts, with
 * velocity and grid velocity `velocity`. Particles are appended and join the
 * cell order at the next move. FSL_ERR_CAPACITY if they do not fit. */
int fsl_add_source(fsl_ctx *ctx, const float lo[3], const float hi[3], const float velocity[3],
                   int ppc, uint32_t seed, uint32_t stream);
/* DeleteFluidInSources / Particles::DeleteParticlesFromRegions
 * (flip_simulation.cc:1268-1277, particles.cc:222-271): deletes the particles
 * whose position lies in any of the n boxes (boxes[6 b .. 6 b + 5] = lo x, y,
 * z, hi x, y, z; bounds inclusive), keeping the others' order. */
int fsl_delete_particles_in_boxes(fsl_ctx *ctx, int n, const float *boxes);
/* Particles::size (particles.h). */
int fsl_num_particles(fsl_ctx *ctx, int64_t *n);
/* Particle data in device order (sorted by cell). Any pointer may be NULL. */
int fsl_get_particles(fsl_ctx *ctx, float *pos, float *vel, float *gv);

/* ---- FlipSimulation stages (flip_simulation.h:133-182) ---- */
/* CFL (flip_simulation.cc:459-488): returns 1/max(sqrt(g/2), v_inf, 1e-2). */
int fsl_cfl(fsl_ctx *ctx, float *inv_speed);
/* MoveParticles + DeleteParticlesInSolids (flip_app_sim.cc:496-517,
 * particles.cc:478-546, 202-219): advect, clamp, delete, sort by cell. */
int fsl_move_particles(fsl_ctx *ctx, float dt);
/* ClearGridData (flip_simulation.cc:379-389). */
int fsl_clear(fsl_ctx *ctx);
/* TransferParticlesToGrid (flip_simulation.cc:808-923), incl. FLUID marking. */
int fsl_p2g(fsl_ctx *ctx);
/* ApplyBoundaryConditions (flip_simulation.cc:656-713). */
int fsl_apply_bc(fsl_ctx *ctx);
/* SaveGridVelocities (flip_simulation.cc:490-496). */
int fsl_save_velocities(fsl_ctx *ctx);
/* DeleteOutsideParticles (particles.cc:273-293). */
int fsl_delete_outside(fsl_ctx *ctx);
/* AddGravity (flip_simulation.cc:505-515). */
int fsl_add_gravity(fsl_ctx *ctx, float dt);
/* ComputeDivergence (flip_simulation.cc:715-736). */
int fsl_divergence(fsl_ctx *ctx);
/* PCGDriver::Solve (solver_driver.cc:13-69): assemble, prepare, iterate,
 * copy solution into the pressure grid. */
int fsl_pcg_solve(fsl_ctx *ctx, const fsl_solve_params *params, fsl_solve_stats *stats);
/* UpdateVelocityFromPressure (flip_simulation.cc:738-778). */
int fsl_pressure_update(fsl_ctx *ctx);
/* GetVelocityUpdate (flip_simulation.cc:780-801). */
int fsl_velocity_update(fsl_ctx *ctx);
/* VectorGrid::ExtrapolateToBoundary on v and dv (vector_grid.cc:302-381). */
int fsl_extrapolate(fsl_ctx *ctx);
/* UpdateParticlesFromGrid (flip_simulation.cc:1024-1189), incl. extrapolation. */
int fsl_g2p(fsl_ctx *ctx);

/* FlipDriver::RunOneStep (flip_app_sim.cc:211-254) with a given dt. The
 * particle count and the leaf lists stay on the device from the advection
 * to the divergence; the host reads them once per step, to size the solve. */
int fsl_step(fsl_ctx *ctx, float dt, const fsl_solve_params *params, fsl_solve_stats *stats);

typedef struct fsl_frame_stats {
  int steps;              /* RunOneStep calls in the frame */
  int pcg_iterations;     /* summed over the frame's solves */
  float dt_last;          /* dt of the frame's last step */
  fsl_solve_stats last_solve;
} fsl_frame_stats;
/* One frame of FlipDriver's loop (flip_app_sim.cc:357-462): While
 * (step + 1e-5 < frame_step) { WhileAndComputeDt; RunOneStep }, with dt
 * computed on the device from the CFL maximum (cfl * CFL(), min with
 * 1/frame_step, floor at frame_step/10, cut and snapped at the frame's end)
 * and read by the advection and gravity there; the elapsed time starts at 0. */
int fsl_advance_frame(fsl_ctx *ctx, int cfl, float frame_step, const fsl_solve_params *params,
                      fsl_frame_stats *stats);
/* With enable set, fsl_advance_frame replays the front half of each step
 * (CFL and dt, the move and sort, clear, P2G, boundary conditions, gravity,
 * divergence: about 25 launches and copies) from a captured HIP graph
 * instead of launching it, and single-reduction solves after the first one
 * enqueue as many iterations as the previous solve took, then the
 * convergence check, and read the solver state once (continuing 32
 * iterations at a time when that was too few), instead of reading it after
 * every batch; the results are the same. No reference counterpart (the
 * reference's transforms are scheduled by Canary). */
int fsl_set_step_graph(fsl_ctx *ctx, int enable);
/* WhileAndComputeDt's dt (flip_app_sim.cc:371-410): cfl * CFL(), min with
 * 1/frame_step, floor at frame_step/10, snap to the frame end; advances *step. */
int fsl_compute_dt(fsl_ctx *ctx, int cfl, float frame_step, float *step, float *dt);

/* ---- grid access in box format ---- */
/* fsl_get_vec / fsl_set_vec also take arrays in device memory of the
 * context's device (for example a torch tensor's storage): those are read or
 * written in place on the context's stream, with no staging copy. The
 * caller's writes to them must be complete (its stream synchronized) before
 * the call; the call returns after its own work on them is complete. With
 * fsl_get_vec, out and a non-null active must both be host or both be
 * device arrays. */
int fsl_get_vec(fsl_ctx *ctx, int which, float *out, uint8_t *active);
int fsl_set_vec(fsl_ctx *ctx, int which, const float *in, const uint8_t *active);
int fsl_get_scalar(fsl_ctx *ctx, int which, float *out);
/* fsl_get_scalar without any allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (serving loops and
 * harnesses that forbid hipMalloc). It stages the box through a fixed
 * device array of 8,388,608 floats (32 MiB, one per GPU, held until the
 * process ends) instead of a buffer allocated per call, so it takes boxes of
 * at most that many voxels (domains up to 200^3) and returns FSL_ERR_ARG
 * beyond. Calls that overlap on one GPU share the array and race. */
int fsl_get_scalar_noalloc(fsl_ctx *ctx, int which, float *out);
int fsl_set_scalar(fsl_ctx *ctx, int which, const float *in);
int fsl_get_int(fsl_ctx *ctx, int which, int32_t *out);
/* Sets the marker grid (labels); SOLID cells must also be solid. */
int fsl_set_marker(fsl_ctx *ctx, const int32_t *in);
/* Active 8^3 leaves (origins, int[3] each, in device list order). */
int fsl_get_active_leaves(fsl_ctx *ctx, int which, int32_t *origins, int64_t cap, int64_t *count);
/* Per-tile FLUID row counts (n_rows of each reference partition). */
int fsl_tile_rows(fsl_ctx *ctx, int64_t *rows_per_tile);

/* ---- PCG stage API (solver.cc:262-366, solver_driver.cc:318-343) ---- */
/* Assembly + PreparePCGSolver: r = b - A x, z = M^-1 r, p = z, rz, rr. */
int fsl_pcg_prepare(fsl_ctx *ctx, const fsl_solve_params *params, double *rr, double *rz);
/* PCGSolver::ComputeLocal1 (solver.cc:296-309): Ap = A p and pAp (the
 * global value, after the driver's reduction, solver_driver.cc:318-343).
 * p = z + beta p of the previous iteration's ComputeLocal3 is applied here,
 * where the pass reads p. The stage calls follow fsl_pcg_prepare in the
 * order Local1, Local2, Local3 (FSL_ERR_STATE otherwise); the While test
 * (solver_driver.cc:318-337) is the caller's. */
int fsl_pcg_local1(fsl_ctx *ctx, double *pAp);
/* PCGSolver::ComputeLocal2 (solver.cc:323-346): alpha = rz / pAp (alpha above
 * 1e5 returns FSL_ERR_ALPHA), x += alpha p, r -= alpha Ap, z = M^-1 r, and the
 * reduced r.r and r.z. */
int fsl_pcg_local2(fsl_ctx *ctx, double *alpha, double *rr, double *rz);
/* PCGSolver::ComputeLocal3 (solver.cc:355-366): beta = rz / rz_old (beta above
 * 1e5 returns FSL_ERR_BETA); the direction update is applied by the next
 * fsl_pcg_local1. */
int fsl_pcg_local3(fsl_ctx *ctx, double *beta);
/* One While-loop body: ComputeLocal1 (Ap, pAp), ComputeLocal2 (alpha, x, r,
 * z, rr, rz), ComputeLocal3 (beta; p = z + beta p is applied lazily by the
 * next ComputeLocal1). Returns the scalars after the iteration. */
int fsl_pcg_iterate(fsl_ctx *ctx, double *pAp, double *alpha, double *rr, double *rz,
                    double *beta);
/* CopySolution (solver.cc:93-99). */
int fsl_pcg_finish(fsl_ctx *ctx, fsl_solve_stats *stats);

/* ---- measurement hook (no reference counterpart) ----
 * When enabled, fsl_step and fsl_pcg_solve record HIP events on the
 * context's stream around each stage and each PCG pass; fsl_get_timing
 * returns the accumulated milliseconds (and launch counts) since enabling. */
#define FSL_T_MOVE 0      /* advection + counting sort */
#define FSL_T_CLEAR 1     /* clear + active-leaf list */
#define FSL_T_P2G 2
#define FSL_T_BC_SAVE 3
#define FSL_T_GRAVITY_BC 4
#define FSL_T_DIVERGENCE 5 /* divergence + labels + fluid-leaf list */
#define FSL_T_SOLVE 6     /* whole PCG solve */
#define FSL_T_UPDATE 7    /* pressure update + velocity update */
#define FSL_T_G2P 8       /* extrapolation + G2P */
#define FSL_T_PASS_A 9    /* PCG pass A (p update + SpMV + p.Ap), summed */
#define FSL_T_PASS_B 10   /* PCG pass B (axpys + r.r, r.z), summed */
#define FSL_T_PASS_A_N 11 /* pass A launches that did work */
#define FSL_T_PASS_B_N 12
#define FSL_T_REDUCE 13   /* PCG scalar reductions (one workgroup each), summed */
#define FSL_T_STEPS 14    /* fsl_step calls timed */
#define FSL_T_FUSED 15    /* single-reduction PCG pass, summed */
#define FSL_T_FUSED_N 16  /* single-reduction passes that did work */
#define FSL_T_COUNT 17
int fsl_set_timing(fsl_ctx *ctx, int enable);
int fsl_get_timing(fsl_ctx *ctx, double *out, int n);

/* ---- differentiable projection (no reference counterpart) ---- */
/* Forward: u_post = P(u) with the current labels, i.e. divergence, PCG,
 * pressure update, on the velocity grid in place. */
int fsl_project_forward(fsl_ctx *ctx, const fsl_solve_params *params, fsl_solve_stats *stats);
/* Backward: given g = dL/du_post (box format, 3 floats per voxel), writes
 * dL/du = P^T g (box format). One adjoint PCG solve on the same operator,
 * with the labels of the last fsl_divergence / fsl_project_forward. The
 * divergence and pressure grids are left as the forward pass wrote them.
 * grad_out and grad_in may be host arrays or device arrays of the context's
 * device (read and written in place, as with fsl_get_vec). */
int fsl_project_backward(fsl_ctx *ctx, const fsl_solve_params *params, const float *grad_out,
                         float *grad_in, fsl_solve_stats *stats);

/* ---- rollouts with checkpointing (no reference counterpart) ----
 * A rollout is a sequence of FLIP substeps (the fsl_step stages). Each
 * substep's particles are checkpointed right after advection; the backward
 * pass recomputes every substep's labels, active faces and P2G weights from
 * its checkpoint and applies the transposed stages in reverse order. The
 * gradient is that of the substeps with the particle positions (hence the
 * labels and active faces) of the recorded trajectory held fixed: under that
 * condition a substep is linear in the particle velocities, the grid
 * velocity and the body force. */
/* Allocates checkpoints for up to max_substeps and forgets earlier ones. */
int fsl_rollout_begin(fsl_ctx *ctx, int max_substeps);
/* One substep (FlipDriver::RunOneStep with dt) with its checkpoint. For the
 * first substep, u0 (box format, may be NULL) replaces the P2G velocity on
 * the faces P2G made active. u_post (may be NULL) receives the velocity right
 * after the pressure update (box format). */
int fsl_rollout_step(fsl_ctx *ctx, float dt, const fsl_solve_params *params, const float *u0,
                     float *u_post, fsl_solve_stats *stats);
/* Re-runs the recorded substeps with the recorded positions (no advection),
 * from the first checkpoint's velocities, with u0 (may be NULL) and body
 * force `force` (in place of fsl_set_params' force); u_post receives the last
 * substep's post-projection velocity. */
int fsl_rollout_replay(fsl_ctx *ctx, const fsl_solve_params *params, const float *u0,
                       const float force[3], float *u_post, fsl_solve_stats *stats);
/* Given dL/du_post of the last substep (box format), writes dL/du0 on the
 * faces active after the first P2G (box format, may be NULL) and dL/dforce
 * (may be NULL). Leaves the context holding the first substep's
 * recomputation. */
int fsl_rollout_backward(fsl_ctx *ctx, const fsl_solve_params *params, const float *loss_grad,
                         float *grad_u0, double grad_force[3], fsl_solve_stats *stats);
/* Checkpoint k: positions and velocities (float[n][3], may be NULL), the sort
 * permutation (new index -> index before advection), counts and dt. */
int fsl_rollout_checkpoint(fsl_ctx *ctx, int k, float *pos, float *vel, uint32_t *perm,
                           int64_t *n, int64_t *n_prev, float *dt);

#ifdef __cplusplus
}
#endif

#endif /* FSL_H_ */
