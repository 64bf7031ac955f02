/*
 * mgf.h — C ABI of the MI355X (gfx950) Eulerian MAC-grid step.
 *
 * The step is the one CGridFluidSolver::update() runs in
 * junwei-jiang/Hybrid-Fluid-Simulation (MPhysics/GridFluidSolver.cpp:336-397):
 *   boundary refresh -> semi-Lagrangian advection of the face-centred velocity
 *   and of the fluid-domain indicator -> body forces -> pressure projection
 *   (markers, 7-point matrix, right-hand side, (P)CG, pressure gradient)
 *   -> velocity extrapolation -> buffer flip.
 *
 * Conventions shared with the reference (kept identical on purpose):
 *   - Grids are cell-centred with resolution (nx, ny, nz), origin and spacing
 *     (MPhysics/Grid.h:5-24).  Storage is structure-of-arrays, x fastest:
 *     idx = z*nx*ny + y*nx + x (MPhysics/EulerMathTool.cu:63-73).
 *   - Face-centred vector fields store three arrays with extents
 *     (nx+1,ny,nz), (nx,ny+1,nz), (nx,ny,nz+1)
 *     (MPhysics/FaceCenteredVectorGrid.cpp:74-114).
 *   - All field values are IEEE float32; UNKNOWN is FLT_MAX
 *     (MPhysics/MMathExtra.h:48-58).
 *
 * Every function returns an mgf_status (0 = success) and never throws.
 * Host buffers passed in are caller-owned and copied; device state is owned by
 * the library.  A scene handle is not thread-safe.  Each scene runs on its
 * own HIP stream (different scenes may be driven from different threads); the
 * mgf_op_* operators run on the context's stream.
 */
#ifndef MGF_H
#define MGF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGF_API __attribute__((visibility("default")))

/* ---- status codes ------------------------------------------------------ */
typedef int32_t mgf_status;
#define MGF_OK                 0
#define MGF_ERR_INVALID        1  /* bad handle, size or parameter            */
#define MGF_ERR_DEVICE         2  /* HIP runtime error                         */
#define MGF_ERR_NOMEM          3  /* device allocation failed                  */
#define MGF_ERR_NOT_CONVERGED  4  /* CG hit its iteration cap (result kept)    */
#define MGF_ERR_NONFINITE      5  /* NaN/Inf found in an input or the solve    */
#define MGF_ERR_UNSUPPORTED    6  /* feature outside this build                */
#define MGF_ERR_BUDGET         7  /* server admission refused: HBM budget full */

/* ---- enumerations ------------------------------------------------------ */
/* Back-trace order; values match EAdvectionAccuracy RK1/RK2/RK3
 * (MPhysics/SemiLagrangian.cpp:188-347). */
#define MGF_RK1 1
#define MGF_RK2 2
#define MGF_RK3 3

/* CG stopping rule.
 * MGF_TOL_ABS_SQRT_NORM: stop when sqrt(||r||_2) < tolerance, the reference
 *   rule (MPhysics/CuMatrixFreePCG.cpp:51-52 with threshold 1e-3,
 *   MPhysics/CuMatrixFreePCG.h:28).
 * MGF_TOL_RELATIVE: stop when ||r||_2 <= tolerance * ||b||_2. */
#define MGF_TOL_ABS_SQRT_NORM 0
#define MGF_TOL_RELATIVE      1

/* Arithmetic of the pressure solve.
 * MGF_PRECISION_PLAIN: the reference's plain FP32 CG (FP32 vectors and
 *   scalars, one kernel per BLAS-1 operation, scalars read by the host).
 * MGF_PRECISION_MIXED: FP32 vectors inside fused kernels, FP64 dot products
 *   and alpha/beta kept on the device, FP64 pressure accumulated over outer
 *   iterative-refinement sweeps whose residual b - A x is formed in FP64. */
#define MGF_PRECISION_PLAIN 0
#define MGF_PRECISION_MIXED 1

/* Preconditioner.  IDENTITY is the reference's (MPhysics/PressureSolver.cpp:
 * 12-15) and the only one of PLAIN precision.  MIXED precision also offers
 * JACOBI (the inverse diagonal recomputed from the markers) and MULTIGRID: one
 * geometric V-cycle per iteration (2x2x2 cell coarsening down to <= 4 cells per
 * axis, damped Jacobi smoothing with mg_sweeps sweeps before and after the
 * coarse correction, restriction = average over the fluid children,
 * prolongation = injection), a symmetric positive definite operator, so the
 * iteration stays a preconditioned CG on the same system and tolerance. */
#define MGF_PRECOND_IDENTITY  0
#define MGF_PRECOND_JACOBI    1
#define MGF_PRECOND_MULTIGRID 2

/* Field identifiers for upload/download. */
#define MGF_FIELD_VEL_X        0  /* (nx+1)*ny*nz  face-centred velocity      */
#define MGF_FIELD_VEL_Y        1  /* nx*(ny+1)*nz                             */
#define MGF_FIELD_VEL_Z        2  /* nx*ny*(nz+1)                             */
#define MGF_FIELD_FLUID_DOMAIN 3  /* nx*ny*nz  fluid indicator (<0 = fluid)   */
#define MGF_FIELD_SOLID_SDF    4  /* nx*ny*nz  total solid marker -k/FLT_MAX
                                     (CBoundarys::getTotalBoundarysSDF); on
                                     upload it replaces the solids' result     */
#define MGF_FIELD_SOLID_VEL_X  5  /* face-centred solid velocity              */
#define MGF_FIELD_SOLID_VEL_Y  6
#define MGF_FIELD_SOLID_VEL_Z  7
#define MGF_FIELD_MARKERS      8  /* nx*ny*nz  2 solid / 1 fluid / 0 air (download) */
#define MGF_FIELD_PRESSURE     9  /* nx*ny*nz  CG solution x (download)       */
#define MGF_FIELD_RHS         10  /* nx*ny*nz  right-hand side b (download)   */
#define MGF_FIELD_MATRIX_A    11  /* 4*nx*ny*nz [diag,+x,+y,+z] (download only;
                                     built from the markers on request)        */
#define MGF_FIELD_DIVERGENCE  12  /* nx*ny*nz  divergence of the velocity
                                     (download only; computed on request)      */
#define MGF_FIELD_COLOR_X     13  /* nx*ny*nz  colour field, one component each
                                     (setColorField / getColorField,
                                     GridFluidSolver.cpp:208-213); advected with
                                     the velocity the step starts from, as
                                     CSemiLagrangian::advect(CCV)
                                     (SemiLagrangian.cpp:152-167).  Held from
                                     the first upload on (zeros before).       */
#define MGF_FIELD_COLOR_Y     14
#define MGF_FIELD_COLOR_Z     15
#define MGF_FIELD_FLUID_SDF   16  /* nx*ny*nz  generateCurFluidSDF
                                     (GridFluidSolver.cpp:571-588,
                                     EulerSolverTool.cu:324-406): layered
                                     distance from the fluid surface over
                                     extrapolation_layers passes (download
                                     only; computed on request)                */
#define MGF_FIELD_FLUID_DENSITY 17 /* nx*ny*nz generateCurFluidDensity
                                     (GridFluidSolver.cpp:590-613,
                                     EulerSolverTool.cu:408-448): 1 in the
                                     fluid, 0.2 at its air surface, 0 elsewhere
                                     (download only; computed on request)      */
#define MGF_FIELD_COUNT       18

/* ---- descriptors -------------------------------------------------------- */
typedef struct mgf_scene_desc {
    int32_t nx, ny, nz;        /* cell resolution (ghost layer included)     */
    float   origin[3];
    float   spacing[3];
    float   dt;                /* CGridFluidSolver m_DeltaT                   */
    float   density;           /* CPressureSolver m_density (1.0)             */
    float   gravity[3];        /* CExternalForcesSolver m_Gravity (0,-9.8,0)  */
    float   external_force[3]; /* CExternalForcesSolver m_ExternalForces      */
    int32_t rk_order;          /* MGF_RK1..3 (reference update() uses RK3)    */
    int32_t tol_mode;          /* MGF_TOL_*                                   */
    double  tolerance;         /* 1e-3 with ABS_SQRT_NORM reproduces the ref. */
    int64_t max_iterations;    /* <= 0: number of cells (PressureSolver.cpp:82) */
    int32_t precision;         /* MGF_PRECISION_*                             */
    int32_t preconditioner;    /* MGF_PRECOND_*                               */
    int32_t extrapolation_layers; /* 1000 in the reference (GridFluidSolver.cpp:176);
                                  above 1, the extrapolation needs
                                  nx + ny + nz + 3 < 65535 (16-bit face
                                  distances) and returns MGF_ERR_UNSUPPORTED
                                  on longer grids                            */
    int32_t max_refinements;   /* MIXED: outer sweeps cap (<= 0: 8)          */
    double  inner_tolerance;   /* MIXED: relative tolerance of one inner CG
                                  (<= 0: 1e-3)                                 */
    int32_t mg_sweeps;         /* MULTIGRID: Jacobi sweeps before and after the
                                  coarse correction (<= 0: 3; at most 8)       */
    float   mg_omega;          /* MULTIGRID: Jacobi damping (<= 0: 0.9; < 2)  */
    int32_t warm_start;        /* MIXED: 1 = start the CG from the previous
                                  step's pressure (fluid cells; 0 elsewhere)
                                  instead of x = 0 (PressureSolver.cpp:145).
                                  Same stop rule, different iterates; off by
                                  default                                    */
    int32_t mg_kernels;        /* MULTIGRID: MGF_MG_* (below); the same
                                  z = M^-1 r bit for bit either way          */
} mgf_scene_desc;

/* Kernels of the multigrid V-cycle (mgf_scene_desc.mg_kernels).  With three
 * sweeps, a 3-D level's descent (three sweeps and the restriction) and ascent
 * (prolongation and three sweeps) can each run as one fused kernel. */
#define MGF_MG_AUTO      0  /* fused on levels of at least 1M cells (default) */
#define MGF_MG_PER_SWEEP 1  /* one kernel per sweep on every level            */
#define MGF_MG_FUSED     2  /* fused on every 3-D level                       */

/* Per-call statistics (summed over the steps of one mgf_step call). */
typedef struct mgf_step_stats {
    int64_t cg_iterations;       /* inner CG iterations (A*p applications)    */
    int32_t refinement_sweeps;   /* MIXED outer sweeps                        */
    int32_t converged;           /* 1 when every solve met its tolerance     */
    double  residual_rel;        /* last solve: ||b - A x||_2 / ||b||_2, FP64 */
    double  residual_recurrence; /* last solve: final ||r||_2 of the recurrence */
    double  b_norm;              /* last solve: ||b||_2                        */
    int32_t extrapolation_layers;/* last step: layers that were non-empty     */
    int32_t steps;               /* steps completed                           */
    double  ms_advect;           /* wall time per stage, summed (host clock,  */
    double  ms_pressure;         /*   only when timing is enabled)            */
    double  ms_extrapolate;
    double  ms_total;
    int64_t fluid_cells;         /* last step: cells with marker 1            */
    double  div_residual;        /* with the divergence check on
                                    (mgf_scene_set_divergence_check): the
                                    largest over the call's solves of
                                    ||D u||_2 / ||b||_2 over the fluid cells,
                                    formed in FP64 from the stored FP32
                                    velocity right after the projection
                                    (D = the face divergence of
                                    FieldMathTool.cu:370-395; b = -D of the
                                    advected velocity with solid faces at the
                                    solid velocity, EulerSolverTool.cu:
                                    128-197); -1 when the check is off      */
    int32_t true_converged;      /* 1 when every solve's FP64 residual
                                    ||b - A x||_2 met the tolerance.  MIXED
                                    stops on it, so this equals converged;
                                    PLAIN stops on the FP32 recurrence
                                    residual as the reference does
                                    (CuMatrixFreePCG.cpp:51-52), and its true
                                    residual can stay above the tolerance    */
    int32_t reserved0;
} mgf_step_stats;

typedef struct mgf_context_s* mgf_context;
typedef struct mgf_scene_s*   mgf_scene;

/* ---- context ------------------------------------------------------------ */
/* Replaces CCudaContextManager::initCudaContext / freeCudaContext
 * (MPhysics/CudaContextManager.cpp:3-40). */
MGF_API mgf_status mgf_init(int32_t device, mgf_context* out);
MGF_API mgf_status mgf_shutdown(mgf_context ctx);
MGF_API const char* mgf_status_string(mgf_status s);
MGF_API const char* mgf_last_error(void);
/* Fills desc with the reference defaults for a grid: RK3, ABS_SQRT_NORM 1e-3,
 * max_iterations = cells, 1000 extrapolation layers, gravity (0,-9.8,0),
 * density 1, PLAIN precision, identity preconditioner. */
MGF_API mgf_status mgf_scene_desc_defaults(mgf_scene_desc* desc, int32_t nx,
        int32_t ny, int32_t nz, const float origin[3], const float spacing[3],
        float dt);

/* ---- scene (CGridFluidSolver) -------------------------------------------- */
/* CGridFluidSolver(dt, res, origin, spacing) + resizeGridFluidSolver
 * (MPhysics/GridFluidSolver.cpp:32-41,161-196).  Fields start at zero. */
MGF_API mgf_status mgf_scene_create(mgf_context ctx, const mgf_scene_desc* desc,
        mgf_scene* out);
MGF_API mgf_status mgf_scene_destroy(mgf_scene scene);
MGF_API mgf_status mgf_scene_get_desc(mgf_scene scene, mgf_scene_desc* out);
/* Changes the solver parameters that do not affect storage (dt, forces,
 * tolerance, precision, preconditioner, order, layers). */
MGF_API mgf_status mgf_scene_set_params(mgf_scene scene, const mgf_scene_desc* desc);
/* Device bytes held by the scene. */
MGF_API mgf_status mgf_scene_device_bytes(mgf_scene scene, int64_t* out);

/* CGridFluidSolver::addSolidBoundary / CBoundarys::addBoundary
 * (MPhysics/GridFluidSolver.cpp:203-206, MPhysics/Boundary.cpp:92-113):
 * a cell-centred SDF of nx*ny*nz floats (negative inside) and its velocity. */
MGF_API mgf_status mgf_scene_add_solid(mgf_scene scene, const float* sdf,
        int64_t count, const float velocity[3]);
/* CGridFluidSolver::generateFluidDomainFromBBox
 * (MPhysics/GridFluidSolver.cpp:551-555, MPhysics/EulerSolverTool.cu:581-608):
 * fluid indicator -500 at cell centres strictly inside (min,max), +500 else. */
MGF_API mgf_status mgf_scene_fluid_box(mgf_scene scene, const float box_min[3],
        const float box_max[3]);

/* setVelocityField / setFluidDomainField / setSolidVelField and the getters
 * (MPhysics/GridFluidSolver.cpp:214-229, 251-324). */
MGF_API mgf_status mgf_scene_upload(mgf_scene scene, int32_t field,
        const float* host, int64_t count);
MGF_API mgf_status mgf_scene_download(mgf_scene scene, int32_t field,
        float* host, int64_t count);
/* The same from / into device memory the caller owns (the fields of the
 * reference live on the device: VectorGrid.h:16-21); same side effects. */
MGF_API mgf_status mgf_scene_upload_device(mgf_scene scene, int32_t field,
        const float* dev, int64_t count);
MGF_API mgf_status mgf_scene_download_device(mgf_scene scene, int32_t field,
        float* dev, int64_t count);
MGF_API mgf_status mgf_scene_field_count(mgf_scene scene, int32_t field,
        int64_t* out);

/* Stages of one step, each backing one reference method. */
/* CBoundarys::updateBoundarys (MPhysics/Boundary.cpp:115-150). */
MGF_API mgf_status mgf_scene_update_boundaries(mgf_scene scene);
/* _computeAdvection (MPhysics/GridFluidSolver.cpp:399-416): velocity and fluid
 * domain advected through the current velocity with desc.rk_order. */
MGF_API mgf_status mgf_scene_advect(mgf_scene scene);
/* _computeExternalForces -> CExternalForcesSolver::applyExternalForces
 * (MPhysics/ExternalForcesSolver.cpp:50-59). */
MGF_API mgf_status mgf_scene_apply_forces(mgf_scene scene);
/* _computePressure -> CPressureSolver::solvePressure
 * (MPhysics/PressureSolver.cpp:85-107): markers, system, CG, gradient. */
MGF_API mgf_status mgf_scene_pressure_solve(mgf_scene scene, mgf_step_stats* stats);
/* _extrapolatingVel -> extrapolatingDataInvoker(FCV)
 * (MPhysics/GridFluidSolver.cpp:527-537, MPhysics/EulerSolverTool.cu:904-969). */
MGF_API mgf_status mgf_scene_extrapolate_velocity(mgf_scene scene,
        int32_t* layers_used);
/* CGridFluidSolver::update() run nsteps times
 * (MPhysics/GridFluidSolver.cpp:336-397). */
MGF_API mgf_status mgf_step(mgf_scene scene, int32_t nsteps, mgf_step_stats* stats);
/* Blocks until the scene's queued work is done. */
MGF_API mgf_status mgf_scene_sync(mgf_scene scene);
/* Host wall-clock timing of stages in mgf_step (adds synchronisation). */
MGF_API mgf_status mgf_scene_set_timing(mgf_scene scene, int32_t enabled);
/* The divergence check (no counterpart in the reference): when enabled, every
 * pressure solve of the scene also forms ||D u||_2 / ||b||_2 over the fluid
 * cells from the projected FP32 velocity, in FP64 (cells on the grid's last
 * layer along an axis are left out: their upper face is UNKNOWN after the
 * projection, EulerSolverTool.cu:318-321), and the stats report the largest
 * value of the call in div_residual.  One extra pass over the velocity per
 * solve, no host synchronisation.  Off by default. */
MGF_API mgf_status mgf_scene_set_divergence_check(mgf_scene scene, int32_t enabled);
/* Diagnostic: whether the mixed solver's CG and fused multigrid kernels skip
 * blocks that hold no fluid cell (1, the default) or run every block (0).
 * The results are the same bit for bit either way. */
MGF_API mgf_status mgf_scene_set_block_skip(mgf_scene scene, int32_t enabled);

/* Kernel probes (measurement; no counterpart in the reference).  While
 * enabled, every kernel the scene's entry points launch records HIP events at
 * the kernel's own start and end (hipExtLaunchKernelGGL on the scene's
 * stream).  mgf_scene_kernel_probes aggregates the launches since the last
 * reset per kernel (template instantiations under one name), in order of
 * first appearance: every launch, as a kernel trace counts them, and the
 * "busy" launches, those that ran at least a quarter of the kernel's longest
 * launch (the others are CG and multigrid launches that returned at once once
 * their solve had converged), with the algorithmic bytes their launch sites
 * gave (each value read or written once; 0 where a site gives none).
 * *count receives the number of kernels (at most cap are written). */
typedef struct mgf_kernel_probe {
    char    name[48];
    int64_t launches;
    double  ms;               /* summed duration of every launch            */
    int64_t busy_launches;
    double  busy_ms;
    double  busy_bytes;
} mgf_kernel_probe;
MGF_API mgf_status mgf_scene_set_kernel_probes(mgf_scene scene, int32_t enabled);
MGF_API mgf_status mgf_scene_kernel_probes(mgf_scene scene, mgf_kernel_probe* out, int32_t cap,
                                           int32_t* count, int32_t reset);

/* The MIXED solver's preconditioner applied once (diagnostic; the reference's
 * CG has none, CuMatrixFreePCG.h:27-28): z = M^-1 r for the markers of the
 * scene's current state (as mgf_scene_pressure_solve forms them), with the
 * scene's preconditioner settings.  r and z are device buffers of nx*ny*nz
 * floats (distinct); *rz (may be null) receives r . z in FP64 as the solver
 * forms it.  MGF_ERR_UNSUPPORTED unless precision is MIXED and the
 * preconditioner is not IDENTITY. */
MGF_API mgf_status mgf_scene_apply_preconditioner(mgf_scene scene, const float* r, float* z, double* rz);

/* ---- batched scenes ------------------------------------------------------ */
/* No counterpart in the reference, which steps one CGridFluidSolver at a time
 * (GridFluidSolver.cpp:336-397): callers that own many independent scenes.
 *
 * mgf_batch_step advances each of the n scenes by nsteps steps.  Every scene
 * runs on its own HIP stream, so the scenes' kernels overlap on the device;
 * `workers` host threads (<= 0: 8; at most 64) submit them, largest scene
 * first.  Each scene computes exactly what mgf_step(scene, nsteps) computes
 * alone (same kernels in the same order on its own buffers): results do not
 * depend on the other scenes.  stats (n entries) and status (n entries) may
 * be null; step_ms (n * nsteps entries, scene-major, may be null) receives the
 * host wall time of each step from its first launch to its completion.
 * Returns MGF_OK when every scene returned MGF_OK, else the first failing
 * scene's status.  The scenes must be distinct and belong to ctx; no other
 * call may use them until this one returns.
 *
 * Lockstep (on by default, mgf_batch_set_lockstep): when two or more of the
 * scenes use MGF_PRECISION_MIXED with MGF_PRECOND_MULTIGRID and no warm
 * start, they step together: every kernel of the step (advection, right-hand
 * side, the CG and the V-cycle, projection, extrapolation) is one launch over
 * all of them (per-scene arguments from a device table), with CG chunks and
 * sweeps in lockstep, per-scene convergence flags on the device and one
 * launch per extrapolation layer depth; only the boundary refresh runs on
 * the scenes' own streams.  A scene's results are bitwise the same either
 * way.  step_ms then holds each batch step's wall time for each of these
 * scenes. */
MGF_API mgf_status mgf_batch_step(mgf_context ctx, mgf_scene* scenes, int32_t n, int32_t nsteps,
                                  int32_t workers, mgf_step_stats* stats, mgf_status* status,
                                  double* step_ms);
/* Enables (default) or disables lockstep in mgf_batch_step for this context. */
MGF_API mgf_status mgf_batch_set_lockstep(mgf_context ctx, int32_t enabled);
/* Autotuner of the lockstep batch (no counterpart in the reference).  Times
 * the V-cycle kernel choices on these scenes, each over one lockstep batch
 * call of trial_steps (>= 2) steps, and keeps the faster on every scene that
 * runs in lockstep (their mg_kernels field): MGF_MG_AUTO (fused descent and
 * ascent on levels of at least 1M cells, one kernel per sweep below) and
 * MGF_MG_FUSED (fused on every 3-D level above the coarse cycle; small
 * levels of many scenes fill the device together).  Every scene advances by
 * 2 * trial_steps steps.  The choice changes only the summation order of
 * r.z (z = M^-1 r is bitwise the same), so the iterates differ within FP64
 * rounding; a scene then computes bitwise what it computes alone with the
 * same mg_kernels.  *chosen (may be null) receives the choice, ms (2
 * entries, may be null) each candidate's fastest step after its first. */
MGF_API mgf_status mgf_batch_autotune(mgf_context ctx, mgf_scene* scenes, int32_t n, int32_t trial_steps,
                                      int32_t* chosen, double* ms);

/* Device bytes a scene with this descriptor would hold once it has stepped
 * (solids excluded); allocates nothing. */
MGF_API mgf_status mgf_scene_bytes_for(mgf_context ctx, const mgf_scene_desc* desc, int64_t* out);

/* ---- served loop ----------------------------------------------------------
 * No counterpart in the reference: callers that stream step requests for
 * resident scenes (continuous batching).
 *
 * A server owns the scenes it admits and runs step requests on `slots` host
 * worker threads, each driving its scene's own stream, so requests that
 * arrive while others run start as soon as a slot is free and overlap with
 * them on the device.  When at least 8 scenes with a pending request could
 * step in lockstep (as in mgf_batch_step; mgf_batch_set_lockstep) and no
 * round is running, a free worker steps all of them together, one step per
 * round; a request of several steps stays at the head of its scene's queue
 * until it has them all.  Results are bitwise those of the scenes stepped
 * alone either way.  Requests for one scene run in submission order, one at
 * a time.  Admission is by HBM budget: mgf_server_open_scene refuses with
 * MGF_ERR_BUDGET when the resident bytes plus the new scene's
 * (mgf_scene_bytes_for) would exceed hbm_budget_bytes; solids added to an
 * admitted scene count from then on.  Between open and the first submit, and
 * while a scene has no request pending, the caller may set the scene up or
 * read it with the ordinary mgf_scene_* calls. */
typedef struct mgf_server_s* mgf_server;
typedef struct mgf_request_result {
    int64_t        ticket;
    mgf_scene      scene;
    mgf_status     status;
    int32_t        steps;        /* steps requested                           */
    double         t_submit_ms;  /* host steady clock (ms) at submit          */
    double         t_start_ms;   /* when a slot started it                    */
    double         t_end_ms;     /* when its last kernel had finished         */
    mgf_step_stats stats;
} mgf_request_result;

MGF_API mgf_status mgf_server_create(mgf_context ctx, int64_t hbm_budget_bytes, int32_t slots,
                                     mgf_server* out);
/* Waits for every pending request, then destroys the admitted scenes. */
MGF_API mgf_status mgf_server_destroy(mgf_server srv);
MGF_API mgf_status mgf_server_open_scene(mgf_server srv, const mgf_scene_desc* desc, mgf_scene* out);
/* Waits for the scene's pending requests, then destroys it (its results stay
 * pollable). */
MGF_API mgf_status mgf_server_close_scene(mgf_server srv, mgf_scene scene);
MGF_API mgf_status mgf_server_submit(mgf_server srv, mgf_scene scene, int32_t nsteps, int64_t* ticket);
/* Moves up to cap finished requests into out (oldest first) and sets *count;
 * waits up to wait_ms (< 0: until one finishes or nothing is pending) when none
 * has finished yet. */
MGF_API mgf_status mgf_server_poll(mgf_server srv, mgf_request_result* out, int32_t cap, int32_t wait_ms,
                                   int32_t* count);
/* Resident device bytes of the admitted scenes, requests waiting for a slot,
 * requests running; any pointer may be null. */
MGF_API mgf_status mgf_server_state(mgf_server srv, int64_t* resident_bytes, int32_t* queued,
                                    int32_t* running);

/* ---- operator entry points on caller device memory ---------------------- */
/* These run one reference operator on device buffers the caller owns (for
 * example buffers from mgf_device_alloc).  All pointers are device pointers.
 * The grid is given as resolution/origin/spacing. */
typedef struct mgf_grid {
    int32_t nx, ny, nz;
    float   origin[3];
    float   spacing[3];
} mgf_grid;

MGF_API mgf_status mgf_device_alloc(mgf_context ctx, int64_t bytes, void** out);
MGF_API mgf_status mgf_device_free(mgf_context ctx, void* ptr);
MGF_API mgf_status mgf_memcpy_d2d(mgf_context ctx, void* dst, const void* src, int64_t bytes);
MGF_API mgf_status mgf_memcpy_h2d(mgf_context ctx, void* dst, const void* src, int64_t bytes);
MGF_API mgf_status mgf_memcpy_d2h(mgf_context ctx, void* dst, const void* src, int64_t bytes);

/* CSemiLagrangian::backTrace (MPhysics/SemiLagrangian.cpp:188-347): moves
 * `count` points (SoA px,py,pz, in place) back through the face-centred
 * velocity (u,v,w on `grid`) by dt with the given RK order. */
MGF_API mgf_status mgf_op_backtrace(mgf_context ctx, const mgf_grid* grid,
        const float* u, const float* v, const float* w, float dt, int32_t rk_order,
        float* px, float* py, float* pz, int64_t count);
/* Trilinear sample of a cell-centred scalar field at `count` points
 * (sampleCellCenteredScalarFieldTrilerp, MPhysics/FieldMathTool.cu:52-104). */
MGF_API mgf_status mgf_op_sample_cc(mgf_context ctx, const mgf_grid* grid,
        const float* field, const float* px, const float* py, const float* pz,
        float* out, int64_t count);
/* CSemiLagrangian::advect for a face-centred field (SemiLagrangian.cpp:169-186):
 * positions are this grid's face centres. Input and output must not alias. */
MGF_API mgf_status mgf_op_advect_fc(mgf_context ctx, const mgf_grid* grid,
        const float* in_u, const float* in_v, const float* in_w,
        const float* vel_u, const float* vel_v, const float* vel_w,
        float dt, int32_t rk_order,
        float* out_u, float* out_v, float* out_w);
/* CSemiLagrangian::advect for a cell-centred scalar (SemiLagrangian.cpp:135-150). */
MGF_API mgf_status mgf_op_advect_cc(mgf_context ctx, const mgf_grid* grid,
        const float* in, const float* vel_u, const float* vel_v, const float* vel_w,
        float dt, int32_t rk_order, float* out);
/* The CG operator callback seam: PressureAxCGProd -> fdmMatrixVectorMul
 * (MPhysics/PressureSolver.cpp:6-10, MPhysics/EulerSolverTool.cu:199-227)
 * with the compact matrix recomputed from cell markers (uint8: 0 air,
 * 1 fluid, 2 solid) and scale = dt/(density*h^2) per axis. */
MGF_API mgf_status mgf_op_pressure_matvec(mgf_context ctx, const mgf_grid* grid,
        const float scale[3], const uint8_t* markers, const float* x, float* y);
/* CExternalForcesSolver::applyExternalForces (MPhysics/ExternalForcesSolver.cpp:
 * 50-59): u_c += gravity_c*dt, then u_c += force_c*dt, on every face; each
 * increment is skipped when all three components are below FLT_EPSILON
 * (FaceCenteredVectorGrid.cpp:162-175). */
MGF_API mgf_status mgf_op_apply_forces(mgf_context ctx, const mgf_grid* grid,
        float* u, float* v, float* w, const float gravity[3], const float force[3], float dt);
/* ---- particles (CEulerParticles, MPhysics/EulerParticles.cpp; kernels
 * MPhysics/EulerParticlesTool.cu) -----------------------------------------
 * The primitives of the PIC / FLIP / mixed advection solvers.  Positions and
 * per-particle vectors are interleaved x,y,z (3*i + k) in caller device
 * memory.  field_kind selects the grid a field lives on: 0 the cells of
 * `grid`, 1/2/3 its x/y/z face grid (extents +1 on that axis, origin -h/2). */
#define MGF_PG_P2GSUM     0   /* EPGTransferAlgorithm (MPhysics/EnumType.h:19-26) */
#define MGF_PG_G2PNEAREST 1
#define MGF_PG_LINEAR     2
#define MGF_PG_QUADRATIC  3
#define MGF_PG_CUBIC      4
/* tranferParticles2{CCS,CCV,FCV}FieldInvoker (EulerParticlesTool.cu:443-711):
 * field and weight (same extents) are zeroed, every particle adds
 * weight*values[span*i + offset] to the clamped corners of its kernel support
 * and the weight to `weight` (P2GSUM: the value to its own cell), then, with
 * normalize, field /= weight where weight > 0.  The sums are FP32 atomics, as
 * in the reference, so their rounding depends on the order they land in. */
MGF_API mgf_status mgf_op_particles_to_grid(mgf_context ctx, const mgf_grid* grid, int32_t field_kind,
        int64_t n, const float* pos, const float* values, int32_t span, int32_t offset, int32_t alg,
        float* field, float* weight, int32_t normalize);
/* tranfer{CCS,CCV,FCV}Field2ParticlesInvoker (EulerParticlesTool.cu:713-868):
 * values[span*i + offset] = the kernel-weighted sum of the field around
 * particle i (G2PNEAREST: the value of its cell). */
MGF_API mgf_status mgf_op_grid_to_particles(mgf_context ctx, const mgf_grid* grid, int32_t field_kind,
        int64_t n, const float* pos, const float* field, int32_t alg, float* values, int32_t span,
        int32_t offset);
/* CEulerParticles::advectParticlesInVelField (EulerParticles.cpp:231-325):
 * substeps of cfl*min(h)/max|v| ending at dt, RK1/2/3 through the face-centred
 * velocity (u,v,w on `grid`); with fix_to_grid the positions are then clamped
 * to the grid box (fixParticlesPosWithBoundarys, applied, as there, to the
 * first min(cells, n) particles).  A non-finite or zero substep returns
 * MGF_ERR_NONFINITE (the reference would not terminate). */
MGF_API mgf_status mgf_op_advect_particles(mgf_context ctx, const mgf_grid* grid, const float* u,
        const float* v, const float* w, int64_t n, float* pos, float dt, float cfl, int32_t order,
        int32_t fix_to_grid);
/* statisticalFluidDensityInvoker (solid == NULL: particles per cell) and
 * buildFluidMarkersInvoker (2 where solid < 0, else 1 where a particle lies,
 * else 0) (EulerParticlesTool.cu:366-441). */
MGF_API mgf_status mgf_op_particle_markers(mgf_context ctx, const mgf_grid* grid, const float* solid,
        int64_t n, const float* pos, float* out);
/* axpyReal / scaleReal (MPhysics/ThrustWapper.cu:281-304): y = a*x + y (one
 * fused multiply-add), x = a*x, on n device floats. */
MGF_API mgf_status mgf_op_axpy(mgf_context ctx, int64_t n, float a, const float* x, float* y);
MGF_API mgf_status mgf_op_scale(mgf_context ctx, int64_t n, float a, float* x);

/* ---- particle systems and particle advection solvers -------------------- */
/* A particle set (CEulerParticles, MPhysics/EulerParticles.h:5-64): positions,
 * velocities and colours (3 floats each per particle, interleaved), the
 * scalar value (1) and vector value (3) the particles carry, and the
 * advection scratch, all in device memory the handle owns.  Not thread-safe;
 * runs on its context's stream. */
typedef struct mgf_particles_s* mgf_particles;
#define MGF_PARTICLE_POS    0
#define MGF_PARTICLE_VEL    1
#define MGF_PARTICLE_COLOR  2
#define MGF_PARTICLE_SCALAR 3   /* n floats  */
#define MGF_PARTICLE_VECTOR 4   /* 3n floats */
/* CEulerParticles::resize (EulerParticles.cpp:20-67): n particles from host
 * arrays of 3n floats; a NULL array gives zeros. */
MGF_API mgf_status mgf_particles_create(mgf_context ctx, int64_t n, const float* pos, const float* vel,
        const float* color, mgf_particles* out);
/* CEulerParticles::generateParticlesInFluid (EulerParticles.cpp:163-201,
 * findFluidGrid EulerParticlesTool.cu:5-25): per_cell particles in every cell
 * with fluid_sdf < 0 and not solid_sdf < 0 (solid_sdf may be NULL), cells in
 * x-fastest order; particle i of a cell at (x + r) * h + o per axis.  The
 * reference draws r with rand() seeded by the clock; here r is a
 * counter-based hash of (seed, particle index, axis) in [0, 1), so a seed
 * reproduces the set exactly.  The fields are device arrays on `grid`. */
MGF_API mgf_status mgf_particles_generate(mgf_context ctx, const mgf_grid* grid, const float* fluid_sdf,
        const float* solid_sdf, int32_t per_cell, uint64_t seed, mgf_particles* out);
MGF_API mgf_status mgf_particles_destroy(mgf_particles p);
MGF_API mgf_status mgf_particles_count(mgf_particles p, int64_t* n);
/* getParticles*GPUPtr (EulerParticles.cpp:99-137): a borrowed device pointer,
 * valid until the handle is destroyed. */
MGF_API mgf_status mgf_particles_data(mgf_particles p, int32_t which, float** dev);
/* setParticlesPos / Vel / Color and the getters, through host memory. */
MGF_API mgf_status mgf_particles_upload(mgf_particles p, int32_t which, const float* host, int64_t count);
MGF_API mgf_status mgf_particles_download(mgf_particles p, int32_t which, float* host, int64_t count);
/* CEulerParticles::advectParticlesInVelField (EulerParticles.cpp:231-325),
 * as mgf_op_advect_particles but with the handle's own scratch. */
MGF_API mgf_status mgf_particles_advect(mgf_particles p, const mgf_grid* grid, const float* u, const float* v,
        const float* w, float dt, float cfl, int32_t order, int32_t fix_to_grid);

/* The particle CAdvectionSolver plug-ins: CParticleInCell
 * (MPhysics/ParticleInCell.cpp:58-106), CFluidImplicitParticle
 * (FluidImplicitParticle.cpp:60-125) and CMixPICAndFLIP
 * (MixPICAndFLIP.cpp:64-134), on a particle set they borrow (it must outlive
 * the solver).  The solver holds the grid-shaped state of the reference class:
 * the previous output velocity (FLIP, MIX), the weight fields, and the flags
 * that make the first scalar / vector call seed the particles' values. */
typedef struct mgf_psolver_s* mgf_psolver;
#define MGF_PSOLVER_PIC  0
#define MGF_PSOLVER_FLIP 1
#define MGF_PSOLVER_MIX  2
/* cfl: the substep CFL number (1.0 in the reference constructors); alg: the
 * particle<->grid kernel, MGF_PG_LINEAR .. MGF_PG_CUBIC; mixing: the PIC share
 * of MIX (0.1 in the reference), ignored by the others. */
MGF_API mgf_status mgf_psolver_create(mgf_context ctx, const mgf_grid* grid, int32_t kind, mgf_particles particles,
        float cfl, int32_t alg, float mixing, mgf_psolver* out);
MGF_API mgf_status mgf_psolver_destroy(mgf_psolver s);
/* advect(CFaceCenteredVectorField): the input velocity (in_*) is
 * transferred to the particles (PIC: replaced; FLIP: the grid change since
 * the last call is added; MIX: the blend (1-mixing) FLIP + mixing PIC), the
 * particles move through vel_* (mgf_particles_advect), and their velocity is
 * splatted into out_* (normalised by the weights).  fix_to_grid stands for a
 * boundary SDF of the grid's resolution being passed (the particles are then
 * clamped to the grid box).  All arrays are device face arrays on `grid`;
 * out_* may not alias in_* or vel_*. */
MGF_API mgf_status mgf_psolver_advect_velocity(mgf_psolver s, const float* in_u, const float* in_v,
        const float* in_w, const float* vel_u, const float* vel_v, const float* vel_w, float dt, int32_t order,
        int32_t fix_to_grid, float* out_u, float* out_v, float* out_w);
/* advect(CCellCenteredScalarField) / advect(CCellCenteredVectorField): the
 * first call samples `in` onto the particles; every call splats the
 * particles' values into `out` (the particles are not moved here). */
MGF_API mgf_status mgf_psolver_advect_scalar(mgf_psolver s, const float* in, float* out);
MGF_API mgf_status mgf_psolver_advect_vector(mgf_psolver s, const float* in_x, const float* in_y,
        const float* in_z, float* out_x, float* out_y, float* out_z);

/* CPressureSolver::solvePressure with particle positions
 * (MPhysics/PressureSolver.cpp:109-134): the markers come from the particles
 * (2 solid, 1 a cell holding a particle, 0 else; EulerParticlesTool.cu:48-77)
 * instead of the fluid domain; the rest of the projection is the same. */
MGF_API mgf_status mgf_scene_pressure_solve_particles(mgf_scene scene, mgf_particles particles,
        mgf_step_stats* stats);
/* CGridFluidSolver::generateFluid's particle branch (GridFluidSolver.cpp:
 * 93-158): mgf_particles_generate on the scene's fluid domain and solid marker
 * (the marker of solids added since the last step is built first, at their
 * current translations). */
MGF_API mgf_status mgf_scene_generate_particles(mgf_scene scene, int32_t per_cell, uint64_t seed,
        mgf_particles* out);
/* CGridFluidSolver::update() with a particle advection solver
 * (GridFluidSolver.cpp:336-370, 399-416, 434-525): boundaries, then the
 * velocity, the colour (when present) and the fluid domain through the
 * solver, body forces, the projection with particle markers, extrapolation;
 * nsteps times.  The solver's grid must be the scene's. */
MGF_API mgf_status mgf_step_particles(mgf_scene scene, mgf_psolver solver, int32_t nsteps, mgf_step_stats* stats);

/* extrapolatingDataInvoker for a face-centred field
 * (MPhysics/EulerSolverTool.cu:450-557, 904-969): in place. */
MGF_API mgf_status mgf_op_extrapolate_fc(mgf_context ctx, const mgf_grid* grid,
        float* u, float* v, float* w, int32_t layers, int32_t* layers_used);

/* mgf_op_extrapolate_fc without allocating: no counterpart in the reference.
 * For callers that cannot allocate device or pinned host memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid it).  Its
 * scratch (face distances, sorted list, masks, sort counters) lives in one
 * static device array of 35.1 MB per device, set up by mgf_init; the layer
 * offsets come back through a stack buffer.  Limits: at most 4194304 faces
 * over the three components ((nx+1)ny nz + nx(ny+1)nz + nx ny(nz+1)), and
 * with more than one layer nx + ny + nz + 3 <= 4093; larger grids return
 * MGF_ERR_INVALID before any launch.  Results are bitwise those of
 * mgf_op_extrapolate_fc.  Every call on one device shares that device's
 * array: calls that overlap on one device (several host threads, or several
 * contexts of the same device) race, so serialise them.  Returns after the
 * context's stream has finished the work. */
MGF_API mgf_status mgf_op_extrapolate_fc_noalloc(mgf_context ctx, const mgf_grid* grid,
        float* u, float* v, float* w, int32_t layers, int32_t* layers_used);

#ifdef __cplusplus
}
#endif

#endif /* MGF_H */
