/*
 * mgsb.h -- C ABI of libmgsb, the MI355X-native multigrid projective-dynamics
 * tetrahedral solver.  Drop-in for the solver path of
 * buuugyellow/MultiGridSoftBody: .msh load, fixed-vertex set, step() and
 * position readback in the Render DLL's layout.
 *
 * Conventions:
 *   - plain pointers and sizes only; caller keeps ownership of what it passes,
 *     arrays are copied;
 *   - every function returns int: 0 on success, a negative MGSB_ERR_* code on
 *     failure, with a message from mgsb_last_error() (per thread);
 *   - no function exits the process or throws across the ABI;
 *   - one solver handle is not thread-safe; distinct handles may be stepped
 *     from different threads (each owns its own HIP stream).
 *
 * Each declaration names the reference interface it replaces (file:line in
 * the reference repository).
 */
#ifndef MGSB_H
#define MGSB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGSB_OK 0
#define MGSB_ERR_ARG (-1)
#define MGSB_ERR_IO (-2)
#define MGSB_ERR_MESH (-3)
#define MGSB_ERR_HIP (-4)
#define MGSB_ERR_STATE (-5)
#define MGSB_ERR_NOMEM (-6)

/* Step schedules. */
#define MGSB_SCHED_SINGLE 0   /* PDSolver::Step, PDSolver.cpp:243-293                */
#define MGSB_SCHED_TWOLEVEL 1 /* PDSolver_MG::Step, PDSolver_MG.cpp:183-253          */
#define MGSB_SCHED_VCYCLE 2   /* residual-correction V-cycle (no reference schedule) */
#define MGSB_SCHED_LEARNED 3  /* V-cycle with a learned coarse warm start            */

/* Local-step variants. */
#define MGSB_IF_PLAIN 0 /* calculateIF, PDSolver.cu:128-183                         */
#define MGSB_IF_AC 1    /* getTetFRV2 + calTetIFV1, PDSolverAc.cu:447-662 (|f|<1e-5 dropped) */
#define MGSB_IF_STABLE 2 /* calculateIF with the Franca polar evaluated from deviatoric     \
                            invariants (k = 1.5 tr C'^2, l = 13.5 det C'): same maths, keeps \
                            small anisotropic strains the reference's FP32 k < 1e-6 branch   \
                            drops; the V-cycle's default                                     */

#define MGSB_IF_LINEAR 3 /* the V-cycle's linear operator on corrections, -F(e) AC k vol with no clamp \
                            (mgsb_tet_forces only; no reference counterpart)                     */

#define MGSB_MAX_SPHERES 8

/* Tet mesh (SoftObject's arrays, SoftObject.h:22-27).  Arrays returned by the
 * loaders/generators are owned by the library: release with mgsb_mesh_free. */
typedef struct mgsb_mesh {
    int32_t nverts;
    int32_t ntets;
    int32_t ntris;
    float* pos;        /* 3*nverts, xyz                                      */
    int32_t* tets;     /* 4*ntets, 0-based vertex ids, positive orientation  */
    uint32_t* tris;    /* 3*ntris surface triangles (TetFaceExtraction)      */
    uint32_t* tri_opp; /* ntris, tet vertex opposite each surface triangle   */
} mgsb_mesh;

/* Solver constants.  mgsb_default_params gives the reference's values
 * (PDSolver::Init, PDSolver.cpp:166-176). */
typedef struct mgsb_params {
    float dt;          /* 1/60                                                   */
    float damping;     /* 0.95                                                   */
    float stiffness;   /* ARAP weight per unit rest volume (8000)                */
    float density;     /* lumped mass = density * volume / 4 (1)                 */
    float rho;         /* Chebyshev spectral radius (0.9992); <=0: estimate      */
    float relax;       /* under-relaxation of the Jacobi step (0.6)              */
    float gravity[3];  /* (0,-9.8,0)                                             */
    float kcol;        /* sphere collision stiffness (1000)                      */
    int32_t iters;        /* PD iterations per step, fine/single level (20)      */
    int32_t iters_coarse; /* two-level schedule: coarse iterations (20)          */
    int32_t schedule;     /* MGSB_SCHED_*                                        */
    int32_t if_mode;      /* MGSB_IF_*                                           */
    int32_t tri_dcd_coarse; /* two-level: triangle DCD on the coarse level (1)   */
    /* V-cycle schedule */
    int32_t pre_sweeps;   /* nonlinear Jacobi sweeps before the coarse correction */
    int32_t post_sweeps;  /* ... after                                           */
    int32_t coarse_sweeps;/* linear Chebyshev-Jacobi sweeps per coarse level     */
    int32_t max_cycles;   /* V-cycles per step at most                           */
    float tol;            /* relative residual target per step (1e-3)           */
    int32_t check_every;  /* V-cycle convergence checks: 0 none (max_cycles run), */
                          /* 1 after every cycle, 2 also after pre-smoothing     */
    int32_t use_graph;    /* capture fixed-count steps in a hipGraph             */
    int32_t block_tet;    /* workgroup size of per-tet kernels (0: tuned)        */
    int32_t block_vert;   /* workgroup size of per-vertex kernels (0: tuned)     */
    int32_t device;       /* HIP device ordinal                                  */
    float rho_smooth;     /* V-cycle: Chebyshev rho of the fine smoother (0: damped Jacobi) */
    float relax_coarse;   /* V-cycle: damped-Jacobi weight on coarse levels (0.6) */
    float rho_coarse;     /* V-cycle: Chebyshev rho on coarse levels; <= 0: power iteration */
    int32_t coarse_direct;/* V-cycle: FP64 Cholesky solve on the coarsest level (1) */
    int32_t warm_start;   /* V-cycle: 0 start from s (the reference's x = s), 1 from the
                           * temporal extrapolation s + 2 u_{n-1} - u_{n-2} of the
                           * previous steps' corrections u = x - s (kept only if it
                           * lowers the residual).  The learned schedule uses the
                           * temporal stage for 0 and 1; 2 runs its MLP stage on s
                           * alone (no temporal stage, the network's own gain).   */
    int32_t learn_gate;   /* learned schedule (1, default; 0: the stage runs every step):
                           * skip the MLP stage after a step that needed at most one
                           * half-cycle (nothing left for it to save), and after 8
                           * consecutive steps whose learned correction was rejected, or
                           * kept without saving a half-cycle against the last step run
                           * without it, for 8, 16, ... 256 steps before trying again */
    int32_t vert_lanes;   /* lanes that gather one vertex's row of corner forces in the
                           * sweep and residual kernels: 1, 2, 4 or 8 (0: 1 with
                           * tile_gather, 4 without; tuned)                          */
    float learn_slack;    /* learned schedule: the MLP's coarse correction is kept when the
                           * residual it leaves is at most learn_slack times the best one so
                           * far (1, default: it must lower the residual; 0 reads as 1).  A
                           * coarse correction is normally followed by smoothing before it is
                           * judged: right after it, the high-frequency error its interpolation
                           * adds can raise the residual while the smooth error it removes is
                           * what the next half-cycle needs gone.  A non-finite residual is
                           * never kept. */
    int32_t coarse_dense; /* V-cycle: the coarse levels below the first (a fixed linear
                           * map of that level's residual) applied as one dense
                           * matrix-vector product, the matrix formed at set-up, when
                           * that level has at most 2048 vertices (1, default); 0: the
                           * sweeps, transfers and coarsest solve run in sequence.  The
                           * two differ by FP32 rounding only. */
    int32_t tile_gather;  /* per-tet kernels sum the corner forces of each tile of 256
                           * consecutive tets per vertex in LDS and write one partial
                           * per (tile, vertex), which the vertex kernels gather (1,
                           * default); 0: one force per (tet, corner) in the vertex-
                           * sorted CSR order.  Both are deterministic and differ by
                           * the grouping of the same sums in FP32. */
    int32_t coarse_first; /* V-cycle, checks after every half-cycle or cycle: a step's
                           * first half-cycle is the coarse correction of the start's
                           * residual with post-smoothing, then pre-smoothing halves and
                           * coarse-correction halves alternate (0, default: pre-smoothing
                           * first, the textbook V-cycle).  After a warm start the error
                           * left is mostly smooth, so the coarse correction pays first. */
    int32_t fuse_sweeps;  /* V-cycle fine-level sweeps with tile_gather: each sweep's vertex
                           * update runs inside the next local-step launch, every tile
                           * updating its own vertices; needs every vertex in a tile of at
                           * most 256 vertices and vert_lanes <= 1.  0: separate launches;
                           * 1 (default): fused on fine levels of at most 65,536 vertices,
                           * where launch latency dominates (each tile repeats the updates
                           * of its ~2.5x shared vertices, which costs more than a launch
                           * on larger meshes); 2: fused at any size.  The results are
                           * identical either way. */
    int32_t stall_guard;  /* V-cycle: when a check's residual exceeds 3 times the previous
                           * check's (or the start's), the fine smoother drops its Chebyshev
                           * extrapolation for the rest of the step (plain damped Jacobi;
                           * 1, default) or keeps it (0).  Near inverted tets the
                           * extrapolated iteration can jump instead of converging; the
                           * residual, its target and the fixed point are unchanged. */
} mgsb_params;

typedef struct mgsb_step_stats {
    double res0;          /* ||g(s)||, FP64, at the start of the last step        */
    double res;           /* ||g(x)|| at its end                                  */
    double rel;           /* res / res0                                           */
    int32_t cycles;       /* V-cycles run in the last step                        */
    int32_t fallback;     /* warm-start stages rejected in the last step: bit 0 the */
                          /* temporal extrapolation, bit 1 the learned correction  */
    int32_t steps;        /* steps taken by this call                             */
    int32_t halves;       /* half-cycles run in the last step (pre-smoothing half, */
                          /* coarse-correction + post-smoothing half)              */
} mgsb_step_stats;

/* Aggregates over the V-cycle steps since the last reset. */
typedef struct mgsb_run_stats_t {
    double max_rel;         /* largest end-of-step ||g|| / ||g(s)||            */
    int64_t total_cycles;   /* V-cycles summed over the steps                  */
    int32_t steps;
    int32_t failed_steps;   /* steps whose final relative residual > tol       */
    int32_t fallbacks;      /* steps with a rejected warm-start stage          */
    int32_t total_halves;   /* half-cycles summed over the steps               */
} mgsb_run_stats_t;

/* Energy diagnostics, runCalEnergy (PDSolver.cu:816-913): Ek, Ep, Ec, Nc, dX. */
typedef struct mgsb_energy_t {
    double Ek, Ep, Ec, Nc, dX;
} mgsb_energy_t;

/* Replaces SoftObject::ReadFromFile's $Nodes/$Elements reader
 * (SoftObject.cpp:85-130) plus SoftObject::TetFaceExtraction
 * (SoftObject.cpp:154-251): gmsh 2.2 ASCII; tets (type 4) kept, other element
 * types skipped by their node counts; surface triangles extracted from tets. */
int mgsb_load_msh(const char* path, mgsb_mesh* out);

/* SoftObject::TetFaceExtraction (SoftObject.cpp:154-251): (re)computes
 * tris/tri_opp from tets; same windings and the same (descending key) order. */
int mgsb_extract_surface(mgsb_mesh* m);

/* Structured box [lo,hi] of nx*ny*nz hex cells, 6 positively oriented tets per
 * hex (Kuhn split), surface extracted.  No reference counterpart (the
 * reference meshes come from gmsh); used by the box workloads. */
int mgsb_make_box(int32_t nx, int32_t ny, int32_t nz, const float lo[3], const float hi[3], mgsb_mesh* out);

/* Voxelised superquadric blob |x/a|^p+|y/b|^p+|z/c|^p <= (1+amp*sum_8 sin)^p,
 * h voxel size, 6 tets per voxel; voxels sit on the lattice h*Z^3, so blobs
 * at h, 3h, 9h nest.  No reference counterpart (organ-like workload). */
int mgsb_make_blob(float a, float b, float c, float p, float amp, uint32_t seed, float h, mgsb_mesh* out);
/* Level `coarsen` (1, 3, 9, ...) of a nested blob hierarchy on the same
 * lattice: cells of size coarsen * h, kept iff any of their cells of size h
 * is inside the shape, so every level covers the finer ones. */
int mgsb_make_blob_level(float a, float b, float c, float p, float amp, uint32_t seed, float h, int32_t coarsen,
                         mgsb_mesh* out);

void mgsb_mesh_free(mgsb_mesh* m);

/* InitVolumeConstraint (PDSolver.cpp:12-132), host only: per tet Dm^-1 (9,
 * row-major), AC (12), rest volume; per vertex lumped mass (density applied)
 * and Jacobi diagonal sum stiffness*vol*AC^2.  Any output may be NULL.
 * Fails with MGSB_ERR_MESH on a degenerate rest tet (|det Dm| < 1e-6), where
 * the reference calls exit(0) (utils.cu:244-248). */
int mgsb_precompute(const mgsb_mesh* m, float stiffness, float density, float* B, float* AC, float* vol,
                    float* mass, float* diag);

/* Barycentric binding of n points into the tets of mesh m, host only.
 * mode 0: PDSolver_MG::interpolate (PDSolver_MG.cpp:5-58), first tet in index
 *         order with all weights in [-1e-5, 1+1e-5];
 * mode 1: coarse-in-fine binding of PDSolver_MG::Init (PDSolver_MG.cpp:133-168),
 *         first tet passing pointInTet, face-area weights.
 * ids/w receive 4 per point; returns the number of unbound points (>= 0) or
 * a negative error. */
int mgsb_bind_points(const mgsb_mesh* m, const float* pts, int32_t n, int32_t mode, int32_t* ids, float* w);

/* The tiled local step's gather layout of mesh m, host only (no counterpart
 * in the reference, whose calculateIF scatters with atomicAdd,
 * PDSolver.cu:170-181): tiles of 256 consecutive tets; for each tile its
 * distinct vertices in ascending order, the position of each one's partial
 * in the vertex-sorted (tile, vertex) order, and the tile's corner array in
 * vertex-sorted (tet, corner) order.  sizes receives {ntiles, npart}
 * (npart = sum of the tiles' vertex counts); each array may be null:
 * prp V+1, tvo ntiles+1, tvs npart, tlr npart+ntiles, tls 2T (two 16-bit
 * corner positions per entry). */
int mgsb_mesh_tiles(const mgsb_mesh* m, int32_t sizes[2], int32_t* prp, int32_t* tvo, int32_t* tvs, int32_t* tlr,
                    uint32_t* tls);

/* PDSolver::Init constants (PDSolver.cpp:166-176) plus V-cycle defaults. */
void mgsb_default_params(mgsb_params* p);

typedef struct mgsb_solver mgsb_solver;

/* PDSolver::Init (PDSolver.cpp:164-200, PDSolver.h:102) for nlevels == 1 and
 * PDSolver_MG::Init (PDSolver_MG.cpp:60-181) for nlevels >= 2.  levels[0] is
 * the finest mesh; levels[i+1] is coarser than levels[i].  The solver
 * precomputes (InitVolumeConstraint, PDSolver.cpp:12-132), builds the
 * barycentric level maps, and uploads everything to the device. */
int mgsb_create(const mgsb_mesh* levels, int32_t nlevels, const mgsb_params* p, mgsb_solver** out);
void mgsb_destroy(mgsb_solver* s);

/* Parameters that may change between steps (iteration counts, schedule,
 * stiffness excluded: stiffness and density are baked at create). */
int mgsb_set_params(mgsb_solver* s, const mgsb_params* p);

/* PDSolver::SetFixedVert (PDSolver.cpp:134-162): fixflag per vertex of a
 * level, 1 = free, 0 = pinned. */
int mgsb_set_fixed(mgsb_solver* s, int32_t level, const float* fixflag);
/* SphereFixer rule (PDSolver.cpp:145-161): pin vertices strictly inside any of
 * n spheres (cx,cy,cz,r) on every level (|X - c| < r in FP32, rest positions),
 * adding to the pins already set. */
int mgsb_fix_spheres(mgsb_solver* s, const float* spheres, int32_t n);
/* Pin vertices with |pos[axis]-value| < tol on every level (the y=10 rule,
 * PDSolver.cpp:137-143). */
int mgsb_fix_plane(mgsb_solver* s, int32_t axis, float value, float tol);

/* The fixflags of a level (V floats, 1 free / 0 pinned): PDSolver's
 * m_tetVertFixed (PDSolver.h:95), host copy. */
int mgsb_read_fixed(mgsb_solver* s, int32_t level, float* fixflag);

/* SphereCollider (Collider.h:5-20): centre, radius, velocity (the centre
 * advances by vel*dt at the start of each step, as UpdateCollider moves it
 * before Step, Simulator.cpp:146-149), active flag. */
int mgsb_set_sphere(mgsb_solver* s, int32_t id, const float center[3], float radius, const float vel[3], int32_t active);
/* Ground plane n.x >= d as a penalty constraint of stiffness k (no reference
 * counterpart). active=0 disables it. */
int mgsb_set_plane(mgsb_solver* s, const float n[3], float d, float k, int32_t active);

/* Per-vertex velocity of the finest level (3*V). */
int mgsb_set_velocity(mgsb_solver* s, const float* v);
/* Positions of a level (3*V). */
int mgsb_set_positions(mgsb_solver* s, int32_t level, const float* x);

/* PDSolver::Step / PDSolver_MG::Step (PDSolver.cpp:243-293,
 * PDSolver_MG.cpp:183-253) run nsteps times. */
int mgsb_step(mgsb_solver* s, int32_t nsteps, mgsb_step_stats* stats);
/* Blocks until the solver's stream is idle. */
int mgsb_sync(mgsb_solver* s);

/* runCpyTetVertForRender (PDSolver.cu:261-268): positions of a level (3*V). */
int mgsb_read_positions(mgsb_solver* s, int32_t level, float* xyz);
int mgsb_read_velocity(mgsb_solver* s, int32_t level, float* v);
/* runUpdateOutsideTetVertNormal (PDSolver.cu:394-460): vertex normals of the
 * finest level (3*V), unit averages of adjacent surface triangle normals. */
int mgsb_read_normals(mgsb_solver* s, float* n);
/* The render repack of main.cpp:251-259: 9 floats per vertex
 * (pos, normal, 0,0,0) for Application::UpdateMesh (application.hpp:56). */
int mgsb_read_render9(mgsb_solver* s, float* v9);
/* Collision flags / depths of the finest level from the last iteration
 * (tetVertIsCollided_d / tetVertCollisionDepth_d, PDSolver.cu:264-265). */
int mgsb_read_collision(mgsb_solver* s, int32_t* flags, float* depth);
/* runCalEnergy (PDSolver.cu:816-913) on the finest level, FP64 sums on device. */
int mgsb_energy(mgsb_solver* s, mgsb_energy_t* e);
/* PDSolver::StepForConvergence (PDSolver.cpp:202-221) on the finest level:
 * from the current state, calculateST (the colliders where the next step
 * will see them), `iters` reference iterations with calculateIF's local step
 * (the reference uses m_iterNumCvg = 128, PDSolver.cpp:167), and
 * runCalEnergy after each; then positions and velocity are restored
 * (runSaveVel / runResetPosVel, PDSolver.cu:270-275), so the next mgsb_step
 * runs from the same state.  E receives iters + 1 totals Ek + Ep + Ec: E[0]
 * at s (after a DCD at s, as PDSolver::Step's energy log does), E[iters]
 * the converged energy (g_conEnergy_V2); last (may be NULL) the five terms
 * of the final iterate. */
int mgsb_step_convergence(mgsb_solver* s, int32_t iters, double* E, mgsb_energy_t* last);
/* The reference's convergence metric for one step (the error of its slides
 * and of PDSolver::Step's energy log, PDSolver.cpp:252-255): runs
 * mgsb_step_convergence(iters), then one mgsb_step of the current schedule,
 * and returns err = (E - E_conv) / (E0 - E_conv) with E = Ek + Ep + Ec at the
 * step's end (collision energy by a DCD at the final positions).  E (may be
 * NULL) receives E0, E_conv, E. */
int mgsb_step_error(mgsb_solver* s, int32_t iters, double* err, double* E);
/* runCalEnergy's per-vertex Ep density (calEveryVertEp, PDSolver.cu:893-924)
 * at the current positions of the finest level: per vertex, the sum over
 * incident tets of k vol/2 |F - R|^2 (negated where det F < 1) divided by the
 * sum of their current volumes (V floats; Simulator::m_tetVertEpDensity). */
int mgsb_read_ep_density(mgsb_solver* s, float* density);
/* mgsb_energy without any allocation: no hipMalloc, hipFree or other
 * allocating call on success (the reduction partials of mgsb_energy grow
 * by hipMalloc on demand).  No counterpart in the reference; it is for
 * callers that cannot allocate (serving loops that read the energy every
 * step, as PDSolver_MG::Step does, harnesses that forbid allocation).  The
 * partials live in a fixed 1 MiB device array per GPU, so the finest level
 * may have at most 8,387,584 vertices and tets (MGSB_ERR_ARG beyond).  Every
 * call on one GPU shares that array: calls that overlap on a GPU (handles
 * stepped from several threads) race and must be serialised by the caller.
 * Results are bit-identical to mgsb_energy. */
int mgsb_energy_noalloc(mgsb_solver* s, mgsb_energy_t* e);
/* ||g(x)|| of the finest level at its current state (FP64). */
int mgsb_residual(mgsb_solver* s, double* res);

/* Run aggregates of the V-cycle schedules (no reference counterpart: the
 * reference has no residual); reset != 0 clears them after reading. */
int mgsb_run_stats(mgsb_solver* s, mgsb_run_stats_t* out, int32_t reset);
/* The relative residuals ||g|| / ||g(s)|| at the last step's checks, in order
 * (the start, the warm-start judge, then one per checked half-cycle; at most
 * 64): count gets their number, rel the first n.  A convergence history for
 * the caller; no reference counterpart (the reference reports energies). */
int mgsb_step_history(mgsb_solver* s, double* rel, int32_t n, int32_t* count);

/* Learned coarse correction (MGSB_SCHED_LEARNED; no reference counterpart).
 * The step starts from the temporal extrapolation x_t = s + 2 u_{n-1} -
 * u_{n-2} (warm_start 0 or 1; with warm_start = 2, x_t = s).  A 4-layer MLP
 * (12 -> 128 -> 128 -> 128 -> 3, ReLU, BF16 MFMA with FP32 accumulation, FP32
 * master weights) maps features
 * of each level-1 vertex (restricted residual at x_t, injected velocity, rest
 * position, previous step's coarse delta) to the coarse correction that x_t
 * still misses; x = x_t + P pred warm-starts the V-cycles.  Each stage is kept
 * only if it lowers the residual (else x returns to the better of s and x_t),
 * and the V-cycles then run to the residual target as usual.
 *
 * record: V-cycle / learned steps store (features, achieved coarse delta
 * inject(u - u_t)) for up to max_steps steps (0 stops); train: Adam on the
 * recorded samples with standardised inputs/outputs, loss_hist receives the
 * mean loss per epoch. */
int mgsb_learn_record(mgsb_solver* s, int32_t max_steps);
int mgsb_learn_samples(mgsb_solver* s, int64_t* nsamples);
int mgsb_learn_clear(mgsb_solver* s);
int mgsb_learn_train(mgsb_solver* s, int32_t epochs, int32_t batch, float lr, uint32_t seed, float* loss_hist);
int mgsb_learn_nparams(mgsb_solver* s, int32_t* n);
/* Flat parameters W1[12][128] b1 W2[128][128] b2 W3[128][128] b3 W4[128][3]
 * b4; set != 0 writes them (and resets Adam state). */
int mgsb_learn_weights(mgsb_solver* s, float* w, int32_t n, int32_t set);
/* Standardisation: in_mu[12] in_sd[12] out_mu[3] out_sd[3]. */
int mgsb_learn_norm(mgsb_solver* s, float* norm, int32_t set);
/* Test entry point: the coarse correction (V1 x 3, after the level-1
 * Dirichlet mask) that the last learned step prolongated into x. */
int mgsb_learn_last(mgsb_solver* s, float* e1);
/* Test entry points: MLP outputs on raw inputs x (n x 12) through the fused
 * inference kernel (path 0) or the training GEMMs (path 1); gradient of the
 * mean squared error against targets t (n x 3). */
int mgsb_mlp_forward(mgsb_solver* s, const float* x, int32_t n, float* y, int32_t path);
int mgsb_mlp_grad(mgsb_solver* s, const float* x, const float* t, int32_t n, float* grad, double* loss);

/* Chebyshev rho used on a coarse level of the V-cycle (power iteration on
 * I - relax_coarse D^-1 A at create, or rho_coarse), and the number of free
 * vertices of the coarsest level's dense FP64 Cholesky factor (-1: none). */
int mgsb_level_rho(mgsb_solver* s, int32_t level, double* rho);
int mgsb_coarse_nfree(mgsb_solver* s, int32_t* nfree);

/* Autotuner (no reference counterpart; the reference hard-codes its launch
 * shapes).  Works on a snapshot of the solver state and restores it.
 * 1. Launch shapes by kernel time: 20 fine-level sweeps (local step +
 *    vertex update, as the schedule runs them) between HIP events for the
 *    current shape, the tiled layout with block_vert in {64,128,256} and
 *    vert_lanes in {1,2}, and the per-corner layout with block_tet and
 *    block_vert in {64,128,256} and vert_lanes in {2,4,8}; the fastest is
 *    kept if it beats the current shape by 2 %.
 * 2. On V-cycle schedules, pre/post sweeps in {3,4,5,6} by steps per second
 *    over trial_steps steps that start trial_steps steps after the saved
 *    state (best of two runs each); a count is kept only if every one of its
 *    steps met the residual target and it beat the current one by 3 %.
 * The chosen values are written into the solver's parameters. */
typedef struct mgsb_tune_result {
    int32_t block_tet, block_vert, pre_sweeps, post_sweeps;
    double ms_per_step;          /* the chosen parameters, over the window      */
    double ms_per_step_default;  /* the parameters before tuning, same window   */
    int32_t trials;
    int32_t vert_lanes;
    int32_t tile_gather;
    double sweep_us;             /* one fine sweep, chosen shape               */
    double sweep_us_default;     /* ... and the shape before tuning            */
    int32_t fuse_sweeps;         /* 1: the chosen shape fuses the fine sweeps  */
} mgsb_tune_result;
int mgsb_autotune(mgsb_solver* s, int32_t trial_steps, mgsb_tune_result* out);

/* Bodies: the connected components of level 0, numbered by their smallest
 * vertex (several soft bodies in one mesh, e.g. a disjoint union of boxes,
 * advance in the same launches).  Each body may have its own stiffness (ARAP
 * weight per unit volume, E); it applies to the body's tets on every level
 * (coarse vertices follow their coarse-in-fine binding) and rebuilds the
 * Jacobi diagonals and the V-cycle's coarse operators.  When every body is a
 * contiguous vertex range, the V-cycle checks each body's own relative
 * residual ||g_b(x)|| / ||g_b(s)|| and a step converges when the largest is at
 * most tol (per_body_checks = 1); mgsb_body_residuals gives res0_b and res_b
 * of the last check.  Batching has no reference counterpart (the reference
 * simulates one body, Simulator.cpp:76-87). */
int mgsb_body_count(mgsb_solver* s, int32_t* nbodies, int32_t* per_body_checks);
int mgsb_set_body_stiffness(mgsb_solver* s, const float* E, int32_t n);
int mgsb_body_residuals(mgsb_solver* s, double* res0, double* res, int32_t n);

/* Sizes of a level. */
int mgsb_level_info(mgsb_solver* s, int32_t level, int32_t* nverts, int32_t* ntets, int32_t* ntris);

/* Kernel timing: when on, HIP events bracket every launch of the dominant
 * kernel -- the per-tet local step on the finest level (kernel 0) -- on the
 * solver's stream (steps then run eagerly, not as a graph);
 * mgsb_kernel_time returns the summed milliseconds and the launch count since
 * timing was switched on. */
int mgsb_set_timing(mgsb_solver* s, int32_t on);
int mgsb_kernel_time(mgsb_solver* s, int32_t kernel, double* ms, int64_t* launches);
/* Throughput timing of kernel 0 as the current schedule launches it (split
 * positions for the V-cycle schedules), or of kernel 1, the fused fine-level
 * sweep (vertex update and local step in one launch; V-cycle schedules with
 * fuse_sweeps on a level that fuses): reps back-to-back launches on the
 * solver's stream between two HIP events; avg_us = elapsed / reps.  The
 * launches overwrite scratch only (corner forces or partials, the spare
 * position buffer), which the next step recomputes. */
int mgsb_time_kernel(mgsb_solver* s, int32_t kernel, int32_t reps, double* avg_us);

/* Test entry point: the per-tet local step of level `level` evaluated at
 * positions x (3*V) in mode if_mode (MGSB_IF_*), the forces each tet adds to
 * its four corners before the gather: cf[12*t + 3*j + i] is component i at
 * corner j of tet t -- calculateIF's temp[] times vol*k (PDSolver.cu:160-181)
 * or calTetIFV1's (PDSolverAc.cu:630-662).  Overwrites the level's corner
 * forces only (the next step recomputes them). */
int mgsb_tet_forces(mgsb_solver* s, int32_t level, int32_t if_mode, const float* x, float* cf);

const char* mgsb_last_error(void);
const char* mgsb_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MGSB_H */
