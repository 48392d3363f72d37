/* od.h — C-ABI of libodhip.so, the MI355X (gfx950) implicit-step hot path.
 *
 * Conventions (no reference counterpart: the reference has no ABI, no error
 * codes and no device memory):
 *   - every call returns 0 on success and a negative code on error; the text
 *     of the last error of the calling thread is od_last_error();
 *   - pointers named *_host are host memory borrowed for the call; pointers
 *     named *_dev are device memory (od_dev_alloc or any HIP allocation on the
 *     context's device);
 *   - each handle is bound to its context's HIP stream; calls are ordered on
 *     that stream and handles must not be used from several threads at once;
 *   - DOF vectors are over free DOFs in the reference's compacted numbering
 *     (core/nodeIndexer.cpp:12-25): a constrained DOF is removed and later
 *     DOFs shift down; they hold ndof entries (od_mesh_info).  Nodes may be
 *     partially constrained (Constrain_X .. Constrain_XZ, core/constrainer.h:
 *     16-22): the kernels then block the free DOFs in triples, the last one
 *     padded, and entry points that take device vectors run on padded copies
 *     (one device-to-device copy in, one out); matrices still upload, download
 *     and export in the reference's layout with its truncated last blocks and
 *     scalar remained columns (core/sparseMatrix.h:90-274).
 *
 * Each entry point cites the reference interface it replaces
 * (paths relative to OpenDeformer/ in the reference tree).
 */
#ifndef OD_H
#define OD_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct od_ctx od_ctx;
typedef struct od_mesh od_mesh;
typedef struct od_matrix od_matrix;
typedef struct od_solver od_solver;
typedef struct od_integrator od_integrator;

enum {
    OD_OK = 0,
    OD_ERR_ARG = -1,       /* invalid argument */
    OD_ERR_HIP = -2,       /* HIP runtime error */
    OD_ERR_STATE = -3,     /* call not valid in the handle's state */
    OD_ERR_UNSUPPORTED = -4
};

/* material kinds per tet */
enum { OD_MAT_COROTATIONAL = 0, OD_MAT_NEOHOOKEAN = 1, OD_MAT_COROT_PLASTIC = 2 };
/* preconditioners (core/preconditioner.h:12-22 family):
 *   OD_PRE_JACOBI  JacobiPreconditioner (solver/jacobiPreconditioner.cpp:6-26)
 *   OD_PRE_ICT     InCholeskyPreconditioner, the reference's threshold IC
 *                  (solver/inCholeskyPreconditioner.cpp:8-162): parameters epsilon and
 *                  relaxedScale (default 1); factored on the host at every solve, applied by
 *                  level-scheduled device sweeps; FP64 solvers only
 *   OD_PRE_IC0     IC(0) on the block pattern of A, factored and applied on the device
 *   OD_PRE_LDLT    InLDLTPreconditioner, SAINV incomplete LDL^T
 *                  (solver/inLDLTPreconditioner.cpp:8-184): parameters sainvEpsilon and
 *                  ldltEpsilon (default 0); host factor, device sweeps; FP64 solvers only */
enum { OD_PRE_NONE = 0, OD_PRE_JACOBI = 1, OD_PRE_ICT = 2, OD_PRE_IC0 = 3, OD_PRE_LDLT = 4 };
/* arithmetic of the inner PCG iterations */
enum { OD_FP64 = 0, OD_MIXED = 1 };

const char *od_last_error(void);
int od_version(void);

/* ---- context and device memory (no reference counterpart) ---------------- */
int od_ctx_create(int device, od_ctx **out);
int od_ctx_destroy(od_ctx *ctx);
int od_ctx_sync(od_ctx *ctx);
/* measurement mode (no reference counterpart): when enabled, every PCG SpMV
 * launch (q = A p) carries start/stop HIP events taken from its own dispatch;
 * od_ctx_profile_read returns their summed duration and count since enabling */
int od_ctx_profile(od_ctx *ctx, int enable);
int od_ctx_profile_read(od_ctx *ctx, double *totalMs_host, long long *launches_host);
int od_dev_alloc(od_ctx *ctx, long long bytes, void **dev);
int od_dev_free(od_ctx *ctx, void *dev);
int od_copy_h2d(od_ctx *ctx, void *dev, const void *host, long long bytes);
int od_copy_d2h(od_ctx *ctx, void *host, const void *dev, long long bytes);
int od_memset_d(od_ctx *ctx, void *dev, int byte, long long bytes);

/* ---- node relabelling (host only: no context, no device) -------------------
 * MeshRelabeler::getNewLables(elementCount, nodePerElementCount, newNodeLables,
 * elements) (core/meshRelabeler.h:113, core/meshRelabler.cpp:57-498): the
 * Gibbs-Poole-Stockmeyer labelling of the mesh graph.  newLabels_host[node] is
 * the node's new label; elements_host (nElems x nodesPerElem) is rewritten with
 * the new labels, sorted by the sum of each element's labels.  Orders the
 * reference leaves to its standard library are fixed: edges in first-insertion
 * order, stable sorts, equal degrees in push order.  A mesh whose graph is not
 * connected is rejected with OD_ERR_ARG. */
int od_relabel_gps(int nNodes, int nElems, int nodesPerElem, int *elements_host, int *newLabels_host);

/* ---- mesh, DOF indexing and the canonical pattern --------------------------
 * Replaces: Mesh/TetMesh set-up (core/mesh.cpp:7-40), NodeIndexer
 * (core/nodeIndexer.cpp:8-74), FullOrderNonlinearMaterial::getMatrixStructure
 * (material/fullOrderNonlinear.h:46-71) with BlockedSymSparseMatrix's
 * constructor (core/sparseMatrix.h:186-274), and getMatrixIndicesPerElement
 * (material/fullOrderNonlinear.h:73-110).
 * X_host: 3*nNodes rest positions; tets_host: 4*nTets node ids;
 * constrained_host: sorted DOF ids 3*node+axis (Constrainer::outputConstrains,
 * core/constrainer.cpp:74-79).  nInstances > 1 concatenates independent
 * meshes: instNodeStart/instTetStart (nInstances+1 entries) give each one's
 * node and tet range (NULL when nInstances == 1); a mesh with partially
 * constrained nodes has one instance. */
int od_mesh_create(od_ctx *ctx, int nNodes, const double *X_host, int nTets, const int *tets_host,
                   int nConstrained, const int *constrained_host, int nInstances,
                   const int *instNodeStart_host, const int *instTetStart_host, od_mesh **out);
int od_mesh_destroy(od_mesh *mesh);
/* info[16]: ndof, nBlockRows, nBlocks, nValues, nLevels, nInstances, nTets,
 * nNodes, nChunks, nPairContrib, nNodeContrib, device bytes, flags (bit 0: the
 * IC(0) levels are consecutive row ranges; bit 1: several instances whose sweeps
 * form the PCG's r.z partials), then the sweep form of the first 8 levels
 * (forward, backward; 4 bits per level: log2 of the lanes per row, or 8 +
 * log2(slots) for the FP32 slot kernel), 1 when some node is partially
 * constrained.  ndof and nValues are the reference's (vector entries, canonical
 * values); nBlockRows and nBlocks count the kernels' blocks. */
int od_mesh_info(od_mesh *mesh, long long *info_host);
/* Canonical layout, identical to BlockedSymSparseMatrix<3,3>'s blockPcol,
 * blockColumnOris, blockRows (core/sparseMatrix.h:627-635; ndof/3 + ndof%3 + 1
 * column entries, blockPcol's last entry of rows), and the 78-entry
 * per-tet map (SparseSymMatrixIndicesPerElementCache,
 * core/sparseMatrix.cpp:97-118), the latter expanded on the device from the
 * block-level map the assembly kernels use.  elemMap78_host may be NULL. */
int od_pattern_export(od_mesh *mesh, int *blockPcol_host, int *blockColumnOris_host, int *blockRows_host,
                      int *elemMap78_host);
/* Launch shapes (no reference counterpart).  The PCG SpMV runs R = 1, 2 or 4
 * chunks of 64 block rows per workgroup (more rows' loads in flight per thread)
 * with 16, 8 or 4 lanes per block row.  The FP32 product also has a slot-value
 * form (lanes 0): the solver keeps an FP32 copy of the matrix in the product's
 * own slot order (36 B per slot, 16 slots per block row), refreshed at every
 * mixed solve, so a wave's value loads read consecutive memory; R is then the
 * slots per thread (1, 2 or 4).  Every shape gives the same bits.  The first
 * solver created on a mesh runs the autotuner, which times each shape for FP64
 * and FP32 values with the matrix out of cache and keeps the fastest, and picks
 * the form of each IC(0) sweep level (od_mesh_sweep_form) the same way.
 * The grouped form (lanes -1; FP64 and FP32) takes up to four chunks that
 * share many blocks (on a 4-colour numbering whose chunks are tiles: one tile's
 * four colour chunks) per workgroup and reads each upper block stored for the
 * group once: a block between two of the group's rows serves both through LDS.
 * R is then the slots of a row per thread (4 or 8).  It needs rows of at most
 * 16 blocks; the solver keeps its own copy of the matrix in the group order,
 * refreshed at every solve that runs it.
 * od_mesh_autotune reruns it (ms_host: 25 median times, [FP64, FP32] x [16, 8,
 * 4 lanes] x [R = 1, 2, 4], then the FP32 slot-value form with 1, 2, 4 slots
 * per thread, then the grouped form [FP64, FP32] x [4, 8 slots per thread] (-1
 * where the mesh has none); chosen_host: R and lanes for FP64, then for FP32;
 * either may be NULL).  od_mesh_spmv_variant sets the SpMV shapes directly (no tuning of the
 * SpMV or the sweeps). */
int od_mesh_autotune(od_mesh *mesh, double *ms_host, int *chosen_host);
int od_mesh_spmv_variant(od_mesh *mesh, int r64, int lanes64, int r32, int lanes32);
/* IC(0) sweep form of every level (no reference counterpart): lanes per row (4, 8
 * or 16) of the lane-per-row kernels, which the FP64 sweeps always run; slots > 0
 * (2, 4, 8 or 16) runs the FP32 sweeps on the slot kernel instead, one 64-row
 * chunk per workgroup with that many of a row's blocks per thread, over the FP32
 * factor's slot-major copy.  Every form gives the same bits.  Like
 * od_mesh_spmv_variant it marks the mesh tuned (no tuning on the next solver). */
int od_mesh_sweep_form(od_mesh *mesh, int lanes, int slots);
/* Rebinds a mesh (and the solvers and integrators built on it) to another
 * context's stream on the same device, after the work pending on its current
 * stream has finished (no reference counterpart; used by the serving loop). */
int od_mesh_bind_ctx(od_mesh *mesh, od_ctx *ctx);
/* Per-instance row range and IC(0) level structure for tests:
 * instRowStart_host (nInstances+1), level_host (nBlockRows). */
int od_mesh_rows(od_mesh *mesh, int *instRowStart_host, int *level_host);

/* Material per tet.  Replaces the constructors of
 * CorotationalHyperelasticMaterial (material/corotationalHyperelastic.cpp:8-18)
 * and InvertibleNeoHookeanMaterial (material/invertibleNeoHookean.cpp:5-9),
 * extended to per-tet (kind, E, nu, rho).  corotThreshold is the compression
 * threshold selecting polar vs QR (element/tetelement.cpp:571-572),
 * neoThreshold the singular-value clamp (material/invertibleHyperelastic.cpp:239-240). */
int od_mesh_set_material(od_mesh *mesh, const int *kind_host, const double *E_host, const double *nu_host,
                         const double *rho_host, double corotThreshold, double neoThreshold);

/* Corotational plastic tets (OD_MAT_COROT_PLASTIC).  Replaces the constructor
 * parameters of CorotationalPlasticMaterial (material/corotationalPlastic.cpp:8-21):
 * per-tet yield criterion, creep ratio and maximum plastic strain; the plastic
 * strain state (9 per tet, corotationalPlastic.cpp:45-46) restarts at zero.  Every
 * stiffness and force evaluation (od_assemble, each integrator step) updates it,
 * as generateMatrixAndVirtualWorks does (corotationalPlastic.cpp:57-114, 124-183).
 * Plastic tets without parameters never yield. */
int od_mesh_set_plastic(od_mesh *mesh, const double *yield_host, const double *creep_host,
                        const double *maxStrain_host);
/* reads (get_host) and/or writes (set_host) the plastic strain state, 9 per tet */
int od_mesh_plastic_strain(od_mesh *mesh, double *get_host, const double *set_host);

/* Free-DOF body-force virtual works (Forcer::addBodyForce + getVirtualWorks,
 * core/forcer.cpp:8-33, 80-96): f_dev = sum over tets of V/4 * body. */
int od_body_force(od_mesh *mesh, const double *bodyPerTet_host /* 3*nTets */, double *f_dev);

/* ---- matrices in the canonical layout ---------------------------------------- */
int od_matrix_create(od_mesh *mesh, od_matrix **out);
int od_matrix_destroy(od_matrix *m);
/* The device values in the kernels' layout: the reference's canonical layout with
 * whole-node constraints; with partially constrained nodes the padded blocks
 * (od_matrix_upload / od_matrix_download convert). */
int od_matrix_values(od_matrix *m, double **values_dev);
int od_matrix_upload(od_matrix *m, const double *values_host);
int od_matrix_download(od_matrix *m, double *values_host);
/* BlockedSymSparseMatrix::setZeros (core/sparseMatrix.h:277-279) */
int od_matrix_zero(od_matrix *m);
/* values = a*values + b*other  (Scale + Add, core/sparseMatrix.h:281-295) */
int od_matrix_axpby(od_matrix *m, double a, double b, const od_matrix *other);
/* MechMaterial::generateMassMatrix (core/mechMaterial.cpp:69-98): M += mass */
int od_mass_matrix(od_mesh *mesh, od_matrix *M);
/* FullOrderNonlinearMaterial::generateMatrixAndVirtualWorks for the per-tet
 * materials (material/corotationalHyperelastic.cpp:43-102,
 * material/invertibleHyperelastic.cpp:34-114): accumulates K += K(u) and
 * fint += f(u) at free-DOF displacement u_dev. */
int od_assemble(od_mesh *mesh, const double *u_dev, od_matrix *K, double *fint_dev);
/* SpMDV (core/numerMethod.h:134-262): y = A x, or y += A x when accumulate. */
int od_spmv(od_matrix *A, const double *x_dev, double *y_dev, int accumulate);
/* FP32 form on an FP32 copy of A refreshed by od_matrix_sync_f32 */
int od_matrix_sync_f32(od_matrix *A);
int od_spmv_f32(od_matrix *A, const float *x_dev, float *y_dev, int accumulate);

/* Measurement helper (no reference counterpart): launches the SpMV `iters`
 * times between two HIP events on the context stream; *avg_ms = mean
 * duration per launch.  fp32 selects the FP32 copy (x, y float). */
int od_spmv_timed(od_matrix *A, const void *x_dev, void *y_dev, int iters, int fp32, double *avg_ms);

/* ---- PCG ---------------------------------------------------------------------
 * Replaces CGSolver::solveLinearSystem (solver/cgSolver.cpp:13-57) with its
 * preconditioner (JacobiPreconditioner, solver/jacobiPreconditioner.cpp:6-26;
 * IC(0) in the reference's Preconditioner role, solver/inCholeskyPreconditioner.cpp):
 * warm start from x_dev, preconditioner rebuilt from A on every solve, stop
 * when ||r||^2 <= tol^2 ||b||^2 or after maxIter iterations, per instance.
 * iters_host / relres_host (nInstances entries each) may be NULL. */
int od_solver_create(od_mesh *mesh, int precond, double precondParam, int precision, od_solver **out);
int od_solver_destroy(od_solver *s);
/* Mixed precision (OD_MIXED) only, no reference counterpart: each FP32 inner
 * solve stops at innerTol relative to its own right-hand side (default 1e-4)
 * or once its residual is below half the outer tolerance tol*|b|, whichever
 * comes first; at most maxOuter inner solves with FP64 residual corrections
 * are made (default 12).  The inner iterations of all sweeps together stop at
 * the solve's maxIter, as the plain solve does. */
int od_solver_set_refinement(od_solver *s, double innerTol, int maxOuter);
/* Graph mode (no reference counterpart): PCG iterations are captured once as a
 * HIP graph of 4 iterations and replayed (iterations after convergence exit
 * at once on the device; the iteration cap is checked on the device).
 * Ignored while od_ctx_profile is on. */
int od_solver_set_graph(od_solver *s, int enable);
/* Preconditioner parameters (the constructor arguments of InCholeskyPreconditioner(epsilon,
 * relaxedScale), inCholeskyPreconditioner.h:14, and InLDLTPreconditioner(sainvEpsilon,
 * ldltEpsilon), inLDLTPreconditioner.h:14-15; IC(0): the pivot epsilon, p2 unused).  The
 * precondParam of od_solver_create / od_newmark_create sets p1 with p2 at its default. */
int od_solver_set_precond_params(od_solver *s, double p1, double p2);
/* info[4] (no reference counterpart): graph captures so far (one per new set of buffers or
 * launch forms), device bytes held by the solver, refinement sweeps and the largest
 * per-instance iteration count of the last solve */
int od_solver_info(od_solver *s, long long *info_host);
int od_pcg_solve(od_solver *s, od_matrix *A, const double *b_dev, double *x_dev, double tol, int maxIter,
                 int *iters_host, double *relres_host);
/* od_pcg_solve without any allocation (no reference counterpart: for callers that cannot
 * allocate device memory, such as code under HIP graph capture or tight serving loops).  The
 * solve's temporaries (five DOF vectors, chunk partials, per-instance scalars and flags) live
 * in a fixed 64 MiB device array per GPU, held until the process ends and looked up when a
 * context is created on that GPU.  Limits: FP64, OD_PRE_NONE or OD_PRE_JACOBI, at most 64
 * instances, 5 x 8 x ndof bytes plus the partials within 64 MiB (about 1.6M DOF); larger
 * systems are rejected with OD_ERR_ARG, meshes with partially constrained nodes (which
 * would need padded copies) with OD_ERR_UNSUPPORTED.  Same results as od_pcg_solve with
 * the same preconditioner.  Calls that overlap on one GPU share the array and race. */
int od_pcg_solve_noalloc(od_matrix *A, int precond, const double *b_dev, double *x_dev, double tol, int maxIter,
                         int *iters_host, double *relres_host);
/* Preconditioner::Preprocess(mat) (core/preconditioner.h:18): builds the factor of A (IC(0),
 * threshold IC, SAINV LDL^T) or D^-1 (Jacobi); and solvePreconditionerSystem(width, rhs,
 * result) (:14): z = P^-1 r with the factor of the last Preprocess.  r and z are device
 * vectors of ndof; od_precond_solve synchronises. */
int od_precond_preprocess(od_solver *s, od_matrix *A);
int od_precond_solve(od_solver *s, const double *r_dev, double *z_dev);
/* Preconditioner::Preprocess + solvePreconditionerSystem once: z = P(A)^-1 r */
int od_precond_apply(od_solver *s, od_matrix *A, const double *r_dev, double *z_dev);
/* IC(0) factor of A in the canonical layout (U = L^T with U^T U ~ A). */
int od_ic0_factor(od_solver *s, od_matrix *A, double *L_host);

/* ---- integrators ---------------------------------------------------------------
 * Full-order linearised Newmark (the reference's Newmark update,
 * intergrator/linearNewmark.cpp:27-30, 57-79, at full order with the assembly
 * structure of intergrator/nonlinearImplicitBackwardEuler.h:82-101) and the
 * reference's NonlinearImplicitBackwardEuler (same file, :48-146), both behind
 * the Intergrator interface (core/intergrator.h:13-35). */
int od_newmark_create(od_mesh *mesh, double beta, double gamma, double dt, double alphaM, double alphaK,
                      int precond, double precondParam, double tol, int maxIter, int precision, int initAccel,
                      od_integrator **out);
int od_backward_euler_create(od_mesh *mesh, double dt, double alphaM, double alphaK, int precond,
                             double precondParam, double tol, int maxIter, od_integrator **out);
int od_integrator_destroy(od_integrator *it);
/* od_solver_set_refinement for the integrator's solver */
int od_integrator_set_refinement(od_integrator *it, double innerTol, int maxOuter);
/* Graph mode: 0 off, 1 the solver's PCG iteration graphs (od_solver_set_graph), 2 also
 * whole-step graphs.  In mode 2 a Newmark step after the
 * first (IC(0), Jacobi or no preconditioner; no impulse pending) is one captured graph: element
 * pass, gather, factor, a fixed number of PCG iterations (FP64) or refinement sweeps of fixed
 * inner length (mixed), each iteration exiting at once on converged instances, and the
 * corrector for the instances the graph finished; one synchronisation reads the statistics.
 * The plan is the previous step's count plus 4, rounded up to a multiple of 4; graphs are
 * cached by plan (8 at most).  Instances the plan does not finish as the eager step would are
 * stepped again eagerly from their untouched state, so results equal the eager path's bits.
 * Mode 2 serves callers that need a step as one launch; measured on 256 batched boxes it is
 * 3-6% slower than mode 1 (the plan's spare iterations each launch every kernel of an
 * iteration over the whole batch). */
int od_integrator_set_graph(od_integrator *it, int enable);
/* info[4] (no reference counterpart): step graphs captured, steps run as a graph, graph steps
 * that completed instances eagerly, graphs cached */
int od_integrator_graph_info(od_integrator *it, long long *info_host);
/* the integrator's solver (owned by the integrator; not to be destroyed by the caller) */
int od_integrator_solver(od_integrator *it, od_solver **s);
/* device bytes held by the integrator, its solver and its mesh */
int od_integrator_device_bytes(od_integrator *it, long long *bytes_host);
/* Intergrator::setExternalVirtualWork (core/intergrator.h:18) */
int od_integrator_set_external(od_integrator *it, const double *f_host);
/* one-step load added to the next step's external force only */
int od_integrator_set_impulse(od_integrator *it, const double *f_host);
/* per-instance mask (nInstances bytes): inactive instances do not advance */
int od_integrator_set_active(od_integrator *it, const unsigned char *active_host);
/* Intergrator::runOneTimeStep x nSteps (core/intergrator.h:25) */
int od_integrator_step(od_integrator *it, int nSteps);
int od_integrator_get_state(od_integrator *it, double *d_host, double *v_host, double *a_host);
int od_integrator_set_state(od_integrator *it, const double *d_host, const double *v_host, const double *a_host);
/* the integrator's effective matrix A = M + gamma h C + beta h^2 K of the last
 * step (owned by the integrator; not to be destroyed by the caller) */
int od_integrator_matrix(od_integrator *it, od_matrix **A);
/* device pointers of the state vectors (d, v, a) for device-side consumers */
int od_integrator_state_dev(od_integrator *it, double **d_dev, double **v_dev, double **a_dev);
/* last step: PCG iterations and relative residual per instance, outer
 * refinement sweeps (mixed precision), initial-acceleration iterations */
int od_integrator_stats(od_integrator *it, int *iters_host, double *relres_host, int *outer_host, int *initIters_host);
/* NonlinearImplicitBackwardEuler::updateMeshVerticesDisplacements
 * (nonlinearImplicitBackwardEuler.h:103-116): u += h v on free DOFs */
int od_backward_euler_update_mesh(od_integrator *it);
/* Simulator::getVertexPositions (core/simulator.cpp:24-33): float X + u per node */
int od_integrator_positions(od_integrator *it, float *pos_host);

/* ---- plane collision damping ----------------------------------------------------
 * MeshPlaneCollisionDampingForcer::setCollisionPair + addDampingMatrix per pair
 * (collision/penaltyDampingForcer.h:25-38, penaltyDampingForcer.cpp:8-61, with the
 * plane-tet intersection and barycentric coordinates of collision/intersection.cpp:
 * 12-143): adds into D, in pair order, the damping matrix of nPairs (plane, tet)
 * pairs.  planes_host: 4 doubles per pair, the plane's normal n and constant c
 * (StaticPlaneShape, collisionShapes.h:17-20; below the plane: n.x + c < 0);
 * tets_host: the pair's tet.  Positions are the rest positions plus disp_dev and
 * the velocities vel_dev (free-DOF vectors; NULL: zero), constrained nodes at
 * rest.  A pair adds coeff dir n̂^T scaled by the barycentric weights of the
 * submerged part's centroid, only when the velocity there moves into the plane;
 * the stored diagonal-block entry (i+j, i) receives t(i, j) as in the reference
 * (penaltyDampingForcer.cpp:40-45).  The sums are in pair order, bit-reproducible.
 * Meshes with partially constrained nodes are rejected (OD_ERR_UNSUPPORTED). */
int od_plane_damping(od_matrix *D, int nPairs, const double *planes_host, const int *tets_host, double coeff,
                     const double *disp_dev, const double *vel_dev);
/* NonlinearImplicitBackwardEuler::addExternalDampingForceMatrix with a
 * MeshPlaneCollisionDampingForcer that reads the integrator's own displacement and
 * velocity (nonlinearImplicitBackwardEuler.h:122-124); the damping matrix enters
 * every later step's tangent (:89) until od_integrator_clear_damping (:126-128).
 * Newmark integrators return OD_ERR_UNSUPPORTED, as the reference's base
 * Intergrator reports the call unimplemented (core/intergrator.h:19-24). */
int od_integrator_add_plane_damping(od_integrator *it, int nPairs, const double *planes_host, const int *tets_host,
                                    double coeff);
int od_integrator_clear_damping(od_integrator *it);
/* The integrator's damping matrix as a matrix handle, for a StaDampingForcer's
 * addDampingMatrix(matrixIndices, mat) (core/forcer.h:22-26), which the reference's
 * addExternalDampingForceMatrix hands its damping matrix (nonlinearImplicitBackwardEuler.h:
 * 122-124): allocated (zero) on the first call, owned by the integrator, in every later step's
 * tangent until od_integrator_clear_damping.  Newmark: OD_ERR_UNSUPPORTED. */
int od_integrator_damping_matrix(od_integrator *it, od_matrix **D);
/* the integrator's damping matrix values (canonical layout; zeros if none) */
int od_integrator_damping(od_integrator *it, double *values_host);

/* ---- serving (no reference counterpart: the reference steps one mesh in one
 * thread, core/simulator.cpp:24-38) ----------------------------------------------
 * A pool of registered objects, each a mesh with its per-tet material, body
 * force per tet and time step.  Objects become device-resident on first use
 * and are evicted least-recently-used (state kept on the host, restored
 * exactly) when a load needs room under hbmBudgetBytes.  nSlots (1..16)
 * worker slots, each a host thread with its own HIP stream, serve requests:
 * request k arrives at arrival_host[k] seconds after od_server_run starts and
 * advances object obj_host[k] by nSteps_host[k] Newmark steps (beta 1/4,
 * gamma 1/2, no damping, PCG to tol with the given preconditioner and
 * precision); impulse forces on the nodes impNodes_host[impStart_host[k] ..
 * impStart_host[k+1]) (3 doubles each in impForces_host) are added to its
 * first step (impStart_host may be NULL).  A request starts once it has
 * arrived, it is its object's oldest waiting request, a slot is free and its
 * object is resident or fits; the slots' work overlaps on the device.  Each
 * object's arithmetic is its own, so its results do not depend on the mix.
 * Outputs per request: start and finish seconds and PCG iterations over its
 * steps; stats_host[4]: peak resident bytes, evictions, loads, resident bytes
 * at the end. */
typedef struct od_server od_server;
int od_server_create(int device, long long hbmBudgetBytes, int nSlots, int precision, int precond, double tol,
                     od_server **out);
int od_server_destroy(od_server *srv);
int od_server_add_object(od_server *srv, int nNodes, const double *X_host, int nTets, const int *tets_host,
                         int nConstrained, const int *constrained_host, const int *kind_host, const double *E_host,
                         const double *nu_host, const double *rho_host, const double *bodyPerTet_host, double dt,
                         int *objId);
int od_server_run(od_server *srv, int nReq, const int *obj_host, const int *nSteps_host, const double *arrival_host,
                  const int *impStart_host, const int *impNodes_host, const double *impForces_host,
                  double *start_host, double *finish_host, int *iters_host, long long *stats_host);
/* free-DOF state of an object (resident or not) */
int od_server_object_state(od_server *srv, int objId, double *d_host, double *v_host, double *a_host);
/* info[5]: ndof, resident (0/1), device bytes (estimate before its first load), steps taken, tets */
int od_server_object_info(od_server *srv, int objId, long long *info_host);

#ifdef __cplusplus
}
#endif
#endif
