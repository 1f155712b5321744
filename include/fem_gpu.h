/*
 * fem_gpu.h -- C ABI of the MI355X (gfx950) Fem3D implicit-step library (libfem_gpu.so).
 *
 * The reference exposes this path as C++ virtual interfaces, not as a C ABI. Each entry
 * point below names the reference interface it replaces (paths relative to the reference
 * root). Conventions:
 *   - plain pointers and sizes only; int status codes (FEM_OK = 0), message via fem_last_error;
 *   - one context = one representation = one device + one HIP stream; not thread-safe per
 *     context (the reference updates one representation per thread-pool task,
 *     Physics/FreeMotion.cpp:52-63);
 *   - DOF layout is node-major interleaved, dof = 3 * node + axis (Math/OdeState.h:38-207);
 *   - host-pointer entry points copy synchronously; *_device entry points take device pointers
 *     and are ordered on the context stream.
 */
#ifndef FEM_GPU_H
#define FEM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fem_ctx fem_ctx;

/* Status codes. Negative values are errors (the reference throws
 * SurgSim::Framework::AssertionFailure, Framework/Assert.h:77-95). */
#define FEM_OK 0
#define FEM_INACTIVE 1          /* state became non-finite; representation deactivated
                                   (Physics/FemRepresentation.cpp:285-293) */
#define FEM_ERR_INVALID (-1)    /* invalid argument or precondition */
#define FEM_ERR_HIP (-2)        /* HIP runtime error */
#define FEM_ERR_ORDER (-3)      /* call out of order (e.g. step before upload) */
#define FEM_ERR_NOT_CONVERGED (-4) /* solver did not reach its tolerance (reported, not thrown,
                                      by Eigen's CG; here only when options ask for it) */

/* Element kinds: the factory names of Physics/Fem3DElementTetrahedron.cpp:47,
 * Physics/Fem3DElementCorotationalTetrahedron.cpp:29-30, Physics/Fem3DElementCube.cpp:32. */
enum {
    FEM_ELEMENT_TET_LINEAR = 0,       /* "SurgSim::Physics::Fem3DElementTetrahedron" */
    FEM_ELEMENT_TET_COROTATIONAL = 1, /* "SurgSim::Physics::Fem3DElementCorotationalTetrahedron" */
    FEM_ELEMENT_HEX = 2               /* "SurgSim::Physics::Fem3DElementCube" */
};

/* Integration schemes (Math/OdeSolver.h:36-49). Every scheme steps tetrahedral and hexahedral meshes and
 * batched instances (each instance with its own time step). */
enum {
    FEM_SCHEME_EULER_IMPLICIT = 0,                 /* INTEGRATIONSCHEME_EULER_IMPLICIT */
    FEM_SCHEME_LINEAR_EULER_IMPLICIT = 1,          /* INTEGRATIONSCHEME_LINEAR_EULER_IMPLICIT */
    FEM_SCHEME_STATIC = 2,                         /* INTEGRATIONSCHEME_STATIC (Math/OdeSolverStatic.cpp) */
    FEM_SCHEME_LINEAR_STATIC = 3,                  /* INTEGRATIONSCHEME_LINEAR_STATIC */
    FEM_SCHEME_EULER_EXPLICIT = 4,                 /* INTEGRATIONSCHEME_EULER_EXPLICIT */
    FEM_SCHEME_LINEAR_EULER_EXPLICIT = 5,          /* INTEGRATIONSCHEME_LINEAR_EULER_EXPLICIT */
    FEM_SCHEME_EULER_EXPLICIT_MODIFIED = 6,        /* INTEGRATIONSCHEME_EULER_EXPLICIT_MODIFIED */
    FEM_SCHEME_LINEAR_EULER_EXPLICIT_MODIFIED = 7, /* INTEGRATIONSCHEME_LINEAR_EULER_EXPLICIT_MODIFIED */
    FEM_SCHEME_RUNGE_KUTTA_4 = 8,                  /* INTEGRATIONSCHEME_RUNGE_KUTTA_4 */
    FEM_SCHEME_LINEAR_RUNGE_KUTTA_4 = 9            /* INTEGRATIONSCHEME_LINEAR_RUNGE_KUTTA_4 */
};

/* Linear solver of the step (Math/LinearSparseSolveAndInverse.h:41-52, same numbering). The C ABI's default is
 * CG; the reference's DeformableRepresentation defaults to LU (Physics/DeformableRepresentation.cpp:53), which
 * the C++ adapters keep. LU: banded factorization without pivoting of the assembled (symmetric) operator in FP64,
 * factored after each assembly (once for the linear schemes, whose dense compliance is then kept when n^2 doubles
 * fit in 1 GiB and applied on every step), whatever the precision option; tetrahedral meshes of one instance. */
enum {
    FEM_LINEAR_SOLVER_LU = 0,   /* LINEARSOLVER_LU (LinearSparseSolveAndInverseLU, .cpp:31-97) */
    FEM_LINEAR_SOLVER_CG = 1    /* LINEARSOLVER_CONJUGATEGRADIENT (LinearSparseSolveAndInverseCG, .cpp:99-129) */
};

/* Arithmetic of the element kernels, assembly and CG. */
enum {
    FEM_PRECISION_FP64 = 0,         /* everything in FP64 */
    FEM_PRECISION_FP32_REFINED = 1  /* FP32 operator copy, SpMV and CG inside FP64 iterative
                                       refinement of the implicit-step residual; element kernels,
                                       polar decomposition and assembly stay FP64 */
};

typedef struct fem_options {
    double dt;                   /* time step (s) */
    double gravity[3];           /* default (0, -9.81, 0), Physics/Representation.cpp:36 */
    int32_t gravity_enabled;     /* default 1, Physics/Representation.cpp:38 */
    int32_t scheme;              /* FEM_SCHEME_* */
    int32_t precision;           /* FEM_PRECISION_* */
    int32_t newton_max_iterations; /* default 1, Math/OdeSolverEulerImplicit.cpp:26 */
    double cg_tolerance;         /* relative residual tolerance (Eigen default: machine epsilon) */
    int64_t cg_max_iterations;   /* < 0: 2 * ndof (Eigen default) */
    double inner_tolerance;      /* FP32 mode: relative tolerance of each inner FP32 CG solve */
    int32_t max_refinements;     /* FP32 mode: maximum FP64 refinement sweeps per step (>= 1) */
    int32_t fail_on_nonconvergence; /* 0: silent like Eigen's CG (default) */
    double rayleigh_mass;        /* Rayleigh damping D = rayleigh_mass M + rayleigh_stiffness K, default 0
                                    (FemRepresentation::setRayleighDampingMass / setRayleighDampingStiffness,
                                    Physics/FemRepresentation.cpp:212-220, 510-518, 565-629) */
    double rayleigh_stiffness;
    int32_t mass_lumping;        /* 1: row-sum lumped mass in the implicit schemes' updates, default 0
                                    (FemRepresentation::setMassLumping, Physics/FemRepresentation.cpp:320-323,
                                    499-508); the other schemes' mass is never lumped (:412-421) */
    double newton_epsilon;       /* Newton-Raphson stops once ||dv - dv_prev||_inf < newton_epsilon, default 1e-5
                                    (OdeSolverEulerImplicit::setNewtonRaphsonEpsilonConvergence,
                                    Math/OdeSolverEulerImplicit.cpp:26, 43-48, 101-112) */
    int32_t linear_solver;       /* FEM_LINEAR_SOLVER_*, default FEM_LINEAR_SOLVER_CG */
} fem_options;

typedef struct fem_stats {
    int64_t steps;                 /* implicit steps taken */
    int64_t last_cg_iterations;    /* CG iterations of the last linear solve (sum of inner solves) */
    double last_cg_error;          /* sqrt(|r|^2 / |b|^2) at exit of the last solve */
    int64_t total_cg_iterations;
    int64_t last_refinements;      /* FP32 mode: refinement sweeps of the last step */
    int64_t det_r_invalid;         /* Physics/Fem3DElementCorotationalTetrahedron.cpp:161-166 */
    int64_t g_singular;            /* Physics/Fem3DElementCorotationalTetrahedron.cpp:211-216 */
    int64_t fp64_polar_fallbacks;  /* FP32 mode: elements whose rotation was redone in FP64 */
    int32_t active;                /* 0 once the state became non-finite */
    int32_t converged;             /* last solve reached its tolerance */
    int64_t last_newton_iterations; /* linear solves of the last step's Newton-Raphson loop
                                       (Math/OdeSolverEulerImplicit.cpp:84-115); 1 for the single-solve schemes */
} fem_stats;

/* Context lifetime. Replaces constructing a Fem3DRepresentation (Physics/Fem3DRepresentation.h:54-121). */
int fem_create(int device, fem_ctx** out);
void fem_destroy(fem_ctx* ctx);
const char* fem_last_error(const fem_ctx* ctx);
/* Run the context's work on an existing hipStream_t (NULL: the context's own stream). */
int fem_set_stream(fem_ctx* ctx, void* hip_stream);

/* Mesh + materials. Replaces Fem3DRepresentation::setFem + FemRepresentation::setFemElementType +
 * FemRepresentation::doInitialize (Physics/Fem3DRepresentation.cpp:77-126, 274-284;
 * Physics/FemRepresentation.cpp:85-168): element initialisation (with the reference's material
 * and degeneracy asserts, Physics/FemElement.cpp:33-43), lumped node masses and the sparsity
 * pattern. connectivity: num_elements x (4 | 8) node ids. young/poisson/density: per element. */
int fem_upload_mesh(fem_ctx* ctx, int element_kind, int64_t num_nodes, const double* x0,
                    int64_t num_elements, const int32_t* connectivity, const double* young,
                    const double* poisson, const double* density);

/* Fixed DOFs. Replaces OdeState::addBoundaryCondition(nodeId[, nodeDofId]) (Math/OdeState.cpp:115-142);
 * duplicates are ignored as in the reference. */
int fem_add_boundary_condition_node(fem_ctx* ctx, int64_t node);
int fem_add_boundary_condition_dof(fem_ctx* ctx, int64_t dof);

/* Options. fem_default_options fills the reference defaults. Replaces the IntegrationScheme /
 * LinearSolver selection of DeformableRepresentation::doInitialize (Physics/DeformableRepresentation.cpp:333-385)
 * and LinearSparseSolveAndInverseCG::setTolerance / setMaxIterations (Math/LinearSparseSolveAndInverse.cpp:99-117). */
void fem_default_options(fem_options* opt);
int fem_set_options(fem_ctx* ctx, const fem_options* opt);

/* External generalized force (ndof values), kept until replaced; NULL clears it.
 * Replaces DeformableRepresentation::addExternalGeneralizedForce (Physics/Fem3DRepresentation.cpp:133-198)
 * for constant nodal loads. */
int fem_set_external_force(fem_ctx* ctx, const double* force);

/* Localized external generalized force for the next step only: force (3) spread over the nodes of
 * `element` with weights coordinate[a] (its natural coordinates, num_nodes of the element), and optional
 * 3x3 row-major stiffness and damping (NULL: none) added on node blocks (a, b) with weight
 * coordinate[a] * coordinate[b]. Calls accumulate; the terms are cleared after the next step, as
 * DeformableRepresentation::afterUpdate clears them. Replaces Fem3DRepresentation::addExternalGeneralizedForce
 * (Physics/Fem3DRepresentation.cpp:133-198; Physics/DeformableRepresentation.cpp:236-253). Stiffness and
 * damping need a scheme that is not one of the linear (constant-operator) variants. */
int fem_add_external_generalized_force(fem_ctx* ctx, int64_t element, const double* coordinate, const double* force,
                                       const double* stiffness, const double* damping);

/* State transfer (x, v of ndof values each; NULL leaves one unchanged). Replaces
 * DeformableRepresentation::setInitialState / getCurrentState (Physics/DeformableRepresentation.cpp:86-119). */
int fem_set_state(fem_ctx* ctx, const double* x, const double* v);

/* out = C b for num_columns column-major columns of ndof values (host pointers), C the compliance of the
 * last step's system matrix: boundary rows of b zeroed, one linear solve per column with the context's
 * solver settings (Eigen CG semantics per column, as Eigen solves a matrix right-hand side column by column),
 * boundary rows of the solution zeroed. Replaces DeformableRepresentation::applyCompliance
 * (Physics/DeformableRepresentation.cpp:175-190) and, with b = H^T, the constraints' getComplianceMatrix() * H^T
 * (Physics/FemConstraintFixedPoint.cpp:107, FemConstraintFrictionlessContact.cpp:108, ...) without the dense
 * inverse of OdeSolver::computeComplianceMatrixFromSystemMatrix (Math/OdeSolver.cpp:69-76). FEM_ERR_ORDER before
 * the first step. The step statistics are left as the last step set them. */
int fem_apply_compliance(fem_ctx* ctx, const double* b, int64_t num_columns, double* out);
/* The same with device pointers (column-major, ndof x num_columns), ordered on the context stream. On a
 * tetrahedral mesh the columns are solved up to 8 at a time by a multi-column FP64 CG that reads the
 * operator once per iteration for all of them (each column keeps its own Eigen CG recurrence and test). */
int fem_apply_compliance_device(fem_ctx* ctx, const double* b, int64_t num_columns, double* out);
/* The same as fem_apply_compliance_device, allocating nothing: no hipMalloc, hipFree, hipHostMalloc or other
 * allocating routine. No counterpart in the reference; it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation). The multi-column CG works in a fixed
 * device workspace of FEM_NOALLOC_SCRATCH_BYTES per GPU, held from fem_create until the process ends, which
 * bounds the mesh to about 256,000 DOFs (larger meshes: FEM_ERR_INVALID with the limit in the message). Every
 * column goes through the multi-column FP64 CG (in FP64 mode the results equal fem_apply_compliance_device's
 * bit for bit; in the FP32 + refinement mode they agree to the CG tolerance). Tetrahedral meshes of one
 * instance only; FEM_ERR_ORDER before the first step and after boundary conditions were added since the last
 * step. All contexts on one GPU share its workspace and each call leaves its contents behind: calls that
 * overlap on one GPU (other streams or threads) race. Blocks the calling thread until the columns are done. */
#define FEM_NOALLOC_SCRATCH_BYTES (96ull << 20)
int fem_apply_compliance_device_noalloc(fem_ctx* ctx, const double* b, int64_t num_columns, double* out);
/* The dense compliance matrix of the last step's system matrix: C = A^-1 with the boundary rows and columns
 * zero, their diagonal entries included (ndof x ndof, column-major). Replaces OdeSolver::getComplianceMatrix after
 * OdeSolver::computeComplianceMatrixFromSystemMatrix (Math/OdeSolver.cpp:69-76), which the linear schemes run on
 * their first step (Math/OdeSolverLinearEulerImplicit.cpp:49). The linear implicit scheme with the direct solver
 * keeps it from its first step (when ndof^2 doubles fit in 1 GiB) and this copies the kept matrix; otherwise it is
 * computed on demand from the band factors of the operator (factored here when the step used CG, whose reference
 * getInverse is the dense inverse, Math/LinearSparseSolveAndInverse.cpp:25-29). Tetrahedral meshes of one
 * instance; FEM_ERR_ORDER before the first step; FEM_ERR_INVALID when the band storage does not fit the device. */
int fem_get_compliance_matrix(fem_ctx* ctx, double* out);         /* host output */
int fem_get_compliance_matrix_device(fem_ctx* ctx, double* out);  /* device output, ordered on the context stream */
/* The compliance terms of a constraint localized in one element, as the constraint builders form them for
 * the MLCP (Physics/FemConstraintFixedPoint.cpp:84-106, Physics/FemConstraintFrictionlessContact.cpp:84-112,
 * consumed by MlcpPhysicsProblem::updateConstraint, Physics/MlcpPhysicsProblem.cpp:46-76). H has
 * num_directions (1..8) rows; row j holds coordinate[a] * directions[3 j + k] * scale at DOF 3 node_a + k for
 * every node a of the element with a non-zero coordinate (coordinate: the element's natural coordinates, one
 * per node; scale: the builder's dt * scale; fixed point: the three unit axes, frictionless contact: the
 * normal). cht (host, ndof x num_directions column-major, or NULL) = C H^T with C applied as
 * fem_apply_compliance does; hcht (host, num_directions^2 row-major, or NULL) = H C H^T. The H^T columns are
 * built and solved on the device (multi-column CG on tetrahedral meshes). FEM_ERR_ORDER before the first step. */
int fem_constraint_compliance(fem_ctx* ctx, int64_t element, const double* coordinate, int num_directions,
                              const double* directions, double scale, double* cht, double* hcht);
int fem_get_state(fem_ctx* ctx, double* x, double* v);
int fem_set_state_device(fem_ctx* ctx, const double* x, const double* v);
int fem_get_state_device(fem_ctx* ctx, double* x, double* v);

/* Advance nsteps implicit steps. Replaces FemRepresentation::update(dt) ->
 * OdeSolver::solve(dt, current, &new, computeCompliance = false) (Physics/FemRepresentation.cpp:232-294,
 * Math/OdeSolverEulerImplicit.cpp:55-161, Math/OdeSolverLinearEulerImplicit.cpp:43-64).
 * Returns FEM_INACTIVE once the state becomes non-finite (further calls do nothing). */
int fem_step(fem_ctx* ctx, int32_t nsteps);

/* fem_step that also records the states: x_hist and v_hist (device, (nsteps + 1) x ndof doubles each, ordered
 * on the context stream) receive the state at the call's start (row 0) and after each step (row i). No
 * counterpart in the reference; it serves the adjoint's recomputation of a checkpoint segment, whose states
 * fem_adjoint_steps reads back. */
int fem_step_record(fem_ctx* ctx, int32_t nsteps, double* x_hist, double* v_hist);
int fem_synchronize(fem_ctx* ctx);
int fem_query(fem_ctx* ctx, fem_stats* stats);

/* Batched independent instances. The uploaded tetrahedral mesh is declared to consist of num_instances
 * disconnected meshes (no element joins two of them): instance i owns nodes
 * [node_offsets[i], node_offsets[i + 1]). Each instance is one representation of the reference advanced
 * by its own step of the configured scheme (implicit Euler with one Newton iteration, or any other scheme)
 * with its own time step dt[i] (NULL: the options' dt for all) and its own Eigen CG (own tolerance test,
 * iteration count and cap), as if FemRepresentation::update ran once per instance
 * (Physics/FreeMotion.cpp:52-63 runs one task per representation). All instances advance in the same launches: per-instance offset tables, per-instance
 * CG scalars with convergence masks, iteration chunks replayed as a captured hipGraph, and row-chunk and
 * lane shapes picked by an autotuner on the first step. Call after fem_upload_mesh. */
int fem_set_instances(fem_ctx* ctx, int32_t num_instances, const int64_t* node_offsets, const double* dt);
/* Per-instance statistics. An instance whose state became non-finite is deactivated and no longer stepped,
 * as Physics/FemRepresentation.cpp:285-293 deactivates one representation; fem_step returns FEM_INACTIVE
 * only once every instance is. */
typedef struct fem_instance_stats {
    int64_t last_iterations;     /* CG iterations of the last step (FP32 mode: inner, summed over sweeps) */
    double last_error;           /* relative residual at the end of the last step's solve */
    int64_t unconverged_steps;   /* steps whose solve stopped at its iteration cap */
    int32_t converged;           /* last step's solve met its tolerance */
    int32_t active;
    int32_t solved_in_fp64;      /* FP32 + refinement mode: 1 once FP32 CG proved inadequate for the instance (its
                                    refinement stopped short of the tolerance, or an inner solve took more than a
                                    quarter of its iteration cap); its later solves run in FP64 */
    int32_t reserved;
} fem_instance_stats;
int fem_query_instances(fem_ctx* ctx, fem_instance_stats* stats);   /* num_instances entries */
/* Fix the chunk shape (chunk_rows > 0 block rows per workgroup, lanes_per_row 4, 8 or 16) or return to
 * autotuning (0, 0); use_graphs 0 launches the CG iterations one by one instead of replaying graphs.
 * No counterpart in the reference. */
int fem_set_batch_tuning(fem_ctx* ctx, int32_t chunk_rows, int32_t lanes_per_row, int32_t use_graphs);
int fem_get_batch_tuning(fem_ctx* ctx, int32_t* chunk_rows, int32_t* lanes_per_row, int32_t* num_chunks,
                         double* ms_per_iteration);

/* Replace the per-element Young's moduli (num_elements values, host pointer), keeping Poisson ratios
 * and densities. Replaces FemElement::setYoungModulus on every element of an initialized
 * representation (Physics/FemElement.h:91-99); operators are rebuilt on the next step. */
int fem_set_young_modulus(fem_ctx* ctx, const double* young);

/* Reverse of one implicit-Euler step (one Newton iteration; the linear scheme for linear tets) taken
 * from the state (x_n, v_n) to velocity v_next, for a loss L of the later states. No counterpart in the
 * reference (it has no gradients); it differentiates Math/OdeSolverEulerImplicit.cpp:55-161 with the
 * elements of Physics/Fem3DElementTetrahedron.cpp and Physics/Fem3DElementCorotationalTetrahedron.cpp.
 * Inputs lam_x, lam_v = dL/dx_{n+1}, dL/dv_{n+1}; outputs lam_x_out, lam_v_out = dL/dx_n, dL/dv_n
 * (may alias the inputs); grad_young (num_elements) is accumulated with dL/dE through this step.
 * All pointers are device pointers of ndof (or num_elements) doubles, ordered on the context stream.
 * The transpose solve reuses the step's operator (symmetric) and the context's CG settings. */
int fem_adjoint_step(fem_ctx* ctx, const double* x_n, const double* v_n, const double* v_next,
                     const double* lam_x, const double* lam_v, double* lam_x_out, double* lam_v_out,
                     double* grad_young);

/* nsteps fem_adjoint_step calls backwards through a recorded history (x_hist, v_hist: rows 0 .. nsteps as
 * fem_step_record writes them): step j = nsteps - 1 .. 0 reverses the step from row j to row j + 1, with
 * lam_x, lam_v updated in place and grad_young accumulated. No counterpart in the reference (see
 * fem_adjoint_step). In FP64 CG mode every transpose solve of the call after the first starts from the
 * solution of the step after it (Eigen's solveWithGuess: same tolerance, relative to the right-hand side);
 * the result depends only on the arguments. */
int fem_adjoint_steps(fem_ctx* ctx, int32_t nsteps, const double* x_hist, const double* v_hist, double* lam_x,
                      double* lam_v, double* grad_young);

/* Lumped node masses (num_nodes values): each element adds rho V / (4 | 8) to its nodes
 * (Physics/FemRepresentation.cpp:102-117); their sum is FemRepresentation::getTotalMass
 * (Physics/FemRepresentation.cpp:192-200). */
int fem_get_node_masses(fem_ctx* ctx, double* mass);

/* Sparsity pattern as scalar CSR (int32). The pattern is structurally symmetric, so these arrays
 * equal the reference's column-major (outerIndex, innerIndex) of Eigen::SparseMatrix<double>
 * built at Physics/FemRepresentation.cpp:121-132. */
int64_t fem_num_nonzeros(fem_ctx* ctx);
/* Hexahedral meshes: the number of distinct FP32 element matrices the inner-solve operator reads from its
 * table (a structured mesh repeats a few), or 0 when it reads one cached matrix per element (more than 32
 * distinct, FEM_HEX_TABLE=0 in the environment, no FP32 operator built yet, or not a hexahedral mesh). No
 * counterpart in the reference, which keeps a matrix per element (Physics/Fem3DElementCube.cpp:143-167). */
int fem_hex_distinct_matrices(fem_ctx* ctx);
int fem_get_pattern(fem_ctx* ctx, int32_t* rowptr, int32_t* colidx);

/* Last assembled system matrix as the CG operator sees it: the lower-triangle self-adjoint view of
 * A = M/dt + D + dt K with boundary conditions (Math/OdeSolverEulerImplicit.cpp:148-152,
 * Math/LinearSparseSolveAndInverse.h:122), values in the pattern's CSR order. Last right-hand side
 * f - dt K v with boundary conditions (Math/OdeSolverEulerImplicit.cpp:155-160). */
int fem_get_system_matrix(fem_ctx* ctx, double* values);
int fem_get_rhs(fem_ctx* ctx, double* rhs);

/* Element evaluation at positions x (ndof, host) for elements [first, first + count): rotation
 * (9, row-major), element stiffness K (row-major, (3n)^2) and force f (3n) as
 * Fem3DElement*::doUpdateFMDK(state, FMDK) leaves m_R, m_K, m_f
 * (Physics/Fem3DElementCorotationalTetrahedron.cpp:77-288, Physics/Fem3DElementTetrahedron.cpp:77-90,
 * Physics/Fem3DElementCube.cpp:129-141). Any output may be NULL. */
int fem_evaluate_elements(fem_ctx* ctx, const double* x, int64_t first, int64_t count, double* rotation,
                          double* stiffness, double* force);

/* Per-phase kernel timers on the context stream. No counterpart in the reference's FEM path; they
 * stand in for its Framework::Timer instrumentation around update(dt) (Framework/Timer.h:31-93,
 * Physics/PhysicsManager.cpp:130-152). Each timed launch carries a start/stop event pair in the
 * dispatch packet of the first kernel its phase issues, so a phase time is that kernel's execution
 * time. spmv counts only the CG operator applications that did work (launches after convergence
 * are reported as spmv_skipped). */
enum {
    FEM_PHASE_ELEMENT = 0, FEM_PHASE_ASSEMBLE, FEM_PHASE_RHS, FEM_PHASE_CG_INIT, FEM_PHASE_SPMV,
    FEM_PHASE_CG_VECTOR, FEM_PHASE_REFINE, FEM_PHASE_UPDATE, FEM_PHASE_SPMV_SKIPPED, FEM_NUM_PHASES
};
typedef struct fem_phase_times {
    double ms[FEM_NUM_PHASES];       /* summed kernel time per phase */
    int64_t launches[FEM_NUM_PHASES];
} fem_phase_times;
int fem_enable_phase_timers(fem_ctx* ctx, int every);   /* instrument one step in `every` (0: off); resets totals */
int fem_get_phase_times(fem_ctx* ctx, fem_phase_times* out);

/* Measurement helper with no counterpart in the reference: `reps` back-to-back launches of the dominant
 * CG kernel of a tetrahedral mesh (the one-kernel iteration: direction update, block SpMV, vector updates
 * and the 7 block sums; or the operator kernel of the split pair under FEM_CG_KERNELS=split), on the CG
 * scratch buffers of the current precision, each with its own start/stop events; average milliseconds of
 * execution per launch. On a hexahedral mesh: the average of `reps` applications of the element-by-element
 * operator over all colours (the FP32 cached operator, or the FP64 matrix-free one), each between its own
 * pair of events. The simulation state is untouched; call after at least one step (single meshes only). */
int fem_time_operator(fem_ctx* ctx, int reps, double* avg_ms);

/* Measurement helper with no counterpart in the reference: `reps` timed copies of `bytes` bytes
 * (rounded down to 16) between two scratch buffers on `device`; average milliseconds per copy.
 * Used to calibrate HBM byte counters and to report achievable copy bandwidth. */
int fem_probe_copy(int device, int64_t bytes, int reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* FEM_GPU_H */
