/*
 * ogsx.h -- C ABI of the MI355X-native implicit FEM step for OGS-5's
 * LIQUID_FLOW and HEAT_TRANSPORT processes (element assembly + Krylov solve).
 *
 * The entry points replace, one for one, the calls that OGS-5's
 * CRFProcess::Execute (FEM/rf_pcs.cpp:4695-4853) makes on the path:
 *
 *   CreateEQS_LinearSolver -> SparseTable/CSparseMatrix  ogsx_pattern_create,
 *                                                       ogsx_system_create
 *   Linear_EQS::ConfigNumerics                           ogsx_solver_config
 *   Linear_EQS::Initialize                               ogsx_system_initialize
 *   GlobalAssembly -> CFiniteElementStd::Assembly        ogsx_assemble
 *   IncorporateSourceTerms                               ogsx_add_rhs
 *   IncorporateBoundaryConditions -> SetKnownX_i         ogsx_set_dirichlet
 *   Linear_EQS::Solver                                   ogsx_solve
 *   Linear_EQS::Write_BIN                                ogsx_write_bin
 *   CalIntegrationPointValue -> Cal_Velocity             ogsx_gauss_velocity
 *
 * Conventions: every function returns int status (0 = OK, < 0 = error; the
 * message is in ogsx_last_error()); no exception crosses the ABI; pointers
 * are host pointers unless a name says _dev; sizes are int64_t.  A handle is
 * owned by its creator and freed with its *_destroy.  Each system owns one
 * HIP stream; calls on one system are not thread-safe, distinct systems are
 * independent.  All device memory of a system is allocated at creation.
 */
#ifndef OGSX_H
#define OGSX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define OGSX_OK 0
#define OGSX_EINVAL (-1)  /* bad argument                                  */
#define OGSX_EHIP (-2)    /* HIP runtime error or no device                */
#define OGSX_ENOMEM (-3)  /* device/host allocation failed                 */
#define OGSX_ESTATE (-4)  /* call out of order (e.g. solve before assemble) */

/* element types: MshElemType::type numbering (MSH/MshElemType.h) */
#define OGSX_LINE2 1
#define OGSX_QUAD4 2
#define OGSX_HEX8 3
#define OGSX_TET4 5

/* processes (a subset of FiniteElement::ProcessType, FEM/FEMEnums.h; own
 * numbering) */
#define OGSX_LIQUID_FLOW 1
#define OGSX_GROUNDWATER_FLOW 2 /* HEAD form: Laplace K*tuf (CalCoefLaplace,
                                   fem_ele_std.cpp:2308-2384), mass = storage
                                   (CalCoefMass :1625-1630), parabolic step as
                                   LIQUID (:9391), no gravity; mu and rho_g
                                   are not used */
#define OGSX_HEAT_TRANSPORT 3

/* Krylov methods: the $LINEAR_SOLVER method ids of Linear_EQS
 * (FEM/equation_class.cpp:191-250) */
#define OGSX_BICGSTAB 2
#define OGSX_CG 5
/* preconditioners (:260-292): 1 Jacobi; 100 (ILU) falls back to Jacobi */
#define OGSX_PRECOND_NONE 0
#define OGSX_PRECOND_JACOBI 1

/* matrix storage/arithmetic of a system */
#define OGSX_FP64 0     /* FP64 matrix, FP64 Krylov (the reference's)         */
#define OGSX_FP32_IR 1  /* FP64 matrix for residuals/BC + FP32 copy for the
                           inner Krylov solve inside FP64 iterative refinement */

/* Dirichlet elimination modes */
#define OGSX_DIRICHLET_ROW 0  /* CSparseMatrix::Diagonize (matrix_class.cpp:1663):
                                 zero the row except the diagonal,
                                 b_i = a_ii * g; the matrix turns nonsymmetric */
#define OGSX_DIRICHLET_SYMMETRIC 1 /* additionally zero the column and lift
                                 b_k -= a_ki * g (the colDEBUG branch :1691);
                                 same solution, keeps A symmetric for CG      */

/* vectors of a system */
#define OGSX_VEC_X 0      /* Linear_EQS::x (initial guess in, solution out)  */
#define OGSX_VEC_B 1      /* Linear_EQS::b                                   */
#define OGSX_VEC_UOLD 2   /* node values at the old time level (idx0)        */
#define OGSX_VEC_IC 3     /* a stored node-value vector, no counterpart in the
                           * reference: e.g. the initial condition, copied
                           * into X / UOLD on the device (ogsx_copy_vector)
                           * to restart a run without a host transfer      */

typedef struct ogsx_pattern ogsx_pattern;
typedef struct ogsx_system ogsx_system;
typedef struct ogsx_group ogsx_group;

const char* ogsx_last_error(void);
/* Number of visible GPUs (0 on a machine without one).  No other call. */
int ogsx_device_count(int* count);
/* Library build string. */
const char* ogsx_version(void);

/* ---- CRS pattern ------------------------------------------------------ */
/* Builds the sparsity pattern exactly as CFEMesh::ConnectedNodes
 * (MSH/msh_mesh.cpp:2931-2979) + SparseTable CRS (FEM/matrix_class.cpp:
 * 743-777): row i holds the distinct nodes of all elements touching node i
 * in ascending order, the equation index being the node index.  Also builds
 * the element colouring (no two elements of one colour share a node) and the
 * element -> CRS slot map used by the atomic-free scatter.  Host work, done
 * once per mesh.  conn: n_elem * npe node indices. */
int ogsx_pattern_create(int64_t n_nodes, int64_t n_elem, int etype,
                        const int64_t* conn, ogsx_pattern** out);
int ogsx_pattern_info(const ogsx_pattern* p, int64_t* n_rows, int64_t* nnz,
                      int* n_colors);
/* num_column_entries[n+1], entry_column[nnz], diag_entry[n] as the reference's
 * `long` arrays (diag = -1 for a node with no element). */
int ogsx_pattern_export(const ogsx_pattern* p, int64_t* row_ptr, int64_t* col,
                        int64_t* diag);
/* Element colour of every element (n_elem ints). */
int ogsx_pattern_colors(const ogsx_pattern* p, int32_t* color);
/* Device bytes that ogsx_system_create + ogsx_system_set_elements (without
 * Gauss-point velocities) allocate for a system on this pattern: what a
 * caller that admits work by HBM footprint reserves before creating it.
 * No counterpart in the reference (its matrices live in host memory). */
int ogsx_pattern_footprint(const ogsx_pattern* p, int precision, int64_t* bytes);
/* Frees the pattern.  Systems keep pointers into their pattern, so it must
 * outlive them: while systems created on it exist the call fails with
 * OGSX_ESTATE and frees nothing. */
int ogsx_pattern_destroy(ogsx_pattern* p);

/* ---- system: Linear_EQS + the element data it is assembled from -------- */
/* Linear_EQS::Linear_EQS (FEM/equation_class.cpp:85): device A (CRS values
 * on the pattern), b, x and Krylov buffers; bNorm starts at 1.0. */
/* Fails with OGSX_EINVAL for an empty pattern or a node that belongs to no
 * element (its CRS row is empty: the reference would read an undefined
 * diagonal there). */
int ogsx_system_create(const ogsx_pattern* p, int precision, ogsx_system** out);
int ogsx_system_destroy(ogsx_system* s);
/* Device bytes the system has allocated so far. */
int ogsx_system_bytes(const ogsx_system* s, int64_t* bytes);

/* Element data for CFiniteElementStd::Assembly (FEM/fem_ele_std.cpp:9367).
 * Uploaded once; later assemblies reuse it. */
typedef struct {
    int process;          /* OGSX_LIQUID_FLOW / _GROUNDWATER_FLOW / _HEAT_TRANSPORT */
    int dim;              /* problem dimension (size of the K tensor)         */
    int n_gauss;          /* $ELE_GAUSS_POINTS, default 3 (rf_num_new.cpp:124) */
    int cond_stride;      /* 1: isotropic scalar per element; dim*dim: tensor */
    const double* x;      /* node coordinates [n_nodes]                       */
    const double* y;
    const double* z;
    const double* mass_coef; /* [n_elem]: LIQUID CalCoefMass storage
                                (fem_ele_std.cpp:1580-1622); HEAT HeatCapacity
                                rho*c (:1693-1697), divided by tuf            */
    const double* cond;   /* [n_elem*cond_stride]: LIQUID permeability tensor
                             (rf_mmp_new.cpp:4517); HEAT conductivity tensor  */
    double tuf;           /* time_unit_factor                                 */
    double mu;            /* LIQUID viscosity divisor (1 for the HEAD form)   */
    double geo_area;      /* MMP geo_area, multiplies |J| for dim != 3 elems  */
    double adv_coef;      /* HEAT: c_f*rho_f (CalCoefAdvection :3633)         */
    const double* vel;    /* HEAT: Gauss-point velocity [n_elem][3][ngp]
                             (ElementValue::Velocity, fem_ele_std.cpp:10309)
                             or NULL for no advection                         */
    double rho_g;         /* LIQUID gravity density*g (Assemble_Gravity
                             :5859); 0 switches the term off                  */
} ogsx_elements;
int ogsx_system_set_elements(ogsx_system* s, const ogsx_elements* e);

/* Linear_EQS::Initialize (equation_class.cpp:325): A = 0, b = 0. */
int ogsx_system_initialize(ogsx_system* s);

/* GlobalAssembly's element loop (rf_pcs.cpp:5738-5763) over every element:
 * CalcMass/CalcLaplace/CalcAdvection at the Gauss points, the parabolic
 * (LIQUID, theta = nls_relaxation) or mixed hyperbolic-parabolic (HEAT,
 * theta = ls_theta) time discretisation, gravity, and the scatter of the
 * local LHS into A and RHS into b.  steady != 0 drops the mass matrix (HEAT
 * tim_type STEADY, fem_ele_std.cpp:8543; for LIQUID the same operator as a
 * zero storage coefficient).  u_old is the system's OGSX_VEC_UOLD vector. */
int ogsx_assemble(ogsx_system* s, double dt, double theta, int steady);

/* Darcy velocity at the Gauss points of every element of a LIQUID_FLOW (or
 * GROUNDWATER_FLOW: v = -K grad h) system from its current solution
 * (OGSX_VEC_X, the new time level):
 * CRFProcess::CalIntegrationPointValue (rf_pcs.cpp:5915-5994) ->
 * CFiniteElementStd::Cal_Velocity (fem_ele_std.cpp:6554-6868),
 * v = -(K/mu)(grad p + rho_g e_dim) per Gauss point, in the layout of
 * ElementValue::Velocity: [n_elem][3][ngp], components >= dim zero.
 *   heat != NULL  writes it into heat's Gauss-point velocity on the device
 *                 (as if passed in ogsx_elements.vel), switching advection on
 *                 for heat's next ogsx_assemble; heat must be a HEAT_TRANSPORT
 *                 system on the same mesh (element e of both is the same
 *                 element) with the same n_gauss.  Ordered after flow's
 *                 pending work; flow's later work waits for the copy.
 *   host_out      if not NULL, also receives the n_elem*3*ngp values
 *                 (blocking).
 * Neither system may have a solve in progress. */
int ogsx_gauss_velocity(ogsx_system* flow, ogsx_system* heat, double* host_out);

/* IncorporateSourceTerms (rf_pcs.cpp:8002; point sources :8447):
 * b[idx[k]] += val[k] in list order. */
int ogsx_add_rhs(ogsx_system* s, int64_t n, const int64_t* idx,
                 const double* val);

/* IncorporateBoundaryConditions -> Linear_EQS::SetKnownX_i for every
 * Dirichlet node (rf_pcs.cpp:6991, equation_class.cpp:926). */
int ogsx_set_dirichlet(ogsx_system* s, int64_t n, const int64_t* idx,
                       const double* val, int mode);

/* Linear_EQS::ConfigNumerics + Solver (equation_class.cpp:191, :857). */
typedef struct {
    int method;        /* OGSX_CG or OGSX_BICGSTAB                            */
    int precond;       /* OGSX_PRECOND_JACOBI (100 accepted) or NONE          */
    int max_iter;      /* ls_max_iterations                                   */
    double tol;        /* ls_error_tolerance: ||r||/||b|| < tol               */
    /* OGSX_FP32_IR systems only */
    double inner_tol;  /* OGSX_IR_RESTART: relative reduction asked of each
                          FP32 inner solve; OGSX_IR_RELIABLE: the drop delta
                          of the FP32 residual that triggers an FP64 update;
                          <= 0: the default (0.1 reliable; restarted 1e-4
                          for CG, 1e-3 for BiCGStab)                        */
    int max_outer;     /* refinement steps (OGSX_IR_RESTART)                  */
    int ir_mode;       /* OGSX_IR_RESTART or OGSX_IR_RELIABLE (CG only)       */
} ogsx_solver_config;

/* FP32_IR refinement schemes */
#define OGSX_IR_RESTART 0  /* classic: x += solve32(A, b - A x) until done     */
#define OGSX_IR_RELIABLE 1 /* CG with reliable updates: one FP32 recurrence,
                              FP64 residual replacement whenever ||r|| drops by
                              delta; the search direction survives updates    */
#define OGSX_IR_RELIABLE_SR 2 /* as OGSX_IR_RELIABLE, with the FP32 recurrence
                              in single-reduction form (Chronopoulos-Gear: one
                              reduction point, two kernels per iteration) on
                              the symmetrically Jacobi-scaled FP32 matrix
                              S A S, S = diag(1/sqrt|a_ii|); CG only          */

typedef struct {
    int iterations;    /* Krylov iterations (inner iterations summed for IR)  */
    int outer;         /* refinement steps (IR), else 0                       */
    int converged;     /* 1 converged, 0 max_iter reached, 2 skipped by the
                          CheckNormRHS test (equation_class.cpp:1104),
                          3 breakdown (rho or omega below DBL_MIN)            */
    double error;      /* ||r||/||b|| at exit (recurrence residual for FP64;
                          true FP64 residual for IR)                          */
    double bnorm;      /* ||b|| (persists as the next solve's bNorm)          */
    int ret;           /* Linear_EQS::Solver's own return value: BiCGStab the
                          iteration count, CG 1/0 (iter <= max_iter)          */
} ogsx_solve_info;

/* Solves A x = b from the current x.  FP64 systems run the reference's
 * recurrences (Jacobi-CG :1129, right-preconditioned BiCGStab :1287) on the
 * device; OGSX_FP32_IR systems run them in FP32 inside FP64 iterative
 * refinement (r = b - A x, norms and the stopping test in FP64). */
int ogsx_solve(ogsx_system* s, const ogsx_solver_config* cfg,
               ogsx_solve_info* info);

/* Host <-> device copies of OGSX_VEC_* vectors (n_rows doubles). */
int ogsx_set_vector(ogsx_system* s, int which, const double* host);
int ogsx_get_vector(ogsx_system* s, int which, double* host);
/* Device-side copy between vectors (the time-level shift u_old <- x of
 * CRFProcess's copy-back, rf_pcs.cpp:5054-5090). */
int ogsx_copy_vector(ogsx_system* s, int dst, int src);
/* CRS values of A (nnz doubles, row-major CRS order) to the host. */
int ogsx_get_matrix(ogsx_system* s, double* values);
/* y = A x on the device with host vectors (CSparseMatrix::multiVec,
 * matrix_class.cpp:1467); for parity tests of the SpMV kernel alone. */
int ogsx_spmv(ogsx_system* s, const double* x, double* y);
/* Linear_EQS::Write_BIN (equation_class.cpp:409; matrix_class.cpp:1403):
 * int64 n; int64 ptr[n+1]; int64 col[nnz]; f64 A[nnz]; f64 b[n]. */
int ogsx_write_bin(ogsx_system* s, const char* path);

/* ---- groups: many systems solved together ----------------------------- */
/* No counterpart in the reference, which solves one Linear_EQS at a time
 * (CRFProcess::Execute, rf_pcs.cpp:4817); for callers that run many
 * independent systems (ensembles, request streams).  A group launches each
 * phase of the Krylov iteration once over the blocks of all its running
 * systems; every system keeps its own grid, partial sums, state and
 * parameters, so its arithmetic, iteration count and solution are
 * bit-identical to ogsx_solve on that system alone.  Members share the
 * group's precision; solves running at the same time share one method.
 * Group work runs on the group's own HIP stream and is ordered after the
 * work already queued on each started system's stream. */
/* use_graphs = 1 replays the batches of iterations as captured HIP graphs. */
int ogsx_group_create(int precision, int use_graphs, ogsx_group** out);
/* Waits for the group's work and releases its members (they stay valid). */
int ogsx_group_destroy(ogsx_group* g);
/* A system belongs to at most one group; while it does, ogsx_solve on it
 * fails (OGSX_ESTATE) and ogsx_system_destroy needs ogsx_group_remove first. */
int ogsx_group_add(ogsx_group* g, ogsx_system* s);
int ogsx_group_remove(ogsx_group* g, ogsx_system* s);  /* not while solving */
/* Starts solves of `count` members with cfg while others may be running
 * (continuous batching).  Each starts from its current x, as ogsx_solve. */
int ogsx_group_start(ogsx_group* g, ogsx_system* const* systems, int count,
                     const ogsx_solver_config* cfg);
/* Runs up to max_iterations Krylov iterations of the running solves (fewer
 * when all finish); *n_active = solves still running. */
int ogsx_group_advance(ogsx_group* g, int max_iterations, int* n_active);
/* Progress of a system's last solve: *active = 1 while it runs in a group;
 * info as ogsx_solve reports it once finished. */
int ogsx_solve_status(ogsx_system* s, int* active, ogsx_solve_info* info);
/* Starts every member with cfg and runs them to completion; infos in the
 * order the members were added (count = members). */
int ogsx_group_solve(ogsx_group* g, const ogsx_solver_config* cfg,
                     ogsx_solve_info* infos);

/* One time step of `count` members in batched launches, then their solves
 * (as ogsx_group_start).  Per member: u_old <- x when shift is set (the
 * copy-back of the previous step, ogsx_copy_vector), Initialize, the
 * element loop of ogsx_assemble (dt, theta, steady), its last source list
 * (ogsx_add_rhs) when sources is set, and its last Dirichlet list
 * (ogsx_set_dirichlet) with mode dirichlet (-1: none) -- the calls of
 * CRFProcess::Execute for that step (rf_pcs.cpp:4695-4853), each phase one
 * launch over all listed members (one per element class for the element
 * passes), with each member's A and b bit-identical to the per-system
 * calls.  The lists are those the members' last ogsx_add_rhs /
 * ogsx_set_dirichlet uploaded.  Members need the two-pass assembly (rows of
 * at most 32 entries).  No counterpart in the reference. */
typedef struct {
    double dt;
    double theta;
    int steady;
    int dirichlet;  /* -1, OGSX_DIRICHLET_ROW or OGSX_DIRICHLET_SYMMETRIC */
    int sources;    /* 1: add the stored source list                      */
    int shift;      /* 1: u_old <- x first                                */
} ogsx_step_params;
int ogsx_group_step(ogsx_group* g, ogsx_system* const* systems, int count,
                    const ogsx_step_params* params, const ogsx_solver_config* cfg);

/* ---- zero-allocation solve -------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the FP64 solve of ogsx_solve (Jacobi-CG or
 * right-preconditioned BiCGStab, the same kernels and grid, so the same
 * iterates bit for bit) on a CRS matrix the caller holds in device memory.
 * row_ptr_dev[n+1], col_dev[nnz] (int32, columns ascending per row as the
 * CRS pattern builds them) and val_dev[nnz], b_dev[n] are read; x_dev[n] is
 * the initial guess in and the solution out.  bnorm_inout carries
 * Linear_EQS::bNorm between solves for the CheckNormRHS test (1.0 for a
 * first solve, 0 switches the test off; NULL means 1.0); it receives ||b||.
 * stream: a hipStream_t, or NULL for the null stream; the call returns once
 * the solve has finished on it.  cfg->method OGSX_CG or OGSX_BICGSTAB;
 * the FP32_IR fields are ignored.
 * Limit: n <= 524288 rows (2^19); larger n fails with OGSX_EINVAL.  The
 * Krylov vectors, partial sums and state live in one fixed device array of
 * about 38 MB per GPU, held until the process ends; the call makes no
 * hipMalloc, hipFree or other allocation.  Every call on a GPU uses that
 * GPU's one array, so calls that overlap on one GPU (other threads, other
 * streams) race and must be serialised by the caller; each call clears the
 * state it relies on. */
int ogsx_solve_crs_noalloc(int64_t n, const int32_t* row_ptr_dev, const int32_t* col_dev,
                           const double* val_dev, const double* b_dev, double* x_dev,
                           double* bnorm_inout, const ogsx_solver_config* cfg,
                           void* stream, ogsx_solve_info* info);

/* ---- measurement ------------------------------------------------------ */
/* Enables HIP-event timing of every SpMV launch that ogsx_solve issues, of
 * ogsx_assemble and of ogsx_gauss_velocity (on the system's stream) and
 * returns the running totals; bytes_per_launch is the kernel's algorithmic
 * HBM traffic per launch (SpMV, velocity; 0 for the assembly). */
int ogsx_profile_enable(ogsx_system* s, int on);
int ogsx_profile_read(ogsx_system* s, int kind, double* total_ms,
                      int64_t* launches, double* bytes_per_launch);
#define OGSX_PROF_SPMV64 0  /* FP64 SpMV kernel                              */
#define OGSX_PROF_SPMV32 1  /* FP32 SpMV kernel (IR inner solve)             */
#define OGSX_PROF_ASSEMBLY 2
#define OGSX_PROF_VELOCITY 3 /* ogsx_gauss_velocity kernel (on the flow system) */
/* SpMV storage/launch variant: 0 = autotune (time variants 3, 5, 6, 7, 10,
 * 11, 12, 13, 14, 17, 18, 21, 22, 19, 20 on this matrix at first use and keep the
 * fastest; all of them sum each row in CRS order over rounded products,
 * bit-identical to CSparseMatrix::multiVec, and hand each row to the same
 * thread of the same block, so the dot products the solve fuses into its
 * SpMV round identically too: the choice never changes results; the same
 * holds for 9); 1 = tiled CRS (blocks stream
 * contiguous nonzero tiles through LDS); 2 = tiled CRS with 16-bit column
 * offsets per tile (when every tile's columns span < 65536); 3 = SELL-64
 * (a device copy of the CRS values in 64-row slices, column-major, one lane
 * per row); 5 = SELL-64 with 16-bit column offsets per slice; 6 = as 5 with
 * 8 entries in flight per lane; 7 = as 6 with non-temporal matrix loads;
 * 10 = 16 entries in flight, non-temporal; 11 = two slices per wavefront,
 * non-temporal; 12 = 7 entries in flight, non-temporal; 13 / 14 = SELL-64
 * with a stencil dictionary instead of column offsets (one pattern id per
 * row, columns = row + the pattern's offsets; when the pattern has at most
 * 255 distinct offset lists), 8 / 7 entries in flight, non-temporal; 17 / 18
 * = as 13 / 14 with cached matrix loads (a matrix that fits the caches stays
 * there between launches); 21 / 22 = as 13 / 17 taking a row's entries 16
 * at a time with the last batch predicated (rows of up to 16 entries in one
 * round trip); 19 / 20 = SELL-64 with one stencil per slice
 * in triples of consecutive column offsets (one coalesced x load per triple,
 * its +1 / +2 neighbours by lane shifts; zero entries where a row lacks a
 * column of its slice's stencil; when every slice needs at most 16 triples
 * and the stored entries are at most twice the nonzeros), 5 / 4 triples in
 * flight, non-temporal / cached; 9 = plain CRS with one thread per row (no
 * layout beyond the CRS arrays; the zero-allocation solve's format); 4, 8,
 * 16, 32, 64 = vector CRS with that many lanes per row (lane-parallel
 * partial sums: a different rounding order).  With 1, 2 and 4-64 the fused
 * dot products of a solve accumulate rows in another order, so solves
 * round differently (within the usual FP64 tolerance).  The CRS arrays stay the system's matrix;
 * SELL copies are refreshed after A changes.  The FP32 matrix of an FP32-IR
 * system that is symmetric by construction (symmetric tensors, symmetric
 * Dirichlet elimination, no advection) on a banded pattern (at most 16
 * positive column offsets) is instead held as format 25: its diagonal and
 * upper diagonals, each lower entry read from its mirror -- half the values.
 * It is chosen by structure, never by timing (so results do not depend on
 * the tuner), is not selectable here, and sums rows in another order than
 * multiVec, which the FP64 refinement absorbs; a format forced here keeps
 * the FP32 matrix off it.  The FP64 matrix of such an FP64 system is held
 * the same way (format 25 for fmt64), summed in multiVec's CRS order with
 * the rows on the threads of the SELL-64 formats, so its solves equal those
 * of every other tuned format bit for bit. */
int ogsx_set_spmv_variant(ogsx_system* s, int variant);
int ogsx_get_spmv_variant(ogsx_system* s, int* variant);
/* The SpMV formats in use for the FP64 and (IR systems, else 0) FP32
 * matrix: the tuner's fastest, a format forced by ogsx_set_spmv_variant, or
 * one shared with the other systems of a group (ogsx_group_start lets a
 * starting system take a format timed within 10 % or 4 us of its fastest,
 * whichever is larger, when that saves the group a distinct format: one
 * SpMV launch per format and Krylov phase).  No counterpart in the
 * reference (one matrix format there). */
int ogsx_get_spmv_formats(ogsx_system* s, int* fmt64, int* fmt32);
/* The autotuner's table: variant and average ms per SpMV launch (FP64 then,
 * for IR systems, FP32 entries); *count = entries available. */
int ogsx_spmv_tuning(ogsx_system* s, int max_entries, int* variants, double* ms,
                     int* count);
/* Runs `reps` SpMV launches of the system's FP64 (fp32 = 0) or FP32 matrix
 * with its tuned variant on device-resident vectors, timed with HIP events on
 * the system's stream: average ms per launch and the launch's algorithmic
 * bytes (stored matrix entries, column data, slice/row tables, x and y). */
int ogsx_bench_spmv(ogsx_system* s, int fp32, int reps, double* avg_ms,
                    double* bytes_per_launch);
/* Bandwidth calibration: a plain 16-byte-per-lane copy kernel over two
 * device buffers of `bytes` each, run `reps` times; *gbs = (read + write)
 * bytes / second / 1e9. */
int ogsx_stream_copy(int64_t bytes, int reps, double* gbs);
/* Synchronises the system's stream. */
int ogsx_synchronize(ogsx_system* s);

/* ---- persistent launches ---------------------------------------------- */
/* No counterpart in the reference.  Solves of small systems (and groups of
 * them) spend their time dispatching the dependent launches of each Krylov
 * iteration; a persistent launch runs a batch of iterations in one
 * cooperative launch of co-resident blocks, with a grid barrier between
 * phases, over the same per-system chunks, so results are bit-identical to
 * the separate launches.  mode: 0 off (default: on MI355X the grid barrier,
 * 12-15 us across the 8 XCDs, costs more than the ~4 us kernel boundary it
 * replaces); 1 auto (running systems of at most 2^19 rows in all, one
 * slice-based or plain CRS SpMV format, no profiling); 2 whenever the
 * launch shape allows.  Library-wide setting. */
int ogsx_set_persistent(int mode);
/* Current mode and whether the device supports cooperative launches. */
int ogsx_get_persistent(int* mode, int* supported);
/* The barrier's waits are bounded by the host: a batch that does not finish
 * within its deadline (20 s) sets a host flag that releases every wait, the
 * launch ends normally and the call fails with OGSX_EHIP.  This diagnostic
 * exercises that path: it launches two blocks that wait for a third (a
 * barrier that cannot complete) with the flag already set; *aborted = 1
 * when the blocks gave up and the launch ended. */
int ogsx_persistent_giveup_probe(int* aborted);

#ifdef __cplusplus
}
#endif
#endif
