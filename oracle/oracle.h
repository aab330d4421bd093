/*
 * oracle.h -- CPU restatement of OGS-5's implicit FEM step for LIQUID_FLOW
 * and HEAT_TRANSPORT (element assembly + Linear_EQS Krylov solve).
 *
 * THIS IS TEST INFRASTRUCTURE.  It is the checker for the HIP library in
 * ogs5_amd/, and the serial CPU baseline that bench.py times beside it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product path never calls into this file.
 *
 * Every function cites the reference file:line whose semantics it restates
 * (paths relative to the ufz/ogs5 source tree).  The arithmetic order inside
 * the Gauss loops follows the reference so that the restatement reproduces
 * the reference's floating-point results on the same inputs.
 *
 * Parity pinning of each function (what checks it besides itself) is listed
 * in the header comment of oracle.c.
 */
#ifndef OGSX_ORACLE_H
#define OGSX_ORACLE_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element types, numbered as MshElemType::type (MSH/MshElemType.h). */
enum { ORC_LINE2 = 1, ORC_QUAD4 = 2, ORC_HEX8 = 3, ORC_TET4 = 5 };
/* Process types on the path. */
enum { ORC_LIQUID_FLOW = 1, ORC_GROUNDWATER_FLOW = 2, ORC_HEAT_TRANSPORT = 3 };

/* ---- a4: Gauss tables and shape functions (FEM/mathlib.cpp) ---------- */
double orc_gauss_pkt(long grd, long pkt);           /* mathlib.cpp:1112 */
double orc_gauss_fkt(long grd, long pkt);           /* mathlib.cpp:1174 */
void orc_sample_tet5(int gp, double* sp);           /* mathlib.cpp:1285 */
int orc_npe(int etype);
int orc_edim(int etype);
int orc_ngp(int etype, int n_gauss);                /* fem_ele.cpp:362  */
/* unit coordinates of Gauss point gp (fem_ele.cpp:813) and the product of
 * quadrature weights that GetGaussData multiplies |J| by (fem_ele.cpp:885). */
void orc_gauss_point(int etype, int n_gauss, int gp, double unit[3],
                     double wfac[3], int* nw);
void orc_shape(int etype, const double* unit, double* N);   /* N[npe]        */
void orc_grad_shape(int etype, const double* unit, double* dN); /* dN[edim*npe] */
/* Shape-function pool: N[ngp*npe], dN[ngp*edim*npe] (ShapeFunctionPool.cpp:109) */
void orc_shape_pool(int etype, int n_gauss, double* N, double* dN);

/* ---- a1/a2: CRS pattern (MSH/msh_mesh.cpp:2931, FEM/matrix_class.cpp:743) */
/* Pass 1: row_ptr[n+1] (and returns nnz).  Pass 2: col[nnz], diag[n]
 * (diag = -1 for an isolated node).  conn is n_elem*npe node indices. */
int64_t orc_pattern_ptr(int64_t n_nodes, int64_t n_elem, int npe,
                        const int64_t* conn, int64_t* row_ptr);
void orc_pattern_fill(int64_t n_nodes, int64_t n_elem, int npe,
                      const int64_t* conn, const int64_t* row_ptr,
                      int64_t* col, int64_t* diag);

/* ---- a5..a16: element assembly ----------------------------------------- */
typedef struct {
    int etype;          /* ORC_LINE2 / QUAD4 / HEX8 / TET4                    */
    int process;        /* ORC_LIQUID_FLOW, ORC_GROUNDWATER_FLOW (HEAD form) or
                           ORC_HEAT_TRANSPORT                                  */
    int dim;            /* problem (mesh) dimension: tensor is dim x dim       */
    int n_gauss;        /* $ELE_GAUSS_POINTS (default 3)                       */
    int64_t n_nodes, n_elem;
    const double* x;    /* node coordinates, n_nodes each                      */
    const double* y;
    const double* z;
    const int64_t* conn;        /* n_elem*npe                                  */
    const double* mass_coef;    /* per element: storage (liquid) / rho*c (heat) */
    const double* cond;         /* per element tensor, cond_stride values      */
    int cond_stride;            /* 1 = isotropic scalar, dim*dim = full tensor */
    double tuf;                 /* time_unit_factor                            */
    double mu;                  /* viscosity divisor (1 for the HEAD form)     */
    double geo_area;            /* multiplies |J| when element dim != 3        */
    double adv_coef;            /* heat: c_f * rho_f                           */
    const double* vel;          /* heat: [n_elem][3][ngp] Gauss-point velocity */
    double rho_g;               /* liquid gravity: density*g, 0 = off          */
    double dt;
    double theta;               /* liquid: nls_relaxation; heat: ls_theta      */
    int steady;                 /* skip the mass matrix (zero storage)         */
    const double* u_old;        /* previous time level, n_nodes                */
} orc_problem;

/* Local matrices of one element (row-major npe x npe, FEM/matrix_class.h:167):
 * mass (CalcMass fem_ele_std.cpp:3753), laplace (CalcLaplace :4948),
 * advection (CalcAdvection :5280, heat only), grav[npe] (Assemble_Gravity
 * :5859, NodalVal before the RHS add).  Any output may be NULL. */
void orc_element_matrices(const orc_problem* p, int64_t e, double* mass,
                          double* laplace, double* advection, double* grav);
/* LHS (StiffMatrix) and RHS vector of one element (AssembleParabolicEquation
 * :7702 / AssembleMixedHyperbolicParabolicEquation :8504); rhs includes the
 * gravity vector (added after the time-discretised RHS in orc_assemble). */
void orc_element_system(const orc_problem* p, int64_t e, double* lhs,
                        double* rhs);

/* Gauss-point Darcy velocity of a LIQUID_FLOW problem from the nodal values
 * u at the new time level (CFiniteElementStd::Cal_Velocity,
 * fem_ele_std.cpp:6554-6868, for every element as
 * CRFProcess::CalIntegrationPointValue rf_pcs.cpp:5966-5993 does):
 * vel[(e*3+i)*ngp+gp] = ElementValue::Velocity(i, gp). */
void orc_gauss_velocity(const orc_problem* p, const double* u, double* vel);

/* JDS storage of the sparse table (matrix_class.cpp:778-861) and its SpMV
 * (:1501); see oracle.c. */
int64_t orc_jds_build(int64_t n, const int64_t* ptr, const int64_t* col, int64_t* n2o,
                      int64_t* ncol, int64_t* jcol, int64_t* src);
void orc_jds_spmv(int64_t n, int64_t maxc, const int64_t* ncol, const int64_t* n2o,
                  const int64_t* jcol, const double* a, const double* x, double* y);

/* ---- Linear_EQS (FEM/equation_class.h:56) ------------------------------- */
typedef struct {
    int64_t n, nnz;
    const int64_t* ptr;  /* num_column_entries */
    const int64_t* col;  /* entry_column       */
    const int64_t* diag; /* diag_entry         */
    double* a;           /* entry[nnz]         */
    double* b;
    double* x;
    double bnorm;        /* persists across solves, starts at 1.0 (:135)       */
    double error;
    int iter;
    int max_iter;
    double tol;
    int method;          /* 2 = BiCGStab, 5 = CG (ConfigNumerics :191)         */
    int precond;         /* 1 = Jacobi (100 maps to 1)                         */
} orc_eqs;

void orc_eqs_init(orc_eqs* q, int64_t n, int64_t nnz, const int64_t* ptr,
                  const int64_t* col, const int64_t* diag, double* a,
                  double* b, double* x);
void orc_eqs_config(orc_eqs* q, int method, int precond, double tol,
                    int max_iter);                   /* :191  */
void orc_eqs_initialize(orc_eqs* q);                 /* :325  A=0,b=0 */
/* Global assembly: loop over elements in index order, local system, then
 * add into A via binary search per entry (add2GlobalMatrixII :8272 and
 * CSparseMatrix::operator() matrix_class.cpp:1178), RHS into b. */
void orc_assemble(orc_eqs* q, const orc_problem* p);
void orc_add_rhs(orc_eqs* q, int64_t n, const int64_t* idx,
                 const double* val);                 /* rf_pcs.cpp:8447 */
void orc_set_known_x(orc_eqs* q, int64_t i, double g);   /* :926 -> Diagonize */
/* Symmetric elimination (zero row and column, lift b); the colDEBUG branch
 * of CSparseMatrix::Diagonize matrix_class.cpp:1691 made the default. */
void orc_set_known_x_sym(orc_eqs* q, int64_t i, double g);
void orc_spmv(const orc_eqs* q, const double* x, double* y);  /* :1467 */
void orc_precond_jacobi(const orc_eqs* q, const double* r, double* z); /* :1786 */
double orc_dot(int64_t n, const double* a, const double* b);  /* :981 */
int orc_cg(orc_eqs* q);          /* :1129 returns iter<=max_iter (or 0 skip) */
int orc_bicgstab(orc_eqs* q);    /* :1287 returns iteration count            */
int orc_solver(orc_eqs* q);      /* :857  dispatch                           */
/* Write_BIN (equation_class.cpp:409, matrix_class.cpp:1403) to a file path. */
int orc_write_bin(const orc_eqs* q, const char* path);

/* Dense local matrix helper: vec_result[i] += fac * sum_j A(i,j) vec[j]
 * (Matrix::multi(const double*, double*, double) matrix_class.cpp:288). */
void orc_dense_multi(const double* A, int nrows, int ncols, const double* vec,
                     double* vec_result, double fac);

#ifdef __cplusplus
}
#endif
#endif
