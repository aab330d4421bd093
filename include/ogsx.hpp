// ogsx.hpp -- C++ host classes with OGS-5's call shape, over the C ABI of
// ogsx.h.  A caller written against Linear_EQS / CFiniteElementStd /
// CRFProcess::Execute keeps its sequence of calls:
//
//   OGS-5 (FEM/rf_pcs.cpp:4695-4853)            here
//   eqs_new->ConfigNumerics(m_num)               LinearEQS::ConfigNumerics
//   eqs_new->Initialize()                        LinearEQS::Initialize
//   GlobalAssembly(): fem->Assembly() per elem   FiniteElementStd::Assembly (all elements)
//   IncorporateSourceTerms: eqs_rhs[i] += q      LinearEQS::AddRHS
//   IncorporateBoundaryConditions:
//       eqs_p->SetKnownX_i(i, g)                 LinearEQS::SetKnownX_i
//   iter_lin = eqs_new->Solver()                 LinearEQS::Solver
//   eqs_new->Write_BIN(os)                       LinearEQS::Write_BIN
//
// Errors surface as ogsx::Error exceptions on the host side of the ABI.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ogsx.h"

namespace ogsx {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc, const char* what)
{
    if (rc != OGSX_OK) throw Error(rc, std::string(what) + ": " + ogsx_last_error());
}

// SparseTable (FEM/matrix_class.h:527): CRS pattern from mesh adjacency.
class SparsePattern {
public:
    SparsePattern(int64_t n_nodes, int etype, const std::vector<int64_t>& conn)
    {
        const int npe = etype == OGSX_LINE2 ? 2 : etype == OGSX_HEX8 ? 8 : 4;
        check(ogsx_pattern_create(n_nodes, (int64_t)conn.size() / npe, etype, conn.data(), &p_),
              "ogsx_pattern_create");
    }
    ~SparsePattern() { ogsx_pattern_destroy(p_); }
    SparsePattern(const SparsePattern&) = delete;
    SparsePattern& operator=(const SparsePattern&) = delete;
    int64_t Rows() const
    {
        int64_t n = 0;
        check(ogsx_pattern_info(p_, &n, nullptr, nullptr), "ogsx_pattern_info");
        return n;
    }
    int64_t Nnz() const
    {
        int64_t z = 0;
        check(ogsx_pattern_info(p_, nullptr, &z, nullptr), "ogsx_pattern_info");
        return z;
    }
    // num_column_entries / entry_column / diag_entry (matrix_class.cpp:743-777)
    void Export(std::vector<int64_t>& ptr, std::vector<int64_t>& col,
                std::vector<int64_t>& diag) const
    {
        ptr.resize(Rows() + 1);
        col.resize(Nnz());
        diag.resize(Rows());
        check(ogsx_pattern_export(p_, ptr.data(), col.data(), diag.data()), "ogsx_pattern_export");
    }
    const ogsx_pattern* handle() const { return p_; }

private:
    ogsx_pattern* p_ = nullptr;
};

// Linear_EQS (FEM/equation_class.h:56) with its A, b, x on the device.
class LinearEQS {
public:
    explicit LinearEQS(const SparsePattern& sp, int precision = OGSX_FP64)
    {
        check(ogsx_system_create(sp.handle(), precision, &s_), "ogsx_system_create");
        n_ = sp.Rows();
    }
    ~LinearEQS() { ogsx_system_destroy(s_); }
    LinearEQS(const LinearEQS&) = delete;
    LinearEQS& operator=(const LinearEQS&) = delete;

    // ConfigNumerics (equation_class.cpp:191): $LINEAR_SOLVER method, precond,
    // tolerance, max iterations.  symmetric_dirichlet selects the symmetric
    // elimination that CG needs (the reference's row-only form otherwise).
    // inner_tol 0 takes the library's default for the refinement scheme.
    void ConfigNumerics(int ls_method, int ls_precond, double ls_error_tolerance,
                        int ls_max_iterations, bool symmetric_dirichlet = false,
                        double inner_tol = 0.0, int max_outer = 50,
                        int ir_mode = OGSX_IR_RESTART)
    {
        cfg_.method = ls_method;
        cfg_.precond = ls_precond;
        cfg_.tol = ls_error_tolerance;
        cfg_.max_iter = ls_max_iterations;
        cfg_.inner_tol = inner_tol;
        cfg_.max_outer = max_outer;
        cfg_.ir_mode = ir_mode;
        sym_ = symmetric_dirichlet;
    }
    // Initialize (equation_class.cpp:325): A = 0, b = 0.
    void Initialize()
    {
        bc_idx_.clear();
        bc_val_.clear();
        check(ogsx_system_initialize(s_), "ogsx_system_initialize");
    }
    // Point source terms (rf_pcs.cpp:8447): b[i] += v.
    void AddRHS(const std::vector<int64_t>& idx, const std::vector<double>& val)
    {
        check(ogsx_add_rhs(s_, (int64_t)idx.size(), idx.data(), val.data()), "ogsx_add_rhs");
    }
    // SetKnownX_i (equation_class.cpp:926).  Calls are collected and applied
    // together, in call order, before the next solve (or by
    // IncorporateBoundaryConditions); nothing touches A in between on the path.
    void SetKnownX_i(int64_t i, double x_i)
    {
        bc_idx_.push_back(i);
        bc_val_.push_back(x_i);
    }
    void IncorporateBoundaryConditions()
    {
        if (bc_idx_.empty()) return;
        check(ogsx_set_dirichlet(s_, (int64_t)bc_idx_.size(), bc_idx_.data(), bc_val_.data(),
                                 sym_ ? OGSX_DIRICHLET_SYMMETRIC : OGSX_DIRICHLET_ROW),
              "ogsx_set_dirichlet");
        bc_idx_.clear();
        bc_val_.clear();
    }
    // Solver (equation_class.cpp:857): BiCGStab returns its iteration count,
    // CG returns iter <= max_iter.
    int Solver()
    {
        IncorporateBoundaryConditions();
        check(ogsx_solve(s_, &cfg_, &info_), "ogsx_solve");
        return info_.ret;
    }
    const ogsx_solve_info& Info() const { return info_; }
    void SetX(const std::vector<double>& x) { check(ogsx_set_vector(s_, OGSX_VEC_X, x.data()), "set x"); }
    std::vector<double> GetX() const { return get(OGSX_VEC_X); }
    std::vector<double> GetRHS() const { return get(OGSX_VEC_B); }
    void SetOldTimeLevel(const std::vector<double>& u)
    {
        check(ogsx_set_vector(s_, OGSX_VEC_UOLD, u.data()), "set u_old");
    }
    // CRFProcess copy-back u(t_n+1) = x (rf_pcs.cpp:5054-5090), on the device.
    void CopyBack() { check(ogsx_copy_vector(s_, OGSX_VEC_UOLD, OGSX_VEC_X), "copy back"); }
    // Write_BIN (equation_class.cpp:409).
    void Write_BIN(const std::string& path) { check(ogsx_write_bin(s_, path.c_str()), "ogsx_write_bin"); }
    int64_t Dim() const { return n_; }
    ogsx_system* handle() { return s_; }

private:
    std::vector<double> get(int which) const
    {
        std::vector<double> v(n_);
        check(ogsx_get_vector(s_, which, v.data()), "ogsx_get_vector");
        return v;
    }
    ogsx_system* s_ = nullptr;
    int64_t n_ = 0;
    ogsx_solver_config cfg_{OGSX_BICGSTAB, OGSX_PRECOND_JACOBI, 10000, 1e-10, 1e-4, 50, OGSX_IR_RESTART};
    ogsx_solve_info info_{};
    bool sym_ = false;
    std::vector<int64_t> bc_idx_;
    std::vector<double> bc_val_;
};

// CFiniteElementStd (FEM/fem_ele_std.h): Assembly() runs the element loop of
// CRFProcess::GlobalAssembly (rf_pcs.cpp:5738-5763) for every element at once.
class FiniteElementStd {
public:
    FiniteElementStd(LinearEQS& eqs, const ogsx_elements& e) : eqs_(eqs)
    {
        check(ogsx_system_set_elements(eqs.handle(), &e), "ogsx_system_set_elements");
    }
    // time_step_length and theta (nls_relaxation for LIQUID_FLOW, ls_theta
    // for HEAT_TRANSPORT), as read by the reference's Assembly().
    void Assembly(double time_step_length, double theta, bool steady = false)
    {
        check(ogsx_assemble(eqs_.handle(), time_step_length, theta, steady ? 1 : 0), "ogsx_assemble");
    }
    // CRFProcess::CalIntegrationPointValue -> Cal_Velocity (rf_pcs.cpp:5915,
    // fem_ele_std.cpp:6554) for this LIQUID_FLOW system's current solution:
    // the Gauss-point velocities become `heat`'s advection velocity on the
    // device; `host` (if not null, n_elem*3*ngp doubles) receives them as
    // [n_elem][3][ngp].
    void Cal_Velocity(FiniteElementStd& heat, double* host = nullptr)
    {
        check(ogsx_gauss_velocity(eqs_.handle(), heat.eqs_.handle(), host), "ogsx_gauss_velocity");
    }

private:
    LinearEQS& eqs_;
};

}  // namespace ogsx
