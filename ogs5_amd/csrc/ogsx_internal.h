// ogsx_internal.h -- internal types shared by the host runtime (api.hip,
// pattern.cpp) and the device kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ogsx.h"

namespace ogsx {

constexpr int kBlock = 256;        // threads per block for n-sized kernels
constexpr int kMaxGaussPts = 64;   // hex8 at n_gauss = 4
constexpr int kNumXcd = 8;         // MI355X: 8 XCDs, blocks dealt round-robin

inline int npe_of(int et)
{
    switch (et) {
        case OGSX_LINE2: return 2;
        case OGSX_QUAD4: return 4;
        case OGSX_HEX8: return 8;
        case OGSX_TET4: return 4;
    }
    return 0;
}
inline int edim_of(int et)
{
    switch (et) {
        case OGSX_LINE2: return 1;
        case OGSX_QUAD4: return 2;
        case OGSX_HEX8: return 3;
        case OGSX_TET4: return 3;
    }
    return 0;
}
inline int ngp_of(int et, int ng)
{
    switch (et) {
        case OGSX_LINE2: return ng;
        case OGSX_QUAD4: return ng * ng;
        case OGSX_HEX8: return ng * ng * ng;
        case OGSX_TET4: return 5;
    }
    return 0;
}

// Host-side pattern: CRS in the reference's layout plus the element
// colouring of the atomic-free scatter.
struct PatternHost {
    int64_t n = 0, nnz = 0, n_elem = 0;
    int etype = 0, npe = 0;
    std::vector<int64_t> ptr, col, diag;   // reference layout (long)
    std::vector<int32_t> conn;             // [n_elem*npe]
    std::vector<int32_t> color;            // [n_elem]
    std::vector<int32_t> color_ptr;        // [n_colors+1]
    std::vector<int32_t> color_elems;      // [n_elem] element ids by colour
    int n_colors = 0;
    // node -> (element, local node) incidence e * npe + i, ascending (an
    // element with a repeated node is listed once per occurrence)
    std::vector<int32_t> inc_ptr;          // [n+1]
    std::vector<int32_t> inc_ele;
    int64_t row_max = 0;                   // longest CRS row
};

// Builds everything in PatternHost; returns OGSX_OK or an error with msg.
int build_pattern(int64_t n_nodes, int64_t n_elem, int etype,
                  const int64_t* conn, PatternHost& out, std::string& msg);

// Shape-function pool for one element type (FEM/ShapeFunctionPool.cpp:109):
// N[gp*npe+a], dN[(gp*edim+k)*npe+a] and the weight factors
// w[gp*3+0..2] (unused factors = 1, see GaussTable::nw).
struct GaussTable {
    int ngp = 0, npe = 0, edim = 0, nw = 0;
    std::vector<double> N, dN, w;
};
int build_gauss_table(int etype, int n_gauss, GaussTable& t);

// Scalars that the Krylov kernels exchange through device memory.  Every
// field is written by exactly one kernel (block 0) and read by others, so
// no kernel reads a field that another block of the same kernel writes.
struct KrylovState {
    double bnorm;     // ||b|| of this solve
    double error;     // ||r||/||b|| (or ||s||/||b||) at exit
    double rr;        // CG: (r,z) of the current iterate
    double rr_old;    // CG: (r,z) of the previous iterate (copied by K2)
    double alpha;
    double omega;
    double rho0;      // BiCGStab: rho of the previous iteration
    double rho1;      // BiCGStab: rho of this iteration
    int iter;         // completed iterations
    int done;         // 0 running, 1 converged, 2 max_iter, 3 breakdown
                      // (BiCGStab: omega), 4 skipped by CheckNormRHS,
                      // 5 reliable-update request, 6 BiCGStab rho breakdown
    int max_iter;
    int pend;         // BiCGStab: converged on s, x updated; done follows
    // single-reduction CG: the scalars an update kernel hands to the next
    // iteration (rr_old, alpha, iter take them over in the next SpMV
    // kernel, so no kernel reads a field that it also writes)
    double sr_gam;    // (r,r) of the iterate the last update started from
    double sr_alp;    // alpha of the last update
    int sr_it;        // iterations completed (>= iter)
    int sr_pad;
};

}  // namespace ogsx
