// pattern.cpp -- host construction of the CRS sparsity pattern, the element
// colouring (the element -> CRS slot map is built on the device).
//
// The CRS arrays reproduce the reference bit for bit:
//   CFEMesh::ConnectedNodes(false)  MSH/msh_mesh.cpp:2931-2979: for every node,
//       the distinct nodes of all elements connected to it, sorted ascending;
//   SparseTable::SparseTable CRS    FEM/matrix_class.cpp:743-777: rows in node
//       order, num_column_entries = running count, diag_entry = position of
//       the node itself, linear elements keep every column < rows.
// Instead of the reference's quadratic dedupe + exchange sort per node this
// gathers candidates from the node -> element incidence and sorts/uniques
// them; the resulting sets and order are the same.
//
// The colouring is greedy first-fit over elements in index order: an element
// takes the lowest colour not used by any element sharing one of its nodes.
// Elements of one colour therefore touch disjoint rows of A and disjoint
// entries of b, so a colour can be scattered without atomics; colours run
// one after the other in a fixed order, which makes the summation order of
// every entry independent of scheduling.
#include <algorithm>
#include <cstring>
#include <thread>

#include "ogsx_internal.h"

namespace ogsx {

namespace {

template <class F>
void parallel_for(int64_t n, F&& f)
{
    unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::min<unsigned>(hw ? hw : 1, 16);
    if (n < 100000 || nt <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) {
        int64_t b = n * t / nt, e = n * (t + 1) / nt;
        th.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& x : th) x.join();
}

}  // namespace

int build_pattern(int64_t n_nodes, int64_t n_elem, int etype,
                  const int64_t* conn, PatternHost& P, std::string& msg)
{
    const int npe = npe_of(etype);
    if (npe == 0) {
        msg = "unsupported element type";
        return OGSX_EINVAL;
    }
    if (n_nodes <= 0 || n_elem < 0 || n_nodes >= (int64_t(1) << 31)) {
        msg = "node count out of range";
        return OGSX_EINVAL;
    }
    if (n_elem > 0 && !conn) {
        msg = "null connectivity";
        return OGSX_EINVAL;
    }
    P = PatternHost();
    P.n = n_nodes;
    P.n_elem = n_elem;
    P.etype = etype;
    P.npe = npe;
    P.conn.resize((size_t)n_elem * npe);
    for (int64_t k = 0; k < n_elem * npe; k++) {
        if (conn[k] < 0 || conn[k] >= n_nodes) {
            msg = "connectivity index out of range";
            return OGSX_EINVAL;
        }
        P.conn[k] = (int32_t)conn[k];
    }
    // node -> element incidence (each element listed once per node)
    std::vector<int64_t> iptr(n_nodes + 1, 0);
    for (int64_t e = 0; e < n_elem; e++)
        for (int a = 0; a < npe; a++) {
            bool dup = false;
            for (int b = 0; b < a; b++) dup |= P.conn[e * npe + b] == P.conn[e * npe + a];
            if (!dup) iptr[P.conn[e * npe + a] + 1]++;
        }
    for (int64_t i = 0; i < n_nodes; i++) iptr[i + 1] += iptr[i];
    std::vector<int32_t> iele(iptr[n_nodes]);
    {
        std::vector<int64_t> fill(iptr.begin(), iptr.end() - 1);
        for (int64_t e = 0; e < n_elem; e++)
            for (int a = 0; a < npe; a++) {
                bool dup = false;
                for (int b = 0; b < a; b++) dup |= P.conn[e * npe + b] == P.conn[e * npe + a];
                if (!dup) iele[fill[P.conn[e * npe + a]]++] = (int32_t)e;
            }
    }
    if (n_elem * npe < (int64_t(1) << 31)) {
        P.inc_ptr.assign(n_nodes + 1, 0);
        for (int64_t k = 0; k < n_elem * npe; k++) P.inc_ptr[P.conn[k] + 1]++;
        for (int64_t i = 0; i < n_nodes; i++) P.inc_ptr[i + 1] += P.inc_ptr[i];
        P.inc_ele.resize(n_elem * npe);
        std::vector<int32_t> fill(P.inc_ptr.begin(), P.inc_ptr.end() - 1);
        for (int64_t k = 0; k < n_elem * npe; k++) P.inc_ele[fill[P.conn[k]]++] = (int32_t)k;
    }
    // row lengths, then columns
    std::vector<int64_t> rowlen(n_nodes);
    auto gather = [&](int64_t i, std::vector<int32_t>& buf) {
        buf.clear();
        for (int64_t k = iptr[i]; k < iptr[i + 1]; k++) {
            const int32_t* en = &P.conn[(int64_t)iele[k] * npe];
            for (int a = 0; a < npe; a++) buf.push_back(en[a]);
        }
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    };
    parallel_for(n_nodes, [&](int64_t b, int64_t e) {
        std::vector<int32_t> buf;
        for (int64_t i = b; i < e; i++) {
            gather(i, buf);
            rowlen[i] = (int64_t)buf.size();
        }
    });
    P.ptr.assign(n_nodes + 1, 0);
    for (int64_t i = 0; i < n_nodes; i++) {
        P.ptr[i + 1] = P.ptr[i] + rowlen[i];
        P.row_max = std::max(P.row_max, rowlen[i]);
    }
    P.nnz = P.ptr[n_nodes];
    if (P.nnz >= (int64_t(1) << 31)) {
        msg = "nnz exceeds the 32-bit index range";
        return OGSX_EINVAL;
    }
    P.col.resize(P.nnz);
    P.diag.assign(n_nodes, -1);
    parallel_for(n_nodes, [&](int64_t b, int64_t e) {
        std::vector<int32_t> buf;
        for (int64_t i = b; i < e; i++) {
            gather(i, buf);
            int64_t w = P.ptr[i];
            for (int32_t c : buf) {
                if (c == i) P.diag[i] = w;
                P.col[w++] = c;
            }
        }
    });
    // greedy colouring with per-node masks of colours already used
    constexpr int kWords = 4;  // up to 256 colours
    std::vector<uint64_t> nodemask((size_t)n_nodes * kWords, 0);
    P.color.resize(n_elem);
    int ncol = 0;
    for (int64_t e = 0; e < n_elem; e++) {
        uint64_t forb[kWords] = {0, 0, 0, 0};
        const int32_t* en = &P.conn[e * npe];
        for (int a = 0; a < npe; a++)
            for (int w = 0; w < kWords; w++) forb[w] |= nodemask[(size_t)en[a] * kWords + w];
        int c = -1;
        for (int w = 0; w < kWords && c < 0; w++)
            if (~forb[w]) c = w * 64 + __builtin_ctzll(~forb[w]);
        if (c < 0) {
            msg = "element colouring needs more than 256 colours";
            return OGSX_EINVAL;
        }
        P.color[e] = c;
        ncol = std::max(ncol, c + 1);
        for (int a = 0; a < npe; a++) nodemask[(size_t)en[a] * kWords + c / 64] |= 1ull << (c % 64);
    }
    P.n_colors = ncol;
    P.color_ptr.assign(ncol + 1, 0);
    for (int64_t e = 0; e < n_elem; e++) P.color_ptr[P.color[e] + 1]++;
    for (int c = 0; c < ncol; c++) P.color_ptr[c + 1] += P.color_ptr[c];
    P.color_elems.resize(n_elem);
    {
        std::vector<int32_t> fill(P.color_ptr.begin(), P.color_ptr.end() - 1);
        for (int64_t e = 0; e < n_elem; e++) P.color_elems[fill[P.color[e]]++] = (int32_t)e;
    }
    return OGSX_OK;
}

// Gauss tables: the reference's literals (FEM/mathlib.cpp:1112-1215, 1285-
// 1322) and shape functions (:1661-2302); point order and weight products as
// CElement::SetGaussPoint / GetGaussData (FEM/fem_ele.cpp:813-933).
namespace {
double gauss_pkt(int grd, int pkt)
{
    static const double p2[2] = {0.577350269189626, -0.577350269189626};
    static const double p3[3] = {0.774596669241483, 0.0, -0.774596669241483};
    static const double p4[4] = {0.861136311594053, 0.339981043584856,
                                 -0.339981043584856, -0.861136311594053};
    if (grd == 1) return 0.0;
    if (grd == 2) return p2[pkt];
    if (grd == 3) return p3[pkt];
    if (grd == 4) return p4[pkt];
    return 0.0;
}
double gauss_fkt(int grd, int pkt)
{
    static const double w3[3] = {0.555555555555556, 0.888888888888889, 0.555555555555556};
    static const double w4[4] = {0.347854845137454, 0.652145154862546,
                                 0.652145154862546, 0.347854845137454};
    if (grd == 1) return 2.0;
    if (grd == 2) return 1.0;
    if (grd == 3) return w3[pkt];
    if (grd == 4) return w4[pkt];
    return 0.0;
}
const double kTet5[5][4] = {
    {0.25, 0.25, 0.25, -0.133333333333333},
    {0.166666666666667, 0.166666666666667, 0.166666666666667, 0.075},
    {0.5, 0.166666666666667, 0.166666666666667, 0.075},
    {0.166666666666667, 0.5, 0.166666666666667, 0.075},
    {0.166666666666667, 0.166666666666667, 0.5, 0.075}};

void shape(int et, const double* u, double* N)
{
    switch (et) {
        case OGSX_LINE2:
            N[0] = 1.0 - u[0];
            N[1] = 1.0 + u[0];
            for (int i = 0; i < 2; i++) N[i] *= 0.5;
            break;
        case OGSX_QUAD4:
            N[0] = (1.0 + u[0]) * (1.0 + u[1]);
            N[1] = (1.0 - u[0]) * (1.0 + u[1]);
            N[2] = (1.0 - u[0]) * (1.0 - u[1]);
            N[3] = (1.0 + u[0]) * (1.0 - u[1]);
            for (int i = 0; i < 4; i++) N[i] *= 0.25;
            break;
        case OGSX_TET4:
            N[0] = 1. - u[0] - u[1] - u[2];
            N[1] = u[0];
            N[2] = u[1];
            N[3] = u[2];
            break;
        case OGSX_HEX8: {
            const double r = u[0], s = u[1], t = u[2];
            const double sr[8] = {1, -1, -1, 1, 1, -1, -1, 1};
            const double ss[8] = {1, 1, -1, -1, 1, 1, -1, -1};
            const double st[8] = {1, 1, 1, 1, -1, -1, -1, -1};
            for (int i = 0; i < 8; i++)
                N[i] = (sr[i] > 0 ? 1.0 + r : 1.0 - r) * (ss[i] > 0 ? 1.0 + s : 1.0 - s) *
                       (st[i] > 0 ? 1.0 + t : 1.0 - t);
            for (int i = 0; i < 8; i++) N[i] *= 0.125;
            break;
        }
    }
}

void grad_shape(int et, const double* u, double* d)
{
    switch (et) {
        case OGSX_LINE2:
            d[0] = -0.5;
            d[1] = 0.5;
            break;
        case OGSX_QUAD4:
            d[0] = +(1.0 + u[1]);
            d[1] = -(1.0 + u[1]);
            d[2] = -(1.0 - u[1]);
            d[3] = +(1.0 - u[1]);
            d[4] = +(1.0 + u[0]);
            d[5] = +(1.0 - u[0]);
            d[6] = -(1.0 - u[0]);
            d[7] = -(1.0 + u[0]);
            for (int i = 0; i < 8; i++) d[i] *= 0.25;
            break;
        case OGSX_TET4: {
            const double g[12] = {-1.0, 1.0, 0.0, 0.0, -1.0, 0.0, 1.0, 0.0, -1.0, 0.0, 0.0, 1.0};
            for (int i = 0; i < 12; i++) d[i] = g[i];
            break;
        }
        case OGSX_HEX8: {
            const double r = u[0], s = u[1], t = u[2];
            // d/dr
            d[0] = +(1.0 + s) * (1.0 + t);
            d[1] = -(1.0 + s) * (1.0 + t);
            d[2] = -(1.0 - s) * (1.0 + t);
            d[3] = +(1.0 - s) * (1.0 + t);
            d[4] = +(1.0 + s) * (1.0 - t);
            d[5] = -(1.0 + s) * (1.0 - t);
            d[6] = -(1.0 - s) * (1.0 - t);
            d[7] = +(1.0 - s) * (1.0 - t);
            // d/ds
            d[8] = +(1.0 + r) * (1.0 + t);
            d[9] = +(1.0 - r) * (1.0 + t);
            d[10] = -(1.0 - r) * (1.0 + t);
            d[11] = -(1.0 + r) * (1.0 + t);
            d[12] = +(1.0 + r) * (1.0 - t);
            d[13] = +(1.0 - r) * (1.0 - t);
            d[14] = -(1.0 - r) * (1.0 - t);
            d[15] = -(1.0 + r) * (1.0 - t);
            // d/dt
            d[16] = +(1.0 + r) * (1.0 + s);
            d[17] = +(1.0 - r) * (1.0 + s);
            d[18] = +(1.0 - r) * (1.0 - s);
            d[19] = +(1.0 + r) * (1.0 - s);
            d[20] = -(1.0 + r) * (1.0 + s);
            d[21] = -(1.0 - r) * (1.0 + s);
            d[22] = -(1.0 - r) * (1.0 - s);
            d[23] = -(1.0 + r) * (1.0 - s);
            for (int i = 0; i < 24; i++) d[i] *= 0.125;
            break;
        }
    }
}
}  // namespace

int build_gauss_table(int et, int ng, GaussTable& t)
{
    if (ng < 1 || ng > 4) return OGSX_EINVAL;
    t.npe = npe_of(et);
    t.edim = edim_of(et);
    t.ngp = ngp_of(et, ng);
    if (!t.npe || t.ngp > kMaxGaussPts) return OGSX_EINVAL;
    t.N.assign(t.ngp * t.npe, 0.0);
    t.dN.assign(t.ngp * t.edim * t.npe, 0.0);
    t.w.assign(t.ngp * 3, 1.0);
    for (int gp = 0; gp < t.ngp; gp++) {
        double u[3] = {0, 0, 0};
        double* w = &t.w[gp * 3];
        switch (et) {
            case OGSX_LINE2:
                u[0] = gauss_pkt(ng, gp);
                w[0] = gauss_fkt(ng, gp);
                t.nw = 1;
                break;
            case OGSX_QUAD4:
                u[0] = gauss_pkt(ng, gp / ng);
                u[1] = gauss_pkt(ng, gp % ng);
                w[0] = gauss_fkt(ng, gp / ng);
                w[1] = gauss_fkt(ng, gp % ng);
                t.nw = 2;
                break;
            case OGSX_HEX8: {
                int r = gp / (ng * ng), s = gp % (ng * ng), tt = s % ng;
                s /= ng;
                u[0] = gauss_pkt(ng, r);
                u[1] = gauss_pkt(ng, s);
                u[2] = gauss_pkt(ng, tt);
                w[0] = gauss_fkt(ng, r);
                w[1] = gauss_fkt(ng, s);
                w[2] = gauss_fkt(ng, tt);
                t.nw = 3;
                break;
            }
            case OGSX_TET4:
                u[0] = kTet5[gp][0];
                u[1] = kTet5[gp][1];
                u[2] = kTet5[gp][2];
                w[0] = kTet5[gp][3];
                t.nw = 1;
                break;
        }
        shape(et, u, &t.N[gp * t.npe]);
        grad_shape(et, u, &t.dN[gp * t.edim * t.npe]);
    }
    return OGSX_OK;
}

}  // namespace ogsx
