// Host checks of what pattern.cpp builds -- the CRS pattern
// (CFEMesh::ConnectedNodes, MSH/msh_mesh.cpp:2931-2979, and SparseTable's CRS
// layout, FEM/matrix_class.cpp:743-777), the element colouring, the node ->
// element incidence and the Gauss tables -- compiled together with
// pattern.cpp under the compiler's address, undefined-behaviour or thread
// sanitizer.  Exit status 0 when every check holds; the sanitizers abort on
// the first report.
//
//   pattern_check            every mesh, pattern checked against brute force
//                            where the mesh is small
//   pattern_check big        one 48^3 hex mesh (117649 nodes), large enough
//                            for the pattern's worker threads
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "ogsx_internal.h"

using namespace ogsx;

namespace {

int fails = 0;

#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);     \
            fails++;                                                              \
        }                                                                         \
    } while (0)

struct Mesh {
    const char* name;
    int64_t n;
    int et;
    std::vector<int64_t> conn;
};

Mesh line_chain(int m)
{
    Mesh s{"line", m + 1, OGSX_LINE2, {}};
    for (int i = 0; i < m; i++) s.conn.insert(s.conn.end(), {i, i + 1});
    return s;
}

// hub node 0 in k line elements, then a chain: row 0 has k + 1 entries
Mesh star(int k)
{
    Mesh s{"star", k + 1, OGSX_LINE2, {}};
    for (int i = 1; i <= k; i++) s.conn.insert(s.conn.end(), {0, i});
    for (int i = 1; i < k; i++) s.conn.insert(s.conn.end(), {i, i + 1});
    return s;
}

Mesh quad_grid(int nx, int ny)
{
    Mesh s{"quad", (int64_t)(nx + 1) * (ny + 1), OGSX_QUAD4, {}};
    auto id = [&](int i, int j) { return (int64_t)j * (nx + 1) + i; };
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
            s.conn.insert(s.conn.end(), {id(i, j), id(i + 1, j), id(i + 1, j + 1), id(i, j + 1)});
    return s;
}

Mesh hex_grid(int n)
{
    const int64_t m = n + 1;
    Mesh s{"hex", m * m * m, OGSX_HEX8, {}};
    auto id = [&](int64_t i, int64_t j, int64_t k) { return (k * m + j) * m + i; };
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++)
                s.conn.insert(s.conn.end(),
                              {id(i, j, k), id(i + 1, j, k), id(i + 1, j + 1, k), id(i, j + 1, k),
                               id(i, j, k + 1), id(i + 1, j, k + 1), id(i + 1, j + 1, k + 1),
                               id(i, j + 1, k + 1)});
    return s;
}

// six tets per cube, each on one monotone path from corner 0 to corner 6
Mesh tet_grid(int n)
{
    const int64_t m = n + 1;
    Mesh s{"tet", m * m * m, OGSX_TET4, {}};
    auto id = [&](int64_t i, int64_t j, int64_t k) { return (k * m + j) * m + i; };
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const int64_t c[8] = {id(i, j, k),         id(i + 1, j, k),
                                      id(i + 1, j + 1, k), id(i, j + 1, k),
                                      id(i, j, k + 1),     id(i + 1, j, k + 1),
                                      id(i + 1, j + 1, k + 1), id(i, j + 1, k + 1)};
                static const int path[6][2] = {{1, 2}, {1, 5}, {3, 2}, {3, 7}, {4, 5}, {4, 7}};
                for (const auto& p : path) s.conn.insert(s.conn.end(), {c[0], c[p[0]], c[p[1]], c[6]});
            }
    return s;
}

// quads with a repeated node (a collapsed edge): the incidence lists the
// element once per occurrence, the pattern once
Mesh degenerate()
{
    Mesh s = quad_grid(3, 2);
    s.name = "degenerate";
    s.conn[2] = s.conn[1];
    return s;
}

void check_mesh(const Mesh& m, bool brute)
{
    const int npe = npe_of(m.et);
    const int64_t ne = (int64_t)m.conn.size() / npe;
    PatternHost P;
    std::string msg;
    const int rc = build_pattern(m.n, ne, m.et, m.conn.data(), P, msg);
    CHECK(rc == OGSX_OK);
    if (rc != OGSX_OK) {
        std::fprintf(stderr, "%s: %s\n", m.name, msg.c_str());
        return;
    }
    CHECK(P.n == m.n && P.n_elem == ne && P.npe == npe);
    CHECK((int64_t)P.ptr.size() == m.n + 1 && P.ptr[0] == 0 && P.ptr[m.n] == P.nnz);
    int64_t row_max = 0;
    for (int64_t i = 0; i < m.n; i++) {
        CHECK(P.ptr[i + 1] >= P.ptr[i]);
        row_max = std::max(row_max, P.ptr[i + 1] - P.ptr[i]);
        for (int64_t k = P.ptr[i]; k < P.ptr[i + 1]; k++) {
            CHECK(P.col[k] >= 0 && P.col[k] < m.n);
            if (k > P.ptr[i]) CHECK(P.col[k] > P.col[k - 1]);  // ascending, distinct
        }
        const int64_t d = P.diag[i];
        if (d >= 0) CHECK(d >= P.ptr[i] && d < P.ptr[i + 1] && P.col[d] == i);
    }
    CHECK(P.row_max == row_max);
    // every element pair is an entry
    for (int64_t e = 0; e < ne; e++)
        for (int a = 0; a < npe; a++)
            for (int b = 0; b < npe; b++) {
                const int64_t r = m.conn[e * npe + a], c = m.conn[e * npe + b];
                const int64_t* lo = P.col.data() + P.ptr[r];
                const int64_t* hi = P.col.data() + P.ptr[r + 1];
                CHECK(std::binary_search(lo, hi, c));
            }
    if (brute) {  // and nothing else
        std::vector<std::set<int64_t>> rows(m.n);
        for (int64_t e = 0; e < ne; e++)
            for (int a = 0; a < npe; a++)
                for (int b = 0; b < npe; b++) rows[m.conn[e * npe + a]].insert(m.conn[e * npe + b]);
        for (int64_t i = 0; i < m.n; i++) {
            CHECK((int64_t)rows[i].size() == P.ptr[i + 1] - P.ptr[i]);
            CHECK(rows[i].empty() == (P.diag[i] < 0));
        }
    }
    // colouring: elements of one colour share no node
    CHECK((int64_t)P.color_ptr.size() == P.n_colors + 1 && P.color_ptr[P.n_colors] == ne);
    std::vector<int32_t> seen(m.n, -1);
    for (int c = 0; c < P.n_colors; c++)
        for (int k = P.color_ptr[c]; k < P.color_ptr[c + 1]; k++) {
            const int32_t e = P.color_elems[k];
            CHECK(P.color[e] == c);
            std::vector<int64_t> nodes(&m.conn[(int64_t)e * npe], &m.conn[(int64_t)e * npe] + npe);
            std::sort(nodes.begin(), nodes.end());
            nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
            for (int64_t v : nodes) {
                CHECK(seen[v] != c);
                seen[v] = c;
            }
        }
    // incidence: node i lists e npe + a for every occurrence, ascending
    CHECK((int64_t)P.inc_ptr.size() == m.n + 1 && P.inc_ptr[m.n] == ne * npe);
    for (int64_t i = 0; i < m.n; i++)
        for (int64_t q = P.inc_ptr[i]; q < P.inc_ptr[i + 1]; q++) {
            CHECK(m.conn[P.inc_ele[q]] == i);
            if (q > P.inc_ptr[i]) CHECK(P.inc_ele[q] > P.inc_ele[q - 1]);
        }
    std::printf("%-10s n=%lld elements=%lld nnz=%lld colours=%d row_max=%lld\n", m.name,
                (long long)m.n, (long long)ne, (long long)P.nnz, P.n_colors, (long long)P.row_max);
}

void check_errors()
{
    PatternHost P;
    std::string msg;
    const std::vector<int64_t> bad = {0, 5};
    CHECK(build_pattern(3, 1, OGSX_LINE2, bad.data(), P, msg) == OGSX_EINVAL);
    CHECK(msg.find("out of range") != std::string::npos);
    CHECK(build_pattern(0, 0, OGSX_LINE2, nullptr, P, msg) == OGSX_EINVAL);
    CHECK(build_pattern(3, 1, 7, bad.data(), P, msg) == OGSX_EINVAL);
    CHECK(build_pattern(3, 1, OGSX_LINE2, nullptr, P, msg) == OGSX_EINVAL);
    CHECK(build_pattern(4, 0, OGSX_QUAD4, nullptr, P, msg) == OGSX_OK);  // no elements
    CHECK(P.nnz == 0 && P.n_colors == 0);
}

void check_gauss()
{
    const int ets[4] = {OGSX_LINE2, OGSX_QUAD4, OGSX_HEX8, OGSX_TET4};
    const double volume[4] = {2.0, 4.0, 8.0, 1.0 / 6.0};
    for (int t = 0; t < 4; t++)
        for (int ng = 1; ng <= 4; ng++) {
            GaussTable G;
            CHECK(build_gauss_table(ets[t], ng, G) == OGSX_OK);
            double vol = 0.0;
            for (int gp = 0; gp < G.ngp; gp++) {
                double sn = 0.0;
                for (int a = 0; a < G.npe; a++) sn += G.N[gp * G.npe + a];
                CHECK(std::fabs(sn - 1.0) < 1e-14);  // partition of unity
                for (int k = 0; k < G.edim; k++) {
                    double sd = 0.0;
                    for (int a = 0; a < G.npe; a++) sd += G.dN[(gp * G.edim + k) * G.npe + a];
                    CHECK(std::fabs(sd) < 1e-14);
                }
                double w = 1.0;
                for (int k = 0; k < G.nw; k++) w *= G.w[gp * 3 + k];
                vol += w;
            }
            CHECK(std::fabs(vol - volume[t]) < 1e-12);
        }
    GaussTable G;
    CHECK(build_gauss_table(OGSX_HEX8, 5, G) == OGSX_EINVAL);
    CHECK(build_gauss_table(OGSX_QUAD4, 0, G) == OGSX_EINVAL);
    CHECK(build_gauss_table(9, 2, G) == OGSX_EINVAL);
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc > 1 && std::strcmp(argv[1], "big") == 0) {
        check_mesh(hex_grid(48), false);
    } else {
        check_errors();
        check_gauss();
        check_mesh(line_chain(50), true);
        check_mesh(star(40), true);
        check_mesh(quad_grid(23, 17), true);
        check_mesh(hex_grid(7), true);
        check_mesh(tet_grid(6), true);
        check_mesh(degenerate(), true);
    }
    if (fails) std::fprintf(stderr, "%d check(s) failed\n", fails);
    return fails ? 1 : 0;
}
