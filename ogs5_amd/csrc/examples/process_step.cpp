// process_step.cpp -- a host program written against the OGS-5 call shape
// (include/ogsx.hpp), i.e. what CRFProcess::Execute does for LIQUID_FLOW or
// HEAT_TRANSPORT (FEM/rf_pcs.cpp:4695-4853), running every step on the GPU.
//
// usage: process_step <input.bin> <output.bin>
// input (little endian): int64 n_nodes, n_elem, etype, process, dim, n_steps,
//   method, symmetric; f64 dt, theta, tol; f64 x[n], y[n], z[n];
//   int64 conn[n_elem*npe]; f64 mass[n_elem], cond[n_elem];
//   int64 n_bc; int64 bc_idx[n_bc]; f64 bc_val[n_bc];
//   int64 n_st; int64 st_idx[n_st]; f64 st_val[n_st]; f64 u0[n]
// output: f64 x[n_steps][n], then int32 iterations[n_steps]
#include <cstdio>
#include <vector>

#include "../../../include/ogsx.hpp"

template <class T>
static bool rd(FILE* f, T* p, size_t n)
{
    return std::fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int64_t hdr[8];
    double par[3];
    if (!rd(f, hdr, 8) || !rd(f, par, 3)) return 3;
    const int64_t n = hdr[0], ne = hdr[1], etype = hdr[2], process = hdr[3], dim = hdr[4];
    const int64_t steps = hdr[5], method = hdr[6], symmetric = hdr[7];
    const int npe = etype == OGSX_LINE2 ? 2 : etype == OGSX_HEX8 ? 8 : 4;
    std::vector<double> x(n), y(n), z(n), mass(ne), cond(ne), u0(n);
    std::vector<int64_t> conn(ne * npe);
    int64_t nbc = 0, nst = 0;
    bool ok = rd(f, x.data(), n) && rd(f, y.data(), n) && rd(f, z.data(), n) &&
              rd(f, conn.data(), conn.size()) && rd(f, mass.data(), ne) && rd(f, cond.data(), ne) &&
              rd(f, &nbc, 1);
    std::vector<int64_t> bci(nbc > 0 ? nbc : 0);
    std::vector<double> bcv(nbc > 0 ? nbc : 0);
    ok = ok && rd(f, bci.data(), bci.size()) && rd(f, bcv.data(), bcv.size()) && rd(f, &nst, 1);
    std::vector<int64_t> sti(nst > 0 ? nst : 0);
    std::vector<double> stv(nst > 0 ? nst : 0);
    ok = ok && rd(f, sti.data(), sti.size()) && rd(f, stv.data(), stv.size()) && rd(f, u0.data(), n);
    std::fclose(f);
    if (!ok) return 3;
    try {
        ogsx::SparsePattern table(n, (int)etype, conn);          // CreateEQS_LinearSolver
        ogsx::LinearEQS eqs(table, OGSX_FP64);                    // new Linear_EQS
        ogsx_elements el{};
        el.process = (int)process;
        el.dim = (int)dim;
        el.n_gauss = 3;
        el.cond_stride = 1;
        el.x = x.data();
        el.y = y.data();
        el.z = z.data();
        el.mass_coef = mass.data();
        el.cond = cond.data();
        el.tuf = 1.0;
        el.mu = 1.0;
        el.geo_area = 1.0;
        ogsx::FiniteElementStd fem(eqs, el);
        eqs.ConfigNumerics((int)method, 1, par[2], 100000, symmetric != 0);
        eqs.SetOldTimeLevel(u0);
        eqs.SetX(u0);
        FILE* o = std::fopen(argv[2], "wb");
        if (!o) return 2;
        std::vector<int> iters;
        for (int64_t k = 0; k < steps; k++) {
            eqs.Initialize();                      // eqs_new->Initialize()
            fem.Assembly(par[0], par[1]);          // GlobalAssembly
            if (nst) eqs.AddRHS(sti, stv);         // IncorporateSourceTerms
            for (int64_t i = 0; i < nbc; i++)      // IncorporateBoundaryConditions
                eqs.SetKnownX_i(bci[i], bcv[i]);
            eqs.Solver();                          // iter_lin = eqs_new->Solver()
            iters.push_back(eqs.Info().iterations);
            const std::vector<double> xs = eqs.GetX();
            std::fwrite(xs.data(), sizeof(double), xs.size(), o);
            eqs.CopyBack();                        // node values <- x
        }
        std::fwrite(iters.data(), sizeof(int), iters.size(), o);
        std::fclose(o);
    } catch (const ogsx::Error& e) {
        std::fprintf(stderr, "ogsx error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
