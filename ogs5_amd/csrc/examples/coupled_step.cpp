// coupled_step.cpp -- LIQUID_FLOW feeding HEAT_TRANSPORT through the C++ host
// classes of include/ogsx.hpp, in the order OGS-5 runs the two processes per
// time step: flow CRFProcess::Execute (FEM/rf_pcs.cpp:4695-4853), then
// CalIntegrationPointValue -> Cal_Velocity (rf_pcs.cpp:5915,
// fem_ele_std.cpp:6554) on the device, then the heat step with that
// velocity.  Flow: CG on the symmetric Dirichlet elimination; heat: BiCGStab
// on the reference's row-only elimination.
//
// usage: coupled_step <input.bin> <output.bin>
// input (little endian): int64 n_nodes, n_elem, etype, dim, n_steps;
//   f64 dt, theta_heat, tol, mu, rho_g, adv_coef;
//   f64 x[n], y[n], z[n]; int64 conn[n_elem*npe];
//   per process (flow, then heat): f64 mass[n_elem], cond[n_elem];
//     int64 n_bc; int64 bc_idx[n_bc]; f64 bc_val[n_bc]; f64 u0[n]
// output: f64 p[n], T[n] after the last step, f64 vel[n_elem*3*ngp] of the
//   last step, int32 iterations_flow[n_steps], iterations_heat[n_steps]
#include <cstdio>
#include <vector>

#include "../../../include/ogsx.hpp"

template <class T>
static bool rd(FILE* f, T* p, size_t n)
{
    return std::fread(p, sizeof(T), n, f) == n;
}

struct ProcessData {
    std::vector<double> mass, cond, bcv, u0;
    std::vector<int64_t> bci;
};

static bool read_process(FILE* f, int64_t n, int64_t ne, ProcessData& d)
{
    d.mass.resize(ne);
    d.cond.resize(ne);
    d.u0.resize(n);
    int64_t nbc = 0;
    if (!rd(f, d.mass.data(), ne) || !rd(f, d.cond.data(), ne) || !rd(f, &nbc, 1) || nbc < 0)
        return false;
    d.bci.resize(nbc);
    d.bcv.resize(nbc);
    return rd(f, d.bci.data(), nbc) && rd(f, d.bcv.data(), nbc) && rd(f, d.u0.data(), n);
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int64_t hdr[5];
    double par[6];
    if (!rd(f, hdr, 5) || !rd(f, par, 6)) return 3;
    const int64_t n = hdr[0], ne = hdr[1], etype = hdr[2], dim = hdr[3], steps = hdr[4];
    const double dt = par[0], theta_heat = par[1], tol = par[2];
    const int npe = etype == OGSX_LINE2 ? 2 : etype == OGSX_HEX8 ? 8 : 4;
    const int ngp = etype == OGSX_TET4 ? 5 : etype == OGSX_LINE2 ? 3 : etype == OGSX_QUAD4 ? 9 : 27;
    std::vector<double> x(n), y(n), z(n);
    std::vector<int64_t> conn(ne * npe);
    ProcessData flow, heat;
    bool ok = rd(f, x.data(), n) && rd(f, y.data(), n) && rd(f, z.data(), n) &&
              rd(f, conn.data(), conn.size()) && read_process(f, n, ne, flow) &&
              read_process(f, n, ne, heat);
    std::fclose(f);
    if (!ok) return 3;
    try {
        ogsx::SparsePattern table(n, (int)etype, conn);  // one mesh, one sparse table
        ogsx::LinearEQS flow_eqs(table, OGSX_FP64), heat_eqs(table, OGSX_FP64);
        auto elements = [&](int process, ProcessData& d) {
            ogsx_elements el{};
            el.process = process;
            el.dim = (int)dim;
            el.n_gauss = 3;
            el.cond_stride = 1;
            el.x = x.data();
            el.y = y.data();
            el.z = z.data();
            el.mass_coef = d.mass.data();
            el.cond = d.cond.data();
            el.tuf = 1.0;
            el.mu = par[3];
            el.geo_area = 1.0;
            el.adv_coef = process == OGSX_HEAT_TRANSPORT ? par[5] : 0.0;
            el.rho_g = process == OGSX_LIQUID_FLOW ? par[4] : 0.0;
            return el;
        };
        ogsx::FiniteElementStd flow_fem(flow_eqs, elements(OGSX_LIQUID_FLOW, flow));
        ogsx::FiniteElementStd heat_fem(heat_eqs, elements(OGSX_HEAT_TRANSPORT, heat));
        flow_eqs.ConfigNumerics(OGSX_CG, 1, tol, 100000, /*symmetric_dirichlet=*/true);
        heat_eqs.ConfigNumerics(OGSX_BICGSTAB, 1, tol, 100000, /*symmetric_dirichlet=*/false);
        for (auto* e : {&flow_eqs, &heat_eqs}) {
            const ProcessData& d = e == &flow_eqs ? flow : heat;
            e->SetOldTimeLevel(d.u0);
            e->SetX(d.u0);
        }
        std::vector<int> it_flow, it_heat;
        std::vector<double> vel((size_t)ne * 3 * ngp);
        for (int64_t k = 0; k < steps; k++) {
            flow_eqs.Initialize();
            flow_fem.Assembly(dt, 0.0);  // nls_relaxation 0 -> implicit Euler
            for (size_t i = 0; i < flow.bci.size(); i++) flow_eqs.SetKnownX_i(flow.bci[i], flow.bcv[i]);
            flow_eqs.Solver();
            it_flow.push_back(flow_eqs.Info().iterations);
            flow_eqs.CopyBack();
            flow_fem.Cal_Velocity(heat_fem, k + 1 == steps ? vel.data() : nullptr);
            heat_eqs.Initialize();
            heat_fem.Assembly(dt, theta_heat);
            for (size_t i = 0; i < heat.bci.size(); i++) heat_eqs.SetKnownX_i(heat.bci[i], heat.bcv[i]);
            heat_eqs.Solver();
            it_heat.push_back(heat_eqs.Info().iterations);
            heat_eqs.CopyBack();
        }
        FILE* o = std::fopen(argv[2], "wb");
        if (!o) return 2;
        const std::vector<double> p = flow_eqs.GetX(), T = heat_eqs.GetX();
        std::fwrite(p.data(), sizeof(double), p.size(), o);
        std::fwrite(T.data(), sizeof(double), T.size(), o);
        std::fwrite(vel.data(), sizeof(double), vel.size(), o);
        std::fwrite(it_flow.data(), sizeof(int), it_flow.size(), o);
        std::fwrite(it_heat.data(), sizeof(int), it_heat.size(), o);
        std::fclose(o);
    } catch (const ogsx::Error& e) {
        std::fprintf(stderr, "ogsx error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
