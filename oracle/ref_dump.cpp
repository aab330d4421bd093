// ref_dump.cpp -- driver linked against the reference's own FEM/mathlib.cpp
// and FEM/matrix_class.cpp (compiled from /root/reference by oracle/Makefile
// into oracle/_ref/).  It prints, as exact hex floats, the reference's Gauss
// tables (mathlib.cpp:1112,1174,1285), its shape functions and local
// gradients at every Gauss point of line2/quad4/hex8/tet4 for n_gauss 1..4
// (mathlib.cpp:1661-2302), and dense Matrix::multi results
// (matrix_class.cpp:288).  tests/test_oracle_ref.py compares the oracle's
// restatement with these values bit for bit.  Test infrastructure only.
#include <cstdio>

#include "matrix_class.h"

double MXPGaussPkt(long grd, long pkt);
double MXPGaussFkt(long grd, long pkt);
void SamplePointTet5(const int nsample, double* SPoint);
void ShapeFunctionLine(double* N1, const double* u);
void GradShapeFunctionLine(double* dN1, const double* u);
void ShapeFunctionQuad(double* N4, const double* u);
void GradShapeFunctionQuad(double* dN4, const double* u);
void ShapeFunctionHex(double* N8, const double* x);
void GradShapeFunctionHex(double* dN8, const double* x);
void ShapeFunctionTet(double* Nt4, const double* x);
void GradShapeFunctionTet(double* dNt4, const double* dummy);

static void unit_point(int et, int ng, int gp, double* u)
{
    u[0] = u[1] = u[2] = 0.0;
    if (et == 1) {
        u[0] = MXPGaussPkt(ng, gp);
    } else if (et == 2) {
        u[0] = MXPGaussPkt(ng, gp / ng);
        u[1] = MXPGaussPkt(ng, gp % ng);
    } else if (et == 3) {
        int r = gp / (ng * ng), s = gp % (ng * ng), t = s % ng;
        s /= ng;
        u[0] = MXPGaussPkt(ng, r);
        u[1] = MXPGaussPkt(ng, s);
        u[2] = MXPGaussPkt(ng, t);
    } else if (et == 5) {
        double sp[4];
        SamplePointTet5(gp, sp);
        u[0] = sp[0];
        u[1] = sp[1];
        u[2] = sp[2];
    }
}

int main()
{
    for (long grd = 1; grd <= 4; grd++)
        for (long pkt = 0; pkt < grd; pkt++)
            std::printf("gauss %ld %ld %a %a\n", grd, pkt, MXPGaussPkt(grd, pkt),
                        MXPGaussFkt(grd, pkt));
    for (int gp = 0; gp < 5; gp++) {
        double sp[4];
        SamplePointTet5(gp, sp);
        std::printf("tet5 %d %a %a %a %a\n", gp, sp[0], sp[1], sp[2], sp[3]);
    }
    const int types[4] = {1, 2, 3, 5};
    const int npe[4] = {2, 4, 8, 4};
    const int edim[4] = {1, 2, 3, 3};
    for (int t = 0; t < 4; t++) {
        const int et = types[t];
        for (int ng = 1; ng <= 4; ng++) {
            int ngp = (et == 1) ? ng : (et == 2) ? ng * ng : (et == 3) ? ng * ng * ng : 5;
            if (et == 5 && ng != 3) continue;
            for (int gp = 0; gp < ngp; gp++) {
                double u[3], N[8], dN[24];
                unit_point(et, ng, gp, u);
                if (et == 1) { ShapeFunctionLine(N, u); GradShapeFunctionLine(dN, u); }
                if (et == 2) { ShapeFunctionQuad(N, u); GradShapeFunctionQuad(dN, u); }
                if (et == 3) { ShapeFunctionHex(N, u); GradShapeFunctionHex(dN, u); }
                if (et == 5) { ShapeFunctionTet(N, u); GradShapeFunctionTet(dN, u); }
                std::printf("shape %d %d %d", et, ng, gp);
                for (int a = 0; a < npe[t]; a++) std::printf(" %a", N[a]);
                for (int a = 0; a < npe[t] * edim[t]; a++) std::printf(" %a", dN[a]);
                std::printf("\n");
            }
        }
    }
    // Dense Matrix::multi on a few row-major matrices with non-trivial values
    // (the element RHS product Aux1 * u_n of AssembleParabolicEquation).
    for (int c = 0; c < 3; c++) {
        const int n = (c == 0) ? 3 : (c == 1) ? 4 : 8;
        Math_Group::Matrix m(n, n);
        double x[8], b[8];
        for (int i = 0; i < n; i++) {
            x[i] = 0.1 * (i + 1) + 1.0 / (3.0 + i + c);
            b[i] = 0.25 * i;
            for (int j = 0; j < n; j++) m(i, j) = 1.0 / (1.0 + i + 2.0 * j + c) - 0.3 * (i == j);
        }
        m.multi(x, b, -0.7);
        std::printf("multi %d", n);
        for (int i = 0; i < n; i++) std::printf(" %a", b[i]);
        std::printf("\n");
    }
    return 0;
}
