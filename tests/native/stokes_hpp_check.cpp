// Compiles include/lfdem_amd_stokes.hpp against stand-ins for LF_DEM's
// vec3d (vec3d.h: x, y, z) and DBlock / ODBlock (MatrixBlocks.h:5-84: the
// packed columns col0..col5), and drives the StokesSolver protocol the way
// System::buildResistanceMatrix and System::computeVelocities* do
// (System.cpp:1456-1472, 1932-1965): reset with drag blocks, one column per
// particle with the rows of a column in a shuffled order, fill, RHS by
// particle, solve, solvingIsDone.
//
//   stokes_hpp_check nodevice   expects init to throw (no GPU visible)
//   stokes_hpp_check solve      solves a 40-particle chain system and checks
//                               R x = b on the host in FP64
// Prints "ok ..." and returns 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

struct vec3d { double x = 0, y = 0, z = 0; };
struct DBlock { double col0[5] = {}, col1[3] = {}, col2[1] = {}, col3[3] = {}, col4[2] = {}, col5[1] = {}; };
struct ODBlock { double col0[5] = {}, col1[3] = {}, col2[1] = {}, col3[5] = {}, col4[3] = {}, col5[1] = {}; };

#include "../../include/lfdem_amd_stokes.hpp"

int main(int argc, char** argv)
{
    const bool nodevice = argc > 1 && std::strcmp(argv[1], "nodevice") == 0;
    lfdem_amd::StokesSolver ss;
    try {
        ss.init(40, 40);
    } catch (const std::runtime_error& e) {
        if (nodevice) { std::printf("ok init threw: %s\n", e.what()); return 0; }
        std::printf("init failed: %s\n", e.what());
        return 1;
    }
    if (nodevice) { std::printf("init did not throw without a device\n"); return 1; }
    const int n = 40;
    // drag: diag(a, a, a, 4/3 a^3 ...) with a = 1
    std::vector<DBlock> drag(n);
    for (auto& d : drag) { d.col0[0] = d.col1[0] = d.col2[0] = 1.0; d.col3[0] = d.col4[0] = d.col5[0] = 4.0/3; }
    ss.resetResistanceMatrix(2*n, 0, 0, drag, true);
    // pair blocks: an isotropic coupling c between i and i+1, i+3 (rows added
    // in the order i+3, i+1 within column i): DBlock c*I on both, ODBlock -c*I
    std::vector<double> A(36*(size_t)n*n, 0.0);
    for (int i = 0; i < n; i++) for (int k = 0; k < 3; k++) {
        A[(6*i+k)*6*n+6*i+k] += 1.0; A[(6*i+3+k)*6*n+6*i+3+k] += 4.0/3;
    }
    for (int i = 0; i < n; i++) {
        ss.startNewColumn();
        for (int step : {3, 1}) {
            int j = i+step;
            if (j >= n) continue;
            double c = 0.3+0.01*i+0.05*step;
            DBlock d;
            d.col0[0] = d.col1[0] = d.col2[0] = c;
            d.col3[0] = d.col4[0] = d.col5[0] = 0.5*c;       // C00, C11, C22 of the packed DBlock
            ODBlock o;
            o.col0[0] = o.col1[0] = o.col2[0] = -c;
            o.col3[2] = o.col4[1] = o.col5[0] = -0.5*c;      // C00, C11, C22 of the packed ODBlock
            ss.addResistanceBlocks(i, j, std::make_pair(d, d), o);
            for (int k = 0; k < 6; k++) {
                double v = k < 3 ? c : 0.5*c;
                A[(6*i+k)*6*n+6*i+k] += v; A[(6*j+k)*6*n+6*j+k] += v;
                A[(6*j+k)*6*n+6*i+k] -= v; A[(6*i+k)*6*n+6*j+k] -= v;
            }
        }
    }
    ss.matrixFillingDone();
    ss.resetRHS();
    std::vector<double> b(6*n);
    for (int i = 0; i < n; i++) {
        vec3d f, t;
        f.x = std::sin(i+1.0); f.y = std::cos(2.0*i); f.z = 0.1*i;
        t.x = 0.3; t.y = -0.2*std::sin(i); t.z = 0.05*i;
        ss.addToRHSForce(i, f);
        ss.addToRHSTorque(i, t);
        b[6*i] = f.x; b[6*i+1] = f.y; b[6*i+2] = f.z; b[6*i+3] = t.x; b[6*i+4] = t.y; b[6*i+5] = t.z;
    }
    std::vector<vec3d> U(n), W(n);
    ss.solve(U, W);
    ss.solvingIsDone();
    std::vector<double> x(6*n);
    for (int i = 0; i < n; i++) {
        x[6*i] = U[i].x; x[6*i+1] = U[i].y; x[6*i+2] = U[i].z;
        x[6*i+3] = W[i].x; x[6*i+4] = W[i].y; x[6*i+5] = W[i].z;
    }
    double rr = 0, bb = 0;
    for (int r = 0; r < 6*n; r++) {
        double s = b[r];
        for (int c = 0; c < 6*n; c++) s -= A[(size_t)r*6*n+c]*x[c];
        rr += s*s; bb += b[r]*b[r];
    }
    const double rel = std::sqrt(rr/bb);
    // an undersized output throws as StokesSolver::solve does (StokesSolver.cpp:683-687)
    bool threw = false;
    std::vector<vec3d> small(n-1);
    try { ss.solve(small, W); } catch (const std::runtime_error&) { threw = true; }
    std::printf("%s residual %.3e, undersized output %s\n", rel <= 1e-11 && threw ? "ok" : "FAILED", rel,
                threw ? "threw" : "did not throw");
    return rel <= 1e-11 && threw ? 0 : 1;
}
