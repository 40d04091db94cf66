// 3-D Taylor-Green vortex driver with the reference's main/main_setup shape
// (main.cpp:9-25, setup.cpp:94-148) on the GPU path.
//   lbm_tgv [N] [nstep] [tout] [Re] [fused|strict]
// Prints "step,KE,enstrophy" like OutputKeEns (output.cpp:272-274) and the
// wall time of the step loop.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "lbm_host.hpp"

int main(int argc, char** argv)
{
    const int N = argc > 1 ? std::atoi(argv[1]) : 64;
    const int nstep = argc > 2 ? std::atoi(argv[2]) : 100;
    const int tout = argc > 3 ? std::atoi(argv[3]) : 10;
    const double Re = argc > 4 ? std::atof(argv[4]) : 1600.0;
    const int mode = (argc > 5 && !std::strcmp(argv[5], "strict")) ? LBM_MODE_STRICT : LBM_MODE_FUSED;
    // 64^3 configuration: R = 1/3, T0 = 1, p0 = 1/3, gamma = 1.4, Pr = 0.71, Ma = 0.1
    const double R = 1.0 / 3.0, T0 = 1.0, p0 = 1.0 / 3.0, gamma = 1.4, Pr = 0.71;
    const double U0 = 0.1 * std::sqrt(gamma * R * T0);
    const double nu = U0 * (N / (2.0 * M_PI)) / Re;
    try {
        LBM lb(N, N, N, nu, mode);
        lb.set_gasconst(R);
        lb.set_gamma(gamma);
        lb.set_prtl(Pr);
        const int Nx = lb.get_Nx(), Ny = lb.get_Ny(), Nz = lb.get_Nz();
        const double a = N, b = N, c = N;
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < Nz; ++k) {
                    const size_t x = lb.idx(i, j, k);
                    if (i == 0 || i == Nx - 1 || j == 0 || j == Ny - 1 || k == 0 || k == Nz - 1) {
                        lb.type[x] = 2;  // TYPE_P
                        continue;
                    }
                    lb.field(LBM_FIELD_U)[x] = U0 * std::sin(2.0 * M_PI / a * i) * std::cos(2.0 * M_PI / b * j) *
                                               std::cos(2.0 * M_PI / c * k);
                    lb.field(LBM_FIELD_V)[x] = -U0 * std::cos(2.0 * M_PI / a * i) * std::sin(2.0 * M_PI / b * j) *
                                               std::cos(2.0 * M_PI / c * k);
                    lb.field(LBM_FIELD_W)[x] = 0.0;
                    lb.field(LBM_FIELD_P)[x] = p0 + 1.0 * U0 * U0 / 16.0 *
                                                        (std::cos(4.0 * M_PI / a * i) + std::cos(4.0 * M_PI / b * j)) *
                                                        (std::cos(4.0 * M_PI / c * k) + 2.0);
                    lb.field(LBM_FIELD_T)[x] = T0;
                }
        auto out = [](int step, LBM& l) {
            lbm_diag d = l.diagnostics();
            std::printf("%d,%.15g,%.15g\n", step, d.kinetic, d.enstrophy);
        };
        auto t0 = std::chrono::steady_clock::now();
        lb.run(nstep, tout, out);
        lbm_sync(lb.context());
        double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("# %d^3, %d steps, %.3f s incl. outputs, %.1f MLUPS\n", N, nstep, sec,
                    (double)N * N * N * nstep / sec / 1e6);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
