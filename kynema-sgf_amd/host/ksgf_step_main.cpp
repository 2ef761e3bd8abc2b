// Standalone C++ driver: Taylor-Green vortex (u = sin x cos y cos z,
// v = -cos x sin y cos z, w = 0 on the periodic unit cube, 2 pi / L as in
// TaylorGreenVortex.cpp:45-51) advanced by the Godunov advect + MAC-project
// step on the GPU.  Prints one JSON line per run.
//   ksgf_step [n] [max_grid_size] [steps] [mixed]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "incflo_godunov.H"

using namespace ksgf::amr;

int main(int argc, char** argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 64;
    const int mgs = argc > 2 ? std::atoi(argv[2]) : 32;
    const int steps = argc > 3 ? std::atoi(argv[3]) : 3;
    const bool mixed = argc > 4 && std::atoi(argv[4]) != 0;
    try {
        Geometry g;
        for (int d = 0; d < 3; ++d) {
            g.domain.lo[d] = 0;
            g.domain.hi[d] = n - 1;
            g.prob_lo[d] = 0.0;
            g.prob_hi[d] = 1.0;
            g.is_periodic[d] = 1;
        }
        std::vector<Box> boxes;
        for (int k = 0; k < n; k += mgs)
            for (int j = 0; j < n; j += mgs)
                for (int i = 0; i < n; i += mgs) {
                    Box b;
                    b.lo = {i, j, k};
                    b.hi = {std::min(i + mgs, n) - 1, std::min(j + mgs, n) - 1, std::min(k + mgs, n) - 1};
                    boxes.push_back(b);
                }
        std::vector<BCRec> bc(3);
        for (auto& b : bc)
            for (int d = 0; d < 3; ++d) b.lo[d] = b.hi[d] = KSGF_BC_INT_DIR;
        ksgf_mlmg_options o;
        ksgf_mlmg_default_options(&o);
        o.precision = mixed ? KSGF_MG_MIXED : KSGF_MG_FP64;
        const auto P = Hydro::LinOpBCType::Periodic;
        GodunovPredictor step(g, boxes, bc, std::vector<double>(18, 0.0), {P, P, P}, {P, P, P}, 1.0, 0.5, 1e-11,
                              2.220446049250313e-14, o);
        const long long N = (long long)n * n * n;
        std::vector<double> vel(3 * N), rho(N, 1.0), gp(3 * N, 0.0);
        const double tp = 2.0 * M_PI, h = 1.0 / n;
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double x = (i + 0.5) * h, y = (j + 0.5) * h, z = (k + 0.5) * h;
                    long long c = i + (long long)n * (j + (long long)n * k);
                    vel[c] = std::sin(tp * x) * std::cos(tp * y) * std::cos(tp * z);
                    vel[N + c] = -std::cos(tp * x) * std::sin(tp * y) * std::cos(tp * z);
                    vel[2 * N + c] = 0.0;
                }
        step.setState(vel.data(), rho.data(), gp.data());
        (void)hipDeviceSynchronize();
        auto t0 = std::chrono::steady_clock::now();
        StepStats last;
        int iters = 0;
        for (int s = 0; s < steps; ++s) {
            last = step.advance();
            iters += last.mlmg.iters;
        }
        (void)hipDeviceSynchronize();
        double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        step.getVelocity(vel.data());
        double ke = 0.0;
        for (long long c = 0; c < 3 * N; ++c) ke += 0.5 * vel[c] * vel[c];
        std::printf("{\"case\": \"tgv%d\", \"boxes\": %zu, \"steps\": %d, \"ms_per_step\": %.4f, "
                    "\"cell_updates_per_s\": %.6e, \"mlmg_iters\": %d, \"dt\": %.10e, \"ke\": %.12e, "
                    "\"precision\": \"%s\"}\n",
                    n, boxes.size(), steps, 1e3 * sec / steps, (double)N * steps / sec, iters, last.dt, ke * h * h * h,
                    mixed ? "mixed" : "fp64");
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ksgf_step: %s\n", e.what());
        return 1;
    }
}
