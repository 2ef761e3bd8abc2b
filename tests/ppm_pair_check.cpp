// CPU check of the PPM reconstruction the Godunov kernels run
// (kynema-sgf_amd/csrc/ksgf_ppm.h, compiled as host C++): the shared-slope
// trace pair of one face must equal the two single-cell traces bit for bit,
// for every limiter, boundary position and advecting-velocity sign, and the
// dual-number evaluation must carry the same values.
#include <cstdio>
#include <cstring>
#include <random>

#include "../kynema-sgf_amd/csrc/ksgf_ppm.h"

using namespace ksgf;

static int g_bad = 0, g_cases = 0;

static bool same(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

int main()
{
    std::mt19937 rng(7);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const int lims[4] = {KSGF_PPM_VANLEER, KSGF_PPM_NOLIMITER, KSGF_PPM_MINMOD, KSGF_PPM_UPWIND};
    const int bcs[4] = {KSGF_BC_INT_DIR, KSGF_BC_EXT_DIR, KSGF_BC_HOEXTRAP, KSGF_BC_FOEXTRAP};
    const int dlo = 0, dhi = 15;
    for (int rep = 0; rep < 4000; ++rep) {
        double s[6];
        for (double& v : s) v = (rep % 7 == 0) ? std::floor(3.0 * U(rng)) : U(rng);  // ties and flat data too
        const double u[4] = {U(rng), U(rng), U(rng), (rep % 5 == 0) ? 0.0 : U(rng)};
        const double dtdx = 0.5 + 0.4 * U(rng);
        for (int lim : lims)
            for (int bl : bcs)
                for (int bh : bcs)
                    for (int x = dlo; x <= dhi + 1; ++x) {
                        double a, b, dmy;
                        ppm_trace(s[0], s[1], s[2], s[3], s[4], lim, bl, bh, x - 1, dlo, dhi, u[0], u[1], dtdx, dmy, a);
                        ppm_trace(s[1], s[2], s[3], s[4], s[5], lim, bl, bh, x, dlo, dhi, u[2], u[3], dtdx, b, dmy);
                        double p, q;
                        ppm_trace_pair(s, lim, bl, bh, x, dlo, dhi, u[0], u[1], u[2], u[3], dtdx, p, q);
                        ++g_cases;
                        if (!same(a, p) || !same(b, q)) {
                            if (g_bad < 5)
                                std::printf("lim %d bc %d/%d x %d: %.17g %.17g vs %.17g %.17g\n", lim, bl, bh, x, a, b, p, q);
                            ++g_bad;
                        }
                        // dual evaluation of the single trace: same value
                        Dual<2> sd[5], ud[2], Id, Ipd;
                        for (int m = 0; m < 5; ++m) sd[m] = dual_seed<2>(s[m], m % 2);
                        ud[0] = dual_const<2>(u[0]);
                        ud[1] = dual_const<2>(u[1]);
                        ppm_trace(sd[0], sd[1], sd[2], sd[3], sd[4], lim, bl, bh, x - 1, dlo, dhi, ud[0], ud[1], dtdx,
                                  Id, Ipd);
                        if (!same(Ipd.v, a)) {
                            if (g_bad < 5) std::printf("dual value differs: %.17g vs %.17g\n", Ipd.v, a);
                            ++g_bad;
                        }
                    }
    }
    if (g_bad) {
        std::printf("FAILED %d of %d\n", g_bad, g_cases);
        return 1;
    }
    std::printf("ok %d\n", g_cases);
    return 0;
}
