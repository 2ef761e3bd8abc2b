// Self-check of the C++ host layer's two-level Hydro::MacProjector, the way
// MacProjOp builds it with two active levels (icns_advection.cpp:88-104:
// MacProjector(Geom(0..1)), initProjector(lpinfo, beta), setDomainBC;
// :331 setUMAC with both levels' face velocities; :353 project).
//
// Level 0: 16x12x16 cells over [0,1]x[0,0.75]x[0,1], periodic in x and z,
// walls in y (MAC Neumann), four boxes.  Level 1: four boxes at ratio 2,
// reaching the periodic x sides and the high y wall.  Face velocities: a
// deterministic hash of the (wrapped) face index, zero on the walls.
// After project(): the composite divergence -- fine cells from the fine
// faces, uncovered coarse cells from the coarse faces (the coarse-fine ones
// hold the fine average) -- is zero to the solver tolerance, and
// getPhiLevels hands back both levels' phi.  Exit status 0 when both hold.
#include <cmath>
#include <cstdio>
#include <vector>

#include "ksgf_hydro.H"

using namespace ksgf::amr;

namespace {

double hashv(int d, int i, int j, int k)
{
    unsigned long long x = 0x9e3779b97f4a7c15ULL * (unsigned long long)(d + 1);
    x ^= (unsigned long long)(unsigned)i * 0xbf58476d1ce4e5b9ULL;
    x ^= (unsigned long long)(unsigned)j * 0x94d049bb133111ebULL;
    x ^= (unsigned long long)(unsigned)k * 0xd6e8feb86659fd93ULL;
    x ^= x >> 31;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    return (double)(x & 0xFFFFFFULL) / (double)0x7FFFFF - 1.0;
}

int wrap(int x, int n) { return ((x % n) + n) % n; }

// host copy of one face MultiFab (1 component, ngrow ng): per box, the FAB
// over the box grown by ng (one more face along d), Fortran order
struct HostFaces {
    std::vector<double> v;
    std::vector<long long> off;
    std::vector<Box> boxes;
    int d, ng;
    HostFaces(const std::vector<Box>& bx, int d_, int ng_) : boxes(bx), d(d_), ng(ng_)
    {
        long long acc = 0;
        for (const Box& b : boxes) {
            off.push_back(acc);
            acc += n(b, 0) * n(b, 1) * n(b, 2);
        }
        v.assign((size_t)acc, 0.0);
    }
    long long n(const Box& b, int e) const { return b.hi[e] - b.lo[e] + 1 + 2 * ng + (e == d); }
    double& at(int b, int i, int j, int k)
    {
        const Box& x = boxes[b];
        return v[(size_t)(off[b] + (i - x.lo[0] + ng) +
                          n(x, 0) * ((j - x.lo[1] + ng) + n(x, 1) * (long long)(k - x.lo[2] + ng)))];
    }
};

}  // namespace

int main()
{
    const int nc[3] = {16, 12, 16};
    Geometry gc, gf;
    for (int e = 0; e < 3; ++e) {
        gc.domain.lo[e] = 0;
        gc.domain.hi[e] = nc[e] - 1;
        gf.domain.lo[e] = 0;
        gf.domain.hi[e] = 2 * nc[e] - 1;
        gc.prob_lo[e] = gf.prob_lo[e] = 0.0;
        gc.is_periodic[e] = gf.is_periodic[e] = (e != 1);
    }
    gc.prob_hi = gf.prob_hi = {1.0, 0.75, 1.0};
    std::vector<Box> bc = {{{0, 0, 0}, {7, 11, 7}}, {{8, 0, 0}, {15, 11, 7}}, {{0, 0, 8}, {7, 11, 15}},
                           {{8, 0, 8}, {15, 11, 15}}};
    std::vector<Box> bf = {{{0, 4, 8}, {7, 11, 23}}, {{24, 8, 8}, {31, 19, 15}}, {{8, 12, 24}, {23, 23, 31}},
                           {{8, 4, 8}, {15, 11, 15}}};
    try {
        BoxArray bac(gc, bc), baf(gf, bf, /*covers_domain*/ false);
        const int ng = 1;
        std::vector<std::unique_ptr<MultiFab>> uc, uf;
        std::vector<HostFaces> hc, hf;
        for (int d = 0; d < 3; ++d) {
            uc.emplace_back(new MultiFab(bac, 1, ng, (IndexType)(KSGF_XFACE + d)));
            uf.emplace_back(new MultiFab(baf, 1, ng, (IndexType)(KSGF_XFACE + d)));
            for (int lev = 0; lev < 2; ++lev) {
                const std::vector<Box>& bx = lev ? bf : bc;
                const int m[3] = {(lev ? 2 : 1) * nc[0], (lev ? 2 : 1) * nc[1], (lev ? 2 : 1) * nc[2]};
                HostFaces h(bx, d, ng);
                for (int b = 0; b < (int)bx.size(); ++b)
                    for (int k = bx[b].lo[2] - ng; k <= bx[b].hi[2] + ng + (d == 2); ++k)
                        for (int j = bx[b].lo[1] - ng; j <= bx[b].hi[1] + ng + (d == 1); ++j)
                            for (int i = bx[b].lo[0] - ng; i <= bx[b].hi[0] + ng + (d == 0); ++i) {
                                const int wi = wrap(i, m[0]), wk = wrap(k, m[2]);
                                double val = hashv(d, wi, j, wk);
                                if (d == 1 && (j <= 0 || j >= m[1])) val = 0.0;  // walls
                                h.at(b, i, j, k) = val;
                            }
                MultiFab& mf = lev ? *uf[d] : *uc[d];
                if (hipMemcpy(mf.data(), h.v.data(), h.v.size() * sizeof(double), hipMemcpyHostToDevice) !=
                    hipSuccess)
                    throw std::runtime_error("upload");
                (lev ? hf : hc).push_back(std::move(h));
            }
        }
        // composite divergence of the host copies (fine cells; uncovered
        // coarse cells), max |.|
        auto comp_div = [&](bool fine_only) {
            double m = 0.0;
            const double hfd[3] = {1.0 / 32, 0.75 / 24, 1.0 / 32}, hcd[3] = {1.0 / 16, 0.75 / 12, 1.0 / 16};
            for (int b = 0; b < (int)bf.size(); ++b)
                for (int k = bf[b].lo[2]; k <= bf[b].hi[2]; ++k)
                    for (int j = bf[b].lo[1]; j <= bf[b].hi[1]; ++j)
                        for (int i = bf[b].lo[0]; i <= bf[b].hi[0]; ++i) {
                            const double dv = (hf[0].at(b, i + 1, j, k) - hf[0].at(b, i, j, k)) / hfd[0] +
                                              (hf[1].at(b, i, j + 1, k) - hf[1].at(b, i, j, k)) / hfd[1] +
                                              (hf[2].at(b, i, j, k + 1) - hf[2].at(b, i, j, k)) / hfd[2];
                            m = std::fmax(m, std::fabs(dv));
                        }
            if (fine_only) return m;
            for (int b = 0; b < (int)bc.size(); ++b)
                for (int k = bc[b].lo[2]; k <= bc[b].hi[2]; ++k)
                    for (int j = bc[b].lo[1]; j <= bc[b].hi[1]; ++j)
                        for (int i = bc[b].lo[0]; i <= bc[b].hi[0]; ++i) {
                            bool covered = false;
                            for (const Box& x : bf)
                                if (2 * i >= x.lo[0] && 2 * i <= x.hi[0] && 2 * j >= x.lo[1] && 2 * j <= x.hi[1] &&
                                    2 * k >= x.lo[2] && 2 * k <= x.hi[2])
                                    covered = true;
                            if (covered) continue;
                            const double dv = (hc[0].at(b, i + 1, j, k) - hc[0].at(b, i, j, k)) / hcd[0] +
                                              (hc[1].at(b, i, j + 1, k) - hc[1].at(b, i, j, k)) / hcd[1] +
                                              (hc[2].at(b, i, j, k + 1) - hc[2].at(b, i, j, k)) / hcd[2];
                            m = std::fmax(m, std::fabs(dv));
                        }
            return m;
        };
        const double div0 = comp_div(true);

        Hydro::MacProjector mp({gc, gf});
        mp.initProjector({&bac, &baf}, Hydro::LPInfo{}, 1.0);
        mp.setDomainBC({Hydro::LinOpBCType::Periodic, Hydro::LinOpBCType::Neumann, Hydro::LinOpBCType::Periodic},
                       {Hydro::LinOpBCType::Periodic, Hydro::LinOpBCType::Neumann, Hydro::LinOpBCType::Periodic});
        mp.setUMAC({{uc[0].get(), uc[1].get(), uc[2].get()}, {uf[0].get(), uf[1].get(), uf[2].get()}});
        mp.project(1e-11, 0.0);
        const ksgf_mlmg_info& info = mp.getMLMG().last;

        for (int d = 0; d < 3; ++d)
            for (int lev = 0; lev < 2; ++lev) {
                HostFaces& h = lev ? hf[d] : hc[d];
                MultiFab& mf = lev ? *uf[d] : *uc[d];
                if (hipMemcpy(h.v.data(), mf.data(), h.v.size() * sizeof(double), hipMemcpyDeviceToHost) !=
                    hipSuccess)
                    throw std::runtime_error("download");
            }
        const double div1 = comp_div(false);
        std::vector<double> phic((size_t)nc[0] * nc[1] * nc[2]);
        double* dphic = nullptr;
        if (hipMalloc(&dphic, phic.size() * sizeof(double)) != hipSuccess) throw std::runtime_error("hipMalloc");
        MultiFab phif(baf, 1, 1);
        mp.getPhiLevels(dphic, phif);
        if (hipMemcpy(phic.data(), dphic, phic.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("download phi");
        (void)hipFree(dphic);
        double pmax = 0.0;
        for (double p : phic) pmax = std::fmax(pmax, std::fabs(p));
        const bool ok = info.converged == 1 && div1 <= 1e-8 * div0 && pmax > 0.0 && std::isfinite(pmax);
        std::printf("levels check %s: %d composite iterations, max composite div %.3e (fine cells before: %.3e), "
                    "max |phi_0| %.3e\n",
                    ok ? "ok" : "FAILED", info.iters, div1, div0, pmax);
        return ok ? 0 : 1;
    } catch (const std::exception& e) {
        std::printf("levels check FAILED: %s\n", e.what());
        return 1;
    }
}
