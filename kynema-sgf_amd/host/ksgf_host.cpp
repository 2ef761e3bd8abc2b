// C entry points of the C++ host layer (libksgf_host.so): the Godunov
// advect + MAC-project step driver for callers that are not C++ (the
// Python bench and tests).  Exceptions never cross this boundary.
#include <cstring>
#include <string>

#include "incflo_godunov.H"
#include "incflo_two_level.H"
#include "ksgf_host.h"

namespace {
thread_local std::string g_err;
}

using ksgf::amr::Box;
using ksgf::amr::Geometry;
using ksgf::amr::GodunovPredictor;
using ksgf::amr::TwoLevelPredictor;

extern "C" {

const char* ksgf_host_last_error(void) { return g_err.c_str(); }

int ksgf_host_stepper_create(const ksgf_geometry* geom, const int* boxes, int nbox, const ksgf_bcrec* vel_bc,
                             const double* vel_bcval, const int* lobc, const int* hibc, double rho0, double cfl,
                             double mg_rtol, double mg_atol, const ksgf_mlmg_options* opts, ksgf_host_stepper** out)
{
    try {
        if (!geom || !boxes || nbox <= 0 || !vel_bc || !vel_bcval || !lobc || !hibc || !out)
            throw std::runtime_error("null argument");
        Geometry g;
        for (int d = 0; d < 3; ++d) {
            g.domain.lo[d] = geom->domain_lo[d];
            g.domain.hi[d] = geom->domain_hi[d];
            g.prob_lo[d] = geom->prob_lo[d];
            g.prob_hi[d] = geom->prob_hi[d];
            g.is_periodic[d] = geom->is_periodic[d];
        }
        std::vector<Box> bx(nbox);
        for (int b = 0; b < nbox; ++b)
            for (int d = 0; d < 3; ++d) {
                bx[b].lo[d] = boxes[6 * b + d];
                bx[b].hi[d] = boxes[6 * b + 3 + d];
            }
        std::vector<ksgf::amr::BCRec> bc(vel_bc, vel_bc + 3);
        std::vector<double> bv(vel_bcval, vel_bcval + 18);
        std::array<Hydro::LinOpBCType, 3> lo, hi;
        for (int d = 0; d < 3; ++d) {
            lo[d] = (Hydro::LinOpBCType)lobc[d];
            hi[d] = (Hydro::LinOpBCType)hibc[d];
        }
        ksgf_mlmg_options o;
        if (opts)
            o = *opts;
        else
            ksgf_mlmg_default_options(&o);
        auto* s = new GodunovPredictor(g, bx, bc, bv, lo, hi, rho0, cfl, mg_rtol, mg_atol, o);
        *out = reinterpret_cast<ksgf_host_stepper*>(s);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_create: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_set_stream(ksgf_host_stepper* s, ksgf_stream_t stream)
{
    if (!s) return KSGF_ERR_INVALID;
    reinterpret_cast<GodunovPredictor*>(s)->setStream((hipStream_t)stream);
    return KSGF_OK;
}

int ksgf_host_stepper_set_state(ksgf_host_stepper* s, const double* vel, const double* rho, const double* gp)
{
    try {
        if (!s || !vel || !rho || !gp) throw std::runtime_error("null argument");
        reinterpret_cast<GodunovPredictor*>(s)->setState(vel, rho, gp);
        return KSGF_OK;
    } catch (const ksgf::amr::UnsupportedError& e) {
        g_err = std::string("ksgf_host_stepper_set_state: ") + e.what();
        return KSGF_ERR_UNSUPPORTED;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_set_state: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_step(ksgf_host_stepper* s, double fixed_dt, double* dt_used, ksgf_mlmg_info* info,
                           double* divmax)
{
    try {
        if (!s) throw std::runtime_error("null stepper");
        auto* g = reinterpret_cast<GodunovPredictor*>(s);
        g->setCheckDivergence(divmax != nullptr);
        auto st = g->advance(fixed_dt);
        if (dt_used) *dt_used = st.dt;
        if (info) *info = st.mlmg;
        if (divmax) *divmax = st.divmax;
        return st.mlmg.converged ? KSGF_OK : KSGF_ERR_NOT_CONVERGED;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_step: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_enable_graph(ksgf_host_stepper* s, int on)
{
    try {
        if (!s) throw std::runtime_error("null stepper");
        reinterpret_cast<GodunovPredictor*>(s)->enableGraph(on != 0);
        return KSGF_OK;
    } catch (const ksgf::amr::UnsupportedError& e) {
        g_err = std::string("ksgf_host_stepper_enable_graph: ") + e.what();
        return KSGF_ERR_UNSUPPORTED;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_enable_graph: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_get_velocity(ksgf_host_stepper* s, double* vel)
{
    try {
        if (!s || !vel) throw std::runtime_error("null argument");
        reinterpret_cast<GodunovPredictor*>(s)->getVelocity(vel);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_get_velocity: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_get_phi(ksgf_host_stepper* s, double* phi)
{
    try {
        if (!s || !phi) throw std::runtime_error("null argument");
        reinterpret_cast<GodunovPredictor*>(s)->getPhi(phi);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_get_phi: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_enable_scalar(ksgf_host_stepper* s, const ksgf_bcrec* bc, const double* bcval)
{
    try {
        if (!s || !bc || !bcval) throw std::runtime_error("null argument");
        reinterpret_cast<GodunovPredictor*>(s)->enableScalar(*bc, std::vector<double>(bcval, bcval + 6));
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_enable_scalar: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_set_scalar(ksgf_host_stepper* s, const double* scal)
{
    try {
        if (!s || !scal) throw std::runtime_error("null argument");
        reinterpret_cast<GodunovPredictor*>(s)->setScalar(scal);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_set_scalar: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_get_scalar(ksgf_host_stepper* s, double* scal)
{
    try {
        if (!s || !scal) throw std::runtime_error("null argument");
        reinterpret_cast<GodunovPredictor*>(s)->getScalar(scal);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_get_scalar: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_write_plotfile(ksgf_host_stepper* s, const char* dir, double time, int step)
{
    try {
        if (!s || !dir) throw std::runtime_error("null argument");
        reinterpret_cast<GodunovPredictor*>(s)->writePlotfile(dir, time, step);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_write_plotfile: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_enable_nodal_projection(ksgf_host_stepper* s, double rtol, double atol,
                                              const ksgf_mlmg_options* opts)
{
    try {
        if (!s) throw std::runtime_error("null stepper");
        ksgf_mlmg_options o;
        if (opts)
            o = *opts;
        else
            ksgf_mlmg_default_options(&o);
        reinterpret_cast<GodunovPredictor*>(s)->enableNodalProjection(rtol, atol, o);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_enable_nodal_projection: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_nodal_info(ksgf_host_stepper* s, ksgf_mlmg_info* info)
{
    if (!s || !info) {
        g_err = "ksgf_host_stepper_nodal_info: null argument";
        return KSGF_ERR_INVALID;
    }
    *info = reinterpret_cast<GodunovPredictor*>(s)->lastNodalInfo();
    return KSGF_OK;
}

int ksgf_host_stepper_get_gradp(ksgf_host_stepper* s, double* gp)
{
    try {
        if (!s || !gp) throw std::runtime_error("null argument");
        reinterpret_cast<GodunovPredictor*>(s)->getGradP(gp);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_get_gradp: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_get_nodal_phi(ksgf_host_stepper* s, double* phi)
{
    try {
        if (!s || !phi) throw std::runtime_error("null argument");
        reinterpret_cast<GodunovPredictor*>(s)->getNodalPhi(phi);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_get_nodal_phi: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_enable_diffusion(ksgf_host_stepper* s, double mu, int type, double rtol, double atol,
                                       const ksgf_mlmg_options* opts, int tensor)
{
    try {
        if (!s) throw std::runtime_error("null stepper");
        ksgf_mlmg_options o;
        if (opts)
            o = *opts;
        else
            ksgf_mlmg_default_options(&o);
        reinterpret_cast<GodunovPredictor*>(s)->enableDiffusion(mu, type, rtol, atol, o, tensor != 0);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_stepper_enable_diffusion: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_stepper_diffusion_info(ksgf_host_stepper* s, ksgf_mlmg_info* info3)
{
    if (!s || !info3) {
        g_err = "ksgf_host_stepper_diffusion_info: null argument";
        return KSGF_ERR_INVALID;
    }
    const auto& st = reinterpret_cast<GodunovPredictor*>(s)->lastStats();
    for (int c = 0; c < 3; ++c) info3[c] = st.diffusion[c];
    return KSGF_OK;
}

int ksgf_host_stepper_destroy(ksgf_host_stepper* s)
{
    delete reinterpret_cast<GodunovPredictor*>(s);
    return KSGF_OK;
}

// ---- two levels -----------------------------------------------------------
int ksgf_host_two_level_create(const ksgf_geometry* geom0, const int* boxes0, int nbox0, const int* boxes1,
                               int nbox1, const ksgf_bcrec* vel_bc, const double* vel_bcval, const int* lobc,
                               const int* hibc, double rho0, double cfl, double mg_rtol, double mg_atol,
                               const ksgf_mlmg_options* opts, int limiter, ksgf_host_two_level** out)
{
    try {
        if (!geom0 || !boxes0 || nbox0 <= 0 || !boxes1 || nbox1 <= 0 || !vel_bc || !vel_bcval || !lobc || !hibc ||
            !out)
            throw std::runtime_error("null argument");
        Geometry g;
        for (int d = 0; d < 3; ++d) {
            g.domain.lo[d] = geom0->domain_lo[d];
            g.domain.hi[d] = geom0->domain_hi[d];
            g.prob_lo[d] = geom0->prob_lo[d];
            g.prob_hi[d] = geom0->prob_hi[d];
            g.is_periodic[d] = geom0->is_periodic[d];
        }
        auto boxes = [](const int* flat, int n) {
            std::vector<Box> bx(n);
            for (int b = 0; b < n; ++b)
                for (int d = 0; d < 3; ++d) {
                    bx[b].lo[d] = flat[6 * b + d];
                    bx[b].hi[d] = flat[6 * b + 3 + d];
                }
            return bx;
        };
        std::vector<ksgf::amr::BCRec> bc(vel_bc, vel_bc + 3);
        std::vector<double> bv(vel_bcval, vel_bcval + 18);
        std::array<Hydro::LinOpBCType, 3> lo, hi;
        for (int d = 0; d < 3; ++d) {
            lo[d] = (Hydro::LinOpBCType)lobc[d];
            hi[d] = (Hydro::LinOpBCType)hibc[d];
        }
        ksgf_mlmg_options o;
        if (opts)
            o = *opts;
        else
            ksgf_mlmg_default_options(&o);
        auto* s = new TwoLevelPredictor(g, boxes(boxes0, nbox0), boxes(boxes1, nbox1), bc, bv, lo, hi, rho0, cfl,
                                        mg_rtol, mg_atol, o, limiter);
        *out = reinterpret_cast<ksgf_host_two_level*>(s);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_two_level_create: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_two_level_set_state(ksgf_host_two_level* s, const double* vel0, const double* vel1, const double* gp0,
                                  const double* gp1, int on_device)
{
    try {
        if (!s || !vel0 || !vel1) throw std::runtime_error("null argument");
        reinterpret_cast<TwoLevelPredictor*>(s)->setState(vel0, vel1, gp0, gp1, on_device != 0);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_two_level_set_state: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_two_level_step(ksgf_host_two_level* s, double fixed_dt, double* dt_used, ksgf_mlmg_info* info)
{
    try {
        if (!s) throw std::runtime_error("null stepper");
        auto st = reinterpret_cast<TwoLevelPredictor*>(s)->advance(fixed_dt);
        if (dt_used) *dt_used = st.dt;
        if (info) *info = st.mlmg;
        return st.mlmg.converged ? KSGF_OK : KSGF_ERR_NOT_CONVERGED;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_two_level_step: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_two_level_get_velocity(ksgf_host_two_level* s, double* vel0, double* vel1)
{
    try {
        if (!s || !vel0 || !vel1) throw std::runtime_error("null argument");
        reinterpret_cast<TwoLevelPredictor*>(s)->getVelocity(vel0, vel1);
        return KSGF_OK;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_two_level_get_velocity: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}

int ksgf_host_two_level_destroy(ksgf_host_two_level* s)
{
    delete reinterpret_cast<TwoLevelPredictor*>(s);
    return KSGF_OK;
}

}  // extern "C"

extern "C" int ksgf_host_mac_project_density(const ksgf_geometry* geom, const int* boxes, int nbox, const int* lobc,
                                             const int* hibc, const double* rho, double factor, double* umac,
                                             double* vmac, double* wmac, double rtol, double atol,
                                             const ksgf_mlmg_options* opts, ksgf_mlmg_info* info, double* phi)
{
    using ksgf::amr::BoxArray;
    using ksgf::amr::IndexType;
    using ksgf::amr::MultiFab;
    try {
        if (!geom || !boxes || nbox <= 0 || !lobc || !hibc || !rho || !umac || !vmac || !wmac)
            throw std::runtime_error("null argument");
        Geometry g;
        for (int d = 0; d < 3; ++d) {
            g.domain.lo[d] = geom->domain_lo[d];
            g.domain.hi[d] = geom->domain_hi[d];
            g.prob_lo[d] = geom->prob_lo[d];
            g.prob_hi[d] = geom->prob_hi[d];
            g.is_periodic[d] = geom->is_periodic[d];
        }
        std::vector<Box> bx(nbox);
        for (int b = 0; b < nbox; ++b)
            for (int d = 0; d < 3; ++d) {
                bx[b].lo[d] = boxes[6 * b + d];
                bx[b].hi[d] = boxes[6 * b + 3 + d];
            }
        BoxArray ba(g, bx);
        // density with one ghost layer: periodic images, first-order
        // extrapolation at walls
        MultiFab rmf(ba, 1, 1);
        rmf.copyFromDomain(rho);
        rmf.FillBoundary(nullptr);
        std::vector<ksgf::amr::BCRec> rbc(1);
        for (int d = 0; d < 3; ++d) {
            rbc[0].lo[d] = g.isPeriodic(d) ? KSGF_BC_INT_DIR : KSGF_BC_FOEXTRAP;
            rbc[0].hi[d] = g.isPeriodic(d) ? KSGF_BC_INT_DIR : KSGF_BC_FOEXTRAP;
        }
        rmf.FillPhysBC(rbc, std::vector<double>(6, 0.0), nullptr);
        MultiFab bxf(ba, 1, 0, IndexType::XFace), byf(ba, 1, 0, IndexType::YFace), bzf(ba, 1, 0, IndexType::ZFace);
        ksgf::amr::face_beta_from_density({&bxf, &byf, &bzf}, rmf, factor, nullptr);
        MultiFab u(ba, 1, 1, IndexType::XFace), v(ba, 1, 1, IndexType::YFace), w(ba, 1, 1, IndexType::ZFace);
        u.copyFromDomain(umac);
        v.copyFromDomain(vmac);
        w.copyFromDomain(wmac);
        Hydro::MacProjector mp({g});
        Hydro::LPInfo lp;
        ksgf_mlmg_options o;
        if (opts)
            o = *opts;
        else
            ksgf_mlmg_default_options(&o);
        lp.do_semicoarsening = o.do_semicoarsening != 0;
        lp.max_coarsening_level = o.max_coarsening_level;
        mp.getMLMG().opts = o;
        mp.initProjector({&ba}, lp, {&bxf, &byf, &bzf});
        std::array<Hydro::LinOpBCType, 3> lo, hi;
        for (int d = 0; d < 3; ++d) {
            lo[d] = (Hydro::LinOpBCType)lobc[d];
            hi[d] = (Hydro::LinOpBCType)hibc[d];
        }
        mp.setDomainBC(lo, hi);
        mp.setUMAC({{&u, &v, &w}});
        mp.project(rtol, atol, nullptr);
        if (info) *info = mp.getMLMG().last;
        if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("synchronize");
        u.copyToDomain(umac);
        v.copyToDomain(vmac);
        w.copyToDomain(wmac);
        if (phi) {
            const size_t n = (size_t)g.domain.numPts();
            double* d = nullptr;
            if (hipMalloc((void**)&d, n * sizeof(double)) != hipSuccess) throw std::runtime_error("hipMalloc");
            mp.getPhi(d, nullptr);
            const hipError_t e = hipMemcpy(phi, d, n * sizeof(double), hipMemcpyDeviceToHost);
            (void)hipFree(d);
            if (e != hipSuccess) throw std::runtime_error("copy phi");
        }
        return mp.getMLMG().last.converged ? KSGF_OK : KSGF_ERR_NOT_CONVERGED;
    } catch (const ksgf::amr::UnsupportedError& e) {
        g_err = std::string("ksgf_host_mac_project_density: ") + e.what();
        return KSGF_ERR_UNSUPPORTED;
    } catch (const std::exception& e) {
        g_err = std::string("ksgf_host_mac_project_density: ") + e.what();
        return KSGF_ERR_INVALID;
    }
}
