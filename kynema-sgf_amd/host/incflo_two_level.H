// The Godunov branch of incflo::ApplyPredictor over two AMR levels (level 1
// at ratio 2, one dt for both levels, no subcycling, constant density), on
// the C++ Hydro layer: every operator loops over the levels as Kynema's
// PDE operators do when two levels are active.
//
//   compute_dt over both levels (convective and forcing CFL, growth cap)
//                                      incflo_compute_dt.cpp:33-220, SimTime.cpp:162-205
//   src = -gp/rho per level, ghosts: FillBoundary + physical BCs on level 0,
//     FillPatchTwoLevels on level 1    icns_ops.H:98-111, FieldFillPatchOps.H:272-331
//   ExtrapVelToFaces + set_inflow per level   icns_advection.H:210-226
//   MacProjector over both levels (composite solve)   icns_advection.cpp:88-95,331,353
//   MAC face ghosts: FillBoundary + physical BCs on level 0, the face
//     fill-patch (fillpatch_sibling_fields) on level 1   icns_advection.H:231-243
//   q = rho u, fq = rho f, ComputeFluxesOnBoxFromState per level   icns_advection.H:281-374
//   the fine fluxes averaged onto the coarse faces (average_down_faces)
//                                      icns_advection.H:401-407
//   ComputeDivergence + predictor_rhs per level   icns_advection.H:418-422, CompRHSOps.H:93-117
//   average_down of the new velocity, velocity ghosts as at the start
//
// It makes the calls kynema-sgf_amd/levels.py makes, in the same order, so
// the two drivers give the same bits.
#pragma once

#include <array>
#include <cmath>
#include <limits>
#include <memory>
#include <stdexcept>
#include <vector>

#include "ksgf_hydro.H"

namespace ksgf::amr {

class TwoLevelPredictor {
public:
    struct StepStats {
        double dt = 0.0;
        ksgf_mlmg_info mlmg{};
    };

    // geom0 / boxes0: level 0 (covers the domain); boxes1: level 1 in the
    // index space of geom0 refined by 2 (aligned to 2)
    TwoLevelPredictor(const Geometry& geom0, std::vector<Box> boxes0, std::vector<Box> boxes1,
                      std::vector<BCRec> vel_bc, std::vector<double> vel_bcval,
                      std::array<Hydro::LinOpBCType, 3> lobc, std::array<Hydro::LinOpBCType, 3> hibc, double rho0,
                      double cfl, double mg_rtol, double mg_atol, const ksgf_mlmg_options& mlmg_opts,
                      int limiter = KSGF_PPM_VANLEER)
        : vel_bc_(std::move(vel_bc)),
          vel_bcval_(std::move(vel_bcval)),
          rho0_(rho0),
          cfl_(cfl),
          rtol_(mg_rtol),
          atol_(mg_atol),
          limiter_(limiter),
          mac_proj_({geom0, refined(geom0)})
    {
        if (vel_bc_.size() != 3 || vel_bcval_.size() != 18) throw std::runtime_error("TwoLevelPredictor: BC sizes");
        lev_[0] = std::make_unique<Level>(geom0, std::move(boxes0), true, rho0_);
        lev_[1] = std::make_unique<Level>(refined(geom0), std::move(boxes1), false, rho0_);
        src_bc_ = vel_bc_;
        for (auto& b : src_bc_)
            for (int d = 0; d < 3; ++d) {
                if (b.lo[d] != KSGF_BC_INT_DIR) b.lo[d] = KSGF_BC_FOEXTRAP;
                if (b.hi[d] != KSGF_BC_INT_DIR) b.hi[d] = KSGF_BC_FOEXTRAP;
            }
        Hydro::LPInfo info;
        info.do_semicoarsening = mlmg_opts.do_semicoarsening != 0;
        info.max_coarsening_level = mlmg_opts.max_coarsening_level;
        mac_proj_.getMLMG().opts = mlmg_opts;
        mac_proj_.initProjector({&lev_[0]->ba, &lev_[1]->ba}, info, 1.0 / rho0_);
        mac_proj_.setDomainBC(lobc, hibc);
        bindUMAC();
    }
    TwoLevelPredictor(const TwoLevelPredictor&) = delete;
    TwoLevelPredictor& operator=(const TwoLevelPredictor&) = delete;

    void setStream(hipStream_t st) { st_ = st; }
    const BoxArray& boxArray(int lev) const { return lev_[lev]->ba; }

    // arrays in the FAB layout (vel with 3 ghost layers, gp with none), in
    // host memory or, with on_device, device memory
    void setState(const double* vel0, const double* vel1, const double* gp0, const double* gp1,
                  bool on_device = false)
    {
        const double* v[2] = {vel0, vel1};
        const double* g[2] = {gp0, gp1};
        for (int l = 0; l < 2; ++l) {
            upload(*lev_[l]->vel, v[l], on_device);
            if (g[l]) upload(lev_[l]->gp, g[l], on_device);
            // the forcing compute_dt reads before the first step (momentum
            // form of the loaded gp)
            check(ksgf_icns_src_term(lev_[l]->ba.handle(), lev_[l]->src_nph.data(), 0, lev_[l]->gp.data(),
                                     lev_[l]->rho.data(), 3, 1, (ksgf_stream_t)st_),
                  "setState: src_term");
        }
        dt_prev_ = 0.0;
    }
    void getVelocity(double* vel0, double* vel1) const
    {
        (void)hipStreamSynchronize(st_);
        double* v[2] = {vel0, vel1};
        for (int l = 0; l < 2; ++l)
            if (hipMemcpy(v[l], lev_[l]->vel->data(), sizeof(double) * (size_t)lev_[l]->vel->size(),
                          hipMemcpyDeviceToHost) != hipSuccess)
                throw std::runtime_error("getVelocity: hipMemcpy");
    }

    StepStats advance(double fixed_dt = -1.0)
    {
        StepStats s;
        Level& L0 = *lev_[0];
        Level& L1 = *lev_[1];
        const ksgf_stream_t st = (ksgf_stream_t)st_;
        fillVelocity(false);
        double dt = fixed_dt;
        if (!(dt > 0.0)) {
            double c[2] = {0.0, 0.0}, f[2] = {0.0, 0.0};
            for (int l = 0; l < 2; ++l) {
                check(ksgf_cfl_max(lev_[l]->ba.handle(), lev_[l]->vel->data(), 3, &c[l], st), "compute_dt");
                check(ksgf_force_cfl_max(lev_[l]->ba.handle(), lev_[l]->src_nph.data(), 0, lev_[l]->rho.data(), 3,
                                         &f[l], st),
                      "compute_dt");
            }
            const double cd = std::max(c[0], c[1]) + 0.0, fc = std::max(f[0], f[1]);
            const double unit = cd + std::sqrt((cd * cd) + (4.0 * fc));
            const double eps = std::numeric_limits<double>::epsilon();
            if (unit < eps) throw std::runtime_error("compute_dt: CFL below machine epsilon, set a fixed dt");
            dt = 2.0 * cfl_ / std::max(unit, eps);
            if (dt_prev_ > 0.0) dt = std::min(dt, (1.0 + dt_growth_) * dt_prev_);
        }
        dt_prev_ = dt;
        s.dt = dt;
        // forcing, velocity form, one ghost layer
        for (auto& L : lev_)
            check(ksgf_icns_src_term(L->ba.handle(), L->src.data(), 1, L->gp.data(), L->rho.data(), 3, 0, st),
                  "compute_source_term");
        const std::vector<double> zero(18, 0.0);
        L0.src.FillBoundary(st_);
        L0.src.FillPhysBC(src_bc_, zero, st_);
        check(ksgf_fillpatch_two_levels(L1.ba.handle(), L1.src.data(), 1, L0.ba.handle(), L0.src.data(), 1, 3, 2,
                                        src_bc_.data(), zero.data(), st),
              "FillPatchTwoLevels (forcing)");
        // preadvect
        for (auto& L : lev_) {
            HydroUtils::ExtrapVelToFaces(*L->vel, L->src, L->umac, L->vmac, L->wmac, vel_bc_, nullptr, L->ba.geom(),
                                         dt, true, false, "Godunov", limiter_, false, st_);
            check(ksgf_set_inflow_velocity(L->ba.handle(), L->umac.data(), L->vmac.data(), L->wmac.data(), 1,
                                           vel_bc_.data(), vel_bcval_.data(), st),
                  "set_inflow_velocity");
        }
        mac_proj_.project(rtol_, atol_, st_);
        s.mlmg = mac_proj_.getMLMG().last;
        MultiFab* m0[3] = {&L0.umac, &L0.vmac, &L0.wmac};
        MultiFab* m1[3] = {&L1.umac, &L1.vmac, &L1.wmac};
        for (int d = 0; d < 3; ++d) {
            std::vector<BCRec> bc1{vel_bc_[d]};
            std::vector<double> v1(6);
            for (int o = 0; o < 6; ++o) v1[o] = vel_bcval_[o * 3 + d];
            m0[d]->FillBoundary(st_);
            m0[d]->FillPhysBC(bc1, v1, st_);
            check(ksgf_fillpatch_faces_two_levels(L1.ba.handle(), m1[d]->data(), KSGF_XFACE + d, 1, L0.ba.handle(),
                                                  m0[d]->data(), 1, 1, bc1.data(), v1.data(), st),
                  "fillpatch_sibling_fields");
        }
        // advection of momentum per level
        const int iconserv[3] = {1, 1, 1};
        for (auto& L : lev_) {
            check(ksgf_scale_by_rho(L->ba.handle(), L->q.data(), 3, L->vel->data(), 3, L->rho.data(), 3, 3, 3, 0, st),
                  "q = rho u");
            check(ksgf_scale_by_rho(L->ba.handle(), L->fq.data(), 1, L->src.data(), 1, L->rho.data(), 3, 3, 1, 0, st),
                  "fq = rho f");
            HydroUtils::ComputeFluxesOnBoxFromState(3, L->q, L->q, L->fx, L->fy, L->fz, L->ex, L->ey, L->ez, false,
                                                    L->umac, L->vmac, L->wmac, nullptr, L->fq, L->ba.geom(), dt,
                                                    vel_bc_, iconserv, true, false, true, false, "Godunov", limiter_,
                                                    false, st_);
        }
        // the fine fluxes onto the coarse faces before the coarse divergence
        check(ksgf_average_down_faces(L1.ba.handle(), L1.fx.data(), L1.fy.data(), L1.fz.data(), 0, L0.ba.handle(),
                                      L0.fx.data(), L0.fy.data(), L0.fz.data(), 0, 3, 2, st),
              "average_down_faces");
        for (auto& L : lev_) {
            HydroUtils::ComputeDivergence(L->conv, L->fx, L->fy, L->fz, 3, L->ba.geom(), -1.0, false, st_);
            check(ksgf_icns_src_term(L->ba.handle(), L->src_nph.data(), 0, L->gp.data(), L->rho.data(), 3, 1, st),
                  "src nph");
            check(ksgf_predictor_update(L->ba.handle(), L->vel_new->data(), 3, L->vel->data(), 3, L->rho.data(), 3,
                                        L->rho.data(), 3, L->conv.data(), L->src_nph.data(), nullptr, 0.0, dt, 3, st),
                  "predictor_rhs");
        }
        check(ksgf_average_down(L1.ba.handle(), L1.vel_new->data(), 3, L0.ba.handle(), L0.vel_new->data(), 3, 3, 2,
                                st),
              "average_down");
        fillVelocity(true);
        for (auto& L : lev_) std::swap(L->vel, L->vel_new);
        return s;
    }

private:
    struct Level {
        BoxArray ba;
        std::unique_ptr<MultiFab> vel, vel_new;
        MultiFab rho, gp, src, src_nph, q, fq, umac, vmac, wmac, fx, fy, fz, ex, ey, ez, conv;
        Level(const Geometry& g, std::vector<Box> boxes, bool covers, double rho0)
            : ba(g, std::move(boxes), covers),
              vel(std::make_unique<MultiFab>(ba, 3, 3)),
              vel_new(std::make_unique<MultiFab>(ba, 3, 3)),
              rho(ba, 1, 3),
              gp(ba, 3, 0),
              src(ba, 3, 1),
              src_nph(ba, 3, 0),
              q(ba, 3, 3),
              fq(ba, 3, 1),
              umac(ba, 1, 1, IndexType::XFace),
              vmac(ba, 1, 1, IndexType::YFace),
              wmac(ba, 1, 1, IndexType::ZFace),
              fx(ba, 3, 0, IndexType::XFace),
              fy(ba, 3, 0, IndexType::YFace),
              fz(ba, 3, 0, IndexType::ZFace),
              ex(ba, 3, 0, IndexType::XFace),
              ey(ba, 3, 0, IndexType::YFace),
              ez(ba, 3, 0, IndexType::ZFace),
              conv(ba, 3, 0)
        {
            std::vector<double> h((size_t)rho.size(), rho0);
            upload(rho, h.data());
        }
    };

    static Geometry refined(const Geometry& g)
    {
        Geometry f = g;
        for (int d = 0; d < 3; ++d) {
            f.domain.lo[d] = 2 * g.domain.lo[d];
            f.domain.hi[d] = 2 * g.domain.hi[d] + 1;
        }
        return f;
    }
    static void upload(MultiFab& m, const double* src, bool on_device = false)
    {
        if (hipMemcpy(m.data(), src, sizeof(double) * (size_t)m.size(),
                      on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("TwoLevelPredictor: hipMemcpy");
    }
    void bindUMAC()
    {
        mac_proj_.setUMAC({{&lev_[0]->umac, &lev_[0]->vmac, &lev_[0]->wmac},
                           {&lev_[1]->umac, &lev_[1]->vmac, &lev_[1]->wmac}});
    }
    // level-0 ghosts from its boxes and the physical BCs; level-1 ghosts by
    // the two-level fill-patch from level 0
    void fillVelocity(bool new_state)
    {
        MultiFab& v0 = new_state ? *lev_[0]->vel_new : *lev_[0]->vel;
        MultiFab& v1 = new_state ? *lev_[1]->vel_new : *lev_[1]->vel;
        v0.FillBoundary(st_);
        v0.FillPhysBC(vel_bc_, vel_bcval_, st_);
        check(ksgf_fillpatch_two_levels(lev_[1]->ba.handle(), v1.data(), 3, lev_[0]->ba.handle(), v0.data(), 3, 3, 2,
                                        vel_bc_.data(), vel_bcval_.data(), (ksgf_stream_t)st_),
              "FillPatchTwoLevels (velocity)");
    }

    std::vector<BCRec> vel_bc_, src_bc_;
    std::vector<double> vel_bcval_;
    double rho0_, cfl_, rtol_, atol_;
    int limiter_;
    double dt_prev_ = 0.0, dt_growth_ = 0.1;
    hipStream_t st_ = nullptr;
    std::unique_ptr<Level> lev_[2];
    Hydro::MacProjector mac_proj_;
};

}  // namespace ksgf::amr
