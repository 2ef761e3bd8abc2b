// Godunov branch of incflo::ApplyPredictor (incflo_advance.cpp:186-438) up to
// the predictor update, on the C++ host layer: the advect + MAC-project step
// without the diffusion solve and the nodal projection.
//
//   compute_dt (convective and forcing CFL, dt growth cap)
//                                      incflo_compute_dt.cpp:33-220, SimTime.cpp:162-205
//   src = -gp/rho (velocity form), 1-ghost fill
//                                      icns_ops.H:98-111, incflo_advance.cpp:283-291
//   AdvectionOp<ICNS,Godunov>::preadvect
//       ExtrapVelToFaces               icns_advection.H:211-216
//       MacProjOp::set_inflow_velocity icns_advection.cpp:157-183
//       MacProjOp::operator()          icns_advection.cpp:206-375
//       face ghost fill                icns_advection.H:233-243
//   AdvectionOp<ICNS,Godunov>::operator()
//       q = rho u, fq = rho f          icns_advection.H:281-314
//       ComputeFluxesOnBoxFromState    icns_advection.H:352-374
//       ComputeDivergence (mult -1)    icns_advection.H:418-422
//   compute_source_term(NPH) + ComputeRHSOp::predictor_rhs
//                                      icns_ops.H:84-121, CompRHSOps.H:93-117
// With enableDiffusion (constant viscosity mu and density rho_0) the step
// carries the viscous term as ApplyPredictor does for Godunov
// (incflo_advance.cpp:246-291, 380-389; DiffusionOps.cpp:154-160):
//   Crank-Nicolson: divtau = mu lap(u^n) (MLMG::apply with scalars (0, -1)),
//     added to the forcing before the division by density, kappa = 1/2 in
//     the predictor, then (rho - dt/2 mu lap) u = rho u*;
//   Implicit: kappa = 0, then (rho - dt mu lap) u = rho u*;
// by default through the tensor operator, as ICNS's DiffusionOp does
// (use_tensor_operator = true, icns_diffusion.H:492-506): MLTensorOp with
// a = rho, the viscosity averaged to the faces (average_velocity_eta_to_faces)
// and divtau = div(mu (grad u + grad u^T) - (2/3) mu (div u) I); or, with the
// scalar operator, one MLABecLaplacian solve per velocity component
// (divtau = mu lap(u)).  Each component takes its BCs from
// get_diffuse_tensor_bc (incflo_diffusion.cpp:11-69: periodic; ext_dir ->
// Dirichlet; foextrap / hoextrap -> Neumann).
// With enableNodalProjection the step ends with incflo::ApplyProjection
// (incflo_advance.cpp:412-414; incflo_apply_nodal_projection.cpp:149-465,
// constant density, not incremental, no small-dt variant):
//   vel += dt gp / rho, vel.setBndry(0), FillBoundary, NodalProjector with
//   sigma = dt / rho_0, project, gp = grad phi, p = phi, velocity fillpatch
// With a passive scalar (enableScalar), the scalar equation runs between the
// MAC projection and the momentum advection, as ApplyPredictor orders the
// scalar equations (incflo_advance.cpp:316-324):
//   AdvectionOp<PDE,Godunov>::operator() with multiply_rho
//       q = rho s, ComputeFluxesOnBoxFromState (ncomp 1, is_velocity false,
//       zero forcing and divu), ComputeDivergence (mult -1)
//                                      AdvOp_Godunov.H:116-300
//   compute_predictor_rhs              CompRHSOps.H:93-117
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <utility>

#include "ksgf_hydro.H"

namespace ksgf::amr {

struct StepStats {
    double dt = 0.0;
    ksgf_mlmg_info mlmg{};
    ksgf_mlmg_info nodal{};  // the nodal projection's solve (when enabled)
    ksgf_mlmg_info diffusion[3]{};  // the velocity components' diffusion solves (when enabled)
    double divmax = -1.0;  // max |div u_mac| after the projection (when asked for)
};

class GodunovPredictor {
public:
    GodunovPredictor(const Geometry& geom, std::vector<Box> boxes, std::vector<BCRec> vel_bc,
                     std::vector<double> vel_bcval, std::array<Hydro::LinOpBCType, 3> lobc,
                     std::array<Hydro::LinOpBCType, 3> hibc, double rho0, double cfl, double mg_rtol, double mg_atol,
                     const ksgf_mlmg_options& mlmg_opts)
        : ba_(geom, std::move(boxes)),
          vel_bc_(std::move(vel_bc)),
          vel_bcval_(std::move(vel_bcval)),
          rho0_(rho0),
          cfl_(cfl),
          rtol_(mg_rtol),
          atol_(mg_atol),
          vel_(std::make_unique<MultiFab>(ba_, 3, 3)),
          vel_new_(std::make_unique<MultiFab>(ba_, 3, 3)),
          rho_(ba_, 1, 3),
          gp_(ba_, 3, 0),
          src_(ba_, 3, 1),
          src_nph_(ba_, 3, 0),
          q_(ba_, 3, 3),
          fq_(ba_, 3, 1),
          umac_(ba_, 1, 1, IndexType::XFace),
          vmac_(ba_, 1, 1, IndexType::YFace),
          wmac_(ba_, 1, 1, IndexType::ZFace),
          fx_(ba_, 3, 0, IndexType::XFace),
          fy_(ba_, 3, 0, IndexType::YFace),
          fz_(ba_, 3, 0, IndexType::ZFace),
          ex_(ba_, 3, 0, IndexType::XFace),
          ey_(ba_, 3, 0, IndexType::YFace),
          ez_(ba_, 3, 0, IndexType::ZFace),
          conv_(ba_, 3, 0),
          mac_proj_({geom})
    {
        if (vel_bc_.size() != 3 || vel_bcval_.size() != 18) throw std::runtime_error("GodunovPredictor: BC sizes");
        // forcing ghosts: first-order extrapolation where the velocity is not periodic
        src_bc_ = vel_bc_;
        for (auto& b : src_bc_)
            for (int d = 0; d < 3; ++d) {
                if (b.lo[d] != KSGF_BC_INT_DIR) b.lo[d] = KSGF_BC_FOEXTRAP;
                if (b.hi[d] != KSGF_BC_INT_DIR) b.hi[d] = KSGF_BC_FOEXTRAP;
            }
        rho_bc_.resize(1);
        for (int d = 0; d < 3; ++d) {
            rho_bc_[0].lo[d] = geom.isPeriodic(d) ? KSGF_BC_INT_DIR : KSGF_BC_FOEXTRAP;
            rho_bc_[0].hi[d] = geom.isPeriodic(d) ? KSGF_BC_INT_DIR : KSGF_BC_FOEXTRAP;
        }
        // MacProjOp::init_projector(Real beta) (icns_advection.cpp:109-155)
        Hydro::LPInfo info;
        info.do_semicoarsening = mlmg_opts.do_semicoarsening != 0;
        info.max_coarsening_level = mlmg_opts.max_coarsening_level;
        mac_proj_.getMLMG().opts = mlmg_opts;
        mac_proj_.initProjector({&ba_}, info, 1.0 / rho0_);
        mac_proj_.setDomainBC(lobc, hibc);
        mac_proj_.setUMAC({{&umac_, &vmac_, &wmac_}});
        lobc_ = lobc;
        hibc_ = hibc;
    }

    ~GodunovPredictor()
    {
        clearGraphs();
        if (gst_) {
            (void)hipStreamSynchronize(gst_);
            (void)hipStreamDestroy(gst_);
        }
    }
    GodunovPredictor(const GodunovPredictor&) = delete;
    GodunovPredictor& operator=(const GodunovPredictor&) = delete;

    void setStream(hipStream_t st)
    {
        if (graph_) throw std::runtime_error("setStream: graph replay owns its stream");
        st_ = st;
    }
    const BoxArray& boxArray() const { return ba_; }

    // domain-wide host arrays: vel (3 x N), rho (N), gp (3 x N).  The density
    // must equal rho_0 everywhere: the projections here use beta = 1 / rho_0
    // and sigma = dt / rho_0, where the reference switches to face and cell
    // coefficients 1 / rho (icns_advection.cpp:234-291,
    // incflo_apply_nodal_projection.cpp:259-290) once density varies.
    void setState(const double* vel, const double* rho, const double* gp)
    {
        const size_t N = (size_t)ba_.geom().domain.numPts();
        for (size_t c = 0; c < N; ++c)
            if (rho[c] != rho0_)
                throw UnsupportedError("setState: density differs from rho_0 (variable-density projections are not "
                                       "provided)");
        vel_->copyFromDomain(vel);
        rho_.copyFromDomain(rho);
        gp_.copyFromDomain(gp);
        vel_->FillBoundary(st_);
        vel_->FillPhysBC(vel_bc_, vel_bcval_, st_);
        rho_.FillBoundary(st_);
        rho_.FillPhysBC(rho_bc_, std::vector<double>(6, 0.0), st_);
        // the forcing compute_dt reads before the first step: the source
        // term (momentum form) of the loaded gp, as the initial iterations
        // leave src_term (incflo_advance.cpp:367, icns_ops.H:98-120)
        check(ksgf_icns_src_term(ba_.handle(), src_nph_.data(), 0, gp_.data(), rho_.data(), 3, 1,
                                 (ksgf_stream_t)st_),
              "setState: src_term");
        dt_prev_ = 0.0;
    }
    // time.max_dt_growth (SimTime.H:248, default 0.1) and time.use_force_cfl
    // (SimTime.H:332, default true)
    void setTimeOptions(double max_dt_growth, bool use_force_cfl)
    {
        dt_growth_ = max_dt_growth;
        use_force_cfl_ = use_force_cfl;
    }
    void getVelocity(double* vel) const
    {
        (void)hipStreamSynchronize(st_);
        vel_->copyToDomain(vel);
    }
    // the MAC projection's phi of the last step, domain-wide (host array)
    void getPhi(double* phi)
    {
        const Box& dom = ba_.geom().domain;
        size_t n = 1;
        for (int d = 0; d < 3; ++d) n *= (size_t)(dom.hi[d] - dom.lo[d] + 1);
        double* d = nullptr;
        if (hipMalloc((void**)&d, n * sizeof(double)) != hipSuccess) throw std::runtime_error("getPhi: hipMalloc");
        try {
            mac_proj_.getPhi(d, st_);
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        const hipError_t e = hipMemcpyAsync(phi, d, n * sizeof(double), hipMemcpyDeviceToHost, st_);
        (void)hipStreamSynchronize(st_);
        (void)hipFree(d);
        if (e != hipSuccess) throw std::runtime_error("getPhi: copy");
    }
    MultiFab& velocity() { return *vel_; }
    MultiFab& umac() { return umac_; }
    MultiFab& vmac() { return vmac_; }
    MultiFab& wmac() { return wmac_; }

    void setCheckDivergence(bool on)
    {
        if (on && graph_) throw ksgf::amr::UnsupportedError("setCheckDivergence: not with graph replay");
        check_div_ = on;
    }

    // diffusion: type 1 Crank-Nicolson, 2 Implicit (incflo.diffusion_type);
    // Dirichlet velocity BCs must hold zero (walls at rest)
    void enableDiffusion(double mu, int type, double rtol, double atol, const ksgf_mlmg_options& opts,
                         bool tensor = true)
    {
        if (graph_) throw ksgf::amr::UnsupportedError("enableDiffusion: not with graph replay");
        if (type != 1 && type != 2) throw std::runtime_error("enableDiffusion: type 1 (Crank-Nicolson) or 2 (Implicit)");
        if (!(mu > 0.0)) throw std::runtime_error("enableDiffusion: mu must be positive");
        Hydro::LPInfo info;
        info.do_semicoarsening = opts.do_semicoarsening != 0;
        info.max_coarsening_level = opts.max_coarsening_level;
        std::array<std::array<Hydro::LinOpBCType, 3>, 3> tlo{}, thi{};
        for (int c = 0; c < 3; ++c) {
            std::array<Hydro::LinOpBCType, 3> lo{}, hi{};
            for (int d = 0; d < 3; ++d) {
                for (int side = 0; side < 2; ++side) {
                    const int t = side ? vel_bc_[c].hi[d] : vel_bc_[c].lo[d];
                    Hydro::LinOpBCType lt = Hydro::LinOpBCType::Neumann;
                    if (t == KSGF_BC_INT_DIR) lt = Hydro::LinOpBCType::Periodic;
                    if (t == KSGF_BC_EXT_DIR) {
                        lt = Hydro::LinOpBCType::Dirichlet;
                        if (vel_bcval_[(side * 3 + d) * 3 + c] != 0.0)
                            throw std::runtime_error("enableDiffusion: inhomogeneous Dirichlet velocity is not provided");
                    }
                    (side ? hi : lo)[d] = lt;
                }
            }
            tlo[c] = lo;
            thi[c] = hi;
            lap_[c].reset();
            if (tensor) continue;
            lap_[c] = std::make_unique<Hydro::MLABecLaplacian>(ba_, info, opts);
            lap_[c]->setDomainBC(lo, hi);
            lap_[c]->setACoeffs(rho0_);
            lap_[c]->setBCoeffs(mu);
        }
        tensor_.reset();
        eta_.reset();
        if (tensor) {
            tensor_ = std::make_unique<Hydro::MLTensorOp>(ba_, info, opts);
            tensor_->setDomainBC(tlo, thi);
            // the (constant) viscosity on the faces
            MultiFab mu_cc(ba_, 1, 1);
            std::vector<double> h((size_t)ba_.geom().domain.numPts(), mu);
            mu_cc.copyFromDomain(h.data());
            mu_cc.FillBoundary(st_);
            mu_cc.FillPhysBC(rho_bc_, std::vector<double>(6, 0.0), st_);
            eta_ = std::make_unique<std::array<MultiFab, 3>>(ksgf::amr::average_velocity_eta_to_faces(mu_cc, st_));
            tensor_->setShearViscosity({&(*eta_)[0], &(*eta_)[1], &(*eta_)[2]}, st_);
            (void)hipStreamSynchronize(st_);
        }
        divtau_ = std::make_unique<MultiFab>(ba_, 3, 0);
        drhs_ = std::make_unique<MultiFab>(ba_, 3, 0);
        diff_type_ = type;
        diff_rtol_ = rtol;
        diff_atol_ = atol;
    }

    // end every step with the nodal projection (nodal_proj.mg_rtol / mg_atol);
    // its domain BCs are the MAC projection's (get_projection_bc maps the
    // pressure BCs the same way, incflo_apply_nodal_projection.cpp:28-58)
    void enableNodalProjection(double rtol, double atol, const ksgf_mlmg_options& opts)
    {
        if (graph_) throw ksgf::amr::UnsupportedError("enableNodalProjection: not with graph replay");
        Hydro::LPInfo info;
        info.do_semicoarsening = opts.do_semicoarsening != 0;
        info.max_coarsening_level = opts.max_coarsening_level;
        nodal_ = std::make_unique<Hydro::NodalProjector>(std::vector<MultiFab*>{vel_.get()}, 1.0,
                                                         std::vector<Geometry>{ba_.geom()}, info);
        nodal_->getMLMG().opts = opts;
        nodal_->setDomainBC(lobc_, hibc_);
        nodal_rtol_ = rtol;
        nodal_atol_ = atol;
    }
    // the last nodal projection's phi (domain-wide nodes, host) and gp (cell, host)
    void getNodalPhi(double* phi)
    {
        if (!nodal_) throw std::runtime_error("getNodalPhi: nodal projection not enabled");
        const auto N = nodal_->nodeDims();
        const size_t n = (size_t)N[0] * N[1] * N[2];
        double* d = nullptr;
        if (hipMalloc((void**)&d, n * sizeof(double)) != hipSuccess) throw std::runtime_error("getNodalPhi: hipMalloc");
        nodal_->getPhi(d, st_);
        const hipError_t e = hipMemcpyAsync(phi, d, n * sizeof(double), hipMemcpyDeviceToHost, st_);
        (void)hipStreamSynchronize(st_);
        (void)hipFree(d);
        if (e != hipSuccess) throw std::runtime_error("getNodalPhi: copy");
    }
    const ksgf_mlmg_info& lastNodalInfo() const { return last_nodal_; }
    const StepStats& lastStats() const { return last_stats_; }
    void getGradP(double* gp) const
    {
        (void)hipStreamSynchronize(st_);
        gp_.copyToDomain(gp);
    }

    // passive scalar: one component, ghost 3, its own BCRec and values
    // (6 orientations)
    void enableScalar(const BCRec& bc, std::vector<double> bcval)
    {
        if (bcval.size() != 6) throw std::runtime_error("enableScalar: 6 BC values");
        sc_ = std::make_unique<Scalar>(ba_);
        sc_->bc = {bc};
        sc_->bcval = std::move(bcval);
    }
    void setScalar(const double* s)
    {
        if (!sc_) throw std::runtime_error("setScalar: no scalar enabled");
        sc_->s->copyFromDomain(s);
        sc_->s->FillBoundary(st_);
        sc_->s->FillPhysBC(sc_->bc, sc_->bcval, st_);
    }
    void getScalar(double* s) const
    {
        if (!sc_) throw std::runtime_error("getScalar: no scalar enabled");
        (void)hipStreamSynchronize(st_);
        sc_->s->copyToDomain(s);
    }

    // IOManager::write_plot_file (IOManager.cpp:136-190) for the fields this
    // step owns: velocity, gp, density and the passive scalar when enabled,
    // in one cell-centred MultiFab (staged through the host, as the
    // reference copies them into one scratch field first)
    void writePlotfile(const std::string& name, double time, int step)
    {
        const auto& dom = ba_.geom().domain;
        const size_t N = (size_t)dom.numPts();
        const int nc = sc_ ? 8 : 7;
        std::vector<double> all((size_t)nc * N);
        (void)hipStreamSynchronize(st_);
        vel_->copyToDomain(all.data());
        gp_.copyToDomain(all.data() + 3 * N);
        rho_.copyToDomain(all.data() + 6 * N);
        if (sc_) sc_->s->copyToDomain(all.data() + 7 * N);
        MultiFab out(ba_, nc, 0);
        out.copyFromDomain(all.data());
        std::vector<std::string> names = {"velocityx", "velocityy", "velocityz", "gpx", "gpy", "gpz", "density"};
        if (sc_) names.push_back("temperature");
        ksgf::amr::WriteSingleLevelPlotfile(name, out, names, time, step, st_);
    }

    StepStats advance(double fixed_dt = -1.0)
    {
        if (graph_ && fixed_dt > 0.0) return advanceGraph(fixed_dt);
        StepStats s;
        const ksgf_boxarray* h = ba_.handle();
        ksgf_stream_t st = (ksgf_stream_t)st_;
        // compute_dt
        double dt = fixed_dt;
        if (!(dt > 0.0)) {
            // incflo::compute_dt: convective CFL of the velocity, forcing CFL
            // of the source term the last step formed; then
            // SimTime::set_current_cfl (no diffusive CFL: Godunov treats
            // viscosity implicitly or by Crank-Nicolson)
            double c = 0.0, f = 0.0;
            check(ksgf_cfl_max(h, vel_->data(), 3, &c, st), "compute_dt");
            if (use_force_cfl_) check(ksgf_force_cfl_max(h, src_nph_.data(), 0, rho_.data(), 3, &f, st), "compute_dt");
            const double cd = c + 0.0;
            const double unit = cd + std::sqrt((cd * cd) + (4.0 * f));
            const double eps = std::numeric_limits<double>::epsilon();
            if (unit < eps) throw std::runtime_error("compute_dt: CFL below machine epsilon, set a fixed dt");
            dt = 2.0 * cfl_ / std::max(unit, eps);
            if (dt_prev_ > 0.0) dt = std::min(dt, (1.0 + dt_growth_) * dt_prev_);
        }
        dt_prev_ = dt;
        s.dt = dt;
        preProjection(dt);
        mac_proj_.project(rtol_, atol_, st_);
        s.mlmg = mac_proj_.getMLMG().last;
        if (check_div_)
            check(ksgf_mac_divergence_max(h, umac_.data(), vmac_.data(), wmac_.data(), 1, &s.divmax, st),
                  "divergence check");
        postProjection(dt, &s);
        swapStates();
        if (nodal_) applyProjection(dt, s);
        last_stats_ = s;
        return s;
    }

    // A fixed-dt step (time.fixed_dt > 0) replayed from hipGraphs: the part
    // before the MAC projection's solve and the solve's first `ncycles`
    // iterations (tested on the device) in one graph, then one host read of
    // the solve's state (an unfinished solve completes on the host path),
    // then the rest of the step in a second graph.  The first fixed-dt step
    // runs eagerly (it builds every launch table and autotunes) and sizes the
    // captured solve at its iteration count (rounded up to even; iterations
    // past convergence cost a dispatch per kernel, a later step that needs
    // more continues on the host path).  Graphs are kept per parity
    // of the state buffers and per dt.  Steps equal the eager ones bit for
    // bit.  Not with diffusion, the nodal projection or the divergence check
    // (they read values on the host inside the step).  Owns a stream of its
    // own from here on.
    void enableGraph(bool on)
    {
        if (on && (diff_type_ > 0 || nodal_ || check_div_))
            throw ksgf::amr::UnsupportedError(
                "enableGraph: not with diffusion, the nodal projection or the divergence check");
        clearGraphs();
        graph_ = on;
        graph_ncyc_ = 0;
        if (on && !gst_) {
            if (hipStreamCreateWithFlags(&gst_, hipStreamNonBlocking) != hipSuccess)
                throw std::runtime_error("enableGraph: hipStreamCreateWithFlags");
            (void)hipStreamSynchronize(st_);
            st_ = gst_;
        }
    }

private:
    struct StepGraph {
        hipGraphExec_t pre = nullptr, post = nullptr;
        double dt = 0.0;
    };

    void clearGraphs()
    {
        for (auto& g : graphs_) {
            if (g.pre) (void)hipGraphExecDestroy(g.pre);
            if (g.post) (void)hipGraphExecDestroy(g.post);
            g = StepGraph{};
        }
    }

    template <typename F>
    hipGraphExec_t capture(F&& body)
    {
        check(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal) == hipSuccess ? 0 : KSGF_ERR_HIP,
              "hipStreamBeginCapture");
        try {
            body();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipGraph_t g = nullptr;
        if (hipStreamEndCapture(st_, &g) != hipSuccess || !g) throw std::runtime_error("hipStreamEndCapture");
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) throw std::runtime_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        return exec;
    }

    StepStats advanceGraph(double dt)
    {
        if (graph_ncyc_ == 0) {
            graph_ = false;
            StepStats s;
            try {
                s = advance(dt);
            } catch (...) {
                graph_ = true;
                throw;
            }
            graph_ = true;
            graph_ncyc_ = std::max(1, s.mlmg.iters);
            return s;
        }
        StepGraph& g = graphs_[parity_];
        if (!g.pre || g.dt != dt) {
            if (g.pre) (void)hipGraphExecDestroy(g.pre);
            if (g.post) (void)hipGraphExecDestroy(g.post);
            g = StepGraph{};
            g.pre = capture([&] {
                preProjection(dt);
                mac_proj_.projectBegin(rtol_, atol_, graph_ncyc_, st_);
            });
            g.post = capture([&] {
                mac_proj_.projectCorrect(st_);
                postProjection(dt, nullptr);
            });
            g.dt = dt;
        }
        StepStats s;
        s.dt = dt;
        dt_prev_ = dt;
        check(hipGraphLaunch(g.pre, st_) == hipSuccess ? 0 : KSGF_ERR_HIP, "hipGraphLaunch");
        mac_proj_.projectFinish(st_);
        s.mlmg = mac_proj_.getMLMG().last;
        check(hipGraphLaunch(g.post, st_) == hipSuccess ? 0 : KSGF_ERR_HIP, "hipGraphLaunch");
        swapStates();
        last_stats_ = s;
        return s;
    }

    // the step up to the MAC projection: explicit viscous term, forcing,
    // ExtrapVelToFaces, inflow faces, beta
    void preProjection(double dt)
    {
        const ksgf_boxarray* h = ba_.handle();
        ksgf_stream_t st = (ksgf_stream_t)st_;
        // explicit viscous term of the old state (Crank-Nicolson)
        const bool cn = (diff_type_ == 1);
        if (cn && tensor_) {
            tensor_->setScalars(0.0, -1.0);
            tensor_->apply(*divtau_, *vel_, st_);
        } else if (cn) {
            for (int c = 0; c < 3; ++c) {
                lap_[c]->setScalars(0.0, -1.0);
                lap_[c]->apply(*divtau_, *vel_, c, st_);
            }
        }
        // forcing in velocity form with one ghost
        if (cn)
            check(ksgf_icns_src_term_divtau(h, src_.data(), 1, gp_.data(), divtau_->data(), rho_.data(), 3, st),
                  "compute_source_term + divtau");
        else
            check(ksgf_icns_src_term(h, src_.data(), 1, gp_.data(), rho_.data(), 3, 0, st), "compute_source_term");
        src_.FillBoundary(st_);
        src_.FillPhysBC(src_bc_, std::vector<double>(18, 0.0), st_);
        // preadvect
        HydroUtils::ExtrapVelToFaces(*vel_, src_, umac_, vmac_, wmac_, vel_bc_, nullptr, ba_.geom(), dt, true, false,
                                     "Godunov", KSGF_PPM_VANLEER, false, st_);
        check(ksgf_set_inflow_velocity(h, umac_.data(), vmac_.data(), wmac_.data(), 1, vel_bc_.data(),
                                       vel_bcval_.data(), st),
              "set_inflow_velocity");
        mac_proj_.updateBeta(1.0 / rho0_);
    }

    // the step after the MAC projection: face fills, scalar, momentum fluxes,
    // divergence, predictor, implicit viscous solve, new-state fills (the
    // states are swapped by the caller)
    void postProjection(double dt, StepStats* s)
    {
        const ksgf_boxarray* h = ba_.handle();
        ksgf_stream_t st = (ksgf_stream_t)st_;
        const bool cn = (diff_type_ == 1);
        MultiFab* mac[3] = {&umac_, &vmac_, &wmac_};
        for (int d = 0; d < 3; ++d) {
            std::vector<BCRec> bc1{vel_bc_[d]};
            std::vector<double> v1(6);
            for (int o = 0; o < 6; ++o) v1[o] = vel_bcval_[o * 3 + d];
            mac[d]->FillBoundary(st_);
            mac[d]->FillPhysBC(bc1, v1, st_);
        }
        if (sc_) advanceScalar(dt);
        // advection of momentum
        check(ksgf_scale_by_rho(h, q_.data(), 3, vel_->data(), 3, rho_.data(), 3, 3, 3, 0, st), "q = rho u");
        check(ksgf_scale_by_rho(h, fq_.data(), 1, src_.data(), 1, rho_.data(), 3, 3, 1, 0, st), "fq = rho f");
        const int iconserv[3] = {1, 1, 1};
        HydroUtils::ComputeFluxesOnBoxFromState(3, q_, q_, fx_, fy_, fz_, ex_, ey_, ez_, false, umac_, vmac_, wmac_,
                                                nullptr, fq_, ba_.geom(), dt, vel_bc_, iconserv, true, false, true,
                                                false, "Godunov", KSGF_PPM_VANLEER, false, st_);
        HydroUtils::ComputeDivergence(conv_, fx_, fy_, fz_, 3, ba_.geom(), -1.0, false, st_);
        // source at n+1/2 (momentum form) and the predictor update
        check(ksgf_icns_src_term(h, src_nph_.data(), 0, gp_.data(), rho_.data(), 3, 1, st), "src nph");
        check(ksgf_predictor_update(h, vel_new_->data(), 3, vel_->data(), 3, rho_.data(), 3, rho_.data(), 3,
                                    conv_.data(), src_nph_.data(), cn ? divtau_->data() : nullptr, cn ? 0.5 : 0.0, dt,
                                    3, st),
              "predictor_rhs");
        if (diff_type_ > 0) {
            // (rho - dt_diff mu lap) u = rho u*, component by component
            const double dt_diff = cn ? 0.5 * dt : dt;
            check(ksgf_scale_by_rho(h, drhs_->data(), 0, vel_new_->data(), 3, rho_.data(), 3, 3, 0, 0, st),
                  "diffusion rhs = rho u*");
            if (tensor_) {
                // one coupled solve of the three components (its statistics in diffusion[0])
                tensor_->setScalars(1.0, dt_diff);
                tensor_->setACoeffs(rho_, st_);
                const ksgf_mlmg_info di = tensor_->solve(*vel_new_, *drhs_, diff_rtol_, diff_atol_, st_);
                if (s) s->diffusion[0] = di;
            } else {
                for (int c = 0; c < 3; ++c) {
                    lap_[c]->setScalars(1.0, dt_diff);
                    const ksgf_mlmg_info di = lap_[c]->solve(*vel_new_, *drhs_, c, diff_rtol_, diff_atol_, st_);
                    if (s) s->diffusion[c] = di;
                }
            }
        }
        vel_new_->FillBoundary(st_);
        vel_new_->FillPhysBC(vel_bc_, vel_bcval_, st_);
    }

    void swapStates()
    {
        std::swap(vel_, vel_new_);
        if (sc_) std::swap(sc_->s, sc_->s_new);
        parity_ ^= 1;
    }

    void applyProjection(double dt, StepStats& s)
    {
        const ksgf_boxarray* h = ba_.handle();
        ksgf_stream_t st = (ksgf_stream_t)st_;
        check(ksgf_incflo_add_gradp(h, vel_->data(), 3, gp_.data(), 0, rho_.data(), 3, dt, st),
              "ApplyProjection: vel += dt gp / rho");
        check(ksgf_set_bndry(h, vel_->data(), KSGF_CELL, 3, 3, 0.0, st), "ApplyProjection: setBndry");
        vel_->FillBoundary(st_);
        nodal_->setVelocity(vel_.get());
        nodal_->setSigma(dt / rho0_);
        nodal_->project(nodal_rtol_, nodal_atol_, &gp_, st_);  // gp = grad phi
        s.nodal = nodal_->getMLMG().last;
        last_nodal_ = s.nodal;
        vel_->FillBoundary(st_);
        vel_->FillPhysBC(vel_bc_, vel_bcval_, st_);
    }

    struct Scalar {
        explicit Scalar(const BoxArray& ba)
            : s(std::make_unique<MultiFab>(ba, 1, 3)),
              s_new(std::make_unique<MultiFab>(ba, 1, 3)),
              q(ba, 1, 3),
              fq(ba, 1, 1),
              src(ba, 1, 0),
              conv(ba, 1, 0),
              fx(ba, 1, 0, IndexType::XFace),
              fy(ba, 1, 0, IndexType::YFace),
              fz(ba, 1, 0, IndexType::ZFace),
              ex(ba, 1, 0, IndexType::XFace),
              ey(ba, 1, 0, IndexType::YFace),
              ez(ba, 1, 0, IndexType::ZFace)
        {
        }
        std::unique_ptr<MultiFab> s, s_new;
        MultiFab q, fq, src, conv, fx, fy, fz, ex, ey, ez;  // fq, src: zero (no scalar forcing)
        std::vector<BCRec> bc;
        std::vector<double> bcval;
    };

    void advanceScalar(double dt)
    {
        const ksgf_boxarray* h = ba_.handle();
        ksgf_stream_t st = (ksgf_stream_t)st_;
        Scalar& S = *sc_;
        check(ksgf_scale_by_rho(h, S.q.data(), 3, S.s->data(), 3, rho_.data(), 3, 1, 3, 0, st), "q = rho s");
        const int iconserv[1] = {1};
        HydroUtils::ComputeFluxesOnBoxFromState(1, S.q, S.q, S.fx, S.fy, S.fz, S.ex, S.ey, S.ez, false, umac_, vmac_,
                                                wmac_, nullptr, S.fq, ba_.geom(), dt, S.bc, iconserv, true, false,
                                                false, false, "Godunov", KSGF_PPM_VANLEER, false, st_);
        HydroUtils::ComputeDivergence(S.conv, S.fx, S.fy, S.fz, 1, ba_.geom(), -1.0, false, st_);
        check(ksgf_predictor_update(h, S.s_new->data(), 3, S.s->data(), 3, rho_.data(), 3, rho_.data(), 3,
                                    S.conv.data(), S.src.data(), nullptr, 0.0, dt, 1, st),
              "scalar predictor_rhs");
        S.s_new->FillBoundary(st_);
        S.s_new->FillPhysBC(S.bc, S.bcval, st_);
    }

    std::unique_ptr<Scalar> sc_;
    std::unique_ptr<Hydro::NodalProjector> nodal_;
    std::unique_ptr<Hydro::MLABecLaplacian> lap_[3];
    std::unique_ptr<Hydro::MLTensorOp> tensor_;
    std::unique_ptr<std::array<MultiFab, 3>> eta_;
    std::unique_ptr<MultiFab> divtau_, drhs_;
    int diff_type_ = 0;
    double diff_rtol_ = 0.0, diff_atol_ = 0.0;
    double nodal_rtol_ = 0.0, nodal_atol_ = 0.0;
    ksgf_mlmg_info last_nodal_{};
    StepStats last_stats_{};
    std::array<Hydro::LinOpBCType, 3> lobc_{}, hibc_{};
    BoxArray ba_;
    std::vector<BCRec> vel_bc_, src_bc_, rho_bc_;
    std::vector<double> vel_bcval_;
    double rho0_, cfl_, rtol_, atol_;
    double dt_prev_ = 0.0, dt_growth_ = 0.1;
    bool use_force_cfl_ = true;
    hipStream_t st_ = nullptr;
    bool check_div_ = false;
    // graph replay of fixed-dt steps
    bool graph_ = false;
    int graph_ncyc_ = 0;
    int parity_ = 0;  // flips with every swap of the state buffers
    StepGraph graphs_[2];
    hipStream_t gst_ = nullptr;
    std::unique_ptr<MultiFab> vel_, vel_new_;
    MultiFab rho_, gp_, src_, src_nph_, q_, fq_;
    MultiFab umac_, vmac_, wmac_;
    MultiFab fx_, fy_, fz_, ex_, ey_, ez_;
    MultiFab conv_;
    Hydro::MacProjector mac_proj_;
};

}  // namespace ksgf::amr
