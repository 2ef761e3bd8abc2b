// C++ host layer: the AMReX / AMReX-Hydro names and signatures the
// reference calls on this path, over the C ABI of include/ksgf.h.
//
//   amrex-like data model   Geometry, Box, BoxArray, BCRec, MultiFab
//                           (device storage in the ksgf MultiFab layout)
//   HydroUtils              ExtrapVelToFaces   (icns_advection.H:211-216)
//                           ComputeFluxesOnBoxFromState (:359-373), one call
//                           for all boxes of the MultiFab (the MFIter loop)
//                           ComputeDivergence  (:418-422)
//   Hydro::MacProjector     MacProjector(geom), initProjector(ba, lpinfo,
//                           beta), setDomainBC, updateBeta, setUMAC, project,
//                           getMLMG (icns_advection.cpp:88-104,143-154,
//                           289-353,372-374)
//   plotfile / VisMF        WriteSingleLevelPlotfile, ReadSingleLevelPlotfile,
//                           VisMF::Write / Read (IOManager.cpp:181-187,
//                           214-219, 249-266)
//
// Errors from the C ABI become std::runtime_error (amrex.throw_exception=1
// semantics, main.cpp:60-71).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ksgf.h"

namespace ksgf::amr {

// a request that is valid in the reference but not provided on this path
// (the C entry points of the host layer return KSGF_ERR_UNSUPPORTED for it)
struct UnsupportedError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(int rc, const char* what)
{
    if (rc < 0) throw std::runtime_error(std::string(what) + ": " + ksgf_last_error());
}

enum class IndexType { Cell = KSGF_CELL, XFace = KSGF_XFACE, YFace = KSGF_YFACE, ZFace = KSGF_ZFACE };

struct Box {
    std::array<int, 3> lo{}, hi{};
    long long numPts() const
    {
        return (long long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
    }
};

struct Geometry {
    Box domain;
    std::array<double, 3> prob_lo{}, prob_hi{};
    std::array<int, 3> is_periodic{};
    double CellSize(int d) const { return (prob_hi[d] - prob_lo[d]) / (domain.hi[d] - domain.lo[d] + 1); }
    bool isPeriodic(int d) const { return is_periodic[d] != 0; }
    ksgf_geometry c() const
    {
        ksgf_geometry g{};
        for (int d = 0; d < 3; ++d) {
            g.domain_lo[d] = domain.lo[d];
            g.domain_hi[d] = domain.hi[d];
            g.prob_lo[d] = prob_lo[d];
            g.prob_hi[d] = prob_hi[d];
            g.is_periodic[d] = is_periodic[d];
        }
        return g;
    }
};

using BCRec = ksgf_bcrec;

// BoxArray bound to its Geometry (the library's device box tables live with it)
class BoxArray {
public:
    // covers_domain = false: a refined AMR level (its boxes cover part of
    // the domain)
    BoxArray(const Geometry& geom, std::vector<Box> boxes, bool covers_domain = true)
        : geom_(geom), boxes_(std::move(boxes))
    {
        std::vector<int> flat;
        for (const auto& b : boxes_)
            for (int v : {b.lo[0], b.lo[1], b.lo[2], b.hi[0], b.hi[1], b.hi[2]}) flat.push_back(v);
        ksgf_geometry g = geom_.c();
        ksgf_boxarray* h = nullptr;
        if (covers_domain)
            check(ksgf_boxarray_create(&g, flat.data(), (int)boxes_.size(), &h), "BoxArray");
        else
            check(ksgf_boxarray_create_level(&g, flat.data(), (int)boxes_.size(), 0, &h), "BoxArray");
        h_ = std::shared_ptr<ksgf_boxarray>(h, [](ksgf_boxarray* p) { ksgf_boxarray_destroy(p); });
    }
    int size() const { return (int)boxes_.size(); }
    const Box& operator[](int b) const { return boxes_[b]; }
    const Geometry& geom() const { return geom_; }
    ksgf_boxarray* handle() const { return h_.get(); }

private:
    Geometry geom_;
    std::vector<Box> boxes_;
    std::shared_ptr<ksgf_boxarray> h_;
};

// Device-resident MultiFab in the ksgf layout (one FAB per box, ghosts
// included, Fortran order, component slowest).
class MultiFab {
public:
    MultiFab(const BoxArray& ba, int ncomp, int ngrow, IndexType it = IndexType::Cell)
        : ba_(ba), ncomp_(ncomp), ngrow_(ngrow), it_(it)
    {
        n_ = ksgf_multifab_size(ba.handle(), (int)it, ncomp, ngrow);
        if (n_ < 0) throw std::runtime_error("MultiFab: bad layout");
        if (hipMalloc(&p_, sizeof(double) * (size_t)n_) != hipSuccess) throw std::runtime_error("MultiFab: hipMalloc");
        if (hipMemset(p_, 0, sizeof(double) * (size_t)n_) != hipSuccess) throw std::runtime_error("MultiFab: hipMemset");
    }
    ~MultiFab()
    {
        if (p_) (void)hipFree(p_);
    }
    MultiFab(const MultiFab&) = delete;
    MultiFab& operator=(const MultiFab&) = delete;
    MultiFab(MultiFab&& o) noexcept : ba_(o.ba_), ncomp_(o.ncomp_), ngrow_(o.ngrow_), it_(o.it_), n_(o.n_), p_(o.p_)
    {
        o.p_ = nullptr;
        o.n_ = 0;
    }

    double* data() { return p_; }
    const double* data() const { return p_; }
    long long size() const { return n_; }
    int nComp() const { return ncomp_; }
    int nGrow() const { return ngrow_; }
    IndexType ixType() const { return it_; }
    const BoxArray& boxArray() const { return ba_; }

    void FillBoundary(hipStream_t st = nullptr)
    {
        check(ksgf_fill_boundary(ba_.handle(), p_, (int)it_, ncomp_, ngrow_, (ksgf_stream_t)st), "FillBoundary");
    }
    void FillPhysBC(const std::vector<BCRec>& bc, const std::vector<double>& bcval, hipStream_t st = nullptr)
    {
        check(ksgf_fill_physbc(ba_.handle(), p_, (int)it_, ncomp_, ngrow_, bc.data(), bcval.data(),
                               (ksgf_stream_t)st),
              "FillPhysBC");
    }
    // copy a domain-wide host array (ncomp x nz x ny x nx, x fastest) into
    // the valid region of every FAB
    void copyFromDomain(const double* host)
    {
        std::vector<double> h((size_t)n_, 0.0);
        const auto& dom = ba_.geom().domain;
        int N[3];
        for (int d = 0; d < 3; ++d) N[d] = dom.hi[d] - dom.lo[d] + 1 + ((int)it_ == d + 1 ? 1 : 0);
        long long off = 0;
        for (int b = 0; b < ba_.size(); ++b) {
            const Box& bx = ba_[b];
            int lo[3], n[3], vhi[3];
            for (int d = 0; d < 3; ++d) {
                int nodal = ((int)it_ == d + 1);
                lo[d] = bx.lo[d] - ngrow_;
                n[d] = bx.hi[d] - bx.lo[d] + 1 + 2 * ngrow_ + nodal;
                vhi[d] = bx.hi[d] + nodal;
            }
            long long cs = (long long)n[0] * n[1] * n[2];
            for (int c = 0; c < ncomp_; ++c)
                for (int k = bx.lo[2]; k <= vhi[2]; ++k)
                    for (int j = bx.lo[1]; j <= vhi[1]; ++j)
                        for (int i = bx.lo[0]; i <= vhi[0]; ++i)
                            h[off + c * cs + (i - lo[0]) + (long long)n[0] * ((j - lo[1]) + (long long)n[1] * (k - lo[2]))] =
                                host[(long long)c * N[0] * N[1] * N[2] + (i - dom.lo[0]) +
                                     (long long)N[0] * ((j - dom.lo[1]) + (long long)N[1] * (k - dom.lo[2]))];
            off += cs * ncomp_;
        }
        if (hipMemcpy(p_, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("MultiFab::copyFromDomain");
    }
    void copyToDomain(double* host) const
    {
        std::vector<double> h((size_t)n_);
        if (hipMemcpy(h.data(), p_, sizeof(double) * h.size(), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("MultiFab::copyToDomain");
        const auto& dom = ba_.geom().domain;
        int N[3];
        for (int d = 0; d < 3; ++d) N[d] = dom.hi[d] - dom.lo[d] + 1 + ((int)it_ == d + 1 ? 1 : 0);
        long long off = 0;
        for (int b = 0; b < ba_.size(); ++b) {
            const Box& bx = ba_[b];
            int lo[3], n[3], vhi[3];
            for (int d = 0; d < 3; ++d) {
                int nodal = ((int)it_ == d + 1);
                lo[d] = bx.lo[d] - ngrow_;
                n[d] = bx.hi[d] - bx.lo[d] + 1 + 2 * ngrow_ + nodal;
                vhi[d] = bx.hi[d] + nodal;
            }
            long long cs = (long long)n[0] * n[1] * n[2];
            for (int c = 0; c < ncomp_; ++c)
                for (int k = bx.lo[2]; k <= vhi[2]; ++k)
                    for (int j = bx.lo[1]; j <= vhi[1]; ++j)
                        for (int i = bx.lo[0]; i <= vhi[0]; ++i)
                            host[(long long)c * N[0] * N[1] * N[2] + (i - dom.lo[0]) +
                                 (long long)N[0] * ((j - dom.lo[1]) + (long long)N[1] * (k - dom.lo[2]))] =
                                h[off + c * cs + (i - lo[0]) + (long long)n[0] * ((j - lo[1]) + (long long)n[1] * (k - lo[2]))];
            off += cs * ncomp_;
        }
    }

private:
    const BoxArray& ba_;
    int ncomp_, ngrow_;
    IndexType it_;
    long long n_ = 0;
    double* p_ = nullptr;
};

// amrex::WriteMultiLevelPlotfile with one level (IOManager::write_plot_file,
// IOManager.cpp:181-187): the valid cells of a cell-centred MultiFab
inline void WriteSingleLevelPlotfile(const std::string& name, const MultiFab& mf,
                                     const std::vector<std::string>& varnames, double time, int step,
                                     hipStream_t st = nullptr)
{
    if ((int)varnames.size() != mf.nComp()) throw std::runtime_error("WriteSingleLevelPlotfile: one name per component");
    std::vector<const char*> names;
    for (const auto& v : varnames) names.push_back(v.c_str());
    check(ksgf_plotfile_write(name.c_str(), mf.boxArray().handle(), mf.data(), mf.nComp(), mf.nGrow(), names.data(),
                              time, step, (ksgf_stream_t)st),
          "WriteSingleLevelPlotfile");
}

// level 0 of a plotfile into the valid cells of mf
inline void ReadSingleLevelPlotfile(const std::string& name, MultiFab& mf, hipStream_t st = nullptr)
{
    check(ksgf_plotfile_read(name.c_str(), mf.boxArray().handle(), mf.data(), mf.nComp(), mf.nGrow(),
                             (ksgf_stream_t)st),
          "ReadSingleLevelPlotfile");
}

// amrex::VisMF::Write / Read of checkpoint fields (IOManager.cpp:214-219, 249-266)
struct VisMF {
    static void Write(const MultiFab& mf, const std::string& prefix, hipStream_t st = nullptr)
    {
        check(ksgf_vismf_write(prefix.c_str(), mf.boxArray().handle(), mf.data(), (int)mf.ixType(), mf.nComp(),
                               mf.nGrow(), (ksgf_stream_t)st),
              "VisMF::Write");
    }
    static void Read(MultiFab& mf, const std::string& prefix, hipStream_t st = nullptr)
    {
        check(ksgf_vismf_read(prefix.c_str(), mf.boxArray().handle(), mf.data(), (int)mf.ixType(), mf.nComp(),
                              mf.nGrow(), (ksgf_stream_t)st),
              "VisMF::Read");
    }
};

}  // namespace ksgf::amr

// ---------------------------------------------------------------------------
// HydroUtils (AMReX-Hydro) MultiFab-level calls
// ---------------------------------------------------------------------------
namespace HydroUtils {

using ksgf::amr::BCRec;
using ksgf::amr::Geometry;
using ksgf::amr::MultiFab;

// the advection_type string is "Godunov" on this path
inline void ExtrapVelToFaces(const MultiFab& vel, const MultiFab& vel_forces, MultiFab& u_mac, MultiFab& v_mac,
                             MultiFab& w_mac, const std::vector<BCRec>& h_bcrec, const BCRec* /*d_bcrec*/,
                             const Geometry& /*geom*/, double dt, bool use_ppm, bool use_forces_in_trans,
                             const std::string& /*advection_type*/, int limiter_type, bool allow_inflow_on_outflow,
                             hipStream_t st = nullptr)
{
    ksgf::amr::check(ksgf_extrap_vel_to_faces(vel.boxArray().handle(), vel.data(), vel_forces.data(), u_mac.data(),
                                              v_mac.data(), w_mac.data(), h_bcrec.data(), dt, use_ppm,
                                              use_forces_in_trans, limiter_type, allow_inflow_on_outflow,
                                              (ksgf_stream_t)st),
                     "ExtrapVelToFaces");
}

inline void ComputeFluxesOnBoxFromState(int ncomp, const MultiFab& q, const MultiFab& q_on_face_bc,
                                        MultiFab& flux_x, MultiFab& flux_y, MultiFab& flux_z, MultiFab& face_x,
                                        MultiFab& face_y, MultiFab& face_z, bool known_edge_state,
                                        const MultiFab& umac, const MultiFab& vmac, const MultiFab& wmac,
                                        const MultiFab* divu, const MultiFab& fq, const Geometry& /*geom*/,
                                        double dt, const std::vector<BCRec>& h_bcrec, const int* iconserv,
                                        bool use_ppm, bool use_forces_in_trans, bool is_velocity,
                                        bool fluxes_are_area_weighted, const std::string& /*advection_type*/,
                                        int limiter_type, bool allow_inflow_on_outflow, hipStream_t st = nullptr)
{
    ksgf::amr::check(ksgf_compute_fluxes_from_state(
                         q.boxArray().handle(), ncomp, q.data(), q_on_face_bc.data(), flux_x.data(), flux_y.data(),
                         flux_z.data(), face_x.data(), face_y.data(), face_z.data(), known_edge_state, umac.data(),
                         vmac.data(), wmac.data(), divu ? divu->data() : nullptr, fq.data(), dt, h_bcrec.data(),
                         iconserv, use_ppm, use_forces_in_trans, is_velocity, fluxes_are_area_weighted,
                         limiter_type, allow_inflow_on_outflow, (ksgf_stream_t)st),
                     "ComputeFluxesOnBoxFromState");
}

inline void ComputeDivergence(MultiFab& div, const MultiFab& fx, const MultiFab& fy, const MultiFab& fz, int ncomp,
                              const Geometry& /*geom*/, double mult, bool fluxes_are_area_weighted,
                              hipStream_t st = nullptr)
{
    ksgf::amr::check(ksgf_compute_divergence(div.boxArray().handle(), div.data(), fx.data(), fy.data(), fz.data(),
                                             ncomp, mult, fluxes_are_area_weighted, (ksgf_stream_t)st),
                     "ComputeDivergence");
}

}  // namespace HydroUtils

// ---------------------------------------------------------------------------
// Hydro::MacProjector
// ---------------------------------------------------------------------------
namespace ksgf::amr {

// amrex::average_cellcenter_to_face (one component)
inline void average_cellcenter_to_face(const std::array<MultiFab*, 3>& faces, const MultiFab& cc,
                                       hipStream_t st = nullptr)
{
    check(ksgf_average_cellcenter_to_face(cc.boxArray().handle(), cc.data(), cc.nGrow(), faces[0]->data(),
                                          faces[1]->data(), faces[2]->data(), faces[0]->nGrow(), (ksgf_stream_t)st),
          "average_cellcenter_to_face");
}

// MacProjOp's variable-density coefficients (icns_advection.cpp:255-283):
// rho_face = average_cellcenter_to_face(rho), then invert(factor): beta on
// the faces = factor / rho_face
inline void face_beta_from_density(const std::array<MultiFab*, 3>& beta, const MultiFab& rho, double factor,
                                   hipStream_t st = nullptr)
{
    average_cellcenter_to_face(beta, rho, st);
    for (int d = 0; d < 3; ++d)
        check(ksgf_invert(rho.boxArray().handle(), beta[d]->data(), KSGF_XFACE + d, 1, beta[d]->nGrow(), 0, factor,
                          (ksgf_stream_t)st),
              "invert");
}

// diffusion::average_velocity_eta_to_faces (incflo_diffusion.cpp:108-128):
// the cell viscosity (1 comp, a filled ghost layer) averaged to the faces,
// the adjacent cell on a non-periodic domain face (fixup_eta_on_domain_faces)
inline std::array<MultiFab, 3> average_velocity_eta_to_faces(const MultiFab& cc_eta, hipStream_t st = nullptr)
{
    const BoxArray& ba = cc_eta.boxArray();
    std::array<MultiFab, 3> r{MultiFab(ba, 1, 0, IndexType::XFace), MultiFab(ba, 1, 0, IndexType::YFace),
                              MultiFab(ba, 1, 0, IndexType::ZFace)};
    check(ksgf_average_velocity_eta_to_faces(ba.handle(), cc_eta.data(), cc_eta.nGrow(), r[0].data(), r[1].data(),
                                             r[2].data(), 0, (ksgf_stream_t)st),
          "average_velocity_eta_to_faces");
    return r;
}

}  // namespace ksgf::amr

namespace Hydro {

enum class LinOpBCType { Periodic = KSGF_LO_PERIODIC, Neumann = KSGF_LO_NEUMANN, Dirichlet = KSGF_LO_DIRICHLET };

// amrex::LPInfo knobs used on this path
struct LPInfo {
    bool do_semicoarsening = false;
    int max_coarsening_level = 30;
};

// what io::print_mlmg_info reads (console_io.cpp:178-186)
struct MLMG {
    ksgf_mlmg_options opts{};
    ksgf_mlmg_info last{};
    MLMG() { ksgf_mlmg_default_options(&opts); }
    void setMaxIter(int n) { opts.maxiter = n; }
    void setPreSmooth(int n) { opts.num_pre_smooth = n; }
    void setPostSmooth(int n) { opts.num_post_smooth = n; }
    void setFinalSmooth(int n) { opts.num_final_smooth = n; }
    void setBottomSmooth(int n) { opts.num_bottom_smooth = n; }
    void setBottomTolerance(double t) { opts.bottom_rtol = t; }
    void setBottomToleranceAbs(double t) { opts.bottom_atol = t; }
    void setBottomMaxIter(int n) { opts.bottom_maxiter = n; }
    void setMixedPrecision(bool on) { opts.precision = on ? KSGF_MG_MIXED : KSGF_MG_FP64; }
    int getNumIters() const { return last.iters; }
    double getInitResidual() const { return last.r0; }
    double getFinalResidual() const { return last.rfinal; }
};

class MacProjector {
public:
    // one level, or two (level 1 at ratio 2: the composite solve)
    explicit MacProjector(const std::vector<ksgf::amr::Geometry>& geom) : geom_(geom)
    {
        if (geom_.size() != 1 && geom_.size() != 2)
            throw std::runtime_error("MacProjector: one or two levels on this path");
    }
    ~MacProjector()
    {
        if (mp_) ksgf_macproj_destroy(mp_);
        if (mpl_) ksgf_macproj_levels_destroy(mpl_);
    }
    void initProjector(const std::vector<const ksgf::amr::BoxArray*>& ba, const LPInfo& info, double beta)
    {
        if (ba.size() != geom_.size()) throw std::runtime_error("MacProjector: one BoxArray per level");
        ba_ = ba.at(0);
        fba_ = (ba.size() > 1) ? ba[1] : nullptr;
        beta_ = beta;
        mlmg_.opts.do_semicoarsening = info.do_semicoarsening ? 1 : 0;
        mlmg_.opts.max_coarsening_level = info.max_coarsening_level;
    }
    // initProjector(LPInfo, Vector<Array<const MultiFab*,3>> beta) (one
    // level): beta on the faces; the reference passes factor / rho_face
    // when density varies (icns_advection.cpp:143-154, 234-291)
    void initProjector(const std::vector<const ksgf::amr::BoxArray*>& ba, const LPInfo& info,
                       const std::array<const ksgf::amr::MultiFab*, 3>& beta)
    {
        if (ba.size() != 1 || geom_.size() != 1)
            throw ksgf::amr::UnsupportedError("MacProjector: face coefficients on one level only");
        initProjector(ba, info, 1.0);
        beta_faces_ = beta;
    }
    void setDomainBC(const std::array<LinOpBCType, 3>& lo, const std::array<LinOpBCType, 3>& hi)
    {
        for (int d = 0; d < 3; ++d) {
            lo_[d] = (int)lo[d];
            hi_[d] = (int)hi[d];
        }
        build();
    }
    void updateBeta(double beta)
    {
        beta_ = beta;
        beta_faces_ = {};
        if (mp_) ksgf::amr::check(ksgf_macproj_set_beta(mp_, beta), "updateBeta");
        if (mpl_) build();
    }
    // updateBeta(Vector<Array<const MultiFab*,3>>) (icns_advection.cpp:289),
    // enqueued on st
    void updateBeta(const std::array<const ksgf::amr::MultiFab*, 3>& beta, hipStream_t st = nullptr)
    {
        if (mpl_ || fba_) throw ksgf::amr::UnsupportedError("MacProjector: face coefficients on one level only");
        beta_faces_ = beta;
        if (!mp_) build();
        setFaces(st);
    }
    void setUMAC(const std::vector<std::array<ksgf::amr::MultiFab*, 3>>& umac)
    {
        if (umac.size() != geom_.size()) throw std::runtime_error("MacProjector::setUMAC: one set per level");
        umac_ = umac.at(0);
        if (umac.size() > 1) umacf_ = umac[1];
    }
    void project(double rtol, double atol, hipStream_t st = nullptr)
    {
        if (!mp_ && !mpl_) build();
        int rc;
        if (mpl_) {
            double* c[3] = {umac_[0]->data(), umac_[1]->data(), umac_[2]->data()};
            double* f[3] = {umacf_[0]->data(), umacf_[1]->data(), umacf_[2]->data()};
            if (umacf_[0]->nGrow() != umac_[0]->nGrow())
                throw std::runtime_error("MacProjector::project: both levels' face MultiFabs need one ngrow");
            rc = ksgf_macproj_levels_project(mpl_, c, f, umac_[0]->nGrow(), rtol, atol, &mlmg_.last,
                                             (ksgf_stream_t)st);
        } else {
            rc = ksgf_macproj_project(mp_, umac_[0]->data(), umac_[1]->data(), umac_[2]->data(), rtol, atol,
                                      &mlmg_.last, (ksgf_stream_t)st);
        }
        ksgf::amr::check(rc, "MacProjector::project");  // KSGF_ERR_NOT_CONVERGED (> 0) is not an abort
    }
    // project split for a step captured into a hipGraph: begin (div u_mac and
    // `ncycles` iterations tested on the device; capturable), finish (reads
    // the solve's state, one host synchronisation, and completes an
    // unfinished solve), correct (u_mac -= beta grad phi; capturable).  One
    // level, one beta.
    void projectBegin(double rtol, double atol, int ncycles, hipStream_t st = nullptr)
    {
        if (mpl_ || fba_ || beta_faces_[0]) throw ksgf::amr::UnsupportedError("MacProjector::projectBegin: one level, one beta");
        if (!mp_) build();
        ksgf::amr::check(ksgf_macproj_project_begin(mp_, umac_[0]->data(), umac_[1]->data(), umac_[2]->data(), rtol,
                                                    atol, ncycles, (ksgf_stream_t)st),
                         "MacProjector::projectBegin");
    }
    void projectFinish(hipStream_t st = nullptr)
    {
        if (!mp_) throw std::runtime_error("MacProjector::projectFinish: no projectBegin");
        ksgf::amr::check(ksgf_macproj_project_finish(mp_, &mlmg_.last, (ksgf_stream_t)st), "MacProjector::projectFinish");
    }
    void projectCorrect(hipStream_t st = nullptr)
    {
        if (!mp_) throw std::runtime_error("MacProjector::projectCorrect: no projectBegin");
        ksgf::amr::check(ksgf_macproj_project_correct(mp_, umac_[0]->data(), umac_[1]->data(), umac_[2]->data(),
                                                      (ksgf_stream_t)st),
                         "MacProjector::projectCorrect");
    }
    MLMG& getMLMG() { return mlmg_; }
    void rebuild() { build(); }
    // the last solution phi, domain-wide (n0*n1*n2 doubles, device), as
    // project(Vector<MultiFab*> phi, ...) hands it back
    void getPhi(double* phi, hipStream_t st = nullptr)
    {
        if (mpl_) throw std::runtime_error("MacProjector::getPhi: two levels hand back phi per level (getPhiLevels)");
        if (!mp_) throw std::runtime_error("MacProjector::getPhi: no projection yet");
        ksgf::amr::check(ksgf_macproj_get_phi(mp_, phi, (ksgf_stream_t)st), "MacProjector::getPhi");
    }

    // two levels: phi_c domain-wide on level 0 (n cells, device), phi_f a
    // level-1 cell MultiFab with one ghost layer
    void getPhiLevels(double* phi_c, ksgf::amr::MultiFab& phi_f, hipStream_t st = nullptr)
    {
        if (!mpl_) throw std::runtime_error("MacProjector::getPhiLevels: no two-level projection yet");
        ksgf::amr::check(ksgf_macproj_levels_get_phi(mpl_, phi_c, phi_f.data(), phi_f.nGrow(), (ksgf_stream_t)st),
                         "MacProjector::getPhiLevels");
    }

private:
    void build()
    {
        if (!ba_) throw std::runtime_error("MacProjector: initProjector first");
        if (mp_) ksgf_macproj_destroy(mp_);
        if (mpl_) ksgf_macproj_levels_destroy(mpl_);
        mp_ = nullptr;
        mpl_ = nullptr;
        if (fba_)
            ksgf::amr::check(ksgf_macproj_levels_create(ba_->handle(), fba_->handle(), lo_, hi_, beta_, &mlmg_.opts,
                                                        &mpl_),
                             "MacProjector");
        else
            ksgf::amr::check(ksgf_macproj_create(ba_->handle(), lo_, hi_, beta_, &mlmg_.opts, &mp_), "MacProjector");
        if (mp_ && beta_faces_[0]) setFaces(nullptr);
    }
    void setFaces(hipStream_t st)
    {
        const auto& b = beta_faces_;
        if (!b[0] || !b[1] || !b[2]) throw std::runtime_error("MacProjector: three face coefficient MultiFabs");
        if (b[0]->nGrow() != b[1]->nGrow() || b[0]->nGrow() != b[2]->nGrow())
            throw std::runtime_error("MacProjector: face coefficients need one ngrow");
        ksgf::amr::check(ksgf_macproj_set_beta_faces(mp_, b[0]->data(), b[1]->data(), b[2]->data(), b[0]->nGrow(),
                                                     (ksgf_stream_t)st),
                         "MacProjector::updateBeta(faces)");
    }
    std::array<const ksgf::amr::MultiFab*, 3> beta_faces_{};
    std::vector<ksgf::amr::Geometry> geom_;
    const ksgf::amr::BoxArray* ba_ = nullptr;
    const ksgf::amr::BoxArray* fba_ = nullptr;
    double beta_ = 1.0;
    int lo_[3] = {0, 0, 0}, hi_[3] = {0, 0, 0};
    std::array<ksgf::amr::MultiFab*, 3> umac_{}, umacf_{};
    MLMG mlmg_;
    ksgf_macproj* mp_ = nullptr;
    ksgf_macproj_levels* mpl_ = nullptr;
};

// amrex::MLABecLaplacian with constant coefficients and its MLMG, one
// component at a time (the implicit diffusion solve of
// DiffSolverIface::linsys_solve, DiffusionOps.cpp:154-160, and the apply of
// compute_diff_term): setScalars(alpha, beta), setACoeffs(a),
// setBCoeffs(b), setDomainBC, solve(sol, rhs, comp), apply(out, in, comp).
class MLABecLaplacian {
public:
    MLABecLaplacian(const ksgf::amr::BoxArray& ba, const LPInfo& info, const ksgf_mlmg_options& opts)
        : ba_(&ba), opts_(opts)
    {
        opts_.do_semicoarsening = info.do_semicoarsening ? 1 : 0;
        opts_.max_coarsening_level = info.max_coarsening_level;
    }
    ~MLABecLaplacian()
    {
        if (ab_) ksgf_abeclap_destroy(ab_);
    }
    MLABecLaplacian(const MLABecLaplacian&) = delete;
    MLABecLaplacian& operator=(const MLABecLaplacian&) = delete;

    void setScalars(double alpha, double beta)
    {
        alpha_ = alpha;
        beta_ = beta;
        if (ab_) ksgf::amr::check(ksgf_abeclap_set_scalars(ab_, alpha, beta), "MLABecLaplacian::setScalars");
    }
    void setACoeffs(double a) { a_ = a, rebuild_ = true; }
    void setBCoeffs(double b) { b_ = b, rebuild_ = true; }
    void setDomainBC(const std::array<LinOpBCType, 3>& lo, const std::array<LinOpBCType, 3>& hi)
    {
        for (int d = 0; d < 3; ++d) {
            lo_[d] = (int)lo[d];
            hi_[d] = (int)hi[d];
        }
        rebuild_ = true;
    }
    // MLMG::solve for component comp of sol (rhs: same component of rhs)
    ksgf_mlmg_info solve(ksgf::amr::MultiFab& sol, const ksgf::amr::MultiFab& rhs, int comp, double rtol,
                         double atol, hipStream_t st = nullptr)
    {
        build();
        ksgf_mlmg_info info{};
        ksgf::amr::check(ksgf_abeclap_solve(ab_, sol.data(), sol.nGrow(), rhs.data(), rhs.nGrow(), sol.nComp(), comp,
                                            rtol, atol, &info, (ksgf_stream_t)st),
                         "MLABecLaplacian::solve");
        return info;
    }
    // MLMG::apply for component comp
    void apply(ksgf::amr::MultiFab& out, const ksgf::amr::MultiFab& in, int comp, hipStream_t st = nullptr)
    {
        build();
        ksgf::amr::check(ksgf_abeclap_apply(ab_, out.data(), out.nGrow(), in.data(), in.nGrow(), in.nComp(), comp,
                                            (ksgf_stream_t)st),
                         "MLABecLaplacian::apply");
    }

private:
    void build()
    {
        if (ab_ && !rebuild_) return;
        if (ab_) ksgf_abeclap_destroy(ab_);
        ab_ = nullptr;
        ksgf::amr::check(ksgf_abeclap_create(ba_->handle(), lo_, hi_, alpha_, a_, beta_, b_, &opts_, &ab_),
                         "MLABecLaplacian");
        rebuild_ = false;
    }
    const ksgf::amr::BoxArray* ba_;
    ksgf_mlmg_options opts_;
    double alpha_ = 1.0, beta_ = 1.0, a_ = 1.0, b_ = 1.0;
    int lo_[3] = {0, 0, 0}, hi_[3] = {0, 0, 0};
    bool rebuild_ = true;
    ksgf_abeclap* ab_ = nullptr;
};

// amrex::MLTensorOp + MLMG as ICNSDiffTensorOp drives it (icns_diffusion.H:
// 18-50; DiffSolverIface::setup_operator and set_bcoeffs, DiffusionOps.cpp:
// 57-110, DiffusionOps.H:54-71): setDomainBC with one LinOpBCType triple per
// velocity component (get_diffuse_tensor_bc), setScalars, setACoeffs(rho),
// setShearViscosity(eta on faces), then MLMG::apply (compute_diff_term) or
// MLMG::solve (linsys_solve) on the three velocity components.
class MLTensorOp {
public:
    MLTensorOp(const ksgf::amr::BoxArray& ba, const LPInfo& info, const ksgf_mlmg_options& opts)
        : ba_(&ba), opts_(opts)
    {
        opts_.do_semicoarsening = info.do_semicoarsening ? 1 : 0;
        opts_.max_coarsening_level = info.max_coarsening_level;
    }
    ~MLTensorOp()
    {
        if (op_) ksgf_tensorop_destroy(op_);
    }
    MLTensorOp(const MLTensorOp&) = delete;
    MLTensorOp& operator=(const MLTensorOp&) = delete;

    void setDomainBC(const std::array<std::array<LinOpBCType, 3>, 3>& lo,
                     const std::array<std::array<LinOpBCType, 3>, 3>& hi)
    {
        int l[9], h[9];
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d) {
                l[c * 3 + d] = (int)lo[c][d];
                h[c * 3 + d] = (int)hi[c][d];
            }
        if (op_) ksgf_tensorop_destroy(op_);
        op_ = nullptr;
        ksgf::amr::check(ksgf_tensorop_create(ba_->handle(), l, h, &opts_, &op_), "MLTensorOp::setDomainBC");
    }
    void setScalars(double alpha, double beta)
    {
        ksgf::amr::check(ksgf_tensorop_set_scalars(need(), alpha, beta), "MLTensorOp::setScalars");
    }
    void setACoeffs(const ksgf::amr::MultiFab& a, hipStream_t st = nullptr)
    {
        ksgf::amr::check(ksgf_tensorop_set_acoeffs(need(), a.data(), a.nGrow(), (ksgf_stream_t)st),
                         "MLTensorOp::setACoeffs");
    }
    void setShearViscosity(const std::array<const ksgf::amr::MultiFab*, 3>& eta, hipStream_t st = nullptr)
    {
        ksgf::amr::check(ksgf_tensorop_set_shear_viscosity(need(), eta[0]->data(), eta[1]->data(), eta[2]->data(),
                                                           eta[0]->nGrow(), (ksgf_stream_t)st),
                         "MLTensorOp::setShearViscosity");
    }
    ksgf_mlmg_info solve(ksgf::amr::MultiFab& sol, const ksgf::amr::MultiFab& rhs, double rtol, double atol,
                         hipStream_t st = nullptr)
    {
        ksgf_mlmg_info info{};
        ksgf::amr::check(ksgf_tensorop_solve(need(), sol.data(), sol.nGrow(), rhs.data(), rhs.nGrow(), rtol, atol,
                                             &info, (ksgf_stream_t)st),
                         "MLTensorOp::solve");
        return info;
    }
    void apply(ksgf::amr::MultiFab& out, const ksgf::amr::MultiFab& in, hipStream_t st = nullptr)
    {
        ksgf::amr::check(ksgf_tensorop_apply(need(), out.data(), out.nGrow(), in.data(), in.nGrow(),
                                             (ksgf_stream_t)st),
                         "MLTensorOp::apply");
    }

private:
    ksgf_tensorop* need() const
    {
        if (!op_) throw std::runtime_error("MLTensorOp: setDomainBC first");
        return op_;
    }
    const ksgf::amr::BoxArray* ba_;
    ksgf_mlmg_options opts_;
    ksgf_tensorop* op_ = nullptr;
};

// Hydro::NodalProjector with a constant sigma, as incflo::ApplyProjection
// builds it for constant density (incflo_apply_nodal_projection.cpp:340-353):
// NodalProjector(vel, scaling_factor / rho_0, geom, lpinfo), setDomainBC,
// the MLMG options, project(rtol, atol), getPhi, getGradPhi.  vel is held by
// pointer (the reference keeps non-owning pointers too); phi is a domain-wide
// node array (N_d = n_d along a periodic direction, n_d + 1 otherwise).
class NodalProjector {
public:
    NodalProjector(const std::vector<ksgf::amr::MultiFab*>& vel, double sigma,
                   const std::vector<ksgf::amr::Geometry>& geom, const LPInfo& info)
        : vel_(vel.at(0)), sigma_(sigma)
    {
        if (geom.size() != 1 || vel.size() != 1) throw std::runtime_error("NodalProjector: one level on this path");
        ksgf_mlmg_default_options(&mlmg_.opts);
        mlmg_.opts.do_semicoarsening = info.do_semicoarsening ? 1 : 0;
        mlmg_.opts.max_coarsening_level = info.max_coarsening_level;
    }
    ~NodalProjector()
    {
        if (np_) ksgf_nodalproj_destroy(np_);
    }
    NodalProjector(const NodalProjector&) = delete;
    NodalProjector& operator=(const NodalProjector&) = delete;

    void setDomainBC(const std::array<LinOpBCType, 3>& lo, const std::array<LinOpBCType, 3>& hi)
    {
        for (int d = 0; d < 3; ++d) {
            lo_[d] = (int)lo[d];
            hi_[d] = (int)hi[d];
        }
        build();
    }
    void setSigma(double sigma)
    {
        sigma_ = sigma;
        if (np_) ksgf::amr::check(ksgf_nodalproj_set_sigma(np_, sigma), "NodalProjector::setSigma");
    }
    // gradphi: cell MultiFab, 3 comp (may be null: not kept)
    void project(double rtol, double atol, ksgf::amr::MultiFab* gradphi = nullptr, hipStream_t st = nullptr)
    {
        if (!np_) build();
        int rc = ksgf_nodalproj_project(np_, vel_->data(), vel_->nGrow(), gradphi ? gradphi->data() : nullptr,
                                        gradphi ? gradphi->nGrow() : 0, rtol, atol, &mlmg_.last, (ksgf_stream_t)st);
        ksgf::amr::check(rc, "NodalProjector::project");
    }
    MLMG& getMLMG() { return mlmg_; }
    // point the projector at another velocity MultiFab on the same BoxArray
    // (the reference builds a new NodalProjector per projection; this keeps
    // the hierarchy and re-targets it)
    void setVelocity(ksgf::amr::MultiFab* vel)
    {
        if (vel->boxArray().handle() != vel_->boxArray().handle())
            throw std::runtime_error("NodalProjector: other BoxArray");
        vel_ = vel;
    }
    void getPhi(double* phi, hipStream_t st = nullptr)
    {
        if (!np_) throw std::runtime_error("NodalProjector::getPhi: no projection yet");
        ksgf::amr::check(ksgf_nodalproj_get_phi(np_, phi, (ksgf_stream_t)st), "NodalProjector::getPhi");
    }
    std::array<int, 3> nodeDims() const
    {
        std::array<int, 3> n{};
        if (np_) ksgf::amr::check(ksgf_nodalproj_node_dims(np_, n.data()), "NodalProjector::nodeDims");
        return n;
    }

private:
    void build()
    {
        if (np_) ksgf_nodalproj_destroy(np_);
        np_ = nullptr;
        ksgf::amr::check(ksgf_nodalproj_create(vel_->boxArray().handle(), lo_, hi_, sigma_, &mlmg_.opts, &np_),
                         "NodalProjector");
    }
    ksgf::amr::MultiFab* vel_;
    double sigma_;
    int lo_[3] = {0, 0, 0}, hi_[3] = {0, 0, 0};
    MLMG mlmg_;
    ksgf_nodalproj* np_ = nullptr;
};

}  // namespace Hydro
