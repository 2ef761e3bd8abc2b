/*
 * ksgf.h -- C ABI of the MI355X (gfx950) implementation of Kynema-SGF's
 * per-step Godunov PPM advection + MAC projection hot path.
 *
 * Everything here is plain C: integers, doubles, raw device pointers and
 * opaque handles; every entry point returns KSGF_OK (0) or a negative error
 * code and never throws.  Device buffers are owned by the caller (as
 * FieldRepo owns MultiFabs, src/core/FieldRepo.H:139-351); the library only
 * allocates its own internal scratch (box tables, per-stage temporaries, the
 * multigrid hierarchy).  Work is enqueued on the caller's hipStream_t;
 * functions that return a host value (CFL number, solver statistics)
 * synchronise that stream.
 *
 * Data layout.  A MultiFab is one flat array of doubles holding one FAB per
 * box of a BoxArray, in box order, without padding.  FAB b covers box b
 * converted to the MultiFab's index type (cell centred, or nodal in one
 * direction for x/y/z faces) and grown by ngrow in every direction; it is
 * stored in Fortran order (i fastest) with the component index slowest --
 * the amrex::Array4 convention.  ksgf_multifab_size() gives the length.
 *
 * The MultiFab-level calls replace the reference's per-box loops: one call
 * processes every box of the BoxArray in a single launch over a ragged box
 * table (the MFIter loops at src/equation_systems/icns/icns_advection.H:349-374
 * and :410-440 become one call each).
 */
#ifndef KSGF_H
#define KSGF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* ksgf_stream_t; /* a hipStream_t; NULL = default stream */

/* status codes */
#define KSGF_OK 0
#define KSGF_ERR_INVALID -1     /* bad argument (null pointer, bad enum, bad box) */
#define KSGF_ERR_UNSUPPORTED -2 /* valid in the reference, not provided here (PLM, WENO) */
#define KSGF_ERR_HIP -3         /* a HIP runtime call failed */
#define KSGF_ERR_NOMEM -4       /* device allocation failed */
#define KSGF_ERR_NOT_CONVERGED 1 /* MLMG stopped at maxiter (not an abort, do_fixed_iters) */
#define KSGF_ERR_DIVERGED -5     /* MLMG residual above 1e20 |rhs| or not finite (AMReX aborts) */

/* index types of a MultiFab */
#define KSGF_CELL 0
#define KSGF_XFACE 1
#define KSGF_YFACE 2
#define KSGF_ZFACE 3

/* amrex::BCType values used on this path (BCInterface.cpp:200-260) */
#define KSGF_BC_INT_DIR 0
#define KSGF_BC_FOEXTRAP 2
#define KSGF_BC_EXT_DIR 3
#define KSGF_BC_HOEXTRAP 4

/* PPM limiter_type (icns_advection.H:195-204: ppm -> default_limiter) */
#define KSGF_PPM_NOLIMITER 0
#define KSGF_PPM_VANLEER 1 /* PPM::default_limiter */
#define KSGF_PPM_WENOZ 2   /* not provided: KSGF_ERR_UNSUPPORTED */
#define KSGF_PPM_WENO_JS 3 /* not provided: KSGF_ERR_UNSUPPORTED */
#define KSGF_PPM_UPWIND 4
#define KSGF_PPM_MINMOD 5

/* linear-operator domain BCs (icns_advection.cpp:21-41) */
#define KSGF_LO_PERIODIC 0
#define KSGF_LO_NEUMANN 1
#define KSGF_LO_DIRICHLET 2

/* multigrid precision */
#define KSGF_MG_FP64 0
#define KSGF_MG_MIXED 1 /* FP32 V-cycles inside an FP64 defect-correction loop */

typedef struct {
    int domain_lo[3], domain_hi[3]; /* cell-centred, inclusive */
    double prob_lo[3], prob_hi[3];
    int is_periodic[3];
} ksgf_geometry;

typedef struct {
    int lo[3];
    int hi[3];
} ksgf_bcrec; /* one per component, amrex::BCRec */

typedef struct {
    int maxiter;          /* mac_proj.maxiter (200) */
    int num_pre_smooth;   /* 2 */
    int num_post_smooth;  /* 2 */
    int num_final_smooth; /* 8: smooths after a failed bottom solve */
    int num_bottom_smooth;/* 0: smooths after a converged bottom solve */
    int bottom_maxiter;   /* 200 */
    double bottom_rtol;   /* eps*1e12 */
    double bottom_atol;   /* eps*1e4 */
    int do_semicoarsening;
    int max_coarsening_level; /* 30 */
    int precision;        /* KSGF_MG_FP64 or KSGF_MG_MIXED */
} ksgf_mlmg_options;

typedef struct {
    int iters;        /* V-cycles (mixed: total inner FP32 V-cycles) */
    int outer_iters;  /* FP64 defect-correction iterations (mixed only) */
    double r0;        /* initial residual, inf norm */
    double rfinal;    /* final residual, inf norm */
    int converged;
    int nlevels;
    int bottom_iters; /* BiCGStab iterations summed over V-cycles */
    int bottom_fail;
} ksgf_mlmg_info; /* what io::print_mlmg_info prints (console_io.cpp:178-186) */

typedef struct ksgf_boxarray ksgf_boxarray;
typedef struct ksgf_macproj ksgf_macproj;

const char* ksgf_last_error(void);
const char* ksgf_version(void);

/* BoxArray + Geometry + DistributionMapping on one GPU.
 * boxes: nbox*6 ints (lo0 lo1 lo2 hi0 hi1 hi2), disjoint, covering the domain.
 * Replaces amrex::BoxArray / Geometry as consumed by MultiFab construction. */
int ksgf_boxarray_create(const ksgf_geometry* geom, const int* boxes, int nbox,
                         ksgf_boxarray** out);
/* A BoxArray of one AMR level: covers_domain = 0 for a refined level whose
 * boxes cover only part of its domain (disjoint boxes inside the domain). */
int ksgf_boxarray_create_level(const ksgf_geometry* geom, const int* boxes, int nbox,
                               int covers_domain, ksgf_boxarray** out);
int ksgf_boxarray_destroy(ksgf_boxarray* ba);
long long ksgf_multifab_size(const ksgf_boxarray* ba, int itype, int ncomp, int ngrow);

/* MultiFab::FillBoundary(periodicity): interior and periodic ghosts
 * (icns_advection.H:239-243). */
int ksgf_fill_boundary(const ksgf_boxarray* ba, double* mf, int itype, int ncomp, int ngrow,
                       ksgf_stream_t stream);

/* Physical-boundary ghost fill (the BC part of FillPatch / fillpatch_sibling_fields,
 * FieldFillPatchOps.H:250-336).  bc: ncomp BCRecs; bcval: [6 orientations
 * xlo,ylo,zlo,xhi,yhi,zhi][ncomp] Dirichlet values (host memory). */
int ksgf_fill_physbc(const ksgf_boxarray* ba, double* mf, int itype, int ncomp, int ngrow,
                     const ksgf_bcrec* bc, const double* bcval, ksgf_stream_t stream);

/* MacProjOp::set_inflow_velocity (icns_advection.cpp:157-183): on ext_dir
 * boundary faces set the normal MAC velocity to the boundary value. */
int ksgf_set_inflow_velocity(const ksgf_boxarray* ba, double* umac, double* vmac, double* wmac,
                             int ngrow_mac, const ksgf_bcrec* bc, const double* bcval,
                             ksgf_stream_t stream);

/* PPM::PredictStateOn{X,Y,Z}Face (test_mflux_schemes.cpp:87-151) for every
 * valid cell: s (cell, ncomp_s, ng_s >= 2), component comp; vel_edge (face of
 * direction dir, 1 comp, ng_v >= 0); Im/Ip (cell, 1 comp, ngrow 0). */
int ksgf_ppm_predict_state_on_face(const ksgf_boxarray* ba, int dir, const double* s,
                                   int ncomp_s, int ng_s, int comp, const double* vel_edge,
                                   int ng_v, const ksgf_bcrec* bc, double dt, int limiter_type,
                                   double* Im, double* Ip, ksgf_stream_t stream);

/* HydroUtils::ExtrapVelToFaces (call icns_advection.H:211-216).
 * vel: cell, 3 comp, ngrow 3 (ghosts filled); vel_forces: cell, 3 comp,
 * ngrow 1; u/v/w_mac: faces, 1 comp, ngrow 1 (valid faces written).
 * bcrec: 3 BCRecs (host). */
int ksgf_extrap_vel_to_faces(const ksgf_boxarray* ba, const double* vel,
                             const double* vel_forces, double* u_mac, double* v_mac,
                             double* w_mac, const ksgf_bcrec* bcrec, double dt, int use_ppm,
                             int use_forces_in_trans, int limiter_type,
                             int allow_inflow_on_outflow, ksgf_stream_t stream);

/* HydroUtils::ComputeFluxesOnBoxFromState (icns_advection.H:359-373,
 * AdvOp_Godunov.H:231-250) over all boxes.  q, q_on_face_bc: cell, ncomp,
 * ngrow 3; fq: cell, ncomp, ngrow 1; u/v/w_mac: faces, ngrow 1; divu: cell,
 * 1 comp, ngrow 1, or NULL (zero); flux_* and face_*: faces, ncomp, ngrow 0.
 * bcrec, iconserv: ncomp entries (host). */
int ksgf_compute_fluxes_from_state(const ksgf_boxarray* ba, int ncomp, const double* q,
                                   const double* q_on_face_bc, double* flux_x, double* flux_y,
                                   double* flux_z, double* face_x, double* face_y,
                                   double* face_z, int known_edge_state, const double* u_mac,
                                   const double* v_mac, const double* w_mac, const double* divu,
                                   const double* fq, double dt, const ksgf_bcrec* bcrec,
                                   const int* iconserv, int use_ppm, int use_forces_in_trans,
                                   int is_velocity, int fluxes_are_area_weighted,
                                   int limiter_type, int allow_inflow_on_outflow,
                                   ksgf_stream_t stream);

/* HydroUtils::ComputeDivergence (icns_advection.H:418-422): div = mult *
 * sum_d (F_d(+) - F_d(-)) / dx_d.  div: cell, ncomp, ngrow 0; fluxes: faces,
 * ncomp, ngrow 0. */
int ksgf_compute_divergence(const ksgf_boxarray* ba, double* div, const double* flux_x,
                            const double* flux_y, const double* flux_z, int ncomp, double mult,
                            int fluxes_are_area_weighted, ksgf_stream_t stream);

/* Hydro::MacProjector (icns_advection.cpp:88-104,143-154,289-353):
 * create = MacProjector(geom) + initProjector(ba, dm, lpinfo, beta) +
 * setDomainBC(lobc, hibc) + MLMGOptions applied; set_beta = updateBeta(Real);
 * project = setUMAC + project(rtol, atol), u/v/w_mac: faces, 1 comp, ngrow 1;
 * the valid faces are corrected in place.  info may be NULL. */
void ksgf_mlmg_default_options(ksgf_mlmg_options* opts);
int ksgf_macproj_create(const ksgf_boxarray* ba, const int* lobc, const int* hibc,
                        double beta, const ksgf_mlmg_options* opts, ksgf_macproj** out);
int ksgf_macproj_set_beta(ksgf_macproj* mp, double beta);
/* initProjector(LPInfo, Vector<Array<const MultiFab*,3>> beta) /
 * updateBeta(faces) (icns_advection.cpp:143-154, 289): beta on the faces,
 * beta_x/y/z: face MultiFabs, 1 comp, ngrow >= 0 (valid faces read); the
 * reference passes factor / rho_face when density varies
 * (icns_advection.cpp:234-291).  Until set_beta(Real) is called again,
 * project / solve / get_phi use a variable-coefficient multigrid in FP64
 * (whatever opts->precision).  Periodic directions need even extents. */
int ksgf_macproj_set_beta_faces(ksgf_macproj* mp, const double* beta_x, const double* beta_y,
                                const double* beta_z, int ngrow, ksgf_stream_t stream);
/* amrex::average_cellcenter_to_face (icns_advection.cpp:259-260): face value =
 * 0.5 (cell below + cell above), on the valid faces of each box; cc: cell,
 * 1 comp, ng_cc >= 1 with its ghosts filled. */
int ksgf_average_cellcenter_to_face(const ksgf_boxarray* ba, const double* cc, int ng_cc, double* fx,
                                    double* fy, double* fz, int ng_face, ksgf_stream_t stream);
/* MultiFab::invert(numerator, nghost) (icns_advection.cpp:282): every point
 * of the valid region grown by nghost becomes numerator / itself. */
int ksgf_invert(const ksgf_boxarray* ba, double* mf, int itype, int ncomp, int ngrow, int nghost,
                double numerator, ksgf_stream_t stream);
int ksgf_macproj_project(ksgf_macproj* mp, double* u_mac, double* v_mac, double* w_mac,
                         double rtol, double atol, ksgf_mlmg_info* info, ksgf_stream_t stream);
/* copy the last solution phi (domain-wide, n0*n1*n2, Fortran order) */
int ksgf_macproj_get_phi(ksgf_macproj* mp, double* phi, ksgf_stream_t stream);
int ksgf_macproj_destroy(ksgf_macproj* mp);

/* MacProjector::project split for a step captured into a hipGraph.  No
 * counterpart in the reference, whose MLMG tests its residual norm on the
 * host after every iteration; for callers that replay a fixed-dt step as a
 * graph and read the solve's state once per step.
 *   begin: div u_mac, the first residual, then `ncycles` iterations (rounded
 *     up to even), each followed by the FP64 residual and a device-side test
 *     (the host loop's: converged, diverging, maxiter) after which the later
 *     iterations' kernels return at once.  No host synchronisation and no
 *     memset: a stream capture records it.
 *   finish: reads the solve's state (one host synchronisation); a solve still
 *     iterating after ncycles continues with host-side tests.  Not capturable.
 *   correct: u_mac -= beta grad phi (capturable).
 * begin + finish + correct give ksgf_macproj_project's faces and phi bit for
 * bit.  One beta (no face coefficients).  Calls on one projector must not
 * overlap. */
int ksgf_macproj_project_begin(ksgf_macproj* mp, const double* u_mac, const double* v_mac,
                               const double* w_mac, double rtol, double atol, int ncycles,
                               ksgf_stream_t stream);
int ksgf_macproj_project_finish(ksgf_macproj* mp, ksgf_mlmg_info* info, ksgf_stream_t stream);
int ksgf_macproj_project_correct(ksgf_macproj* mp, double* u_mac, double* v_mac, double* w_mac,
                                 ksgf_stream_t stream);
/* Standalone Poisson solve div(beta grad phi) = rhs on the projector's
 * domain (the MLMG solve inside project); rhs/phi domain-wide (n0*n1*n2). */
int ksgf_macproj_solve(ksgf_macproj* mp, const double* rhs, double* phi, double rtol,
                       double atol, ksgf_mlmg_info* info, ksgf_stream_t stream);

/* Hydro::MacProjector over two AMR levels (icns_advection.cpp:88-95 builds
 * it on Geom(0..finest_level), :331 setUMAC(mac_vec) with every level's face
 * velocities, :353 project(rtol, atol)): the composite solve of
 * div(beta grad phi) = div(u*) on the coarse level outside the fine level
 * and on the fine level (ratio 2), then u -= beta grad phi on both levels and
 * the fine average onto the coarse faces under and around the fine level.
 * crse_ba covers the domain (origin 0); fine_ba: a level created with
 * ksgf_boxarray_create_level, boxes aligned to 2.  Coarse-fine ghosts are
 * quadratic through the tangentially interpolated coarse value; uncovered
 * coarse cells next to the fine level use the averaged fine fluxes (reflux).
 * opts: maxiter (composite iterations), num_pre/post_smooth (smoothing
 * steps on the fine level, each a pair of damped-Jacobi sweeps); each
 * iteration takes one FP64 V-cycle on the coarse grid.
 * project: umac_c / umac_f: the x, y, z face MultiFabs of each level, ngrow
 * ng, corrected in place; returns KSGF_ERR_NOT_CONVERGED (after correcting)
 * if maxiter is reached.  get_phi: the coarse solution domain-wide (n cells,
 * covered cells hold the average of their children) and the fine one as a
 * cell MultiFab with one ghost layer (ng_f must be 1).  The handle holds the
 * iterate between calls: calls on one handle must not overlap, and both
 * BoxArrays must outlive it. */
typedef struct ksgf_macproj_levels ksgf_macproj_levels;
int ksgf_macproj_levels_create(const ksgf_boxarray* crse_ba, const ksgf_boxarray* fine_ba,
                               const int* lobc, const int* hibc, double beta,
                               const ksgf_mlmg_options* opts, ksgf_macproj_levels** out);
int ksgf_macproj_levels_project(ksgf_macproj_levels* mp, double* const* umac_c,
                                double* const* umac_f, int ng, double rtol, double atol,
                                ksgf_mlmg_info* info, ksgf_stream_t stream);
int ksgf_macproj_levels_get_phi(ksgf_macproj_levels* mp, double* phi_c, double* phi_f, int ng_f,
                                ksgf_stream_t stream);
int ksgf_macproj_levels_destroy(ksgf_macproj_levels* mp);

/* Many independent Poisson problems solved together.  The reference solves
 * each with its own MLMG (MLMG::solve behind MacProjector::project,
 * icns_advection.cpp:331, one after another); here every multigrid stage of
 * every problem is one launch over a table of (problem, tile) blocks, and
 * one iteration is captured once into a hipGraph and replayed.  Each result
 * equals ksgf_macproj_solve on that problem alone (same iterations).
 * Problem p: n[3p..3p+2] cells over one box, dx[3p..], lobc/hibc[3p..]
 * (KSGF_LO_*), beta[p].  Needs an even x extent on every grid level and at
 * least one pre- and one post-smoothing sweep (else KSGF_ERR_UNSUPPORTED).
 * solve: rhs[p], phi[p] domain-wide device arrays (n0*n1*n2, Fortran
 * order); infos: np entries or NULL; use_graph 0 launches eagerly.
 * Synchronises the stream once per iteration.  Returns KSGF_ERR_DIVERGED
 * if any problem diverged, else KSGF_ERR_NOT_CONVERGED if any did not
 * converge. */
/* Reverse mode of the advect + project step (differentiation workload; no
 * counterpart in the reference, which has no adjoint).  A triply periodic
 * one-box domain of n[0..2] >= 8 cells over [0, prob_hi], density 1, zero
 * pressure gradient, a passive scalar s and the velocity u advanced at a
 * fixed dt by the step of ksgf_host_stepper / GodunovPredictor:
 *   u_mac = P ExtrapVelToFaces(u);  s' = s - dt div F(s);  u' = u - dt div F(u)
 * (conservative PPM fluxes with the given limiter, MAC projection to
 * mg_rtol / mg_atol with opts).  Fields are domain-wide device arrays,
 * Fortran order, velocity component-major (3*N), scalar N.
 * forward: nsteps steps from (vel0, s0), keeping a checkpoint of the state
 *   every checkpoint_every steps; writes the final fields (vel_out, s_out may
 *   be NULL) and, if mlmg_iters is not NULL, the V-cycles of each step (host).
 * backward: given the adjoint (gradient of a scalar function) of the final
 *   fields, returns the adjoint of the initial fields of the last forward
 *   call, recomputing each checkpoint segment and reversing its steps
 *   (limiter and upwind branches frozen at the forward state; the MAC
 *   projection's adjoint is the projection).  Atomic accumulation: the
 *   result is reproducible to rounding. */
typedef struct ksgf_adjoint ksgf_adjoint;
int ksgf_adjoint_create(const int* n, const double* prob_hi, int limiter_type, double mg_rtol,
                        double mg_atol, const ksgf_mlmg_options* opts, ksgf_adjoint** out);
int ksgf_adjoint_forward(ksgf_adjoint* a, const double* vel0, const double* s0, int nsteps,
                         double dt, int checkpoint_every, double* vel_out, double* s_out,
                         int* mlmg_iters, ksgf_stream_t stream);
int ksgf_adjoint_backward(ksgf_adjoint* a, const double* vel_bar, const double* s_bar,
                          double* vel0_bar, double* s0_bar, ksgf_stream_t stream);
int ksgf_adjoint_destroy(ksgf_adjoint* a);

typedef struct ksgf_poisson_batch ksgf_poisson_batch;
int ksgf_poisson_batch_create(int nprob, const int* n, const double* dx, const int* lobc,
                              const int* hibc, const double* beta,
                              const ksgf_mlmg_options* opts, ksgf_poisson_batch** out);
int ksgf_poisson_batch_solve(ksgf_poisson_batch* pb, const double* const* rhs,
                             double* const* phi, double rtol, double atol, int use_graph,
                             ksgf_mlmg_info* infos, ksgf_stream_t stream);
int ksgf_poisson_batch_destroy(ksgf_poisson_batch* pb);

/* amrex::MLABecLaplacian with constant coefficients solved by MLMG: the
 * implicit (or Crank-Nicolson) diffusion solve of DiffSolverIface::linsys_solve
 * (src/equation_systems/DiffusionOps.cpp:154-160: setScalars(1, dt_diff),
 * setACoeffs(rho), setBCoeffs(mu on faces), MLMG::solve(field, rho*field)),
 * one component at a time as ICNSDiffScalarOp solves the velocity
 * (icns_diffusion.H:60-140):  alpha a phi - beta b lap(phi) = rhs.
 * Domain BCs are KSGF_LO_* per direction (get_diffuse_tensor_bc); Dirichlet
 * sides are homogeneous with the value on the face (walls at rest).  The MAC
 * projector's multigrid with the diagonal grown by alpha a; phi starts from
 * zero.  solve: sol and rhs are cell MultiFabs of ncomp components (valid
 * cells of component comp are read / written).  apply = MLMG::apply
 * (compute_diff_term uses alpha = 0, beta = -1: out = b lap(in)); apply
 * reuses the solve's level-0 scratch.  set_scalars takes any finite values;
 * a solve needs beta b > 0 and alpha a >= 0 (else KSGF_ERR_INVALID). */
typedef struct ksgf_abeclap ksgf_abeclap;
int ksgf_abeclap_create(const ksgf_boxarray* ba, const int* lobc, const int* hibc, double alpha,
                        double a, double beta, double b, const ksgf_mlmg_options* opts,
                        ksgf_abeclap** out);
int ksgf_abeclap_set_scalars(ksgf_abeclap* ab, double alpha, double beta);
int ksgf_abeclap_solve(ksgf_abeclap* ab, double* sol, int ng_sol, const double* rhs, int ng_rhs,
                       int ncomp, int comp, double rtol, double atol, ksgf_mlmg_info* info,
                       ksgf_stream_t stream);
int ksgf_abeclap_apply(ksgf_abeclap* ab, double* out, int ng_out, const double* in, int ng_in,
                       int ncomp, int comp, ksgf_stream_t stream);
int ksgf_abeclap_destroy(ksgf_abeclap* ab);

/* amrex::MLTensorOp solved by MLMG: the default ICNS viscous operator
 * (ICNSDiffTensorOp, src/equation_systems/icns/icns_diffusion.H:18-50, chosen
 * by use_tensor_operator = true at :492-506; set up by
 * DiffSolverIface::setup_operator, DiffusionOps.cpp:57-110: setScalars,
 * setLevelBC, setACoeffs(rho), setShearViscosity(eta faces) at
 * DiffusionOps.H:54-71; no bulk viscosity):
 *   L u = alpha a u - beta div(eta (grad u + grad u^T) - (2/3) eta (div u) I).
 * create: lobc / hibc hold 9 KSGF_LO_* values, [component * 3 + direction]
 * (get_diffuse_tensor_bc's Vector<Array<LinOpBCType,3>>, incflo_diffusion.cpp:
 * 11-69); Dirichlet and Neumann sides are homogeneous (walls at rest);
 * periodic exactly in the geometry's periodic directions, with even extents.
 * Coefficients start as MLABecLaplacian's (a = 0, eta = 1); the scalars must
 * be set before apply / solve.  set_acoeffs: cell MultiFab, 1 comp;
 * set_shear_viscosity: face MultiFabs (ksgf_average_velocity_eta_to_faces of
 * the cell viscosity).  apply = MLMG::apply (compute_diff_term: scalars
 * (0, -1), out = div tau); solve = MLMG::solve (linsys_solve: scalars
 * (1, dt_diff), rhs = rho u): 3-component cell MultiFabs, valid cells read /
 * written, sol starts from zero; FP64 (alpha >= 0, beta > 0; a singular
 * operator is KSGF_ERR_INVALID). */
typedef struct ksgf_tensorop ksgf_tensorop;
int ksgf_tensorop_create(const ksgf_boxarray* ba, const int* lobc, const int* hibc,
                         const ksgf_mlmg_options* opts, ksgf_tensorop** out);
int ksgf_tensorop_set_scalars(ksgf_tensorop* op, double alpha, double beta);
int ksgf_tensorop_set_acoeffs(ksgf_tensorop* op, const double* a, int ngrow, ksgf_stream_t stream);
int ksgf_tensorop_set_shear_viscosity(ksgf_tensorop* op, const double* eta_x, const double* eta_y,
                                      const double* eta_z, int ngrow, ksgf_stream_t stream);
int ksgf_tensorop_apply(ksgf_tensorop* op, double* out, int ng_out, const double* in, int ng_in,
                        ksgf_stream_t stream);
int ksgf_tensorop_solve(ksgf_tensorop* op, double* sol, int ng_sol, const double* rhs, int ng_rhs,
                        double rtol, double atol, ksgf_mlmg_info* info, ksgf_stream_t stream);
int ksgf_tensorop_destroy(ksgf_tensorop* op);
/* diffusion::average_velocity_eta_to_faces (src/diffusion/incflo_diffusion.cpp:
 * 108-128): average_cellcenter_to_face of the cell viscosity, then
 * fixup_eta_on_domain_faces (:130-200): a face on a non-periodic domain side
 * takes the one cell beside it.  cc: cell, 1 comp, ng_cc >= 1 filled. */
int ksgf_average_velocity_eta_to_faces(const ksgf_boxarray* ba, const double* cc, int ng_cc, double* fx,
                                       double* fy, double* fz, int ng_face, ksgf_stream_t stream);

/* Two-level averaging, fine level at refinement ratio `ratio` (2 or 4; the fine
 * domain is the coarse domain refined, the fine boxes coarsen exactly):
 * amrex::average_down of cell data (gp after the nodal projection,
 * incflo_apply_nodal_projection.cpp:457-460): every coarse cell under fine
 * data gets the mean of its ratio^3 children; amrex::average_down_faces
 * (advective fluxes before the convective term, icns_advection.H:401-407):
 * every coarse face under fine faces gets the mean of its ratio^2 children,
 * in each coarse FAB that holds the face (and its periodic image).  Other
 * coarse data is left as it is.  fine: ng_f ghosts, crse: ng_c ghosts. */
int ksgf_average_down(const ksgf_boxarray* fine_ba, const double* fine, int ng_f,
                      const ksgf_boxarray* crse_ba, double* crse, int ng_c, int ncomp, int ratio,
                      ksgf_stream_t stream);
int ksgf_average_down_faces(const ksgf_boxarray* fine_ba, const double* fine_x,
                            const double* fine_y, const double* fine_z, int ng_f,
                            const ksgf_boxarray* crse_ba, double* crse_x, double* crse_y,
                            double* crse_z, int ng_c, int ncomp, int ratio, ksgf_stream_t stream);

/* amrex::FillPatchTwoLevels for the ghost cells of a refined level's cell
 * MultiFab, as Field::fillpatch fills level > 0 (FieldFillPatchOps.H:272-336),
 * without time interpolation: ghost cells inside the domain take the coarse data
 * interpolated (conservative linear, the default CellConsLinear: MC-limited
 * slopes of the parent, scaled so no child leaves the range of the parent and
 * its face neighbours), then the fine level's own data where a fine box holds
 * them (FillBoundary), then the physical BCs (bc / bcval as ksgf_fill_physbc).
 * crse: coarse cell MultiFab with ng_c >= 1 ghosts already filled. */
int ksgf_fillpatch_two_levels(const ksgf_boxarray* fine_ba, double* fine, int ng,
                              const ksgf_boxarray* crse_ba, const double* crse, int ng_c,
                              int ncomp, int ratio, const ksgf_bcrec* bc, const double* bcval,
                              ksgf_stream_t stream);

/* FillPatchTwoLevels of a refined level's face MultiFab (itype x/y/z
 * faces), as fillpatch_sibling_fields fills the MAC velocities' ghost faces
 * after the projection (icns_advection.H:231-237): face-linear interpolation
 * at ratio 2 (AMReX's FaceLinear: a fine face on a coarse face takes its
 * value, one midway between two coarse faces their mean, piecewise constant
 * across the face); a ghost face that is a valid face of another fine box
 * (a fine cell on either side, e.g. that box's high face on the coarse-fine
 * boundary) takes that box's value; then FillBoundary and the physical BCs
 * (bc / bcval as ksgf_fill_physbc).  crse: the coarse face MultiFab (valid
 * faces used). */
int ksgf_fillpatch_faces_two_levels(const ksgf_boxarray* fine_ba, double* fine, int itype, int ng,
                                    const ksgf_boxarray* crse_ba, const double* crse, int ng_c,
                                    int ncomp, const ksgf_bcrec* bc, const double* bcval,
                                    ksgf_stream_t stream);

/* Hydro::NodalProjector with a constant sigma (incflo::ApplyProjection,
 * src/projection/incflo_apply_nodal_projection.cpp:340-353 for the constant-
 * density branch, options from MLMGOptions::operator()(NodalProjector&),
 * src/core/MLMGOptions.cpp:118-130): create = NodalProjector(vel,
 * scaling_factor/rho_0, geom, lpinfo) + setDomainBC(lobc, hibc) (KSGF_LO_*;
 * inflow gives KSGF_ERR_UNSUPPORTED) + the options; set_sigma changes sigma.
 * project = project(rtol, atol): vel is a cell MultiFab, 3 comp, ngrow
 * ng_vel >= 1, whose ghosts are as ApplyProjection leaves them (setBndry(0)
 * then FillBoundary: zero on physical boundaries, filled elsewhere); the
 * nodal divergence of vel is solved for and vel -= sigma grad phi on the valid
 * cells.  gradphi (getGradPhi; cell, 3 comp, ngrow ng_gp, may be NULL)
 * receives grad phi on the valid cells.  Node arrays (getPhi, solve,
 * compute_rhs) are domain-wide, Fortran order, N_d = n_d nodes along a
 * periodic direction and n_d + 1 otherwise (node_dims). */
typedef struct ksgf_nodalproj ksgf_nodalproj;
int ksgf_nodalproj_create(const ksgf_boxarray* ba, const int* lobc, const int* hibc,
                          double sigma, const ksgf_mlmg_options* opts, ksgf_nodalproj** out);
int ksgf_nodalproj_set_sigma(ksgf_nodalproj* np, double sigma);
int ksgf_nodalproj_node_dims(const ksgf_nodalproj* np, int* N);
int ksgf_nodalproj_project(ksgf_nodalproj* np, double* vel, int ng_vel, double* gradphi, int ng_gp,
                           double rtol, double atol, ksgf_mlmg_info* info, ksgf_stream_t stream);
int ksgf_nodalproj_get_phi(ksgf_nodalproj* np, double* phi, ksgf_stream_t stream);
/* NodalProjector::computeRHS: nodal divergence of vel (ghosts as for project) */
int ksgf_nodalproj_compute_rhs(ksgf_nodalproj* np, const double* vel, int ng_vel, double* rhs,
                               ksgf_stream_t stream);
/* the MLMG solve inside project, for a nodal right-hand side */
int ksgf_nodalproj_solve(ksgf_nodalproj* np, const double* rhs, double* phi, double rtol,
                         double atol, ksgf_mlmg_info* info, ksgf_stream_t stream);
int ksgf_nodalproj_destroy(ksgf_nodalproj* np);

/* The non-incremental "add (grad p / rho) back" loop of incflo::ApplyProjection
 * (incflo_apply_nodal_projection.cpp:209-233, no mesh mapping): on valid cells
 * vel += gp * (scaling_factor / rho).  vel, gp: cell, 3 comp; rho: 1 comp. */
int ksgf_incflo_add_gradp(const ksgf_boxarray* ba, double* vel, int ng_vel, const double* gp,
                          int ng_gp, const double* rho, int ng_rho, double scaling_factor,
                          ksgf_stream_t stream);

/* MultiFab::setBndry(val): every stored point outside the valid box of each
 * FAB (incflo_apply_nodal_projection.cpp:310). */
int ksgf_set_bndry(const ksgf_boxarray* ba, double* mf, int itype, int ncomp, int ngrow, double val,
                   ksgf_stream_t stream);

/* incflo::compute_dt convective part (incflo_compute_dt.cpp:69-91):
 * max over valid cells of max_d |u_d|/dx_d (floor -1).  Synchronises. */
int ksgf_cfl_max(const ksgf_boxarray* ba, const double* vel, int ngrow, double* out,
                 ksgf_stream_t stream);

/* No counterpart in the reference.  ksgf_cfl_max for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same result, with the reduction's
 * partial maxima in a file-scope device workspace of 8 MiB per device (held
 * until the process ends) instead of library-allocated memory.  It calls no
 * hipMalloc or hipFree.  Limits: at most 268,435,456 valid cells, and the
 * velocity's ngrow within 0..3 (the launch tables ksgf_boxarray_create
 * builds); beyond either it returns KSGF_ERR_INVALID.  Every call on one
 * device shares that device's workspace: calls that overlap on one device
 * (other threads, other streams) race.  Synchronises. */
int ksgf_cfl_max_noalloc(const ksgf_boxarray* ba, const double* vel, int ngrow, double* out,
                         ksgf_stream_t stream);

/* incflo::compute_dt forcing part (incflo_compute_dt.cpp:168-198, taken when
 * time.use_force_cfl, the default): max over valid cells of
 * max_d |src_d| / dx_d / rho, with src the ICNS source term the step last
 * formed (momentum form).  src: cell, 3 comp; rho: 1 comp.  Synchronises.
 * SimTime::set_current_cfl (SimTime.cpp:162-205) then takes
 * dt = 2 cfl / (C + sqrt(C^2 + 4 F)), capped at (1 + max_dt_growth) dt_prev. */
int ksgf_force_cfl_max(const ksgf_boxarray* ba, const double* src, int ng_src, const double* rho,
                       int ng_rho, double* out, ksgf_stream_t stream);

/* max over valid cells of |div u_mac| (the post-projection check the MAC
 * projector's convergence implies; diagnostic, synchronises).  u/v/w_mac:
 * faces, 1 comp, ngrow ngrow_mac. */
int ksgf_mac_divergence_max(const ksgf_boxarray* ba, const double* u_mac, const double* v_mac,
                            const double* w_mac, int ngrow_mac, double* out, ksgf_stream_t stream);

/* ComputeRHSOp::predictor_rhs, multiply_rho branch (CompRHSOps.H:93-117):
 * vel_new = (rho_old*vel_old + dt*(conv + src + kappa*diff)) / rho_new.
 * diff may be NULL. conv/src/diff: cell, ncomp, ngrow 0. */
int ksgf_predictor_update(const ksgf_boxarray* ba, double* vel_new, int ng_new,
                          const double* vel_old, int ng_old, const double* rho_old, int ng_ro,
                          const double* rho_new, int ng_rn, const double* conv,
                          const double* src, const double* diff, double kappa, double dt,
                          int ncomp, ksgf_stream_t stream);

/* SrcTermOp<ICNS> (icns_ops.H:98-111) in velocity form: src = -gp / rho,
 * then (scale_by_rho) src *= rho (PDEOps multiply_rho).  gp: ngrow 0. */
int ksgf_icns_src_term(const ksgf_boxarray* ba, double* src, int ng_src, const double* gp,
                       const double* rho, int ng_rho, int scale_by_rho, ksgf_stream_t stream);

/* The same with the explicit viscous term divtau (cell, 3 comp, ngrow 0)
 * added to the momentum-form forcing before the division by density, as
 * ApplyPredictor adds it for Godunov (incflo_advance.cpp:246-291):
 * src = (-gp + divtau) / rho. */
int ksgf_icns_src_term_divtau(const ksgf_boxarray* ba, double* src, int ng_src, const double* gp,
                              const double* divtau, const double* rho, int ng_rho,
                              ksgf_stream_t stream);

/* out(comp n) = a(comp n) * rho over valid + ngrow cells (MultiFab::Copy +
 * MultiFab::Multiply, icns_advection.H:281-314).  a/out: ncomp, ng_a/ng_out
 * >= ngrow; rho: 1 comp, ng_rho >= ngrow. */
int ksgf_scale_by_rho(const ksgf_boxarray* ba, double* out, int ng_out, const double* a,
                      int ng_a, const double* rho, int ng_rho, int ncomp, int ngrow,
                      int divide, ksgf_stream_t stream);

/* ---- AMReX plotfile / VisMF files of MultiFabs in HBM (host I/O) ----
 * Text layouts as the reference's readers parse them (tools/amrex_plotfile.py,
 * tools/utilities/CheckpointToCSV/CheckpointFileUtil.cpp); FAB payloads in IEEE
 * double, this host's byte order.  The MultiFab is copied to or from the host
 * once (the stream is synchronised). */

/* amrex::WriteMultiLevelPlotfile with one level (IOManager::write_plot_file,
 * IOManager.cpp:181-187): <dir>/Header, <dir>/Level_0/Cell_H and
 * <dir>/Level_0/Cell_D_00000, valid cells of the cell-centred MultiFab mf
 * (ncomp components, ngrow ghosts in memory); names: ncomp variable names. */
int ksgf_plotfile_write(const char* dir, const ksgf_boxarray* ba, const double* mf, int ncomp,
                        int ngrow, const char* const* names, double time, int step,
                        ksgf_stream_t stream);

/* Level 0 of a plotfile into the valid cells of mf (ghosts untouched): every file
 * box's data is copied where it overlaps the BoxArray's boxes (ParallelCopy).
 * IEEE double or float data in either byte order; other formats give
 * KSGF_ERR_UNSUPPORTED, a component count other than ncomp KSGF_ERR_INVALID. */
int ksgf_plotfile_read(const char* dir, const ksgf_boxarray* ba, double* mf, int ncomp, int ngrow,
                       ksgf_stream_t stream);

/* amrex::VisMF::Write of one MultiFab, ghost cells included (checkpoint fields,
 * IOManager.cpp:214-219): <prefix>_H and <prefix>_D_00000. */
int ksgf_vismf_write(const char* prefix, const ksgf_boxarray* ba, const double* mf, int itype,
                     int ncomp, int ngrow, ksgf_stream_t stream);

/* amrex::VisMF::Read (IOManager.cpp:249-266): with the file's boxes equal to the
 * BoxArray's, each FAB's points held on both sides, ghosts included; otherwise the
 * file's valid data into the valid boxes it overlaps. */
int ksgf_vismf_read(const char* prefix, const ksgf_boxarray* ba, double* mf, int itype, int ncomp,
                    int ngrow, ksgf_stream_t stream);

/* Checkpoint <dir>/Header of one level (IOManager::write_header,
 * IOManager.cpp:308-356); creates <dir>/Level_0 for the VisMF fields. */
int ksgf_checkpoint_write_header(const char* dir, const ksgf_boxarray* ba, int step, double time,
                                 double dt, double dt_nm1, double dt_nm2);

/* Live timing of the finest-level smoother launches (bench instrumentation;
 * no counterpart in the reference): reset starts recording HIP events around
 * each launch on its stream; read synchronises on them and returns the
 * summed duration, launch count, algorithmic bytes and kernel name, then
 * stops recording. */
int ksgf_prof_reset(void);
int ksgf_prof_read(double* ms, long long* launches, double* bytes, char* name, int name_len);

#ifdef __cplusplus
}
#endif
#endif
