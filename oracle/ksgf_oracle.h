/*
 * ksgf_oracle.h -- CPU FP64 restatement of Kynema-SGF's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked
 * against (and the CPU baseline timed beside it).  Only tests/, the smoke
 * check and the bench's cpu_baseline leg may load it; the product library
 * never links or calls it.
 *
 * The reference delegates this arithmetic to AMReX-Hydro and AMReX, which are
 * empty submodules in the reference snapshot (.gitmodules:1-9).  The oracle
 * therefore restates the published methods the reference's call sites select:
 *   - PPM edge values and traces: Colella & Woodward, JCP 54 (1984), with the
 *     van Leer / MC slope limiter of the 4th-order edge interpolant.
 *   - Unsplit corner-transport (CTU) Godunov predictor in 3-D: Colella,
 *     JCP 87 (1990); Saltzman, JCP 115 (1994); as used for incompressible
 *     flow by Almgren, Bell, Colella, Howell & Welcome, JCP 142 (1998)
 *     (cited at docs/sphinx/theory/include/discretization.rst:7-9).
 *   - MAC projection: cell-centred 7-point Poisson, geometric multigrid
 *     V-cycles with red-black Gauss-Seidel, average restriction, piecewise
 *     constant prolongation, unpreconditioned BiCGStab bottom solve
 *     (van der Vorst, SIAM J. Sci. Stat. Comput. 13 (1992)).
 * Call structure, options and defaults follow the reference call sites:
 *   src/equation_systems/icns/icns_advection.H:172-442,
 *   src/equation_systems/icns/icns_advection.cpp:21-375,
 *   src/equation_systems/CompRHSOps.H:32-141,
 *   src/core/MLMGOptions.{H,cpp}, src/incflo_compute_dt.cpp:33-220.
 *
 * Data layout (shared with the HIP library): a MultiFab is one flat double
 * array holding one FAB per box, boxes in BoxArray order, no padding.  FAB b
 * spans the box converted to the index type (cell, or nodal in one direction
 * for x/y/z faces) and grown by ngrow, stored Fortran order (i fastest) with
 * the component index slowest.
 */
#ifndef KSGF_ORACLE_H
#define KSGF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* amrex::BCType numeric values */
enum { KO_BC_INT_DIR = 0, KO_BC_FOEXTRAP = 2, KO_BC_EXT_DIR = 3, KO_BC_HOEXTRAP = 4 };
/* PPM limiter selector */
enum { KO_PPM_NOLIMITER = 0, KO_PPM_VANLEER = 1, KO_PPM_UPWIND = 4, KO_PPM_MINMOD = 5 };
/* index types */
enum { KO_CELL = 0, KO_XFACE = 1, KO_YFACE = 2, KO_ZFACE = 3 };
/* linear-operator domain BCs (LinOpBCType subset) */
enum { KO_LO_PERIODIC = 0, KO_LO_NEUMANN = 1, KO_LO_DIRICHLET = 2 };

typedef struct {
    int nbox;
    const int* boxes; /* nbox*6: lo0 lo1 lo2 hi0 hi1 hi2, cell-centred inclusive */
    int dlo[3], dhi[3];
    double dx[3];
    int periodic[3];
} ko_geom;

typedef struct { int lo[3]; int hi[3]; } ko_bcrec;

/* Planted mistakes, for tests that show their check fails on a wrong
 * oracle (0 = none; bits documented in ksgf_oracle.c). */
void ko_set_plant(int bits);

typedef struct {
    int maxiter, nu1, nu2, nuf, nub;
    int bottom_maxiter;
    double bottom_rtol, bottom_atol;
    int semicoarsening;
    int max_coarsening_level;
} ko_mlmg_opts;

typedef struct {
    int iters;
    double r0, rfinal;
    int converged;
    int nlevels;
    int bottom_iters;   /* total over all V-cycles */
    int bottom_fail;    /* V-cycles whose bottom solve did not converge */
} ko_mlmg_info;

long ko_mf_size(const ko_geom* g, int itype, int ncomp, int ngrow);

void ko_fill_boundary(const ko_geom* g, double* mf, int itype, int ncomp, int ngrow);
void ko_fill_physbc(const ko_geom* g, double* mf, int itype, int ncomp, int ngrow,
                    const ko_bcrec* bc, const double* bcval);

void ko_ppm_predict_state(const ko_geom* g, int dir, const double* s, int ncomp_s, int ng_s,
                          int comp, const double* vel_edge, int ng_v, const ko_bcrec* bc,
                          double dt, int limiter, double* Im, double* Ip);

int ko_extrap_vel_to_faces(const ko_geom* g, const double* vel, const double* f,
                           double* umac, double* vmac, double* wmac, const ko_bcrec* bc,
                           double dt, int use_ppm, int use_forces_in_trans, int limiter,
                           int allow_inflow_on_outflow);

int ko_compute_fluxes_from_state(const ko_geom* g, int ncomp, const double* q,
                                 const double* q_bc, double* flux_x, double* flux_y,
                                 double* flux_z, double* face_x, double* face_y,
                                 double* face_z, int known_edge_state, const double* umac,
                                 const double* vmac, const double* wmac, const double* divu,
                                 const double* fq, double dt, const ko_bcrec* bc,
                                 const int* iconserv, int use_ppm, int use_forces_in_trans,
                                 int is_velocity, int area_weighted, int limiter,
                                 int allow_inflow_on_outflow);

void ko_compute_divergence(const ko_geom* g, double* div, const double* fx, const double* fy,
                           const double* fz, int ncomp, double mult, int area_weighted);

void ko_mlmg_default_opts(ko_mlmg_opts* o);

int ko_mac_project(const ko_geom* g, double* umac, double* vmac, double* wmac,
                   const int* lobc, const int* hibc, double beta, double rtol, double atol,
                   const ko_mlmg_opts* opts, ko_mlmg_info* info, double* phi_out);

/* mg solves return 0 (converged), 1 (maxiter reached) or KO_MG_DIVERGED
 * (residual above KO_MG_DIVERGE times |rhs|_inf, or not finite) */
#define KO_MG_DIVERGE 1.0e20
#define KO_MG_DIVERGED (-5)

/* div(beta grad phi) = rhs with face coefficients (MacProjector's
 * Array<MultiFab,3> beta): bx (n0+1, n1, n2), by, bz domain-wide */
int ko_vb_solve(const int* n, const double* dx, const int* lobc, const int* hibc, const double* bx,
                const double* by, const double* bz, const double* rhs_in, double* phi_out, double rtol,
                double atol, const ko_mlmg_opts* opts, ko_mlmg_info* info);
int ko_poisson_solve(const int* n, const double* dx, const int* lobc, const int* hibc,
                     double beta, const double* rhs, double* phi, double rtol, double atol,
                     const ko_mlmg_opts* opts, ko_mlmg_info* info);

/* Nodal projection with a constant sigma (NodalProjector / MLNodeLaplacian;
 * src/projection/incflo_apply_nodal_projection.cpp:149-465).  Node arrays
 * are domain-wide, Fortran order, N_d = n_d nodes in a periodic direction and
 * n_d + 1 otherwise (ko_nodal_dims). */
int ko_nodal_dims(const int* n, const int* lobc, int* N);
void ko_nodal_rhs(const ko_geom* g, const double* vel, int ng, const int* lobc, const int* hibc,
                  double* rhs);
int ko_nodal_solve(const int* n, const double* dx, const int* lobc, const int* hibc, double sigma,
                   const double* rhs, double* phi, double rtol, double atol,
                   const ko_mlmg_opts* opts, ko_mlmg_info* info);
void ko_nodal_correct(const ko_geom* g, double* vel, int ng, double* gp, const double* phi,
                      const int* lobc, double sigma);
int ko_nodal_project(const ko_geom* g, double* vel, int ng, const int* lobc, const int* hibc,
                     double sigma, double rtol, double atol, const ko_mlmg_opts* opts,
                     ko_mlmg_info* info, double* phi_out, double* gp_out);

/* MLABecLaplacian, constant coefficients (implicit diffusion,
 * DiffusionOps.cpp:154-160): alpha_a phi - beta_b lap(phi) = rhs, and its
 * apply; domain-wide cell arrays, Fortran order */
int ko_abeclap_solve(const int* n, const double* dx, const int* lobc, const int* hibc, double alpha_a,
                     double beta_b, const double* rhs, double* phi, double rtol, double atol,
                     const ko_mlmg_opts* opts, ko_mlmg_info* info);
void ko_abeclap_apply(const int* n, const double* dx, const int* lobc, const int* hibc, double alpha_a,
                      double beta_b, const double* x, double* y);

/* MLTensorOp (the default ICNS viscous operator): out = alpha a u - beta
 * div(eta (grad u + grad u^T) - (2/3) eta (div u) I), domain-wide, u / out
 * 3 components (component slowest), a cell (NULL = 1), ex / ey / ez face
 * viscosities; lobc / hibc [component * 3 + direction] (KO_LO_*) */
void ko_tensor_apply(const int* n, const double* dx, const int* lobc, const int* hibc, double alpha,
                     const double* a, double beta, const double* ex, const double* ey, const double* ez,
                     const double* u, double* out);
int ko_tensor_solve(const int* n, const double* dx, const int* lobc, const int* hibc, double alpha,
                    const double* a, double beta, const double* ex, const double* ey, const double* ez,
                    const double* rhs, double* sol, double rtol, double atol, int maxiter, ko_mlmg_info* info);

/* amrex::average_down / average_down_faces at refinement ratio `ratio` */
void ko_average_down(const ko_geom* fg, const double* fine, int ng_f, const ko_geom* cg, double* crse,
                     int ng_c, int ncomp, int ratio);
void ko_average_down_faces(const ko_geom* fg, const double* const* fine, int ng_f, const ko_geom* cg,
                           double* const* crse, int ng_c, int ncomp, int ratio);

/* FillPatchTwoLevels of a refined level's cell MultiFab (ghosts): coarse data
 * by conservative linear interpolation, then FillBoundary, then physical BCs */
void ko_fillpatch_two_levels(const ko_geom* fg, double* fine, int ng, const ko_geom* cg, const double* crse,
                             int ng_c, int ncomp, int ratio, const ko_bcrec* bc, const double* bcval);

/* FillPatchTwoLevels of a refined level's face MultiFab (itype: x/y/z
 * faces), face-linear interpolation at ratio 2, then FillBoundary and
 * physical BCs (the MAC velocities' fillpatch_sibling_fields) */
void ko_fillpatch_faces_two_levels(const ko_geom* fg, double* fine, int itype, int ng, const ko_geom* cg,
                                   const double* crse, int ng_c, int ncomp, const ko_bcrec* bc, const double* bcval);

/* Composite two-level MAC projection (MacProjector over levels 0 and 1 at
 * ratio 2).  Domain-wide arrays: coarse cells n, fine cells 2n (valid where a
 * fine box covers), faces with one more entry along their direction.  fboxes:
 * fine boxes in fine indices, aligned to 2.  umc/umf: x, y, z face arrays of
 * each level, corrected in place; phic/phif (nullable): the solution. */
#define KO_TL_OMEGA (6.0 / 7.0)
void ko_tl_apply(const int* n, const double* dxc, const int* lobc, const int* hibc, const int* fboxes, int nfbox,
                 double beta, const double* phic, const double* phif, double* Lc, double* Lf);
int ko_mac_project_two_levels(const int* n, const double* dxc, const int* lobc, const int* hibc, const int* fboxes,
                              int nfbox, double beta, double* const* umc, double* const* umf, double rtol,
                              double atol, const ko_mlmg_opts* opts, ko_mlmg_info* info, double* phic_out,
                              double* phif_out);

double ko_cfl_max(const ko_geom* g, const double* vel, int ng);
double ko_force_cfl_max(const ko_geom* g, const double* src, int ng_src, const double* rho, int ng_rho);

void ko_predictor_update(const ko_geom* g, double* vel_new, int ng_new, const double* vel_old,
                         int ng_old, const double* rho_old, int ng_ro, const double* rho_new,
                         int ng_rn, const double* conv, const double* src, const double* diff,
                         double kappa, double dt, int ncomp);

#ifdef __cplusplus
}
#endif
#endif
