/* C entry points of the C++ host layer (libksgf_host.so).
 * The advect + MAC-project step of incflo::ApplyPredictor's Godunov branch
 * (incflo_advance.cpp:186-438, without diffusion and nodal projection),
 * driven by the C++ GodunovPredictor over the ksgf C ABI. */
#ifndef KSGF_HOST_H
#define KSGF_HOST_H

#include "../../include/ksgf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ksgf_host_stepper ksgf_host_stepper;

const char* ksgf_host_last_error(void);
/* vel_bc: 3 BCRecs; vel_bcval: [6 orientations][3 comps]; lobc/hibc: LinOp BCs */
int ksgf_host_stepper_create(const ksgf_geometry* geom, const int* boxes, int nbox, const ksgf_bcrec* vel_bc,
                             const double* vel_bcval, const int* lobc, const int* hibc, double rho0, double cfl,
                             double mg_rtol, double mg_atol, const ksgf_mlmg_options* opts, ksgf_host_stepper** out);
int ksgf_host_stepper_set_stream(ksgf_host_stepper* s, ksgf_stream_t stream);
/* domain-wide host arrays, x fastest: vel 3*N, rho N, gp 3*N */
int ksgf_host_stepper_set_state(ksgf_host_stepper* s, const double* vel, const double* rho, const double* gp);
/* fixed_dt <= 0: dt from the convective CFL; divmax (may be NULL): max
 * |div u_mac| after the MAC projection */
int ksgf_host_stepper_step(ksgf_host_stepper* s, double fixed_dt, double* dt_used, ksgf_mlmg_info* info,
                           double* divmax);
/* fixed-dt steps (fixed_dt > 0) replayed from hipGraphs with one host read
 * per step (GodunovPredictor::enableGraph); equal to the eager steps bit for
 * bit; KSGF_ERR_UNSUPPORTED with diffusion, the nodal projection or the
 * divergence check */
int ksgf_host_stepper_enable_graph(ksgf_host_stepper* s, int on);
int ksgf_host_stepper_get_velocity(ksgf_host_stepper* s, double* vel);
/* the MAC projection's phi of the last step, domain-wide host array (n0*n1*n2) */
int ksgf_host_stepper_get_phi(ksgf_host_stepper* s, double* phi);
/* passive scalar advected with the step (AdvOp_Godunov.H:116-300):
 * bc: its BCRec, bcval: 6 orientation values; then set / get domain-wide */
int ksgf_host_stepper_enable_scalar(ksgf_host_stepper* s, const ksgf_bcrec* bc, const double* bcval);
int ksgf_host_stepper_set_scalar(ksgf_host_stepper* s, const double* scal);
int ksgf_host_stepper_get_scalar(ksgf_host_stepper* s, double* scal);
/* plotfile of the step's fields (IOManager::write_plot_file): velocity, gp,
 * density (and the scalar when enabled) */
int ksgf_host_stepper_write_plotfile(ksgf_host_stepper* s, const char* dir, double time, int step);
/* end every step with incflo::ApplyProjection (nodal projection, constant
 * density; incflo_advance.cpp:412-414) to nodal_proj rtol/atol with opts
 * (NULL: defaults); nodal_info: the last step's nodal solve */
int ksgf_host_stepper_enable_nodal_projection(ksgf_host_stepper* s, double rtol, double atol,
                                              const ksgf_mlmg_options* opts);
int ksgf_host_stepper_nodal_info(ksgf_host_stepper* s, ksgf_mlmg_info* info);
/* gp (3*N, domain-wide host) and the nodal phi (domain-wide nodes, host) */
int ksgf_host_stepper_get_gradp(ksgf_host_stepper* s, double* gp);
int ksgf_host_stepper_get_nodal_phi(ksgf_host_stepper* s, double* phi);
/* viscous term with constant viscosity mu: type 1 Crank-Nicolson, 2 Implicit
 * (incflo.diffusion_type); tensor 1: the tensor operator (MLTensorOp, the
 * reference's default, use_tensor_operator = true), 0: the scalar operator
 * (one MLABecLaplacian per component); diffusion_info: the last step's three
 * component solves (the tensor operator's one coupled solve in [0]) */
int ksgf_host_stepper_enable_diffusion(ksgf_host_stepper* s, double mu, int type, double rtol, double atol,
                                       const ksgf_mlmg_options* opts, int tensor);
int ksgf_host_stepper_diffusion_info(ksgf_host_stepper* s, ksgf_mlmg_info* info3);
int ksgf_host_stepper_destroy(ksgf_host_stepper* s);

/* MacProjOp::operator() with variable density (icns_advection.cpp:234-299)
 * on the Hydro layer, for one level: rho_face = average_cellcenter_to_face
 * of rho (cell MultiFab, one ghost layer filled by the host layer's
 * FillBoundary / foextrap walls), beta = factor / rho_face (invert), then
 * initProjector(LPInfo, beta faces), setDomainBC, setUMAC, project.  rho:
 * domain-wide host (N); umac/vmac/wmac: domain-wide host face arrays,
 * corrected in place; phi (may be NULL): domain-wide host (N). */
int ksgf_host_mac_project_density(const ksgf_geometry* geom, const int* boxes, int nbox, const int* lobc,
                                  const int* hibc, const double* rho, double factor, double* umac, double* vmac,
                                  double* wmac, double rtol, double atol, const ksgf_mlmg_options* opts,
                                  ksgf_mlmg_info* info, double* phi);

/* Two AMR levels (level 1 at ratio 2, boxes in geom0's index space refined
 * by 2, aligned to 2): the Godunov advect + composite MAC-project step of
 * ksgf::amr::TwoLevelPredictor (incflo_two_level.H).  set_state / get_velocity
 * take arrays in the FAB layout of each level's MultiFab (velocity with 3
 * ghost layers; gp with none, may be NULL), in host memory (set_state:
 * device memory with on_device).  limiter: KSGF_PPM_*. */
typedef struct ksgf_host_two_level ksgf_host_two_level;
int ksgf_host_two_level_create(const ksgf_geometry* geom0, const int* boxes0, int nbox0, const int* boxes1,
                               int nbox1, const ksgf_bcrec* vel_bc, const double* vel_bcval, const int* lobc,
                               const int* hibc, double rho0, double cfl, double mg_rtol, double mg_atol,
                               const ksgf_mlmg_options* opts, int limiter, ksgf_host_two_level** out);
int ksgf_host_two_level_set_state(ksgf_host_two_level* s, const double* vel0, const double* vel1, const double* gp0,
                                  const double* gp1, int on_device);
int ksgf_host_two_level_step(ksgf_host_two_level* s, double fixed_dt, double* dt_used, ksgf_mlmg_info* info);
int ksgf_host_two_level_get_velocity(ksgf_host_two_level* s, double* vel0, double* vel1);
int ksgf_host_two_level_destroy(ksgf_host_two_level* s);

#ifdef __cplusplus
}
#endif
#endif
