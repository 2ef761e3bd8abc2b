// Streaming operations around the advection/projection path:
//   CFL reduction              incflo_compute_dt.cpp:69-91
//   predictor update           CompRHSOps.H:93-117 (multiply_rho branch)
//   ICNS source term           icns_ops.H:98-111
//   momentum formation         icns_advection.H:281-314 (Copy + Multiply)
#include <cstring>

#include "ksgf_internal.h"

namespace ksgf {

__global__ void __launch_bounds__(256) k_cfl(CMFDev v, RegDev r, const int* __restrict__ box, double dxi,
                                             double dyi, double dzi, unsigned long long* __restrict__ out)
{
    __shared__ double sh[4];
    double m = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < r.total; t += (long long)gridDim.x * blockDim.x) {
        int b, i, j, k;
        region_locate_flat(r, box, t, b, i, j, k);
        CFab f = fab_at(v, box, b);
        double a = fabs(f(i, j, k, 0)) * dxi;
        double bb = fabs(f(i, j, k, 1)) * dyi;
        double e = fabs(f(i, j, k, 2)) * dzi;
        m = fmax(m, fmax(a, fmax(bb, e)));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double q = sh[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) q = fmax(q, sh[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(q));
    }
}

__global__ void __launch_bounds__(256) k_predictor(MFDev vn, CMFDev vo, CMFDev ro, CMFDev rn, CMFDev conv,
                                                   CMFDev src, CMFDev diff, int has_diff, double kappa, double dt,
                                                   int ncomp, RegDev r, const int* __restrict__ box)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    Fab fn = fab_at(vn, box, b);
    CFab fo = fab_at(vo, box, b), fro = fab_at(ro, box, b), frn = fab_at(rn, box, b);
    CFab fc = fab_at(conv, box, b), fs = fab_at(src, box, b);
    double rho_o = fro(i, j, k), rho_n = frn(i, j, k);
    for (int n = 0; n < ncomp; ++n) {
        double df = 0.0;
        if (has_diff) {
            CFab fd = fab_at(diff, box, b);
            df = fd(i, j, k, n);
        }
        double tt = rho_o * fo(i, j, k, n) + dt * (fc(i, j, k, n) + fs(i, j, k, n) + kappa * df);
        fn(i, j, k, n) = tt / rho_n;
    }
}

__global__ void __launch_bounds__(256) k_src_term(MFDev src, CMFDev gp, CMFDev rho, int scale, RegDev r,
                                                  const int* __restrict__ box)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    Fab fs = fab_at(src, box, b);
    CFab fg = fab_at(gp, box, b), fr = fab_at(rho, box, b);
    double rv = fr(i, j, k);
    double rhoinv = 1.0 / rv;
    for (int n = 0; n < 3; ++n) {
        double s = -(1.0 * fg(i, j, k, n)) * rhoinv;
        if (scale) s = s * rv;
        fs(i, j, k, n) = s;
    }
}

__global__ void __launch_bounds__(256) k_src_term_divtau(MFDev src, CMFDev gp, CMFDev tau, CMFDev rho, RegDev r,
                                                         const int* __restrict__ box)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    Fab fs = fab_at(src, box, b);
    CFab fg = fab_at(gp, box, b), ft = fab_at(tau, box, b), fr = fab_at(rho, box, b);
    double rhoinv = 1.0 / fr(i, j, k);
    for (int n = 0; n < 3; ++n) fs(i, j, k, n) = (-(1.0 * fg(i, j, k, n)) + ft(i, j, k, n)) * rhoinv;
}

__global__ void __launch_bounds__(256) k_scale_rho(MFDev out, CMFDev a, CMFDev rho, int ncomp, int divide, RegDev r,
                                                   const int* __restrict__ box)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    Fab fo = fab_at(out, box, b);
    CFab fa = fab_at(a, box, b), fr = fab_at(rho, box, b);
    double rv = fr(i, j, k);
    for (int n = 0; n < ncomp; ++n) fo(i, j, k, n) = divide ? fa(i, j, k, n) / rv : fa(i, j, k, n) * rv;
}

__global__ void __launch_bounds__(256) k_mac_divmax(CMFDev u, CMFDev v, CMFDev w, RegDev r, const int* __restrict__ box,
                                                    double dxi, double dyi, double dzi, unsigned long long* __restrict__ out)
{
    __shared__ double sh[4];
    double m = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < r.total; t += (long long)gridDim.x * blockDim.x) {
        int b, i, j, k;
        region_locate_flat(r, box, t, b, i, j, k);
        CFab fu = fab_at(u, box, b), fv = fab_at(v, box, b), fw = fab_at(w, box, b);
        double d = (fu(i + 1, j, k) - fu(i, j, k)) * dxi + (fv(i, j + 1, k) - fv(i, j, k)) * dyi +
                   (fw(i, j, k + 1) - fw(i, j, k)) * dzi;
        m = fmax(m, fabs(d));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double q = sh[0];
        for (int w2 = 1; w2 < (int)(blockDim.x >> 6); ++w2) q = fmax(q, sh[w2]);
        atomicMax(out, (unsigned long long)__double_as_longlong(q));
    }
}

}  // namespace ksgf

using namespace ksgf;

extern "C" {

int ksgf_cfl_max(const ksgf_boxarray* ba, const double* vel, int ngrow, double* out, ksgf_stream_t stream)
{
    if (!ba || !vel || !out || ngrow < 0) {
        set_error("ksgf_cfl_max: bad argument");
        return KSGF_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    CMFDev v;
    int e = cmfdev(ba, vel, KSGF_CELL, 3, ngrow, &v);
    if (e) return e;
    int err = KSGF_OK;
    int z[3] = {0, 0, 0};
    const Region* r = get_region(ba, z, z, &err);
    if (!r) return err;
    unsigned long long* d = nullptr;
    double* h = nullptr;
    int* scratch_err = &err;
    d = (unsigned long long*)get_scratch(ba, 40, 1, scratch_err);
    if (!d) return err;
    KSGF_HIP_CHECK(hipMemsetAsync(d, 0, sizeof(unsigned long long), st));
    int nb = grid_for(r->total, 256);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(k_cfl, dim3(nb), dim3(256), 0, st, v, regdev(r, ba->nbox), ba->dbox,
                       1.0 / ba->dx[0], 1.0 / ba->dx[1], 1.0 / ba->dx[2], d);
    KSGF_LAUNCH_CHECK();
    unsigned long long bits = 0;
    KSGF_HIP_CHECK(hipMemcpyAsync(&bits, d, sizeof(bits), hipMemcpyDeviceToHost, st));
    KSGF_HIP_CHECK(hipStreamSynchronize(st));
    double m;
    std::memcpy(&m, &bits, sizeof(double));
    (void)h;
    // ParReduce floor of -1 (incflo_compute_dt.cpp:86-90): an empty or zero
    // field reports 0 here, which the caller treats as "no constraint"
    *out = m;
    return KSGF_OK;
}

int ksgf_mac_divergence_max(const ksgf_boxarray* ba, const double* u_mac, const double* v_mac, const double* w_mac,
                            int ngrow_mac, double* out, ksgf_stream_t stream)
{
    if (!ba || !u_mac || !v_mac || !w_mac || !out || ngrow_mac < 0) {
        set_error("ksgf_mac_divergence_max: bad argument");
        return KSGF_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    CMFDev u, v, w;
    int e;
    if ((e = cmfdev(ba, u_mac, KSGF_XFACE, 1, ngrow_mac, &u)) || (e = cmfdev(ba, v_mac, KSGF_YFACE, 1, ngrow_mac, &v)) ||
        (e = cmfdev(ba, w_mac, KSGF_ZFACE, 1, ngrow_mac, &w)))
        return e;
    int err = KSGF_OK;
    int z[3] = {0, 0, 0};
    const Region* r = get_region(ba, z, z, &err);
    if (!r) return err;
    unsigned long long* d = (unsigned long long*)get_scratch(ba, 41, 1, &err);
    if (!d) return err;
    KSGF_HIP_CHECK(hipMemsetAsync(d, 0, sizeof(unsigned long long), st));
    int nb = grid_for(r->total, 256);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(k_mac_divmax, dim3(nb), dim3(256), 0, st, u, v, w, regdev(r, ba->nbox), ba->dbox, 1.0 / ba->dx[0],
                       1.0 / ba->dx[1], 1.0 / ba->dx[2], d);
    KSGF_LAUNCH_CHECK();
    unsigned long long bits = 0;
    KSGF_HIP_CHECK(hipMemcpyAsync(&bits, d, sizeof(bits), hipMemcpyDeviceToHost, st));
    KSGF_HIP_CHECK(hipStreamSynchronize(st));
    std::memcpy(out, &bits, sizeof(double));
    return KSGF_OK;
}

int ksgf_predictor_update(const ksgf_boxarray* ba, double* vel_new, int ng_new, const double* vel_old, int ng_old,
                          const double* rho_old, int ng_ro, const double* rho_new, int ng_rn, const double* conv,
                          const double* src, const double* diff, double kappa, double dt, int ncomp,
                          ksgf_stream_t stream)
{
    if (!ba || !vel_new || !vel_old || !rho_old || !rho_new || !conv || !src || ncomp <= 0) {
        set_error("ksgf_predictor_update: null argument");
        return KSGF_ERR_INVALID;
    }
    MFDev vn;
    CMFDev vo, ro, rn, cv, sv, dv;
    int e;
    if ((e = mfdev(ba, vel_new, KSGF_CELL, ncomp, ng_new, &vn)) || (e = cmfdev(ba, vel_old, KSGF_CELL, ncomp, ng_old, &vo)) ||
        (e = cmfdev(ba, rho_old, KSGF_CELL, 1, ng_ro, &ro)) || (e = cmfdev(ba, rho_new, KSGF_CELL, 1, ng_rn, &rn)) ||
        (e = cmfdev(ba, conv, KSGF_CELL, ncomp, 0, &cv)) || (e = cmfdev(ba, src, KSGF_CELL, ncomp, 0, &sv)))
        return e;
    if (diff) {
        if ((e = cmfdev(ba, diff, KSGF_CELL, ncomp, 0, &dv))) return e;
    } else {
        dv = cv;
    }
    int err = KSGF_OK;
    int z[3] = {0, 0, 0};
    const Region* r = get_region(ba, z, z, &err);
    if (!r) return err;
    hipLaunchKernelGGL(k_predictor, dim3(r->nchunks), dim3(256), 0, (hipStream_t)stream, vn, vo, ro, rn,
                       cv, sv, dv, diff ? 1 : 0, kappa, dt, ncomp, regdev(r, ba->nbox), ba->dbox);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

int ksgf_icns_src_term(const ksgf_boxarray* ba, double* src, int ng_src, const double* gp, const double* rho,
                       int ng_rho, int scale_by_rho, ksgf_stream_t stream)
{
    if (!ba || !src || !gp || !rho || ng_src < 0 || ng_rho < 0) {
        set_error("ksgf_icns_src_term: bad argument");
        return KSGF_ERR_INVALID;
    }
    MFDev s;
    CMFDev g, rr;
    int e;
    if ((e = mfdev(ba, src, KSGF_CELL, 3, ng_src, &s)) || (e = cmfdev(ba, gp, KSGF_CELL, 3, 0, &g)) ||
        (e = cmfdev(ba, rho, KSGF_CELL, 1, ng_rho, &rr)))
        return e;
    int err = KSGF_OK;
    int z[3] = {0, 0, 0};
    const Region* r = get_region(ba, z, z, &err);
    if (!r) return err;
    hipLaunchKernelGGL(k_src_term, dim3(r->nchunks), dim3(256), 0, (hipStream_t)stream, s, g, rr,
                       scale_by_rho, regdev(r, ba->nbox), ba->dbox);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

int ksgf_scale_by_rho(const ksgf_boxarray* ba, double* out, int ng_out, const double* a, int ng_a, const double* rho,
                      int ng_rho, int ncomp, int ngrow, int divide, ksgf_stream_t stream)
{
    if (!ba || !out || !a || !rho || ncomp <= 0 || ngrow < 0 || ng_out < ngrow || ng_a < ngrow || ng_rho < ngrow) {
        set_error("ksgf_scale_by_rho: bad argument");
        return KSGF_ERR_INVALID;
    }
    MFDev o;
    CMFDev av, rv;
    int e;
    if ((e = mfdev(ba, out, KSGF_CELL, ncomp, ng_out, &o)) || (e = cmfdev(ba, a, KSGF_CELL, ncomp, ng_a, &av)) ||
        (e = cmfdev(ba, rho, KSGF_CELL, 1, ng_rho, &rv)))
        return e;
    int err = KSGF_OK;
    int g[3] = {ngrow, ngrow, ngrow};
    const Region* r = get_region(ba, g, g, &err);
    if (!r) return err;
    hipLaunchKernelGGL(k_scale_rho, dim3(r->nchunks), dim3(256), 0, (hipStream_t)stream, o, av, rv, ncomp,
                       divide, regdev(r, ba->nbox), ba->dbox);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

int ksgf_icns_src_term_divtau(const ksgf_boxarray* ba, double* src, int ng_src, const double* gp, const double* divtau,
                              const double* rho, int ng_rho, ksgf_stream_t stream)
{
    if (!ba || !src || !gp || !divtau || !rho) {
        set_error("ksgf_icns_src_term_divtau: null argument");
        return KSGF_ERR_INVALID;
    }
    MFDev s;
    CMFDev g, t, rr;
    int e;
    if ((e = mfdev(ba, src, KSGF_CELL, 3, ng_src, &s)) || (e = cmfdev(ba, gp, KSGF_CELL, 3, 0, &g)) ||
        (e = cmfdev(ba, divtau, KSGF_CELL, 3, 0, &t)) || (e = cmfdev(ba, rho, KSGF_CELL, 1, ng_rho, &rr)))
        return e;
    int err = KSGF_OK;
    int z[3] = {0, 0, 0};
    const Region* r = get_region(ba, z, z, &err);
    if (!r) return err;
    hipLaunchKernelGGL(k_src_term_divtau, dim3(r->nchunks), dim3(256), 0, (hipStream_t)stream, s, g, t, rr,
                       regdev(r, ba->nbox), ba->dbox);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

}  // extern "C"
