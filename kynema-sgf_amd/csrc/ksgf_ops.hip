// Streaming operations around the advection/projection path:
//   CFL reduction              incflo_compute_dt.cpp:69-91
//   predictor update           CompRHSOps.H:93-117 (multiply_rho branch)
//   ICNS source term           icns_ops.H:98-111
//   momentum formation         icns_advection.H:281-314 (Copy + Multiply)
#include <cstring>
#include <mutex>
#include <string>

#include <climits>

#include "ksgf_internal.h"

namespace ksgf {

// ---------------------------------------------------------------------------
// ksgf_cfl_max_noalloc: the CFL reduction with its workspace in a file-scope
// device array instead of library-allocated memory.  One partial maximum per
// 256-cell chunk of the valid region, then one workgroup reduces the
// partials into the last slot.  Every call on a device shares that device's
// array: calls that overlap on one device race.
// ---------------------------------------------------------------------------
constexpr int kNoAllocPartials = 1 << 20;  // chunks of 256 cells: up to 268,435,456 valid cells per call
static __device__ __align__(256) double scratch[kNoAllocPartials + 1];

__global__ void __launch_bounds__(256) k_cfl_partial(CMFDev v, RegDev r, const int* __restrict__ box, double dxi,
                                                     double dyi, double dzi, double* __restrict__ part)
{
    __shared__ double sh[4];
    int b, i, j, k;
    double m = 0.0;
    if (region_locate(r, box, b, i, j, k)) {
        CFab f = fab_at(v, box, b);
        double a = fabs(f(i, j, k, 0)) * dxi;
        double bb = fabs(f(i, j, k, 1)) * dyi;
        double e = fabs(f(i, j, k, 2)) * dzi;
        m = fmax(a, fmax(bb, e));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// one workgroup of kMaxPartialsThreads: out = max(part[0 .. n)), eight
// independent loads in flight per thread
constexpr int kMaxPartialsThreads = 1024;
__global__ void __launch_bounds__(kMaxPartialsThreads) k_max_partials(const double* __restrict__ part, int n,
                                                                      double* __restrict__ out)
{
    __shared__ double sh[kMaxPartialsThreads / 64];
    double m = 0.0;
    int t = threadIdx.x;
    for (; t + 7 * kMaxPartialsThreads < n; t += 8 * kMaxPartialsThreads) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part[t + u * kMaxPartialsThreads];
#pragma unroll
        for (int u = 0; u < 8; ++u) m = fmax(m, v[u]);
    }
    for (; t < n; t += kMaxPartialsThreads) m = fmax(m, part[t]);
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double q = sh[0];
        for (int w = 1; w < kMaxPartialsThreads / 64; ++w) q = fmax(q, sh[w]);
        *out = q;
    }
}

// device address of `scratch` on each device, found once per device with that
// device current (each device holds its own copy of the array)
static double* noalloc_scratch(int* err)
{
    static std::mutex mu;
    static double* addr[64] = {};
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        set_error("ksgf_cfl_max_noalloc: no current device");
        *err = KSGF_ERR_HIP;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        double* p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("ksgf_cfl_max_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            *err = KSGF_ERR_HIP;
            return nullptr;
        }
        addr[dev] = p;
    }
    return addr[dev];
}

// the forcing CFL per 256-cell chunk (one workgroup per chunk of the valid
// region: no box search), reduced by k_max_partials
__global__ void __launch_bounds__(256) k_force_cfl_partial(CMFDev s, CMFDev rho, RegDev r, const int* __restrict__ box,
                                                           double dxi, double dyi, double dzi,
                                                           double* __restrict__ part)
{
    __shared__ double sh[4];
    int b, i, j, k;
    double m = 0.0;
    if (region_locate(r, box, b, i, j, k)) {
        CFab f = fab_at(s, box, b);
        const double rr = fab_at(rho, box, b)(i, j, k);
        double a = fabs(f(i, j, k, 0)) * dxi / rr;
        double bb = fabs(f(i, j, k, 1)) * dyi / rr;
        double e = fabs(f(i, j, k, 2)) * dzi / rr;
        m = fmax(a, fmax(bb, e));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// MultiFab::invert(numerator, nghost): a = numerator / a on every point of a region
__global__ void __launch_bounds__(256) k_invert(MFDev m, int ncomp, RegDev r, const int* __restrict__ box,
                                                double numer)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    Fab f = fab_at(m, box, b);
    for (int n = 0; n < ncomp; ++n) f(i, j, k, n) = numer / f(i, j, k, n);
}

// amrex::average_cellcenter_to_face: the mean of the two cells of each face
template <int D>
__global__ void __launch_bounds__(256) k_cc_to_face(CMFDev cc, MFDev fc, RegDev r, const int* __restrict__ box,
                                                    int fix_lo, int fix_hi)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    CFab c = fab_at(cc, box, b);
    int qi = i, qj = j, qk = k;
    if (D == 0) qi -= 1;
    if (D == 1) qj -= 1;
    if (D == 2) qk -= 1;
    const int x = (D == 0) ? i : (D == 1) ? j : k;
    double v;
    if (x == fix_lo)
        v = c(i, j, k);  // a non-periodic domain face: the one cell beside it
    else if (x == fix_hi)
        v = c(qi, qj, qk);
    else
        v = 0.5 * (c(qi, qj, qk) + c(i, j, k));
    fab_at(fc, box, b)(i, j, k) = v;
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

struct GeomFP {
    int fdlo[3], fdhi[3], per[3], cdlo[3];
    int cn0, cn1;
    int cdhi2;  // last coarse cell along z
};

// amrex::average_down: one thread per coarse cell of a coarsened fine box
// (launched over the coarsened copy of the fine BoxArray, box b = fine box b)
__global__ void __launch_bounds__(256) k_avgdown(CMFDev fine, const int* __restrict__ fbox, MFDev crse,
                                                 const int* __restrict__ cbox, const int* __restrict__ owner,
                                                 int ncomp, int ratio, double vf, int dlo0, int dlo1, int dlo2, int n0,
                                                 int n1, RegDev r, const int* __restrict__ rbox)
{
    int b, I, J, K;
    if (!region_locate(r, rbox, b, I, J, K)) return;
    const int cb = owner[(I - dlo0) + n0 * ((J - dlo1) + n1 * (K - dlo2))];
    if (cb < 0) return;
    CFab f = fab_at(fine, fbox, b);
    Fab c = fab_at(crse, cbox, cb);
    for (int n = 0; n < ncomp; ++n) {
        double sum = 0.0;
        for (int kk = 0; kk < ratio; ++kk)
            for (int jj = 0; jj < ratio; ++jj)
                for (int ii = 0; ii < ratio; ++ii) sum += f(I * ratio + ii, J * ratio + jj, K * ratio + kk, n);
        c(I, J, K, n) = sum * vf;
    }
}

// amrex::average_down_faces for faces of direction D: the coarse face goes to
// the FAB of the cell above it, the FAB of the cell below it (when another
// box), and on a periodic direction the image face's FAB
template <int D>
__global__ void __launch_bounds__(256) k_avgdown_faces(CMFDev fine, const int* __restrict__ fbox, MFDev crse,
                                                       const int* __restrict__ cbox, const int* __restrict__ owner,
                                                       int ncomp, int ratio, double af, int dlo0, int dlo1, int dlo2,
                                                       int dhiD, int n0, int n1, int periodic, RegDev r,
                                                       const int* __restrict__ rbox)
{
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    int b, I, J, K;
    if (!region_locate(r, rbox, b, I, J, K)) return;
    CFab f = fab_at(fine, fbox, b);
    const int C[3] = {I, J, K};
    const int dlo[3] = {dlo0, dlo1, dlo2};
    const int nd = dhiD - dlo[D] + 1;
    auto own = [&](int a0, int a1, int a2) {
        return owner[(a0 - dlo0) + n0 * ((a1 - dlo1) + n1 * (a2 - dlo2))];
    };
    // targets: (box, face coordinates)
    int tb[3], tc[3][3], nt = 0;
    if (C[D] <= dhiD) {
        tb[nt] = own(I, J, K);
        for (int e = 0; e < 3; ++e) tc[nt][e] = C[e];
        ++nt;
    }
    if (C[D] > dlo[D]) {
        int q[3] = {I, J, K};
        q[D] -= 1;
        const int ob = own(q[0], q[1], q[2]);
        if (nt == 0 || ob != tb[0]) {
            tb[nt] = ob;
            for (int e = 0; e < 3; ++e) tc[nt][e] = C[e];
            ++nt;
        }
    }
    if (periodic && (C[D] == dlo[D] || C[D] == dhiD + 1)) {
        int p[3] = {I, J, K};
        int q[3] = {I, J, K};
        if (C[D] == dlo[D]) {  // image at the high side, held by the cell below it
            p[D] = C[D] + nd;
            q[D] = dhiD;
        } else {  // image at the low side, held by the cell above it
            p[D] = C[D] - nd;
            q[D] = dlo[D];
        }
        tb[nt] = own(q[0], q[1], q[2]);
        for (int e = 0; e < 3; ++e) tc[nt][e] = p[e];
        ++nt;
    }
    for (int n = 0; n < ncomp; ++n) {
        double sum = 0.0;
        for (int b2 = 0; b2 < ratio; ++b2)
            for (int b1 = 0; b1 < ratio; ++b1) {
                int g3[3] = {C[0] * ratio, C[1] * ratio, C[2] * ratio};
                g3[T1] += b1;
                g3[T2] += b2;
                sum += f(g3[0], g3[1], g3[2], n);
            }
        const double v = sum * af;
        for (int t = 0; t < nt; ++t) {
            if (tb[t] < 0) continue;
            Fab c = fab_at(crse, cbox, tb[t]);
            c(tc[t][0], tc[t][1], tc[t][2], n) = v;
        }
    }
}

// conservative linear interpolation of a refined level's ghost cells from the
// coarse level (FillPatchTwoLevels' interpolation step): MC-limited slopes of
// the parent coarse cell, scaled by one alpha <= 1 that keeps every child in
// [min, max] of the parent and its face neighbours
__device__ __forceinline__ double fp_mc_slope(double um, double uc, double up)
{
    const double dc = 0.5 * (up - um), dl = uc - um, dr = up - uc;
    if (dl * dr > 0.0) return copysign(fmin(fabs(dc), fmin(2.0 * fabs(dl), 2.0 * fabs(dr))), dc);
    return 0.0;
}

__global__ void __launch_bounds__(256) k_fp_interp(MFDev fine, const int* __restrict__ fbox, CMFDev crse,
                                                   const int* __restrict__ cbox, const int* __restrict__ cowner,
                                                   int ncomp, int ratio, double omax, GeomFP g, RegDev r)
{
    int b, i, j, k;
    if (!region_locate(r, fbox, b, i, j, k)) return;
    const int* bx = fbox + 6 * b;
    if (i >= bx[0] && i <= bx[3] && j >= bx[1] && j <= bx[4] && k >= bx[2] && k <= bx[5]) return;
    const int p[3] = {i, j, k};
    int P[3];
    double off[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int L = g.fdhi[d] - g.fdlo[d] + 1;
        int x = p[d];
        if (g.per[d])
            x = g.fdlo[d] + (((x - g.fdlo[d]) % L) + L) % L;
        else if (x < g.fdlo[d] || x > g.fdhi[d])
            return;  // outside a non-periodic domain: physical BC
        P[d] = (x >= 0) ? x / ratio : -((-x + ratio - 1) / ratio);
        off[d] = ((double)(x - ratio * P[d]) + 0.5) / (double)ratio - 0.5;
    }
    const int cb = cowner[(P[0] - g.cdlo[0]) + g.cn0 * ((P[1] - g.cdlo[1]) + g.cn1 * (P[2] - g.cdlo[2]))];
    if (cb < 0) return;
    CFab c = fab_at(crse, cbox, cb);
    Fab f = fab_at(fine, fbox, b);
    const int c0 = c.idx(P[0], P[1], P[2]);
    const int st[3] = {1, c.n0, c.n0 * c.n1};
    for (int n = 0; n < ncomp; ++n) {
        const double* cs = c.p + n * c.cs;
        const double uc = cs[c0];
        double sl[3], umax = uc, umin = uc, dev = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double um = cs[c0 - st[d]], up = cs[c0 + st[d]];
            sl[d] = fp_mc_slope(um, uc, up);
            umax = fmax(umax, fmax(um, up));
            umin = fmin(umin, fmin(um, up));
            dev += fabs(sl[d]) * omax;
        }
        double alpha = 1.0;
        if (dev > 0.0) alpha = fmin(1.0, fmin((umax - uc) / dev, (uc - umin) / dev));
        f(i, j, k, n) = uc + alpha * ((sl[0] * off[0] + sl[1] * off[1]) + sl[2] * off[2]);
    }
}

// face-linear interpolation of a refined level's ghost faces of direction D
// from the coarse faces (ratio 2): on a coarse face its value, midway between
// two coarse faces their mean, piecewise constant across the face
template <int D>
__global__ void __launch_bounds__(256) k_fp_faces(MFDev fine, const int* __restrict__ fbox, CMFDev crse,
                                                  const int* __restrict__ cbox, const int* __restrict__ cowner,
                                                  const int* __restrict__ fowner, int ncomp, GeomFP g, RegDev r)
{
    int b, i, j, k;
    if (!region_locate(r, fbox, b, i, j, k)) return;
    const int* bx = fbox + 6 * b;
    const int p[3] = {i, j, k};
    bool valid = true;
#pragma unroll
    for (int e = 0; e < 3; ++e)
        if (p[e] < bx[e] || p[e] > bx[3 + e] + (e == D)) valid = false;
    if (valid) return;
    int w[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const int L = g.fdhi[e] - g.fdlo[e] + 1;
        const int hi = g.fdhi[e] + (e == D);
        int x = p[e];
        if (g.per[e])
            x = g.fdlo[e] + (((x - g.fdlo[e]) % L) + L) % L;
        else if (x < g.fdlo[e] || x > hi)
            return;  // outside a non-periodic domain: physical BC
        w[e] = x;
    }
    // a valid face of another fine box (a fine cell on either side) keeps
    // that box's value
    {
        const int fn0 = g.fdhi[0] - g.fdlo[0] + 1, fn1 = g.fdhi[1] - g.fdlo[1] + 1, L = g.fdhi[D] - g.fdlo[D] + 1;
        int fb2 = -1;
        if (w[D] <= g.fdhi[D])
            fb2 = fowner[(w[0] - g.fdlo[0]) + fn0 * ((w[1] - g.fdlo[1]) + fn1 * (w[2] - g.fdlo[2]))];
        if (fb2 < 0) {
            int lo[3] = {w[0], w[1], w[2]};
            lo[D] -= 1;
            if (lo[D] < g.fdlo[D]) lo[D] = g.per[D] ? g.fdhi[D] : g.fdlo[D] - 1;
            if (lo[D] >= g.fdlo[D])
                fb2 = fowner[(lo[0] - g.fdlo[0]) + fn0 * ((lo[1] - g.fdlo[1]) + fn1 * (lo[2] - g.fdlo[2]))];
        }
        if (fb2 >= 0) {
            const int* ob = fbox + 6 * fb2;
            int q[3] = {w[0], w[1], w[2]};
            if (q[D] > ob[3 + D] + 1) q[D] -= L;
            if (q[D] < ob[D]) q[D] += L;
            Fab src = fab_at(fine, fbox, fb2);
            Fab f = fab_at(fine, fbox, b);
            for (int n = 0; n < ncomp; ++n) f(i, j, k, n) = src(q[0], q[1], q[2], n);
            return;
        }
    }
    int C[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) C[e] = (w[e] >= 0) ? w[e] / 2 : -((-w[e] + 1) / 2);
    const bool even = (w[D] - 2 * C[D]) == 0;
    const int chi = (D == 0) ? g.cdlo[0] + g.cn0 - 1 : (D == 1 ? g.cdlo[1] + g.cn1 - 1 : g.cdhi2);
    if (even && C[D] > chi) C[D] -= 1;  // the domain's high face: the cell below holds it
    const int cb = cowner[(C[0] - g.cdlo[0]) + g.cn0 * ((C[1] - g.cdlo[1]) + g.cn1 * (C[2] - g.cdlo[2]))];
    if (cb < 0) return;
    CFab c = fab_at(crse, cbox, cb);
    Fab f = fab_at(fine, fbox, b);
    int F0[3] = {C[0], C[1], C[2]};
    if (even) F0[D] = w[D] / 2;
    int F1[3] = {C[0], C[1], C[2]};
    F1[D] = C[D] + 1;
    for (int n = 0; n < ncomp; ++n) {
        double v;
        if (even)
            v = c(F0[0], F0[1], F0[2], n);
        else
            v = 0.5 * (c(C[0], C[1], C[2], n) + c(F1[0], F1[1], F1[2], n));
        f(i, j, k, n) = v;
    }
}

}  // namespace ksgf

using namespace ksgf;

namespace {

// the fine BoxArray coarsened by ratio on the coarse geometry (a level that
// need not cover the domain); checks that the two levels match
int coarsened_level(const ksgf_boxarray* fba, const ksgf_boxarray* cba, int ratio, ksgf_boxarray** out)
{
    if (ratio != 2 && ratio != 4) {
        set_error("average_down: refinement ratio 2 or 4");
        return KSGF_ERR_INVALID;
    }
    for (int d = 0; d < 3; ++d)
        if (fba->dlo[d] != ratio * cba->dlo[d] || fba->dhi[d] + 1 != ratio * (cba->dhi[d] + 1) ||
            fba->periodic[d] != cba->periodic[d]) {
            set_error("average_down: the fine domain is not the coarse domain refined by ratio");
            return KSGF_ERR_INVALID;
        }
    std::vector<int> cb(6 * fba->nbox);
    for (int b = 0; b < fba->nbox; ++b)
        for (int d = 0; d < 3; ++d) {
            const int lo = fba->hbox[6 * b + d], hi = fba->hbox[6 * b + 3 + d];
            if (lo % ratio != 0 || (hi + 1) % ratio != 0) {
                set_error("average_down: a fine box does not coarsen by ratio");
                return KSGF_ERR_INVALID;
            }
            cb[6 * b + d] = lo / ratio;
            cb[6 * b + 3 + d] = (hi + 1) / ratio - 1;
        }
    return ksgf_boxarray_create_level(&cba->geom, cb.data(), fba->nbox, 0, out);
}

}  // namespace

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
    // one partial maximum per 256-cell chunk, then one workgroup reduces them
    // (the maximum does not depend on the order)
    double* part = get_scratch(ba, 43, (size_t)r->nchunks + 1, &err);
    if (!part) return err;
    if (r->nchunks > 0)
        hipLaunchKernelGGL(k_cfl_partial, dim3(r->nchunks), dim3(256), 0, st, v, regdev(r, ba->nbox), ba->dbox,
                           1.0 / ba->dx[0], 1.0 / ba->dx[1], 1.0 / ba->dx[2], part);
    hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(kMaxPartialsThreads), 0, st, (const double*)part, r->nchunks,
                       part + r->nchunks);
    KSGF_LAUNCH_CHECK();
    double m = 0.0;
    KSGF_HIP_CHECK(hipMemcpyAsync(&m, part + r->nchunks, sizeof(m), hipMemcpyDeviceToHost, st));
    KSGF_HIP_CHECK(hipStreamSynchronize(st));
    // ParReduce floor of -1 (incflo_compute_dt.cpp:86-90): an empty or zero
    // field reports 0 here, which the caller treats as "no constraint"
    *out = m;
    return KSGF_OK;
}

int ksgf_invert(const ksgf_boxarray* ba, double* mf, int itype, int ncomp, int ngrow, int nghost, double numerator,
                ksgf_stream_t stream)
{
    if (!ba || !mf || itype < KSGF_CELL || itype > KSGF_ZFACE || ncomp <= 0 || ngrow < 0 || nghost < 0 ||
        nghost > ngrow) {
        set_error("ksgf_invert: bad argument");
        return KSGF_ERR_INVALID;
    }
    MFDev m;
    int e = mfdev(ba, mf, itype, ncomp, ngrow, &m);
    if (e) return e;
    int err = KSGF_OK;
    int glo[3], ghi[3];
    for (int d = 0; d < 3; ++d) {
        glo[d] = nghost;
        ghi[d] = nghost + (itype == d + 1 ? 1 : 0);
    }
    const Region* r = get_region(ba, glo, ghi, &err);
    if (!r) return err;
    hipLaunchKernelGGL(k_invert, dim3(r->nchunks), dim3(256), 0, (hipStream_t)stream, m, ncomp, regdev(r, ba->nbox),
                       ba->dbox, numerator);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

// fixup: faces on a non-periodic domain side take the one cell beside them
static int cc_to_face(const ksgf_boxarray* ba, const double* cc, int ng_cc, double* fx, double* fy, double* fz,
                      int ng_face, int fixup, hipStream_t st)
{
    CMFDev c;
    int e = cmfdev(ba, cc, KSGF_CELL, 1, ng_cc, &c);
    if (e) return e;
    double* out[3] = {fx, fy, fz};
    for (int d = 0; d < 3; ++d) {
        MFDev m;
        if ((e = mfdev(ba, out[d], KSGF_XFACE + d, 1, ng_face, &m))) return e;
        int err = KSGF_OK;
        int glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
        ghi[d] = 1;
        const Region* r = get_region(ba, glo, ghi, &err);
        if (!r) return err;
        dim3 g(r->nchunks);
        const bool fix = fixup && !ba->periodic[d];
        const int flo = fix ? ba->dlo[d] : INT_MIN, fhi = fix ? ba->dhi[d] + 1 : INT_MIN;
        if (d == 0)
            hipLaunchKernelGGL(k_cc_to_face<0>, g, dim3(256), 0, st, c, m, regdev(r, ba->nbox), ba->dbox, flo, fhi);
        else if (d == 1)
            hipLaunchKernelGGL(k_cc_to_face<1>, g, dim3(256), 0, st, c, m, regdev(r, ba->nbox), ba->dbox, flo, fhi);
        else
            hipLaunchKernelGGL(k_cc_to_face<2>, g, dim3(256), 0, st, c, m, regdev(r, ba->nbox), ba->dbox, flo, fhi);
        KSGF_LAUNCH_CHECK();
    }
    return KSGF_OK;
}

int ksgf_average_cellcenter_to_face(const ksgf_boxarray* ba, const double* cc, int ng_cc, double* fx, double* fy,
                                    double* fz, int ng_face, ksgf_stream_t stream)
{
    if (!ba || !cc || !fx || !fy || !fz || ng_cc < 1 || ng_face < 0) {
        set_error("ksgf_average_cellcenter_to_face: bad argument (the cell MultiFab needs a filled ghost layer)");
        return KSGF_ERR_INVALID;
    }
    return cc_to_face(ba, cc, ng_cc, fx, fy, fz, ng_face, 0, (hipStream_t)stream);
}

int ksgf_average_velocity_eta_to_faces(const ksgf_boxarray* ba, const double* cc, int ng_cc, double* fx, double* fy,
                                       double* fz, int ng_face, ksgf_stream_t stream)
{
    if (!ba || !cc || !fx || !fy || !fz || ng_cc < 1 || ng_face < 0) {
        set_error("ksgf_average_velocity_eta_to_faces: bad argument (the cell MultiFab needs a filled ghost layer)");
        return KSGF_ERR_INVALID;
    }
    return cc_to_face(ba, cc, ng_cc, fx, fy, fz, ng_face, 1, (hipStream_t)stream);
}

int ksgf_cfl_max_noalloc(const ksgf_boxarray* ba, const double* vel, int ngrow, double* out, ksgf_stream_t stream)
{
    if (!ba || !vel || !out || ngrow < 0) {
        set_error("ksgf_cfl_max_noalloc: bad argument");
        return KSGF_ERR_INVALID;
    }
    // the BoxArray's launch tables for the valid cells and for the velocity's
    // storage are built when it is created (ngrow 0..3); never build one here
    const int z[3] = {0, 0, 0}, g[3] = {ngrow, ngrow, ngrow};
    const Region* r = find_region(ba, z, z);
    const Region* rv = find_region(ba, g, g);
    if (!r || !rv) {
        set_error("ksgf_cfl_max_noalloc: velocity ngrow " + std::to_string(ngrow) +
                  " has no launch table (0..3 are built with the BoxArray)");
        return KSGF_ERR_INVALID;
    }
    if (r->nchunks > kNoAllocPartials) {
        set_error("ksgf_cfl_max_noalloc: " + std::to_string(r->total) + " valid cells exceed the workspace (at most " +
                  std::to_string((long long)kNoAllocPartials * kChunk) + ")");
        return KSGF_ERR_INVALID;
    }
    int err = KSGF_OK;
    double* part = noalloc_scratch(&err);
    if (!part) return err;
    CMFDev v;
    v.p = vel;
    v.pre = rv->pre;
    v.itype = KSGF_CELL;
    v.ngrow = ngrow;
    v.ncomp = 3;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_cfl_partial, dim3(r->nchunks), dim3(256), 0, st, v, regdev(r, ba->nbox), ba->dbox,
                       1.0 / ba->dx[0], 1.0 / ba->dx[1], 1.0 / ba->dx[2], part);
    KSGF_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(kMaxPartialsThreads), 0, st, part, r->nchunks,
                       part + kNoAllocPartials);
    KSGF_LAUNCH_CHECK();
    double m = 0.0;
    KSGF_HIP_CHECK(hipMemcpyAsync(&m, part + kNoAllocPartials, sizeof(double), hipMemcpyDeviceToHost, st));
    KSGF_HIP_CHECK(hipStreamSynchronize(st));
    *out = m;
    return KSGF_OK;
}

int ksgf_force_cfl_max(const ksgf_boxarray* ba, const double* src, int ng_src, const double* rho, int ng_rho,
                       double* out, ksgf_stream_t stream)
{
    if (!ba || !src || !rho || !out || ng_src < 0 || ng_rho < 0) {
        set_error("ksgf_force_cfl_max: bad argument");
        return KSGF_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    CMFDev s, rh;
    int e = cmfdev(ba, src, KSGF_CELL, 3, ng_src, &s);
    if (e) return e;
    if ((e = cmfdev(ba, rho, KSGF_CELL, 1, ng_rho, &rh))) return e;
    int err = KSGF_OK;
    int z[3] = {0, 0, 0};
    const Region* r = get_region(ba, z, z, &err);
    if (!r) return err;
    double* part = get_scratch(ba, 44, (size_t)r->nchunks + 1, &err);
    if (!part) return err;
    if (r->nchunks > 0)
        hipLaunchKernelGGL(k_force_cfl_partial, dim3(r->nchunks), dim3(256), 0, st, s, rh, regdev(r, ba->nbox),
                           ba->dbox, 1.0 / ba->dx[0], 1.0 / ba->dx[1], 1.0 / ba->dx[2], part);
    hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(kMaxPartialsThreads), 0, st, (const double*)part, r->nchunks,
                       part + r->nchunks);
    KSGF_LAUNCH_CHECK();
    double m = 0.0;
    KSGF_HIP_CHECK(hipMemcpyAsync(&m, part + r->nchunks, sizeof(m), hipMemcpyDeviceToHost, st));
    KSGF_HIP_CHECK(hipStreamSynchronize(st));
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

int ksgf_average_down(const ksgf_boxarray* fine_ba, const double* fine, int ng_f, const ksgf_boxarray* crse_ba,
                      double* crse, int ng_c, int ncomp, int ratio, ksgf_stream_t stream)
{
    if (!fine_ba || !fine || !crse_ba || !crse || ncomp <= 0 || ng_f < 0 || ng_c < 0) {
        set_error("ksgf_average_down: bad argument");
        return KSGF_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    CMFDev f;
    MFDev c;
    int e;
    if ((e = cmfdev(fine_ba, fine, KSGF_CELL, ncomp, ng_f, &f)) || (e = mfdev(crse_ba, crse, KSGF_CELL, ncomp, ng_c, &c)))
        return e;
    const int* owner = nullptr;
    if ((e = owner_map(crse_ba, st, &owner))) return e;
    ksgf_boxarray* cl = nullptr;
    if ((e = coarsened_level(fine_ba, crse_ba, ratio, &cl))) return e;
    int err = KSGF_OK;
    const int z[3] = {0, 0, 0};
    const Region* r = get_region(cl, z, z, &err);
    if (r)
        hipLaunchKernelGGL(k_avgdown, dim3(r->nchunks), dim3(256), 0, st, f, fine_ba->dbox, c, crse_ba->dbox, owner, ncomp,
                           ratio, 1.0 / (double)(ratio * ratio * ratio), crse_ba->dlo[0], crse_ba->dlo[1],
                           crse_ba->dlo[2], crse_ba->dhi[0] - crse_ba->dlo[0] + 1, crse_ba->dhi[1] - crse_ba->dlo[1] + 1,
                           regdev(r, cl->nbox), cl->dbox);
    hipError_t le = hipGetLastError();
    // the coarsened copy's tables may be in use by the launch: release after it
    (void)hipStreamSynchronize(st);
    ksgf_boxarray_destroy(cl);
    if (!r) return err;
    if (le != hipSuccess) {
        set_error(std::string("ksgf_average_down: ") + hipGetErrorString(le));
        return KSGF_ERR_HIP;
    }
    return KSGF_OK;
}

int ksgf_average_down_faces(const ksgf_boxarray* fine_ba, const double* fine_x, const double* fine_y,
                            const double* fine_z, int ng_f, const ksgf_boxarray* crse_ba, double* crse_x,
                            double* crse_y, double* crse_z, int ng_c, int ncomp, int ratio, ksgf_stream_t stream)
{
    const double* fp[3] = {fine_x, fine_y, fine_z};
    double* cp[3] = {crse_x, crse_y, crse_z};
    if (!fine_ba || !crse_ba || !fine_x || !fine_y || !fine_z || !crse_x || !crse_y || !crse_z || ncomp <= 0 ||
        ng_f < 0 || ng_c < 0) {
        set_error("ksgf_average_down_faces: bad argument");
        return KSGF_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    int e;
    const int* owner = nullptr;
    if ((e = owner_map(crse_ba, st, &owner))) return e;
    ksgf_boxarray* cl = nullptr;
    if ((e = coarsened_level(fine_ba, crse_ba, ratio, &cl))) return e;
    const double af = 1.0 / (double)(ratio * ratio);
    const int n0 = crse_ba->dhi[0] - crse_ba->dlo[0] + 1, n1 = crse_ba->dhi[1] - crse_ba->dlo[1] + 1;
    int ret = KSGF_OK;
    for (int d = 0; d < 3 && ret == KSGF_OK; ++d) {
        CMFDev f;
        MFDev c;
        if ((ret = cmfdev(fine_ba, fp[d], KSGF_XFACE + d, ncomp, ng_f, &f)) ||
            (ret = mfdev(crse_ba, cp[d], KSGF_XFACE + d, ncomp, ng_c, &c)))
            break;
        int err = KSGF_OK;
        int glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
        ghi[d] = 1;
        const Region* r = get_region(cl, glo, ghi, &err);
        if (!r) {
            ret = err;
            break;
        }
        const RegDev rd = regdev(r, cl->nbox);
        if (d == 0)
            hipLaunchKernelGGL(k_avgdown_faces<0>, dim3(r->nchunks), dim3(256), 0, st, f, fine_ba->dbox, c,
                               crse_ba->dbox, owner, ncomp, ratio, af, crse_ba->dlo[0], crse_ba->dlo[1],
                               crse_ba->dlo[2], crse_ba->dhi[0], n0, n1, crse_ba->periodic[0], rd, cl->dbox);
        else if (d == 1)
            hipLaunchKernelGGL(k_avgdown_faces<1>, dim3(r->nchunks), dim3(256), 0, st, f, fine_ba->dbox, c,
                               crse_ba->dbox, owner, ncomp, ratio, af, crse_ba->dlo[0], crse_ba->dlo[1],
                               crse_ba->dlo[2], crse_ba->dhi[1], n0, n1, crse_ba->periodic[1], rd, cl->dbox);
        else
            hipLaunchKernelGGL(k_avgdown_faces<2>, dim3(r->nchunks), dim3(256), 0, st, f, fine_ba->dbox, c,
                               crse_ba->dbox, owner, ncomp, ratio, af, crse_ba->dlo[0], crse_ba->dlo[1],
                               crse_ba->dlo[2], crse_ba->dhi[2], n0, n1, crse_ba->periodic[2], rd, cl->dbox);
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) {
            set_error(std::string("ksgf_average_down_faces: ") + hipGetErrorString(le));
            ret = KSGF_ERR_HIP;
        }
    }
    (void)hipStreamSynchronize(st);
    ksgf_boxarray_destroy(cl);
    return ret;
}

int ksgf_fillpatch_two_levels(const ksgf_boxarray* fine_ba, double* fine, int ng, const ksgf_boxarray* crse_ba,
                              const double* crse, int ng_c, int ncomp, int ratio, const ksgf_bcrec* bc,
                              const double* bcval, ksgf_stream_t stream)
{
    if (!fine_ba || !fine || !crse_ba || !crse || !bc || !bcval || ncomp <= 0 || ncomp > 8 || ng < 0 || ng_c < 1 ||
        (ratio != 2 && ratio != 4)) {
        set_error("ksgf_fillpatch_two_levels: bad argument (coarse data needs one ghost layer)");
        return KSGF_ERR_INVALID;
    }
    for (int d = 0; d < 3; ++d)
        if (fine_ba->dlo[d] != ratio * crse_ba->dlo[d] || fine_ba->dhi[d] + 1 != ratio * (crse_ba->dhi[d] + 1) ||
            fine_ba->periodic[d] != crse_ba->periodic[d]) {
            set_error("ksgf_fillpatch_two_levels: the fine domain is not the coarse domain refined by ratio");
            return KSGF_ERR_INVALID;
        }
    if (ng == 0) return KSGF_OK;
    hipStream_t st = (hipStream_t)stream;
    MFDev f;
    CMFDev c;
    int e;
    if ((e = mfdev(fine_ba, fine, KSGF_CELL, ncomp, ng, &f)) || (e = cmfdev(crse_ba, crse, KSGF_CELL, ncomp, ng_c, &c)))
        return e;
    const int* owner = nullptr;
    if ((e = owner_map(crse_ba, st, &owner))) return e;
    int err = KSGF_OK;
    const Region* r = fab_region(fine_ba, KSGF_CELL, ng, &err);
    if (!r) return err;
    GeomFP g;
    for (int d = 0; d < 3; ++d) {
        g.fdlo[d] = fine_ba->dlo[d];
        g.fdhi[d] = fine_ba->dhi[d];
        g.per[d] = fine_ba->periodic[d];
        g.cdlo[d] = crse_ba->dlo[d];
    }
    g.cn0 = crse_ba->dhi[0] - crse_ba->dlo[0] + 1;
    g.cn1 = crse_ba->dhi[1] - crse_ba->dlo[1] + 1;
    g.cdhi2 = crse_ba->dhi[2];
    hipLaunchKernelGGL(k_fp_interp, dim3(r->nchunks), dim3(256), 0, st, f, fine_ba->dbox, c, crse_ba->dbox, owner, ncomp,
                       ratio, 0.5 - 0.5 / (double)ratio, g, regdev(r, fine_ba->nbox));
    KSGF_LAUNCH_CHECK();
    if ((e = ksgf_fill_boundary(fine_ba, fine, KSGF_CELL, ncomp, ng, stream))) return e;
    return ksgf_fill_physbc(fine_ba, fine, KSGF_CELL, ncomp, ng, bc, bcval, stream);
}

int ksgf_fillpatch_faces_two_levels(const ksgf_boxarray* fine_ba, double* fine, int itype, int ng,
                                    const ksgf_boxarray* crse_ba, const double* crse, int ng_c, int ncomp,
                                    const ksgf_bcrec* bc, const double* bcval, ksgf_stream_t stream)
{
    if (!fine_ba || !fine || !crse_ba || !crse || !bc || !bcval || ncomp <= 0 || ncomp > 8 || ng < 0 || ng_c < 0 ||
        itype < KSGF_XFACE || itype > KSGF_ZFACE) {
        set_error("ksgf_fillpatch_faces_two_levels: bad argument");
        return KSGF_ERR_INVALID;
    }
    for (int d = 0; d < 3; ++d)
        if (fine_ba->dlo[d] != 2 * crse_ba->dlo[d] || fine_ba->dhi[d] + 1 != 2 * (crse_ba->dhi[d] + 1) ||
            fine_ba->periodic[d] != crse_ba->periodic[d]) {
            set_error("ksgf_fillpatch_faces_two_levels: the fine domain is not the coarse domain refined by 2");
            return KSGF_ERR_INVALID;
        }
    if (ng == 0) return KSGF_OK;
    hipStream_t st = (hipStream_t)stream;
    MFDev f;
    CMFDev c;
    int e;
    if ((e = mfdev(fine_ba, fine, itype, ncomp, ng, &f)) || (e = cmfdev(crse_ba, crse, itype, ncomp, ng_c, &c)))
        return e;
    const int *owner = nullptr, *fowner = nullptr;
    if ((e = owner_map(crse_ba, st, &owner)) || (e = owner_map(fine_ba, st, &fowner))) return e;
    int err = KSGF_OK;
    const Region* r = fab_region(fine_ba, itype, ng, &err);
    if (!r) return err;
    GeomFP g;
    for (int d = 0; d < 3; ++d) {
        g.fdlo[d] = fine_ba->dlo[d];
        g.fdhi[d] = fine_ba->dhi[d];
        g.per[d] = fine_ba->periodic[d];
        g.cdlo[d] = crse_ba->dlo[d];
    }
    g.cn0 = crse_ba->dhi[0] - crse_ba->dlo[0] + 1;
    g.cn1 = crse_ba->dhi[1] - crse_ba->dlo[1] + 1;
    g.cdhi2 = crse_ba->dhi[2];
    const RegDev rd = regdev(r, fine_ba->nbox);
    switch (itype) {
    case KSGF_XFACE:
        hipLaunchKernelGGL(k_fp_faces<0>, dim3(r->nchunks), dim3(256), 0, st, f, fine_ba->dbox, c, crse_ba->dbox,
                           owner, fowner, ncomp, g, rd);
        break;
    case KSGF_YFACE:
        hipLaunchKernelGGL(k_fp_faces<1>, dim3(r->nchunks), dim3(256), 0, st, f, fine_ba->dbox, c, crse_ba->dbox,
                           owner, fowner, ncomp, g, rd);
        break;
    default:
        hipLaunchKernelGGL(k_fp_faces<2>, dim3(r->nchunks), dim3(256), 0, st, f, fine_ba->dbox, c, crse_ba->dbox,
                           owner, fowner, ncomp, g, rd);
        break;
    }
    KSGF_LAUNCH_CHECK();
    if ((e = ksgf_fill_boundary(fine_ba, fine, itype, ncomp, ng, stream))) return e;
    return ksgf_fill_physbc(fine_ba, fine, itype, ncomp, ng, bc, bcval, stream);
}

}  // extern "C"
