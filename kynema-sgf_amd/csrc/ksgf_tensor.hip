// amrex::MLTensorOp as the default ICNS viscous operator uses it
// (ICNSDiffTensorOp, icns_diffusion.H:18-50; DiffSolverIface::setup_operator
// and set_bcoeffs, DiffusionOps.cpp:57-110, DiffusionOps.H:54-71):
//   L u = alpha a u - beta div(tau),
//   tau = eta (grad u + grad u^T) - (2/3) eta (div u) I   (no bulk viscosity).
// apply is MLMG::apply (compute_diff_term: scalars (0, -1), divtau = div tau);
// solve is MLMG::solve of linsys_solve (scalars (1, dt_diff), a = rho).
//
// Face stress of face f in direction d (between cells f - e_d and f), for
// component n:  tau_dn = eta_d(f) g,
//   d == n: g = (4/3) N_n - (2/3) sum_{m != n, ascending} T_mm
//   d != n: g = N_n + T_dn
// N_n: the difference of u_n across the face over dx_d; T_qm: the derivative
// of u_q along m on the face, the mean of the central differences of the two
// cells beside it.  Per component and domain side (LinOpBCType, homogeneous):
// Dirichlet puts 0 on the face (ghost -u; derivatives of the boundary value
// along the face are 0), Neumann gives N = 0 and takes the derivatives along
// the face from the interior cell alone, periodic wraps.  Every expression is
// evaluated in the CPU restatement's operand order (no contraction), so apply
// agrees with it bit for bit.
//
// The solve is MLMG's outer iteration in FP64 on domain-wide arrays:
//   r = rhs - L u (the full tensor operator, one kernel, with |r|_inf),
//   c_n = one V-cycle of the component's own block (alpha a - beta div(eta_n
//   grad)) with eta_n = eta, 4/3 eta on the faces normal to n (the variable-
//   coefficient multigrid with a zeroth-order term), u += c,
// until |r|_inf <= max(atol, rtol |rhs|_inf).  The blocks carry the 4/3
// normal coefficient; the cross terms enter through the residual only (with
// exact blocks the iteration contracts by at most 1/3 per step for constant
// eta).  u starts from zero.
#include <cmath>
#include <cstring>

#include "ksgf_internal.h"

using namespace ksgf;

namespace {

constexpr int kTnBlocks = 1024;

struct TnDev {
    int n[3];
    double dxi[3];
    int lo[9], hi[9];  // [component * 3 + direction]
    const double* u;   // 3 components, component slowest
    long long nc;
};

__device__ __forceinline__ double tn_at(const TnDev& T, int q, const int* p)
{
    return T.u[q * T.nc + p[0] + (long long)T.n[0] * (p[1] + (long long)T.n[1] * p[2])];
}

__device__ __forceinline__ double tn_nb(const TnDev& T, int q, const int* p, int m, int s)
{
    int c[3] = {p[0], p[1], p[2]};
    c[m] += s;
    if (c[m] < 0 || c[m] >= T.n[m]) {
        const int bc = (s < 0) ? T.lo[q * 3 + m] : T.hi[q * 3 + m];
        if (bc != KSGF_LO_PERIODIC) {
            const double v = tn_at(T, q, p);
            return (bc == KSGF_LO_NEUMANN) ? v : -v;
        }
        c[m] = (c[m] < 0) ? c[m] + T.n[m] : c[m] - T.n[m];
    }
    return tn_at(T, q, c);
}

__device__ __forceinline__ double tn_diff(const TnDev& T, int q, int m, const int* p)
{
    return tn_nb(T, q, p, m, 1) - tn_nb(T, q, p, m, -1);
}

// T_qm on face f of direction d (m != d)
__device__ double tn_tr(const TnDev& T, int q, int m, int d, const int* f)
{
    int a[3] = {f[0], f[1], f[2]}, b[3] = {f[0], f[1], f[2]};
    b[d] -= 1;
    const bool at_lo = (f[d] == 0), at_hi = (f[d] == T.n[d]);
    if (at_lo || at_hi) {
        const int bc = at_lo ? T.lo[q * 3 + d] : T.hi[q * 3 + d];
        if (bc == KSGF_LO_DIRICHLET) return 0.0;
        if (bc == KSGF_LO_NEUMANN) return tn_diff(T, q, m, at_lo ? a : b) * (0.5 * T.dxi[m]);
        if (at_lo)
            b[d] = T.n[d] - 1;
        else
            a[d] = 0;
    }
    return (tn_diff(T, q, m, a) + tn_diff(T, q, m, b)) * (0.25 * T.dxi[m]);
}

// N_q across face f of direction d
__device__ double tn_nd(const TnDev& T, int q, int d, const int* f)
{
    int a[3] = {f[0], f[1], f[2]}, b[3] = {f[0], f[1], f[2]};
    b[d] -= 1;
    const bool at_lo = (f[d] == 0), at_hi = (f[d] == T.n[d]);
    if (at_lo || at_hi) {
        const int bc = at_lo ? T.lo[q * 3 + d] : T.hi[q * 3 + d];
        if (bc == KSGF_LO_NEUMANN) return 0.0;
        if (bc == KSGF_LO_DIRICHLET) {
            const double v = tn_at(T, q, at_lo ? a : b);
            return at_lo ? (v - (-v)) * T.dxi[d] : ((-v) - v) * T.dxi[d];
        }
        if (at_lo)
            b[d] = T.n[d] - 1;
        else
            a[d] = 0;
    }
    return (tn_at(T, q, a) - tn_at(T, q, b)) * T.dxi[d];
}

__device__ __forceinline__ long long tn_fidx(const TnDev& T, int d, const int* f)
{
    const int m0 = T.n[0] + (d == 0), m1 = T.n[1] + (d == 1);
    return f[0] + (long long)m0 * (f[1] + (long long)m1 * f[2]);
}

__device__ double tn_tau(const TnDev& T, const double* __restrict__ eta, int n, int d, const int* f)
{
    double g;
    if (d == n) {
        g = (4.0 / 3.0) * tn_nd(T, n, d, f);
        double s = 0.0;
        for (int m = 0; m < 3; ++m)
            if (m != n) s += tn_tr(T, m, m, d, f);
        g = g - (2.0 / 3.0) * s;
    } else {
        g = tn_nd(T, n, d, f);
        g = g + tn_tr(T, d, n, d, f);
    }
    return eta[tn_fidx(T, d, f)] * g;
}

__device__ __forceinline__ double tn_cell(const TnDev& T, int q, int i, int j, int k, double alpha, double av,
                                          double beta, const double* __restrict__ ex, const double* __restrict__ ey,
                                          const double* __restrict__ ez)
{
    const int c[3] = {i, j, k};
    const double* eta[3] = {ex, ey, ez};
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int f1[3] = {i, j, k};
        f1[d] += 1;
        const double t = (tn_tau(T, eta[d], q, d, f1) - tn_tau(T, eta[d], q, d, c)) * T.dxi[d];
        s = (d == 0) ? t : s + t;
    }
    return alpha * av * tn_at(T, q, c) - beta * s;
}

// MODE 0: out = L u; MODE 1: r = rhs - L u, f_q = -r_q (the blocks' level-0
// right-hand sides) and the block maxima of |r| (NaN as +inf)
template <int MODE>
__global__ void __launch_bounds__(256) k_tn_apply(TnDev T, const double* __restrict__ a, double alpha, double beta,
                                                  const double* __restrict__ ex, const double* __restrict__ ey,
                                                  const double* __restrict__ ez, double* __restrict__ out,
                                                  const double* __restrict__ rhs, double* __restrict__ f0,
                                                  double* __restrict__ f1, double* __restrict__ f2,
                                                  double* __restrict__ part)
{
    __shared__ double sh[4];
    const long long tot = 3 * T.nc;
    double mx = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
         t += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(t / T.nc);
        const long long c = t - (long long)q * T.nc;
        const long long qq = c / T.n[0];
        const int i = (int)(c - qq * T.n[0]), j = (int)(qq % T.n[1]), k = (int)(qq / T.n[1]);
        const double av = a ? a[c] : 0.0;
        const double v = tn_cell(T, q, i, j, k, alpha, av, beta, ex, ey, ez);
        if (MODE == 0) {
            out[t] = v;
        } else {
            const double r = rhs[t] - v;
            double* f = (q == 0) ? f0 : (q == 1) ? f1 : f2;
            f[c] = -r;
            const double ar = fabs(r);
            mx = (ar == ar) ? fmax(mx, ar) : INFINITY;
        }
    }
    if (MODE == 1) {
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_down(mx, off, 64);
            mx = (o == o) ? fmax(mx, o) : INFINITY;
        }
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = sh[0];
            for (int w = 1; w < 4; ++w) m = (sh[w] == sh[w]) ? fmax(m, sh[w]) : INFINITY;
            part[blockIdx.x] = m;
        }
    }
}

__global__ void __launch_bounds__(256) k_tn_absmax(const double* __restrict__ x, long long n,
                                                   double* __restrict__ part)
{
    __shared__ double sh[4];
    double mx = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const double ar = fabs(x[t]);
        mx = (ar == ar) ? fmax(mx, ar) : INFINITY;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(mx, off, 64);
        mx = (o == o) ? fmax(mx, o) : INFINITY;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sh[0];
        for (int w = 1; w < 4; ++w) m = (sh[w] == sh[w]) ? fmax(m, sh[w]) : INFINITY;
        part[blockIdx.x] = m;
    }
}

__global__ void k_tn_max_partials(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    if (threadIdx.x != 0) return;
    double m = 0.0;
    for (int b = 0; b < nb; ++b) m = (part[b] == part[b]) ? fmax(m, part[b]) : INFINITY;
    *out = m;
}

// u_q += e_q
__global__ void __launch_bounds__(256) k_tn_add3(double* __restrict__ u, const double* __restrict__ e0,
                                                 const double* __restrict__ e1, const double* __restrict__ e2,
                                                 long long nc)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nc) return;
    const int q = (int)(t / nc);
    const long long c = t - (long long)q * nc;
    const double* e = (q == 0) ? e0 : (q == 1) ? e1 : e2;
    u[t] += e[c];
}

__global__ void __launch_bounds__(256) k_tn_fill(double* __restrict__ x, long long n, double v)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] = v;
}

// valid cells of ncomp components of a cell MultiFab <-> domain-wide arrays
// (component slowest)
template <int TO_DOMAIN>
__global__ void __launch_bounds__(256) k_tn_cells(MFDev m, CMFDev cm, RegDev r, const int* __restrict__ box, int dlo0,
                                                  int dlo1, int dlo2, int n0, int n1, long long nc, int ncomp,
                                                  double* __restrict__ dom, const double* __restrict__ cdom)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    const long long c = (i - dlo0) + (long long)n0 * ((j - dlo1) + (long long)n1 * (k - dlo2));
    if (TO_DOMAIN) {
        CFab f = fab_at(cm, box, b);
        for (int q = 0; q < ncomp; ++q) dom[q * nc + c] = f(i, j, k, q);
    } else {
        Fab f = fab_at(m, box, b);
        for (int q = 0; q < ncomp; ++q) f(i, j, k, q) = cdom[q * nc + c];
    }
}

// valid faces of direction D of a face MultiFab into the domain-wide array
template <int D>
__global__ void __launch_bounds__(256) k_tn_faces(CMFDev m, RegDev r, const int* __restrict__ box, int dlo0, int dlo1,
                                                  int dlo2, int n0, int n1, double* __restrict__ out)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    const int m0 = n0 + (D == 0), m1 = n1 + (D == 1);
    out[(i - dlo0) + (long long)m0 * ((j - dlo1) + (long long)m1 * (k - dlo2))] = fab_at(m, box, b)(i, j, k);
}

int grid_of(long long n) { return (int)((n + 255) / 256); }

}  // namespace

struct ksgf_tensorop {
    const ksgf_boxarray* ba = nullptr;
    int n[3];
    double dx[3];
    int lo[9], hi[9];
    double alpha = NAN, beta = NAN;  // MLABecLaplacian's scalars start unset
    ksgf_mlmg_options opt;
    long long nc = 0;
    double *u = nullptr, *rhs = nullptr, *a = nullptr, *eta[3] = {nullptr, nullptr, nullptr};
    double *part = nullptr, *hbuf = nullptr;
    int have_a = 0;
    int coefs_dirty = 1;
    VbSolver* vb[3] = {nullptr, nullptr, nullptr};
};

static void tn_free(ksgf_tensorop* op)
{
    if (!op) return;
    for (int c = 0; c < 3; ++c)
        if (op->vb[c]) vb_destroy(op->vb[c]);
    double* arrs[] = {op->u, op->rhs, op->a, op->eta[0], op->eta[1], op->eta[2], op->part};
    for (double* p : arrs)
        if (p) (void)hipFree(p);
    if (op->hbuf) (void)hipHostFree(op->hbuf);
    delete op;
}

static TnDev tn_dev(const ksgf_tensorop* op, const double* u)
{
    TnDev T;
    for (int d = 0; d < 3; ++d) {
        T.n[d] = op->n[d];
        T.dxi[d] = 1.0 / op->dx[d];
    }
    for (int t = 0; t < 9; ++t) {
        T.lo[t] = op->lo[t];
        T.hi[t] = op->hi[t];
    }
    T.u = u;
    T.nc = op->nc;
    return T;
}

static int tn_cells(const ksgf_tensorop* op, double* mf, const double* cmf, int ng, int to_domain, double* dom,
                    const double* cdom, int ncomp, hipStream_t st)
{
    const ksgf_boxarray* ba = op->ba;
    MFDev m{};
    CMFDev cm{};
    int e = to_domain ? cmfdev(ba, cmf, KSGF_CELL, ncomp, ng, &cm) : mfdev(ba, mf, KSGF_CELL, ncomp, ng, &m);
    if (e) return e;
    int err = KSGF_OK;
    const int z[3] = {0, 0, 0};
    const Region* r = get_region(ba, z, z, &err);
    if (!r) return err;
    if (to_domain)
        hipLaunchKernelGGL(k_tn_cells<1>, dim3(r->nchunks), dim3(256), 0, st, m, cm, regdev(r, ba->nbox), ba->dbox,
                           ba->dlo[0], ba->dlo[1], ba->dlo[2], op->n[0], op->n[1], op->nc, ncomp, dom, cdom);
    else
        hipLaunchKernelGGL(k_tn_cells<0>, dim3(r->nchunks), dim3(256), 0, st, m, cm, regdev(r, ba->nbox), ba->dbox,
                           ba->dlo[0], ba->dlo[1], ba->dlo[2], op->n[0], op->n[1], op->nc, ncomp, dom, cdom);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

static int tn_ready(const ksgf_tensorop* op, const char* who)
{
    if (!(std::isfinite(op->alpha) && std::isfinite(op->beta))) {
        set_error(std::string(who) + ": call ksgf_tensorop_set_scalars first");
        return KSGF_ERR_INVALID;
    }
    return KSGF_OK;
}

extern "C" {

int ksgf_tensorop_create(const ksgf_boxarray* ba, const int* lobc, const int* hibc, const ksgf_mlmg_options* opts,
                         ksgf_tensorop** out)
{
    if (!ba || !lobc || !hibc || !out) {
        set_error("ksgf_tensorop_create: bad argument");
        return KSGF_ERR_INVALID;
    }
    *out = nullptr;
    for (int t = 0; t < 9; ++t)
        if (lobc[t] < KSGF_LO_PERIODIC || lobc[t] > KSGF_LO_DIRICHLET || hibc[t] < KSGF_LO_PERIODIC ||
            hibc[t] > KSGF_LO_DIRICHLET) {
            set_error("ksgf_tensorop_create: a domain BC is not KSGF_LO_PERIODIC / NEUMANN / DIRICHLET");
            return KSGF_ERR_INVALID;
        }
    for (int d = 0; d < 3; ++d) {
        const bool per = ba->periodic[d] != 0;
        for (int c = 0; c < 3; ++c)
            if ((lobc[c * 3 + d] == KSGF_LO_PERIODIC) != per || (hibc[c * 3 + d] == KSGF_LO_PERIODIC) != per) {
                set_error("ksgf_tensorop_create: periodic BCs must be exactly the geometry's periodic directions");
                return KSGF_ERR_INVALID;
            }
    }
    ksgf_tensorop* op = new ksgf_tensorop();
    op->ba = ba;
    if (opts)
        op->opt = *opts;
    else
        ksgf_mlmg_default_options(&op->opt);
    for (int d = 0; d < 3; ++d) {
        op->n[d] = ba->dhi[d] - ba->dlo[d] + 1;
        op->dx[d] = ba->dx[d];
    }
    std::memcpy(op->lo, lobc, sizeof(op->lo));
    std::memcpy(op->hi, hibc, sizeof(op->hi));
    op->nc = (long long)op->n[0] * op->n[1] * op->n[2];
    for (int c = 0; c < 3; ++c) {
        int err = KSGF_OK;
        op->vb[c] = vb_create(ba, lobc + 3 * c, hibc + 3 * c, op->opt, &err, 1);
        if (!op->vb[c]) {
            tn_free(op);
            return err;
        }
    }
    auto alloc = [&](double** p, long long n) {
        return hipMalloc(p, sizeof(double) * (size_t)n) == hipSuccess &&
               hipMemset(*p, 0, sizeof(double) * (size_t)n) == hipSuccess;
    };
    bool ok = alloc(&op->u, 3 * op->nc) && alloc(&op->rhs, 3 * op->nc) && alloc(&op->a, op->nc) &&
              alloc(&op->part, kTnBlocks + 1);
    for (int d = 0; d < 3 && ok; ++d) ok = alloc(&op->eta[d], op->nc / op->n[d] * (op->n[d] + 1));
    if (ok) ok = hipHostMalloc((void**)&op->hbuf, sizeof(double)) == hipSuccess;
    if (!ok) {
        tn_free(op);
        set_error("ksgf_tensorop_create: out of device memory");
        return KSGF_ERR_NOMEM;
    }
    // MLABecLaplacian's defaults: a = 0, the face coefficients 1
    for (int d = 0; d < 3; ++d) {
        const long long m = op->nc / op->n[d] * (op->n[d] + 1);
        hipLaunchKernelGGL(k_tn_fill, dim3(grid_of(m)), dim3(256), 0, 0, op->eta[d], m, 1.0);
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        tn_free(op);
        set_error("ksgf_tensorop_create: initialising the coefficients failed");
        return KSGF_ERR_HIP;
    }
    *out = op;
    return KSGF_OK;
}

int ksgf_tensorop_set_scalars(ksgf_tensorop* op, double alpha, double beta)
{
    if (!op || !std::isfinite(alpha) || !std::isfinite(beta)) {
        set_error("ksgf_tensorop_set_scalars: bad argument");
        return KSGF_ERR_INVALID;
    }
    if (alpha != op->alpha || beta != op->beta) op->coefs_dirty = 1;
    op->alpha = alpha;
    op->beta = beta;
    return KSGF_OK;
}

int ksgf_tensorop_set_acoeffs(ksgf_tensorop* op, const double* a, int ngrow, ksgf_stream_t stream)
{
    if (!op || !a || ngrow < 0) {
        set_error("ksgf_tensorop_set_acoeffs: bad argument");
        return KSGF_ERR_INVALID;
    }
    const int e = tn_cells(op, nullptr, a, ngrow, 1, op->a, nullptr, 1, (hipStream_t)stream);
    if (e) return e;
    op->have_a = 1;
    op->coefs_dirty = 1;
    return KSGF_OK;
}

int ksgf_tensorop_set_shear_viscosity(ksgf_tensorop* op, const double* eta_x, const double* eta_y,
                                      const double* eta_z, int ngrow, ksgf_stream_t stream)
{
    if (!op || !eta_x || !eta_y || !eta_z || ngrow < 0) {
        set_error("ksgf_tensorop_set_shear_viscosity: bad argument");
        return KSGF_ERR_INVALID;
    }
    const ksgf_boxarray* ba = op->ba;
    hipStream_t st = (hipStream_t)stream;
    const double* src[3] = {eta_x, eta_y, eta_z};
    for (int d = 0; d < 3; ++d) {
        CMFDev m;
        int e = cmfdev(ba, src[d], KSGF_XFACE + d, 1, ngrow, &m);
        if (e) return e;
        int glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
        ghi[d] = 1;
        int err = KSGF_OK;
        const Region* r = get_region(ba, glo, ghi, &err);
        if (!r) return err;
        dim3 g(r->nchunks);
        if (d == 0)
            hipLaunchKernelGGL(k_tn_faces<0>, g, dim3(256), 0, st, m, regdev(r, ba->nbox), ba->dbox, ba->dlo[0],
                               ba->dlo[1], ba->dlo[2], op->n[0], op->n[1], op->eta[0]);
        else if (d == 1)
            hipLaunchKernelGGL(k_tn_faces<1>, g, dim3(256), 0, st, m, regdev(r, ba->nbox), ba->dbox, ba->dlo[0],
                               ba->dlo[1], ba->dlo[2], op->n[0], op->n[1], op->eta[1]);
        else
            hipLaunchKernelGGL(k_tn_faces<2>, g, dim3(256), 0, st, m, regdev(r, ba->nbox), ba->dbox, ba->dlo[0],
                               ba->dlo[1], ba->dlo[2], op->n[0], op->n[1], op->eta[2]);
        KSGF_LAUNCH_CHECK();
    }
    op->coefs_dirty = 1;
    return KSGF_OK;
}

int ksgf_tensorop_apply(ksgf_tensorop* op, double* out, int ng_out, const double* in, int ng_in,
                        ksgf_stream_t stream)
{
    if (!op || !out || !in || ng_out < 0 || ng_in < 0) {
        set_error("ksgf_tensorop_apply: bad argument");
        return KSGF_ERR_INVALID;
    }
    int e = tn_ready(op, "ksgf_tensorop_apply");
    if (e) return e;
    hipStream_t st = (hipStream_t)stream;
    if ((e = tn_cells(op, nullptr, in, ng_in, 1, op->u, nullptr, 3, st))) return e;
    const TnDev T = tn_dev(op, op->u);
    hipLaunchKernelGGL(k_tn_apply<0>, dim3(grid_of(3 * op->nc)), dim3(256), 0, st, T, (const double*)op->a,
                       op->alpha, op->beta, (const double*)op->eta[0], (const double*)op->eta[1],
                       (const double*)op->eta[2], op->rhs, (const double*)nullptr, (double*)nullptr,
                       (double*)nullptr, (double*)nullptr, (double*)nullptr);
    KSGF_LAUNCH_CHECK();
    return tn_cells(op, out, nullptr, ng_out, 0, nullptr, op->rhs, 3, st);
}

int ksgf_tensorop_solve(ksgf_tensorop* op, double* sol, int ng_sol, const double* rhs, int ng_rhs, double rtol,
                        double atol, ksgf_mlmg_info* info, ksgf_stream_t stream)
{
    if (!op || !sol || !rhs || ng_sol < 0 || ng_rhs < 0 || !(rtol >= 0.0) || !(atol >= 0.0)) {
        set_error("ksgf_tensorop_solve: bad argument");
        return KSGF_ERR_INVALID;
    }
    int e = tn_ready(op, "ksgf_tensorop_solve");
    if (e) return e;
    if (!(op->beta > 0.0) || op->alpha < 0.0) {
        set_error("ksgf_tensorop_solve: needs beta > 0 and alpha >= 0");
        return KSGF_ERR_INVALID;
    }
    if (op->alpha == 0.0 || !op->have_a)
        for (int c = 0; c < 3; ++c) {
            bool dir = false;
            for (int d = 0; d < 3; ++d)
                dir = dir || op->lo[c * 3 + d] == KSGF_LO_DIRICHLET || op->hi[c * 3 + d] == KSGF_LO_DIRICHLET;
            if (!dir) {
                set_error("ksgf_tensorop_solve: singular (alpha a = 0 and a component without a Dirichlet side)");
                return KSGF_ERR_INVALID;
            }
        }
    hipStream_t st = (hipStream_t)stream;
    if (op->coefs_dirty) {
        for (int c = 0; c < 3; ++c) {
            const double* b3[3] = {op->eta[0], op->eta[1], op->eta[2]};
            double scale[3];
            for (int d = 0; d < 3; ++d) scale[d] = op->beta * (d == c ? 4.0 / 3.0 : 1.0);
            if ((e = vb_set_coefs(op->vb[c], b3, scale, op->a, op->alpha, st))) return e;
        }
        op->coefs_dirty = 0;
    }
    if ((e = tn_cells(op, nullptr, rhs, ng_rhs, 1, op->rhs, nullptr, 3, st))) return e;
    const long long n3 = 3 * op->nc;
    const int nb = (int)std::min<long long>(kTnBlocks, grid_of(n3));
    hipLaunchKernelGGL(k_tn_absmax, dim3(nb), dim3(256), 0, st, (const double*)op->rhs, n3, op->part);
    hipLaunchKernelGGL(k_tn_max_partials, dim3(1), dim3(64), 0, st, (const double*)op->part, nb,
                       op->part + kTnBlocks);
    hipLaunchKernelGGL(k_tn_fill, dim3(grid_of(n3)), dim3(256), 0, st, op->u, n3, 0.0);
    KSGF_LAUNCH_CHECK();
    auto read = [&](double* v) -> int {
        KSGF_HIP_CHECK(hipMemcpyAsync(op->hbuf, op->part + kTnBlocks, sizeof(double), hipMemcpyDeviceToHost, st));
        KSGF_HIP_CHECK(hipStreamSynchronize(st));
        *v = *op->hbuf;
        return KSGF_OK;
    };
    double rhsnorm;
    if ((e = read(&rhsnorm))) return e;
    const double target = std::fmax(atol, rtol * rhsnorm);
    ksgf_mlmg_info inf{};
    inf.r0 = rhsnorm;
    inf.nlevels = vb_nlevels(op->vb[0]);
    double r = rhsnorm;
    inf.converged = (r <= target);
    bool diverged = false;
    const TnDev T = tn_dev(op, op->u);
    auto residual = [&]() {
        hipLaunchKernelGGL(k_tn_apply<1>, dim3(nb), dim3(256), 0, st, T, (const double*)op->a, op->alpha, op->beta,
                           (const double*)op->eta[0], (const double*)op->eta[1], (const double*)op->eta[2],
                           (double*)nullptr, (const double*)op->rhs, vb_level0_f(op->vb[0]),
                           vb_level0_f(op->vb[1]), vb_level0_f(op->vb[2]), op->part);
        hipLaunchKernelGGL(k_tn_max_partials, dim3(1), dim3(64), 0, st, (const double*)op->part, nb,
                           op->part + kTnBlocks);
    };
    if (!inf.converged) residual();
    for (int it = 1; !inf.converged && it <= op->opt.maxiter; ++it) {
        for (int c = 0; c < 3; ++c) vb_vcycle(op->vb[c], st);
        hipLaunchKernelGGL(k_tn_add3, dim3(grid_of(n3)), dim3(256), 0, st, op->u, vb_level0_e(op->vb[0]),
                           vb_level0_e(op->vb[1]), vb_level0_e(op->vb[2]), op->nc);
        residual();
        KSGF_LAUNCH_CHECK();
        if ((e = read(&r))) return e;
        inf.iters = it;
        if (r <= target) {
            inf.converged = 1;
            break;
        }
        if (!(r <= 1e20 * rhsnorm)) {
            diverged = true;
            break;
        }
    }
    inf.rfinal = r;
    if (info) *info = inf;
    if (diverged) {
        set_error("MLMG (tensor operator): the residual exceeds 1e20 |rhs| (diverging)");
        return KSGF_ERR_DIVERGED;
    }
    if ((e = tn_cells(op, sol, nullptr, ng_sol, 0, nullptr, op->u, 3, st))) return e;
    if (!inf.converged) set_error("MLMG (tensor operator): not converged within maxiter");
    return inf.converged ? KSGF_OK : KSGF_ERR_NOT_CONVERGED;
}

int ksgf_tensorop_destroy(ksgf_tensorop* op)
{
    if (op) (void)hipDeviceSynchronize();
    tn_free(op);
    return KSGF_OK;
}

}  // extern "C"
