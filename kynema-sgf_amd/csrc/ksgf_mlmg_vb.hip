// MAC projection with face coefficients: Hydro::MacProjector's
// initProjector(LPInfo, Array<const MultiFab*,3> beta) / updateBeta(faces)
// (icns_advection.cpp:143-154, 234-291: beta = factor / rho_face when the
// density varies).  div(beta grad phi) = div u_mac, then
// u_mac -= beta_face grad phi.
//
// Geometric multigrid in FP64 on domain-wide arrays (cell index
// i + n0 (j + n1 k), no ghost cells; domain sides in the stencil):
//   red-black Gauss-Seidel on the variable 7-point operator, one thread per
//   cell of the colour; residual restricted by the cell average; face
//   coefficients coarsened by the average of the fine faces on each coarse
//   face; piecewise-constant prolongation; red-black sweeps on the coarsest
//   level as the bottom solve (one workgroup when the level is small);
//   the FP64 correction form of the solve: phi += e, r = rhs - A phi, the
//   inf norm against max(atol, rtol |rhs|) as the constant-coefficient path.
// Every neighbour term is summed in the order x-, x+, y-, y+, z-, z+, the
// order of the CPU restatement, so the sweeps agree with it bit for bit; the
// rhs mean (singular problems) is a fixed-order sum of block partials.
// The coefficient path runs in FP64 whatever the projector's precision.
#include <cstring>

#include "ksgf_internal.h"

namespace ksgf {

constexpr int kVbMaxLev = 40;
constexpr int kVbBlocks = 1024;       // blocks of the partial sums and maxima
constexpr int kVbBottomCells = 8192;  // coarsest levels up to this size run in one workgroup

struct VbLev {
    int n[3];
    int r[3];  // coarsening ratio to the next level
    double ih2[3];
    double *b[3], *e, *f;
    double* ac = nullptr;  // zeroth-order coefficient (alpha a), when set
    long long nc;
};

struct VbDevLev {  // kernel argument: one level
    int n0, n1, n2;
    double ih2x, ih2y, ih2z;
    const double *bx, *by, *bz;
    const double* ac;  // nullptr: no zeroth-order term
};

struct VbSolver {
    int n[3];
    double dx[3];
    int lobc[3], hibc[3];
    int nlev = 0, nbot = 0, singular = 1;
    ksgf_mlmg_options opt;
    VbLev L[kVbMaxLev];
    double *phi = nullptr, *rhs = nullptr, *r = nullptr, *part = nullptr;
    double* hbuf = nullptr;  // pinned: one double
};

__host__ __device__ inline long long vb_fidx(int d, int n0, int n1, int i, int j, int k)
{
    const int m0 = n0 + (d == 0), m1 = n1 + (d == 1);
    return i + (long long)m0 * (j + (long long)m1 * k);
}

struct VbBc {
    int lo[3], hi[3];
};

// neighbour sum and diagonal of cell (i, j, k)
__device__ __forceinline__ void vb_cell(const VbDevLev& L, const VbBc& bc, const double* __restrict__ x, int i, int j,
                                        int k, double& acc, double& dg)
{
    double a = 0.0, g = 0.0;
    const int c[3] = {i, j, k};
    const int nn[3] = {L.n0, L.n1, L.n2};
    const double* bb[3] = {L.bx, L.by, L.bz};
    const double ih2[3] = {L.ih2x, L.ih2y, L.ih2z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            int f[3] = {i, j, k};
            if (side) f[d] += 1;
            const double b = bb[d][vb_fidx(d, L.n0, L.n1, f[0], f[1], f[2])] * ih2[d];
            const int t = side ? bc.hi[d] : bc.lo[d];
            const bool edge = side ? (c[d] == nn[d] - 1) : (c[d] == 0);
            if (edge && t == KSGF_LO_NEUMANN) continue;
            if (edge && t == KSGF_LO_DIRICHLET) {
                g += 2.0 * b;
                continue;
            }
            int p[3] = {i, j, k};
            p[d] += side ? 1 : -1;
            if (p[d] < 0) p[d] += nn[d];
            if (p[d] >= nn[d]) p[d] -= nn[d];
            a += b * x[p[0] + (long long)L.n0 * (p[1] + (long long)L.n1 * p[2])];
            g += b;
        }
    }
    if (L.ac) g += L.ac[i + (long long)L.n0 * (j + (long long)L.n1 * k)];
    acc = a;
    dg = g;
}

__device__ __forceinline__ void vb_decode(long long t, int n0, int n1, int& i, int& j, int& k)
{
    const long long q = t / n0;
    i = (int)(t - q * n0);
    j = (int)(q % n1);
    k = (int)(q / n1);
}

__global__ void __launch_bounds__(256) k_vb_fill(double* __restrict__ a, long long n, double v)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

__global__ void __launch_bounds__(256) k_vb_sweep(VbDevLev L, VbBc bc, double* __restrict__ e,
                                                  const double* __restrict__ f, int color)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)L.n0 * L.n1 * L.n2) return;
    int i, j, k;
    vb_decode(t, L.n0, L.n1, i, j, k);
    if (((i + j + k) & 1) != color) return;
    double a, g;
    vb_cell(L, bc, e, i, j, k, a, g);
    e[t] = (a - f[t]) / g;
}

// the coarsest level: nbot red-black sweeps from zero, one workgroup
__global__ void __launch_bounds__(1024) k_vb_bottom(VbDevLev L, VbBc bc, double* __restrict__ e,
                                                    const double* __restrict__ f, int nbot)
{
    const int nc = L.n0 * L.n1 * L.n2;
    for (int t = threadIdx.x; t < nc; t += blockDim.x) e[t] = 0.0;
    __syncthreads();
    for (int s = 0; s < nbot; ++s)
        for (int color = 0; color < 2; ++color) {
            for (int t = threadIdx.x; t < nc; t += blockDim.x) {
                int i, j, k;
                vb_decode(t, L.n0, L.n1, i, j, k);
                if (((i + j + k) & 1) != color) continue;
                double a, g;
                vb_cell(L, bc, e, i, j, k, a, g);
                e[t] = (a - f[t]) / g;
            }
            __syncthreads();
        }
}

// coarse right-hand side: the average of the fine residuals f - A e
__global__ void __launch_bounds__(256) k_vb_resid_restrict(VbDevLev F, VbBc bc, const double* __restrict__ e,
                                                           const double* __restrict__ f, int r0, int r1, int r2,
                                                           int c0, int c1, int c2, double* __restrict__ out)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)c0 * c1 * c2) return;
    int I, J, K;
    vb_decode(t, c0, c1, I, J, K);
    double s = 0.0;
    for (int dk = 0; dk < r2; ++dk)
        for (int dj = 0; dj < r1; ++dj)
            for (int di = 0; di < r0; ++di) {
                const int i = I * r0 + di, j = J * r1 + dj, k = K * r2 + dk;
                double a, g;
                vb_cell(F, bc, e, i, j, k, a, g);
                const long long c = i + (long long)F.n0 * (j + (long long)F.n1 * k);
                s += f[c] - (a - g * e[c]);
            }
    out[t] = s * (1.0 / (double)(r0 * r1 * r2));
}

__global__ void __launch_bounds__(256) k_vb_prolong_add(int n0, int n1, int n2, double* __restrict__ e,
                                                        const double* __restrict__ ec, int r0, int r1, int r2,
                                                        int c0, int c1)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n0 * n1 * n2) return;
    int i, j, k;
    vb_decode(t, n0, n1, i, j, k);
    e[t] += ec[i / r0 + (long long)c0 * (j / r1 + (long long)c1 * (k / r2))];
}

// coarse face coefficients of direction D: the average of the fine faces
template <int D>
__global__ void __launch_bounds__(256) k_vb_coarsen(const double* __restrict__ bf, int f0, int f1, double* __restrict__ bc,
                                                    int c0, int c1, int c2, int r0, int r1, int r2)
{
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    const int m0 = c0 + (D == 0), m1 = c1 + (D == 1), m2 = c2 + (D == 2);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)m0 * m1 * m2) return;
    int cI[3];
    vb_decode(t, m0, m1, cI[0], cI[1], cI[2]);
    const int r[3] = {r0, r1, r2};
    const int ra = r[T1], rb = r[T2];
    double s = 0.0;
    for (int a2 = 0; a2 < rb; ++a2)
        for (int a1 = 0; a1 < ra; ++a1) {
            int fI[3];
            fI[D] = cI[D] * r[D];
            fI[T1] = cI[T1] * ra + a1;
            fI[T2] = cI[T2] * rb + a2;
            s += bf[vb_fidx(D, f0, f1, fI[0], fI[1], fI[2])];
        }
    bc[t] = s * (1.0 / (double)(ra * rb));
}

// level 0: phi += e
__global__ void __launch_bounds__(256) k_vb_add(double* __restrict__ phi, const double* __restrict__ e, long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) phi[t] += e[t];
}

// level 0: r = rhs - A phi and the block maxima of |r| (NaN as +inf)
__global__ void __launch_bounds__(256) k_vb_resid(VbDevLev L, VbBc bc, const double* __restrict__ phi,
                                                  const double* __restrict__ rhs, double* __restrict__ r,
                                                  double* __restrict__ part)
{
    __shared__ double sh[4];
    const long long n = (long long)L.n0 * L.n1 * L.n2;
    double m = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        vb_decode(t, L.n0, L.n1, i, j, k);
        double a, g;
        vb_cell(L, bc, phi, i, j, k, a, g);
        const double v = rhs[t] - (a - g * phi[t]);
        r[t] = v;
        const double av = fabs(v);
        m = (av == av) ? fmax(m, av) : INFINITY;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(m, off, 64);
        m = (o == o) ? fmax(m, o) : INFINITY;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double q = sh[0];
        for (int w = 1; w < 4; ++w) q = (sh[w] == sh[w]) ? fmax(q, sh[w]) : INFINITY;
        part[blockIdx.x] = q;
    }
}

// |a|_inf of a plain array into part[] blocks
__global__ void __launch_bounds__(256) k_vb_absmax(const double* __restrict__ a, long long n, double* __restrict__ part)
{
    __shared__ double sh[4];
    double m = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const double av = fabs(a[t]);
        m = (av == av) ? fmax(m, av) : INFINITY;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(m, off, 64);
        m = (o == o) ? fmax(m, o) : INFINITY;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double q = sh[0];
        for (int w = 1; w < 4; ++w) q = (sh[w] == sh[w]) ? fmax(q, sh[w]) : INFINITY;
        part[blockIdx.x] = q;
    }
}

// one workgroup: out = max(part[0..nb)) (mode 0) or sum in block order (mode 1)
__global__ void __launch_bounds__(256) k_vb_finish(const double* __restrict__ part, int nb, int mode,
                                                   double* __restrict__ out)
{
    if (threadIdx.x != 0) return;
    double q = (mode == 0) ? 0.0 : 0.0;
    for (int b = 0; b < nb; ++b) {
        if (mode == 0)
            q = (part[b] == part[b]) ? fmax(q, part[b]) : INFINITY;
        else
            q += part[b];
    }
    *out = q;
}

// block partial sums (fixed order within a block: thread-strided, then a tree)
__global__ void __launch_bounds__(256) k_vb_psum(const double* __restrict__ a, long long n, double* __restrict__ part)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        s += a[t];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) k_vb_sub_mean(double* __restrict__ a, long long n,
                                                     const double* __restrict__ sum)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] -= *sum / (double)n;
}

// valid faces of direction D of a face MultiFab into the domain-wide array
template <int D>
__global__ void __launch_bounds__(256) k_vb_gather_faces(CMFDev m, RegDev r, const int* __restrict__ box, int dlo0,
                                                         int dlo1, int dlo2, int n0, int n1,
                                                         double* __restrict__ out)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    out[vb_fidx(D, n0, n1, i - dlo0, j - dlo1, k - dlo2)] = fab_at(m, box, b)(i, j, k);
}

// rhs = div u_mac on valid cells (k_mac_rhs's expression)
__global__ void __launch_bounds__(256) k_vb_mac_rhs(CMFDev u, CMFDev v, CMFDev w, RegDev r,
                                                    const int* __restrict__ box, int dlo0, int dlo1, int dlo2, int n0,
                                                    int n1, double dxi, double dyi, double dzi,
                                                    double* __restrict__ rhs)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    CFab fu = fab_at(u, box, b), fv = fab_at(v, box, b), fw = fab_at(w, box, b);
    rhs[(i - dlo0) + (long long)n0 * ((j - dlo1) + (long long)n1 * (k - dlo2))] =
        (fu(i + 1, j, k) - fu(i, j, k)) * dxi + (fv(i, j + 1, k) - fv(i, j, k)) * dyi +
        (fw(i, j, k + 1) - fw(i, j, k)) * dzi;
}

// u_mac -= (beta_face (1/dx)) (phi_hi - phi_lo) on the valid faces of
// direction D (periodic wrap; Neumann sides unchanged; Dirichlet: the ghost
// beyond the face is -phi)
template <int D>
__global__ void __launch_bounds__(256) k_vb_correct(MFDev u, RegDev r, const int* __restrict__ box, int dlo0, int dlo1,
                                                    int dlo2, int n0, int n1, int n2, VbBc bc,
                                                    const double* __restrict__ phi, const double* __restrict__ bf,
                                                    double dxi)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    const int c[3] = {i - dlo0, j - dlo1, k - dlo2};
    const int nn[3] = {n0, n1, n2};
    const int x = c[D];
    int lo[3] = {c[0], c[1], c[2]}, hi[3] = {c[0], c[1], c[2]};
    lo[D] = x - 1;
    const bool at_lo = (x == 0), at_hi = (x == nn[D]);
    if ((at_lo && bc.lo[D] == KSGF_LO_NEUMANN) || (at_hi && bc.hi[D] == KSGF_LO_NEUMANN)) return;
    auto P = [&](const int* p) { return phi[p[0] + (long long)n0 * (p[1] + (long long)n1 * p[2])]; };
    double ph, pl;
    if (at_lo) {
        ph = P(hi);
        if (bc.lo[D] == KSGF_LO_DIRICHLET) {
            pl = -ph;
        } else {
            lo[D] = nn[D] - 1;
            pl = P(lo);
        }
    } else if (at_hi) {
        lo[D] = nn[D] - 1;
        pl = P(lo);
        if (bc.hi[D] == KSGF_LO_DIRICHLET) {
            ph = -pl;
        } else {
            hi[D] = 0;
            ph = P(hi);
        }
    } else {
        ph = P(hi);
        pl = P(lo);
    }
    const double bd = bf[vb_fidx(D, n0, n1, c[0], c[1], c[2])] * dxi;
    Fab fu = fab_at(u, box, b);
    fu(i, j, k) = fu(i, j, k) - bd * (ph - pl);
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
static VbDevLev vb_dev(const VbLev& L)
{
    VbDevLev d;
    d.n0 = L.n[0];
    d.n1 = L.n[1];
    d.n2 = L.n[2];
    d.ih2x = L.ih2[0];
    d.ih2y = L.ih2[1];
    d.ih2z = L.ih2[2];
    d.bx = L.b[0];
    d.by = L.b[1];
    d.bz = L.b[2];
    d.ac = L.ac;
    return d;
}

static VbBc vb_bc(const VbSolver* S)
{
    VbBc b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = S->lobc[d];
        b.hi[d] = S->hibc[d];
    }
    return b;
}

static int vb_grid(long long n) { return (int)((n + 255) / 256); }

static int vb_coarsen_faces(VbSolver* S, hipStream_t st);

void vb_destroy(VbSolver* S)
{
    if (!S) return;
    for (int l = 0; l < S->nlev; ++l) {
        for (int d = 0; d < 3; ++d)
            if (S->L[l].b[d]) (void)hipFree(S->L[l].b[d]);
        if (S->L[l].e) (void)hipFree(S->L[l].e);
        if (S->L[l].f) (void)hipFree(S->L[l].f);
        if (S->L[l].ac) (void)hipFree(S->L[l].ac);
    }
    if (S->phi) (void)hipFree(S->phi);
    if (S->rhs) (void)hipFree(S->rhs);
    if (S->r) (void)hipFree(S->r);
    if (S->part) (void)hipFree(S->part);
    if (S->hbuf) (void)hipHostFree(S->hbuf);
    delete S;
}

VbSolver* vb_create(const ksgf_boxarray* ba, const int* lobc, const int* hibc, const ksgf_mlmg_options& opt, int* err,
                    int inner_only)
{
    VbSolver* S = new VbSolver();
    S->opt = opt;
    for (int d = 0; d < 3; ++d) {
        S->n[d] = ba->dhi[d] - ba->dlo[d] + 1;
        S->dx[d] = ba->dx[d];
        S->lobc[d] = lobc[d];
        S->hibc[d] = hibc[d];
        if (lobc[d] == KSGF_LO_DIRICHLET || hibc[d] == KSGF_LO_DIRICHLET) S->singular = 0;
        if (lobc[d] == KSGF_LO_PERIODIC && S->n[d] % 2) {
            set_error("face-coefficient MAC projection: an odd periodic extent has no red-black ordering");
            *err = KSGF_ERR_UNSUPPORTED;
            delete S;
            return nullptr;
        }
    }
    // levels: the coarsening rule of the constant-coefficient hierarchy, a
    // periodic extent kept even
    int cn[3] = {S->n[0], S->n[1], S->n[2]};
    double ch[3] = {S->dx[0], S->dx[1], S->dx[2]};
    int nl = 0;
    for (;;) {
        VbLev& L = S->L[nl];
        for (int d = 0; d < 3; ++d) {
            L.n[d] = cn[d];
            L.r[d] = 1;
            L.ih2[d] = 1.0 / (ch[d] * ch[d]);
        }
        ++nl;
        if (nl >= kVbMaxLev || nl > opt.max_coarsening_level) break;
        int can[3], any = 0;
        for (int d = 0; d < 3; ++d)
            can[d] = (cn[d] % 2 == 0 && cn[d] / 2 >= 2 && (lobc[d] != KSGF_LO_PERIODIC || (cn[d] / 2) % 2 == 0));
        if (!opt.do_semicoarsening) {
            if (!(can[0] && can[1] && can[2])) break;
            for (int d = 0; d < 3; ++d) L.r[d] = 2;
            any = 1;
        } else {
            double hmin = 1e300;
            for (int d = 0; d < 3; ++d)
                if (ch[d] < hmin) hmin = ch[d];
            for (int d = 0; d < 3; ++d)
                if (can[d] && ch[d] <= 1.5 * hmin) {
                    L.r[d] = 2;
                    any = 1;
                }
            if (!any) break;
        }
        for (int d = 0; d < 3; ++d) {
            cn[d] /= L.r[d];
            ch[d] *= L.r[d];
        }
    }
    S->nlev = nl;
    const VbLev& C = S->L[nl - 1];
    int m = std::max(C.n[0], std::max(C.n[1], C.n[2]));
    S->nbot = std::min(2 * m * m + 8, 400);
    auto alloc = [&](double** p, long long n) {
        if (hipMalloc(p, sizeof(double) * (size_t)n) != hipSuccess || hipMemset(*p, 0, sizeof(double) * (size_t)n) != hipSuccess) {
            set_error("face-coefficient MAC projection: hipMalloc");
            return false;
        }
        return true;
    };
    bool ok = true;
    for (int l = 0; l < nl && ok; ++l) {
        VbLev& L = S->L[l];
        L.nc = (long long)L.n[0] * L.n[1] * L.n[2];
        ok = alloc(&L.e, L.nc) && alloc(&L.f, L.nc);
        for (int d = 0; d < 3 && ok; ++d) ok = alloc(&L.b[d], L.nc / L.n[d] * (L.n[d] + 1));
    }
    const long long n0 = S->L[0].nc;
    if (!inner_only) ok = ok && alloc(&S->phi, n0) && alloc(&S->rhs, n0) && alloc(&S->r, n0);
    ok = ok && alloc(&S->part, kVbBlocks + 1);
    if (ok && hipHostMalloc((void**)&S->hbuf, sizeof(double)) != hipSuccess) {
        set_error("face-coefficient MAC projection: pinned buffer");
        ok = false;
    }
    if (!ok) {
        vb_destroy(S);
        *err = KSGF_ERR_NOMEM;
        return nullptr;
    }
    return S;
}

// level-0 face coefficients from face MultiFabs (valid faces), then every
// coarser level's by face averaging
int vb_set_faces(VbSolver* S, const ksgf_boxarray* ba, const double* const* b3, int ng, hipStream_t st)
{
    for (int d = 0; d < 3; ++d) {
        CMFDev m;
        int e = cmfdev(ba, b3[d], KSGF_XFACE + d, 1, ng, &m);
        if (e) return e;
        int glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
        ghi[d] = 1;
        int err = KSGF_OK;
        const Region* r = get_region(ba, glo, ghi, &err);
        if (!r) return err;
        dim3 gr(r->nchunks);
        if (d == 0)
            hipLaunchKernelGGL(k_vb_gather_faces<0>, gr, dim3(256), 0, st, m, regdev(r, ba->nbox), ba->dbox,
                               ba->dlo[0], ba->dlo[1], ba->dlo[2], S->n[0], S->n[1], S->L[0].b[0]);
        else if (d == 1)
            hipLaunchKernelGGL(k_vb_gather_faces<1>, gr, dim3(256), 0, st, m, regdev(r, ba->nbox), ba->dbox,
                               ba->dlo[0], ba->dlo[1], ba->dlo[2], S->n[0], S->n[1], S->L[0].b[1]);
        else
            hipLaunchKernelGGL(k_vb_gather_faces<2>, gr, dim3(256), 0, st, m, regdev(r, ba->nbox), ba->dbox,
                               ba->dlo[0], ba->dlo[1], ba->dlo[2], S->n[0], S->n[1], S->L[0].b[2]);
        KSGF_LAUNCH_CHECK();
    }
    return vb_coarsen_faces(S, st);
}

static int vb_coarsen_faces(VbSolver* S, hipStream_t st)
{
    for (int l = 0; l + 1 < S->nlev; ++l) {
        const VbLev& F = S->L[l];
        const VbLev& C = S->L[l + 1];
        for (int d = 0; d < 3; ++d) {
            const long long m = C.nc / C.n[d] * (C.n[d] + 1);
            const dim3 g(vb_grid(m));
            if (d == 0)
                hipLaunchKernelGGL(k_vb_coarsen<0>, g, dim3(256), 0, st, (const double*)F.b[0], F.n[0], F.n[1], C.b[0],
                                   C.n[0], C.n[1], C.n[2], F.r[0], F.r[1], F.r[2]);
            else if (d == 1)
                hipLaunchKernelGGL(k_vb_coarsen<1>, g, dim3(256), 0, st, (const double*)F.b[1], F.n[0], F.n[1], C.b[1],
                                   C.n[0], C.n[1], C.n[2], F.r[0], F.r[1], F.r[2]);
            else
                hipLaunchKernelGGL(k_vb_coarsen<2>, g, dim3(256), 0, st, (const double*)F.b[2], F.n[0], F.n[1], C.b[2],
                                   C.n[0], C.n[1], C.n[2], F.r[0], F.r[1], F.r[2]);
            KSGF_LAUNCH_CHECK();
        }
    }
    return KSGF_OK;
}

static void vb_sweeps(VbSolver* S, int l, int nsweep, hipStream_t st)
{
    const VbLev& L = S->L[l];
    const VbDevLev d = vb_dev(L);
    const VbBc bc = vb_bc(S);
    for (int s = 0; s < nsweep; ++s)
        for (int color = 0; color < 2; ++color)
            hipLaunchKernelGGL(k_vb_sweep, dim3(vb_grid(L.nc)), dim3(256), 0, st, d, bc, L.e, (const double*)L.f,
                               color);
}

void vb_vcycle(VbSolver* S, hipStream_t st)
{
    const VbBc bc = vb_bc(S);
    const int last = S->nlev - 1;
    for (int l = 0; l < last; ++l) {
        VbLev& L = S->L[l];
        VbLev& C = S->L[l + 1];
        hipLaunchKernelGGL(k_vb_fill, dim3(vb_grid(L.nc)), dim3(256), 0, st, L.e, L.nc, 0.0);
        vb_sweeps(S, l, S->opt.num_pre_smooth, st);
        hipLaunchKernelGGL(k_vb_resid_restrict, dim3(vb_grid(C.nc)), dim3(256), 0, st, vb_dev(L), bc,
                           (const double*)L.e, (const double*)L.f, L.r[0], L.r[1], L.r[2], C.n[0], C.n[1], C.n[2],
                           C.f);
    }
    VbLev& B = S->L[last];
    if (B.nc <= kVbBottomCells) {
        hipLaunchKernelGGL(k_vb_bottom, dim3(1), dim3(1024), 0, st, vb_dev(B), bc, B.e, (const double*)B.f, S->nbot);
    } else {
        hipLaunchKernelGGL(k_vb_fill, dim3(vb_grid(B.nc)), dim3(256), 0, st, B.e, B.nc, 0.0);
        vb_sweeps(S, last, S->nbot, st);
    }
    for (int l = last - 1; l >= 0; --l) {
        VbLev& L = S->L[l];
        VbLev& C = S->L[l + 1];
        hipLaunchKernelGGL(k_vb_prolong_add, dim3(vb_grid(L.nc)), dim3(256), 0, st, L.n[0], L.n[1], L.n[2], L.e,
                           (const double*)C.e, L.r[0], L.r[1], L.r[2], C.n[0], C.n[1]);
        vb_sweeps(S, l, S->opt.num_post_smooth, st);
    }
}

// dst = s * src (src == nullptr: dst = s)
__global__ void __launch_bounds__(256) k_vb_scale(double* __restrict__ dst, const double* __restrict__ src, long long n,
                                                  double s)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src ? s * src[t] : s;
}

// coarse cell coefficient: the average of the fine cells it covers
__global__ void __launch_bounds__(256) k_vb_coarsen_cells(const double* __restrict__ af, int f0, int f1,
                                                          double* __restrict__ ac, int c0, int c1, int c2, int r0,
                                                          int r1, int r2)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)c0 * c1 * c2) return;
    int I, J, K;
    vb_decode(t, c0, c1, I, J, K);
    double s = 0.0;
    for (int dk = 0; dk < r2; ++dk)
        for (int dj = 0; dj < r1; ++dj)
            for (int di = 0; di < r0; ++di)
                s += af[(I * r0 + di) + (long long)f0 * ((J * r1 + dj) + (long long)f1 * (K * r2 + dk))];
    ac[t] = s * (1.0 / (double)(r0 * r1 * r2));
}

// level-0 coefficients from domain-wide arrays: face d = scale[d] * b3[d],
// zeroth-order term ac_scale * a0 (a0 == nullptr: the constant ac_scale);
// then every coarser level by averaging
int vb_set_coefs(VbSolver* S, const double* const* b3, const double* scale, const double* a0, double ac_scale,
                 hipStream_t st)
{
    for (int l = 0; l < S->nlev; ++l)
        if (!S->L[l].ac && hipMalloc(&S->L[l].ac, sizeof(double) * (size_t)S->L[l].nc) != hipSuccess) {
            set_error("face-coefficient multigrid: hipMalloc");
            return KSGF_ERR_NOMEM;
        }
    const VbLev& L0 = S->L[0];
    for (int d = 0; d < 3; ++d) {
        const long long m = L0.nc / L0.n[d] * (L0.n[d] + 1);
        hipLaunchKernelGGL(k_vb_scale, dim3(vb_grid(m)), dim3(256), 0, st, L0.b[d], b3[d], m, scale[d]);
    }
    hipLaunchKernelGGL(k_vb_scale, dim3(vb_grid(L0.nc)), dim3(256), 0, st, L0.ac, a0, L0.nc, ac_scale);
    KSGF_LAUNCH_CHECK();
    int e = vb_coarsen_faces(S, st);
    if (e) return e;
    for (int l = 0; l + 1 < S->nlev; ++l) {
        const VbLev& F = S->L[l];
        const VbLev& C = S->L[l + 1];
        hipLaunchKernelGGL(k_vb_coarsen_cells, dim3(vb_grid(C.nc)), dim3(256), 0, st, (const double*)F.ac, F.n[0],
                           F.n[1], C.ac, C.n[0], C.n[1], C.n[2], F.r[0], F.r[1], F.r[2]);
    }
    KSGF_LAUNCH_CHECK();
    S->singular = 0;
    return KSGF_OK;
}

double* vb_level0_f(VbSolver* S) { return S->L[0].f; }
const double* vb_level0_e(const VbSolver* S) { return S->L[0].e; }
int vb_nlevels(const VbSolver* S) { return S->nlev; }

static int vb_read(VbSolver* S, const double* dsrc, hipStream_t st, double* out)
{
    KSGF_HIP_CHECK(hipMemcpyAsync(S->hbuf, dsrc, sizeof(double), hipMemcpyDeviceToHost, st));
    KSGF_HIP_CHECK(hipStreamSynchronize(st));
    *out = *S->hbuf;
    return KSGF_OK;
}

// solve with the level-0 rhs in S->rhs; phi in S->phi
static int vb_solve_rhs(VbSolver* S, double rtol, double atol, ksgf_mlmg_info* info, hipStream_t st)
{
    VbLev& L0 = S->L[0];
    const long long n = L0.nc;
    const int nb = (int)std::min<long long>(kVbBlocks, vb_grid(n));
    if (S->singular) {
        hipLaunchKernelGGL(k_vb_psum, dim3(nb), dim3(256), 0, st, (const double*)S->rhs, n, S->part);
        hipLaunchKernelGGL(k_vb_finish, dim3(1), dim3(64), 0, st, (const double*)S->part, nb, 1, S->part + kVbBlocks);
        hipLaunchKernelGGL(k_vb_sub_mean, dim3(vb_grid(n)), dim3(256), 0, st, S->rhs, n,
                           (const double*)(S->part + kVbBlocks));
    }
    hipLaunchKernelGGL(k_vb_fill, dim3(vb_grid(n)), dim3(256), 0, st, S->phi, n, 0.0);
    hipLaunchKernelGGL(k_vb_absmax, dim3(nb), dim3(256), 0, st, (const double*)S->rhs, n, S->part);
    hipLaunchKernelGGL(k_vb_finish, dim3(1), dim3(64), 0, st, (const double*)S->part, nb, 0, S->part + kVbBlocks);
    KSGF_LAUNCH_CHECK();
    double rhsnorm;
    int e = vb_read(S, S->part + kVbBlocks, st, &rhsnorm);
    if (e) return e;
    const double target = std::fmax(atol, rtol * rhsnorm);
    ksgf_mlmg_info inf{};
    inf.r0 = rhsnorm;
    inf.nlevels = S->nlev;
    double r = rhsnorm;
    inf.converged = (r <= target);
    KSGF_HIP_CHECK(hipMemcpyAsync(S->r, S->rhs, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st));
    bool diverged = false;
    const VbBc bc = vb_bc(S);
    for (int it = 1; !inf.converged && it <= S->opt.maxiter; ++it) {
        KSGF_HIP_CHECK(hipMemcpyAsync(L0.f, S->r, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st));
        vb_vcycle(S, st);
        hipLaunchKernelGGL(k_vb_add, dim3(vb_grid(n)), dim3(256), 0, st, S->phi, (const double*)L0.e, n);
        hipLaunchKernelGGL(k_vb_resid, dim3(nb), dim3(256), 0, st, vb_dev(L0), bc, (const double*)S->phi,
                           (const double*)S->rhs, S->r, S->part);
        hipLaunchKernelGGL(k_vb_finish, dim3(1), dim3(64), 0, st, (const double*)S->part, nb, 0, S->part + kVbBlocks);
        KSGF_LAUNCH_CHECK();
        e = vb_read(S, S->part + kVbBlocks, st, &r);
        if (e) return e;
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
        set_error("MLMG (face coefficients): the residual exceeds 1e20 |rhs| (diverging)");
        return KSGF_ERR_DIVERGED;
    }
    return inf.converged ? KSGF_OK : KSGF_ERR_NOT_CONVERGED;
}

int vb_project(VbSolver* S, const ksgf_boxarray* ba, double* const* mac, int ng, double rtol, double atol,
               ksgf_mlmg_info* info, hipStream_t st)
{
    CMFDev cu[3];
    MFDev mu[3];
    for (int d = 0; d < 3; ++d) {
        int e = cmfdev(ba, mac[d], KSGF_XFACE + d, 1, ng, &cu[d]);
        if (!e) e = mfdev(ba, mac[d], KSGF_XFACE + d, 1, ng, &mu[d]);
        if (e) return e;
    }
    int err = KSGF_OK;
    int z[3] = {0, 0, 0};
    const Region* rv = get_region(ba, z, z, &err);
    if (!rv) return err;
    hipLaunchKernelGGL(k_vb_mac_rhs, dim3(rv->nchunks), dim3(256), 0, st, cu[0], cu[1], cu[2], regdev(rv, ba->nbox),
                       ba->dbox, ba->dlo[0], ba->dlo[1], ba->dlo[2], S->n[0], S->n[1], 1.0 / ba->dx[0],
                       1.0 / ba->dx[1], 1.0 / ba->dx[2], S->rhs);
    KSGF_LAUNCH_CHECK();
    const int ret = vb_solve_rhs(S, rtol, atol, info, st);
    if (ret < 0) return ret;
    const VbBc bc = vb_bc(S);
    for (int d = 0; d < 3; ++d) {
        int glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
        ghi[d] = 1;
        const Region* rf = get_region(ba, glo, ghi, &err);
        if (!rf) return err;
        dim3 gr(rf->nchunks);
        const double dxi = 1.0 / ba->dx[d];
        if (d == 0)
            hipLaunchKernelGGL(k_vb_correct<0>, gr, dim3(256), 0, st, mu[0], regdev(rf, ba->nbox), ba->dbox,
                               ba->dlo[0], ba->dlo[1], ba->dlo[2], S->n[0], S->n[1], S->n[2], bc,
                               (const double*)S->phi, (const double*)S->L[0].b[0], dxi);
        else if (d == 1)
            hipLaunchKernelGGL(k_vb_correct<1>, gr, dim3(256), 0, st, mu[1], regdev(rf, ba->nbox), ba->dbox,
                               ba->dlo[0], ba->dlo[1], ba->dlo[2], S->n[0], S->n[1], S->n[2], bc,
                               (const double*)S->phi, (const double*)S->L[0].b[1], dxi);
        else
            hipLaunchKernelGGL(k_vb_correct<2>, gr, dim3(256), 0, st, mu[2], regdev(rf, ba->nbox), ba->dbox,
                               ba->dlo[0], ba->dlo[1], ba->dlo[2], S->n[0], S->n[1], S->n[2], bc,
                               (const double*)S->phi, (const double*)S->L[0].b[2], dxi);
        KSGF_LAUNCH_CHECK();
    }
    return ret;
}

// domain-wide rhs in, domain-wide phi out
int vb_solve(VbSolver* S, const double* rhs, double* phi, double rtol, double atol, ksgf_mlmg_info* info,
             hipStream_t st)
{
    const size_t n = sizeof(double) * (size_t)S->L[0].nc;
    KSGF_HIP_CHECK(hipMemcpyAsync(S->rhs, rhs, n, hipMemcpyDeviceToDevice, st));
    const int ret = vb_solve_rhs(S, rtol, atol, info, st);
    if (ret < 0) return ret;
    KSGF_HIP_CHECK(hipMemcpyAsync(phi, S->phi, n, hipMemcpyDeviceToDevice, st));
    return ret;
}

int vb_get_phi(VbSolver* S, double* phi, hipStream_t st)
{
    KSGF_HIP_CHECK(hipMemcpyAsync(phi, S->phi, sizeof(double) * (size_t)S->L[0].nc, hipMemcpyDeviceToDevice, st));
    return KSGF_OK;
}

}  // namespace ksgf
