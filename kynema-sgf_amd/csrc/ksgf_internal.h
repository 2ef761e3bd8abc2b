// Internal declarations shared by the HIP translation units of libksgf.
//
// Device-side conventions
// -----------------------
// * A MultiFab is addressed through its base pointer plus a "region prefix"
//   table pre[0..nbox] (int64, device): pre[b] is the number of points per
//   component in FABs 0..b-1, so FAB b starts at ncomp*pre[b] and its
//   component stride is pre[b+1]-pre[b].
// * Launches over ragged box sets use a chunk table per region: one
//   256-thread workgroup per 256 consecutive points of one box.
// * A region is box b grown by glo[d] below and ghi[d] above in direction d.
//   FAB storage of (itype, ngrow) is the region glo = ngrow,
//   ghi = ngrow + (itype == d+1).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ksgf.h"
#include "ksgf_mg_grid.h"

namespace ksgf {

void set_error(const std::string& msg);

#define KSGF_HIP_CHECK(call)                                                        \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ::ksgf::set_error(std::string(#call) + ": " + hipGetErrorString(e_));   \
            return KSGF_ERR_HIP;                                                    \
        }                                                                           \
    } while (0)

#define KSGF_LAUNCH_CHECK()                                                         \
    do {                                                                            \
        hipError_t e_ = hipGetLastError();                                          \
        if (e_ != hipSuccess) {                                                     \
            ::ksgf::set_error(std::string("kernel launch: ") + hipGetErrorString(e_)); \
            return KSGF_ERR_HIP;                                                    \
        }                                                                           \
    } while (0)

struct RegionKey {
    int glo[3], ghi[3];
    bool operator<(const RegionKey& o) const
    {
        for (int d = 0; d < 3; ++d) {
            if (glo[d] != o.glo[d]) return glo[d] < o.glo[d];
            if (ghi[d] != o.ghi[d]) return ghi[d] < o.ghi[d];
        }
        return false;
    }
};

struct Region {
    long long* pre = nullptr;  // device, nbox+1
    int* chunks = nullptr;     // device, 2 ints per 256-point chunk: box, first point
    std::vector<long long> hpre;
    long long total = 0;
    int nchunks = 0;
    RegionKey key;
};

constexpr int kChunk = 256;  // points per chunk = threads per workgroup of region launches

// ghost-exchange tables for one (itype, ngrow) storage
struct HaloTable {
    int* fb = nullptr;  // device: 4 ints per entry (dst box, dst local, src box, src local)
    long long nfb = 0;
    int* pb[3] = {nullptr, nullptr, nullptr};  // device: 2 ints per entry (box, local)
    long long npb[3] = {0, 0, 0};
};

}  // namespace ksgf

struct ksgf_boxarray {
    ksgf_geometry geom;
    int dlo[3], dhi[3];
    double dx[3];
    int periodic[3];
    int nbox = 0;
    std::vector<int> hbox;  // 6*nbox
    int* dbox = nullptr;    // device copy
    mutable std::mutex mu;
    mutable std::map<ksgf::RegionKey, ksgf::Region*> regions;
    mutable std::map<std::pair<int, int>, ksgf::HaloTable*> halos;
    // per-stage scratch for the advection kernels, grown on demand
    mutable std::vector<double*> scratch;
    mutable std::vector<size_t> scratch_len;
    // box index of every domain cell (-1 where no box), built on first use
    mutable int* owner = nullptr;
    bool covers_domain = true;
};

namespace ksgf {

// region helpers (host)
const Region* get_region(const ksgf_boxarray* ba, const int glo[3], const int ghi[3], int* err);
const Region* fab_region(const ksgf_boxarray* ba, int itype, int ngrow, int* err);
// a launch table that already exists, or null (never allocates)
const Region* find_region(const ksgf_boxarray* ba, const int glo[3], const int ghi[3]);
const HaloTable* get_halo(const ksgf_boxarray* ba, int itype, int ngrow, int* err);
double* get_scratch(const ksgf_boxarray* ba, int slot, size_t n, int* err);
int owner_map(const ksgf_boxarray* ba, hipStream_t st, const int** out);
// V-cycles of a single-level projector's solve, phi written even when the
// tolerance is not met (ksgf_mlmg.hip)
int macproj_vcycles(ksgf_macproj* mp, const double* rhs, double* phi, ksgf_mlmg_info* info, hipStream_t st);

inline int grid_for(long long n, int block) { return (int)((n + block - 1) / block); }

// ----------------------------------------------------------------------------
// device helpers
// ----------------------------------------------------------------------------
struct BoxDev {
    const int* box;  // 6*nbox
    int nbox;
};

// FAB view: base of the component-0 data of one FAB plus its shape
struct Fab {
    double* p;
    int lo0, lo1, lo2;
    int n0, n1;
    int cs;  // component stride (a FAB holds fewer than 2^31 points)
    __device__ __forceinline__ int idx(int i, int j, int k) const
    {
        return (i - lo0) + n0 * ((j - lo1) + n1 * (k - lo2));
    }
    __device__ __forceinline__ double& operator()(int i, int j, int k, int n = 0) const
    {
        return p[idx(i, j, k) + n * cs];
    }
};

struct CFab {
    const double* p;
    int lo0, lo1, lo2;
    int n0, n1;
    int cs;
    __device__ __forceinline__ int idx(int i, int j, int k) const
    {
        return (i - lo0) + n0 * ((j - lo1) + n1 * (k - lo2));
    }
    __device__ __forceinline__ double operator()(int i, int j, int k, int n = 0) const
    {
        return p[idx(i, j, k) + n * cs];
    }
};

// MultiFab handle passed to kernels: base + prefix table + storage shape
struct MFDev {
    double* p;
    const long long* pre;
    int itype, ngrow, ncomp;
};
struct CMFDev {
    const double* p;
    const long long* pre;
    int itype, ngrow, ncomp;
};

__device__ __forceinline__ Fab fab_at(const MFDev& m, const int* box, int b)
{
    Fab f;
    const int* bx = box + 6 * b;
    f.lo0 = bx[0] - m.ngrow;
    f.lo1 = bx[1] - m.ngrow;
    f.lo2 = bx[2] - m.ngrow;
    f.n0 = bx[3] - bx[0] + 1 + 2 * m.ngrow + (m.itype == 1);
    f.n1 = bx[4] - bx[1] + 1 + 2 * m.ngrow + (m.itype == 2);
    long long a = m.pre[b], e = m.pre[b + 1];
    f.cs = (int)(e - a);
    f.p = m.p + (long long)m.ncomp * a;
    return f;
}
__device__ __forceinline__ CFab fab_at(const CMFDev& m, const int* box, int b)
{
    CFab f;
    const int* bx = box + 6 * b;
    f.lo0 = bx[0] - m.ngrow;
    f.lo1 = bx[1] - m.ngrow;
    f.lo2 = bx[2] - m.ngrow;
    f.n0 = bx[3] - bx[0] + 1 + 2 * m.ngrow + (m.itype == 1);
    f.n1 = bx[4] - bx[1] + 1 + 2 * m.ngrow + (m.itype == 2);
    long long a = m.pre[b], e = m.pre[b + 1];
    f.cs = (int)(e - a);
    f.p = m.p + (long long)m.ncomp * a;
    return f;
}

// region iteration.  Launches over a region use one 256-thread workgroup
// per chunk of 256 consecutive points of one box (chunk table: box, first
// point), so a thread finds its box without a search and decodes (i,j,k)
// with 32-bit arithmetic; x is fastest, consecutive lanes read consecutive x.
struct RegDev {
    const long long* pre;
    const int* chunks;
    int nbox;
    int glo0, glo1, glo2, ghi0, ghi1, ghi2;
    long long total;
};

__device__ __forceinline__ bool decode_in_box(const RegDev& r, const int* box, int b, unsigned loc, int& i, int& j,
                                              int& k)
{
    const int* bx = box + 6 * b;
    int l0 = bx[0] - r.glo0, l1 = bx[1] - r.glo1, l2 = bx[2] - r.glo2;
    unsigned n0 = (unsigned)(bx[3] + r.ghi0 - l0 + 1);
    unsigned n1 = (unsigned)(bx[4] + r.ghi1 - l1 + 1);
    unsigned n2 = (unsigned)(bx[5] + r.ghi2 - l2 + 1);
    if (loc >= n0 * n1 * n2) return false;
    unsigned q = loc / n0;
    i = l0 + (int)(loc - q * n0);
    unsigned kk = q / n1;
    j = l1 + (int)(q - kk * n1);
    k = l2 + (int)kk;
    return true;
}

// chunked launch: grid = nchunks, block = kChunk
__device__ __forceinline__ bool region_locate(const RegDev& r, const int* box, int& b, int& i, int& j, int& k)
{
    b = r.chunks[2 * blockIdx.x];
    unsigned loc = (unsigned)r.chunks[2 * blockIdx.x + 1] + threadIdx.x;
    return decode_in_box(r, box, b, loc, i, j, k);
}

// flat index (grid-stride loops): binary search of the box
__device__ __forceinline__ bool region_locate_flat(const RegDev& r, const int* box, long long t, int& b, int& i,
                                                   int& j, int& k)
{
    if (t >= r.total) return false;
    int lo = 0, hi = r.nbox - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (r.pre[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    b = lo;
    return decode_in_box(r, box, b, (unsigned)(t - r.pre[b]), i, j, k);
}

inline RegDev regdev(const Region* r, int nbox)
{
    RegDev d;
    d.pre = r->pre;
    d.chunks = r->chunks;
    d.nbox = nbox;
    d.glo0 = r->key.glo[0];
    d.glo1 = r->key.glo[1];
    d.glo2 = r->key.glo[2];
    d.ghi0 = r->key.ghi[0];
    d.ghi1 = r->key.ghi[1];
    d.ghi2 = r->key.ghi[2];
    d.total = r->total;
    return d;
}

// host helper to build kernel MultiFab handles
int mfdev(const ksgf_boxarray* ba, double* p, int itype, int ncomp, int ngrow, MFDev* out);
int cmfdev(const ksgf_boxarray* ba, const double* p, int itype, int ncomp, int ngrow, CMFDev* out);

// Intermediate face arrays of the three Godunov stages (faces of each
// direction, ngrow 1, in per-BoxArray scratch that the next Godunov call on
// the same BoxArray overwrites): PPM face states lo/hi, upwinded normal
// states ed, advective (Burgers) velocities uad, corner-coupled states
// crn[normal][transverse].
struct GodunovStages {
    MFDev lo[3], hi[3], ed[3], uad[3], crn[3][3];
};

// ksgf_extrap_vel_to_faces / ksgf_compute_fluxes_from_state that also hand
// back their stage arrays (keep may be null)
int extrap_vel_to_faces_keep(const ksgf_boxarray* ba, const double* vel, const double* vel_forces, double* u_mac,
                             double* v_mac, double* w_mac, const ksgf_bcrec* bcrec, double dt, int use_ppm,
                             int use_forces_in_trans, int limiter_type, int allow_inflow_on_outflow,
                             ksgf_stream_t stream, GodunovStages* keep);
int compute_fluxes_keep(const ksgf_boxarray* ba, int ncomp, const double* q, const double* q_on_face_bc,
                        double* flux_x, double* flux_y, double* flux_z, double* face_x, double* face_y,
                        double* face_z, int known_edge_state, const double* u_mac, const double* v_mac,
                        const double* w_mac, const double* divu, const double* fq, double dt,
                        const ksgf_bcrec* bcrec, const int* iconserv, int use_ppm, int use_forces_in_trans,
                        int is_velocity, int fluxes_are_area_weighted, int limiter_type,
                        int allow_inflow_on_outflow, ksgf_stream_t stream, GodunovStages* keep);

// MAC projection with face coefficients (ksgf_mlmg_vb.hip)
struct VbSolver;
// inner_only: no level-0 phi / rhs / residual arrays (the caller drives
// vb_vcycle on the level-0 f and e itself)
VbSolver* vb_create(const ksgf_boxarray* ba, const int* lobc, const int* hibc, const ksgf_mlmg_options& opt, int* err,
                    int inner_only = 0);
void vb_destroy(VbSolver* S);
int vb_set_faces(VbSolver* S, const ksgf_boxarray* ba, const double* const* b3, int ng, hipStream_t st);
int vb_project(VbSolver* S, const ksgf_boxarray* ba, double* const* mac, int ng, double rtol, double atol,
               ksgf_mlmg_info* info, hipStream_t st);
int vb_solve(VbSolver* S, const double* rhs, double* phi, double rtol, double atol, ksgf_mlmg_info* info,
             hipStream_t st);
int vb_get_phi(VbSolver* S, double* phi, hipStream_t st);
// level-0 coefficients from domain-wide arrays: faces scale[d] * b3[d] and the
// zeroth-order term ac_scale * a0 (a0 == nullptr: constant), coarsened to
// every level; the operator becomes div(b grad e) - ac e
int vb_set_coefs(VbSolver* S, const double* const* b3, const double* scale, const double* a0, double ac_scale,
                 hipStream_t st);
// one V-cycle from zero: level-0 e from level-0 f (div(b grad e) - ac e = f)
void vb_vcycle(VbSolver* S, hipStream_t st);
double* vb_level0_f(VbSolver* S);
const double* vb_level0_e(const VbSolver* S);
int vb_nlevels(const VbSolver* S);

}  // namespace ksgf
