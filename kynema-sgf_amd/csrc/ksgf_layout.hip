// BoxArray / MultiFab layout, ghost exchange (FillBoundary) and physical
// boundary fills on one GPU.
//
// FillBoundary (icns_advection.H:239-243) and the BC part of FillPatch
// (FieldFillPatchOps.H:250-336) become table-driven gather kernels: the host
// enumerates, once per (index type, ngrow), every ghost point of every FAB
// together with the FAB and point that own it (interior or periodic image),
// and the kernel copies all components of all ghosts in one launch.
#include "ksgf_internal.h"

#include <algorithm>
#include <cstring>

namespace ksgf {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

static int region_key_check(const int glo[3], const int ghi[3])
{
    for (int d = 0; d < 3; ++d)
        if (glo[d] < 0 || ghi[d] < 0 || glo[d] > 8 || ghi[d] > 9) return 0;
    return 1;
}

const Region* get_region(const ksgf_boxarray* ba, const int glo[3], const int ghi[3], int* err)
{
    RegionKey key;
    for (int d = 0; d < 3; ++d) {
        key.glo[d] = glo[d];
        key.ghi[d] = ghi[d];
    }
    std::lock_guard<std::mutex> lk(ba->mu);
    auto it = ba->regions.find(key);
    if (it != ba->regions.end()) return it->second;
    if (!region_key_check(glo, ghi)) {
        set_error("region growth out of range");
        *err = KSGF_ERR_INVALID;
        return nullptr;
    }
    Region* r = new Region();
    r->key = key;
    r->hpre.resize(ba->nbox + 1);
    long long acc = 0;
    for (int b = 0; b < ba->nbox; ++b) {
        r->hpre[b] = acc;
        long long s = 1;
        for (int d = 0; d < 3; ++d) s *= (long long)(ba->hbox[6 * b + 3 + d] - ba->hbox[6 * b + d] + 1 + glo[d] + ghi[d]);
        acc += s;
    }
    r->hpre[ba->nbox] = acc;
    r->total = acc;
    std::vector<int> ch;
    for (int b = 0; b < ba->nbox; ++b) {
        long long n = r->hpre[b + 1] - r->hpre[b];
        for (long long s0 = 0; s0 < n; s0 += kChunk) {
            ch.push_back(b);
            ch.push_back((int)s0);
        }
    }
    r->nchunks = (int)(ch.size() / 2);
    if (hipMalloc(&r->chunks, sizeof(int) * ch.size()) != hipSuccess ||
        hipMemcpy(r->chunks, ch.data(), sizeof(int) * ch.size(), hipMemcpyHostToDevice) != hipSuccess) {
        delete r;
        set_error("chunk table upload");
        *err = KSGF_ERR_NOMEM;
        return nullptr;
    }
    if (hipMalloc(&r->pre, sizeof(long long) * (ba->nbox + 1)) != hipSuccess) {
        delete r;
        set_error("hipMalloc region table");
        *err = KSGF_ERR_NOMEM;
        return nullptr;
    }
    if (hipMemcpy(r->pre, r->hpre.data(), sizeof(long long) * (ba->nbox + 1), hipMemcpyHostToDevice) !=
        hipSuccess) {
        (void)hipFree(r->pre);
        delete r;
        set_error("hipMemcpy region table");
        *err = KSGF_ERR_HIP;
        return nullptr;
    }
    ba->regions[key] = r;
    return r;
}

const Region* find_region(const ksgf_boxarray* ba, const int glo[3], const int ghi[3])
{
    RegionKey key;
    for (int d = 0; d < 3; ++d) {
        key.glo[d] = glo[d];
        key.ghi[d] = ghi[d];
    }
    std::lock_guard<std::mutex> lk(ba->mu);
    auto it = ba->regions.find(key);
    return it == ba->regions.end() ? nullptr : it->second;
}

const Region* fab_region(const ksgf_boxarray* ba, int itype, int ngrow, int* err)
{
    int glo[3], ghi[3];
    for (int d = 0; d < 3; ++d) {
        glo[d] = ngrow;
        ghi[d] = ngrow + (itype == d + 1 ? 1 : 0);
    }
    return get_region(ba, glo, ghi, err);
}

int mfdev(const ksgf_boxarray* ba, double* p, int itype, int ncomp, int ngrow, MFDev* out)
{
    int err = KSGF_OK;
    const Region* r = fab_region(ba, itype, ngrow, &err);
    if (!r) return err;
    out->p = p;
    out->pre = r->pre;
    out->itype = itype;
    out->ngrow = ngrow;
    out->ncomp = ncomp;
    return KSGF_OK;
}
int cmfdev(const ksgf_boxarray* ba, const double* p, int itype, int ncomp, int ngrow, CMFDev* out)
{
    int err = KSGF_OK;
    const Region* r = fab_region(ba, itype, ngrow, &err);
    if (!r) return err;
    out->p = p;
    out->pre = r->pre;
    out->itype = itype;
    out->ngrow = ngrow;
    out->ncomp = ncomp;
    return KSGF_OK;
}

double* get_scratch(const ksgf_boxarray* ba, int slot, size_t n, int* err)
{
    std::lock_guard<std::mutex> lk(ba->mu);
    if ((int)ba->scratch.size() <= slot) {
        ba->scratch.resize(slot + 1, nullptr);
        ba->scratch_len.resize(slot + 1, 0);
    }
    if (ba->scratch_len[slot] < n) {
        if (ba->scratch[slot]) (void)hipFree(ba->scratch[slot]);
        ba->scratch[slot] = nullptr;
        ba->scratch_len[slot] = 0;
        if (hipMalloc(&ba->scratch[slot], n * sizeof(double)) != hipSuccess) {
            set_error("hipMalloc scratch");
            *err = KSGF_ERR_NOMEM;
            return nullptr;
        }
        ba->scratch_len[slot] = n;
    }
    return ba->scratch[slot];
}

// ---------------------------------------------------------------------------
// halo tables
// ---------------------------------------------------------------------------
const HaloTable* get_halo(const ksgf_boxarray* ba, int itype, int ngrow, int* err)
{
    {
        std::lock_guard<std::mutex> lk(ba->mu);
        auto it = ba->halos.find({itype, ngrow});
        if (it != ba->halos.end()) return it->second;
    }
    const Region* fr = fab_region(ba, itype, ngrow, err);
    if (!fr) return nullptr;
    const int* dlo = ba->dlo;
    const int* dhi = ba->dhi;
    int L[3], n[3];
    for (int d = 0; d < 3; ++d) {
        L[d] = dhi[d] - dlo[d] + 1;
        n[d] = L[d];
    }
    std::vector<int> owner((size_t)n[0] * n[1] * n[2], -1);
    for (int b = 0; b < ba->nbox; ++b) {
        const int* bx = &ba->hbox[6 * b];
        for (int k = bx[2]; k <= bx[5]; ++k)
            for (int j = bx[1]; j <= bx[4]; ++j)
                for (int i = bx[0]; i <= bx[3]; ++i)
                    owner[(size_t)(i - dlo[0]) + (size_t)n[0] * ((j - dlo[1]) + (size_t)n[1] * (k - dlo[2]))] = b;
    }
    std::vector<int> fb;
    std::vector<int> pb[3];
    for (int b = 0; b < ba->nbox; ++b) {
        const int* bx = &ba->hbox[6 * b];
        int flo[3], fn[3], vlo[3], vhi[3];
        for (int d = 0; d < 3; ++d) {
            int nodal = (itype == d + 1);
            flo[d] = bx[d] - ngrow;
            fn[d] = bx[3 + d] - bx[d] + 1 + 2 * ngrow + nodal;
            vlo[d] = bx[d];
            vhi[d] = bx[3 + d] + nodal;
        }
        for (int k = flo[2]; k < flo[2] + fn[2]; ++k)
            for (int j = flo[1]; j < flo[1] + fn[1]; ++j)
                for (int i = flo[0]; i < flo[0] + fn[0]; ++i) {
                    int p[3] = {i, j, k};
                    if (p[0] >= vlo[0] && p[0] <= vhi[0] && p[1] >= vlo[1] && p[1] <= vhi[1] &&
                        p[2] >= vlo[2] && p[2] <= vhi[2])
                        continue;
                    int local = (i - flo[0]) + fn[0] * ((j - flo[1]) + fn[1] * (k - flo[2]));
                    int w[3], ok = 1;
                    for (int d = 0; d < 3; ++d) {
                        int nodal = (itype == d + 1);
                        int x = p[d];
                        if (ba->periodic[d]) {
                            x = dlo[d] + (((x - dlo[d]) % L[d]) + L[d]) % L[d];
                        } else {
                            int top = dhi[d] + (nodal ? 1 : 0);
                            if (x < dlo[d] || x > top) {
                                ok = 0;
                                pb[d].push_back(b);
                                pb[d].push_back(local);
                            }
                        }
                        w[d] = x;
                    }
                    if (!ok) continue;
                    int c[3] = {w[0], w[1], w[2]};
                    if (itype >= 1) {
                        int d = itype - 1;
                        if (c[d] > dhi[d]) c[d] -= 1;
                    }
                    int src = owner[(size_t)(c[0] - dlo[0]) +
                                    (size_t)n[0] * ((c[1] - dlo[1]) + (size_t)n[1] * (c[2] - dlo[2]))];
                    if (src < 0) continue;
                    const int* sb = &ba->hbox[6 * src];
                    int slo[3], sn[3];
                    for (int d = 0; d < 3; ++d) {
                        int nodal = (itype == d + 1);
                        slo[d] = sb[d] - ngrow;
                        sn[d] = sb[3 + d] - sb[d] + 1 + 2 * ngrow + nodal;
                    }
                    int slocal = (w[0] - slo[0]) + sn[0] * ((w[1] - slo[1]) + sn[1] * (w[2] - slo[2]));
                    fb.push_back(b);
                    fb.push_back(local);
                    fb.push_back(src);
                    fb.push_back(slocal);
                }
    }
    HaloTable* h = new HaloTable();
    h->nfb = (long long)fb.size() / 4;
    if (h->nfb > 0) {
        if (hipMalloc(&h->fb, fb.size() * sizeof(int)) != hipSuccess ||
            hipMemcpy(h->fb, fb.data(), fb.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("halo table upload");
            *err = KSGF_ERR_NOMEM;
            delete h;
            return nullptr;
        }
    }
    for (int d = 0; d < 3; ++d) {
        h->npb[d] = (long long)pb[d].size() / 2;
        if (h->npb[d] > 0) {
            if (hipMalloc(&h->pb[d], pb[d].size() * sizeof(int)) != hipSuccess ||
                hipMemcpy(h->pb[d], pb[d].data(), pb[d].size() * sizeof(int), hipMemcpyHostToDevice) !=
                    hipSuccess) {
                set_error("physbc table upload");
                *err = KSGF_ERR_NOMEM;
                delete h;
                return nullptr;
            }
        }
    }
    std::lock_guard<std::mutex> lk(ba->mu);
    auto it = ba->halos.find({itype, ngrow});
    if (it != ba->halos.end()) {
        // another thread built it first
        if (h->fb) (void)hipFree(h->fb);
        for (int d = 0; d < 3; ++d)
            if (h->pb[d]) (void)hipFree(h->pb[d]);
        delete h;
        return it->second;
    }
    ba->halos[{itype, ngrow}] = h;
    return h;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fill_boundary(double* __restrict__ p, const long long* __restrict__ pre,
                                                       int ncomp, const int* __restrict__ tab, long long n)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int db = tab[4 * t + 0], dl = tab[4 * t + 1], sb = tab[4 * t + 2], sl = tab[4 * t + 3];
    long long da = pre[db], dcs = pre[db + 1] - da;
    long long sa = pre[sb], scs = pre[sb + 1] - sa;
    double* dst = p + ncomp * da + dl;
    const double* src = p + ncomp * sa + sl;
    for (int c = 0; c < ncomp; ++c) dst[c * dcs] = src[c * scs];
}

struct PhysBCArgs {
    int bclo[8], bchi[8];   // per component BC type along d
    double vlo[8], vhi[8];  // per component Dirichlet value
};

// one direction d of the physical-boundary fill (see ko_fill_physbc)
__global__ void __launch_bounds__(256) k_fill_physbc(double* __restrict__ p, const long long* __restrict__ pre,
                                                     const int* __restrict__ box, int ncomp, int itype,
                                                     int ngrow, int d, int dlo, int dhi_nodal, int nin,
                                                     const int* __restrict__ tab, long long n, PhysBCArgs a)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int b = tab[2 * t], local = tab[2 * t + 1];
    const int* bx = box + 6 * b;
    int fn0 = bx[3] - bx[0] + 1 + 2 * ngrow + (itype == 1);
    int fn1 = bx[4] - bx[1] + 1 + 2 * ngrow + (itype == 2);
    int li = local % fn0, lr = local / fn0;
    int lj = lr % fn1, lk = lr / fn1;
    int pos[3] = {bx[0] - ngrow + li, bx[1] - ngrow + lj, bx[2] - ngrow + lk};
    long long st[3] = {1, fn0, (long long)fn0 * fn1};
    int x = pos[d];
    int side;
    if (x < dlo)
        side = 0;
    else if (x > dhi_nodal)
        side = 1;
    else
        return;
    long long base = pre[b], cs = pre[b + 1] - base;
    double* f = p + ncomp * base;
    int m = side ? (x - dhi_nodal) : (dlo - x);
    long long sgn = side ? -1 : 1;
    long long at = local;
    long long bnd = at + sgn * (long long)m * st[d];
    for (int c = 0; c < ncomp; ++c) {
        double* v = f + c * cs;
        int tt = side ? a.bchi[c] : a.bclo[c];
        if (tt == KSGF_BC_EXT_DIR) {
            v[at] = side ? a.vhi[c] : a.vlo[c];
        } else if (tt == KSGF_BC_FOEXTRAP || (tt == KSGF_BC_HOEXTRAP && itype >= 1)) {
            v[at] = v[bnd];
        } else if (tt == KSGF_BC_HOEXTRAP) {
            if (m == 1) {
                if (nin >= 3)
                    v[at] = 0.125 * (15.0 * v[bnd] - 10.0 * v[bnd + sgn * st[d]] + 3.0 * v[bnd + 2 * sgn * st[d]]);
                else
                    v[at] = 0.5 * (3.0 * v[bnd] - v[bnd + sgn * st[d]]);
            } else {
                v[at] = v[bnd];
            }
        }
    }
}

// normal MAC velocity on ext_dir boundary faces
__global__ void __launch_bounds__(256) k_set_inflow(MFDev m, RegDev r, const int* __restrict__ box, int d,
                                                    int dlo, int dhi1, int tlo, int thi, double vlo,
                                                    double vhi)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    int p[3] = {i, j, k};
    Fab f = fab_at(m, box, b);
    if (p[d] == dlo && tlo == KSGF_BC_EXT_DIR) f(i, j, k) = vlo;
    if (p[d] == dhi1 && thi == KSGF_BC_EXT_DIR) f(i, j, k) = vhi;
}

}  // namespace ksgf

namespace ksgf {

__global__ void __launch_bounds__(256) k_owner(int* __restrict__ owner, RegDev r, const int* __restrict__ box, int dlo0,
                                               int dlo1, int dlo2, int n0, int n1)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    owner[(i - dlo0) + n0 * ((j - dlo1) + n1 * (k - dlo2))] = b;
}

__global__ void __launch_bounds__(256) k_fill_int(int* __restrict__ a, long long n, int v)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t < n) a[t] = v;
}

// box index of every domain cell (-1 where the BoxArray has no box), kept
// with the BoxArray after the first call
int owner_map(const ksgf_boxarray* ba, hipStream_t st, const int** out)
{
    int err = KSGF_OK;
    const int z[3] = {0, 0, 0};
    const Region* r = get_region(ba, z, z, &err);  // takes the BoxArray's lock itself
    if (!r) return err;
    std::lock_guard<std::mutex> lk(ba->mu);
    if (!ba->owner) {
        const long long n = (long long)(ba->dhi[0] - ba->dlo[0] + 1) * (ba->dhi[1] - ba->dlo[1] + 1) *
                            (ba->dhi[2] - ba->dlo[2] + 1);
        int* o = nullptr;
        if (hipMalloc(&o, sizeof(int) * n) != hipSuccess) {
            set_error("hipMalloc owner map");
            return KSGF_ERR_NOMEM;
        }
        hipLaunchKernelGGL(k_fill_int, dim3(grid_for(n, 256)), dim3(256), 0, st, o, n, -1);
        hipLaunchKernelGGL(k_owner, dim3(r->nchunks), dim3(kChunk), 0, st, o, regdev(r, ba->nbox), ba->dbox, ba->dlo[0],
                           ba->dlo[1], ba->dlo[2], ba->dhi[0] - ba->dlo[0] + 1, ba->dhi[1] - ba->dlo[1] + 1);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            (void)hipFree(o);
            set_error(std::string("owner map: ") + hipGetErrorString(e));
            return KSGF_ERR_HIP;
        }
        ba->owner = o;
    }
    *out = ba->owner;
    return KSGF_OK;
}

}  // namespace ksgf

using namespace ksgf;

extern "C" {

const char* ksgf_last_error(void) { return g_last_error.c_str(); }
const char* ksgf_version(void) { return "ksgf-gfx950 0.1"; }

int ksgf_boxarray_create(const ksgf_geometry* geom, const int* boxes, int nbox, ksgf_boxarray** out)
{
    return ksgf_boxarray_create_level(geom, boxes, nbox, 1, out);
}

int ksgf_boxarray_create_level(const ksgf_geometry* geom, const int* boxes, int nbox, int covers_domain,
                               ksgf_boxarray** out)
{
    if (!geom || !boxes || nbox <= 0 || !out) {
        set_error("ksgf_boxarray_create: null argument or empty box list");
        return KSGF_ERR_INVALID;
    }
    ksgf_boxarray* ba = new ksgf_boxarray();
    ba->geom = *geom;
    long long ncell = 1, covered = 0;
    for (int d = 0; d < 3; ++d) {
        ba->dlo[d] = geom->domain_lo[d];
        ba->dhi[d] = geom->domain_hi[d];
        int nd = geom->domain_hi[d] - geom->domain_lo[d] + 1;
        if (nd <= 0 || !(geom->prob_hi[d] > geom->prob_lo[d])) {
            delete ba;
            set_error("ksgf_boxarray_create: bad geometry");
            return KSGF_ERR_INVALID;
        }
        ba->dx[d] = (geom->prob_hi[d] - geom->prob_lo[d]) / nd;
        ba->periodic[d] = geom->is_periodic[d] ? 1 : 0;
        ncell *= nd;
    }
    ba->nbox = nbox;
    ba->hbox.assign(boxes, boxes + 6 * nbox);
    for (int b = 0; b < nbox; ++b) {
        long long c = 1;
        for (int d = 0; d < 3; ++d) {
            int lo = boxes[6 * b + d], hi = boxes[6 * b + 3 + d];
            if (lo > hi || lo < ba->dlo[d] || hi > ba->dhi[d]) {
                delete ba;
                set_error("ksgf_boxarray_create: box outside the domain or empty");
                return KSGF_ERR_INVALID;
            }
            c *= (hi - lo + 1);
        }
        covered += c;
    }
    // disjointness: boxes sorted by their low z, each tested against the
    // boxes that start below its top (with disjoint boxes, a cell count equal
    // to the domain's then means an exact tiling)
    {
        std::vector<int> order(nbox);
        for (int b = 0; b < nbox; ++b) order[b] = b;
        std::sort(order.begin(), order.end(), [&](int a, int c) { return boxes[6 * a + 2] < boxes[6 * c + 2]; });
        for (int x = 0; x < nbox; ++x) {
            const int* a = boxes + 6 * order[x];
            for (int y = x + 1; y < nbox; ++y) {
                const int* c = boxes + 6 * order[y];
                if (c[2] > a[5]) break;
                bool meet = true;
                for (int d = 0; d < 3; ++d) meet = meet && !(c[3 + d] < a[d] || a[3 + d] < c[d]);
                if (meet) {
                    delete ba;
                    set_error("ksgf_boxarray_create: boxes " + std::to_string(order[x]) + " and " +
                              std::to_string(order[y]) + " overlap");
                    return KSGF_ERR_INVALID;
                }
            }
        }
    }
    ba->covers_domain = covers_domain != 0;
    if (covers_domain ? (covered != ncell) : (covered > ncell)) {
        delete ba;
        set_error("ksgf_boxarray_create: boxes do not tile the domain");
        return KSGF_ERR_INVALID;
    }
    if (hipMalloc(&ba->dbox, sizeof(int) * 6 * nbox) != hipSuccess ||
        hipMemcpy(ba->dbox, boxes, sizeof(int) * 6 * nbox, hipMemcpyHostToDevice) != hipSuccess) {
        delete ba;
        set_error("ksgf_boxarray_create: device box table");
        return KSGF_ERR_HIP;
    }
    // launch tables of the cell MultiFabs a step uses (ngrow 0..3), built
    // here so that entry points which must not allocate find them
    for (int g = 0; g <= 3; ++g) {
        int err = KSGF_OK;
        if (!fab_region(ba, KSGF_CELL, g, &err)) {
            ksgf_boxarray_destroy(ba);
            return err;
        }
    }
    *out = ba;
    return KSGF_OK;
}

int ksgf_boxarray_destroy(ksgf_boxarray* ba)
{
    if (!ba) return KSGF_OK;
    for (auto& kv : ba->regions) {
        if (kv.second->pre) (void)hipFree(kv.second->pre);
        if (kv.second->chunks) (void)hipFree(kv.second->chunks);
        delete kv.second;
    }
    for (auto& kv : ba->halos) {
        if (kv.second->fb) (void)hipFree(kv.second->fb);
        for (int d = 0; d < 3; ++d)
            if (kv.second->pb[d]) (void)hipFree(kv.second->pb[d]);
        delete kv.second;
    }
    for (double* s : ba->scratch)
        if (s) (void)hipFree(s);
    if (ba->dbox) (void)hipFree(ba->dbox);
    if (ba->owner) (void)hipFree(ba->owner);
    delete ba;
    return KSGF_OK;
}

long long ksgf_multifab_size(const ksgf_boxarray* ba, int itype, int ncomp, int ngrow)
{
    if (!ba || itype < 0 || itype > 3 || ncomp <= 0 || ngrow < 0) return -1;
    int err = KSGF_OK;
    const Region* r = fab_region(ba, itype, ngrow, &err);
    if (!r) return -1;
    return r->total * ncomp;
}

int ksgf_fill_boundary(const ksgf_boxarray* ba, double* mf, int itype, int ncomp, int ngrow,
                       ksgf_stream_t stream)
{
    if (!ba || !mf || itype < 0 || itype > 3 || ncomp <= 0 || ngrow < 0) {
        set_error("ksgf_fill_boundary: bad argument");
        return KSGF_ERR_INVALID;
    }
    if (ngrow == 0) return KSGF_OK;
    int err = KSGF_OK;
    const Region* r = fab_region(ba, itype, ngrow, &err);
    if (!r) return err;
    const HaloTable* h = get_halo(ba, itype, ngrow, &err);
    if (!h) return err;
    if (h->nfb == 0) return KSGF_OK;
    hipLaunchKernelGGL(k_fill_boundary, dim3(grid_for(h->nfb, 256)), dim3(256), 0, (hipStream_t)stream, mf,
                       r->pre, ncomp, h->fb, h->nfb);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

int ksgf_fill_physbc(const ksgf_boxarray* ba, double* mf, int itype, int ncomp, int ngrow,
                     const ksgf_bcrec* bc, const double* bcval, ksgf_stream_t stream)
{
    if (!ba || !mf || !bc || !bcval || itype < 0 || itype > 3 || ncomp <= 0 || ncomp > 8 || ngrow < 0) {
        set_error("ksgf_fill_physbc: bad argument");
        return KSGF_ERR_INVALID;
    }
    if (ngrow == 0) return KSGF_OK;
    int err = KSGF_OK;
    const Region* r = fab_region(ba, itype, ngrow, &err);
    if (!r) return err;
    const HaloTable* h = get_halo(ba, itype, ngrow, &err);
    if (!h) return err;
    for (int d = 0; d < 3; ++d) {
        if (ba->periodic[d] || h->npb[d] == 0) continue;
        PhysBCArgs a;
        for (int c = 0; c < ncomp; ++c) {
            a.bclo[c] = bc[c].lo[d];
            a.bchi[c] = bc[c].hi[d];
            a.vlo[c] = bcval[d * ncomp + c];
            a.vhi[c] = bcval[(3 + d) * ncomp + c];
        }
        int nodal = (itype == d + 1);
        hipLaunchKernelGGL(k_fill_physbc, dim3(grid_for(h->npb[d], 256)), dim3(256), 0, (hipStream_t)stream,
                           mf, r->pre, ba->dbox, ncomp, itype, ngrow, d, ba->dlo[d], ba->dhi[d] + nodal,
                           ba->dhi[d] - ba->dlo[d] + 1, h->pb[d], h->npb[d], a);
        KSGF_LAUNCH_CHECK();
    }
    return KSGF_OK;
}

int ksgf_set_inflow_velocity(const ksgf_boxarray* ba, double* umac, double* vmac, double* wmac,
                             int ngrow_mac, const ksgf_bcrec* bc, const double* bcval,
                             ksgf_stream_t stream)
{
    if (!ba || !umac || !vmac || !wmac || !bc || !bcval || ngrow_mac < 0) {
        set_error("ksgf_set_inflow_velocity: bad argument");
        return KSGF_ERR_INVALID;
    }
    double* mac[3] = {umac, vmac, wmac};
    for (int d = 0; d < 3; ++d) {
        if (ba->periodic[d]) continue;
        int tlo = bc[d].lo[d], thi = bc[d].hi[d];
        if (tlo != KSGF_BC_EXT_DIR && thi != KSGF_BC_EXT_DIR) continue;
        MFDev m;
        int e = mfdev(ba, mac[d], KSGF_XFACE + d, 1, ngrow_mac, &m);
        if (e) return e;
        int glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
        ghi[d] = 1;
        int err = KSGF_OK;
        const Region* r = get_region(ba, glo, ghi, &err);
        if (!r) return err;
        hipLaunchKernelGGL(k_set_inflow, dim3(r->nchunks), dim3(256), 0, (hipStream_t)stream, m,
                           regdev(r, ba->nbox), ba->dbox, d, ba->dlo[d], ba->dhi[d] + 1, tlo, thi,
                           bcval[d * 3 + d], bcval[(3 + d) * 3 + d]);
        KSGF_LAUNCH_CHECK();
    }
    return KSGF_OK;
}

}  // extern "C"
