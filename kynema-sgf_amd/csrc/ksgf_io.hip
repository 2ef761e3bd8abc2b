// AMReX plotfile and VisMF files of MultiFabs that live in HBM (host code):
//   amrex::WriteMultiLevelPlotfile, one level   (IOManager::write_plot_file,
//                                                IOManager.cpp:136-190)
//   amrex::VisMF::Write / VisMF::Read           (checkpoint fields,
//                                                IOManager.cpp:214-219, 249-266)
//   checkpoint Header                           (IOManager::write_header,
//                                                IOManager.cpp:308-356)
// The text layouts are the ones the reference's own readers parse
// (tools/amrex_plotfile.py: parse_header / write_header;
// tools/utilities/CheckpointToCSV/CheckpointFileUtil.cpp:40-120).
//
// A FAB of this library's MultiFab is one contiguous block, Fortran order
// with the component slowest -- the byte layout of a FAB's data on disk -- so
// a write copies the MultiFab to the host once and streams each FAB behind
// its header line; a read parses the headers, converts the payload (IEEE
// double or float, either byte order) and copies it into the FABs it
// overlaps, then uploads once.
#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "ksgf_internal.h"

namespace ksgf {
namespace {

struct IBox {
    int lo[3], hi[3], t[3];  // inclusive, index type per direction (1 = nodal)
    long long npts() const
    {
        long long n = 1;
        for (int d = 0; d < 3; ++d) n *= (long long)(hi[d] - lo[d] + 1);
        return n;
    }
};

std::string box_text(const IBox& b)
{
    std::ostringstream o;
    o << "((" << b.lo[0] << ',' << b.lo[1] << ',' << b.lo[2] << ") (" << b.hi[0] << ',' << b.hi[1] << ','
      << b.hi[2] << ") (" << b.t[0] << ',' << b.t[1] << ',' << b.t[2] << "))";
    return o.str();
}

// valid box b of the BoxArray in index type itype (faces: one more point
// along their direction), optionally grown
IBox fab_box(const ksgf_boxarray* ba, int b, int itype, int ngrow)
{
    IBox x;
    for (int d = 0; d < 3; ++d) {
        x.t[d] = (itype == d + 1) ? 1 : 0;
        x.lo[d] = ba->hbox[6 * b + d] - ngrow;
        x.hi[d] = ba->hbox[6 * b + 3 + d] + x.t[d] + ngrow;
    }
    return x;
}

// the cell-centred box b, in index type itype as stored in a BoxArray text
// (AMReX keeps the cell box and the index type)
IBox ba_box(const ksgf_boxarray* ba, int b, int itype)
{
    IBox x;
    for (int d = 0; d < 3; ++d) {
        x.t[d] = (itype == d + 1) ? 1 : 0;
        x.lo[d] = ba->hbox[6 * b + d];
        x.hi[d] = ba->hbox[6 * b + 3 + d] + x.t[d];
    }
    return x;
}

int fail(const std::string& what, int code = KSGF_ERR_INVALID)
{
    set_error(what);
    return code;
}

int make_dir(const std::string& p)
{
    if (mkdir(p.c_str(), 0755) == 0 || errno == EEXIST) return KSGF_OK;
    return fail("cannot create directory " + p);
}

// the FAB's real-number descriptor for IEEE double in this host's byte order
std::string native_descriptor()
{
    const uint16_t one = 1;
    const bool little = *(const uint8_t*)&one == 1;
    return little ? "(8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1))"
                  : "(8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8))";
}

// host copy of a whole MultiFab and its per-box offsets (in points)
struct HostMF {
    std::vector<double> v;
    std::vector<long long> pre;
};

int download(const ksgf_boxarray* ba, const double* mf, int itype, int ncomp, int ngrow, hipStream_t st, HostMF& h)
{
    int err = KSGF_OK;
    const Region* r = fab_region(ba, itype, ngrow, &err);
    if (!r) return err;
    h.pre = r->hpre;
    h.v.resize((size_t)r->total * ncomp);
    if (hipMemcpyAsync(h.v.data(), mf, h.v.size() * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail("MultiFab download", KSGF_ERR_HIP);
    return KSGF_OK;
}

int prepare(const ksgf_boxarray* ba, int itype, int ncomp, int ngrow, HostMF& h)
{
    int err = KSGF_OK;
    const Region* r = fab_region(ba, itype, ngrow, &err);
    if (!r) return err;
    h.pre = r->hpre;
    h.v.assign((size_t)r->total * ncomp, 0.0);
    return KSGF_OK;
}

// per-component minima and maxima of a box of a stored FAB
void fab_minmax(const double* p, const IBox& stored, const IBox& part, int ncomp, std::vector<double>& mn,
                std::vector<double>& mx)
{
    const long long n0 = stored.hi[0] - stored.lo[0] + 1, n1 = stored.hi[1] - stored.lo[1] + 1;
    const long long cs = stored.npts();
    mn.assign(ncomp, 0.0);
    mx.assign(ncomp, 0.0);
    for (int n = 0; n < ncomp; ++n) {
        bool first = true;
        double a = 0.0, b = 0.0;
        for (int k = part.lo[2]; k <= part.hi[2]; ++k)
            for (int j = part.lo[1]; j <= part.hi[1]; ++j)
                for (int i = part.lo[0]; i <= part.hi[0]; ++i) {
                    const double x =
                        p[n * cs + (i - stored.lo[0]) + n0 * ((j - stored.lo[1]) + n1 * (k - stored.lo[2]))];
                    if (first) {
                        a = b = x;
                        first = false;
                    } else {
                        a = std::min(a, x);
                        b = std::max(b, x);
                    }
                }
        mn[n] = a;
        mx[n] = b;
    }
}

// copy the points of `part` from FAB (src, sbox) to FAB (dst, dbox), all components
void copy_part(const double* src, const IBox& sbox, double* dst, const IBox& dbox, const IBox& part, int ncomp)
{
    const long long s0 = sbox.hi[0] - sbox.lo[0] + 1, s1 = sbox.hi[1] - sbox.lo[1] + 1, scs = sbox.npts();
    const long long d0 = dbox.hi[0] - dbox.lo[0] + 1, d1 = dbox.hi[1] - dbox.lo[1] + 1, dcs = dbox.npts();
    const int len = part.hi[0] - part.lo[0] + 1;
    for (int n = 0; n < ncomp; ++n)
        for (int k = part.lo[2]; k <= part.hi[2]; ++k)
            for (int j = part.lo[1]; j <= part.hi[1]; ++j) {
                const double* s =
                    src + n * scs + (part.lo[0] - sbox.lo[0]) + s0 * ((j - sbox.lo[1]) + s1 * (k - sbox.lo[2]));
                double* d = dst + n * dcs + (part.lo[0] - dbox.lo[0]) + d0 * ((j - dbox.lo[1]) + d1 * (k - dbox.lo[2]));
                std::memcpy(d, s, sizeof(double) * len);
            }
}

bool intersect(const IBox& a, const IBox& b, IBox& out)
{
    for (int d = 0; d < 3; ++d) {
        out.lo[d] = std::max(a.lo[d], b.lo[d]);
        out.hi[d] = std::min(a.hi[d], b.hi[d]);
        out.t[d] = a.t[d];
        if (out.lo[d] > out.hi[d]) return false;
    }
    return true;
}

// VisMF: <prefix>_H and one data file <prefix>_D_00000.  write_box(b) gives
// the part of FAB b that goes to disk (the stored FAB, or its valid box for
// a plotfile) and the box the header's BoxArray lists.
int vismf_write(const std::string& prefix, const ksgf_boxarray* ba, const HostMF& h, int itype, int ncomp,
                int ngrow_stored, int ngrow_out)
{
    const std::string dname = prefix + "_D_00000";
    const std::string dbase = dname.substr(dname.find_last_of('/') + 1);
    std::ofstream df(dname, std::ios::binary | std::ios::trunc);
    if (!df) return fail("cannot write " + dname);
    const std::string desc = native_descriptor();
    std::vector<long long> offs(ba->nbox);
    std::vector<std::vector<double>> mins(ba->nbox), maxs(ba->nbox);
    std::vector<double> buf;
    for (int b = 0; b < ba->nbox; ++b) {
        const IBox stored = fab_box(ba, b, itype, ngrow_stored);
        const IBox out = fab_box(ba, b, itype, ngrow_out);
        const double* p = h.v.data() + (size_t)ncomp * h.pre[b];
        offs[b] = (long long)df.tellp();
        df << "FAB (" << desc << ")" << box_text(out) << ' ' << ncomp << '\n';
        if (ngrow_out == ngrow_stored) {
            df.write((const char*)p, (std::streamsize)(sizeof(double) * ncomp * stored.npts()));
        } else {
            buf.assign((size_t)ncomp * out.npts(), 0.0);
            copy_part(p, stored, buf.data(), out, out, ncomp);
            df.write((const char*)buf.data(), (std::streamsize)(sizeof(double) * buf.size()));
        }
        fab_minmax(p, stored, out, ncomp, mins[b], maxs[b]);
    }
    df.close();
    if (!df) return fail("write failed: " + dname);
    std::ofstream hf(prefix + "_H", std::ios::trunc);
    if (!hf) return fail("cannot write " + prefix + "_H");
    hf << std::setprecision(17);
    hf << 1 << '\n' << 1 << '\n' << ncomp << '\n' << ngrow_out << '\n';
    hf << '(' << ba->nbox << " 0\n";
    for (int b = 0; b < ba->nbox; ++b) hf << box_text(ba_box(ba, b, itype)) << '\n';
    hf << ")\n";
    hf << ba->nbox << '\n';
    for (int b = 0; b < ba->nbox; ++b) hf << "FabOnDisk: " << dbase << ' ' << offs[b] << '\n';
    hf << '\n';
    for (int pass = 0; pass < 2; ++pass) {
        hf << ba->nbox << ',' << ncomp << '\n';
        for (int b = 0; b < ba->nbox; ++b) {
            for (int n = 0; n < ncomp; ++n) hf << (pass ? maxs[b][n] : mins[b][n]) << ',';
            hf << '\n';
        }
        hf << '\n';
    }
    hf.close();
    if (!hf) return fail("write failed: " + prefix + "_H");
    return KSGF_OK;
}

// ---------------------------------------------------------------------------
// reading
// ---------------------------------------------------------------------------
bool parse_box(std::istream& is, IBox& b)
{
    char c;
    auto expect = [&](char want) {
        if (!(is >> c) || c != want) return false;
        return true;
    };
    auto triple = [&](int* v) {
        if (!expect('(')) return false;
        for (int d = 0; d < 3; ++d) {
            if (!(is >> v[d])) return false;
            if (d < 2 && !expect(',')) return false;
        }
        return expect(')');
    };
    return expect('(') && triple(b.lo) && triple(b.hi) && triple(b.t) && expect(')');
}

struct FileFab {
    std::string file;
    long long offset;
};

struct VisMFHeader {
    int ncomp = 0, ngrow = 0;
    std::vector<IBox> boxes;
    std::vector<FileFab> fabs;
};

int read_vismf_header(const std::string& path, VisMFHeader& H)
{
    std::ifstream is(path);
    if (!is) return fail("cannot open " + path);
    int vers, how;
    if (!(is >> vers >> how >> H.ncomp)) return fail("bad VisMF header " + path);
    if (vers != 1) return fail("VisMF header version " + std::to_string(vers) + " not supported: " + path,
                               KSGF_ERR_UNSUPPORTED);
    // nGrow: one integer, or an IntVect (g,g,g) with equal entries
    is >> std::ws;
    if (is.peek() == '(') {
        char c;
        int g[3];
        is >> c >> g[0] >> c >> g[1] >> c >> g[2] >> c;
        if (g[0] != g[1] || g[0] != g[2]) return fail("unequal nGrow in " + path, KSGF_ERR_UNSUPPORTED);
        H.ngrow = g[0];
    } else if (!(is >> H.ngrow)) {
        return fail("bad nGrow in " + path);
    }
    char c;
    int nb, hash;
    if (!(is >> c) || c != '(' || !(is >> nb >> hash) || nb < 0) return fail("bad BoxArray in " + path);
    H.boxes.resize(nb);
    for (int b = 0; b < nb; ++b)
        if (!parse_box(is, H.boxes[b])) return fail("bad box in " + path);
    if (!(is >> c) || c != ')') return fail("bad BoxArray end in " + path);
    int nf;
    if (!(is >> nf) || nf != nb) return fail("FabOnDisk count differs from the box count in " + path);
    H.fabs.resize(nf);
    for (int f = 0; f < nf; ++f) {
        std::string tag;
        if (!(is >> tag >> H.fabs[f].file >> H.fabs[f].offset) || tag != "FabOnDisk:")
            return fail("bad FabOnDisk line in " + path);
    }
    return KSGF_OK;
}

// read one FAB at `offset` of `file`: its box (stored, i.e. grown), component
// count and data converted to double
int read_fab(const std::string& file, long long offset, IBox& box, int& ncomp, std::vector<double>& out)
{
    std::ifstream is(file, std::ios::binary);
    if (!is) return fail("cannot open " + file);
    is.seekg(offset);
    std::string line;
    if (!std::getline(is, line) || line.compare(0, 4, "FAB ") != 0) return fail("bad FAB header in " + file);
    // FAB ((nb, (fmt..)),(nb, (order..)))((lo) (hi) (t)) ncomp
    const size_t close = line.find(")))");
    if (close == std::string::npos) return fail("bad FAB descriptor in " + file);
    const std::string desc = line.substr(4, close + 3 - 4);
    std::istringstream bs(line.substr(close + 3));
    if (!parse_box(bs, box) || !(bs >> ncomp) || ncomp <= 0) return fail("bad FAB box in " + file);
    int nbytes = 0;
    bool swap = false;
    auto has = [&](const char* s) { return desc.find(s) != std::string::npos; };
    if (has("(8, (64 11 52 0 1 12 0 1023))")) {
        nbytes = 8;
        if (has("(8, (8 7 6 5 4 3 2 1))"))
            swap = false;
        else if (has("(8, (1 2 3 4 5 6 7 8))"))
            swap = true;
        else
            nbytes = 0;
    } else if (has("(4, (32 8 23 0 1 9 0 127))")) {
        nbytes = 4;
        if (has("(4, (4 3 2 1))"))
            swap = false;
        else if (has("(4, (1 2 3 4))"))
            swap = true;
        else
            nbytes = 0;
    }
    if (!nbytes) return fail("FAB real format not supported (IEEE double or float): " + desc, KSGF_ERR_UNSUPPORTED);
    {  // the native byte order is little-endian here; "swap" means big-endian data
        const uint16_t one = 1;
        if (*(const uint8_t*)&one != 1) swap = !swap;
    }
    const long long n = box.npts() * ncomp;
    std::vector<unsigned char> raw((size_t)n * nbytes);
    if (!is.read((char*)raw.data(), (std::streamsize)raw.size())) return fail("short FAB data in " + file);
    out.resize((size_t)n);
    for (long long q = 0; q < n; ++q) {
        unsigned char* p = raw.data() + q * nbytes;
        if (swap) std::reverse(p, p + nbytes);
        if (nbytes == 8) {
            std::memcpy(&out[q], p, 8);
        } else {
            float f;
            std::memcpy(&f, p, 4);
            out[q] = (double)f;
        }
    }
    return KSGF_OK;
}

// Read the MultiFab <prefix>_H into h (layout of ba, itype, ngrow).  With the
// same boxes in the same order, every FAB's stored points that both sides
// hold are copied (ghosts included); otherwise the valid parts of the file's
// boxes are copied into every FAB's valid box they overlap (a ParallelCopy).
int vismf_read(const std::string& prefix, const ksgf_boxarray* ba, HostMF& h, int itype, int ncomp, int ngrow)
{
    VisMFHeader H;
    int e = read_vismf_header(prefix + "_H", H);
    if (e) return e;
    if (H.ncomp != ncomp)
        return fail("file has " + std::to_string(H.ncomp) + " components, the MultiFab " + std::to_string(ncomp));
    const std::string dir = prefix.substr(0, prefix.find_last_of('/') + 1);
    bool same = (int)H.boxes.size() == ba->nbox;
    for (int b = 0; same && b < ba->nbox; ++b) {
        const IBox mine = ba_box(ba, b, itype);
        for (int d = 0; d < 3; ++d)
            if (mine.lo[d] != H.boxes[b].lo[d] || mine.hi[d] != H.boxes[b].hi[d] || mine.t[d] != H.boxes[b].t[d])
                same = false;
    }
    for (size_t f = 0; f < H.fabs.size(); ++f) {
        IBox fb;
        int nc;
        std::vector<double> data;
        if ((e = read_fab(dir + H.fabs[f].file, H.fabs[f].offset, fb, nc, data))) return e;
        if (nc != ncomp) return fail("FAB component count differs from the header in " + prefix);
        for (int d = 0; d < 3; ++d)
            if (fb.t[d] != ((itype == d + 1) ? 1 : 0)) return fail("index type differs from the MultiFab's: " + prefix);
        IBox fvalid = H.boxes[f];
        for (int b = 0; b < ba->nbox; ++b) {
            if (same && b != (int)f) continue;
            const IBox stored = fab_box(ba, b, itype, ngrow);
            IBox part;
            const bool hit = same ? intersect(fb, stored, part)
                                  : intersect(fvalid, fab_box(ba, b, itype, 0), part);
            if (!hit) continue;
            copy_part(data.data(), fb, h.v.data() + (size_t)ncomp * h.pre[b], stored, part, ncomp);
        }
    }
    return KSGF_OK;
}

int upload(const HostMF& h, double* mf, hipStream_t st)
{
    if (hipMemcpyAsync(mf, h.v.data(), h.v.size() * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail("MultiFab upload", KSGF_ERR_HIP);
    return KSGF_OK;
}

}  // namespace
}  // namespace ksgf

using namespace ksgf;

extern "C" {

int ksgf_plotfile_write(const char* dir, const ksgf_boxarray* ba, const double* mf, int ncomp, int ngrow,
                        const char* const* names, double time, int step, ksgf_stream_t stream)
{
    if (!dir || !ba || !mf || !names || ncomp <= 0 || ngrow < 0) return fail("ksgf_plotfile_write: bad argument");
    for (int n = 0; n < ncomp; ++n)
        if (!names[n] || !names[n][0]) return fail("ksgf_plotfile_write: empty variable name");
    HostMF h;
    int e;
    if ((e = download(ba, mf, KSGF_CELL, ncomp, ngrow, (hipStream_t)stream, h))) return e;
    const std::string root(dir);
    if ((e = make_dir(root)) || (e = make_dir(root + "/Level_0"))) return e;
    if ((e = vismf_write(root + "/Level_0/Cell", ba, h, KSGF_CELL, ncomp, ngrow, 0))) return e;
    const ksgf_geometry& g = ba->geom;
    std::ofstream hf(root + "/Header", std::ios::trunc);
    if (!hf) return fail("cannot write " + root + "/Header");
    hf << std::setprecision(17);
    hf << "HyperCLaw-V1.1\n" << ncomp << '\n';
    for (int n = 0; n < ncomp; ++n) hf << names[n] << '\n';
    hf << 3 << '\n' << time << '\n' << 0 << '\n';
    for (int d = 0; d < 3; ++d) hf << g.prob_lo[d] << ' ';
    hf << '\n';
    for (int d = 0; d < 3; ++d) hf << g.prob_hi[d] << ' ';
    hf << '\n' << '\n';  // refinement ratios: none for one level
    IBox dom;
    for (int d = 0; d < 3; ++d) {
        dom.lo[d] = g.domain_lo[d];
        dom.hi[d] = g.domain_hi[d];
        dom.t[d] = 0;
    }
    hf << box_text(dom) << " \n" << step << " \n";
    for (int d = 0; d < 3; ++d) hf << ba->dx[d] << ' ';
    hf << '\n' << 0 << '\n' << 0 << '\n';  // Cartesian coordinates, no boundary width
    hf << 0 << ' ' << ba->nbox << ' ' << time << '\n' << step << '\n';
    for (int b = 0; b < ba->nbox; ++b)
        for (int d = 0; d < 3; ++d) {
            const double lo = g.prob_lo[d] + ba->dx[d] * (ba->hbox[6 * b + d] - g.domain_lo[d]);
            const double hi = g.prob_lo[d] + ba->dx[d] * (ba->hbox[6 * b + 3 + d] + 1 - g.domain_lo[d]);
            hf << lo << ' ' << hi << '\n';
        }
    hf << "Level_0/Cell\n";
    hf.close();
    if (!hf) return fail("write failed: " + root + "/Header");
    return KSGF_OK;
}

int ksgf_plotfile_read(const char* dir, const ksgf_boxarray* ba, double* mf, int ncomp, int ngrow,
                       ksgf_stream_t stream)
{
    if (!dir || !ba || !mf || ncomp <= 0 || ngrow < 0) return fail("ksgf_plotfile_read: bad argument");
    const std::string root(dir);
    std::ifstream is(root + "/Header");
    if (!is) return fail("cannot open " + root + "/Header");
    std::string line;
    std::getline(is, line);
    if (line.compare(0, 9, "HyperCLaw") != 0) return fail("not a plotfile header: " + root + "/Header");
    int nc = 0;
    is >> nc;
    if (nc != ncomp) return fail("plotfile has " + std::to_string(nc) + " components, the MultiFab " + std::to_string(ncomp));
    std::getline(is, line);
    for (int n = 0; n < nc; ++n) std::getline(is, line);
    int dim, finest;
    double t;
    is >> dim >> t >> finest;
    if (dim != 3) return fail("plotfile is not three-dimensional", KSGF_ERR_UNSUPPORTED);
    // the level-0 MultiFab path is the last line of the level-0 block
    std::string mfpath;
    int nlev_seen = -1;
    while (std::getline(is, line)) {
        if (line.compare(0, 6, "Level_") == 0) {
            mfpath = line;
            if (++nlev_seen == 0) break;
        }
    }
    if (mfpath.empty()) return fail("no level-0 MultiFab in " + root + "/Header");
    while (!mfpath.empty() && (mfpath.back() == '\r' || mfpath.back() == ' ')) mfpath.pop_back();
    HostMF h;
    int e;
    if ((e = download(ba, mf, KSGF_CELL, ncomp, ngrow, (hipStream_t)stream, h))) return e;
    if ((e = vismf_read(root + "/" + mfpath, ba, h, KSGF_CELL, ncomp, ngrow))) return e;
    return upload(h, mf, (hipStream_t)stream);
}

int ksgf_vismf_write(const char* prefix, const ksgf_boxarray* ba, const double* mf, int itype, int ncomp, int ngrow,
                     ksgf_stream_t stream)
{
    if (!prefix || !ba || !mf || itype < 0 || itype > 3 || ncomp <= 0 || ngrow < 0)
        return fail("ksgf_vismf_write: bad argument");
    HostMF h;
    int e;
    if ((e = download(ba, mf, itype, ncomp, ngrow, (hipStream_t)stream, h))) return e;
    return vismf_write(prefix, ba, h, itype, ncomp, ngrow, ngrow);
}

int ksgf_vismf_read(const char* prefix, const ksgf_boxarray* ba, double* mf, int itype, int ncomp, int ngrow,
                    ksgf_stream_t stream)
{
    if (!prefix || !ba || !mf || itype < 0 || itype > 3 || ncomp <= 0 || ngrow < 0)
        return fail("ksgf_vismf_read: bad argument");
    HostMF h;
    int e;
    if ((e = download(ba, mf, itype, ncomp, ngrow, (hipStream_t)stream, h))) return e;
    if ((e = vismf_read(prefix, ba, h, itype, ncomp, ngrow))) return e;
    return upload(h, mf, (hipStream_t)stream);
}

int ksgf_checkpoint_write_header(const char* dir, const ksgf_boxarray* ba, int step, double time, double dt,
                                 double dt_nm1, double dt_nm2)
{
    if (!dir || !ba) return fail("ksgf_checkpoint_write_header: bad argument");
    const std::string root(dir);
    int e;
    if ((e = make_dir(root)) || (e = make_dir(root + "/Level_0"))) return e;
    std::ofstream hf(root + "/Header", std::ios::trunc | std::ios::binary);
    if (!hf) return fail("cannot write " + root + "/Header");
    hf << std::setprecision(17);
    hf << "Checkpoint version: 1\n" << 0 << '\n' << step << '\n' << time << '\n' << dt << '\n' << dt_nm1 << '\n'
       << dt_nm2 << '\n';
    for (int d = 0; d < 3; ++d) hf << ba->geom.prob_lo[d] << ' ';
    hf << '\n';
    for (int d = 0; d < 3; ++d) hf << ba->geom.prob_hi[d] << ' ';
    hf << '\n';
    hf << '(' << ba->nbox << " 0\n";
    for (int b = 0; b < ba->nbox; ++b) hf << box_text(ba_box(ba, b, KSGF_CELL)) << '\n';
    hf << ")\n";
    hf.close();
    if (!hf) return fail("write failed: " + root + "/Header");
    return KSGF_OK;
}

}  // extern "C"
