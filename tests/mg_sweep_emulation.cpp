// CPU check of the packed red-black multigrid kernels.
//
// The per-cell bodies the HIP kernels run (ksgf_mg_grid.h: pk_relax,
// pk_restrict_cell) are driven here cell by cell on packed arrays and
// compared bit for bit with a plain natural-layout implementation: in-place
// half sweeps after explicit ghost fills, and residual + average restriction
// from a ghost-filled array.  Every sweep mode (plain, zero initial guess,
// fused prolongation), both colours, periodic / Neumann / Dirichlet sides,
// semicoarsened ratios, FP32 and FP64; the four-cell form of the plain and
// zero-guess half sweeps against the one-cell form.  Built with
// AddressSanitizer, so every array access is bounds checked.
//
// Prints "ok <n>" and exits 0 when every case matches.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#include "../kynema-sgf_amd/csrc/ksgf_mg_grid.h"

using namespace ksgf;

static LevelGeom make_level(int n0, int n1, int n2, const int lo[3], const int hi[3], double h0, double h1, double h2,
                            bool packed)
{
    LevelGeom L{};
    L.n0 = n0;
    L.n1 = n1;
    L.n2 = n2;
    level_layout(L, packed);
    L.cx = 1.0 / (h0 * h0);
    L.cy = 1.0 / (h1 * h1);
    L.cz = 1.0 / (h2 * h2);
    L.dg = 2.0 * (L.cx + L.cy + L.cz);
    L.idg = 1.0 / L.dg;
    for (int d = 0; d < 3; ++d) {
        L.lo[d] = lo[d];
        L.hi[d] = hi[d];
    }
    L.ghost_mode = 0;
    return L;
}

template <typename T>
static T ghost_of(int bc, T near_v, T far_v)
{
    return bc == KSGF_LO_PERIODIC ? far_v : (bc == KSGF_LO_NEUMANN ? near_v : -near_v);
}

template <typename T>
static void fill_ghosts(std::vector<T>& a, const LevelGeom& L)
{
    for (int k = 0; k < L.n2; ++k)
        for (int j = 0; j < L.n1; ++j) {
            a[LIdx(L, -1, j, k)] = ghost_of<T>(L.lo[0], a[LIdx(L, 0, j, k)], a[LIdx(L, L.n0 - 1, j, k)]);
            a[LIdx(L, L.n0, j, k)] = ghost_of<T>(L.hi[0], a[LIdx(L, L.n0 - 1, j, k)], a[LIdx(L, 0, j, k)]);
        }
    for (int k = 0; k < L.n2; ++k)
        for (int i = 0; i < L.n0; ++i) {
            a[LIdx(L, i, -1, k)] = ghost_of<T>(L.lo[1], a[LIdx(L, i, 0, k)], a[LIdx(L, i, L.n1 - 1, k)]);
            a[LIdx(L, i, L.n1, k)] = ghost_of<T>(L.hi[1], a[LIdx(L, i, L.n1 - 1, k)], a[LIdx(L, i, 0, k)]);
        }
    for (int j = 0; j < L.n1; ++j)
        for (int i = 0; i < L.n0; ++i) {
            a[LIdx(L, i, j, -1)] = ghost_of<T>(L.lo[2], a[LIdx(L, i, j, 0)], a[LIdx(L, i, j, L.n2 - 1)]);
            a[LIdx(L, i, j, L.n2)] = ghost_of<T>(L.hi[2], a[LIdx(L, i, j, L.n2 - 1)], a[LIdx(L, i, j, 0)]);
        }
}

// in-place half sweep of one colour on a natural array
template <typename T>
static void reference_half(std::vector<T>& y, const std::vector<T>& b, const LevelGeom& L, int color)
{
    const T cx = (T)L.cx, cy = (T)L.cy, cz = (T)L.cz, idg = (T)L.idg;
    fill_ghosts(y, L);
    for (int k = 0; k < L.n2; ++k)
        for (int j = 0; j < L.n1; ++j)
            for (int i = 0; i < L.n0; ++i) {
                if (((i + j + k + color) & 1) != 0) continue;
                const long long c = LIdx(L, i, j, k);
                y[c] = ((cx * (y[c - 1] + y[c + 1]) + cy * (y[c - L.sy] + y[c + L.sy]) +
                         cz * (y[c - L.sz] + y[c + L.sz])) -
                        b[c]) *
                       idg;
            }
}

static int g_cases = 0;

template <typename T>
static void fail(const char* what, int n0, int n1, int n2, int mode, int color, long long bad)
{
    std::printf("FAIL %s %dx%dx%d mode %d colour %d T%zu: %lld bad\n", what, n0, n1, n2, mode, color, sizeof(T), bad);
    std::exit(1);
}

template <typename T, int MODE>
static void run_half(int n0, int n1, int n2, const int lo[3], const int hi[3], const int ratio[3], int color,
                     bool coarse_packed, unsigned seed)
{
    const LevelGeom N = make_level(n0, n1, n2, lo, hi, 1.0, 1.7, 0.45, false);
    const LevelGeom P = make_level(n0, n1, n2, lo, hi, 1.0, 1.7, 0.45, true);
    const LevelGeom C = make_level(n0 / ratio[0], n1 / ratio[1], n2 / ratio[2], lo, hi, 2.0, 3.4, 0.9,
                                   coarse_packed && (n0 / ratio[0]) % 2 == 0);
    const T nan = std::numeric_limits<T>::quiet_NaN();
    std::vector<T> xn(N.tot, nan), bn(N.tot, nan), xp(P.tot, nan), bp(P.tot, nan), xc(C.tot, nan);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                const T x = (T)U(rng), b = (T)U(rng);
                xn[LIdx(N, i, j, k)] = x;
                bn[LIdx(N, i, j, k)] = b;
                xp[PIdx(P, i, j, k)] = x;
                bp[PIdx(P, i, j, k)] = b;
            }
    for (int k = 0; k < C.n2; ++k)
        for (int j = 0; j < C.n1; ++j)
            for (int i = 0; i < C.n0; ++i) xc[AnyIdx(C, i, j, k)] = (T)U(rng);
    // reference: the natural array the mode stands for, then the half sweep
    std::vector<T> y(N.tot, nan);
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                const long long c = LIdx(N, i, j, k);
                const bool mine = ((i + j + k) & 1) == color;
                const T pro = xn[c] + xc[AnyIdx(C, i / ratio[0], j / ratio[1], k / ratio[2])];
                if (MODE == kPkZero)
                    y[c] = (color == 0 || mine) ? (T)0 : xn[c];  // red: all zero; black: blacks zero
                else if (MODE == kPkProlong)
                    y[c] = (color == 0 || mine) ? pro : xn[c];   // red: all prolonged; black: blacks
                else
                    y[c] = xn[c];
            }
    reference_half(y, bn, N, color);
    PkArgs<T> a{};
    a.x = xp.data();
    a.b = bp.data();
    a.xc = xc.data();
    a.L = P;
    a.C = C;
    a.s0 = ratio[0] == 2;
    a.s1 = ratio[1] == 2;
    a.s2 = ratio[2] == 2;
    // the four-cell form (plain and zero-guess modes, rows of a multiple of
    // 4 packed cells) on 64-byte aligned copies of the same arrays
    const bool quads = (MODE == kPkPlain || MODE == kPkZero) && n0 % 8 == 0;
    T *xq = nullptr, *bq = nullptr;
    if (quads) {
        const size_t bytes = ((size_t)P.tot * sizeof(T) + 63) / 64 * 64;
        xq = (T*)std::aligned_alloc(64, bytes);
        bq = (T*)std::aligned_alloc(64, bytes);
        std::memcpy(xq, xp.data(), (size_t)P.tot * sizeof(T));
        std::memcpy(bq, bp.data(), (size_t)P.tot * sizeof(T));
        PkArgs<T> q = a;
        q.x = xq;
        q.b = bq;
        for (int k = 0; k < n2; ++k)
            for (int j = 0; j < n1; ++j)
                for (int hq = 0; hq < n0 / 8; ++hq)
                    pk_relax4<T, MODE == kPkZero ? kPkZero : kPkPlain, false>(q, color, hq, j, k);
    }
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int h = 0; h < n0 / 2; ++h) {
                if (C.packed)
                    pk_relax<T, MODE, true>(a, color, h, j, k);
                else
                    pk_relax<T, MODE, false>(a, color, h, j, k);
            }
    long long bad = 0;
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                const bool mine = ((i + j + k) & 1) == color;
                const T g = xp[PIdx(P, i, j, k)];
                const T r = mine ? y[LIdx(N, i, j, k)] : xn[LIdx(N, i, j, k)];  // the other colour is not written
                if (std::memcmp(&g, &r, sizeof(T)) != 0) {
                    if (bad < 3) std::printf("  (%d,%d,%d): %.17g vs %.17g\n", i, j, k, (double)g, (double)r);
                    ++bad;
                }
            }
    if (bad) fail<T>("half sweep", n0, n1, n2, MODE, color, bad);
    ++g_cases;
    if (quads) {
        if (std::memcmp(xq, xp.data(), (size_t)P.tot * sizeof(T)) != 0) fail<T>("four-cell half sweep", n0, n1, n2, MODE, color, 1);
        ++g_cases;
        std::free(xq);
        std::free(bq);
    }
}

// the first post-smoothing sweep in two forms: on-the-fly prolongation
// (red: kPkProlong, black: kPkProlong) against the correction added to the
// black cells first (pk_prolong_color), then red kPkProlongWalls and black
// kPkPlain.  Both colours must come out bit for bit the same.
template <typename T>
static void run_split_prolong(int n0, int n1, int n2, const int lo[3], const int hi[3], const int ratio[3],
                              bool coarse_packed, unsigned seed)
{
    const LevelGeom P = make_level(n0, n1, n2, lo, hi, 1.0, 1.7, 0.45, true);
    const LevelGeom C = make_level(n0 / ratio[0], n1 / ratio[1], n2 / ratio[2], lo, hi, 2.0, 3.4, 0.9,
                                   coarse_packed && (n0 / ratio[0]) % 2 == 0);
    const T nan = std::numeric_limits<T>::quiet_NaN();
    std::vector<T> x1(P.tot, nan), bp(P.tot, nan), xc(C.tot, nan);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                x1[PIdx(P, i, j, k)] = (T)U(rng);
                bp[PIdx(P, i, j, k)] = (T)U(rng);
            }
    for (int k = 0; k < C.n2; ++k)
        for (int j = 0; j < C.n1; ++j)
            for (int i = 0; i < C.n0; ++i) xc[AnyIdx(C, i, j, k)] = (T)U(rng);
    std::vector<T> x2 = x1;
    const std::vector<T> x0 = x1;
    auto args = [&](std::vector<T>& x) {
        PkArgs<T> a{};
        a.x = x.data();
        a.b = bp.data();
        a.xc = xc.data();
        a.L = P;
        a.C = C;
        a.s0 = ratio[0] == 2;
        a.s1 = ratio[1] == 2;
        a.s2 = ratio[2] == 2;
        return a;
    };
    auto each = [&](auto&& f) {
        for (int k = 0; k < n2; ++k)
            for (int j = 0; j < n1; ++j)
                for (int h = 0; h < n0 / 2; ++h) f(h, j, k);
    };
    const PkArgs<T> a1 = args(x1), a2 = args(x2);
    const bool cp = C.packed;
    each([&](int h, int j, int k) { cp ? pk_relax<T, kPkProlong, true>(a1, 0, h, j, k) : pk_relax<T, kPkProlong, false>(a1, 0, h, j, k); });
    each([&](int h, int j, int k) { cp ? pk_relax<T, kPkProlong, true>(a1, 1, h, j, k) : pk_relax<T, kPkProlong, false>(a1, 1, h, j, k); });
    each([&](int h, int j, int k) { cp ? pk_prolong_color<T, true>(a2, 1, h, j, k) : pk_prolong_color<T, false>(a2, 1, h, j, k); });
    each([&](int h, int j, int k) { cp ? pk_relax<T, kPkProlongWalls, true>(a2, 0, h, j, k) : pk_relax<T, kPkProlongWalls, false>(a2, 0, h, j, k); });
    each([&](int h, int j, int k) { pk_relax<T, kPkPlain, false>(a2, 1, h, j, k); });
    long long bad = 0;
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                const T g = x2[PIdx(P, i, j, k)], r = x1[PIdx(P, i, j, k)];
                if (std::memcmp(&g, &r, sizeof(T)) != 0) ++bad;
            }
    // the split sweep in the four-cell form, on 64-byte aligned copies
    if (n0 % 8 == 0) {
        const size_t bytes = ((size_t)P.tot * sizeof(T) + 63) / 64 * 64;
        T* xq = (T*)std::aligned_alloc(64, bytes);
        T* bq = (T*)std::aligned_alloc(64, bytes);
        std::memcpy(xq, x0.data(), (size_t)P.tot * sizeof(T));
        std::memcpy(bq, bp.data(), (size_t)P.tot * sizeof(T));
        PkArgs<T> q = a2;
        q.x = xq;
        q.b = bq;
        auto quads = [&](auto&& f) {
            for (int k = 0; k < n2; ++k)
                for (int j = 0; j < n1; ++j)
                    for (int hq = 0; hq < n0 / 8; ++hq) f(hq, j, k);
        };
        quads([&](int h, int j, int k) { cp ? pk_prolong_color4<T, true>(q, 1, h, j, k) : pk_prolong_color4<T, false>(q, 1, h, j, k); });
        quads([&](int h, int j, int k) { cp ? pk_relax4<T, kPkProlongWalls, true>(q, 0, h, j, k) : pk_relax4<T, kPkProlongWalls, false>(q, 0, h, j, k); });
        quads([&](int h, int j, int k) { pk_relax4<T, kPkPlain, false>(q, 1, h, j, k); });
        if (std::memcmp(xq, x2.data(), (size_t)P.tot * sizeof(T)) != 0) ++bad;
        std::free(xq);
        std::free(bq);
    }
    if (bad) fail<T>("split prolongation sweep", n0, n1, n2, kPkProlongWalls, -1, bad);
    ++g_cases;
}

template <typename T>
static void run_restrict(int n0, int n1, int n2, const int lo[3], const int hi[3], const int ratio[3],
                         unsigned seed)
{
    const LevelGeom N = make_level(n0, n1, n2, lo, hi, 1.0, 1.7, 0.45, false);
    const LevelGeom P = make_level(n0, n1, n2, lo, hi, 1.0, 1.7, 0.45, true);
    const T nan = std::numeric_limits<T>::quiet_NaN();
    std::vector<T> xn(N.tot, nan), bn(N.tot, nan), xp(P.tot, nan), bp(P.tot, nan);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                const T x = (T)U(rng), b = (T)U(rng);
                xn[LIdx(N, i, j, k)] = x;
                bn[LIdx(N, i, j, k)] = b;
                xp[PIdx(P, i, j, k)] = x;
                bp[PIdx(P, i, j, k)] = b;
            }
    fill_ghosts(xn, N);
    // 64-byte aligned copies of the packed arrays for the four-cell form
    const size_t pbytes = ((size_t)P.tot * sizeof(T) + 63) / 64 * 64;
    T* xa = (T*)std::aligned_alloc(64, pbytes);
    T* ba = (T*)std::aligned_alloc(64, pbytes);
    std::memcpy(xa, xp.data(), (size_t)P.tot * sizeof(T));
    std::memcpy(ba, bp.data(), (size_t)P.tot * sizeof(T));
    const T cx = (T)N.cx, cy = (T)N.cy, cz = (T)N.cz;
    auto st = [&](long long c, int, int, int, T) {
        return cx * (xn[c - 1] + xn[c + 1]) + cy * (xn[c - N.sy] + xn[c + N.sy]) + cz * (xn[c - N.sz] + xn[c + N.sz]);
    };
    const T w = (T)(1.0 / (double)(ratio[0] * ratio[1] * ratio[2]));
    long long bad = 0;
    for (int K = 0; K < n2 / ratio[2]; ++K)
        for (int J = 0; J < n1 / ratio[1]; ++J)
            for (int I = 0; I < n0 / ratio[0]; ++I) {
                const T g = pk_restrict_cell<T>(xp.data(), bp.data(), P, I, J, K, ratio[0], ratio[1], ratio[2], w);
                const T r = restrict_cell<T>(xn.data(), bn.data(), N, I, J, K, ratio[0], ratio[1], ratio[2], w, st);
                if (std::memcmp(&g, &r, sizeof(T)) != 0) ++bad;
                // the four-cell form (x ratio 1) at the first cell of each quad
                if (ratio[0] == 1 && n0 % 4 == 0 && I % 4 == 0) {
                    T q[4];
                    pk_restrict_quad<T>(xa, ba, P, I, J, K, ratio[1], ratio[2], w, q);
                    for (int m = 0; m < 4; ++m) {
                        const T gm = pk_restrict_cell<T>(xp.data(), bp.data(), P, I + m, J, K, 1, ratio[1], ratio[2], w);
                        if (std::memcmp(&q[m], &gm, sizeof(T)) != 0) ++bad;
                    }
                }
            }
    std::free(xa);
    std::free(ba);
    if (bad) fail<T>("restriction", n0, n1, n2, -1, -1, bad);
    ++g_cases;
}

template <typename T>
static void run_all(int n0, int n1, int n2, const int lo[3], const int hi[3], const int ratio[3], unsigned seed)
{
    for (int color = 0; color < 2; ++color)
        for (int cp = 0; cp < 2; ++cp) {
            run_half<T, kPkPlain>(n0, n1, n2, lo, hi, ratio, color, cp, seed + 10 * color + cp);
            run_half<T, kPkZero>(n0, n1, n2, lo, hi, ratio, color, cp, seed + 10 * color + cp + 2);
            run_half<T, kPkProlong>(n0, n1, n2, lo, hi, ratio, color, cp, seed + 10 * color + cp + 4);
        }
    run_restrict<T>(n0, n1, n2, lo, hi, ratio, seed + 100);
    for (int cp = 0; cp < 2; ++cp) run_split_prolong<T>(n0, n1, n2, lo, hi, ratio, cp, seed + 200 + cp);
}

int main()
{
    const int P = KSGF_LO_PERIODIC, N = KSGF_LO_NEUMANN, D = KSGF_LO_DIRICHLET;
    const int allp[3] = {P, P, P};
    const int abl[3] = {P, P, N};
    const int r222[3] = {2, 2, 2}, r112[3] = {1, 1, 2}, r221[3] = {2, 2, 1};
    const int mixlo[3] = {D, N, P}, mixhi[3] = {N, D, P};
    const int wall_lo[3] = {D, N, N}, wall_hi[3] = {D, N, D};
    run_all<double>(16, 8, 12, allp, allp, r222, 11u);
    run_all<float>(16, 8, 12, allp, allp, r222, 21u);
    run_all<double>(12, 10, 8, abl, abl, r112, 31u);
    run_all<float>(12, 10, 8, abl, abl, r112, 41u);
    run_all<double>(10, 6, 5, wall_lo, wall_hi, r221, 51u);
    run_all<double>(6, 10, 6, mixlo, mixhi, r222, 61u);
    run_all<float>(6, 10, 6, mixlo, mixhi, r222, 71u);
    run_all<double>(2, 2, 2, allp, allp, r222, 81u);
    run_all<double>(4, 18, 8, abl, abl, r221, 91u);
    // rows of 4 and 8 packed cells with x walls and x periodic (four-cell form)
    run_all<double>(16, 6, 5, wall_lo, wall_hi, r221, 101u);
    run_all<float>(16, 6, 5, wall_lo, wall_hi, r221, 111u);
    run_all<double>(8, 10, 6, mixlo, mixhi, r222, 121u);
    run_all<float>(16, 10, 6, mixhi, mixlo, r222, 131u);
    run_all<float>(24, 8, 4, allp, allp, r112, 141u);
    // semicoarsened restriction of rows of 4 and 8 cells next to x walls
    const int r121[3] = {1, 2, 1};
    run_all<double>(16, 6, 6, wall_lo, wall_hi, r112, 151u);
    run_all<float>(8, 10, 6, mixlo, mixhi, r112, 161u);
    run_all<double>(8, 10, 6, mixhi, mixlo, r121, 171u);
    std::printf("ok %d\n", g_cases);
    return 0;
}
