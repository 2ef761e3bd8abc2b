// CPU emulation of the fused red-black sweep kernel (k_rb_sweep).
//
// The workgroup-phase bodies in ksgf_mg_grid.h are driven here thread by
// thread -- every thread of the red phase, then every thread of the black
// phase, for each plane and each workgroup, exactly the order the barriers
// impose on the GPU -- and the result is compared bit for bit with a plain
// in-place half-sweep red-black Gauss-Seidel (explicit ghost fills before
// each colour).  Ghost layers and row padding of the input hold NaN, so a
// read of anything but an interior cell shows up in the output.  Built with
// AddressSanitizer, every array access (including the LDS planes,
// which are separate heap arrays here) is bounds checked.
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

static LevelGeom make_level(int n0, int n1, int n2, const int lo[3], const int hi[3], double h0, double h1, double h2)
{
    LevelGeom L{};
    L.n0 = n0;
    L.n1 = n1;
    L.n2 = n2;
    level_layout(L);
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

// in-place half sweeps (red then black) from the initial iterate y
template <typename T>
static void reference_sweep(std::vector<T>& y, const std::vector<T>& b, const LevelGeom& L)
{
    const T cx = (T)L.cx, cy = (T)L.cy, cz = (T)L.cz, idg = (T)L.idg;
    for (int color = 0; color < 2; ++color) {
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
}

template <typename T, int MODE>
static void emulate(const SweepArgs<T>& a0, int kz)
{
    SweepArgs<T> a = a0;
    const int nb = sweep_blocks(a, kz);
    const T nan = std::numeric_limits<T>::quiet_NaN();
    std::vector<std::vector<T>> X(4, std::vector<T>(kSwXPlane)), B(3, std::vector<T>(kSwPlane)),
        R(3, std::vector<T>(kSwPlane));
    std::vector<std::vector<T>> xa(kSwThreads, std::vector<T>(kSwXPer)), ba(kSwThreads, std::vector<T>(kSwBPer));
    std::vector<std::vector<T>> xb(kSwThreads, std::vector<T>(kSwXPer)), bb(kSwThreads, std::vector<T>(kSwBPer));
    std::vector<bool> seen(nb, false);
    for (int bid = 0; bid < nb; ++bid) {
        const int lid = xcd_remap(bid, nb);
        if (lid < 0 || lid >= nb || seen[lid]) {
            std::printf("xcd_remap not a permutation (nb %d)\n", nb);
            std::exit(1);
        }
        seen[lid] = true;
        int tx, ty, kc;
        sweep_tile(lid, a.ntx, a.nty, tx, ty, kc);
        for (auto& p : X) std::fill(p.begin(), p.end(), nan);
        for (auto& p : B) std::fill(p.begin(), p.end(), nan);
        for (auto& p : R) std::fill(p.begin(), p.end(), nan);
        const int k0 = kc * a.kz;
        const int k1 = std::min(k0 + a.kz, a.L.n2);
        auto xs = [&](int p) { return (p - k0 + 2) & 3; };
        auto bs = [&](int p) { return (p - k0 + 1) % 3; };
        // the kernel's pipeline: two register sets, loads two planes ahead
        std::vector<SwThread> th(kSwThreads);
        for (int tid = 0; tid < kSwThreads; ++tid) {
            sw_thread_init<T>(a, th[tid], tid, tx, ty);
            sw_load_x<T, MODE>(a, th[tid], xa[tid].data(), k0 - 2);
            sw_store_x<T>(X[0].data(), xa[tid].data(), tid);
            sw_load_x<T, MODE>(a, th[tid], xa[tid].data(), k0 - 1);
            sw_store_x<T>(X[1].data(), xa[tid].data(), tid);
            sw_load_x<T, MODE>(a, th[tid], xa[tid].data(), k0);
            sw_load_b<T>(a, th[tid], ba[tid].data(), k0 - 1);
            sw_load_x<T, MODE>(a, th[tid], xb[tid].data(), k0 + 1);
            sw_load_b<T>(a, th[tid], bb[tid].data(), k0);
        }
        auto step = [&](std::vector<std::vector<T>>& xv, std::vector<std::vector<T>>& bv, int k) {
            for (int tid = 0; tid < kSwThreads; ++tid) {
                sw_store_x<T>(X[xs(k + 1)].data(), xv[tid].data(), tid);
                sw_store_b<T>(B[bs(k)].data(), bv[tid].data(), tid);
                if (k + 3 <= k1 + 1) sw_load_x<T, MODE>(a, th[tid], xv[tid].data(), k + 3);
                if (k + 2 <= k1) sw_load_b<T>(a, th[tid], bv[tid].data(), k + 2);
            }
            for (int tid = 0; tid < kSwThreads; ++tid)
                sw_phase_red<T>(a.L, th[tid], X[xs(k - 1)].data(), X[xs(k)].data(), X[xs(k + 1)].data(),
                                B[bs(k)].data(), R[bs(k)].data(), k);
            if (k > k0) {
                const int kb = k - 1;
                for (int tid = 0; tid < kSwThreads; ++tid) {
                    long long at;
                    T v0, v1;
                    if (sw_phase_black<T>(a.L, th[tid], R[bs(kb - 1)].data(), R[bs(kb)].data(), R[bs(k)].data(),
                                          X[xs(kb)].data(), B[bs(kb)].data(), kb, at, v0, v1)) {
                        if ((at & 1) != 0) {
                            std::printf("unaligned pair store at %lld\n", at);
                            std::exit(1);
                        }
                        a.xout[at] = v0;
                        a.xout[at + 1] = v1;
                    }
                }
            }
        };
        for (int k = k0 - 1; k <= k1; k += 2) {
            step(xa, ba, k);
            if (k + 1 <= k1) step(xb, bb, k + 1);
        }
    }
}

static int g_cases = 0;

template <typename T, int MODE>
static void run_case(int n0, int n1, int n2, const int lo[3], const int hi[3], const int ratio[3], int kz,
                     unsigned seed)
{
    const LevelGeom L = make_level(n0, n1, n2, lo, hi, 1.0, 1.7, 0.45);
    const LevelGeom C = make_level(n0 / ratio[0], n1 / ratio[1], n2 / ratio[2], lo, hi, 2.0, 3.4, 0.9);
    const T nan = std::numeric_limits<T>::quiet_NaN();
    std::vector<T> x(L.tot, nan), b(L.tot, nan), xc(C.tot, nan), out(L.tot, nan);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                x[LIdx(L, i, j, k)] = (T)U(rng);
                b[LIdx(L, i, j, k)] = (T)U(rng);
            }
    for (int k = 0; k < C.n2; ++k)
        for (int j = 0; j < C.n1; ++j)
            for (int i = 0; i < C.n0; ++i) xc[LIdx(C, i, j, k)] = (T)U(rng);
    // reference
    std::vector<T> y(L.tot, nan);
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                const long long c = LIdx(L, i, j, k);
                if (MODE == kSwZero)
                    y[c] = (T)0;
                else if (MODE == kSwProlong)
                    y[c] = x[c] + xc[LIdx(C, i / ratio[0], j / ratio[1], k / ratio[2])];
                else
                    y[c] = x[c];
            }
    reference_sweep(y, b, L);
    SweepArgs<T> a{};
    a.xin = x.data();
    a.xout = out.data();
    a.b = b.data();
    a.xc = xc.data();
    a.L = L;
    a.C = C;
    a.s0 = ratio[0] == 2;
    a.s1 = ratio[1] == 2;
    a.s2 = ratio[2] == 2;
    emulate<T, MODE>(a, kz);
    long long bad = 0;
    for (long long c = 0; c < L.tot; ++c) {
        // interior cells must match bit for bit, everything else stays untouched
        const long long q = c / L.sz - 1, r = c % L.sz, jj = r / L.sy - 1, ii = r % L.sy - kXOff;
        const bool inside = q >= 0 && q < n2 && jj >= 0 && jj < n1 && ii >= 0 && ii < n0;
        if (inside) {
            if (std::memcmp(&out[c], &y[c], sizeof(T)) != 0) {
                if (bad < 5)
                    std::printf("mismatch at (%lld,%lld,%lld): %.17g vs %.17g\n", ii, jj, q, (double)out[c],
                                (double)y[c]);
                ++bad;
            }
        } else if (!std::isnan((double)out[c])) {
            if (bad < 5) std::printf("write outside the interior at %lld\n", c);
            ++bad;
        }
    }
    if (bad) {
        std::printf("FAIL case %dx%dx%d mode %d kz %d T%zu bc lo %d%d%d hi %d%d%d: %lld bad\n", n0, n1, n2, MODE, kz,
                    sizeof(T), lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], bad);
        std::exit(1);
    }
    ++g_cases;
}

template <typename T>
static void run_all_modes(int n0, int n1, int n2, const int lo[3], const int hi[3], const int ratio[3], int kz,
                          unsigned seed)
{
    run_case<T, kSwRead>(n0, n1, n2, lo, hi, ratio, kz, seed);
    run_case<T, kSwZero>(n0, n1, n2, lo, hi, ratio, kz, seed + 1);
    run_case<T, kSwProlong>(n0, n1, n2, lo, hi, ratio, kz, seed + 2);
}

int main()
{
    const int P = KSGF_LO_PERIODIC, N = KSGF_LO_NEUMANN, D = KSGF_LO_DIRICHLET;
    const int allp[3] = {P, P, P};
    const int abl_lo[3] = {P, P, N}, abl_hi[3] = {P, P, N};
    const int r222[3] = {2, 2, 2}, r112[3] = {1, 1, 2}, r221[3] = {2, 2, 1};
    const int mixlo[3] = {D, N, P}, mixhi[3] = {N, D, P};
    const int wall_lo[3] = {D, N, N}, wall_hi[3] = {D, N, D};
    for (int kz : {1, 3, 4, 16}) {
        run_all_modes<double>(128, 64, 32, allp, allp, r222, kz, 11u + kz);
        run_all_modes<float>(128, 64, 32, allp, allp, r222, kz, 21u + kz);
        run_all_modes<double>(96, 40, 20, abl_lo, abl_hi, r112, kz, 31u + kz);
        run_all_modes<float>(96, 40, 20, abl_lo, abl_hi, r112, kz, 41u + kz);
    }
    for (int kz : {2, 5}) {
        run_all_modes<double>(34, 6, 5, wall_lo, wall_hi, r221, kz, 51u + kz);
        run_all_modes<double>(66, 10, 6, mixlo, mixhi, r222, kz, 61u + kz);
        run_all_modes<float>(66, 10, 6, mixlo, mixhi, r222, kz, 71u + kz);
        run_all_modes<double>(2, 2, 2, allp, allp, r222, kz, 81u + kz);
        run_all_modes<double>(4, 18, 8, abl_lo, abl_hi, r222, kz, 91u + kz);
    }
    std::printf("ok %d\n", g_cases);
    return 0;
}
