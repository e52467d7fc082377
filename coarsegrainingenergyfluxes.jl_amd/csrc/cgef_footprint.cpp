// Host-side footprint construction (FP64) for libcgef_hip.so.
//
// The banded footprint is stored as RUNS: for every band (one per latitude row on the sphere, one
// per level on a stretched 3-D grid, a single band on a translation-invariant grid) and every
// source-row offset (dy, dz) the in-support columns form one contiguous range lo..hi, because the
// distance grows monotonically with |di| at fixed (dy, dz). Entries, order and weights are those of
// the reference's CSR footprint (Filtering.jl:636-732 for 2-D, 2172-2203 / 2268-2360 for 3-D):
// dz, then dy, then di ascending, weight kernel_weight(d)·measure. Runs let a kernel walk the
// footprint with wave-uniform loop bounds and weights, and let each source row be staged once.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <numeric>
#include <thread>

#include "cgef_internal.hpp"

namespace cgef {

static double kernel_weight(int kind, double alpha, double d, double ell) {
    // Kernels.jl:318-324
    if (kind == CGEF_TOPHAT) return d <= ell / 2.0 ? 1.0 : 0.0;
    double r = d / ell;
    return std::exp(-alpha * (r * r));
}

static int64_t iceil(double v) { return (int64_t)std::ceil(v); }

namespace {
// Runs and weights are collected per band (woff relative to the band), so bands can be built in
// parallel and concatenated in band order afterwards.
struct Builder {
    Plan &p;
    std::vector<std::vector<Run>> band_runs;
    std::vector<std::vector<double>> band_w;
    explicit Builder(Plan &pl) : p(pl) {}
    void resize(int nb) {
        band_runs.assign(nb, {});
        band_w.assign(nb, {});
    }

    // Appends the run of columns in [-lim_lo, lim_hi] whose distance passes d <= rad: the reference's
    // outward scan from di = 0, which stops at the first column that fails (Filtering.jl:665-671,
    // 705-720). dist(di) grows with |di|, so a binary search on each side finds where the scan would
    // stop; the weight pass then evaluates each column's distance once, from di = 0 outward, and still
    // cuts the run at the first failure it meets.
    template <class Dist, class Wt>
    void add_run(int band, int dy, int dz, int64_t lim_lo, int64_t lim_hi, Dist dist, Wt weight) {
        const double d0 = dist(0);
        if (d0 > p.rad) return;
        auto last_in = [&](int64_t sgn, int64_t lim) {  // largest x ≤ lim with dist(sgn·x) ≤ rad
            int64_t a = 0, b = lim;
            while (a < b) {
                const int64_t mid = a + (b - a + 1) / 2;
                if (dist(sgn * mid) <= p.rad) a = mid;
                else b = mid - 1;
            }
            return a;
        };
        thread_local std::vector<double> dpos, dneg;  // distances at di = +x and di = −x
        auto side = [&](int64_t sgn, int64_t lim, std::vector<double> &dv) {
            const int64_t cand = last_in(sgn, lim);
            dv.assign(1, d0);
            for (int64_t x = 1; x <= cand; x++) {
                const double d = dist(sgn * x);
                if (d > p.rad) break;
                dv.push_back(d);
            }
            return (int64_t)dv.size() - 1;
        };
        const int64_t hi = side(1, lim_hi, dpos);
        const int64_t lo = -side(-1, lim_lo, dneg);
        Run r{};
        r.dy = dy;
        r.dz = dz;
        r.lo = (int32_t)lo;
        r.hi = (int32_t)hi;
        std::vector<double> &w = band_w[band];
        r.woff = (int32_t)w.size();
        for (int64_t di = lo; di < 0; di++) w.push_back(weight(dneg[-di]));
        for (int64_t di = 0; di <= hi; di++) w.push_back(weight(dpos[di]));
        band_runs[band].push_back(r);
    }
};

// Runs fn(band) for band = 0..nb-1 on the host's cores (bands are independent).
template <class F>
void parallel_bands(int nb, F fn) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int nthreads = (int)std::min<unsigned>(hw, (unsigned)std::max(nb, 1));
    if (nthreads <= 1 || nb < 8) {
        for (int b = 0; b < nb; b++) fn(b);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++)
        pool.emplace_back([&] {
            for (int b = next++; b < nb; b = next++) fn(b);
        });
    for (auto &th : pool) th.join();
}
}  // namespace

void build_banded(Plan &p, bool need_fp32) {
    const Grid &g = *p.grid;
    const int64_t Nx = g.n[0], Ny = g.n[1], Nz = g.n[2];
    Builder B(p);
    p.edge_x = !g.periodic[0];
    if (g.geometry == CGEF_SPHERICAL) {
        // Filtering.jl:674-731 — one band per latitude row; rows clipped to 1:Ny; longitude window
        // capped at one turn, the antipodal column counted once (Filtering.jl:713).
        if (!g.uniform[0] || !g.uniform[1])
            fail(CGEF_ERR_UNSUPPORTED, "spherical banded footprint needs uniform (Range) lon/lat axes");
        const double R = g.R, dlam = g.coord[0][1] - g.coord[0][0], dphi = g.coord[1][1] - g.coord[1][0];
        const int64_t dj_lim = dphi > 0 ? iceil(p.rad / (R * dphi)) : 0;
        const int64_t half = Nx / 2;
        p.nbands = (int)Ny;
        B.resize((int)Ny);
        std::vector<double> area(Ny);
        for (int64_t jj = 0; jj < Ny; jj++) area[jj] = g.cell_area(0, jj, 0);
        parallel_bands((int)Ny, [&](int jb) {
            const int64_t j = jb;
            const double phi = g.coord[1][j];
            // A row whose support reaches the antipodal column covers the whole ring; the window is
            // then -Nx÷2 .. Nx÷2-1 on an even ring. A row that does not reach it is unaffected by
            // the cap, so the capped limits serve every row.
            const int64_t lim_lo = half;
            const int64_t lim_hi = (2 * half + 1 > Nx) ? half - 1 : half;
            for (int64_t ddj = -dj_lim; ddj <= dj_lim; ddj++) {
                int64_t jj = j + ddj;
                if (jj < 0 || jj >= Ny) continue;
                const double phi2 = g.coord[1][jj];
                const double A = area[jj];
                // haversine(R, phi, di·dλ, phi2) with its per-run factors taken out of the column loop
                // (the same operations in the same order, so the same values)
                const double s1 = std::sin((phi2 - phi) / 2.0);
                const double s1s1 = s1 * s1, c12 = std::cos(phi) * std::cos(phi2);
                auto dist = [&](int64_t di) {
                    const double s2 = std::sin((double)di * dlam / 2.0);
                    double a = s1s1 + c12 * (s2 * s2);
                    if (a > 1.0) a = 1.0;
                    return 2.0 * R * std::asin(std::sqrt(a));
                };
                auto wt = [&](double d) { return kernel_weight(p.kind, p.alpha, d, p.scale) * A; };
                B.add_run((int)j, (int)ddj, 0, lim_lo, lim_hi, dist, wt);
            }
        });
    } else if (g.ndim == 2) {
        // Filtering.jl:653-673 — Cartesian: one translation-invariant band, disk d <= rad.
        if (!g.uniform[0] || !g.uniform[1])
            fail(CGEF_ERR_UNSUPPORTED, "Cartesian banded footprint needs uniform (Range) axes");
        const double dx = g.coord[0][1] - g.coord[0][0], dy = g.coord[1][1] - g.coord[1][0];
        const double A = g.cell_area(0, 0, 0);
        const int64_t di_lim = dx > 0 ? iceil(p.rad / dx) : 0, dj_lim = dy > 0 ? iceil(p.rad / dy) : 0;
        p.nbands = 1;
        B.resize(1);
        for (int64_t ddj = -dj_lim; ddj <= dj_lim; ddj++) {
            const double b = (double)ddj * dy;
            auto dist = [&](int64_t di) {
                double a = (double)di * dx;
                return std::sqrt(a * a + b * b);
            };
            auto wt = [&](double d) { return kernel_weight(p.kind, p.alpha, d, p.scale) * A; };
            B.add_run(0, (int)ddj, 0, di_lim, di_lim, dist, wt);
        }
    } else {
        // 3-D Cartesian ball (Filtering.jl:2172-2203 uniform; 2268-2360 stretched): offsets in
        // CartesianIndices order (di fastest, then dj, then dk), d <= rad, weight
        // kernel_weight(d)·cell volume. x and y must be uniform: the footprint is then
        // translation-invariant in (i, j) and depends on the target level only through z.
        if (!g.uniform[0] || !g.uniform[1])
            fail(CGEF_ERR_UNSUPPORTED, "3-D banded footprint needs uniform (Range) x and y axes");
        const double dx = g.coord[0][1] - g.coord[0][0], dy = g.coord[1][1] - g.coord[1][0];
        const int64_t lx = dx > 0 ? iceil(p.rad / dx) : 0, ly = dy > 0 ? iceil(p.rad / dy) : 0;
        double gapz = INFINITY;
        for (int64_t k = 1; k < Nz; k++) gapz = std::min(gapz, std::fabs(g.coord[2][k] - g.coord[2][k - 1]));
        const int64_t lz = (std::isfinite(gapz) && gapz > 0) ? iceil(p.rad / gapz) : 0;
        const bool all_uniform = g.uniform[2] != 0;
        const double A = g.cell_area(0, 0, 0);
        p.nbands = all_uniform ? 1 : (int)Nz;
        B.resize(p.nbands);
        parallel_bands(p.nbands, [&](int band) {
            const int64_t k = band;  // target level (uniform: any level, offsets only)
            for (int64_t ddk = -lz; ddk <= lz; ddk++) {
                double ddz;
                double meas = A;
                if (all_uniform) {
                    ddz = (double)ddk * (Nz > 1 ? g.coord[2][1] - g.coord[2][0] : 0.0);
                } else {
                    int64_t kk = k + ddk;
                    double sh = 0.0;
                    if (kk < 0 || kk >= Nz) {
                        if (!g.periodic[2]) continue;
                        sh = kk < 0 ? -g.period[2] : g.period[2];
                        kk = ((kk % Nz) + Nz) % Nz;
                    }
                    ddz = (g.coord[2][kk] + sh) - g.coord[2][k];
                    meas = g.cell_area(0, 0, kk);
                }
                for (int64_t ddj = -ly; ddj <= ly; ddj++) {
                    const double b = (double)ddj * dy;
                    auto dist = [&](int64_t di) {
                        double a = (double)di * dx;
                        return std::sqrt(a * a + b * b + ddz * ddz);
                    };
                    auto wt = [&](double d) { return kernel_weight(p.kind, p.alpha, d, p.scale) * meas; };
                    B.add_run(band, (int)ddj, (int)ddk, lx, lx, dist, wt);
                }
            }
        });
    }

    // flatten runs and weights in band order (woff becomes global), band pointers
    size_t total = 0;
    for (int b = 0; b < p.nbands; b++) total += B.band_w[b].size();
    if (total >= (size_t)INT32_MAX) fail(CGEF_ERR_UNSUPPORTED, "footprint table exceeds 2^31 entries");
    std::vector<size_t> wbase(p.nbands + 1, 0);
    p.run_ptr.assign(p.nbands + 1, 0);
    for (int b = 0; b < p.nbands; b++) {
        wbase[b + 1] = wbase[b] + B.band_w[b].size();
        p.run_ptr[b + 1] = p.run_ptr[b] + (int32_t)B.band_runs[b].size();
    }
    p.w64.resize(total);
    p.runs.resize(p.run_ptr[p.nbands]);
    parallel_bands(p.nbands, [&](int b) {
        Run *dst = p.runs.data() + p.run_ptr[b];
        for (Run r : B.band_runs[b]) {
            r.woff += (int32_t)wbase[b];
            *dst++ = r;
        }
        std::copy(B.band_w[b].begin(), B.band_w[b].end(), p.w64.begin() + (ptrdiff_t)wbase[b]);
        std::vector<double>().swap(B.band_w[b]);
    });
    p.nentries = (int64_t)p.w64.size();

    // FP32 weights scaled per band to O(1) magnitude; the FP32 path normalizes with sums of the
    // rounded weights so that it filters exactly the operator its weights represent. (The ring
    // engine works in FP64 only and skips them.)
    if (need_fp32) {
        p.w32.resize(p.w64.size());
        p.w32d.resize(p.w64.size());
    }
    for (int b = 0; b < p.nbands && need_fp32; b++) {
        double tot = 0.0;
        for (int32_t q = p.run_ptr[b]; q < p.run_ptr[b + 1]; q++) {
            const Run &r = p.runs[q];
            for (int32_t e = 0; e <= r.hi - r.lo; e++) tot += p.w64[r.woff + e];
        }
        const double s = tot > 0 ? 1.0 / tot : 1.0;
        for (int32_t q = p.run_ptr[b]; q < p.run_ptr[b + 1]; q++) {
            const Run &r = p.runs[q];
            for (int32_t e = 0; e <= r.hi - r.lo; e++) {
                float f = (float)(p.w64[r.woff + e] * s);
                p.w32[r.woff + e] = f;
                p.w32d[r.woff + e] = (double)f;
            }
        }
    }

    // rows: target row r = j + Ny·k; per-row in-range denominators and tap counts
    const int64_t nrows = Ny * Nz;
    p.band_of_row.resize(nrows);
    p.D64_row.assign(nrows, 0.0);
    p.D32_row.assign(nrows, 0.0);
    std::vector<int64_t> cost(nrows, 0);
    const int64_t rchunk = 64;  // rows per task: every row's sums keep their order
    parallel_bands((int)((nrows + rchunk - 1) / rchunk), [&](int rb) {
    for (int64_t r = rb * rchunk; r < std::min<int64_t>(nrows, (rb + 1) * rchunk); r++) {
        const int64_t j = r % Ny, k = r / Ny;
        int band = (g.geometry == CGEF_SPHERICAL) ? (int)j : (p.nbands == 1 ? 0 : (int)k);
        p.band_of_row[r] = band;
        double d64 = 0.0, d32 = 0.0;
        int64_t taps = 0;
        for (int32_t q = p.run_ptr[band]; q < p.run_ptr[band + 1]; q++) {
            const Run &ru = p.runs[q];
            const int64_t jj = j + ru.dy, kk = k + ru.dz;
            if ((jj < 0 || jj >= Ny) && !g.periodic[1]) continue;
            if ((kk < 0 || kk >= Nz) && !(g.ndim == 3 && g.periodic[2])) continue;
            for (int32_t e = 0; e <= ru.hi - ru.lo; e++) d64 += p.w64[ru.woff + e];
            if (need_fp32)
                for (int32_t e = 0; e <= ru.hi - ru.lo; e++) d32 += p.w32d[ru.woff + e];
            taps += ru.hi - ru.lo + 1;
        }
        p.D64_row[r] = d64;
        p.D32_row[r] = d32;
        cost[r] = taps;
    }
    });
    // heaviest rows first: the hardware dispatches workgroups in order, so this is an LPT schedule
    p.row_order.resize(nrows);
    std::iota(p.row_order.begin(), p.row_order.end(), 0);
    std::stable_sort(p.row_order.begin(), p.row_order.end(), [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
    double tot = 0.0;
    p.max_taps = 0;
    p.active_taps = 0.0;
    for (int64_t r = 0; r < nrows; r++) {
        tot += (double)cost[r];
        p.max_taps = std::max<int64_t>(p.max_taps, cost[r]);
        int64_t nact = Nx;
        if (g.has_mask) {
            nact = 0;
            for (int64_t i = 0; i < Nx; i++) nact += g.mask_h[i + Nx * r] ? 1 : 0;
        }
        p.active_taps += (double)cost[r] * (double)nact;
    }
    p.mean_taps = nrows ? tot / (double)nrows : 0.0;
}

// Filtering.jl:1382-1441 _build_separable_footprint: per-axis 1-D Gaussian samples at offset·step
// (square truncation lim = ceil(rad/step)), ZeroFill profiles Σ valid weights per index, and for
// Deformable on a masked grid 1/(separable convolution of the mask).
void build_separable(Plan &p) {
    const Grid &g = *p.grid;
    if (g.ndim != 2 || g.geometry != CGEF_CARTESIAN || p.kind != CGEF_GAUSSIAN || !g.uniform[0] || !g.uniform[1])
        fail(CGEF_ERR_ARGUMENT, "the separable engine needs a Gaussian kernel on a uniform 2-D Cartesian grid");
    const int64_t N[2] = {g.n[0], g.n[1]};
    std::vector<double> gw[2], prof[2];
    for (int ax = 0; ax < 2; ax++) {
        const double d = std::fabs(g.coord[ax][1] - g.coord[ax][0]);
        const int64_t lim = (std::isfinite(d) && d > 0) ? iceil(p.rad / d) : 0;
        p.sep_lim[ax] = (int)lim;
        gw[ax].resize(2 * lim + 1);
        for (int64_t k = -lim; k <= lim; k++) gw[ax][k + lim] = kernel_weight(CGEF_GAUSSIAN, p.alpha, std::fabs((double)k * d), p.scale);
        prof[ax].assign(N[ax], 0.0);
        for (int64_t i = 0; i < N[ax]; i++) {
            double s = 0.0;
            for (int64_t dd = -lim; dd <= lim; dd++) {
                int64_t ii = i + dd;
                if ((ii >= 0 && ii < N[ax]) || g.periodic[ax]) s += gw[ax][dd + lim];
            }
            prof[ax][i] = s;
        }
        CGEF_HIP(hipMalloc(&p.sep_g_d[ax], gw[ax].size() * sizeof(double)));
        CGEF_HIP(hipMemcpy(p.sep_g_d[ax], gw[ax].data(), gw[ax].size() * sizeof(double), hipMemcpyHostToDevice));
        CGEF_HIP(hipMalloc(&p.sep_prof_d[ax], N[ax] * sizeof(double)));
        CGEF_HIP(hipMemcpy(p.sep_prof_d[ax], prof[ax].data(), N[ax] * sizeof(double), hipMemcpyHostToDevice));
    }
    if (g.has_mask && p.strategy == CGEF_DEFORMABLE) {
        const int64_t Nx = N[0], Ny = N[1];
        const int lx = p.sep_lim[0], ly = p.sep_lim[1];
        std::vector<double> rp(Nx * Ny, 0.0), den(Nx * Ny, 0.0);
        for (int64_t j = 0; j < Ny; j++)
            for (int64_t ddi = -lx; ddi <= lx; ddi++)
                for (int64_t i = 0; i < Nx; i++) {
                    int64_t ii = i + ddi;
                    if (ii < 0 || ii >= Nx) {
                        if (!g.periodic[0]) continue;
                        ii = ((ii % Nx) + Nx) % Nx;
                    }
                    rp[i + Nx * j] += gw[0][ddi + lx] * (g.mask_h[ii + Nx * j] ? 1.0 : 0.0);
                }
        for (int64_t j = 0; j < Ny; j++)
            for (int64_t ddj = -ly; ddj <= ly; ddj++) {
                int64_t jj = j + ddj;
                if (jj < 0 || jj >= Ny) {
                    if (!g.periodic[1]) continue;
                    jj = ((jj % Ny) + Ny) % Ny;
                }
                for (int64_t i = 0; i < Nx; i++) den[i + Nx * j] += gw[1][ddj + ly] * rp[i + Nx * jj];
            }
        for (auto &v : den) v = v > 1e-15 ? 1.0 / v : 0.0;
        CGEF_HIP(hipMalloc(&p.sep_invden_d, den.size() * sizeof(double)));
        CGEF_HIP(hipMemcpy(p.sep_invden_d, den.data(), den.size() * sizeof(double), hipMemcpyHostToDevice));
    }
}

}  // namespace cgef
