// Host side of the ring engine: factorization, twiddles, (dz, dy) offset table and the device-built
// Fourier weights Ŵ[m][offset][band].
#include <algorithm>
#include <cmath>
#include <map>
#include <string>

#include "cgef_ring.hpp"

namespace cgef {

bool ring_factor(int n, RingFft &F) {
    F = RingFft{};
    F.n = n;
    if (n < 2) return false;
    int m = n;
    const int order[] = {4, 2, 3, 5, 7, 11, 13};
    for (int r : order) {
        while (m % r == 0) {
            if (F.nstage == RING_MAX_STAGES) return false;
            F.radix[F.nstage++] = r;
            m /= r;
        }
    }
    return m == 1;
}

template <class T>
static T *up(const std::vector<T> &v) {
    T *d = nullptr;
    if (v.empty()) return d;
    CGEF_HIP(hipMalloc(&d, v.size() * sizeof(T)));
    CGEF_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

// Whether the ring engine can serve a grid: a periodic, uniform x axis whose length factors into the
// supported radices.
bool ring_grid_ok(const Grid &g) {
    if (!g.periodic[0] || !g.uniform[0] || g.n[0] < 2) return false;
    RingFft F;
    return ring_factor((int)g.n[0], F);
}

// ... and a plan: its runs must be symmetric modulo the ring (so Ŵ is real), the contraction's LDS
// slab (RING_CONTRACT_CC columns × (256 + dy range) complex values) must fit a workgroup, and Ŵ
// (mh × offsets × bands FP64, each dz group's dy range filled) must stay within the plan budget.
bool ring_supported(const Plan &p, std::string *why) {
    auto no = [&](const std::string &m) {
        if (why) *why = m;
        return false;
    };
    if (!ring_grid_ok(*p.grid))
        return no("the ring engine needs a periodic, uniform x axis whose length factors into 2,3,5,7,11,13");
    const Grid &g = *p.grid;
    std::map<int, std::pair<int, int>> dy_range;
    int dymin = 0, dymax = 0;
    for (const Run &r : p.runs) {
        if (!(r.lo == -r.hi || (int64_t)(r.hi - r.lo + 1) == g.n[0]))
            return no("the footprint's longitude runs are not symmetric about the target column");
        auto it = dy_range.find(r.dz);
        if (it == dy_range.end()) dy_range[r.dz] = {r.dy, r.dy};
        else it->second = {std::min(it->second.first, r.dy), std::max(it->second.second, r.dy)};
        dymin = std::min(dymin, r.dy);
        dymax = std::max(dymax, r.dy);
    }
    int64_t noff = 0;
    for (const auto &kv : dy_range) noff += kv.second.second - kv.second.first + 1;
    const size_t slab = (size_t)RING_CONTRACT_CC * (size_t)(256 + dymax - dymin) * sizeof(double2);
    if (slab > LDS_BYTES_PER_CU)
        return no("the footprint spans " + std::to_string(dymax - dymin + 1) +
                  " source rows: the ring contraction's LDS slab would exceed 160 KiB");
    const double what = (double)(g.n[0] / 2 + 1) * (double)noff * (double)p.nbands * sizeof(double);
    if (what > RING_WHAT_BUDGET_BYTES)
        return no("the ring engine's Fourier weights would take " + std::to_string(what / 1073741824.0) +
                  " GiB, above the 32 GiB per-plan budget");
    return true;
}

// Butterflies per thread a staged stage of radix R may hold (the radix cases compiled into the staged
// kernels, about 100 VGPRs together); 0 for radices the staged FFT does not evaluate (the primes 7,
// 11, 13 take the LDS ping-pong FFT).
static int staged_mb_max(int R) {
    switch (R) {
        case 2: return 6;
        case 3: return 4;
        case 4: return 3;
        case 5: case 6: case 8: return 2;
        case 9: case 10: case 12: case 15: case 16: return 1;
        default: return 0;
    }
}

static bool staged_dfs(int n, int m, int depth, int *rad, int k) {
    if (m == 1) return true;
    if (depth == 0) return false;
    static const int pref[] = {16, 12, 10, 9, 15, 8, 6, 5, 4, 3, 2};
    for (int R : pref) {
        if (m % R != 0 || (n / R + 255) / 256 > staged_mb_max(R)) continue;
        rad[k] = R;
        if (staged_dfs(n, m / R, depth - 1, rad, k + 1)) return true;
    }
    return false;
}

bool ring_staged_factor(RingFft &F) {
    const int n = F.n;
    F.snst = 0;
    F.lpsh = 31;
    if (n < 2 || n > 4096 || (size_t)(n + n / 16 + 1) * sizeof(double2) > 80 * 1024) return false;
    F.fixed = 0;
    // lengths the row kernels also carry at compile time (fewer registers than the radix switch):
    // 1440 as 12·12·10 rather than 9·16·10, 256 as 4·4·4·4 (the factorization found below anyway)
    if (n == 1440 || n == 256) {
        const int r1440[3] = {12, 12, 10}, r256[4] = {4, 4, 4, 4};
        F.snst = n == 1440 ? 3 : 4;
        for (int st = 0; st < F.snst; st++) F.srad[st] = n == 1440 ? r1440[st] : r256[st];
        F.lpsh = 4;  // even first radix: padded LDS
        F.sstr = n + n / 16 + 1;
        F.nsq = 1;
        F.fixed = n == 1440 ? RING_FIXED_1440 : RING_FIXED_256;
        return true;
    }
    int rad[RING_MAX_STAGES];
    for (int d = 1; d <= RING_MAX_STAGES; d++) {
        if (!staged_dfs(n, n, d, rad, 0)) continue;
        int cnt = 0, m = n;
        while (m > 1) m /= rad[cnt++];
        // odd radices first: the first stage's stride-R writes then fall on distinct banks
        std::stable_sort(rad, rad + cnt, [](int a, int b) { return (a & 1) > (b & 1); });
        F.snst = cnt;
        for (int st = 0; st < cnt; st++) F.srad[st] = rad[st];
        F.lpsh = (rad[0] & 1) ? 31 : 4;
        F.sstr = n + n / 16 + 1;
        F.nsq = 1;
        return true;
    }
    return false;
}

// Row pairs per block of the staged row transforms: doubled while the block's LDS stays within 32 KiB
// and every stage's butterflies (over all the block's sequences) fit the per-thread budget, up to 8
// pairs — short rows then keep all 256 threads busy (a 256-point row has 64 radix-4 butterflies).
static void ring_rows_nsq(RingFft &F) {
    F.nsq = 1;
    if (F.nseq == 0) return;
    for (int q = 2; q <= 8; q *= 2) {
        if ((size_t)q * F.sstr * sizeof(double2) > 32 * 1024) break;
        bool ok = true;
        for (int st = 0; st < F.snst; st++)
            ok = ok && (F.n / F.srad[st] * q + 255) / 256 <= staged_mb_max(F.srad[st]);
        if (!ok) break;
        F.nsq = q;
    }
}

// Twiddle tables of a factored length: exp(−2πik/N) and the per-stage table of the register-staged
// FFT (stage of radix R after Ns points: exp(−2πi·r·k/(Ns·R)) at twoff + k·(R−1) + r−1).
void fft_tables(RingFft &F, double2 *&tw_d, double2 *&tws_d) {
    const int n = F.n;
    std::vector<double2> tw(n);
    for (int k = 0; k < n; k++) {
        const double a = 2.0 * M_PI * (double)k / (double)n;
        tw[k] = make_double2(std::cos(a), -std::sin(a));
    }
    tw_d = up(tw);
    F.nseq = ring_staged_factor(F) ? 1 : 0;
    std::vector<double2> tws;
    int Ns = 1;
    for (int st = 0; st < F.snst; st++) {
        const int Rr = F.srad[st];
        F.twoff[st] = (int)tws.size();
        for (int k = 0; k < Ns; k++)
            for (int r = 1; r < Rr; r++) {
                const double a = 2.0 * M_PI * (double)r * (double)k / ((double)Ns * (double)Rr);
                tws.push_back(make_double2(std::cos(a), -std::sin(a)));
            }
        Ns *= Rr;
    }
    if (tws.empty()) tws.push_back(make_double2(1.0, 0.0));
    F.twsize = (int)tws.size();
    tws_d = up(tws);
}

// The spectral engine: row transforms of the ring engine along x, register-staged transforms along y.
bool spectral_supported(const Grid &g, int kind) {
    if (g.ndim != 2 || g.geometry != CGEF_CARTESIAN || (kind != CGEF_GAUSSIAN && kind != CGEF_TOPHAT)) return false;
    // the transform needs only the periods (k = 2π·m/L), not the coordinates (FFTWExt.jl:138-142)
    if (!g.periodic[0] || !g.periodic[1] || g.n[0] < 2 || g.n[1] < 2) return false;
    RingFft fx, fy;
    if (!ring_factor((int)g.n[0], fx) || !ring_factor((int)g.n[1], fy)) return false;
    // each transform runs in one workgroup's LDS: register-staged, else the ping-pong pair
    auto fits = [](const RingFft &F) {
        RingFft t = F;
        return ring_staged_factor(t) || (size_t)2 * F.n * sizeof(double2) <= 160 * 1024;
    };
    return fits(fx) && fits(fy);
}

void build_spectral(Plan &p, const GeomArgs &G, hipStream_t s) {
    Grid &g = *p.grid;
    if (!spectral_supported(g, p.kind))
        fail(CGEF_ERR_UNSUPPORTED,
             "the spectral engine needs a doubly periodic 2-D Cartesian grid whose sizes factor into "
             "2,3,5,7,11,13, at most 5120 points each");
    RingPlan &R = p.ring;
    ring_factor((int)g.n[0], R.fft);
    ring_factor((int)g.n[1], R.fy);
    fft_tables(R.fft, R.tw_d, R.tws_d);
    ring_rows_nsq(R.fft);
    fft_tables(R.fy, R.twy_d, R.twsy_d);
    R.mh = g.n[0] / 2 + 1;
    RingArgs &A = R.args;
    A.ny = g.n[1];
    A.nz = 1;
    A.nrows = g.n[1];
    A.mh = R.mh;
    A.py = 1;
    R.spectral = true;
    R.kind = p.kind;
    R.alpha = p.alpha;
    R.ell = p.scale;
    // the transform's wavenumbers 2π·m/L (FFTWExt.jl:138-146)
    R.dkx = 2.0 * M_PI / g.period[0];
    R.dky = 2.0 * M_PI / g.period[1];
    R.ok = true;
    if (g.has_mask && p.strategy == CGEF_DEFORMABLE) {
        // FFTWExt.jl:162-176: the mask's own filtered value, inverted where ≥ 0.01, built once
        struct Tmp {
            void *p = nullptr;
            ~Tmp() {
                if (p) (void)hipFree(p);
            }
        } xin, xf, xo;
        const size_t nxf = (size_t)R.mh * g.n[1];
        CGEF_HIP(hipMalloc(&xin.p, (size_t)g.npts * sizeof(double)));
        CGEF_HIP(hipMalloc(&xf.p, nxf * sizeof(double2)));
        CGEF_HIP(hipMalloc(&xo.p, nxf * sizeof(double2)));
        CGEF_HIP(hipMalloc(&R.invrenorm_d, (size_t)g.npts * sizeof(double)));
        ring_prep_fields(G, CGEF_F64, nullptr, 0, 1, (double *)xin.p, s);
        ring_forward(R, (const double *)xin.p, (double2 *)xf.p, 1, s);
        ring_spectral_cols(R, (const double2 *)xf.p, (double2 *)xo.p, 1, s);
        ring_inverse(R, (const double2 *)xo.p, (double *)xin.p, 1, s);
        ring_spectral_invert(G, (const double *)xin.p, R.invrenorm_d, s);
        CGEF_HIP(hipGetLastError());
        CGEF_HIP(hipStreamSynchronize(s));
    }
}

// Requires p.runs / p.run_ptr / p.w64 from build_banded (before any run splitting) and the device
// copies runs_d / w64_d / band_of_row_d.
void build_ring(Plan &p, hipStream_t s) {
    Grid &g = *p.grid;
    RingPlan &R = p.ring;
    if (!ring_factor((int)g.n[0], R.fft)) fail(CGEF_ERR_UNSUPPORTED, "row length does not factor into radices 2,3,5,7,11,13");
    const int n = (int)g.n[0];
    R.mh = n / 2 + 1;
    // offsets (dz, dy), sorted
    std::map<std::pair<int, int>, int> offs;
    for (const Run &r : p.runs) offs[{r.dz, r.dy}] = 0;
    // the contraction walks each dz group as one contiguous dy range: fill any gap with an offset
    // whose Ŵ stays zero
    std::map<int, std::pair<int, int>> dy_range;
    for (const auto &kv : offs) {
        auto it = dy_range.find(kv.first.first);
        if (it == dy_range.end()) dy_range[kv.first.first] = {kv.first.second, kv.first.second};
        else it->second = {std::min(it->second.first, kv.first.second), std::max(it->second.second, kv.first.second)};
    }
    for (const auto &kv : dy_range)
        for (int dy = kv.second.first; dy <= kv.second.second; dy++) offs[{kv.first, dy}] = 0;
    std::vector<int32_t> off_dy, dz_val, dz_off0;
    int idx = 0, dymin = 0, dymax = 0;
    int last_dz = INT32_MIN;
    for (auto &kv : offs) {
        kv.second = idx;
        const int dz = kv.first.first, dy = kv.first.second;
        if (dz != last_dz) {
            dz_val.push_back(dz);
            dz_off0.push_back(idx);
            last_dz = dz;
        }
        off_dy.push_back(dy);
        dymin = std::min(dymin, dy);
        dymax = std::max(dymax, dy);
        idx++;
    }
    dz_off0.push_back(idx);
    std::vector<int32_t> run_off(2 * p.runs.size());
    for (int b = 0; b < p.nbands; b++)
        for (int32_t q = p.run_ptr[b]; q < p.run_ptr[b + 1]; q++) {
            run_off[2 * q] = b;
            run_off[2 * q + 1] = offs[{p.runs[q].dz, p.runs[q].dy}];
        }
    fft_tables(R.fft, R.tw_d, R.tws_d);
    ring_rows_nsq(R.fft);
    R.off_dy_d = up(off_dy);
    R.dz_val_d = up(dz_val);
    R.dz_off0_d = up(dz_off0);
    int32_t *run_off_d = up(run_off);
    RingArgs &A = R.args;
    A.ny = g.n[1];
    A.nz = g.n[2];
    A.nrows = g.n[1] * g.n[2];
    A.mh = R.mh;
    A.py = g.periodic[1];
    A.pz = g.ndim == 3 ? g.periodic[2] : 0;
    A.noff = idx;
    A.nbands = p.nbands;
    A.ndz = (int32_t)dz_val.size();
    A.dy_min = dymin;
    A.dy_max = dymax;
    A.band_of_row = p.band_of_row_d;
    A.dz_val = R.dz_val_d;
    A.dz_off0 = R.dz_off0_d;
    A.off_dy = R.off_dy_d;
    CGEF_HIP(hipMalloc(&R.what_d, (size_t)R.mh * A.noff * A.nbands * sizeof(double)));
    A.what = R.what_d;
    ring_build_what(R, p.runs_d, run_off_d, p.w64_d, (int64_t)p.runs.size(), s);
    if (A.ndz == 1 && A.noff >= RING_MFMA_MIN_OFFSETS) {
        // the matrix-core contraction's lane-ordered copy of Ŵ (a wide 2-D footprint)
        R.wt_steps = (15 + A.noff + 3) / 4;
        R.wt_tiles = A.nbands == 1 ? 1 : (int32_t)((A.ny + 15) / 16);
        CGEF_HIP(hipMalloc(&R.wt_d, (size_t)R.mh * R.wt_tiles * R.wt_steps * 64 * sizeof(double)));
        ring_build_wtiles(R, s);
    }
    CGEF_HIP(hipGetLastError());
    CGEF_HIP(hipStreamSynchronize(s));
    (void)hipFree(run_off_d);
    ring_build_mode_cut(p, R, s);
    R.ok = true;
}

// The mode cut (RING_MODE_TOL, RING_MODE_SLOPE_TOL): per band the modes kept, from the device-built Ŵ;
// per row the band's value; and the contraction work lists over the (mode, row block) pairs some row
// keeps.
void ring_build_mode_cut(const Plan &p, RingPlan &R, hipStream_t s) {
    const RingArgs &A = R.args;
    const Grid &g = *p.grid;
    const std::vector<int32_t> &band_of_row = p.band_of_row;
    // x spacing per row (the ring's uniform step; on the sphere R·cosφ·Δλ), the smallest per band
    const double step = g.coord[0].size() > 1 ? std::fabs(g.coord[0][1] - g.coord[0][0]) : 1.0;
    std::vector<double> dxb(A.nbands, HUGE_VAL);
    for (int64_t r = 0; r < A.nrows; r++) {
        const int64_t j = r % A.ny;
        const double dx = g.geometry == CGEF_SPHERICAL ? g.R * std::cos(g.coord[1][j]) * step : step;
        double &d = dxb[band_of_row[r]];
        d = std::min(d, std::max(dx, 1e-300));
    }
    struct Tmp {
        void *p = nullptr;
        ~Tmp() {
            if (p) (void)hipFree(p);
        }
    } S, mb;
    CGEF_HIP(hipMalloc(&S.p, (size_t)A.mh * A.nbands * sizeof(double)));
    CGEF_HIP(hipMalloc(&mb.p, (size_t)A.nbands * sizeof(int32_t)));
    double *dxb_d = up(dxb);
    ring_mode_band_cut(R, p.scale, dxb_d, (double *)S.p, (int32_t *)mb.p, s);
    CGEF_HIP(hipStreamSynchronize(s));
    (void)hipFree(dxb_d);
    CGEF_HIP(hipGetLastError());
    std::vector<int32_t> mband(A.nbands);
    CGEF_HIP(hipMemcpyAsync(mband.data(), mb.p, mband.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CGEF_HIP(hipStreamSynchronize(s));
    const int64_t Ny = A.ny, nrows = A.nrows;
    std::vector<int32_t> mrow(nrows);
    int64_t kept = 0;
    for (int64_t r = 0; r < nrows; r++) {
        mrow[r] = mband[band_of_row[r]];
        kept += mrow[r];
    }
    R.modes_kept = (double)kept / ((double)nrows * (double)A.mh);
    if (kept == nrows * A.mh) return;   // nothing to drop
    // per row block (rows of one level), the most modes any of its rows keeps; the work list holds the
    // (m, block) pairs below that, in ascending m as the full grid orders them
    auto work_list = [&](int64_t rows_per_block, int64_t &count) {
        const int64_t nyb = (Ny + rows_per_block - 1) / rows_per_block, nblk = nyb * A.nz;
        if ((int64_t)A.mh * nblk >= ((int64_t)1 << 31)) fail(CGEF_ERR_UNSUPPORTED, "mode-cut work list exceeds 32 bits");
        std::vector<int32_t> mblk(nblk, 0), list;
        for (int64_t r = 0; r < nrows; r++) {
            const int64_t k = r / Ny, j = r - k * Ny;
            int32_t &x = mblk[k * nyb + j / rows_per_block];
            x = std::max(x, mrow[r]);
        }
        for (int64_t m = 0; m < A.mh; m++)
            for (int64_t q = 0; q < nblk; q++)
                if (m < mblk[q]) list.push_back((int32_t)(m * nblk + q));
        count = (int64_t)list.size();
        return up(list);
    };
    R.citems_d = work_list(256, R.ncitems);
    if (R.wt_d) {
        if (A.nz != 1) fail(CGEF_ERR_UNSUPPORTED, "matrix-core contraction on a 3-D plan");
        R.mitems_d = work_list(64, R.nmitems);
    }
    R.mcut_d = up(mrow);
    R.mcut_h = std::move(mrow);
}

void free_ring(RingPlan &R) {
    if (R.twy_d) (void)hipFree(R.twy_d);
    if (R.twsy_d) (void)hipFree(R.twsy_d);
    if (R.invrenorm_d) (void)hipFree(R.invrenorm_d);
    if (R.tw_d) (void)hipFree(R.tw_d);
    if (R.tws_d) (void)hipFree(R.tws_d);
    if (R.what_d) (void)hipFree(R.what_d);
    if (R.wt_d) (void)hipFree(R.wt_d);
    if (R.mcut_d) (void)hipFree(R.mcut_d);
    if (R.citems_d) (void)hipFree(R.citems_d);
    if (R.mitems_d) (void)hipFree(R.mitems_d);
    if (R.dz_val_d) (void)hipFree(R.dz_val_d);
    if (R.dz_off0_d) (void)hipFree(R.dz_off0_d);
    if (R.off_dy_d) (void)hipFree(R.off_dy_d);
    R = RingPlan{};
}

}  // namespace cgef
