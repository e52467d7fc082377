// Host side of the prefix-sum top-hat engine (Filtering.jl:751-1141): the per-(row, row offset)
// support description, the extended axis-1 coordinates and the mask-only prefix tables.
#include <algorithm>
#include <cmath>
#include <limits>

#include "cgef_internal.hpp"

namespace cgef {

template <class T>
static T *up(const std::vector<T> &v) {
    T *d = nullptr;
    if (v.empty()) return d;
    CGEF_HIP(hipMalloc(&d, v.size() * sizeof(T)));
    CGEF_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

// The reference builds this engine for a top-hat on any rectilinear 2-D grid (Filtering.jl:1115-1141);
// here axis 1 must also be stored ascending (the extended coordinates are then the axis itself).
bool prefix_supported(const Grid &g, int kind) {
    if (kind != CGEF_TOPHAT || g.ndim != 2) return false;
    const std::vector<double> &x = g.coord[0];
    for (size_t i = 1; i < x.size(); i++)
        if (!(x[i] > x[i - 1])) return false;
    return true;
}

// Half-width along axis 1 of the support between rows at y1 and y2 (Filtering.jl:764-766), or < 0
// when they are farther apart than rad: Cartesian √(rad² − Δy²); sphere
// acos((cos(rad/R) − sinφ₁sinφ₂)/(cosφ₁cosφ₂)), all longitudes at a pole row inside the cap.
static double metric_band(const Grid &g, double y1, double y2, double rad) {
    if (g.geometry != CGEF_SPHERICAL) {
        const double dy = y2 - y1;
        const double r2 = rad * rad - dy * dy;
        return r2 < 0.0 ? -1.0 : std::sqrt(r2);
    }
    const double c = std::cos(rad / g.R);
    const double cc = std::cos(y1) * std::cos(y2), ss = std::sin(y1) * std::sin(y2);
    if (cc <= 0.0) return (c <= ss) ? std::acos(-1.0) : -1.0;
    double a = (c - ss) / cc;
    if (a > 1.0) return -1.0;
    if (a < -1.0) a = -1.0;
    return std::acos(a);
}

void build_prefix(Plan &p) {
    const Grid &g = *p.grid;
    if (!prefix_supported(g, p.kind))
        fail(CGEF_ERR_UNSUPPORTED, "the prefix-sum engine needs a top-hat kernel on a 2-D grid with ascending x");
    const int64_t Nx = g.n[0], Ny = g.n[1];
    const std::vector<double> &x = g.coord[0], &y = g.coord[1];
    const double rad = p.rad;
    const bool px = g.periodic[0] && Nx > 1, py = g.periodic[1] && Ny > 1;
    const double xper = px ? g.period[0] : 0.0, yper = py ? g.period[1] : 0.0;
    const bool sph = g.geometry == CGEF_SPHERICAL;
    // measure factors: area(i, j) = wx[i]·wy[j]
    std::vector<double> wx(Nx), wy(Ny);
    for (int64_t i = 0; i < Nx; i++) wx[i] = g.width[0][i];
    for (int64_t j = 0; j < Ny; j++) wy[j] = (sph ? g.R * g.R * std::cos(y[j]) : 1.0) * g.width[1][j];
    // band bound from the axis-2 minimum gap (Filtering.jl:848-851)
    double min_dy = std::numeric_limits<double>::infinity();
    if (Ny >= 2) {
        if (g.uniform[1]) min_dy = std::fabs(y[1] - y[0]);
        else
            for (int64_t j = 1; j < Ny; j++) min_dy = std::min(min_dy, std::fabs(y[j] - y[j - 1]));
    }
    const double dyp = sph ? g.R * min_dy : min_dy;
    int64_t djl = (std::isfinite(dyp) && dyp > 0) ? (int64_t)std::ceil(rad / dyp) : 0;
    djl = std::min<int64_t>(djl, Ny - 1);
    const int nb = (int)std::min<int64_t>(2 * djl + 1, Ny);
    const int64_t ne = px ? 3 * Nx : Nx;
    std::vector<double> xe(ne);
    for (int64_t k = 0; k < ne; k++) xe[k] = x[k % Nx] + (px ? (double)(k / Nx - 1) * xper : 0.0);
    const bool uniform_fast = !px && g.uniform[0] && Nx > 1 && x[1] - x[0] > 0.0;
    const double step = Nx > 1 ? x[1] - x[0] : 0.0;
    std::vector<int32_t> jjt((size_t)Ny * nb, -1), wlt((size_t)Ny * nb, -1);
    std::vector<double> wyt((size_t)Ny * nb, 0.0), hwt((size_t)Ny * nb, 0.0);
    double taps = 0.0, active_taps = 0.0;
    int64_t max_taps = 0;
    for (int64_t j = 0; j < Ny; j++) {
        int64_t row_taps = 0;
        const double yt = y[j];
        for (int b = 0; b < nb; b++) {
            const int64_t jr = j - djl + b;
            if (!py && (jr < 0 || jr >= Ny)) continue;
            const int64_t jj = ((jr % Ny) + Ny) % Ny;
            double dy = y[jj] - yt;
            if (py && yper > 0.0) dy -= yper * std::round(dy / yper);
            const double hw = metric_band(g, yt, yt + dy, rad);
            if (hw < 0.0) continue;
            const size_t e = (size_t)j * nb + b;
            jjt[e] = (int32_t)jj;
            wyt[e] = wy[jj];
            hwt[e] = hw;
            if (px && 2.0 * hw >= xper) {
                wlt[e] = -2;
                row_taps += Nx;
            } else if (uniform_fast) {
                int64_t w = 0;
                while ((double)(w + 1) * step <= hw) w++;
                wlt[e] = (int32_t)std::min<int64_t>(w, Nx);
                row_taps += std::min<int64_t>(2 * w + 1, Nx);
            } else {
                wlt[e] = -1;
                const double h = Nx > 1 ? (x[Nx - 1] - x[0]) / (double)(Nx - 1) : 1.0;
                row_taps += std::min<int64_t>(2 * (int64_t)(hw / h) + 1, Nx);
            }
        }
        taps += (double)row_taps * Nx;
        max_taps = std::max(max_taps, row_taps);
        int64_t act = Nx;
        if (g.has_mask) {
            act = 0;
            for (int64_t i = 0; i < Nx; i++) act += g.mask_h[i + Nx * j] ? 1 : 0;
        }
        active_taps += (double)row_taps * act;
    }
    std::vector<double> pwx(Nx + 1, 0.0);
    for (int64_t i = 0; i < Nx; i++) pwx[i + 1] = pwx[i] + wx[i];
    const bool deform = p.strategy == CGEF_DEFORMABLE && g.has_mask;
    std::vector<double> pd;
    if (deform) {
        pd.assign((size_t)(Nx + 1) * Ny, 0.0);
        for (int64_t j = 0; j < Ny; j++) {
            double acc = 0.0;
            for (int64_t i = 0; i < Nx; i++) {
                acc += g.mask_h[i + Nx * j] ? wx[i] : 0.0;
                pd[(size_t)j * (Nx + 1) + i + 1] = acc;
            }
        }
    }
    PrefixPlan &P = p.pre;
    P.jj_d = up(jjt);
    P.wl_d = up(wlt);
    P.wy_d = up(wyt);
    P.hw_d = up(hwt);
    P.wx_d = up(wx);
    P.xe_d = up(xe);
    P.pwx_d = up(pwx);
    P.pd_d = up(pd);
    PrefixArgs &A = P.args;
    A.nx = Nx;
    A.ny = Ny;
    A.ne = ne;
    A.nb = nb;
    A.uniform_x = g.uniform[0] && Nx > 1 ? 1 : 0;
    A.deform = deform ? 1 : 0;
    A.x0 = xe[0];
    A.dx = step;
    A.jj = P.jj_d;
    A.wl = P.wl_d;
    A.wy = P.wy_d;
    A.hw = P.hw_d;
    A.wx = P.wx_d;
    A.xe = P.xe_d;
    A.pwx = P.pwx_d;
    A.pd = P.pd_d;
    P.ok = true;
    p.nbands = nb;
    p.nentries = (int64_t)Ny * nb;
    p.mean_taps = taps / (double)(Nx * Ny);
    p.max_taps = max_taps;
    p.active_taps = active_taps;
}

void free_prefix(PrefixPlan &P) {
    for (int32_t *q : {P.jj_d, P.wl_d})
        if (q) (void)hipFree(q);
    for (double *q : {P.wy_d, P.hw_d, P.wx_d, P.xe_d, P.pwx_d, P.pd_d})
        if (q) (void)hipFree(q);
    P = PrefixPlan{};
}

}  // namespace cgef
