/*
 * cgef_oracle.c — FP64 CPU restatement of CoarseGrainingEnergyFluxes.jl's real-space filter and
 * Π pipeline.
 *
 * TEST INFRASTRUCTURE ONLY: the checker and the CPU baseline ("port"), never the product. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Each function cites the reference file:line it restates. Where the arithmetic lives in the
 * un-vendored FlowGeometries.jl (distance, cell area, derivative stencils, vector rotations), the
 * choice made here is stated in a comment and in the repository's design notes ("parity
 * unpinned" items); the product kernels make the same choice, and the tests pin these functions
 * against closed forms, invariants and brute force instead.
 */
#define _DEFAULT_SOURCE /* j1 (Bessel function of the first kind) from <math.h> */
#include "cgef_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GAUSSIAN_TRUNCATION_TOL 1e-10 /* Kernels.jl:307 */

/* Planted mistakes, for showing that the pinning tests catch them (0 = none, the restatement):
 * OG_MUTATE_FORWARD_WIDTH — forward-difference widths on non-uniform axes; OG_MUTATE_RUNEND_2NODE —
 * two-node stencils at domain and mask-run ends instead of the shifted three-node window. */
static int g_mutation = 0;
void og_set_mutation(int m) { g_mutation = m; }


/* ---------------------------------------------------------------- kernels (Kernels.jl) ----- */

/* Kernels.jl:318-324 — unnormalized weights. */
double og_kernel_weight(int kind, double alpha, double d, double ell) {
    if (kind == OG_TOPHAT) return d <= ell / 2.0 ? 1.0 : 0.0;
    double r = d / ell;
    return exp(-alpha * (r * r));
}

/* Kernels.jl:354-359 — truncation radius. */
double og_kernel_radius(int kind, double alpha, double ell) {
    if (kind == OG_TOPHAT) return ell / 2.0;
    return ell * sqrt(-log(GAUSSIAN_TRUNCATION_TOL) / alpha);
}

/* Great-circle (haversine) distance: Filtering.jl:265 names the metric; the formula is
 * FlowGeometries' (un-vendored): d = 2R asin(sqrt(sin²(Δφ/2) + cosφ1 cosφ2 sin²(Δλ/2))). */
double og_haversine(double R, double lam1, double phi1, double lam2, double phi2) {
    double s1 = sin((phi2 - phi1) / 2.0);
    double s2 = sin((lam2 - lam1) / 2.0);
    double a = s1 * s1 + cos(phi1) * cos(phi2) * (s2 * s2);
    if (a > 1.0) a = 1.0;
    return 2.0 * R * asin(sqrt(a));
}

/* Cell width along one axis. A uniform (Range) axis has the constant step everywhere; a
 * non-uniform axis uses midpoint widths, half cells at non-periodic ends (FlowGeometries' measure
 * is un-vendored; the choice is pinned by tiling, by the step inside a collected uniform axis,
 * test_grids.jl:203-206, and by 2nd-order convergence of the weighted filter). */
double og_cell_width(const og_grid *g, int axis, int64_t i) {
    int64_t n = g->n[axis];
    const double *x = g->x[axis];
    if (n < 2) return 1.0;
    if (g->uniform[axis]) return fabs(x[1] - x[0]);
    if (g_mutation == OG_MUTATE_FORWARD_WIDTH) return i < n - 1 ? fabs(x[i + 1] - x[i]) : fabs(x[n - 1] - x[n - 2]);
    if (g->periodic[axis]) {
        double lo = (i == 0) ? x[n - 1] - g->period[axis] : x[i - 1];
        double hi = (i == n - 1) ? x[0] + g->period[axis] : x[i + 1];
        return fabs(hi - lo) / 2.0;
    }
    if (i == 0) return fabs(x[1] - x[0]) / 2.0;
    if (i == n - 1) return fabs(x[n - 1] - x[n - 2]) / 2.0;
    return fabs(x[i + 1] - x[i - 1]) / 2.0;
}

/* Cell measure: Cartesian product of widths; sphere R² cosφ Δλ Δφ (∝ cosφ on a uniform grid). */
double og_area(const og_grid *g, int64_t i, int64_t j, int64_t k) {
    if (g->spherical) {
        double phi = g->x[1][j];
        return g->R * g->R * cos(phi) * og_cell_width(g, 0, i) * og_cell_width(g, 1, j);
    }
    double a = og_cell_width(g, 0, i) * og_cell_width(g, 1, j);
    if (g->ndim == 3) a *= og_cell_width(g, 2, k);
    return a;
}

static inline int active_at(const og_grid *g, int64_t lin) {
    return g->mask == NULL ? 1 : (g->mask[lin] != 0);
}

static inline int64_t wrap(int64_t a, int64_t n) {
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

/* ------------------------------------------------ banded footprint (Filtering.jl:636-732) ---- */

typedef struct {
    int64_t cap, n;
    int32_t *di, *dj;
    double *w;
} vec_t;

static void vpush(vec_t *v, int32_t di, int32_t dj, double w) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 1024;
        v->di = (int32_t *)realloc(v->di, v->cap * sizeof(int32_t));
        v->dj = (int32_t *)realloc(v->dj, v->cap * sizeof(int32_t));
        v->w = (double *)realloc(v->w, v->cap * sizeof(double));
    }
    v->di[v->n] = di;
    v->dj[v->n] = dj;
    v->w[v->n] = w;
    v->n++;
}

og_footprint *og_footprint_build(const og_grid *g, int kind, double alpha, double ell) {
    return og_footprint_build_bands(g, kind, alpha, ell, NULL);
}

/* bandsel (spherical grids; NULL = all): bands it leaves unmarked are built empty, so a check on a
 * few latitude rows builds only their bands. A marked band's entries are those of the full build. */
og_footprint *og_footprint_build_bands(const og_grid *g, int kind, double alpha, double ell, const uint8_t *bandsel) {
    og_footprint *fp = (og_footprint *)calloc(1, sizeof(og_footprint));
    vec_t v = {0};
    double rad = og_kernel_radius(kind, alpha, ell);
    int64_t Nx = g->n[0], Ny = g->n[1];
    if (!g->spherical) {
        /* Filtering.jl:653-673: one translation-invariant band, disk d <= rad. */
        double dx = g->x[0][1] - g->x[0][0];
        double dy = g->x[1][1] - g->x[1][0];
        double A = og_area(g, 0, 0, 0);
        int64_t di_lim = dx > 0 ? (int64_t)ceil(rad / dx) : 0;
        int64_t dj_lim = dy > 0 ? (int64_t)ceil(rad / dy) : 0;
        for (int64_t ddj = -dj_lim; ddj <= dj_lim; ddj++)
            for (int64_t ddi = -di_lim; ddi <= di_lim; ddi++) {
                double a = (double)ddi * dx, b = (double)ddj * dy;
                double d = sqrt(a * a + b * b);
                if (d <= rad) vpush(&v, (int32_t)ddi, (int32_t)ddj, og_kernel_weight(kind, alpha, d, ell) * A);
            }
        fp->nbands = 1;
        fp->ptr = (int64_t *)malloc(2 * sizeof(int64_t));
        fp->ptr[0] = 0;
        fp->ptr[1] = v.n;
    } else {
        /* Filtering.jl:674-731: one band per latitude row. */
        double R = g->R;
        double dlam = g->x[0][1] - g->x[0][0];
        double dphi = g->x[1][1] - g->x[1][0];
        int64_t dj_lim = dphi > 0 ? (int64_t)ceil(rad / (R * dphi)) : 0;
        int64_t half = Nx / 2;
        int64_t *dil = (int64_t *)malloc(Ny * sizeof(int64_t));
        /* Longitude window per band: the reference's metric_window is a superset of the support
         * over every row the ball reaches, capped at one turn (Filtering.jl:688-692). Any superset
         * yields the same entries in the same order, so the exact support plus one is used. */
        for (int64_t j = 0; j < Ny; j++) {
            int64_t kmax = -1;
            if (bandsel && !bandsel[j]) {
                dil[j] = 0;
                continue;
            }
            for (int64_t ddj = -dj_lim; ddj <= dj_lim; ddj++) {
                int64_t jj = j + ddj;
                if (jj < 0 || jj >= Ny) continue;
                for (int64_t ddi = 0; ddi <= half; ddi++) {
                    double d = og_haversine(R, 0.0, g->x[1][j], (double)ddi * dlam, g->x[1][jj]);
                    if (d > rad) break;
                    if (ddi > kmax) kmax = ddi;
                }
            }
            int64_t lim = kmax + 1;
            dil[j] = lim < half ? lim : half;
        }
        fp->nbands = Ny;
        fp->ptr = (int64_t *)malloc((Ny + 1) * sizeof(int64_t));
        fp->ptr[0] = 0;
        for (int64_t j = 0; j < Ny; j++) {
            double phi = g->x[1][j];
            int64_t di_lim = dil[j];
            if (bandsel && !bandsel[j]) {
                fp->ptr[j + 1] = v.n;
                continue;
            }
            for (int64_t ddj = -dj_lim; ddj <= dj_lim; ddj++) {
                int64_t jj = j + ddj;
                if (jj < 0 || jj >= Ny) continue;
                double phi2 = g->x[1][jj];
                double A = og_area(g, 0, jj, 0);
                /* Filtering.jl:713: each ring column visited once. */
                int64_t hi_i = (2 * di_lim + 1 > Nx) ? di_lim - 1 : di_lim;
                for (int64_t ddi = -di_lim; ddi <= hi_i; ddi++) {
                    double d = og_haversine(R, 0.0, phi, (double)ddi * dlam, phi2);
                    if (d <= rad) vpush(&v, (int32_t)ddi, (int32_t)ddj, og_kernel_weight(kind, alpha, d, ell) * A);
                }
            }
            fp->ptr[j + 1] = v.n;
        }
        free(dil);
    }
    fp->nentries = v.n;
    fp->di = v.di;
    fp->dj = v.dj;
    fp->w = v.w;
    return fp;
}

void og_footprint_free(og_footprint *fp) {
    if (!fp) return;
    free(fp->di);
    free(fp->dj);
    free(fp->w);
    free(fp->ptr);
    free(fp);
}
int64_t og_footprint_nentries(const og_footprint *fp) { return fp->nentries; }
int64_t og_footprint_nbands(const og_footprint *fp) { return fp->nbands; }
void og_footprint_copy(const og_footprint *fp, int32_t *di, int32_t *dj, double *w, int64_t *ptr) {
    memcpy(di, fp->di, fp->nentries * sizeof(int32_t));
    memcpy(dj, fp->dj, fp->nentries * sizeof(int32_t));
    memcpy(w, fp->w, fp->nentries * sizeof(double));
    memcpy(ptr, fp->ptr, (fp->nbands + 1) * sizeof(int64_t));
}

/* Filtering.jl:2536-2598 (apply_footprint_row_batch!) and 1593-1634 (_footprint_point): one
 * neighbour walk per point shared by K fields, k ascending; ZeroFill counts inactive neighbours
 * in the denominator only, Deformable drops them; ws/wn if wn > 1e-15, inactive targets 0.
 * Rows in parallel with a dynamic schedule, as OhMyThreadsExt.jl:20-43, 97-123. */
static void apply_banded_rows(const og_grid *g, const og_footprint *fp, int strategy, int K,
                              const double *const *in, double *const *out, const uint8_t *rowsel, int nthreads) {
    int64_t Nx = g->n[0], Ny = g->n[1];
    int px = g->periodic[0], py = g->periodic[1];
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
#endif
    for (int64_t j = 0; j < Ny; j++) {
        if (rowsel && !rowsel[j]) continue;
        double ws[16];
        int64_t b = fp->nbands == 1 ? 0 : j;
        int64_t lo = fp->ptr[b], hi = fp->ptr[b + 1];
        for (int64_t i = 0; i < Nx; i++) {
            int64_t t = i + Nx * j;
            if (!active_at(g, t)) {
                for (int m = 0; m < K; m++) out[m][t] = 0.0;
                continue;
            }
            double wn = 0.0;
            for (int m = 0; m < K; m++) ws[m] = 0.0;
            for (int64_t k = lo; k < hi; k++) {
                int64_t jj = j + fp->dj[k];
                if (jj < 0 || jj >= Ny) {
                    if (!py) continue;
                    jj = wrap(jj, Ny);
                }
                int64_t ii = i + fp->di[k];
                if (ii < 0 || ii >= Nx) {
                    if (!px) continue;
                    ii = wrap(ii, Nx);
                }
                int64_t s = ii + Nx * jj;
                int act = active_at(g, s);
                double w = fp->w[k];
                if (strategy == OG_ZEROFILL) {
                    wn += w;
                    if (act)
                        for (int m = 0; m < K; m++) ws[m] += w * in[m][s];
                } else if (act) {
                    wn += w;
                    for (int m = 0; m < K; m++) ws[m] += w * in[m][s];
                }
            }
            for (int m = 0; m < K; m++) out[m][t] = wn > 1e-15 ? ws[m] / wn : 0.0;
        }
    }
}

void og_apply_banded(const og_grid *g, const og_footprint *fp, int strategy, int K,
                     const double *const *in, double *const *out, int nthreads) {
    apply_banded_rows(g, fp, strategy, K, in, out, NULL, nthreads);
}

/* ------------------------------------- separable Gaussian (Filtering.jl:1143-1561) ---------- */

static void sep_conv(const double *src, double *dst, double *rp, int64_t Nx, int64_t Ny,
                     const double *gx, int64_t lx, const double *gy, int64_t ly, int px, int py) {
    /* Row pass, taps outermost (Filtering.jl:1210-1254): per output point the order is ddi asc. */
    for (int64_t j = 0; j < Ny; j++) {
        double *r = rp + Nx * j;
        const double *s = src + Nx * j;
        for (int64_t i = 0; i < Nx; i++) r[i] = 0.0;
        for (int64_t ddi = -lx; ddi <= lx; ddi++) {
            double g = gx[ddi + lx];
            for (int64_t i = 0; i < Nx; i++) {
                int64_t ii = i + ddi;
                if (ii < 0 || ii >= Nx) {
                    if (!px) continue;
                    ii = wrap(ii, Nx);
                }
                r[i] += g * s[ii];
            }
        }
    }
    /* Column pass (Filtering.jl:1256-1279). */
    for (int64_t j = 0; j < Ny; j++) {
        double *d = dst + Nx * j;
        for (int64_t i = 0; i < Nx; i++) d[i] = 0.0;
        for (int64_t ddj = -ly; ddj <= ly; ddj++) {
            int64_t jj = j + ddj;
            if (jj < 0 || jj >= Ny) {
                if (!py) continue;
                jj = wrap(jj, Ny);
            }
            double wt = gy[ddj + ly];
            const double *r = rp + Nx * jj;
            for (int64_t i = 0; i < Nx; i++) d[i] += wt * r[i];
        }
    }
}

static void sep_profile(const double *g, int64_t lim, int64_t N, int per, double *prof) {
    /* Filtering.jl:1448-1460 */
    for (int64_t i = 0; i < N; i++) {
        double s = 0.0;
        for (int64_t dd = -lim; dd <= lim; dd++) {
            int64_t ii = i + dd;
            int valid = (ii >= 0 && ii < N) || per;
            if (valid) s += g[dd + lim];
        }
        prof[i] = s;
    }
}

void og_apply_separable(const og_grid *g, double alpha, double ell, int strategy, int K,
                        const double *const *in, double *const *out) {
    int64_t Nx = g->n[0], Ny = g->n[1], N = Nx * Ny;
    double rad = og_kernel_radius(OG_GAUSSIAN, alpha, ell);
    double dx = fabs(g->x[0][1] - g->x[0][0]), dy = fabs(g->x[1][1] - g->x[1][0]);
    int64_t lx = dx > 0 ? (int64_t)ceil(rad / dx) : 0, ly = dy > 0 ? (int64_t)ceil(rad / dy) : 0;
    double *gx = (double *)malloc((2 * lx + 1) * sizeof(double));
    double *gy = (double *)malloc((2 * ly + 1) * sizeof(double));
    /* Filtering.jl:1364-1371 — point samples of the 1-D profile (uniform axis, no measure). */
    for (int64_t k = -lx; k <= lx; k++) gx[k + lx] = og_kernel_weight(OG_GAUSSIAN, alpha, fabs((double)k * dx), ell);
    for (int64_t k = -ly; k <= ly; k++) gy[k + ly] = og_kernel_weight(OG_GAUSSIAN, alpha, fabs((double)k * dy), ell);
    int fully_active = 1;
    if (g->mask)
        for (int64_t t = 0; t < N; t++)
            if (!g->mask[t]) { fully_active = 0; break; }
    double *rp = (double *)malloc(N * sizeof(double));
    double *mi = (double *)malloc(N * sizeof(double));
    double *inv = NULL, *px_prof = NULL, *py_prof = NULL;
    if (fully_active || strategy == OG_ZEROFILL) {
        px_prof = (double *)malloc(Nx * sizeof(double));
        py_prof = (double *)malloc(Ny * sizeof(double));
        sep_profile(gx, lx, Nx, g->periodic[0], px_prof);
        sep_profile(gy, ly, Ny, g->periodic[1], py_prof);
    } else {
        inv = (double *)malloc(N * sizeof(double));
        for (int64_t t = 0; t < N; t++) mi[t] = active_at(g, t) ? 1.0 : 0.0;
        sep_conv(mi, inv, rp, Nx, Ny, gx, lx, gy, ly, g->periodic[0], g->periodic[1]);
        for (int64_t t = 0; t < N; t++) inv[t] = inv[t] > 1e-15 ? 1.0 / inv[t] : 0.0;
    }
    for (int m = 0; m < K; m++) {
        /* Filtering.jl:1471-1481 */
        for (int64_t t = 0; t < N; t++) mi[t] = (active_at(g, t) ? 1.0 : 0.0) * in[m][t];
        sep_conv(mi, out[m], rp, Nx, Ny, gx, lx, gy, ly, g->periodic[0], g->periodic[1]);
        /* Filtering.jl:1510-1528 */
        if (inv) {
            for (int64_t t = 0; t < N; t++) out[m][t] *= inv[t];
        } else {
            for (int64_t j = 0; j < Ny; j++)
                for (int64_t i = 0; i < Nx; i++) {
                    double den = px_prof[i] * py_prof[j];
                    int64_t t = i + Nx * j;
                    out[m][t] = den > 1e-15 ? out[m][t] / den : 0.0;
                }
        }
        for (int64_t t = 0; t < N; t++)
            if (!active_at(g, t)) out[m][t] = 0.0;
    }
    free(gx); free(gy); free(rp); free(mi); free(inv); free(px_prof); free(py_prof);
}

/* ------------------------- spectral filter (FFTW extension, FFTWExt.jl:121-208) ------------ */

/* Kernels.jl:405-406 (Gaussian) and the SpecialFunctions extension's disk transform for the top-hat:
 * Ĝ(k) = exp(−k²ℓ²/(4α)); 2·J₁(kR)/(kR) with R = ℓ/2; Ĝ(0) = 1. */
double og_spectral_transfer(int kind, double alpha, double k, double ell) {
    if (kind == OG_GAUSSIAN) return exp(-k * k * ell * ell / (4.0 * alpha));
    double x = k * ell / 2.0;
    return x == 0.0 ? 1.0 : 2.0 * j1(x) / x;
}

/* Periodic uniform 2-D Cartesian grid: out = irfft2(rfft2(mask·f) · Ĝ(|k|)), with |k| from the
 * transform's wavenumbers 2π·m/L (FFTWExt.jl:138-146); Deformable multiplies by 1/filter(mask) where
 * |filter(mask)| ≥ 0.01, else 0 (FFTWExt.jl:162-176). No masking of the output. Direct DFTs along
 * each axis (an independent restatement of the transform, O(N·(Nx + Ny)) per field). */
static void spectral_one(const og_grid *g, const double *G, const double *cx, const double *sx, const double *cy,
                         const double *sy, const double *a, double *o) {
    const int64_t Nx = g->n[0], Ny = g->n[1], mh = Nx / 2 + 1;
    double *Xr = (double *)malloc(mh * Ny * sizeof(double)), *Xi = (double *)malloc(mh * Ny * sizeof(double));
    double *Yr = (double *)malloc(Ny * sizeof(double)), *Yi = (double *)malloc(Ny * sizeof(double));
    for (int64_t j = 0; j < Ny; j++)
        for (int64_t m = 0; m < mh; m++) {
            double re = 0.0, im = 0.0;
            for (int64_t i = 0; i < Nx; i++) {
                int64_t q = (m * i) % Nx;
                re += a[i + Nx * j] * cx[q];
                im -= a[i + Nx * j] * sx[q];
            }
            Xr[m + mh * j] = re;
            Xi[m + mh * j] = im;
        }
    for (int64_t m = 0; m < mh; m++) {
        for (int64_t n = 0; n < Ny; n++) {
            double re = 0.0, im = 0.0;
            for (int64_t j = 0; j < Ny; j++) {
                int64_t q = (n * j) % Ny;
                double xr = Xr[m + mh * j], xi = Xi[m + mh * j];
                re += xr * cy[q] + xi * sy[q];
                im += xi * cy[q] - xr * sy[q];
            }
            Yr[n] = re * G[m + mh * n];
            Yi[n] = im * G[m + mh * n];
        }
        for (int64_t j = 0; j < Ny; j++) {
            double re = 0.0, im = 0.0;
            for (int64_t n = 0; n < Ny; n++) {
                int64_t q = (n * j) % Ny;
                re += Yr[n] * cy[q] - Yi[n] * sy[q];
                im += Yi[n] * cy[q] + Yr[n] * sy[q];
            }
            Xr[m + mh * j] = re / (double)Ny;
            Xi[m + mh * j] = im / (double)Ny;
        }
    }
    for (int64_t j = 0; j < Ny; j++)
        for (int64_t i = 0; i < Nx; i++) {
            double acc = Xr[mh * j]; /* m = 0: real part only */
            for (int64_t m = 1; m < mh; m++) {
                int64_t q = (m * i) % Nx;
                double term = Xr[m + mh * j] * cx[q] - Xi[m + mh * j] * sx[q];
                acc += (Nx % 2 == 0 && m == Nx / 2) ? term : 2.0 * term;
            }
            o[i + Nx * j] = acc / (double)Nx;
        }
    free(Xr); free(Xi); free(Yr); free(Yi);
}

void og_apply_spectral(const og_grid *g, int kind, double alpha, double ell, int strategy, int K,
                       const double *const *in, double *const *out) {
    const int64_t Nx = g->n[0], Ny = g->n[1], mh = Nx / 2 + 1, N = Nx * Ny;
    const double tau = 2.0 * acos(-1.0);
    double *cx = (double *)malloc(Nx * sizeof(double)), *sx = (double *)malloc(Nx * sizeof(double));
    double *cy = (double *)malloc(Ny * sizeof(double)), *sy = (double *)malloc(Ny * sizeof(double));
    for (int64_t q = 0; q < Nx; q++) { cx[q] = cos(tau * q / Nx); sx[q] = sin(tau * q / Nx); }
    for (int64_t q = 0; q < Ny; q++) { cy[q] = cos(tau * q / Ny); sy[q] = sin(tau * q / Ny); }
    const double dkx = tau / g->period[0], dky = tau / g->period[1];
    double *G = (double *)malloc(mh * Ny * sizeof(double));
    for (int64_t n = 0; n < Ny; n++) {
        int64_t nn = n <= (Ny - 1) / 2 ? n : n - Ny;
        for (int64_t m = 0; m < mh; m++) {
            double kx = dkx * (double)m, ky = dky * (double)nn;
            G[m + mh * n] = og_spectral_transfer(kind, alpha, sqrt(kx * kx + ky * ky), ell);
        }
    }
    int masked = 0;
    if (g->mask)
        for (int64_t t = 0; t < N; t++) if (!g->mask[t]) { masked = 1; break; }
    double *a = (double *)malloc(N * sizeof(double)), *ir = NULL;
    if (masked && strategy == OG_DEFORMABLE) {
        ir = (double *)malloc(N * sizeof(double));
        for (int64_t t = 0; t < N; t++) a[t] = active_at(g, t) ? 1.0 : 0.0;
        spectral_one(g, G, cx, sx, cy, sy, a, ir);
        for (int64_t t = 0; t < N; t++) ir[t] = fabs(ir[t]) >= 0.01 ? 1.0 / ir[t] : 0.0;
    }
    for (int q = 0; q < K; q++) {
        for (int64_t t = 0; t < N; t++) a[t] = masked ? (active_at(g, t) ? in[q][t] : 0.0) : in[q][t];
        spectral_one(g, G, cx, sx, cy, sy, a, out[q]);
        if (ir)
            for (int64_t t = 0; t < N; t++) out[q][t] *= ir[t];
    }
    free(cx); free(sx); free(cy); free(sy); free(G); free(a); free(ir);
}

/* ------------------------------ 3-D ball (Filtering.jl:2172-2203, 2268-2458) ---------------- */

static double min_gap(const og_grid *g, int axis) {
    int64_t n = g->n[axis];
    const double *x = g->x[axis];
    if (n < 2) return INFINITY;
    if (g->uniform[axis]) return fabs(x[1] - x[0]);
    double m = INFINITY;
    for (int64_t i = 1; i < n; i++) {
        double d = fabs(x[i] - x[i - 1]);
        if (d < m) m = d;
    }
    return m;
}

/* ----------------------------- prefix-sum top-hat (Filtering.jl:751-1097) --------------------- */

/* Half-width along axis 1 of the top-hat support between a target on row coordinate y1 and a source
 * row at y2 = y1 + dy, or < 0 when the rows are farther apart than rad (Filtering.jl:764-766:
 * Cartesian |Δx| ≤ √(rad² − Δy²); sphere |Δλ| ≤ acos((cos(rad/R) − sinφ₁ sinφ₂)/(cosφ₁ cosφ₂))). This
 * is FlowGeometries' metric_band, restated from those formulas (parity unpinned beyond them). */
double og_metric_band(const og_grid *g, double y1, double y2, double rad) {
    if (!g->spherical) {
        double dy = y2 - y1;
        double r2 = rad * rad - dy * dy;
        return r2 < 0.0 ? -1.0 : sqrt(r2);
    }
    double c = cos(rad / g->R);
    double cc = cos(y1) * cos(y2), ss = sin(y1) * sin(y2);
    if (cc <= 0.0) return (c <= ss) ? acos(-1.0) : -1.0;   /* a pole row: all longitudes or none */
    double a = (c - ss) / cc;
    if (a > 1.0) return -1.0;
    if (a < -1.0) a = -1.0;
    return acos(a);
}

/* Filtering.jl:833-906 (_build_prefixsum_tophat) and 913-1097 (fill + apply): per-row prefix sums of
 * mask·f·wx along axis 1 (tripled when periodic), then for each target and each row in the band an
 * interval sum found by comparing coordinates (two monotone pointers), or the constant ±w window
 * on a uniform non-periodic axis, or the whole row when the support spans a full period. */
void og_apply_prefix(const og_grid *g, double ell, int strategy, int K, const double *const *in,
                     double *const *out, int nthreads) {
    const int64_t Nx = g->n[0], Ny = g->n[1];
    const double *x = g->x[0], *y = g->x[1];
    const double rad = ell / 2.0; /* Kernels.jl:354 */
    const int px = g->periodic[0] && Nx > 1, py = g->periodic[1] && Ny > 1;
    const double xper = px ? g->period[0] : 0.0, yper = py ? g->period[1] : 0.0;
    double *wx = (double *)malloc(Nx * sizeof(double)), *wy = (double *)malloc(Ny * sizeof(double));
    for (int64_t i = 0; i < Nx; i++) wx[i] = og_cell_width(g, 0, i);
    for (int64_t j = 0; j < Ny; j++)
        wy[j] = (g->spherical ? g->R * g->R * cos(y[j]) : 1.0) * og_cell_width(g, 1, j);
    double min_dy = min_gap(g, 1);
    double dyp = g->spherical ? g->R * min_dy : min_dy;
    int64_t djl = (isfinite(dyp) && dyp > 0) ? (int64_t)ceil(rad / dyp) : 0;
    if (djl > Ny - 1) djl = Ny - 1;
    const int nrep = px ? 3 : 1;
    const int64_t ne = nrep * Nx;
    /* ascending order of axis 1 with the real index of each extended entry */
    int64_t *ord = (int64_t *)malloc(Nx * sizeof(int64_t));
    for (int64_t i = 0; i < Nx; i++) ord[i] = i;
    for (int64_t a = 1; a < Nx; a++) { /* insertion sort by coordinate (stable) */
        int64_t t = ord[a], b = a - 1;
        while (b >= 0 && x[ord[b]] > x[t]) { ord[b + 1] = ord[b]; b--; }
        ord[b + 1] = t;
    }
    double *xe = (double *)malloc(ne * sizeof(double));
    int64_t *src = (int64_t *)malloc(ne * sizeof(int64_t));
    for (int r = 0; r < nrep; r++)
        for (int64_t t = 0; t < Nx; t++) {
            xe[r * Nx + t] = x[ord[t]] + (nrep == 3 ? (r - 1) * xper : 0.0);
            src[r * Nx + t] = ord[t];
        }
    double *Pwx = (double *)malloc((ne + 1) * sizeof(double));
    Pwx[0] = 0.0;
    for (int64_t k = 0; k < ne; k++) Pwx[k + 1] = Pwx[k] + wx[src[k]];
    int masked = 0;
    if (g->mask)
        for (int64_t t = 0; t < Nx * Ny; t++) if (!g->mask[t]) { masked = 1; break; }
    const int use_den = strategy == OG_DEFORMABLE && masked;
    double *Pd = NULL;
    if (use_den) {
        Pd = (double *)malloc((ne + 1) * Ny * sizeof(double));
        for (int64_t j = 0; j < Ny; j++) {
            double acc = 0.0;
            Pd[j * (ne + 1)] = 0.0;
            for (int64_t k = 0; k < ne; k++) {
                int64_t i = src[k];
                acc += active_at(g, i + Nx * j) ? wx[i] : 0.0;
                Pd[j * (ne + 1) + k + 1] = acc;
            }
        }
    }
    const int uniform_fast = !px && ne == Nx && g->uniform[0] && Nx > 1 && x[1] - x[0] > 0.0;
    const double step = Nx > 1 ? x[1] - x[0] : 0.0;
    double *Pn = (double *)malloc((ne + 1) * Ny * sizeof(double));
    double *den = (double *)malloc(Nx * Ny * sizeof(double));
    for (int q = 0; q < K; q++) {
        const double *f = in[q];
        double *o = out[q];
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        for (int64_t j = 0; j < Ny; j++) {
            double acc = 0.0;
            Pn[j * (ne + 1)] = 0.0;
            for (int64_t k = 0; k < ne; k++) {
                int64_t i = src[k];
                acc += active_at(g, i + Nx * j) ? f[i + Nx * j] * wx[i] : 0.0;
                Pn[j * (ne + 1) + k + 1] = acc;
            }
        }
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        for (int64_t j = 0; j < Ny; j++) {
            double *orow = o + Nx * j, *drow = den + Nx * j;
            for (int64_t i = 0; i < Nx; i++) orow[i] = drow[i] = 0.0;
            const double yt = y[j];
            int64_t nb = 2 * djl + 1 < Ny ? 2 * djl + 1 : Ny;
            for (int64_t b = 0; b < nb; b++) {
                int64_t jr = j - djl + b;
                if (!py && (jr < 0 || jr >= Ny)) continue;
                int64_t jj = wrap(jr, Ny);
                double dy = y[jj] - yt;
                if (py && yper > 0.0) dy -= yper * round(dy / yper);
                double hw = og_metric_band(g, yt, yt + dy, rad);
                if (hw < 0.0) continue;
                const double wyj = wy[jj];
                const double *P = Pn + jj * (ne + 1);
                const double *D = use_den ? Pd + jj * (ne + 1) : Pwx;
                if (px && ne > Nx && 2.0 * hw >= xper) {
                    double na = P[Nx] - P[0], da = D[Nx] - D[0];
                    for (int64_t i = 0; i < Nx; i++) { orow[i] += wyj * na; drow[i] += wyj * da; }
                } else if (uniform_fast) {
                    int64_t w = 0;
                    while ((double)(w + 1) * step <= hw) w++;
                    for (int64_t i = 0; i < Nx; i++) {
                        int64_t lo = i - w < 0 ? 0 : i - w, hi = i + w > Nx - 1 ? Nx - 1 : i + w;
                        orow[i] += wyj * (P[hi + 1] - P[lo]);
                        drow[i] += wyj * (D[hi + 1] - D[lo]);
                    }
                } else {
                    int64_t lo = 0, hi = -1;
                    for (int64_t t = 0; t < Nx; t++) {
                        int64_t i = src[t];
                        double xlo = x[i] - hw, xhi = x[i] + hw;
                        while (lo < ne && xe[lo] < xlo) lo++;
                        while (hi < ne - 1 && xe[hi + 1] <= xhi) hi++;
                        if (hi >= lo) {
                            orow[i] += wyj * (P[hi + 1] - P[lo]);
                            drow[i] += wyj * (D[hi + 1] - D[lo]);
                        }
                    }
                }
            }
            for (int64_t i = 0; i < Nx; i++)
                orow[i] = (active_at(g, i + Nx * j) && drow[i] > 1e-15) ? orow[i] / drow[i] : 0.0;
        }
    }
    free(wx); free(wy); free(ord); free(xe); free(src); free(Pwx); free(Pd); free(Pn); free(den);
}


typedef struct {
    int32_t o0, o1, kk;
    double w;
} nd_entry;

/* Per target level the candidate list is translation-invariant in (i, j) when x and y are
 * uniform, so it is built once per level in the reference's enumeration order (offset axis 0
 * fastest, CartesianIndices order; Filtering.jl:2346-2360) and walked per point. Distances
 * follow the scattered engine (coordinate difference with a one-period shift, Filtering.jl:
 * 2352-2357); the uniform engine (all axes Ranges, Filtering.jl:2172-2203) uses offset·step,
 * which is the same number for exactly representable coordinates, and a constant measure. */
void og_apply_nd3(const og_grid *g, int kind, double alpha, double ell, int strategy, int K,
                  const double *const *in, double *const *out, int nthreads) {
    int64_t Nx = g->n[0], Ny = g->n[1], Nz = g->n[2];
    double rad = og_kernel_radius(kind, alpha, ell);
    int64_t lim[3];
    for (int d = 0; d < 3; d++) {
        double gap = min_gap(g, d);
        lim[d] = (isfinite(gap) && gap > 0) ? (int64_t)ceil(rad / gap) : 0;
    }
    int all_uniform = g->uniform[0] && g->uniform[1] && g->uniform[2];
    double dx = g->x[0][1] - g->x[0][0], dy = g->x[1][1] - g->x[1][0];
    double A = og_area(g, 0, 0, 0);
    int64_t cap = (2 * lim[0] + 1) * (2 * lim[1] + 1) * (2 * lim[2] + 1);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        nd_entry *lst = (nd_entry *)malloc(cap * sizeof(nd_entry));
        double ws[16];
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
        for (int64_t k = 0; k < Nz; k++) {
            int64_t n = 0;
            double zk = g->x[2][k];
            for (int64_t o2 = -lim[2]; o2 <= lim[2]; o2++) {
                int64_t kk = k + o2;
                double sh = 0.0;
                if (kk < 0 || kk >= Nz) {
                    if (!g->periodic[2]) continue;
                    sh = kk < 0 ? -g->period[2] : g->period[2];
                    kk = wrap(kk, Nz);
                }
                double ddz = all_uniform ? (double)o2 * (g->x[2][1] - g->x[2][0]) : (g->x[2][kk] + sh) - zk;
                for (int64_t o1 = -lim[1]; o1 <= lim[1]; o1++)
                    for (int64_t o0 = -lim[0]; o0 <= lim[0]; o0++) {
                        double a = (double)o0 * dx, b = (double)o1 * dy;
                        double d = sqrt(a * a + b * b + ddz * ddz);
                        if (d > rad) continue;
                        double meas = all_uniform ? A : og_area(g, 0, 0, kk);
                        lst[n].o0 = (int32_t)o0;
                        lst[n].o1 = (int32_t)o1;
                        lst[n].kk = (int32_t)kk;
                        lst[n].w = og_kernel_weight(kind, alpha, d, ell) * meas;
                        n++;
                    }
            }
            for (int64_t j = 0; j < Ny; j++)
                for (int64_t i = 0; i < Nx; i++) {
                    int64_t t = i + Nx * (j + Ny * k);
                    if (!active_at(g, t)) {
                        for (int m = 0; m < K; m++) out[m][t] = 0.0;
                        continue;
                    }
                    double wn = 0.0;
                    for (int m = 0; m < K; m++) ws[m] = 0.0;
                    for (int64_t e = 0; e < n; e++) {
                        int64_t ii = i + lst[e].o0, jj = j + lst[e].o1;
                        if (ii < 0 || ii >= Nx) {
                            if (!g->periodic[0]) continue;
                            ii = wrap(ii, Nx);
                        }
                        if (jj < 0 || jj >= Ny) {
                            if (!g->periodic[1]) continue;
                            jj = wrap(jj, Ny);
                        }
                        int64_t s = ii + Nx * (jj + Ny * (int64_t)lst[e].kk);
                        int act = active_at(g, s);
                        double w = lst[e].w;
                        if (strategy == OG_ZEROFILL) {
                            wn += w;
                            if (act)
                                for (int m = 0; m < K; m++) ws[m] += w * in[m][s];
                        } else if (act) {
                            wn += w;
                            for (int m = 0; m < K; m++) ws[m] += w * in[m][s];
                        }
                    }
                    for (int m = 0; m < K; m++) out[m][t] = wn > 1e-15 ? ws[m] / wn : 0.0;
                }
        }
        free(lst);
    }
}

/* ------------------------------------------------------------- derivatives (Derivatives.jl) - */

/* Lagrange first-derivative weights at x for nodes (a, b, c) — the 3-node Fornberg stencil. */
static void lag3(double x, double a, double b, double c, double w[3]) {
    w[0] = (2 * x - b - c) / ((a - b) * (a - c));
    w[1] = (2 * x - a - c) / ((b - a) * (b - c));
    w[2] = (2 * x - a - b) / ((c - a) * (c - b));
}

/* Derivatives.jl:26-31 → FlowGeometries.Discretization.derivative!(order=1, nodes=3, masked=0,
 * ReduceInRun()). Stated choice for the un-vendored stencil: centred 3-node window when both
 * neighbours are active; otherwise the window shifts inward within the run of active cells
 * (3 nodes, CHANGELOG "Bounded-domain end points are now 2nd order"), then 2 nodes for a run of
 * two, and 0 for an isolated cell; periodic axes wrap index and coordinate. Spherical metric:
 * 1/(R cosφ) ∂λ (0 where |cosφ| < √eps) and 1/R ∂φ (Derivatives.jl:16-19). */
void og_deriv(const og_grid *g, int axis, const double *f, double *df) {
    int64_t n0 = g->n[0], n1 = g->n[1], n2 = g->ndim == 3 ? g->n[2] : 1;
    int64_t N = n0 * n1 * n2;
    int64_t stride = axis == 0 ? 1 : (axis == 1 ? n0 : n0 * n1);
    int64_t n = g->n[axis];
    const double *x = g->x[axis];
    int per = g->periodic[axis];
    double L = g->period[axis];
    for (int64_t t = 0; t < N; t++) {
        if (!active_at(g, t)) { df[t] = 0.0; continue; }
        int64_t p = (t / stride) % n;
        int64_t base = t - p * stride;
        if (n < 2) { df[t] = 0.0; continue; }
        /* neighbour q = p + off: index and coordinate (shifted on a periodic wrap) */
        int ok[5];
        double xc[5];
        int64_t ix[5];
        for (int o = -2; o <= 2; o++) {
            int64_t q = p + o;
            double sh = 0.0;
            int valid = 1;
            if (q < 0 || q >= n) {
                if (!per || n < 3) valid = 0;
                else {
                    sh = q < 0 ? -L : L;
                    q = wrap(q, n);
                }
            }
            ok[o + 2] = valid && active_at(g, base + q * stride);
            ix[o + 2] = q;
            xc[o + 2] = valid ? x[q] + sh : 0.0;
        }
        double w3[3];
        double v;
        if (ok[1] && ok[3]) {
            lag3(xc[2], xc[1], xc[2], xc[3], w3);
            v = w3[0] * f[base + ix[1] * stride] + w3[1] * f[t] + w3[2] * f[base + ix[3] * stride];
        } else if (ok[3] && ok[4] && g_mutation != OG_MUTATE_RUNEND_2NODE) {
            lag3(xc[2], xc[2], xc[3], xc[4], w3);
            v = w3[0] * f[t] + w3[1] * f[base + ix[3] * stride] + w3[2] * f[base + ix[4] * stride];
        } else if (ok[1] && ok[0] && g_mutation != OG_MUTATE_RUNEND_2NODE) {
            lag3(xc[2], xc[0], xc[1], xc[2], w3);
            v = w3[0] * f[base + ix[0] * stride] + w3[1] * f[base + ix[1] * stride] + w3[2] * f[t];
        } else if (ok[3]) {
            v = (f[base + ix[3] * stride] - f[t]) / (xc[3] - xc[2]);
        } else if (ok[1]) {
            v = (f[t] - f[base + ix[1] * stride]) / (xc[2] - xc[1]);
        } else {
            v = 0.0;
        }
        if (g->spherical) {
            if (axis == 0) {
                int64_t j = (t / n0) % n1;
                double c = cos(g->x[1][j]);
                v = fabs(c) < sqrt(DBL_EPSILON) ? 0.0 : v / (g->R * c);
            } else {
                v = v / g->R;
            }
        }
        df[t] = v;
    }
}

/* ------------------------------------------------------------- rotations (FlowGeometries) ---- */

/* Local (east, north, radial) ↔ planetary Cartesian: e_east = (−sinλ, cosλ, 0),
 * e_north = (−sinφ cosλ, −sinφ sinλ, cosφ), e_r = (cosφ cosλ, cosφ sinλ, sinφ). */
void og_vector_to_cartesian(double u, double v, double w, double lam, double phi, double out[3]) {
    double sl = sin(lam), cl = cos(lam), sp = sin(phi), cp = cos(phi);
    out[0] = -sl * u - sp * cl * v + cp * cl * w;
    out[1] = cl * u - sp * sl * v + cp * sl * w;
    out[2] = cp * v + sp * w;
}

void og_vector_from_cartesian(double X, double Y, double Z, double lam, double phi, double out[3]) {
    double sl = sin(lam), cl = cos(lam), sp = sin(phi), cp = cos(phi);
    out[0] = -sl * X + cl * Y;
    out[1] = -sp * cl * X - sp * sl * Y + cp * Z;
    out[2] = cp * cl * X + cp * sl * Y + sp * Z;
}

/* t = (xx, xy, xz, yy, yz, zz) → (ee, en, er, nn, nr, rr) = e_aᵀ T e_b (Diagnostics.jl:449-455). */
void og_tensor_to_local(const double t[6], double lam, double phi, double out[6]) {
    double sl = sin(lam), cl = cos(lam), sp = sin(phi), cp = cos(phi);
    double e[3][3] = {{-sl, cl, 0.0}, {-sp * cl, -sp * sl, cp}, {cp * cl, cp * sl, sp}};
    double T[3][3] = {{t[0], t[1], t[2]}, {t[1], t[3], t[4]}, {t[2], t[4], t[5]}};
    double Te[3][3];
    for (int a = 0; a < 3; a++)
        for (int r = 0; r < 3; r++) Te[r][a] = T[r][0] * e[a][0] + T[r][1] * e[a][1] + T[r][2] * e[a][2];
    int idx[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int q = 0; q < 6; q++) {
        int a = idx[q][0], b = idx[q][1];
        out[q] = e[a][0] * Te[0][b] + e[a][1] * Te[1][b] + e[a][2] * Te[2][b];
    }
}

/* ------------------------------------------------------------------------ Π (Diagnostics.jl) - */

/* rowsel (2-D banded engine only; NULL = every row): filter only the rows it marks. */
static void filt(const og_grid *g, const og_filter_spec *s, const og_footprint *fp, int K,
                 const double *const *in, double *const *out, const uint8_t *rowsel) {
    if (s->engine == 1) og_apply_separable(g, s->alpha, s->ell, s->strategy, K, in, out);
    else if (s->engine == 2) og_apply_nd3(g, s->kind, s->alpha, s->ell, s->strategy, K, in, out, s->nthreads);
    else if (s->engine == 3) og_apply_prefix(g, s->ell, s->strategy, K, in, out, s->nthreads);
    else if (s->engine == 4) og_apply_spectral(g, s->kind, s->alpha, s->ell, s->strategy, K, in, out);
    else apply_banded_rows(g, fp, s->strategy, K, in, out, rowsel, s->nthreads);
}

static void compute_pi_impl(const og_grid *g, const og_filter_spec *spec, const og_footprint *fp,
                            const double *u, const double *v, const double *w, const uint8_t *rowsel,
                            double *Pi, double *ubar, double *vbar, double *wbar);

void og_compute_pi(const og_grid *g, const og_filter_spec *spec, const og_footprint *fp,
                   const double *u, const double *v, const double *w,
                   double *Pi, double *ubar, double *vbar, double *wbar) {
    compute_pi_impl(g, spec, fp, u, v, w, NULL, Pi, ubar, vbar, wbar);
}

/* The same Π at the rows `rows` marks only (2-D grids, banded footprint): the filter runs on those
 * rows and the two on either side (all the derivative stencils reach), the rest is elementwise or
 * O(N). Π, ū, v̄ are NaN on unmarked rows. The values on marked rows are those og_compute_pi gives,
 * bit for bit: every row is filtered by the same per-row code. */
void og_compute_pi_rows(const og_grid *g, const og_filter_spec *spec, const og_footprint *fp,
                        const double *u, const double *v, const uint8_t *rows,
                        double *Pi, double *ubar, double *vbar) {
    int64_t Nx = g->n[0], Ny = g->n[1];
    uint8_t *need = (uint8_t *)calloc(Ny, 1);
    for (int64_t j = 0; j < Ny; j++)
        if (rows[j])
            for (int64_t d = -2; d <= 2; d++)
                if (j + d >= 0 && j + d < Ny) need[j + d] = 1;
    compute_pi_impl(g, spec, fp, u, v, NULL, need, Pi, ubar, vbar, NULL);
    for (int64_t j = 0; j < Ny; j++)
        if (!rows[j])
            for (int64_t i = 0; i < Nx; i++) {
                Pi[i + Nx * j] = NAN;
                if (ubar) ubar[i + Nx * j] = NAN;
                if (vbar) vbar[i + Nx * j] = NAN;
            }
    free(need);
}

static void compute_pi_impl(const og_grid *g, const og_filter_spec *spec, const og_footprint *fp,
                            const double *u, const double *v, const double *w, const uint8_t *rowsel,
                            double *Pi, double *ubar, double *vbar, double *wbar) {
    int64_t N = g->n[0] * g->n[1] * (g->ndim == 3 ? g->n[2] : 1);
    int64_t Nx = g->n[0], Ny = g->n[1];
    /* zeroed, so rows a row selection leaves unfiltered hold finite values */
    double *buf = (double *)calloc(24 * N, sizeof(double));
    double *B[24];
    for (int q = 0; q < 24; q++) B[q] = buf + q * N;
    if (g->ndim == 3) {
        /* Diagnostics.jl:792-857 — true 3-D Cartesian. */
        double *uf = B[0], *vf = B[1], *wf = B[2];
        const double *vin[3] = {u, v, w};
        double *vout[3] = {uf, vf, wf};
        filt(g, spec, fp, 3, vin, vout, rowsel);
        for (int64_t t = 0; t < N; t++) {
            B[3][t] = u[t] * u[t]; B[4][t] = u[t] * v[t]; B[5][t] = u[t] * w[t];
            B[6][t] = v[t] * v[t]; B[7][t] = v[t] * w[t]; B[8][t] = w[t] * w[t];
        }
        const double *pin[6] = {B[3], B[4], B[5], B[6], B[7], B[8]};
        double *pout[6] = {B[9], B[10], B[11], B[12], B[13], B[14]}; /* uu uv uw vv vw ww */
        filt(g, spec, fp, 6, pin, pout, rowsel);
        double *Sxx = B[15], *Syy = B[16], *Szz = B[17], *Sxy = B[18], *Sxz = B[19], *Syz = B[20], *sc = B[21];
        og_deriv(g, 0, uf, Sxx);
        og_deriv(g, 1, vf, Syy);
        og_deriv(g, 2, wf, Szz);
        og_deriv(g, 1, uf, Sxy); og_deriv(g, 0, vf, sc);
        for (int64_t t = 0; t < N; t++) Sxy[t] = 0.5 * (Sxy[t] + sc[t]);
        og_deriv(g, 2, uf, Sxz); og_deriv(g, 0, wf, sc);
        for (int64_t t = 0; t < N; t++) Sxz[t] = 0.5 * (Sxz[t] + sc[t]);
        og_deriv(g, 2, vf, Syz); og_deriv(g, 1, wf, sc);
        for (int64_t t = 0; t < N; t++) Syz[t] = 0.5 * (Syz[t] + sc[t]);
        for (int64_t t = 0; t < N; t++) {
            double txx = B[9][t] - uf[t] * uf[t], txy = B[10][t] - uf[t] * vf[t], txz = B[11][t] - uf[t] * wf[t];
            double tyy = B[12][t] - vf[t] * vf[t], tyz = B[13][t] - vf[t] * wf[t], tzz = B[14][t] - wf[t] * wf[t];
            double c = Sxx[t] * txx + 2.0 * Sxy[t] * txy + Syy[t] * tyy + 2.0 * Sxz[t] * txz +
                       2.0 * Syz[t] * tyz + Szz[t] * tzz;
            Pi[t] = active_at(g, t) ? -c : 0.0;
        }
        if (ubar) memcpy(ubar, uf, N * sizeof(double));
        if (vbar) memcpy(vbar, vf, N * sizeof(double));
        if (wbar) memcpy(wbar, wf, N * sizeof(double));
    } else if (!g->spherical) {
        /* Diagnostics.jl:485-537, 683-698 — Cartesian 2-D (no w). */
        double *uf = B[0], *vf = B[1];
        const double *vin[2] = {u, v};
        double *vout[2] = {uf, vf};
        filt(g, spec, fp, 2, vin, vout, rowsel);
        for (int64_t t = 0; t < N; t++) {
            B[2][t] = u[t] * u[t]; B[3][t] = u[t] * v[t]; B[4][t] = v[t] * v[t];
        }
        const double *pin[3] = {B[2], B[3], B[4]};
        double *pout[3] = {B[5], B[6], B[7]};
        filt(g, spec, fp, 3, pin, pout, rowsel);
        double *Sxx = B[8], *Sxy = B[9], *Syy = B[10], *sc = B[11];
        og_deriv(g, 0, uf, Sxx);
        og_deriv(g, 1, uf, Sxy);
        og_deriv(g, 0, vf, sc);
        og_deriv(g, 1, vf, Syy);
        for (int64_t t = 0; t < N; t++) Sxy[t] = 0.5 * (Sxy[t] + sc[t]);
        for (int64_t t = 0; t < N; t++) {
            double txx = B[5][t] - uf[t] * uf[t], txy = B[6][t] - uf[t] * vf[t], tyy = B[7][t] - vf[t] * vf[t];
            double c = Sxx[t] * txx + 2.0 * Sxy[t] * txy + Syy[t] * tyy;
            Pi[t] = active_at(g, t) ? -c : 0.0;
        }
        if (ubar) memcpy(ubar, uf, N * sizeof(double));
        if (vbar) memcpy(vbar, vf, N * sizeof(double));
    } else {
        /* Diagnostics.jl:550-681 — spherical: planetary-Cartesian filtering, local frame back. */
        double *ux = B[0], *uy = B[1], *uz = B[2];
        for (int64_t j = 0; j < Ny; j++)
            for (int64_t i = 0; i < Nx; i++) {
                int64_t t = i + Nx * j;
                if (active_at(g, t)) {
                    double p[3];
                    og_vector_to_cartesian(u[t], v[t], 0.0, g->x[0][i], g->x[1][j], p);
                    ux[t] = p[0]; uy[t] = p[1]; uz[t] = p[2];
                } else {
                    ux[t] = uy[t] = uz[t] = 0.0;
                }
            }
        double *uxf = B[3], *uyf = B[4], *uzf = B[5];
        const double *vin[3] = {ux, uy, uz};
        double *vout[3] = {uxf, uyf, uzf};
        filt(g, spec, fp, 3, vin, vout, rowsel);
        for (int64_t t = 0; t < N; t++) {
            B[6][t] = ux[t] * ux[t]; B[7][t] = ux[t] * uy[t]; B[8][t] = ux[t] * uz[t];
            B[9][t] = uy[t] * uy[t]; B[10][t] = uy[t] * uz[t]; B[11][t] = uz[t] * uz[t];
        }
        const double *pin[6] = {B[6], B[7], B[8], B[9], B[10], B[11]};
        double *pout[6] = {B[12], B[13], B[14], B[15], B[16], B[17]}; /* uu uv uw vv vw ww */
        filt(g, spec, fp, 6, pin, pout, rowsel);
        double *uf = B[6], *vf = B[7], *tee = B[8], *ten = B[9], *tnn = B[10];
        for (int64_t j = 0; j < Ny; j++)
            for (int64_t i = 0; i < Nx; i++) {
                int64_t t = i + Nx * j;
                if (!active_at(g, t)) {
                    uf[t] = vf[t] = tee[t] = ten[t] = tnn[t] = 0.0;
                    if (wbar) wbar[t] = 0.0;
                    continue;
                }
                double lam = g->x[0][i], phi = g->x[1][j];
                double l[3];
                og_vector_from_cartesian(uxf[t], uyf[t], uzf[t], lam, phi, l);
                uf[t] = l[0]; vf[t] = l[1];
                if (wbar) wbar[t] = l[2];
                double tc[6] = {B[12][t] - uxf[t] * uxf[t], B[13][t] - uxf[t] * uyf[t],
                                B[14][t] - uxf[t] * uzf[t], B[15][t] - uyf[t] * uyf[t],
                                B[16][t] - uyf[t] * uzf[t], B[17][t] - uzf[t] * uzf[t]};
                double tl[6];
                og_tensor_to_local(tc, lam, phi, tl);
                tee[t] = tl[0]; ten[t] = tl[1]; tnn[t] = tl[3];
            }
        double *Sxx = B[18], *Sxy = B[19], *Syy = B[20], *sc = B[21];
        og_deriv(g, 0, uf, Sxx);
        og_deriv(g, 1, uf, Sxy);
        og_deriv(g, 0, vf, sc);
        og_deriv(g, 1, vf, Syy);
        double R = g->R;
        for (int64_t j = 0; j < Ny; j++) {
            double phi = g->x[1][j];
            double sp = sin(phi), cp = cos(phi);
            double tf = fabs(cp) > 1e-12 ? sp / (R * cp) : 0.0;
            for (int64_t i = 0; i < Nx; i++) {
                int64_t t = i + Nx * j;
                if (!active_at(g, t)) continue;
                Sxx[t] -= vf[t] * tf;
                Sxy[t] = 0.5 * (Sxy[t] + sc[t] + uf[t] * tf);
            }
        }
        for (int64_t t = 0; t < N; t++) {
            double c = Sxx[t] * tee[t] + 2.0 * Sxy[t] * ten[t] + Syy[t] * tnn[t];
            Pi[t] = active_at(g, t) ? -c : 0.0;
        }
        if (ubar) memcpy(ubar, uf, N * sizeof(double));
        if (vbar) memcpy(vbar, vf, N * sizeof(double));
    }
    free(buf);
}

/* Diagnostics.jl:1051-1059 */
double og_active_area(const og_grid *g) {
    int64_t n0 = g->n[0], n1 = g->n[1], n2 = g->ndim == 3 ? g->n[2] : 1;
    double tot = 0.0;
    for (int64_t k = 0; k < n2; k++)
        for (int64_t j = 0; j < n1; j++)
            for (int64_t i = 0; i < n0; i++)
                if (active_at(g, i + n0 * (j + n1 * k))) tot += og_area(g, i, j, k);
    return tot;
}

/* Diagnostics.jl:1125-1134 — ½ Σ_active |ū|² A / Σ_active A. */
double og_energy(const og_grid *g, const double *ub, const double *vb, const double *wb) {
    int64_t n0 = g->n[0], n1 = g->n[1], n2 = g->ndim == 3 ? g->n[2] : 1;
    double e = 0.0;
    for (int64_t k = 0; k < n2; k++)
        for (int64_t j = 0; j < n1; j++)
            for (int64_t i = 0; i < n0; i++) {
                int64_t t = i + n0 * (j + n1 * k);
                if (!active_at(g, t)) continue;
                double v2 = ub[t] * ub[t] + vb[t] * vb[t];
                if (wb) v2 += wb[t] * wb[t];
                e += v2 * og_area(g, i, j, k);
            }
    return 0.5 * e / og_active_area(g);
}

/* Diagnostics.jl:1319-1333 */
void og_spectral_density(int64_t n, const double *C, const double *k, double *g) {
    if (n < 2) {
        for (int64_t i = 0; i < n; i++) g[i] = 0.0;
        return;
    }
    for (int64_t i = 0; i < n; i++) {
        if (i == 0) g[i] = (C[1] - C[0]) / (k[1] - k[0]);
        else if (i == n - 1) g[i] = (C[n - 1] - C[n - 2]) / (k[n - 1] - k[n - 2]);
        else g[i] = (C[i + 1] - C[i - 1]) / (k[i + 1] - k[i - 1]);
    }
}
