/*
 * cgef_oracle.h — FP64 CPU restatement of the reference's real-space filter + Π pipeline.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker, never the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The product path
 * (libcgef_hip.so) neither links nor calls anything declared here.
 *
 * Layout convention (the reference's): column-major, i (x / longitude) fastest, then j
 * (y / latitude), then k (z), then trailing batch. 0-based indices here.
 *
 * Pinning: tests/test_oracle.py checks these functions against the reference's own known answers,
 * closed forms, invariants, the identities its test suite asserts, and brute force. Two choices the
 * un-vendored FlowGeometries.jl makes are pinned by what the reference states and the mathematics
 * fixes, each test also failing on the plausible wrong choice planted with og_set_mutation:
 *   - og_deriv at domain and mask-run ends: the 3-node window shifted inward (CHANGELOG.md:257-259),
 *     so a quadratic is differentiated exactly in every run of three or more cells; 2 nodes for a run
 *     of two, 0 for an isolated cell;
 *   - og_cell_width on non-uniform axes: midpoint widths with half cells at bounded ends — the widths
 *     tile the axis, equal the step inside a collected uniform axis (test_grids.jl:203-206), and make
 *     the measure-weighted filter converge at 2nd order on a smoothly stretched axis.
 * Still "parity unpinned", with the choice made here (the HIP kernels make the same one):
 *   - periodicity: explicit flags (no auto-detection);
 *   - og_haversine rounding at a spherical top-hat rim;
 *   - inputs are promoted to FP64 (no mixed-precision promotion rules).
 */
#ifndef CGEF_ORACLE_H
#define CGEF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t ndim;          /* 2 or 3 */
    int32_t spherical;     /* 0 Cartesian, 1 spherical lon-lat (ndim 2) */
    int64_t n[3];          /* Nx, Ny, Nz */
    const double *x[3];    /* axis coordinates; spherical: lon, lat in radians */
    int32_t uniform[3];    /* 1: axis is a Range (constant step) */
    int32_t periodic[3];
    double period[3];
    double R;              /* sphere radius */
    const uint8_t *mask;   /* NULL = every cell active; else one byte per cell */
} og_grid;

enum { OG_TOPHAT = 0, OG_GAUSSIAN = 1 };
enum { OG_ZEROFILL = 0, OG_DEFORMABLE = 1 };

/* Test-only: plant a known mistake so a pinning test can show it fails on it (0 restores). */
enum { OG_MUTATE_NONE = 0, OG_MUTATE_FORWARD_WIDTH = 1, OG_MUTATE_RUNEND_2NODE = 2 };
void og_set_mutation(int m);

double og_kernel_weight(int kind, double alpha, double d, double ell);
double og_kernel_radius(int kind, double alpha, double ell);
double og_haversine(double R, double lam1, double phi1, double lam2, double phi2);
double og_area(const og_grid *g, int64_t i, int64_t j, int64_t k);
double og_cell_width(const og_grid *g, int axis, int64_t i);

/* Banded 2-D footprint (CSR): entries di, dj, w; band b owns ptr[b]..ptr[b+1]-1. */
typedef struct {
    int64_t nentries;
    int64_t nbands;
    int32_t *di;
    int32_t *dj;
    double *w;
    int64_t *ptr;
} og_footprint;

og_footprint *og_footprint_build(const og_grid *g, int kind, double alpha, double ell);
/* Spherical grids: only the latitude bands bandsel marks (the others empty); NULL = all. */
og_footprint *og_footprint_build_bands(const og_grid *g, int kind, double alpha, double ell, const uint8_t *bandsel);
void og_footprint_free(og_footprint *fp);
int64_t og_footprint_nentries(const og_footprint *fp);
int64_t og_footprint_nbands(const og_footprint *fp);
void og_footprint_copy(const og_footprint *fp, int32_t *di, int32_t *dj, double *w, int64_t *ptr);

/* Apply a banded footprint to K fields sharing one neighbour walk (2-D grids). */
void og_apply_banded(const og_grid *g, const og_footprint *fp, int strategy, int K,
                     const double *const *in, double *const *out, int nthreads);

/* Separable Gaussian engine (2-D Cartesian, uniform axes): square truncation. */
void og_apply_separable(const og_grid *g, double alpha, double ell, int strategy, int K,
                        const double *const *in, double *const *out);

/* Top-hat by per-row prefix sums (2-D rectilinear grids; the reference's top-hat engine). */
double og_metric_band(const og_grid *g, double y1, double y2, double rad);
void og_apply_prefix(const og_grid *g, double ell, int strategy, int K, const double *const *in,
                     double *const *out, int nthreads);

/* Spectral filter on a periodic uniform 2-D Cartesian grid (the reference's FFTW extension). */
double og_spectral_transfer(int kind, double alpha, double k, double ell);
void og_apply_spectral(const og_grid *g, int kind, double alpha, double ell, int strategy, int K,
                       const double *const *in, double *const *out);

/* 3-D Cartesian ball footprint applied per point (uniform or stretched axes). */
void og_apply_nd3(const og_grid *g, int kind, double alpha, double ell, int strategy, int K,
                  const double *const *in, double *const *out, int nthreads);

/* First derivative along `axis` w.r.t. distance (3-node stencil, mask-run aware). */
void og_deriv(const og_grid *g, int axis, const double *f, double *df);

/* Planetary-Cartesian rotations (spherical). */
void og_vector_to_cartesian(double u, double v, double w, double lam, double phi, double out[3]);
void og_vector_from_cartesian(double X, double Y, double Z, double lam, double phi, double out[3]);
void og_tensor_to_local(const double t[6], double lam, double phi, double out[6]);

/* engine: 0 banded footprint, 1 separable (Cartesian Gaussian only), 2 3-D ball, 3 prefix-sum
 * top-hat (2-D), 4 spectral (periodic uniform 2-D Cartesian). */
typedef struct {
    int32_t kind;
    double alpha;
    double ell;
    int32_t strategy;
    int32_t engine;
    int32_t nthreads;
} og_filter_spec;

/* Π = -S̄:τ for one slice (no ρ₀). Outputs ubar/vbar/wbar (may be NULL) hold the filtered
 * velocity the reference leaves in the workspace (local frame on the sphere). */
void og_compute_pi(const og_grid *g, const og_filter_spec *spec, const og_footprint *fp,
                   const double *u, const double *v, const double *w,
                   double *Pi, double *ubar, double *vbar, double *wbar);

/* Π at the rows rows[j] != 0 marks only (2-D grids, banded footprint; NaN elsewhere), identical on
 * those rows to og_compute_pi: for checking wide footprints at full size on a few rows. */
void og_compute_pi_rows(const og_grid *g, const og_filter_spec *spec, const og_footprint *fp,
                        const double *u, const double *v, const uint8_t *rows,
                        double *Pi, double *ubar, double *vbar);

double og_active_area(const og_grid *g);
double og_energy(const og_grid *g, const double *ub, const double *vb, const double *wb);
void og_spectral_density(int64_t n, const double *C, const double *k, double *g);

#ifdef __cplusplus
}
#endif
#endif
