// Kernel argument blocks and launch entry points shared by the host API and the device code.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cgef_internal.hpp"

namespace cgef {

constexpr int TI = 256;        // targets per workgroup along x (one per lane)
constexpr int MAX_RUN = 1536;  // runs wider than this are split (keeps the LDS segment bounded)

// Geometry as the kernels see it (all device pointers).
struct GeomArgs {
    int64_t n[3];          // Nx, Ny, Nz (Nz = 1 in 2-D)
    int64_t npts;
    int32_t periodic[3];
    double period[3];
    double R;
    int32_t mode;          // 0 Cartesian 2-D, 1 spherical, 2 Cartesian 3-D
    const uint8_t *mask;   // one byte per cell, never null
    const uint16_t *scode; // derivative stencil code per cell (STENCIL_* kind per axis, STENCIL_ACTIVE)
    const double *coord[3];
    const double *width[3];
    const double *trig;    // spherical: sinλ[Nx], cosλ[Nx], sinφ[Ny], cosφ[Ny], then per row the strain's
                           // 1/(R cosφ) [Ny] and tanφ/R [Ny] (0 within √ε, resp. 1e-12, of a pole)
    const double *dcoef[3];  // derivative stencil weights, DCOEF_SLOTS arrays of n[axis] each
    double active_area;
};

struct BandArgs {
    const Run *runs;
    const int32_t *run_ptr;
    const int32_t *band_of_row;
    const int32_t *row_order;
    const float *w32;      // FP32 weights (FP32 path, pass 2)
    const double *wd;      // FP64 weights: exact copies of w32 (FP32 path) or the FP64 weights
    const double *D_row;   // per-row in-range weight sum matching wd
    int32_t strategy;
    int32_t max_run;       // widest run (hi - lo + 1)
};

GeomArgs geom_args(const Grid &g);
BandArgs band_args(const Plan &p, bool fp32);

// ---- launches (cgef_kernels.hip) ----
void launch_mean(const GeomArgs &G, int dtype, const void *u, const void *v, const void *w, int64_t nbatch,
                 double *partials, int64_t npartials, double *cmean, hipStream_t s);
void launch_prep(const GeomArgs &G, int dtype, const void *u, const void *v, const void *w, int64_t nbatch,
                 const double *cmean, void *qpack, hipStream_t s);
void launch_pi_filter(const GeomArgs &G, const BandArgs &B, int dtype, bool masked, bool edge_x,
                      const void *qpack, const double *cmean, double *ubar, void *tau, int64_t nbatch,
                      hipStream_t s);
void launch_strain_pi(const GeomArgs &G, int dtype, const double *ubar, const void *tau, void *Pi,
                      double rho0, int64_t nbatch, double *partials, int64_t npartials, hipStream_t s,
                      int64_t pi_stride = 0);
void launch_energy_final(int64_t nbatch, int64_t nblocks, const double *partials, double active_area,
                         double *E_dev, hipStream_t s);
int64_t strain_blocks(const GeomArgs &G);   // partial sums written by the strain / Π kernel per slice
int64_t energy_blocks(const GeomArgs &G);   // … and by the energy-of-fields kernel (at least as many)
void launch_filter_lin(const GeomArgs &G, const BandArgs &B, int dtype, bool masked, bool edge_x,
                       const void *in, void *out, int64_t nslices, hipStream_t s);
void launch_baseline_filter(const GeomArgs &G, const Plan &p, const double *in, double *out, int64_t nslices,
                            hipStream_t s);
void launch_separable(const GeomArgs &G, const Plan &p, int dtype, const void *in, void *out, int64_t nslices,
                      double *rowpass, hipStream_t s);
// separable Π path (2-D Cartesian): rp holds 5 planes per slice; ū/v̄/w̄ also leave in the caller's
// type when uo/vo/wo are given
void launch_separable_pi(const GeomArgs &G, const Plan &p, int dtype, const void *u, const void *v, int64_t nbatch,
                         double *rp, double *ubar, void *tau, void *uo, void *vo, void *wo, hipStream_t s);
// prefix-sum top-hat (cgef_prefix.hip): Pbuf holds nslices × Ny × (Nx + 1) doubles
void launch_prefix(const GeomArgs &G, const Plan &p, int dtype, const void *in, void *out, int64_t nslices,
                   double *Pbuf, hipStream_t s);
// Π planes scanned straight from (u, v[, w]) (Pbuf: nf·nbatch × Ny × (Nx + 1) doubles), then ū and τ
// per point; ū/v̄/w̄ also leave in the caller's type when uo/vo/wo are given
void launch_prefix_pi(const GeomArgs &G, const Plan &p, int dtype, const void *u, const void *v, const void *w,
                      int64_t nbatch, double *Pbuf, double *ubar, void *tau, void *uo, void *vo, void *wo,
                      hipStream_t s);
void launch_unfused_prep(const GeomArgs &G, int dtype, const void *u, const void *v, const void *w, int64_t nbatch,
                         double *vel, hipStream_t s);
void launch_products(const GeomArgs &G, const double *vel, double *prod, int64_t nbatch, hipStream_t s);
void launch_unfused_tau(const GeomArgs &G, const double *velf, const double *prodf, double *ubar, double *tau,
                        int64_t nbatch, hipStream_t s);
void launch_derivative(const GeomArgs &G, int axis, int dtype, const void *in, void *out, int64_t nbatch,
                       hipStream_t s);
void launch_energy_fields(const GeomArgs &G, int dtype, const void *ub, const void *vb, const void *wb,
                          int64_t nbatch, double *partials, int64_t npartials, hipStream_t s);
void launch_cast(int dtype_out, const double *in, void *out, int64_t n, hipStream_t s);

}  // namespace cgef
