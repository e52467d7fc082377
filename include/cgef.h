/*
 * cgef.h — C ABI of libcgef_hip.so, the MI355X (gfx950) backend for CoarseGrainingEnergyFluxes.jl's
 * real-space footprint filter and Π pipeline.
 *
 * Plain pointers and sizes only. Field pointers are DEVICE pointers (HBM-resident), laid out as the
 * reference lays out its arrays: column-major (Nx, Ny[, Nz], trailing batch) with i fastest.
 * Descriptor arrays (axis coordinates, mask) are HOST pointers copied at creation.
 *
 * Every entry point returns a cgef_status; CGEF_ERR_ARGUMENT corresponds to the reference's
 * ArgumentError and CGEF_ERR_DIMENSION to its DimensionMismatch. cgef_last_error() returns the
 * message of the calling thread's last failure. All compute entry points are stream-ordered on the
 * given hipStream_t (NULL = the null stream) and return after enqueueing unless stated otherwise;
 * entry points that return host values (energies) synchronize that stream.
 */
#ifndef CGEF_H
#define CGEF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CGEF_OK = 0,
    CGEF_ERR_ARGUMENT = 1,    /* reference: ArgumentError */
    CGEF_ERR_DIMENSION = 2,   /* reference: DimensionMismatch */
    CGEF_ERR_DEVICE = 3,      /* a HIP runtime call failed */
    CGEF_ERR_ALLOC = 4,       /* device or host allocation failed */
    CGEF_ERR_UNSUPPORTED = 5  /* valid request this backend has no kernel for */
} cgef_status;

typedef enum {
    CGEF_F32 = 0,
    CGEF_F64 = 1,
    CGEF_BF16 = 2  /* learned-closure tensors only (stored as uint16 bit patterns) */
} cgef_dtype;
typedef enum { CGEF_CARTESIAN = 0, CGEF_SPHERICAL = 1 } cgef_geometry;
typedef enum { CGEF_TOPHAT = 0, CGEF_GAUSSIAN = 1 } cgef_kernel_kind;
typedef enum { CGEF_ZEROFILL = 0, CGEF_DEFORMABLE = 1 } cgef_mask_strategy;
typedef enum {
    CGEF_ENGINE_AUTO = 0,      /* the reference's default per grid/kernel */
    CGEF_ENGINE_BANDED = 1,    /* disk/cap/ball footprint direct sum (fused, LDS-tiled) */
    CGEF_ENGINE_SEPARABLE = 2, /* square-truncated two-pass Gaussian (Cartesian 2-D) */
    CGEF_ENGINE_BASELINE = 3,  /* plain FP64 direct sum, one thread per point, CSR in HBM */
    CGEF_ENGINE_RING = 4,      /* the same banded footprint evaluated in Fourier space along a
                                  periodic x axis (FP64 FFT + banded contraction over rows) */
    CGEF_ENGINE_PREFIX = 5,    /* top-hat on a 2-D rectilinear grid by per-row prefix sums
                                  (Filtering.jl:751-1141), the reference's top-hat engine */
    CGEF_ENGINE_SPECTRAL = 6   /* method = Spectral() on a doubly periodic uniform 2-D Cartesian
                                  grid: irfft2(Ĝ(|k|)·rfft2(mask·f)), FFTWExt.jl:121-208; never
                                  chosen by AUTO (the reference's default method is RealSpace) */
} cgef_engine;

/* Grid description. Reference: FlowGeometries.Grids.StructuredGrid (un-vendored), as consumed by
 * Filtering.jl:636-732 (axes, spacing, periodicity, radius) and Diagnostics.jl:1051-1059 (area). */
typedef struct {
    int32_t ndim;             /* 2 or 3 */
    int32_t geometry;         /* cgef_geometry; spherical requires ndim == 2 */
    int64_t n[3];             /* Nx, Ny, Nz (Nz ignored for ndim 2) */
    const double *coord[3];   /* host: axis coordinates; spherical lon/lat in radians */
    int32_t uniform[3];       /* 1 = the axis is a Range (constant step) */
    int32_t periodic[3];
    double period[3];         /* wrap length of a periodic axis */
    double radius;            /* sphere radius (spherical only) */
    const uint8_t *mask;      /* host: one byte per cell, nonzero = active; NULL = all active */
} cgef_grid_desc;

/* Kernels.jl:28, 59-62 — TopHatKernel / GaussianKernel(α). */
typedef struct {
    int32_t kind;             /* cgef_kernel_kind */
    double alpha;             /* Gaussian α (ignored for the top-hat) */
} cgef_kernel_desc;

typedef struct cgef_grid cgef_grid;
typedef struct cgef_plan cgef_plan;
typedef struct cgef_workspace cgef_workspace;

typedef struct {
    int32_t engine;           /* resolved cgef_engine */
    int32_t nbands;
    int64_t nentries;         /* footprint table entries (in-support taps summed over bands) */
    int64_t nruns;
    double mean_taps;         /* in-support taps per grid point, averaged over rows */
    int64_t max_taps;
    double radius;            /* kernel_radius(kernel, scale) */
    double active_taps;       /* in-support taps summed over ACTIVE target points */
    int64_t ring_offsets;     /* ring engine: distinct (dz, dy) source-row offsets; 0 otherwise */
    int64_t ring_modes;       /* ring engine: Fourier modes kept along x (N/2 + 1); 0 otherwise */
    int64_t ring_pi_fields;   /* ring engine: planes filtered per snapshot by compute_pi; 0 otherwise */
    int64_t ring_contract_kernel; /* ring engine: the contraction kernel kept by measurement for the
                                     widest call so far (0-2 VALU weight-batch variants, 3 matrix
                                     cores), −1 before any measured call */
    double ring_modes_kept;   /* ring engine: fraction of the (row, mode) pairs FP32 calls contract
                                 under the mode cut (1 = none dropped; FP64 calls keep every mode):
                                 row r keeps modes m < M(r), the smallest M at which the weight tail
                                 2·Σ_{m≥M} Σ_offsets |Ŵ(m)| is ≤ 3e-8 of its weight sum and the same
                                 tail weighted by max(1, 4ℓ·sin(πm/N)/Δx) ≤ 3e-6 of it, so under
                                 ZeroFill a filtered value moves by ≤ 3e-8·max|q| and an x-difference
                                 of one by ≤ 3e-6·max|q|/ℓ; 0 for other engines */
} cgef_plan_info;

const char *cgef_version(void);
const char *cgef_last_error(void);

/* Device-resident grid (mask, axes, cell measures, rotation tables). Replaces the grid argument
 * the reference threads through plan_filter / compute_Π! (Filtering.jl:2931; Diagnostics.jl:389). */
cgef_status cgef_grid_create(const cgef_grid_desc *desc, int32_t device, cgef_grid **out);
cgef_status cgef_grid_destroy(cgef_grid *grid);
/* Diagnostics.jl:1051-1059 active_area; ArgumentError when every cell is masked. */
cgef_status cgef_active_area(const cgef_grid *grid, double *out);

/* Filtering.jl:2931-2954 plan_filter (+ the GPU residency hook prepare_workspace, 2910-2912 and
 * KernelAbstractionsExt:474-571): builds the footprint once on the host and makes it resident.
 * ArgumentError for a non-finite or non-positive scale (Filtering.jl:2917-2922). */
cgef_status cgef_plan_create(cgef_grid *grid, const cgef_kernel_desc *kernel, double scale,
                             int32_t mask_strategy, int32_t engine, cgef_plan **out);
cgef_status cgef_plan_destroy(cgef_plan *plan);
cgef_status cgef_plan_query(const cgef_plan *plan, cgef_plan_info *info);

/* ΠWorkspace (Diagnostics.jl:145-194): caller-owned scratch sized for nbatch slices, for the
 * given element type. Passing NULL to the entry points below uses a scratch the grid owns. */
cgef_status cgef_workspace_create(const cgef_grid *grid, int64_t nbatch, int32_t dtype, cgef_workspace **out);
cgef_status cgef_workspace_destroy(cgef_workspace *ws);

/* Filtering.jl:3242-3256 filter_apply!, 3360-3385 filter_apply_batch!, 3286-3305
 * filter_apply_batched! and the GPU hooks gpu_filter_field!/gpu_filter_field_batched!
 * (Filtering.jl:195-197, 3342-3344). Filters nfields × nbatch slices of prod(n) elements each,
 * slice s of `in` at in + s·prod(n) (field-major, then batch), into the same layout of `out`. */
cgef_status cgef_filter_apply(cgef_plan *plan, void *out, const void *in, int64_t nfields,
                              int64_t nbatch, int32_t dtype, void *stream);

/* Diagnostics.jl:389-416, 472-702 (2-D, Cartesian and spherical) and 792-857 (3-D Cartesian):
 * Π = -rho0 · S̄:τ (rho0 = 1 reproduces the reference, which applies no density) for nbatch
 * slices. w may be NULL for 2-D grids (required for 3-D). ubar/vbar/wbar (device, may be NULL)
 * receive the filtered velocity the reference leaves in ws.u_filt/v_filt/w_filt (local frame on the
 * sphere). E_out (host, nbatch entries, may be NULL) receives energy_from_filtered(!)
 * (Diagnostics.jl:1086-1121) for each slice; requesting it synchronizes the stream. */
cgef_status cgef_compute_pi(cgef_plan *plan, cgef_workspace *ws, void *Pi, const void *u,
                            const void *v, const void *w, int64_t nbatch, int32_t dtype,
                            double rho0, double *E_out, void *ubar, void *vbar, void *wbar,
                            void *stream);

/* Pipeline.jl:412-468 coarse_grain! with the GPU batch driver of Pipeline.jl:519-563: Π for nscales
 * plans of one grid over nbatch slices in one call. Pi is (nbatch, nscales, grid…) in C order — the
 * reference's column-major (grid…, Nscales, batch…) — and E_out (host, may be NULL) receives
 * E(ℓ) as (nbatch × nscales, scale fastest). With ring-engine plans the velocity and its products
 * are transformed once for all scales. Synchronizes the stream when E_out is requested. */
cgef_status cgef_coarse_grain(cgef_plan *const *plans, int64_t nscales, cgef_workspace *ws, void *Pi,
                              const void *u, const void *v, const void *w, int64_t nbatch, int32_t dtype,
                              double rho0, double *E_out, void *stream);

/* No counterpart in the reference. For callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc): cgef_energy with its scratch — the per-block partial sums
 * and the per-slice energies — in one fixed device array of 1,052,672 doubles (8 MiB) that the
 * library holds on each GPU it runs on until the process ends, instead of the grid's workspace that
 * cgef_energy allocates on first use. Allocates nothing. Limits: nbatch ≤ 4096 and
 * nbatch · ceil(Nx/256)·Ny·Nz ≤ 1,048,576 (0.25° sphere: 242 slices), else CGEF_ERR_DIMENSION
 * before any launch. Calls on one GPU share the array, so calls that overlap on one GPU race.
 * Synchronizes the stream. */
cgef_status cgef_energy_noalloc(const cgef_grid *grid, const void *ubar, const void *vbar, const void *wbar,
                                int64_t nbatch, int32_t dtype, double *E_out, void *stream);

/* Derivatives.jl:20-144 ddx!/ddy!/ddz!: derivative w.r.t. distance along axis 0/1/2 of nbatch
 * slices (3-node stencil, masked cells 0). */
cgef_status cgef_derivative(const cgef_grid *grid, int32_t axis, void *out, const void *in,
                            int64_t nbatch, int32_t dtype, void *stream);

/* Diagnostics.jl:1125-1134 _energy_over: ½ Σ_active |ū|² A / Σ_active A per slice (wbar may be
 * NULL). Synchronizes the stream. */
cgef_status cgef_energy(const cgef_grid *grid, const void *ubar, const void *vbar, const void *wbar,
                        int64_t nbatch, int32_t dtype, double *E_out, void *stream);

/* Diagnostics.jl:1144-1217 cumulative_energy!: filters the raw (u, v[, w]) with plans[s] and
 * reduces E per scale; E_out is host (nscales × nbatch, scale fastest). Synchronizes the stream. */
cgef_status cgef_cumulative_energy(cgef_plan *const *plans, int64_t nscales, cgef_workspace *ws,
                                   const void *u, const void *v, const void *w, int64_t nbatch,
                                   int32_t dtype, double *E_out, void *stream);

/* Kernel timing (no counterpart in the reference; for measurement): while enabled (on = 1), every
 * launch of the plan's dominant filter step is bracketed by HIP events recorded on the launch stream
 * (banded: the fused filter kernel; ring: the contraction; separable and prefix-sum: their two
 * filter kernels together); on = 2 (ring plans) times the Π planes' inverse transform + epilogue
 * instead; on = 0 turns timing off.
 * cgef_profile_read synchronizes on those events and returns the summed kernel time (ms) and the
 * number of launches since enabling; enabling again resets the counters. */
cgef_status cgef_profile_enable(cgef_plan *plan, int32_t on);
/* The ring engine's Fourier-mode cut (no counterpart in the reference; for measurement and
 * inspection): modes[r] (host, nrows = Ny·Nz entries) receives the number of Fourier modes along x
 * that row r keeps in FP32 calls (N/2 + 1 where the cut drops none; see ring_modes_kept).
 * CGEF_ERR_UNSUPPORTED for plans of other engines. */
cgef_status cgef_plan_mode_cut(const cgef_plan *plan, int32_t *modes, int64_t nrows);
cgef_status cgef_profile_read(cgef_plan *plan, double *total_ms, int64_t *nlaunches);

/* Diagnostics.jl:1319-1333 spectral_density!: host arrays of length n. */
cgef_status cgef_spectral_density(int64_t n, const double *C, const double *k, double *g);

/* ---------------------------------------------------------------------------------------------
 * Learned closure for τ_ℓ. No counterpart in the reference, which obtains τ only from the exact
 * filter: these entry points train and run a 4-layer 5×5 convolutional network (BF16 weights and
 * activations, FP32 accumulation) whose targets are the exact τ this library's filter produces,
 * and the exact Π remains the check and the fallback. Tensors are device (n, channels, rows, cols), cols
 * fastest; BF16 tensors are uint16 bit patterns. All are stream-ordered and do not synchronize.
 * --------------------------------------------------------------------------------------------- */

/* One 5×5 convolution layer. pad 0: valid, out = (n, cout, hin−4, win−4); pad 4: full, out =
 * (n, cout, hin+4, win+4) with zeros outside the input (the data gradient uses it with the
 * backward-data weight layout). Input BF16 (in_gelu: GELU applied while loading, i.e. the input
 * is the previous layer's pre-activation) or FP32. w is BF16 [cin][25][ocs], k = 5·ky + kx, ocs a
 * multiple of the output-channel block (4 for cout ≤ 4, 16 for cout ≤ 16, else 64). Without zgrad:
 * out = conv + bias as BF16 or FP32; with a BF16 output, act (may be NULL, shaped as out) receives
 * GELU(out) in BF16, the next layer's input. With zgrad (BF16, shaped as out): out = conv ·
 * GELU'(zgrad), FP32, bias unused. */
cgef_status cgef_conv5(const void *in, int32_t in_dtype, int32_t in_gelu, int64_t n, int64_t cin, int64_t hin,
                       int64_t win, const uint16_t *w, int64_t ocs, const float *bias, int64_t cout, int32_t pad,
                       void *out, int32_t out_dtype, const uint16_t *zgrad, uint16_t *act, void *stream);

/* The same layer for 64 input and 64 output channels on the matrix cores (BF16 MFMA, FP32
 * accumulation; an FP32 input is rounded to BF16 as it is staged). wm is the layout written by
 * cgef_conv5_pack_mfma: forward [25][64 out][64 in], or backward-data [25][64][64] for pad 4.
 * Other arguments as in cgef_conv5. */
cgef_status cgef_conv5_mfma(const void *in, int32_t in_dtype, int32_t in_gelu, int64_t n, int64_t hin, int64_t win,
                            const uint16_t *wm, const float *bias, int32_t pad, void *out, int32_t out_dtype,
                            const uint16_t *zgrad, uint16_t *act, void *stream);

/* BF16 matrix-core layouts of a 64 × 64 master weight tensor W[o][i][25]: forward wmf[k][o][i] and
 * backward-data wmb[24 − k][i][o] (either may be NULL), 64·64·25 entries each. */
cgef_status cgef_conv5_pack_mfma(const float *w, uint16_t *wmf, uint16_t *wmb, void *stream);

/* cgef_conv5_mfma for 1 to 64 input and output channels (the end layers of the closure network:
 * 4 → 64, 64 → 3 and the data gradient through the latter). On the matrix cores the input channels
 * are padded to cin_pad = 32 (cin ≤ 32) or 64 and the output channels to cout_pad = 16 (cout ≤ 16)
 * or 64, with zeros; wm is the layout of cgef_conv5_pack_mfma_ex for this (cin, cout) — its forward
 * layout for pad 0, its backward-data layout (a layer cout → cin) for pad 4. in_gelu only with 33 to
 * 64 channels on both sides. Other arguments as in cgef_conv5. */
cgef_status cgef_conv5_mfma_ex(const void *in, int32_t in_dtype, int32_t in_gelu, int64_t n, int64_t cin, int64_t hin,
                               int64_t win, const uint16_t *wm, const float *bias, int64_t cout, int32_t pad, void *out,
                               int32_t out_dtype, const uint16_t *zgrad, uint16_t *act, void *stream);

/* BF16 matrix-core layouts of a cout × cin master weight tensor W[o][i][25] (1 to 64 each):
 * forward wmf[k][o][i] of 25 · pad16(cout) · pad32(cin) entries and backward-data wmb[24 − k][i][o]
 * of 25 · pad16(cin) · pad32(cout) entries, where pad16(c) = 16 for c ≤ 16 else 64 and pad32(c) = 32
 * for c ≤ 32 else 64; padding entries are zero. Either output may be NULL. */
cgef_status cgef_conv5_pack_mfma_ex(const float *w, int64_t cout, int64_t cin, uint16_t *wmf, uint16_t *wmb,
                                    void *stream);

/* Weight and bias gradients of a valid 5×5 layer: dw[o][i][k] = Σ dz[n][o][p]·A[n][i][p+k],
 * db[o] = Σ dz[n][o][p], with A the layer input (n, cin, hout+4, wout+4) as in cgef_conv5 and dz
 * FP32 (n, cout, hout, wout). mfma = 1 (64 → 64 layers only) runs the GEMMs on the matrix cores with
 * dz and A rounded to BF16; mfma = 0 uses FP32 VALU products of the stored values. Deterministic:
 * per-block partial sums go to `scratch` (at least the floats cgef_conv5_wgrad_scratch returns for
 * the same arguments) and are added in a fixed order. Overwrites dw, db. */
cgef_status cgef_conv5_wgrad_scratch(int64_t n, int64_t cin, int64_t cout, int64_t hout, int64_t wout, int32_t mfma,
                                     int64_t *floats);
cgef_status cgef_conv5_wgrad(const float *dz, const void *in, int32_t in_dtype, int32_t in_gelu, int64_t n,
                             int64_t cin, int64_t cout, int64_t hout, int64_t wout, int32_t mfma, float *dw, float *db,
                             float *scratch, int64_t scratch_floats, void *stream);

/* BF16 layouts of an FP32 master weight tensor W[cout][cin][25]: forward wf[cin][25][ocs_f] and
 * backward-data wb[cout][25][ocs_b] with wb[o][24−k][i] = W[o][i][k] (either may be NULL). Entries
 * of the padding channels are left as they are (zero them once at allocation). */
cgef_status cgef_conv5_pack_weights(const float *w, int64_t cout, int64_t cin, uint16_t *wf, int64_t ocs_f,
                                    uint16_t *wb, int64_t ocs_b, void *stream);

/* One Adam step (bias-corrected; step counts from 1) on an FP32 parameter tensor with FP32 moments;
 * if wf / wb are given the updated weights are also written as BF16 layouts as in
 * cgef_conv5_pack_weights (then count = cout·cin·25). */
cgef_status cgef_adam_update(float *param, float *m, float *v, const float *grad, int64_t count, double lr,
                             double beta1, double beta2, double eps, int64_t step, int64_t cout, int64_t cin,
                             uint16_t *wf, int64_t ocs_f, uint16_t *wb, int64_t ocs_b, void *stream);

/* Mean-squared-error gradient dpred = 2(pred − target)/count (FP32); Σ(pred − target)² is added to
 * the device scalar *loss_sum when it is not NULL. */
cgef_status cgef_mse_grad(const float *pred, const float *target, int64_t count, float *dpred, float *loss_sum,
                          void *stream);

/* Closure loss with a Π term: L = mean over the n·3·npix τ values of (pred − target)² + λ·mean over
 * the n·npix pixels of (Π(pred) − Π(target))², Π(τ) = −(a0 τxx + a1 τxy + a2 τyy) with a0 = g0·coef[0]
 * ·coef[4], a1 = (g1·coef[1] + g2·coef[2])·coef[5], a2 = g3·coef[3]·coef[6] from feat = g (n, 4, npix)
 * at the same pixels (coef: 7 device floats). pred, target, dpred (n, 3, npix) FP32; dpred = ∂L/∂pred.
 * loss (2 device floats, may be NULL) += (Σ τ part, Σ Π part) before the means. */
cgef_status cgef_closure_loss_grad(const float *pred, const float *target, const float *feat, const float *coef,
                                   double lambda_pi, int64_t n, int64_t npix, float *dpred, float *loss, void *stream);

/* Periodic crops: dst[i][c][y][x] = chan_scale[c]·src[img][c][(y0+y) mod H][(x0+x) mod W] with
 * origins[i] = (img, y0, x0) (device int32, n × 3); src FP32 (nsrc, c, H, W); dst FP32 or BF16
 * (n, c, h, w); chan_scale may be NULL. An origin whose image index is outside [0, nsrc) gives zeros. */
cgef_status cgef_closure_crop(const void *src, int32_t src_dtype, int64_t nsrc, int64_t c, int64_t H, int64_t W,
                              const int32_t *origins, int64_t n, int64_t h, int64_t w, const float *chan_scale,
                              void *dst, int32_t dst_dtype, void *stream);

/* The planes the exact τ is filtered from: out5 = (nb, 5, npts) FP64 holding u, v, u², u·v, v². */
cgef_status cgef_closure_products(const void *u, const void *v, int32_t dtype, int64_t npts, int64_t nb, double *out5,
                                  void *stream);

/* From the filtered planes (nb, 5, npts) and the gradients (nb, 4, npts) of ū, v̄ (∂x ū, ∂y ū,
 * ∂x v̄, ∂y v̄), all FP64: tau3 = (τxx, τxy, τyy) = F[q_a q_b] − q̄_a q̄_b and feat4 = the gradients,
 * both FP32. With tau3 NULL only feat4 is written and filt5 may be NULL. */
cgef_status cgef_closure_targets(const double *filt5, const double *grad4, int64_t npts, int64_t nb, float *tau3,
                                 float *feat4, void *stream);

/* Π = −(S̄xx τxx + 2 S̄xy τxy + S̄yy τyy) (Diagnostics.jl:460-466 contraction) from feat4, with
 * τ = tau3 · tau_scale[c] + sigma2 · G Gᵀ (tau_scale may be NULL; sigma2 = 0 uses tau3 alone; G the
 * gradient matrix in feat4: the learned residual plus the gradient model), FP32 (nb, npts). */
cgef_status cgef_closure_pi(const float *feat4, const float *tau3, const float *tau_scale, double sigma2, int64_t npts,
                            int64_t nb, float *Pi, void *stream);

/* The closure without learning, the gradient model τ_ab = σ² Σ_k ∂_k ū_a ∂_k ū_b, plus base3 when
 * it is not NULL (base3 = exact τ with σ² negated gives the residual the network learns). */
cgef_status cgef_closure_gradient_model(const float *feat4, double sigma2, const float *base3, int64_t npts,
                                        int64_t nb, float *tau3, void *stream);

#ifdef __cplusplus
}
#endif
#endif
