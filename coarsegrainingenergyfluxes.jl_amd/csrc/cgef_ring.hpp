// Ring engine (Fourier-space evaluation of the banded footprint along the periodic x axis): plan
// structures and launches. See cgef_ring.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "cgef_launch.hpp"

namespace cgef {

bool ring_factor(int n, RingFft &F);
// The register-staged factorization of F.n (F.snst, F.srad, F.lpsh): the fewest stages whose
// butterflies fit the per-thread register budget of a 256-thread block, odd radices first. False when
// there is none or the length exceeds 4096 (then the LDS ping-pong FFT serves).
bool ring_staged_factor(RingFft &F);
bool ring_supported(const Plan &p, std::string *why = nullptr);
void fft_tables(RingFft &F, double2 *&tw_d, double2 *&tws_d);
bool spectral_supported(const Grid &g, int kind);
void build_spectral(Plan &p, const GeomArgs &G, hipStream_t s);
// spectral engine launches: per (mode along x, plane column) transform along y, × Ĝ, transform back
void ring_spectral_cols(const RingPlan &R, const double2 *XF, double2 *OUT, int64_t ncol, hipStream_t s);
// the same split for sweeps over scales: the forward y transform once, then × Ĝ and back per scale
void ring_spectral_analyze_cols(const RingPlan &R, const double2 *XF, double2 *OUT, int64_t ncol, hipStream_t s);
void ring_spectral_synth_cols(const RingPlan &R, const double2 *XF, double2 *OUT, int64_t ncol, hipStream_t s);
void ring_spectral_invert(const GeomArgs &G, const double *renorm, double *invrenorm, hipStream_t s);
void ring_spectral_normalize(const GeomArgs &G, int dtype, const double *yout, const double *invrenorm, int64_t nsl,
                             void *out, hipStream_t s);
void ring_spectral_epilogue(const GeomArgs &G, int dtype, const double *yout, const double *invrenorm, double *ubar,
                            void *tau, void *uo, void *vo, void *wo, int64_t nbatch, hipStream_t s);
void build_ring(Plan &p, hipStream_t s);
void free_ring(RingPlan &R);
void ring_build_what(const RingPlan &R, const Run *runs_d, const int32_t *run_off_d, const double *w64_d, int64_t nruns,
                     hipStream_t s);
void ring_forward(const RingPlan &R, const double *xin, double2 *XF, int64_t ncol, hipStream_t s);
// The mode cut's per-row table for a call in `dtype`: the plan's cut for FP32 calls when it drops any
// mode (CGEF_MODE_CUT=0 turns it off), none (null) for FP64 calls.
const int32_t *ring_mode_cut(const RingPlan &R, int dtype);
// Builds the plan's mode cut (RING_MODE_TOL, RING_MODE_SLOPE_TOL) from Ŵ: mcut_d and the contraction
// work lists.
void ring_build_mode_cut(const Plan &p, RingPlan &R, hipStream_t s);
// Per band the modes kept (dxb: the band's smallest x spacing, device; S: mh × bands scratch; mband:
// bands).
void ring_mode_band_cut(const RingPlan &R, double ell, const double *dxb, double *S, int32_t *mband, hipStream_t s);
// out_rowmajor: OUT as [m][row][col] (what the fused inverse + epilogue reads) instead of [m][col][row].
// mcut (from ring_mode_cut, may be null): contract only the (mode, row block) pairs of the cut's work
// lists; the consumers of OUT then read row r's modes m < mcut[r] only.
void ring_contract(const RingPlan &R, const double2 *XF, double2 *OUT, int64_t ncol, hipStream_t s,
                   bool out_rowmajor = false, const int32_t *mcut = nullptr);
// fills R.wt_d (allocated by the caller: mh × wt_tiles × wt_steps × 64 doubles) from Ŵ
void ring_build_wtiles(RingPlan &R, hipStream_t s);
void ring_inverse(const RingPlan &R, const double2 *OUT, double *yout, int64_t ncol, hipStream_t s,
                  const int32_t *mcut = nullptr);
// Π inputs: per slice nf planes (frame velocity components, their products, then the mask when
// `withmask`); the epilogue turns the filtered planes into ū (FP64) and τ, optionally exporting ū in T.
void ring_prep(const GeomArgs &G, int dtype, const void *u, const void *v, const void *w, int nf, int withmask,
               double *xin, int64_t nbatch, hipStream_t s);
// Whether the inverse row transforms of the Π planes can run fused with the epilogue for this plan,
// plane set and batch (row lengths with a compiled-in factorization); if not, ring_inverse +
// ring_epilogue. The fused kernels read the contracted spectra as [m][row][col] (ring_contract with
// out_rowmajor). spectral_rotation: on the 1440-point sphere (except Deformable with a radial
// export), rotate to the local frame in Fourier space (the caller has checked that longitude ascends
// by 2π/N from trig[0]), else per point after the transforms.
bool ring_inverse_epilogue_serves(const RingPlan &R, const GeomArgs &G, int nf, int withmask, int strategy,
                                  int64_t nbatch);
void ring_inverse_epilogue(const RingPlan &R, const GeomArgs &G, int dtype, const double2 *OUT, const double *D_row,
                           int nf, int withmask, int strategy, double *ubar, void *tau, void *uo, void *vo, void *wo,
                           int64_t nbatch, bool spectral_rotation, hipStream_t s, const int32_t *mcut = nullptr);
void ring_epilogue(const GeomArgs &G, int dtype, const double *yout, const double *D_row, int nf, int withmask,
                   int strategy, double *ubar, void *tau, void *uo, void *vo, void *wo, int64_t nbatch, hipStream_t s);
void ring_prep_fields(const GeomArgs &G, int dtype, const void *in, int64_t nsl, int withmask, double *xin, hipStream_t s);
void ring_normalize(const GeomArgs &G, int dtype, const double *yout, const double *D_row, int64_t nsl, int withmask,
                    int strategy, void *out, hipStream_t s);

}  // namespace cgef
