// Internal structures of libcgef_hip.so: device-resident grid, footprint plan and workspace.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cgef.h"

namespace cgef {

// Error plumbing: the C ABI returns a status and keeps a per-thread message.
void set_error(const std::string &msg);
struct Error {
    cgef_status status;
    std::string msg;
};
[[noreturn]] void fail(cgef_status s, const std::string &msg);
void hip_check(hipError_t e, const char *what);
#define CGEF_HIP(call) ::cgef::hip_check((call), #call)

// One run of the banded footprint: source row (dy, dz) relative to the target row, columns
// di = lo..hi relative to the target column, weights at w[woff + (di - lo)].
struct alignas(32) Run {
    int32_t dy, dz, lo, hi;
    int32_t woff, pad0, pad1, pad2;
};

struct Workspace;

// First-derivative stencil weights of an axis of n points, DCOEF_SLOTS arrays of n doubles (slot s of
// index p at s·n + p, so lanes on consecutive p read consecutive words): slots 0-2 centred nodes
// p−1, p, p+1; 3-5 forward p, p+1, p+2; 6-8 backward p−2, p−1, p; 9-10 two-node forward p, p+1;
// 11-12 two-node backward p−1, p. Periodic axes (n ≥ 3) wrap index and coordinate.
constexpr int DCOEF_SLOTS = 13;

// First-derivative stencil kinds of a point along one axis (3 bits each in the point's stencil code):
// none (isolated cell), centred (slots 0-2), forward and backward 3-node (3-5, 6-8), forward and
// backward 2-node (9-10, 11-12). STENCIL_ACTIVE marks an active point.
enum : int {
    STENCIL_NONE = 0,
    STENCIL_CENTRED = 1,
    STENCIL_FORWARD3 = 2,
    STENCIL_BACKWARD3 = 3,
    STENCIL_FORWARD2 = 4,
    STENCIL_BACKWARD2 = 5,
};
constexpr uint16_t STENCIL_ACTIVE = 0x8000;

// Ring engine plan (see cgef_ring.hip): mixed-radix factorization of the row length, the (dz, dy)
// offsets of the footprint and the Fourier weights Ŵ[m][offset][band].
constexpr int RING_MAX_STAGES = 24;
// Row lengths whose staged factorization the row kernels also carry at compile time (RingFft::fixed):
// 1440 (0.25° in longitude) as 12·12·10, 256 as 4·4·4·4.
constexpr int RING_FIXED_1440 = 1;
constexpr int RING_FIXED_256 = 2;

struct RingFft {
    int n = 0;
    int nstage = 0;
    int radix[RING_MAX_STAGES] = {};
    // register-staged factorization (radices up to 16, composite ones evaluated in registers):
    // snst stages of radix srad[st]; the staged FFT's LDS index shift lpsh (4: padded, 31: not)
    int snst = 0;
    int srad[RING_MAX_STAGES] = {};
    int lpsh = 31;
    int nsq = 1;                       // staged row transforms: sequences (row pairs) per block
    int sstr = 0;                      // staged LDS stride between sequences (padded length)
    int twoff[RING_MAX_STAGES] = {};  // offset of staged stage st in the per-stage twiddle table
    int twsize = 0;                    // entries of the per-stage twiddle table
    int nseq = 0;                      // 1: the register-staged FFT serves this length (one sequence per block)
                                       // (0: not supported for this length, the LDS ping-pong FFT serves)
    int fixed = 0;                     // RING_FIXED_*: srad is that compile-time factorization (0: none)
};

// Columns (planes) per block of the ring engine's banded contraction, whose LDS slab holds
// RING_CONTRACT_CC × (256 + the footprint's dy range) complex FP64 values.
constexpr int RING_CONTRACT_CC = 5;
// Largest Ŵ table (mh × offsets × bands FP64) a ring plan may hold; wider footprints go to the
// banded direct sum under engine = AUTO.
constexpr double RING_WHAT_BUDGET_BYTES = 32.0 * 1024 * 1024 * 1024;
// Offsets from which a 2-D ring plan also carries the matrix-core contraction's weight table.
constexpr int RING_MFMA_MIN_OFFSETS = 48;
// LDS per CU on gfx950, which is also the most one workgroup may take.
constexpr size_t LDS_BYTES_PER_CU = 160 * 1024;
// Fourier-mode cut of the ring engine for FP32 calls: target row r keeps the modes m < M(r), the
// smallest M for which both tails below hold, with S(m) = Σ_offsets |Ŵ_band(r)(m)|, S0 the row's
// in-range weight sum Σ_offsets Ŵ_band(r)(0) and c_m = 2 (1 for modes 0 and N/2):
//   values:  Σ_{m≥M} c_m·S(m)                                ≤ RING_MODE_TOL · S0,
//   slopes:  Σ_{m≥M} c_m·S(m)·max(1, 4ℓ·sin(πm/N)/Δx_r)      ≤ RING_MODE_SLOPE_TOL · S0.
// Each |DFT coefficient| of a row is at most N·max|q|, and every x-difference stencil of the strain
// (centred, one-sided two- and three-node) multiplies mode m by at most 4·sin(πm/N)/Δx. So under
// ZeroFill, whose denominator is S0, every filtered value moves by at most RING_MODE_TOL·max|q| (the
// size of the FP32 output's rounding) and every x-difference of one by at most
// RING_MODE_SLOPE_TOL·max|q|/ℓ (Deformable: both times S0 / filtered mask). The slope condition is what
// binds near the poles, where Δx = R·cosφ·Δλ is tiny against ℓ. The bounds sit above the floor that
// the footprint's truncation at G = 1e-10 leaves in every mode (its edge step puts ~1e-12 of S0 into
// each |Ŵ(m)|; summed over modes about 1e-8 for values, 1e-6 for slopes on the 0.25° sphere), so
// the cut follows the kernel's own spectral decay (ℓ = 2000 km on that sphere: 5 modes at the poles, 55 at
// the equator, of 721).
constexpr double RING_MODE_TOL = 3e-8;
constexpr double RING_MODE_SLOPE_TOL = 3e-6;

struct RingArgs {
    int64_t ny = 1, nz = 1, nrows = 1;
    int64_t mh = 0;                    // Fourier modes kept (N/2 + 1)
    int32_t py = 0, pz = 0;
    int32_t noff = 0, nbands = 0;
    int32_t ndz = 0;
    int32_t dy_min = 0, dy_max = 0;
    const int32_t *band_of_row = nullptr;
    const double *what = nullptr;      // [m][offset][band]
    const int32_t *dz_val = nullptr;   // per dz group
    const int32_t *dz_off0 = nullptr;  // offsets of group g: dz_off0[g] .. dz_off0[g+1]-1
    const int32_t *off_dy = nullptr;   // per offset
};

// Measured choice of the banded-contraction kernel variant per column count (see ring_contract).
struct ContractTune {
    std::mutex mu;
    std::vector<std::pair<int64_t, int>> choice;   // (4·columns + 2·mode cut + output layout, variant)
};

struct RingPlan {
    bool ok = false;
    std::shared_ptr<ContractTune> tune = std::make_shared<ContractTune>();
    // spectral engine (FFTWExt.jl:121-208): the same row transforms, then per Fourier mode along x a
    // transform along y, the transfer function Ĝ(|k|), and back — instead of the banded contraction
    bool spectral = false;
    RingFft fy;
    double2 *twy_d = nullptr, *twsy_d = nullptr;
    int kind = CGEF_GAUSSIAN;
    double alpha = 6.0, ell = 0.0, dkx = 0.0, dky = 0.0;
    double *invrenorm_d = nullptr;     // Deformable on a masked grid: 1/filter(mask) where ≥ 0.01, else 0
    RingFft fft;
    int64_t mh = 0;                    // N/2 + 1 Fourier modes kept
    double2 *tw_d = nullptr;           // exp(−2πik/N), k < N
    double2 *tws_d = nullptr;          // per stage (radix R after Ns points): exp(−2πi·r·k/(Ns·R)) at
                                       // twoff + k·(R−1) + r−1, 0 ≤ k < Ns, 1 ≤ r < R
    double *what_d = nullptr;
    int32_t *dz_val_d = nullptr, *dz_off0_d = nullptr, *off_dy_d = nullptr;
    // Ŵ in the A-operand order of the matrix-core contraction (k_contract_mfma), built for 2-D plans
    // with at least RING_MFMA_MIN_OFFSETS offsets: per (m, 16-row tile, 4-row step) the 64 lane values
    // A[i = lane%16][k = lane/16] = Ŵ[m][4·step + k − i][band(16·tile + i)] (0 outside the footprint).
    // A grid with one band (translation invariant along y) stores one tile for all.
    double *wt_d = nullptr;
    int32_t wt_steps = 0, wt_tiles = 0;
    // The mode cut (RING_MODE_TOL, RING_MODE_SLOPE_TOL), used by FP32 calls; mcut_d is null when every
    // row keeps every mode. mcut_d: per row (Ny·Nz) the modes kept. citems_d: the (m, 256-row block of one level)
    // pairs, m · (blocks) + block in ascending m, with m below the cut of some row of the block — the
    // VALU contraction's work list; mitems_d the same over 64-row blocks for the matrix-core one.
    int32_t *mcut_d = nullptr, *citems_d = nullptr, *mitems_d = nullptr;
    int64_t ncitems = 0, nmitems = 0;
    std::vector<int32_t> mcut_h;       // host copy of the per-row modes kept (empty: every mode)
    double modes_kept = 1.0;           // Σ_rows M(r) / (rows · mh)
    RingArgs args;
};

// Prefix-sum top-hat plan (2-D rectilinear grids): per target row j and row offset b the source row
// jj[j·nb + b] (−1: no support), its measure factor wy, and how the axis-1 interval is found:
// wl ≥ 0 a ±wl-cell window (uniform, non-periodic axis), −2 the whole row (support spans a period),
// −1 by comparing the extended coordinates xe with x_i ± hw.
struct PrefixArgs {
    int64_t nx = 0, ny = 0, ne = 0;     // ne: extended axis length (3·nx when periodic, else nx)
    int32_t nb = 0;
    int32_t uniform_x = 0;              // constant step: the search starts from an exact guess
    int32_t deform = 0;                 // Deformable on a masked grid: denominator from pd
    double x0 = 0.0, dx = 0.0;          // xe[0] and the step (uniform_x)
    const int32_t *jj = nullptr, *wl = nullptr;
    const double *wy = nullptr, *hw = nullptr;
    const double *wx = nullptr;         // axis-1 measure factor per index
    const double *xe = nullptr;         // extended ascending coordinates
    const double *pwx = nullptr;        // prefix of wx over the real axis (nx + 1)
    const double *pd = nullptr;         // per row prefix of mask·wx (nx + 1 each), Deformable
};

struct PrefixPlan {
    bool ok = false;
    PrefixArgs args;
    int32_t *jj_d = nullptr, *wl_d = nullptr;
    double *wy_d = nullptr, *hw_d = nullptr, *wx_d = nullptr, *xe_d = nullptr, *pwx_d = nullptr, *pd_d = nullptr;
};

struct Grid {
    int device = 0;
    int ndim = 2, geometry = CGEF_CARTESIAN;
    int64_t n[3] = {1, 1, 1};
    int periodic[3] = {0, 0, 0}, uniform[3] = {1, 1, 1};
    double period[3] = {0, 0, 0};
    double R = 1.0;
    std::vector<double> coord[3], width[3];
    std::vector<uint8_t> mask_h;  // empty = all active
    bool has_mask = false;
    double active_area = 0.0;
    int64_t npts = 0, nrows = 0;  // npts = Nx*Ny*Nz, nrows = Ny*Nz
    // device
    uint8_t *mask_d = nullptr;    // always allocated (all ones when unmasked)
    uint16_t *scode_d = nullptr;  // per-point derivative stencil code (STENCIL_* per axis, STENCIL_ACTIVE)
    double *coord_d[3] = {nullptr, nullptr, nullptr};
    double *width_d[3] = {nullptr, nullptr, nullptr};
    double *trig_d = nullptr;     // sphere: [sinλ(Nx) cosλ(Nx) sinφ(Ny) cosφ(Ny)]
    double *dcoef_d[3] = {nullptr, nullptr, nullptr};  // derivative stencil weights (DCOEF_SLOTS × n)
    std::atomic<int> refs{1};
    std::mutex ws_mu;
    Workspace *internal_ws = nullptr;
    double cell_area(int64_t i, int64_t j, int64_t k) const;
    ~Grid();
};

// Identity of a plan for the lifetime of the process (addresses can be reused after a destroy).
inline uint64_t next_plan_uid() {
    static std::atomic<uint64_t> n{0};
    return ++n;
}

struct Plan {
    uint64_t uid = next_plan_uid();
    Grid *grid = nullptr;
    int kind = CGEF_GAUSSIAN;
    double alpha = 6.0, scale = 0.0, rad = 0.0;
    int strategy = CGEF_ZEROFILL;
    int engine = CGEF_ENGINE_BANDED;
    bool edge_x = false;          // non-periodic x: per-point in-range denominator
    // banded tables (host)
    int nbands = 0;
    std::vector<int32_t> band_of_row, run_ptr, row_order;
    std::vector<Run> runs;
    std::vector<double> w64, w32d, D64_row, D32_row;
    std::vector<float> w32;
    int64_t nentries = 0;
    double mean_taps = 0.0;
    int64_t max_taps = 0;
    int32_t max_run = 1;          // widest run after splitting (LDS segment sizing)
    double active_taps = 0.0;     // Σ over active target points of their in-support taps
    // banded tables (device)
    Run *runs_d = nullptr;
    int32_t *run_ptr_d = nullptr, *band_of_row_d = nullptr, *row_order_d = nullptr;
    double *w64_d = nullptr, *w32d_d = nullptr, *D64_row_d = nullptr, *D32_row_d = nullptr;
    float *w32_d = nullptr;
    // plain CSR of the baseline engine (device), built on first use
    int32_t *csr_di_d = nullptr, *csr_dj_d = nullptr, *csr_dk_d = nullptr;
    int64_t *csr_ptr_d = nullptr;
    // separable engine (device): per-axis 1-D weights and ZeroFill profiles / Deformable 1/den
    int sep_lim[2] = {0, 0};
    double *sep_g_d[2] = {nullptr, nullptr};
    double *sep_prof_d[2] = {nullptr, nullptr};
    double *sep_invden_d = nullptr;
    // ring engine
    RingPlan ring;
    // prefix-sum top-hat engine
    PrefixPlan pre;
    // optional kernel timing: event pairs around each dominant-kernel launch
    int profiling = 0;            // which step cgef_profile_enable times (0: none)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_events;
    size_t prof_used = 0;
    // brackets a launch of step `what` (1: the filter step, 2: the ring engine's Π inverse transform +
    // epilogue) with events when that step is being timed
    void prof_begin(hipStream_t s, int what = 1);
    void prof_end(hipStream_t s, int what = 1);
    ~Plan();
};

struct Workspace {
    Grid *grid = nullptr;
    int64_t nbatch = 0;
    int dtype = CGEF_F32;
    // fused path
    void *qpack = nullptr;        // T4 per point per slice
    double *ubar = nullptr;       // 3 planes per slice (local frame)
    void *tau = nullptr;          // 6 planes per slice, element type T
    double *cmean = nullptr;      // 3 per slice
    double *partials = nullptr;   // reduction scratch
    int64_t npartials = 0;
    double *E_dev = nullptr;      // per-slice energies
    // unfused (baseline / separable) path: FP64 planes, grown on demand
    double *scratch64 = nullptr;
    int64_t nscratch64 = 0;
    void *fbuf = nullptr;         // filtered-field / row-pass scratch, grown on demand
    size_t nfbuf = 0;
    // ring engine scratch: spatial planes (in and out share), forward spectra, contracted spectra
    double *rxy = nullptr;
    double2 *rxf = nullptr, *rout = nullptr;
    size_t nrxy = 0, nrxf = 0;
    // Π of one scale for nbatch slices before it is scattered into a multi-scale result
    void *pistage = nullptr;
    size_t npistage = 0;
    // multi-scale sweeps (cgef_coarse_grain, ring and spectral plans) replayed as one HIP graph: the
    // arguments of the last eager sweep (seen_key) and of the captured one (graph_key)
    std::vector<uint64_t> seen_key, graph_key;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t cap_stream = nullptr;
    bool graph_broken = false;
    ~Workspace();
};

// Footprint construction (host, FP64), cgef_footprint.cpp
void build_banded(Plan &p, bool need_fp32);
// Prefix-sum top-hat tables (host, FP64), cgef_prefix_host.cpp
bool prefix_supported(const Grid &g, int kind);
void build_prefix(Plan &p);
void free_prefix(PrefixPlan &P);
bool ring_grid_ok(const Grid &g);
void build_separable(Plan &p);

}  // namespace cgef
