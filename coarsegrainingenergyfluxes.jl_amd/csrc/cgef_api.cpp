// C ABI of libcgef_hip.so (declared in include/cgef.h).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "cgef_launch.hpp"
#include "cgef_ring.hpp"

#define CGEF_VERSION_STRING "cgef-mi355x 0.1.0 (gfx950)"

namespace cgef {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

[[noreturn]] void fail(cgef_status s, const std::string &msg) { throw Error{s, msg}; }

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) fail(CGEF_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
static cgef_status guard(F &&f) {
    try {
        f();
        return CGEF_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.status;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return CGEF_ERR_ALLOC;
    } catch (const std::exception &e) {
        set_error(e.what());
        return CGEF_ERR_ARGUMENT;
    }
}

template <class P>
static void dfree(P *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

template <class V>
static auto *upload(const V &v) {
    using T = typename V::value_type;
    T *d = nullptr;
    if (v.empty()) return d;
    CGEF_HIP(hipMalloc(&d, v.size() * sizeof(T)));
    CGEF_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

double Grid::cell_area(int64_t i, int64_t j, int64_t k) const {
    if (geometry == CGEF_SPHERICAL) return R * R * std::cos(coord[1][j]) * width[0][i] * width[1][j];
    double a = width[0][i] * width[1][j];
    if (ndim == 3) a *= width[2][k];
    return a;
}

Grid::~Grid() {
    delete internal_ws;
    dfree(mask_d);
    dfree(scode_d);
    for (int d = 0; d < 3; d++) {
        dfree(coord_d[d]);
        dfree(width_d[d]);
    }
    dfree(trig_d);
    for (int d = 0; d < 3; d++) dfree(dcoef_d[d]);
}

void Plan::prof_begin(hipStream_t s, int what) {
    if (profiling != what) return;
    if (prof_used == prof_events.size()) {
        hipEvent_t a, b;
        CGEF_HIP(hipEventCreate(&a));
        CGEF_HIP(hipEventCreate(&b));
        prof_events.emplace_back(a, b);
    }
    CGEF_HIP(hipEventRecord(prof_events[prof_used].first, s));
}

void Plan::prof_end(hipStream_t s, int what) {
    if (profiling != what) return;
    CGEF_HIP(hipEventRecord(prof_events[prof_used].second, s));
    prof_used++;
}

Plan::~Plan() {
    for (auto &e : prof_events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    dfree(runs_d);
    dfree(run_ptr_d);
    dfree(band_of_row_d);
    dfree(row_order_d);
    dfree(w64_d);
    dfree(w32d_d);
    dfree(D64_row_d);
    dfree(D32_row_d);
    dfree(w32_d);
    dfree(csr_di_d);
    dfree(csr_dj_d);
    dfree(csr_dk_d);
    dfree(csr_ptr_d);
    for (int a = 0; a < 2; a++) {
        dfree(sep_g_d[a]);
        dfree(sep_prof_d[a]);
    }
    dfree(sep_invden_d);
    free_ring(ring);
    free_prefix(pre);
    if (grid && --grid->refs == 0) delete grid;
}

Workspace::~Workspace() {
    dfree(qpack);
    dfree(ubar);
    dfree(tau);
    dfree(cmean);
    dfree(partials);
    dfree(scratch64);
    dfree(fbuf);
    dfree(E_dev);
    dfree(rxy);
    dfree(rxf);
    dfree(rout);
    dfree(pistage);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
}

// Lagrange weights of the derivative at x of the quadratic through (a, b, c).
static void lag3_weights(double x, double a, double b, double c, double *w) {
    w[0] = (2 * x - b - c) / ((a - b) * (a - c));
    w[1] = (2 * x - a - c) / ((b - a) * (b - c));
    w[2] = (2 * x - a - b) / ((c - a) * (c - b));
}

// Stencil weights of the 3-node first derivative along one axis (Derivatives.jl:26-31 → the
// FlowGeometries derivative with nodes = 3): which of the five stencils applies at a point depends
// on the mask (stencil_codes); the weights depend only on the axis index.
static std::vector<double> deriv_table(const Grid &g, int axis) {
    const int64_t n = g.n[axis];
    std::vector<double> tab((size_t)n * DCOEF_SLOTS, 0.0);
    if (n < 2) return tab;
    const bool per = g.periodic[axis] && n >= 3;
    const std::vector<double> &x = g.coord[axis];
    for (int64_t p = 0; p < n; p++) {
        double xc[5];
        bool ok[5];
        for (int o = -2; o <= 2; o++) {
            int64_t q = p + o;
            double sh = 0.0;
            ok[o + 2] = true;
            if (q < 0 || q >= n) {
                if (!per) {
                    ok[o + 2] = false;
                    xc[o + 2] = 0.0;
                    continue;
                }
                sh = q < 0 ? -g.period[axis] : g.period[axis];
                q = q < 0 ? q + n : q - n;
            }
            xc[o + 2] = x[q] + sh;
        }
        double w[DCOEF_SLOTS] = {};
        if (ok[1] && ok[3]) lag3_weights(xc[2], xc[1], xc[2], xc[3], w);
        if (ok[3] && ok[4]) lag3_weights(xc[2], xc[2], xc[3], xc[4], w + 3);
        if (ok[0] && ok[1]) lag3_weights(xc[2], xc[0], xc[1], xc[2], w + 6);
        if (ok[3]) {
            w[10] = 1.0 / (xc[3] - xc[2]);
            w[9] = -w[10];
        }
        if (ok[1]) {
            w[12] = 1.0 / (xc[2] - xc[1]);
            w[11] = -w[12];
        }
        for (int sl = 0; sl < DCOEF_SLOTS; sl++) tab[(size_t)sl * n + p] = w[sl];
    }
    return tab;
}

// Which first-derivative stencil applies at each point, per axis (Derivatives.jl:26-31 → FlowGeometries
// derivative!, nodes = 3, ReduceInRun): centred 3-node inside a run of active cells, the window shifted
// inward at a domain or mask edge, two nodes for a run of two, none for an isolated cell. The choice
// depends only on the mask and the axis periodicity, so it is made once per grid: STENCIL_* kinds in
// bits 3·axis .. 3·axis+2 of the point's code, STENCIL_ACTIVE set on active points (inactive: 0).
static std::vector<uint16_t> stencil_codes(const Grid &g, const std::vector<uint8_t> &m) {
    std::vector<uint16_t> code((size_t)g.npts, 0);
    const int64_t stride[3] = {1, g.n[0], g.n[0] * g.n[1]};
    for (int64_t k = 0; k < g.n[2]; k++)
        for (int64_t j = 0; j < g.n[1]; j++)
            for (int64_t i = 0; i < g.n[0]; i++) {
                const int64_t t = i + g.n[0] * (j + g.n[1] * k);
                if (!m[t]) continue;
                const int64_t pos[3] = {i, j, k};
                uint16_t c = STENCIL_ACTIVE;
                for (int d = 0; d < g.ndim; d++) {
                    const int64_t n = g.n[d], p = pos[d];
                    if (n < 2) continue;
                    const bool per = g.periodic[d] && n >= 3;
                    auto ok = [&](int o) {  // neighbour p+o in range (or wrapped) and active
                        int64_t q = p + o;
                        if (q < 0 || q >= n) {
                            if (!per) return false;
                            q = q < 0 ? q + n : q - n;
                        }
                        return m[t + (q - p) * stride[d]] != 0;
                    };
                    const bool m1 = ok(-1), p1 = ok(1);
                    int kind = STENCIL_NONE;
                    if (m1 && p1) kind = STENCIL_CENTRED;
                    else if (p1 && ok(2)) kind = STENCIL_FORWARD3;
                    else if (m1 && ok(-2)) kind = STENCIL_BACKWARD3;
                    else if (p1) kind = STENCIL_FORWARD2;
                    else if (m1) kind = STENCIL_BACKWARD2;
                    c |= (uint16_t)(kind << (3 * d));
                }
                code[t] = c;
            }
    return code;
}

GeomArgs geom_args(const Grid &g) {
    GeomArgs a{};
    for (int d = 0; d < 3; d++) {
        a.n[d] = g.n[d];
        a.periodic[d] = g.periodic[d];
        a.period[d] = g.period[d];
        a.coord[d] = g.coord_d[d];
        a.width[d] = g.width_d[d];
        a.dcoef[d] = g.dcoef_d[d];
    }
    a.npts = g.npts;
    a.R = g.R;
    a.mode = g.geometry == CGEF_SPHERICAL ? 1 : (g.ndim == 3 ? 2 : 0);
    a.mask = g.mask_d;
    a.scode = g.scode_d;
    a.trig = g.trig_d;
    a.active_area = g.active_area;
    return a;
}

BandArgs band_args(const Plan &p, bool fp32) {
    BandArgs b{};
    b.runs = p.runs_d;
    b.run_ptr = p.run_ptr_d;
    b.band_of_row = p.band_of_row_d;
    b.row_order = p.row_order_d;
    b.w32 = p.w32_d;
    b.wd = fp32 ? p.w32d_d : p.w64_d;
    b.D_row = fp32 ? p.D32_row_d : p.D64_row_d;
    b.strategy = p.strategy;
    b.max_run = p.max_run;
    return b;
}

static size_t esize(int dtype) { return dtype == CGEF_F32 ? 4 : 8; }

static void check_dtype(int dtype) {
    if (dtype != CGEF_F32 && dtype != CGEF_F64) fail(CGEF_ERR_ARGUMENT, "dtype must be CGEF_F32 or CGEF_F64");
}

// Workspace sizing: fused-path buffers for nbatch slices of the given dtype.
static void ws_reserve(Workspace &w, int64_t nbatch, int dtype) {
    const Grid &g = *w.grid;
    if (w.nbatch >= nbatch && w.dtype == dtype && w.qpack) return;
    dfree(w.qpack);
    dfree(w.ubar);
    dfree(w.tau);
    dfree(w.cmean);
    dfree(w.partials);
    dfree(w.E_dev);
    const int64_t P = g.npts;
    const size_t es = esize(dtype);
    CGEF_HIP(hipMalloc(&w.qpack, (size_t)P * nbatch * 4 * es));
    CGEF_HIP(hipMalloc(&w.ubar, (size_t)P * nbatch * 3 * sizeof(double)));
    CGEF_HIP(hipMalloc(&w.tau, (size_t)P * nbatch * 6 * es));
    CGEF_HIP(hipMalloc(&w.cmean, (size_t)nbatch * 3 * sizeof(double)));
    const int64_t nblk = (g.n[0] + 255) / 256 * g.n[1] * g.n[2];  // energy_blocks ≥ strain_blocks
    w.npartials = nbatch * std::max<int64_t>(nblk, 3 * 1024);
    CGEF_HIP(hipMalloc(&w.partials, (size_t)w.npartials * sizeof(double)));
    CGEF_HIP(hipMalloc(&w.E_dev, (size_t)nbatch * sizeof(double)));
    w.nbatch = nbatch;
    w.dtype = dtype;
}

static double *ws_scratch64(Workspace &w, int64_t nelems) {
    if (w.nscratch64 < nelems) {
        dfree(w.scratch64);
        CGEF_HIP(hipMalloc(&w.scratch64, (size_t)nelems * sizeof(double)));
        w.nscratch64 = nelems;
    }
    return w.scratch64;
}

static void *ws_fbuf(Workspace &w, size_t bytes) {
    if (w.nfbuf < bytes) {
        dfree(w.fbuf);
        CGEF_HIP(hipMalloc(&w.fbuf, bytes));
        w.nfbuf = bytes;
    }
    return w.fbuf;
}

static Workspace &resolve_ws(Grid &g, cgef_workspace *ws, int64_t nbatch, int dtype) {
    Workspace *w = reinterpret_cast<Workspace *>(ws);
    if (w) {
        if (w->grid != &g) fail(CGEF_ERR_ARGUMENT, "workspace was created for a different grid");
        if (w->nbatch < nbatch)
            fail(CGEF_ERR_DIMENSION, "workspace holds " + std::to_string(w->nbatch) + " slices, call needs " +
                                         std::to_string(nbatch));
        if (w->dtype != dtype) fail(CGEF_ERR_ARGUMENT, "workspace was created for a different dtype");
        return *w;
    }
    std::lock_guard<std::mutex> lk(g.ws_mu);
    if (!g.internal_ws) {
        g.internal_ws = new Workspace();
        g.internal_ws->grid = &g;
    }
    ws_reserve(*g.internal_ws, std::max<int64_t>(nbatch, g.internal_ws->nbatch), dtype);
    return *g.internal_ws;
}

static void split_wide_runs(Plan &p) {
    std::vector<Run> out;
    std::vector<int32_t> ptr(p.nbands + 1, 0);
    for (int b = 0; b < p.nbands; b++) {
        ptr[b] = (int32_t)out.size();
        for (int32_t q = p.run_ptr[b]; q < p.run_ptr[b + 1]; q++) {
            Run r = p.runs[q];
            int32_t lo = r.lo;
            while (r.hi - lo + 1 > MAX_RUN) {
                Run piece = r;
                piece.lo = lo;
                piece.hi = lo + MAX_RUN - 1;
                piece.woff = r.woff + (lo - r.lo);
                out.push_back(piece);
                lo += MAX_RUN;
            }
            Run last = r;
            last.lo = lo;
            last.woff = r.woff + (lo - r.lo);
            out.push_back(last);
        }
    }
    ptr[p.nbands] = (int32_t)out.size();
    p.runs.swap(out);
    p.run_ptr.swap(ptr);
}

static void build_csr(Plan &p) {
    std::vector<int32_t> di, dj, dk;
    std::vector<int64_t> ptr(p.nbands + 1, 0);
    di.reserve(p.w64.size());
    dj.reserve(p.w64.size());
    dk.reserve(p.w64.size());
    for (int b = 0; b < p.nbands; b++) {
        ptr[b] = (int64_t)di.size();
        for (int32_t q = p.run_ptr[b]; q < p.run_ptr[b + 1]; q++) {
            const Run &r = p.runs[q];
            for (int32_t x = r.lo; x <= r.hi; x++) {
                di.push_back(x);
                dj.push_back(r.dy);
                dk.push_back(r.dz);
            }
        }
    }
    ptr[p.nbands] = (int64_t)di.size();
    p.csr_di_d = upload(di);
    p.csr_dj_d = upload(dj);
    p.csr_dk_d = upload(dk);
    p.csr_ptr_d = upload(ptr);
}

static hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Ring engine scratch: ncol spatial planes (FP64, input and output share) and two spectra sets.
static void ring_reserve(Workspace &w, const Plan &p, int64_t ncol) {
    const Grid &g = *p.grid;
    const size_t nxy = (size_t)ncol * g.npts;
    const size_t nxf = (size_t)ncol * p.ring.mh * g.nrows;
    if (w.nrxy < nxy) {
        dfree(w.rxy);
        CGEF_HIP(hipMalloc(&w.rxy, nxy * sizeof(double)));
        w.nrxy = nxy;
    }
    if (w.nrxf < nxf) {
        dfree(w.rxf);
        dfree(w.rout);
        CGEF_HIP(hipMalloc(&w.rxf, nxf * sizeof(double2)));
        CGEF_HIP(hipMalloc(&w.rout, nxf * sizeof(double2)));
        w.nrxf = nxf;
    }
}

// forward transform of w.rxy (ncol planes), banded contraction, inverse back into w.rxy (FP32 calls
// under the plan's mode cut)
static void ring_filter_planes(Plan &p, Workspace &w, int64_t ncol, int dtype, hipStream_t s) {
    const int32_t *mcut = ring_mode_cut(p.ring, dtype);
    ring_forward(p.ring, w.rxy, w.rxf, ncol, s);
    p.prof_begin(s);
    ring_contract(p.ring, w.rxf, w.rout, ncol, s, false, mcut);
    p.prof_end(s);
    ring_inverse(p.ring, w.rout, w.rxy, ncol, s, mcut);
}

// Longitude ascends from coord[0][0] by 2π/N (to 1e-9 of a step): the condition under which the
// sphere's east/north weights are the mode shifts the Fourier-space rotation applies.
static bool lon_ascends_by_ring_step(const Grid &g) {
    const int64_t N = g.n[0];
    if (g.geometry != CGEF_SPHERICAL || (int64_t)g.coord[0].size() != N || N < 2) return false;
    const double step = 2.0 * M_PI / (double)N;
    for (int64_t i = 0; i < N; i++)
        if (std::fabs(g.coord[0][i] - g.coord[0][0] - step * (double)i) > 1e-9 * step) return false;
    return true;
}

// Whether the Π planes' inverse transform runs fused with the epilogue (the row length allows it and
// CGEF_FUSE is not 0); the contraction then writes its output as [m][row][col].
static bool ring_pi_fused(const Plan &p, const GeomArgs &G, int nf, int wm, int64_t nbatch) {
    const char *fe = std::getenv("CGEF_FUSE");
    if (fe && fe[0] == '0') return false;
    return ring_inverse_epilogue_serves(p.ring, G, nf, wm, p.strategy, nbatch);
}

// The contracted spectra of the Π planes (w.rout) to ū and τ: the fused inverse transform + epilogue
// when `fused` (from ring_pi_fused, before the contraction), else the inverse into w.rxy and the
// separate epilogue. On the sphere the fused kernel rotates to the local frame in Fourier space where
// it can; CGEF_FUSE=1 keeps the rotation per point (testing).
static void ring_pi_inverse_epilogue(Plan &p, Workspace &w, const GeomArgs &G, int dtype, int nf, int wm, void *uo,
                                     void *vo, void *wo, int64_t nbatch, bool fused, hipStream_t s) {
    const int32_t *mcut = ring_mode_cut(p.ring, dtype);   // what the contraction ran under
    p.prof_begin(s, 2);
    if (fused) {
        const char *fe = std::getenv("CGEF_FUSE");
        const bool spectral = !(fe && fe[0] == '1') && lon_ascends_by_ring_step(*p.grid);
        ring_inverse_epilogue(p.ring, G, dtype, w.rout, p.D64_row_d, nf, wm, p.strategy, w.ubar, w.tau, uo, vo, wo,
                              nbatch, spectral, s, mcut);
    } else {
        ring_inverse(p.ring, w.rout, w.rxy, nf * nbatch, s, mcut);
        ring_epilogue(G, dtype, w.rxy, p.D64_row_d, nf, wm, p.strategy, w.ubar, w.tau, uo, vo, wo, nbatch, s);
    }
    p.prof_end(s, 2);
}

// spectral engine: forward row transforms, per-mode column transform × Ĝ and back, row inverse
static void spectral_filter_planes(Plan &p, Workspace &w, int64_t ncol, hipStream_t s) {
    ring_forward(p.ring, w.rxy, w.rxf, ncol, s);
    p.prof_begin(s);
    ring_spectral_cols(p.ring, w.rxf, w.rout, ncol, s);
    p.prof_end(s);
    ring_inverse(p.ring, w.rout, w.rxy, ncol, s);
}

static void spectral_filter_apply(Plan &p, Workspace &w, int dtype, const void *in, void *out, int64_t nsl,
                                  hipStream_t s) {
    const GeomArgs G = geom_args(*p.grid);
    ring_reserve(w, p, nsl);
    ring_prep_fields(G, dtype, in, nsl, 0, w.rxy, s);
    spectral_filter_planes(p, w, nsl, s);
    ring_spectral_normalize(G, dtype, w.rxy, p.ring.invrenorm_d, nsl, out, s);
}

// Planes the ring engine filters per slice for Π: the frame velocity components and their products,
// plus the mask when a Deformable plan on a masked grid needs its filtered value as the denominator.
static int ring_withmask(const Grid &g, int strategy) { return g.has_mask && strategy == CGEF_DEFORMABLE ? 1 : 0; }
static int ring_nfields(const Grid &g, int withmask) {
    return (g.ndim == 2 && g.geometry == CGEF_CARTESIAN ? 5 : 9) + withmask;
}

// filter nsl plain slices with the ring engine
static void ring_filter_apply(Plan &p, Workspace &w, int dtype, const void *in, void *out, int64_t nsl, hipStream_t s) {
    Grid &g = *p.grid;
    const GeomArgs G = geom_args(g);
    const int withmask = (g.has_mask && p.strategy == CGEF_DEFORMABLE) ? 1 : 0;
    const int64_t ncol = nsl + withmask;
    ring_reserve(w, p, ncol);
    ring_prep_fields(G, dtype, in, nsl, withmask, w.rxy, s);
    ring_filter_planes(p, w, ncol, dtype, s);
    ring_normalize(G, dtype, w.rxy, p.D64_row_d, nsl, withmask, p.strategy, out, s);
}

static void export_planes(const Grid &g, const double *ubar, int plane, void *out, int64_t nbatch, int dtype, hipStream_t s) {
    const int64_t P = g.npts;
    for (int64_t b = 0; b < nbatch; b++)
        launch_cast(dtype, ubar + (b * 3 + plane) * P, (char *)out + (size_t)b * P * esize(dtype), P, s);
}

static void fetch_energy(Workspace &w, int64_t nbatch, int64_t nblk, double active_area, double *E_out, hipStream_t s) {
    if (!(active_area > 0)) fail(CGEF_ERR_ARGUMENT, "grid has no active cells (all masked out)");
    launch_energy_final(nbatch, nblk, w.partials, active_area, w.E_dev, s);
    CGEF_HIP(hipMemcpyAsync(E_out, w.E_dev, (size_t)nbatch * sizeof(double), hipMemcpyDeviceToHost, s));
    CGEF_HIP(hipStreamSynchronize(s));
}

// per-row prefix sums of nsl slices (prefix-sum engine)
static size_t prefix_bytes(const Grid &g, int64_t nsl) {
    return (size_t)nsl * g.n[1] * (g.n[0] + 1) * sizeof(double);
}

// Unfused Π for the baseline engine (the reference's plain per-field kernels restated): FP64 planes of
// the frame velocity and its products, one filter launch per group of planes, then τ, ū and the strain.
static void compute_pi_unfused(Plan &p, Workspace &w, void *Pi, const void *u, const void *v, const void *wv,
                               int64_t nbatch, int dtype, double rho0, double *E_out, hipStream_t s) {
    Grid &g = *p.grid;
    const GeomArgs G = geom_args(g);
    const int64_t P = g.npts;
    const int NC = G.mode == 0 ? 2 : 3;
    double *base = ws_scratch64(w, P * nbatch * (3 + 3 + 6 + 6 + 6 + 3 + 1));
    double *vel = base, *velf = vel + 3 * P * nbatch, *prod = velf + 3 * P * nbatch, *prodf = prod + 6 * P * nbatch;
    double *tau = prodf + 6 * P * nbatch, *ubar = tau + 6 * P * nbatch, *pi64 = ubar + 3 * P * nbatch;
    launch_unfused_prep(G, dtype, u, v, wv, nbatch, vel, s);
    auto filt = [&](const double *in, double *out, int64_t nsl) { launch_baseline_filter(G, p, in, out, nsl, s); };
    for (int64_t b = 0; b < nbatch; b++) filt(vel + b * 3 * P, velf + b * 3 * P, NC);
    launch_products(G, vel, prod, nbatch, s);
    // the Cartesian 2-D product set is uu, uv, vv: planes 0, 1, 3 of the six
    for (int64_t b = 0; b < nbatch; b++) {
        if (NC == 2) {
            filt(prod + b * 6 * P, prodf + b * 6 * P, 2);
            filt(prod + b * 6 * P + 3 * P, prodf + b * 6 * P + 3 * P, 1);
        } else {
            filt(prod + b * 6 * P, prodf + b * 6 * P, 6);
        }
    }
    launch_unfused_tau(G, velf, prodf, ubar, tau, nbatch, s);
    launch_strain_pi(G, CGEF_F64, ubar, tau, pi64, rho0, nbatch, w.partials, w.npartials, s);
    launch_cast(dtype, pi64, Pi, P * nbatch, s);
    if (E_out) fetch_energy(w, nbatch, strain_blocks(G), g.active_area, E_out, s);
    // leave the local-frame ū in the workspace's plane layout for export
    CGEF_HIP(hipMemcpyAsync(w.ubar, ubar, (size_t)P * nbatch * 3 * sizeof(double), hipMemcpyDeviceToDevice, s));
}

}  // namespace cgef

using namespace cgef;

extern "C" {

const char *cgef_version(void) { return CGEF_VERSION_STRING; }
const char *cgef_last_error(void) { return g_last_error.c_str(); }

cgef_status cgef_grid_create(const cgef_grid_desc *desc, int32_t device, cgef_grid **out) {
    return guard([&] {
        if (!desc || !out) fail(CGEF_ERR_ARGUMENT, "null descriptor or output pointer");
        *out = nullptr;
        if (desc->ndim != 2 && desc->ndim != 3) fail(CGEF_ERR_ARGUMENT, "ndim must be 2 or 3");
        if (desc->geometry != CGEF_CARTESIAN && desc->geometry != CGEF_SPHERICAL)
            fail(CGEF_ERR_ARGUMENT, "unknown geometry");
        if (desc->geometry == CGEF_SPHERICAL && desc->ndim != 2)
            fail(CGEF_ERR_UNSUPPORTED, "spherical grids are 2-D (lon, lat) here");
        if (desc->geometry == CGEF_SPHERICAL && !(desc->radius > 0)) fail(CGEF_ERR_ARGUMENT, "sphere radius must be > 0");
        CGEF_HIP(hipSetDevice(device));
        Grid *g = new Grid();
        try {
            g->device = device;
            g->ndim = desc->ndim;
            g->geometry = desc->geometry;
            g->R = desc->geometry == CGEF_SPHERICAL ? desc->radius : 1.0;
            for (int d = 0; d < 3; d++) {
                g->n[d] = d < g->ndim ? desc->n[d] : 1;
                if (g->n[d] < 1) fail(CGEF_ERR_DIMENSION, "every axis needs at least one point");
                if (d < g->ndim) {
                    if (!desc->coord[d]) fail(CGEF_ERR_ARGUMENT, "null axis coordinates");
                    g->coord[d].assign(desc->coord[d], desc->coord[d] + g->n[d]);
                    g->uniform[d] = desc->uniform[d] ? 1 : 0;
                    g->periodic[d] = desc->periodic[d] ? 1 : 0;
                    g->period[d] = desc->period[d];
                    if (g->periodic[d] && !(g->period[d] > 0)) fail(CGEF_ERR_ARGUMENT, "a periodic axis needs period > 0");
                } else {
                    g->coord[d].assign(1, 0.0);
                }
                // cell widths: constant step on a uniform axis; midpoint widths (half cells at
                // non-periodic ends) on a non-uniform one
                const int64_t n = g->n[d];
                const auto &x = g->coord[d];
                g->width[d].assign(n, 1.0);
                if (n >= 2) {
                    for (int64_t i = 0; i < n; i++) {
                        double wdt;
                        if (g->uniform[d]) wdt = std::fabs(x[1] - x[0]);
                        else if (g->periodic[d]) {
                            double lo = i == 0 ? x[n - 1] - g->period[d] : x[i - 1];
                            double hi = i == n - 1 ? x[0] + g->period[d] : x[i + 1];
                            wdt = std::fabs(hi - lo) / 2.0;
                        } else if (i == 0) wdt = std::fabs(x[1] - x[0]) / 2.0;
                        else if (i == n - 1) wdt = std::fabs(x[n - 1] - x[n - 2]) / 2.0;
                        else wdt = std::fabs(x[i + 1] - x[i - 1]) / 2.0;
                        g->width[d][i] = wdt;
                    }
                }
            }
            g->npts = g->n[0] * g->n[1] * g->n[2];
            g->nrows = g->n[1] * g->n[2];
            std::vector<uint8_t> m(g->npts, 1);
            if (desc->mask) {
                g->has_mask = true;
                for (int64_t t = 0; t < g->npts; t++) m[t] = desc->mask[t] ? 1 : 0;
                g->mask_h = m;
            }
            double tot = 0.0;  // Diagnostics.jl:1051-1059, i fastest
            for (int64_t k = 0; k < g->n[2]; k++)
                for (int64_t j = 0; j < g->n[1]; j++)
                    for (int64_t i = 0; i < g->n[0]; i++)
                        if (m[i + g->n[0] * (j + g->n[1] * k)]) tot += g->cell_area(i, j, k);
            g->active_area = tot;
            g->mask_d = upload(m);
            g->scode_d = upload(stencil_codes(*g, m));
            for (int d = 0; d < 3; d++) {
                g->coord_d[d] = upload(g->coord[d]);
                g->width_d[d] = upload(g->width[d]);
                g->dcoef_d[d] = upload(deriv_table(*g, d));
            }
            if (g->geometry == CGEF_SPHERICAL) {
                const int64_t Nx = g->n[0], Ny = g->n[1];
                std::vector<double> tr(2 * Nx + 4 * Ny);
                for (int64_t i = 0; i < Nx; i++) {
                    tr[i] = std::sin(g->coord[0][i]);
                    tr[Nx + i] = std::cos(g->coord[0][i]);
                }
                for (int64_t j = 0; j < Ny; j++) {
                    tr[2 * Nx + j] = std::sin(g->coord[1][j]);
                    tr[2 * Nx + Ny + j] = std::cos(g->coord[1][j]);
                    // the strain's metric factors of the row: 1/(R cosφ) and tanφ/R (0 at the poles)
                    const double sp = tr[2 * Nx + j], cp = tr[2 * Nx + Ny + j];
                    tr[2 * Nx + 2 * Ny + j] = std::fabs(cp) < std::sqrt(DBL_EPSILON) ? 0.0 : 1.0 / (g->R * cp);
                    tr[2 * Nx + 3 * Ny + j] = std::fabs(cp) > 1e-12 ? sp / (g->R * cp) : 0.0;
                }
                g->trig_d = upload(tr);
            }
        } catch (...) {
            delete g;
            throw;
        }
        *out = reinterpret_cast<cgef_grid *>(g);
    });
}

cgef_status cgef_grid_destroy(cgef_grid *grid) {
    return guard([&] {
        Grid *g = reinterpret_cast<Grid *>(grid);
        if (g && --g->refs == 0) delete g;
    });
}

cgef_status cgef_active_area(const cgef_grid *grid, double *out) {
    return guard([&] {
        const Grid *g = reinterpret_cast<const Grid *>(grid);
        if (!g || !out) fail(CGEF_ERR_ARGUMENT, "null grid or output");
        if (!(g->active_area > 0)) fail(CGEF_ERR_ARGUMENT, "grid has no active cells (all masked out)");
        *out = g->active_area;
    });
}

cgef_status cgef_plan_create(cgef_grid *grid, const cgef_kernel_desc *kernel, double scale, int32_t mask_strategy,
                             int32_t engine, cgef_plan **out) {
    return guard([&] {
        Grid *g = reinterpret_cast<Grid *>(grid);
        if (!g || !kernel || !out) fail(CGEF_ERR_ARGUMENT, "null grid, kernel or output pointer");
        *out = nullptr;
        // Filtering.jl:2917-2922
        if (!(std::isfinite(scale) && scale > 0))
            fail(CGEF_ERR_ARGUMENT, "filter scale must be finite and positive, got " + std::to_string(scale));
        if (kernel->kind != CGEF_TOPHAT && kernel->kind != CGEF_GAUSSIAN) fail(CGEF_ERR_ARGUMENT, "unknown kernel kind");
        if (kernel->kind == CGEF_GAUSSIAN && !(kernel->alpha > 0)) fail(CGEF_ERR_ARGUMENT, "Gaussian alpha must be > 0");
        if (mask_strategy != CGEF_ZEROFILL && mask_strategy != CGEF_DEFORMABLE)
            fail(CGEF_ERR_ARGUMENT, "mask_strategy must be ZeroFill or Deformable");
        CGEF_HIP(hipSetDevice(g->device));
        Plan *p = new Plan();
        try {
            p->grid = g;
            g->refs++;
            p->kind = kernel->kind;
            p->alpha = kernel->alpha;
            p->scale = scale;
            p->strategy = mask_strategy;
            // Kernels.jl:354-359
            p->rad = p->kind == CGEF_TOPHAT ? scale / 2.0 : scale * std::sqrt(-std::log(1e-10) / p->alpha);
            int eng = engine;
            bool sep = false;
            if (eng == CGEF_ENGINE_AUTO) {
                // the reference's dispatch: a top-hat on a 2-D grid takes the prefix-sum engine
                // (Filtering.jl:1115-1141), a Gaussian on a uniform 2-D Cartesian grid the separable
                // engine (Filtering.jl:1533-1540), everything else the banded footprint — evaluated
                // by the ring engine where x is periodic and factorable
                sep = g->ndim == 2 && g->geometry == CGEF_CARTESIAN && p->kind == CGEF_GAUSSIAN && g->uniform[0] &&
                      g->uniform[1];
                if (prefix_supported(*g, p->kind)) eng = CGEF_ENGINE_PREFIX;
            }
            if (eng != CGEF_ENGINE_AUTO && eng != CGEF_ENGINE_BANDED && eng != CGEF_ENGINE_SEPARABLE &&
                eng != CGEF_ENGINE_BASELINE && eng != CGEF_ENGINE_RING && eng != CGEF_ENGINE_PREFIX &&
                eng != CGEF_ENGINE_SPECTRAL)
                fail(CGEF_ERR_ARGUMENT, "unknown engine");
            if (eng == CGEF_ENGINE_SPECTRAL) {
                p->engine = CGEF_ENGINE_SPECTRAL;
                build_spectral(*p, geom_args(*g), nullptr);
            } else if (eng == CGEF_ENGINE_PREFIX) {
                p->engine = CGEF_ENGINE_PREFIX;
                build_prefix(*p);
            } else if (eng == CGEF_ENGINE_SEPARABLE || sep) {
                p->engine = CGEF_ENGINE_SEPARABLE;
                build_separable(*p);
            } else {
                const bool auto_eng = eng == CGEF_ENGINE_AUTO;
                if (auto_eng) eng = ring_grid_ok(*g) ? CGEF_ENGINE_RING : CGEF_ENGINE_BANDED;
                build_banded(*p, eng != CGEF_ENGINE_RING);
                std::string why;
                if (eng == CGEF_ENGINE_RING && !ring_supported(*p, &why)) {
                    if (!auto_eng) fail(CGEF_ERR_UNSUPPORTED, why);
                    // AUTO: the same footprint by the fused direct sum (its FP32 tables are needed now)
                    eng = CGEF_ENGINE_BANDED;
                    build_banded(*p, true);
                }
                p->engine = eng;
                p->band_of_row_d = upload(p->band_of_row);
                p->D64_row_d = upload(p->D64_row);
                if (eng == CGEF_ENGINE_RING) {
                    // Ŵ is built on the device from the unsplit runs; the direct-sum tables are then dropped
                    p->runs_d = upload(p->runs);
                    p->w64_d = upload(p->w64);
                    build_ring(*p, nullptr);
                    dfree(p->runs_d);
                    dfree(p->w64_d);
                } else {
                    split_wide_runs(*p);
                    p->runs_d = upload(p->runs);
                    p->run_ptr_d = upload(p->run_ptr);
                    p->row_order_d = upload(p->row_order);
                    p->w64_d = upload(p->w64);
                    p->w32_d = upload(p->w32);
                    p->w32d_d = upload(p->w32d);
                    p->D32_row_d = upload(p->D32_row);
                    if (eng == CGEF_ENGINE_BASELINE) build_csr(*p);
                }
                int mr = 1;
                for (const Run &r : p->runs) mr = std::max(mr, r.hi - r.lo + 1);
                p->max_run = mr;
                // the host copies of the weight tables are not needed once resident
                std::vector<double>().swap(p->w64);
                std::vector<double>().swap(p->w32d);
                std::vector<float>().swap(p->w32);
            }
        } catch (...) {
            delete p;
            throw;
        }
        *out = reinterpret_cast<cgef_plan *>(p);
    });
}

cgef_status cgef_plan_destroy(cgef_plan *plan) {
    return guard([&] { delete reinterpret_cast<Plan *>(plan); });
}

cgef_status cgef_plan_query(const cgef_plan *plan, cgef_plan_info *info) {
    return guard([&] {
        const Plan *p = reinterpret_cast<const Plan *>(plan);
        if (!p || !info) fail(CGEF_ERR_ARGUMENT, "null plan or info");
        info->engine = p->engine;
        info->nbands = p->nbands;
        info->nentries = p->nentries;
        info->nruns = (int64_t)p->runs.size();
        info->mean_taps = p->mean_taps;
        info->max_taps = p->max_taps;
        info->active_taps = p->active_taps;
        info->radius = p->rad;
        info->ring_offsets = p->ring.ok ? p->ring.args.noff : 0;
        info->ring_modes = p->ring.ok ? p->ring.mh : 0;
        info->ring_modes_kept = p->ring.ok && !p->ring.spectral ? p->ring.modes_kept : 0.0;
        info->ring_pi_fields = !p->ring.ok ? 0 : p->ring.spectral ? 5 : ring_nfields(*p->grid, ring_withmask(*p->grid, p->strategy));
        info->ring_contract_kernel = -1;
        if (p->ring.ok && !p->ring.spectral) {
            std::lock_guard<std::mutex> lk(p->ring.tune->mu);
            int64_t widest = -1;
            for (const auto &c : p->ring.tune->choice)
                if (c.first > widest) {
                    widest = c.first;
                    info->ring_contract_kernel = c.second;
                }
        }
    });
}

cgef_status cgef_workspace_create(const cgef_grid *grid, int64_t nbatch, int32_t dtype, cgef_workspace **out) {
    return guard([&] {
        Grid *g = const_cast<Grid *>(reinterpret_cast<const Grid *>(grid));
        if (!g || !out) fail(CGEF_ERR_ARGUMENT, "null grid or output");
        if (nbatch < 1) fail(CGEF_ERR_DIMENSION, "nbatch must be >= 1");
        check_dtype(dtype);
        CGEF_HIP(hipSetDevice(g->device));
        Workspace *w = new Workspace();
        w->grid = g;
        try {
            ws_reserve(*w, nbatch, dtype);
        } catch (...) {
            delete w;
            throw;
        }
        *out = reinterpret_cast<cgef_workspace *>(w);
    });
}

cgef_status cgef_workspace_destroy(cgef_workspace *ws) {
    return guard([&] { delete reinterpret_cast<Workspace *>(ws); });
}

cgef_status cgef_filter_apply(cgef_plan *plan, void *out, const void *in, int64_t nfields, int64_t nbatch, int32_t dtype,
                              void *stream) {
    return guard([&] {
        Plan *p = reinterpret_cast<Plan *>(plan);
        if (!p || !out || !in) fail(CGEF_ERR_ARGUMENT, "null plan or field pointer");
        if (nfields < 1 || nbatch < 1) fail(CGEF_ERR_DIMENSION, "nfields and nbatch must be >= 1");
        check_dtype(dtype);
        Grid &g = *p->grid;
        CGEF_HIP(hipSetDevice(g.device));
        const GeomArgs G = geom_args(g);
        const int64_t nsl = nfields * nbatch;
        hipStream_t s = S(stream);
        if (p->engine == CGEF_ENGINE_BANDED) {
            launch_filter_lin(G, band_args(*p, dtype == CGEF_F32), dtype, g.has_mask, p->edge_x, in, out, nsl, s);
        } else if (p->engine == CGEF_ENGINE_RING) {
            ring_filter_apply(*p, resolve_ws(g, nullptr, 1, dtype), dtype, in, out, nsl, s);
        } else if (p->engine == CGEF_ENGINE_SEPARABLE) {
            Workspace &w = resolve_ws(g, nullptr, 1, dtype);
            double *rp = (double *)ws_fbuf(w, (size_t)g.npts * nsl * sizeof(double));
            launch_separable(G, *p, dtype, in, out, nsl, rp, s);
        } else if (p->engine == CGEF_ENGINE_PREFIX) {
            Workspace &w = resolve_ws(g, nullptr, 1, dtype);
            launch_prefix(G, *p, dtype, in, out, nsl, (double *)ws_fbuf(w, prefix_bytes(g, nsl)), s);
        } else if (p->engine == CGEF_ENGINE_SPECTRAL) {
            spectral_filter_apply(*p, resolve_ws(g, nullptr, 1, dtype), dtype, in, out, nsl, s);
        } else {
            if (dtype != CGEF_F64) fail(CGEF_ERR_ARGUMENT, "the baseline engine filters FP64 fields");
            launch_baseline_filter(G, *p, (const double *)in, (double *)out, nsl, s);
        }
        CGEF_HIP(hipGetLastError());
    });
}

cgef_status cgef_compute_pi(cgef_plan *plan, cgef_workspace *ws, void *Pi, const void *u, const void *v, const void *w,
                            int64_t nbatch, int32_t dtype, double rho0, double *E_out, void *ubar, void *vbar, void *wbar,
                            void *stream) {
    return guard([&] {
        Plan *p = reinterpret_cast<Plan *>(plan);
        if (!p || !Pi || !u || !v) fail(CGEF_ERR_ARGUMENT, "null plan or field pointer");
        if (nbatch < 1) fail(CGEF_ERR_DIMENSION, "nbatch must be >= 1");
        check_dtype(dtype);
        Grid &g = *p->grid;
        if (g.ndim == 3 && !w) fail(CGEF_ERR_ARGUMENT, "3-D Π needs the vertical velocity w");
        if (g.ndim == 2 && w)
            fail(CGEF_ERR_UNSUPPORTED, "the 2.5-D (u, v, w) layer Π is not provided by this backend; pass w = NULL");
        CGEF_HIP(hipSetDevice(g.device));
        hipStream_t s = S(stream);
        Workspace &W = resolve_ws(g, ws, nbatch, dtype);
        const GeomArgs G = geom_args(g);
        if (p->engine == CGEF_ENGINE_BANDED) {
            launch_mean(G, dtype, u, v, w, nbatch, W.partials, W.npartials, W.cmean, s);
            launch_prep(G, dtype, u, v, w, nbatch, W.cmean, W.qpack, s);
            p->prof_begin(s);
            launch_pi_filter(G, band_args(*p, dtype == CGEF_F32), dtype, g.has_mask, p->edge_x, W.qpack, W.cmean, W.ubar,
                             W.tau, nbatch, s);
            p->prof_end(s);
            launch_strain_pi(G, dtype, W.ubar, W.tau, Pi, rho0, nbatch, W.partials, W.npartials, s);
            CGEF_HIP(hipGetLastError());
            if (E_out) fetch_energy(W, nbatch, strain_blocks(G), g.active_area, E_out, s);
        } else if (p->engine == CGEF_ENGINE_RING) {
            const int wm = ring_withmask(g, p->strategy), nf = ring_nfields(g, wm);
            ring_reserve(W, *p, nf * nbatch);
            ring_prep(G, dtype, u, v, w, nf, wm, W.rxy, nbatch, s);
            ring_forward(p->ring, W.rxy, W.rxf, nf * nbatch, s);
            const bool fused = ring_pi_fused(*p, G, nf, wm, nbatch);
            p->prof_begin(s);
            ring_contract(p->ring, W.rxf, W.rout, nf * nbatch, s, fused, ring_mode_cut(p->ring, dtype));
            p->prof_end(s);
            ring_pi_inverse_epilogue(*p, W, G, dtype, nf, wm, ubar, vbar, wbar, nbatch, fused, s);
            launch_strain_pi(G, dtype, W.ubar, W.tau, Pi, rho0, nbatch, W.partials, W.npartials, s);
            CGEF_HIP(hipGetLastError());
            if (E_out) fetch_energy(W, nbatch, strain_blocks(G), g.active_area, E_out, s);
        } else if (p->engine == CGEF_ENGINE_SEPARABLE) {
            double *rp = (double *)ws_fbuf(W, (size_t)5 * g.npts * nbatch * sizeof(double));
            p->prof_begin(s);
            launch_separable_pi(G, *p, dtype, u, v, nbatch, rp, W.ubar, W.tau, ubar, vbar, wbar, s);
            p->prof_end(s);
            launch_strain_pi(G, dtype, W.ubar, W.tau, Pi, rho0, nbatch, W.partials, W.npartials, s);
            CGEF_HIP(hipGetLastError());
            if (E_out) fetch_energy(W, nbatch, strain_blocks(G), g.active_area, E_out, s);
        } else if (p->engine == CGEF_ENGINE_PREFIX) {
            const int nf = g.ndim == 2 && g.geometry == CGEF_CARTESIAN ? 5 : 9;
            double *pb = (double *)ws_fbuf(W, prefix_bytes(g, nf * nbatch));
            p->prof_begin(s);
            launch_prefix_pi(G, *p, dtype, u, v, w, nbatch, pb, W.ubar, W.tau, ubar, vbar, wbar, s);
            p->prof_end(s);
            launch_strain_pi(G, dtype, W.ubar, W.tau, Pi, rho0, nbatch, W.partials, W.npartials, s);
            CGEF_HIP(hipGetLastError());
            if (E_out) fetch_energy(W, nbatch, strain_blocks(G), g.active_area, E_out, s);
        } else if (p->engine == CGEF_ENGINE_SPECTRAL) {
            ring_reserve(W, *p, 5 * nbatch);
            ring_prep(G, dtype, u, v, w, 5, 0, W.rxy, nbatch, s);
            spectral_filter_planes(*p, W, 5 * nbatch, s);
            ring_spectral_epilogue(G, dtype, W.rxy, p->ring.invrenorm_d, W.ubar, W.tau, ubar, vbar, wbar, nbatch, s);
            launch_strain_pi(G, dtype, W.ubar, W.tau, Pi, rho0, nbatch, W.partials, W.npartials, s);
            CGEF_HIP(hipGetLastError());
            if (E_out) fetch_energy(W, nbatch, strain_blocks(G), g.active_area, E_out, s);
        } else {
            compute_pi_unfused(*p, W, Pi, u, v, w, nbatch, dtype, rho0, E_out, s);
        }
        if (p->engine == CGEF_ENGINE_BANDED || p->engine == CGEF_ENGINE_BASELINE) {  // the others' epilogues write them
            if (ubar) export_planes(g, W.ubar, 0, ubar, nbatch, dtype, s);
            if (vbar) export_planes(g, W.ubar, 1, vbar, nbatch, dtype, s);
            if (wbar) export_planes(g, W.ubar, 2, wbar, nbatch, dtype, s);
        }
        CGEF_HIP(hipGetLastError());
    });
}

cgef_status cgef_coarse_grain(cgef_plan *const *plans, int64_t nscales, cgef_workspace *ws, void *Pi, const void *u,
                              const void *v, const void *w, int64_t nbatch, int32_t dtype, double rho0, double *E_out,
                              void *stream) {
    return guard([&] {
        if (!plans || !Pi || !u || !v) fail(CGEF_ERR_ARGUMENT, "null plans or field pointer");
        if (nscales < 0 || nbatch < 1) fail(CGEF_ERR_DIMENSION, "bad nscales / nbatch");
        check_dtype(dtype);
        if (nscales == 0) return;
        Grid &g = *reinterpret_cast<Plan *>(plans[0])->grid;
        bool all_ring = true, all_spec = true;
        for (int64_t q = 0; q < nscales; q++) {
            Plan *p = reinterpret_cast<Plan *>(plans[q]);
            if (!p || p->grid != &g) fail(CGEF_ERR_ARGUMENT, "every plan must belong to the same grid");
            all_ring = all_ring && p->engine == CGEF_ENGINE_RING;
            all_spec = all_spec && p->engine == CGEF_ENGINE_SPECTRAL;
        }
        if (g.ndim == 3 && !w) fail(CGEF_ERR_ARGUMENT, "3-D Π needs the vertical velocity w");
        if (g.ndim == 2 && w)
            fail(CGEF_ERR_UNSUPPORTED, "the 2.5-D (u, v, w) layer Π is not provided by this backend; pass w = NULL");
        CGEF_HIP(hipSetDevice(g.device));
        hipStream_t s = S(stream);
        Workspace &W = resolve_ws(g, ws, nbatch, dtype);
        const GeomArgs G = geom_args(g);
        const int64_t P = g.npts;
        const size_t es = esize(dtype);
        const int64_t nblk = strain_blocks(G);
        // E(ℓ) per scale and slice, then each scale's partial sums in a slice of their own: one energy
        // reduction after the last scale instead of one launch per scale
        double *Eall = (double *)ws_fbuf(W, (size_t)nscales * nbatch * (1 + nblk) * sizeof(double));
        double *Pall = Eall + nscales * nbatch;
        if (all_ring || all_spec) {
            Plan &p0 = *reinterpret_cast<Plan *>(plans[0]);
            int wm = 0;
            if (all_ring)
                for (int64_t q = 0; q < nscales; q++) wm |= ring_withmask(g, reinterpret_cast<Plan *>(plans[q])->strategy);
            const int nf = all_ring ? ring_nfields(g, wm) : 5;
            ring_reserve(W, p0, nf * nbatch);
            auto sweep = [&](hipStream_t st) {
                if (all_ring) {
                    // one forward transform for every scale (Diagnostics.jl:283-301's analyze-once idea,
                    // here for the real-space footprint): only the contraction and the inverse depend on ℓ
                    ring_prep(G, dtype, u, v, w, nf, wm, W.rxy, nbatch, st);
                    ring_forward(p0.ring, W.rxy, W.rxf, nf * nbatch, st);
                    for (int64_t q = 0; q < nscales; q++) {
                        Plan &p = *reinterpret_cast<Plan *>(plans[q]);
                        const bool fused = ring_pi_fused(p, G, nf, wm, nbatch);
                        p.prof_begin(st);
                        ring_contract(p.ring, W.rxf, W.rout, nf * nbatch, st, fused, ring_mode_cut(p.ring, dtype));
                        p.prof_end(st);
                        ring_pi_inverse_epilogue(p, W, G, dtype, nf, wm, nullptr, nullptr, nullptr, nbatch, fused, st);
                        launch_strain_pi(G, dtype, W.ubar, W.tau, (char *)Pi + (size_t)q * P * es, rho0, nbatch,
                                         Pall + q * nbatch * nblk, nbatch * nblk, st, nscales * P);
                    }
                } else {
                    // the spectrum of the masked inputs is taken once and reused by every scale
                    // (FFTWExt.jl:100-120 filter_analyze! / filter_synthesize!)
                    ring_prep(G, dtype, u, v, w, 5, 0, W.rxy, nbatch, st);
                    ring_forward(p0.ring, W.rxy, W.rxf, 5 * nbatch, st);
                    ring_spectral_analyze_cols(p0.ring, W.rxf, W.rout, 5 * nbatch, st);
                    for (int64_t q = 0; q < nscales; q++) {
                        Plan &p = *reinterpret_cast<Plan *>(plans[q]);
                        p.prof_begin(st);
                        ring_spectral_synth_cols(p.ring, W.rout, W.rxf, 5 * nbatch, st);
                        p.prof_end(st);
                        ring_inverse(p.ring, W.rxf, W.rxy, 5 * nbatch, st);
                        ring_spectral_epilogue(G, dtype, W.rxy, p.ring.invrenorm_d, W.ubar, W.tau, nullptr, nullptr,
                                               nullptr, nbatch, st);
                        launch_strain_pi(G, dtype, W.ubar, W.tau, (char *)Pi + (size_t)q * P * es, rho0, nbatch,
                                         Pall + q * nbatch * nblk, nbatch * nblk, st, nscales * P);
                    }
                }
                if (E_out) launch_energy_final(nscales * nbatch, nblk, Pall, g.active_area, Eall, st);
                CGEF_HIP(hipGetLastError());
            };
            // Optionally the sweep as one HIP graph (CGEF_GRAPH=1): the first sweep with a given set
            // of arguments runs eagerly (and settles the contraction's measured variants), the second
            // is captured on the workspace's own stream and launched, later ones replay it. A graph
            // holds no memset node. Off by default: on MI355X the replayed 8-snapshot × 24-scale
            // sweep on the 0.25° sphere measures no faster than launched eagerly (12.44–12.60 ms
            // against 12.29–12.44 ms; its launches run 5 µs–0.3 ms each). Plans being profiled
            // always run eagerly. Calls that share a workspace must not overlap (as
            // for every entry point using it).
            bool profiling = false;
            for (int64_t q = 0; q < nscales; q++) profiling = profiling || reinterpret_cast<Plan *>(plans[q])->profiling;
            const char *ge = std::getenv("CGEF_GRAPH");
            const bool use_graph = (ge && ge[0] == '1') && !profiling && !W.graph_broken;
            std::vector<uint64_t> key;
            if (use_graph) {
                auto u64 = [](const void *ptr) { return (uint64_t)(uintptr_t)ptr; };
                uint64_t rb;
                std::memcpy(&rb, &rho0, sizeof rb);
                key = {(uint64_t)all_ring, (uint64_t)nscales, u64(Pi), u64(u), u64(v), u64(w), (uint64_t)nbatch,
                       (uint64_t)dtype, rb, (uint64_t)(E_out != nullptr), u64(W.rxy), u64(W.rxf), u64(W.rout),
                       u64(W.ubar), u64(W.tau), u64(W.partials), u64(Eall)};
                for (int64_t q = 0; q < nscales; q++) key.push_back(reinterpret_cast<Plan *>(plans[q])->uid);
            }
            if (use_graph && W.graph_exec && key == W.graph_key) {
                CGEF_HIP(hipGraphLaunch(W.graph_exec, s));
            } else if (use_graph && key == W.seen_key) {
                if (!W.cap_stream) CGEF_HIP(hipStreamCreateWithFlags(&W.cap_stream, hipStreamNonBlocking));
                hipGraph_t graph = nullptr;
                hipGraphExec_t exec = nullptr;
                bool ok = hipStreamBeginCapture(W.cap_stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
                if (ok) {
                    try {
                        sweep(W.cap_stream);
                    } catch (const Error &) {
                        ok = false;
                    }
                    ok = hipStreamEndCapture(W.cap_stream, &graph) == hipSuccess && ok && graph != nullptr;
                }
                if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
                if (graph) (void)hipGraphDestroy(graph);
                if (ok) {
                    if (W.graph_exec) (void)hipGraphExecDestroy(W.graph_exec);
                    W.graph_exec = exec;
                    W.graph_key = key;
                    CGEF_HIP(hipGraphLaunch(W.graph_exec, s));
                } else {
                    (void)hipGetLastError();
                    W.graph_broken = true;   // capture not possible here: stay eager for this workspace
                    sweep(s);
                }
            } else {
                sweep(s);
                if (use_graph) W.seen_key = key;
            }
        } else {
            // other engines: one compute_Π! per scale, written into the scale's strided slots — directly
            // for one slice, else through the workspace's Π staging buffer (allocated on first use and
            // kept, as every workspace buffer is)
            void *pis = nullptr;
            if (nbatch > 1) {
                if (W.npistage < (size_t)nbatch * P * es) {
                    dfree(W.pistage);
                    CGEF_HIP(hipMalloc(&W.pistage, (size_t)nbatch * P * es));
                    W.npistage = (size_t)nbatch * P * es;
                }
                pis = W.pistage;
            }
            std::vector<double> Eq(nbatch);
            for (int64_t q = 0; q < nscales; q++) {
                void *dst = nbatch > 1 ? pis : (void *)((char *)Pi + (size_t)q * P * es);
                cgef_status st = cgef_compute_pi(plans[q], reinterpret_cast<cgef_workspace *>(&W), dst, u, v, w, nbatch,
                                                 dtype, rho0, E_out ? Eq.data() : nullptr, nullptr, nullptr, nullptr, stream);
                if (st != CGEF_OK) fail(st, g_last_error);
                if (nbatch > 1)
                    CGEF_HIP(hipMemcpy2DAsync((char *)Pi + (size_t)q * P * es, (size_t)nscales * P * es, pis,
                                              (size_t)P * es, (size_t)P * es, (size_t)nbatch, hipMemcpyDeviceToDevice, s));
                if (E_out)
                    for (int64_t b = 0; b < nbatch; b++) E_out[q + nscales * b] = Eq[b];
            }
            return;
        }
        if (E_out) {
            std::vector<double> Eh((size_t)nscales * nbatch);
            CGEF_HIP(hipMemcpyAsync(Eh.data(), Eall, Eh.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            CGEF_HIP(hipStreamSynchronize(s));
            if (!(g.active_area > 0)) fail(CGEF_ERR_ARGUMENT, "grid has no active cells (all masked out)");
            for (int64_t q = 0; q < nscales; q++)
                for (int64_t b = 0; b < nbatch; b++) E_out[q + nscales * b] = Eh[q * nbatch + b];
        }
    });
}

cgef_status cgef_derivative(const cgef_grid *grid, int32_t axis, void *out, const void *in, int64_t nbatch, int32_t dtype,
                            void *stream) {
    return guard([&] {
        const Grid *g = reinterpret_cast<const Grid *>(grid);
        if (!g || !out || !in) fail(CGEF_ERR_ARGUMENT, "null grid or field pointer");
        if (axis < 0 || axis >= g->ndim) fail(CGEF_ERR_ARGUMENT, "axis out of range for this grid");
        if (nbatch < 1) fail(CGEF_ERR_DIMENSION, "nbatch must be >= 1");
        check_dtype(dtype);
        CGEF_HIP(hipSetDevice(g->device));
        launch_derivative(geom_args(*g), axis, dtype, in, out, nbatch, S(stream));
        CGEF_HIP(hipGetLastError());
    });
}

cgef_status cgef_energy(const cgef_grid *grid, const void *ubar, const void *vbar, const void *wbar, int64_t nbatch,
                        int32_t dtype, double *E_out, void *stream) {
    return guard([&] {
        Grid *g = const_cast<Grid *>(reinterpret_cast<const Grid *>(grid));
        if (!g || !ubar || !vbar || !E_out) fail(CGEF_ERR_ARGUMENT, "null grid, field or output pointer");
        if (nbatch < 1) fail(CGEF_ERR_DIMENSION, "nbatch must be >= 1");
        check_dtype(dtype);
        CGEF_HIP(hipSetDevice(g->device));
        hipStream_t s = S(stream);
        Workspace &W = resolve_ws(*g, nullptr, nbatch, dtype);
        const GeomArgs G = geom_args(*g);
        launch_energy_fields(G, dtype, ubar, vbar, wbar, nbatch, W.partials, W.npartials, s);
        fetch_energy(W, nbatch, energy_blocks(G), g->active_area, E_out, s);
    });
}

cgef_status cgef_cumulative_energy(cgef_plan *const *plans, int64_t nscales, cgef_workspace *ws, const void *u,
                                   const void *v, const void *w, int64_t nbatch, int32_t dtype, double *E_out,
                                   void *stream) {
    return guard([&] {
        if (!plans || !u || !v || !E_out) fail(CGEF_ERR_ARGUMENT, "null plans, field or output pointer");
        if (nscales < 0 || nbatch < 1) fail(CGEF_ERR_DIMENSION, "bad nscales / nbatch");
        check_dtype(dtype);
        if (nscales == 0) return;
        Grid &g = *reinterpret_cast<Plan *>(plans[0])->grid;
        for (int64_t q = 0; q < nscales; q++)
            if (!plans[q] || reinterpret_cast<Plan *>(plans[q])->grid != &g)
                fail(CGEF_ERR_ARGUMENT, "every plan must belong to the same grid");
        if (!(g.active_area > 0)) fail(CGEF_ERR_ARGUMENT, "grid has no active cells (all masked out)");
        CGEF_HIP(hipSetDevice(g.device));
        hipStream_t s = S(stream);
        Workspace &W = resolve_ws(g, ws, nbatch, dtype);
        const int64_t P = g.npts;
        const size_t es = esize(dtype);
        char *fb = (char *)ws_fbuf(W, (size_t)P * nbatch * 3 * es + (size_t)P * nbatch * 8);
        void *uf = fb, *vf = fb + (size_t)P * nbatch * es, *wf = w ? fb + 2 * (size_t)P * nbatch * es : nullptr;
        const GeomArgs G = geom_args(g);
        std::vector<double> E(nbatch);
        bool all_spec = true;
        for (int64_t q = 0; q < nscales; q++) all_spec = all_spec && reinterpret_cast<Plan *>(plans[q])->engine == CGEF_ENGINE_SPECTRAL;
        if (all_spec) {
            // analyze the masked velocities once, synthesize per scale (Diagnostics.jl:270-301)
            Plan &p0 = *reinterpret_cast<Plan *>(plans[0]);
            const int64_t nsl = (w ? 3 : 2) * nbatch;
            ring_reserve(W, p0, nsl);
            ring_prep_fields(G, dtype, u, nbatch, 0, W.rxy, s);
            ring_prep_fields(G, dtype, v, nbatch, 0, W.rxy + nbatch * P, s);
            if (w) ring_prep_fields(G, dtype, w, nbatch, 0, W.rxy + 2 * nbatch * P, s);
            ring_forward(p0.ring, W.rxy, W.rxf, nsl, s);
            ring_spectral_analyze_cols(p0.ring, W.rxf, W.rout, nsl, s);
            for (int64_t q = 0; q < nscales; q++) {
                Plan &p = *reinterpret_cast<Plan *>(plans[q]);
                ring_spectral_synth_cols(p.ring, W.rout, W.rxf, nsl, s);
                ring_inverse(p.ring, W.rxf, W.rxy, nsl, s);
                ring_spectral_normalize(G, dtype, W.rxy, p.ring.invrenorm_d, nsl, fb, s);
                launch_energy_fields(G, dtype, uf, vf, wf, nbatch, W.partials, W.npartials, s);
                fetch_energy(W, nbatch, energy_blocks(G), g.active_area, E.data(), s);
                for (int64_t b = 0; b < nbatch; b++) E_out[q + nscales * b] = E[b];
            }
            return;
        }
        for (int64_t q = 0; q < nscales; q++) {
            Plan *p = reinterpret_cast<Plan *>(plans[q]);
            auto filt = [&](const void *in, void *out) {
                if (p->engine == CGEF_ENGINE_BANDED) {
                    launch_filter_lin(G, band_args(*p, dtype == CGEF_F32), dtype, g.has_mask, p->edge_x, in, out, nbatch, s);
                } else if (p->engine == CGEF_ENGINE_RING) {
                    ring_filter_apply(*p, W, dtype, in, out, nbatch, s);
                } else if (p->engine == CGEF_ENGINE_SEPARABLE) {
                    double *rp = ws_scratch64(W, P * nbatch);
                    launch_separable(G, *p, dtype, in, out, nbatch, rp, s);
                } else if (p->engine == CGEF_ENGINE_PREFIX) {
                    double *pb = ws_scratch64(W, (int64_t)(prefix_bytes(g, nbatch) / sizeof(double)));
                    launch_prefix(G, *p, dtype, in, out, nbatch, pb, s);
                } else if (p->engine == CGEF_ENGINE_SPECTRAL) {
                    spectral_filter_apply(*p, W, dtype, in, out, nbatch, s);
                } else {
                    if (dtype != CGEF_F64) fail(CGEF_ERR_ARGUMENT, "the baseline engine filters FP64 fields");
                    launch_baseline_filter(G, *p, (const double *)in, (double *)out, nbatch, s);
                }
            };
            filt(u, uf);
            filt(v, vf);
            if (w) filt(w, wf);
            launch_energy_fields(G, dtype, uf, vf, wf, nbatch, W.partials, W.npartials, s);
            fetch_energy(W, nbatch, energy_blocks(G), g.active_area, E.data(), s);
            for (int64_t b = 0; b < nbatch; b++) E_out[q + nscales * b] = E[b];
        }
    });
}

cgef_status cgef_profile_enable(cgef_plan *plan, int32_t on) {
    return guard([&] {
        Plan *p = reinterpret_cast<Plan *>(plan);
        if (!p) fail(CGEF_ERR_ARGUMENT, "null plan");
        if (on < 0 || on > 2) fail(CGEF_ERR_ARGUMENT, "profile step must be 0, 1 or 2");
        if (on == 2 && p->engine != CGEF_ENGINE_RING) fail(CGEF_ERR_UNSUPPORTED, "step 2 is the ring engine's");
        p->profiling = on;
        p->prof_used = 0;
    });
}

cgef_status cgef_profile_read(cgef_plan *plan, double *total_ms, int64_t *nlaunches) {
    return guard([&] {
        Plan *p = reinterpret_cast<Plan *>(plan);
        if (!p || !total_ms || !nlaunches) fail(CGEF_ERR_ARGUMENT, "null plan or output");
        double tot = 0.0;
        for (size_t q = 0; q < p->prof_used; q++) {
            CGEF_HIP(hipEventSynchronize(p->prof_events[q].second));
            float ms = 0.f;
            CGEF_HIP(hipEventElapsedTime(&ms, p->prof_events[q].first, p->prof_events[q].second));
            tot += ms;
        }
        *total_ms = tot;
        *nlaunches = (int64_t)p->prof_used;
    });
}

cgef_status cgef_plan_mode_cut(const cgef_plan *plan, int32_t *modes, int64_t nrows) {
    return guard([&] {
        const Plan *p = reinterpret_cast<const Plan *>(plan);
        if (!p || !modes) fail(CGEF_ERR_ARGUMENT, "null plan or output");
        if (!p->ring.ok || p->ring.spectral) fail(CGEF_ERR_UNSUPPORTED, "the mode cut belongs to the ring engine");
        const int64_t n = p->grid->nrows;
        if (nrows != n) fail(CGEF_ERR_DIMENSION, "expected " + std::to_string(n) + " rows (Ny·Nz)");
        for (int64_t r = 0; r < n; r++) modes[r] = p->ring.mcut_h.empty() ? (int32_t)p->ring.mh : p->ring.mcut_h[r];
    });
}

cgef_status cgef_spectral_density(int64_t n, const double *C, const double *k, double *g) {
    return guard([&] {
        if (n < 0) fail(CGEF_ERR_DIMENSION, "negative length");
        if (n > 0 && (!C || !k || !g)) fail(CGEF_ERR_ARGUMENT, "null array");
        // Diagnostics.jl:1319-1333
        if (n < 2) {
            for (int64_t i = 0; i < n; i++) g[i] = 0.0;
            return;
        }
        for (int64_t i = 0; i < n; i++) {
            if (i == 0) g[i] = (C[1] - C[0]) / (k[1] - k[0]);
            else if (i == n - 1) g[i] = (C[n - 1] - C[n - 2]) / (k[n - 1] - k[n - 2]);
            else g[i] = (C[i + 1] - C[i - 1]) / (k[i + 1] - k[i - 1]);
        }
    });
}

}  // extern "C"
