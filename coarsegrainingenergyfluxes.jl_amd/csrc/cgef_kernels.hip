// Device code of libcgef_hip.so (gfx950 / CDNA4).
//
// Hot path: k_pi_filter, the fused footprint filter of the (mean-shifted) velocity components and
// their centred quadratic products for Π. One workgroup = one target row × 256 consecutive target
// columns, one target per lane. Per footprint run (one source row) the workgroup stages the source
// segment (velocity components + mask flag, 16 B per cell in FP32) in LDS once, then every lane
// walks the run with wave-uniform loop bounds and wave-uniform (scalar-loaded) weights, so a weight
// costs no VGPR traffic and each staged cell is reused by up to 256 targets.
//
// Precision (FP32 path): pass 1 accumulates Σ w·q in FP64 (v_fma_f64 runs at the FP32 VALU rate on
// CDNA4), so ū carries no accumulation error beyond its FP32 inputs. Pass 2 forms the CENTRED
// products w·(q_a − q̄_a)(q_b − q̄_b) in FP32 (no cancellation: an error in the centre enters τ only
// at second order) and sums them in two levels, FP32 over chunks of 32 taps, FP64 across chunks. τ = F[(q−q̄)(q−q̄)] + (1−m)[…] with m = F[1]
// (ZeroFill near a mask) is algebraically F[q_a q_b] − q̄_a q̄_b, i.e. the reference's τ.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <type_traits>

#include "cgef_pi_point.hpp"

namespace cgef {

template <typename T> struct V4;
template <> struct V4<float> { using type = float4; };
template <> struct V4<double> { using type = double4; };

__device__ __forceinline__ int64_t wrapi(int64_t a, int64_t n) {
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

template <typename T>
__device__ __forceinline__ T ld(const void *p, int64_t i) {
    return static_cast<const T *>(p)[i];
}

// Wave/block reductions in FP64 (deterministic tree order).
__device__ __forceinline__ double warp_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

template <int NT>
__device__ double block_sum(double v, double *sh) {
    v = warp_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int q = 0; q < NT / 64; q++) r += sh[q];
    }
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------ velocity mean (the shift) ----

__device__ __forceinline__ double cell_area(const GeomArgs &G, int64_t i, int64_t j, int64_t k) {
    if (G.mode == 1) {
        const double cphi = G.trig[2 * G.n[0] + G.n[1] + j];
        return G.R * G.R * cphi * G.width[0][i] * G.width[1][j];
    }
    double a = G.width[0][i] * G.width[1][j];
    if (G.mode == 2) a *= G.width[2][k];
    return a;
}

// Velocity in the frame the filter works in: planetary Cartesian on the sphere (Diagnostics.jl:
// 555-573, e_east = (−sinλ, cosλ, 0), e_north = (−sinφcosλ, −sinφsinλ, cosφ)), as given otherwise.
template <typename T>
__device__ __forceinline__ void frame_velocity(const GeomArgs &G, const void *u, const void *v, const void *w,
                                               int64_t off, int64_t t, int64_t i, int64_t j, double U[3]) {
    const double uu = (double)ld<T>(u, off + t), vv = (double)ld<T>(v, off + t);
    if (G.mode == 1) {
        const double sl = G.trig[i], cl = G.trig[G.n[0] + i];
        const double sp = G.trig[2 * G.n[0] + j], cp = G.trig[2 * G.n[0] + G.n[1] + j];
        U[0] = -sl * uu - sp * cl * vv;
        U[1] = cl * uu - sp * sl * vv;
        U[2] = cp * vv;
    } else {
        U[0] = uu;
        U[1] = vv;
        U[2] = (G.mode == 2) ? (double)ld<T>(w, off + t) : 0.0;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_mean_partial(GeomArgs G, const void *u, const void *v, const void *w,
                                                      double *partials, int64_t nblk) {
    __shared__ double sh[4];
    const int64_t b = blockIdx.y;
    double acc[3] = {0.0, 0.0, 0.0};
    const int64_t Nx = G.n[0], Ny = G.n[1];
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < G.npts; t += (int64_t)gridDim.x * 256) {
        if (!G.mask[t]) continue;
        const int64_t i = t % Nx, j = (t / Nx) % Ny, k = t / (Nx * Ny);
        double U[3];
        frame_velocity<T>(G, u, v, w, b * G.npts, t, i, j, U);
        const double A = cell_area(G, i, j, k);
        for (int a = 0; a < 3; a++) acc[a] += A * U[a];
    }
    for (int a = 0; a < 3; a++) {
        double s = block_sum<256>(acc[a], sh);
        if (threadIdx.x == 0) partials[(b * 3 + a) * nblk + blockIdx.x] = s;
    }
}

__global__ void k_mean_final(const double *partials, int64_t nblk, double inv_area, double *cmean) {
    __shared__ double sh[4];
    const int64_t b = blockIdx.x;
    for (int a = 0; a < 3; a++) {
        double acc = 0.0;
        for (int64_t q = threadIdx.x; q < nblk; q += 256) acc += partials[(b * 3 + a) * nblk + q];
        double s = block_sum<256>(acc, sh);
        if (threadIdx.x == 0) cmean[b * 3 + a] = s * inv_area;
    }
}

void launch_mean(const GeomArgs &G, int dtype, const void *u, const void *v, const void *w, int64_t nbatch,
                 double *partials, int64_t npartials, double *cmean, hipStream_t s) {
    int64_t nblk = (G.npts + 255) / 256;
    if (nblk > 1024) nblk = 1024;
    if (nblk * 3 * nbatch > npartials) nblk = npartials / (3 * nbatch);
    dim3 grid((unsigned)nblk, (unsigned)nbatch);
    if (dtype == CGEF_F32)
        hipLaunchKernelGGL(k_mean_partial<float>, grid, dim3(256), 0, s, G, u, v, w, partials, nblk);
    else
        hipLaunchKernelGGL(k_mean_partial<double>, grid, dim3(256), 0, s, G, u, v, w, partials, nblk);
    hipLaunchKernelGGL(k_mean_final, dim3((unsigned)nbatch), dim3(256), 0, s, partials, nblk, 1.0 / G.active_area, cmean);
}

// ------------------------------------------------------------- pack: shifted, masked components ----

// qpack cell = (q0, q1, q2, flag): q = U − c at active cells, 0 at inactive; flag 0 active, 1 inactive.
template <typename T>
__global__ __launch_bounds__(256) void k_prep(GeomArgs G, const void *u, const void *v, const void *w,
                                              const double *cmean, typename V4<T>::type *qpack) {
    const int64_t b = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= G.npts) return;
    using T4 = typename V4<T>::type;
    T4 o;
    if (G.mask[t]) {
        const int64_t Nx = G.n[0], Ny = G.n[1];
        const int64_t i = t % Nx, j = (t / Nx) % Ny;
        double U[3];
        frame_velocity<T>(G, u, v, w, b * G.npts, t, i, j, U);
        const double *c = cmean + 3 * b;
        o.x = (T)(U[0] - c[0]);
        o.y = (T)(U[1] - c[1]);
        o.z = (T)(U[2] - c[2]);
        o.w = (T)0;
    } else {
        o.x = o.y = o.z = (T)0;
        o.w = (T)1;
    }
    qpack[b * G.npts + t] = o;
}

void launch_prep(const GeomArgs &G, int dtype, const void *u, const void *v, const void *w, int64_t nbatch,
                 const double *cmean, void *qpack, hipStream_t s) {
    dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)nbatch);
    if (dtype == CGEF_F32)
        hipLaunchKernelGGL(k_prep<float>, grid, dim3(256), 0, s, G, u, v, w, cmean, (float4 *)qpack);
    else
        hipLaunchKernelGGL(k_prep<double>, grid, dim3(256), 0, s, G, u, v, w, cmean, (double4 *)qpack);
}

// ----------------------------------------------------------------------- the fused Π filter ----

// Stage source cells g0 .. g0+L-1 of one row (x wrapped on a periodic axis; flag 2 outside).
template <typename T4, typename T>
__device__ __forceinline__ void stage(T4 *seg, const T4 *__restrict__ q, int64_t rowbase, int64_t g0, int L,
                                      int64_t Nx, int px) {
    for (int s = threadIdx.x; s < L; s += TI) {
        int64_t gx = g0 + s;
        T4 val;
        if (gx >= 0 && gx < Nx) {
            val = q[rowbase + gx];
        } else if (px) {
            gx = wrapi(gx, Nx);
            val = q[rowbase + gx];
        } else {
            val.x = val.y = val.z = (T)0;
            val.w = (T)2;
        }
        seg[s] = val;
    }
}

template <typename T, int NC, bool MASKED, bool EDGE_X, bool SPHERE>
__global__ __launch_bounds__(TI) void k_pi_filter(GeomArgs G, BandArgs B, const typename V4<T>::type *__restrict__ qpack,
                                                  const double *__restrict__ cmean, double *__restrict__ ubar,
                                                  T *__restrict__ tau) {
    using T4 = typename V4<T>::type;
    extern __shared__ unsigned char smem_raw[];
    T4 *seg = reinterpret_cast<T4 *>(smem_raw);
    const int tid = threadIdx.x;
    const int64_t Nx = G.n[0], Ny = G.n[1], Nz = G.n[2];
    const int32_t r = B.row_order[blockIdx.y];
    const int64_t j = r % Ny, k = r / Ny;
    const int64_t b = blockIdx.z;
    const int64_t i0 = (int64_t)blockIdx.x * TI;
    const int64_t i = i0 + tid;
    const bool inb = i < Nx;
    const int64_t tpt = i + Nx * (int64_t)r;
    const bool tact = inb && (!MASKED || G.mask[tpt] != 0);
    const T4 *q = qpack + b * G.npts;
    const int band = B.band_of_row[r];
    const int rb = B.run_ptr[band], re = B.run_ptr[band + 1];
    const T *wt = std::is_same<T, float>::value ? reinterpret_cast<const T *>(B.w32) : reinterpret_cast<const T *>(B.wd);

    // ---- pass 1: Σ w q (FP64), Σ w over inactive (mo) and in-range (din) cells
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, mo = 0.0, din = 0.0;
    for (int ru = rb; ru < re; ++ru) {
        const Run R = B.runs[ru];
        int64_t jj = j + R.dy, kk = k + R.dz;
        if (jj < 0 || jj >= Ny) {
            if (!G.periodic[1]) continue;
            jj = wrapi(jj, Ny);
        }
        if (kk < 0 || kk >= Nz) {
            if (!G.periodic[2]) continue;
            kk = wrapi(kk, Nz);
        }
        const int L = TI + R.hi - R.lo;
        __syncthreads();
        stage<T4, T>(seg, q, Nx * (jj + Ny * kk), i0 + R.lo, L, Nx, G.periodic[0]);
        __syncthreads();
        if (tact) {
            const double *wd = B.wd + R.woff;
            const T4 *sp = seg + tid;
            const int n = R.hi - R.lo + 1;
#pragma unroll 4
            for (int e = 0; e < n; ++e) {
                const double w = wd[e];
                const T4 c = sp[e];
                s0 = fma(w, (double)c.x, s0);
                s1 = fma(w, (double)c.y, s1);
                if (NC == 3) s2 = fma(w, (double)c.z, s2);
                if (EDGE_X) {
                    din += (c.w < (T)1.5) ? w : 0.0;
                    mo += (c.w > (T)0.5 && c.w < (T)1.5) ? w : 0.0;
                } else if (MASKED) {
                    mo = fma(w, (double)c.w, mo);
                }
            }
        }
    }
    const double D = EDGE_X ? din : B.D_row[r];
    const double den = (B.strategy == CGEF_DEFORMABLE) ? D - mo : D;
    const bool ok = tact && den > 1e-15;
    const double qb0 = ok ? s0 / den : 0.0, qb1 = ok ? s1 / den : 0.0, qb2 = ok ? s2 / den : 0.0;
    const T qc0 = (T)qb0, qc1 = (T)qb1, qc2 = (T)qb2;

    // ---- pass 2: centred products Σ w (q_a − q̄_a)(q_b − q̄_b) over active in-range cells, summed
    // in two levels: T partials over chunks of at most P2_CHUNK taps, added into FP64 totals. A
    // single FP32 running sum over a footprint of N taps carries a relative error of about
    // 2e-8·√N (N up to 25,000 at the poles of the 0.25° sphere at ℓ = 200 km), which the
    // contraction −S̄:τ amplifies where τ is nearly isotropic; the chunked sum keeps it near 1e-8.
    constexpr int P2_CHUNK = 32;
    double c00 = 0, c01 = 0, c02 = 0, c11 = 0, c12 = 0, c22 = 0;
    for (int ru = rb; ru < re; ++ru) {
        const Run R = B.runs[ru];
        int64_t jj = j + R.dy, kk = k + R.dz;
        if (jj < 0 || jj >= Ny) {
            if (!G.periodic[1]) continue;
            jj = wrapi(jj, Ny);
        }
        if (kk < 0 || kk >= Nz) {
            if (!G.periodic[2]) continue;
            kk = wrapi(kk, Nz);
        }
        const int L = TI + R.hi - R.lo;
        __syncthreads();
        stage<T4, T>(seg, q, Nx * (jj + Ny * kk), i0 + R.lo, L, Nx, G.periodic[0]);
        __syncthreads();
        if (tact) {
            const T *wr = wt + R.woff;
            const T4 *sp = seg + tid;
            const int n = R.hi - R.lo + 1;
            for (int e0 = 0; e0 < n; e0 += P2_CHUNK) {
                const int e1 = n - e0 < P2_CHUNK ? n : e0 + P2_CHUNK;
                T p00 = 0, p01 = 0, p02 = 0, p11 = 0, p12 = 0, p22 = 0;
#pragma unroll 4
                for (int e = e0; e < e1; ++e) {
                    const T4 c = sp[e];
                    T tw = wr[e];
                    if (EDGE_X) tw = (c.w < (T)0.5) ? tw : (T)0;
                    else if (MASKED) tw = tw * ((T)1 - c.w);
                    const T d0 = c.x - qc0, d1 = c.y - qc1;
                    const T t0 = tw * d0, t1 = tw * d1;
                    p00 = fma(t0, d0, p00);
                    p01 = fma(t0, d1, p01);
                    p11 = fma(t1, d1, p11);
                    if (NC == 3) {
                        const T d2 = c.z - qc2;
                        const T t2 = tw * d2;
                        p02 = fma(t0, d2, p02);
                        p12 = fma(t1, d2, p12);
                        p22 = fma(t2, d2, p22);
                    }
                }
                c00 += (double)p00;
                c01 += (double)p01;
                c11 += (double)p11;
                if (NC == 3) {
                    c02 += (double)p02;
                    c12 += (double)p12;
                    c22 += (double)p22;
                }
            }
        }
    }
    if (!inb) return;

    // ---- epilogue: τ (frame of the filter), ū, then the local frame on the sphere
    const double *cm = cmean + 3 * b;
    double U[3] = {0.0, 0.0, 0.0}, tc[6] = {0, 0, 0, 0, 0, 0};
    if (ok) {
        const double om = (B.strategy == CGEF_ZEROFILL) ? mo / D : 0.0;  // 1 − m, m = F[1]
        const double m = 1.0 - om;
        const double qb[3] = {qb0, qb1, qb2};
        const double cv[6] = {c00, c01, c02, c11, c12, c22};
        const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
        for (int a = 0; a < 3; a++) U[a] = qb[a] + cm[a] * ((B.strategy == CGEF_ZEROFILL) ? m : 1.0);
        for (int q6 = 0; q6 < 6; q6++) {
            const int a = ia[q6], bb = ib[q6];
            tc[q6] = cv[q6] / den + om * ((qb[a] + cm[a]) * (qb[bb] + cm[bb]) - om * cm[a] * cm[bb]);
        }
    }
    const int64_t P = G.npts;
    double *ub = ubar + b * 3 * P;
    T *tb = tau + b * 6 * P;
    if (SPHERE) {
        const double sl = G.trig[i], cl = G.trig[Nx + i];
        const double sp = G.trig[2 * Nx + j], cp = G.trig[2 * Nx + Ny + j];
        const double ee[3] = {-sl, cl, 0.0}, en[3] = {-sp * cl, -sp * sl, cp};
        const double er[3] = {cp * cl, cp * sl, sp};
        // Diagnostics.jl:596-610 (ū back to local) and 619-650 (τ to local east/north).
        ub[tpt] = ee[0] * U[0] + ee[1] * U[1] + ee[2] * U[2];
        ub[P + tpt] = en[0] * U[0] + en[1] * U[1] + en[2] * U[2];
        ub[2 * P + tpt] = er[0] * U[0] + er[1] * U[1] + er[2] * U[2];
        const double T3[3][3] = {{tc[0], tc[1], tc[2]}, {tc[1], tc[3], tc[4]}, {tc[2], tc[4], tc[5]}};
        double Te[3], Tn[3];
        for (int rr = 0; rr < 3; rr++) {
            Te[rr] = T3[rr][0] * ee[0] + T3[rr][1] * ee[1] + T3[rr][2] * ee[2];
            Tn[rr] = T3[rr][0] * en[0] + T3[rr][1] * en[1] + T3[rr][2] * en[2];
        }
        tb[tpt] = (T)(ee[0] * Te[0] + ee[1] * Te[1] + ee[2] * Te[2]);
        tb[P + tpt] = (T)(ee[0] * Tn[0] + ee[1] * Tn[1] + ee[2] * Tn[2]);
        tb[2 * P + tpt] = (T)(en[0] * Tn[0] + en[1] * Tn[1] + en[2] * Tn[2]);
    } else if (NC == 2) {
        ub[tpt] = U[0];
        ub[P + tpt] = U[1];
        tb[tpt] = (T)tc[0];
        tb[P + tpt] = (T)tc[1];
        tb[2 * P + tpt] = (T)tc[3];
    } else {
        ub[tpt] = U[0];
        ub[P + tpt] = U[1];
        ub[2 * P + tpt] = U[2];
        for (int q6 = 0; q6 < 6; q6++) tb[q6 * P + tpt] = (T)tc[q6];
    }
}

template <typename T, int NC, bool SPHERE>
static void pi_filter_dispatch(const GeomArgs &G, const BandArgs &B, bool masked, bool edge_x, const void *qpack,
                               const double *cmean, double *ubar, void *tau, int64_t nbatch, hipStream_t s) {
    using T4 = typename V4<T>::type;
    const size_t shmem = (size_t)(TI + B.max_run - 1) * sizeof(T4);
    dim3 grid((unsigned)((G.n[0] + TI - 1) / TI), (unsigned)(G.n[1] * G.n[2]), (unsigned)nbatch);
    const T4 *qp = (const T4 *)qpack;
    T *tp = (T *)tau;
    if (edge_x)
        hipLaunchKernelGGL((k_pi_filter<T, NC, true, true, SPHERE>), grid, dim3(TI), shmem, s, G, B, qp, cmean, ubar, tp);
    else if (masked)
        hipLaunchKernelGGL((k_pi_filter<T, NC, true, false, SPHERE>), grid, dim3(TI), shmem, s, G, B, qp, cmean, ubar, tp);
    else
        hipLaunchKernelGGL((k_pi_filter<T, NC, false, false, SPHERE>), grid, dim3(TI), shmem, s, G, B, qp, cmean, ubar, tp);
}

void launch_pi_filter(const GeomArgs &G, const BandArgs &B, int dtype, bool masked, bool edge_x, const void *qpack,
                      const double *cmean, double *ubar, void *tau, int64_t nbatch, hipStream_t s) {
    if (dtype == CGEF_F32) {
        if (G.mode == 1) pi_filter_dispatch<float, 3, true>(G, B, masked, edge_x, qpack, cmean, ubar, tau, nbatch, s);
        else if (G.mode == 0) pi_filter_dispatch<float, 2, false>(G, B, masked, edge_x, qpack, cmean, ubar, tau, nbatch, s);
        else pi_filter_dispatch<float, 3, false>(G, B, masked, edge_x, qpack, cmean, ubar, tau, nbatch, s);
    } else {
        if (G.mode == 1) pi_filter_dispatch<double, 3, true>(G, B, masked, edge_x, qpack, cmean, ubar, tau, nbatch, s);
        else if (G.mode == 0) pi_filter_dispatch<double, 2, false>(G, B, masked, edge_x, qpack, cmean, ubar, tau, nbatch, s);
        else pi_filter_dispatch<double, 3, false>(G, B, masked, edge_x, qpack, cmean, ubar, tau, nbatch, s);
    }
}

// ------------------------------------------------------------ derivatives, strain, Π, energy ----

// First-derivative stencil of the given kind (STENCIL_*, chosen per point from the mask once per
// grid) along `axis` at linear index t (axis index p, index stride `stride`): the derivative is
// c0·f[q0] + c1·f[q1] + c2·f[q2], weights from the grid's per-index table (Derivatives.jl:26-31 →
// FlowGeometries derivative!, nodes = 3, ReduceInRun). A kind only leaves the axis range on a
// periodic axis, where the index wraps; STENCIL_NONE gives a zero derivative.
struct Stencil {
    int64_t q0, q1, q2;
    double c0, c1, c2;
};

__device__ __forceinline__ Stencil stencil_of(const GeomArgs &G, int axis, int kind, int64_t t, int64_t p,
                                              int64_t stride) {
    Stencil S{t, t, t, 0.0, 0.0, 0.0};
    if (kind == STENCIL_NONE) return S;
    const int64_t n = G.n[axis];
    const int o0 = kind == STENCIL_BACKWARD3 ? -2 : ((kind == STENCIL_CENTRED || kind == STENCIL_BACKWARD2) ? -1 : 0);
    const int sb = kind == STENCIL_CENTRED ? 0 : kind == STENCIL_FORWARD3 ? 3 : kind == STENCIL_BACKWARD3 ? 6
                 : kind == STENCIL_FORWARD2 ? 9 : 11;
    auto at = [&](int o) {
        int64_t q = p + o;
        q = q < 0 ? q + n : (q >= n ? q - n : q);
        return t + (q - p) * stride;
    };
    const double *w = G.dcoef[axis] + p;  // slot s at w[s·n]
    S.q0 = at(o0);
    S.q1 = at(o0 + 1);
    S.c0 = w[sb * n];
    S.c1 = w[(sb + 1) * n];
    if (kind <= STENCIL_BACKWARD3) {
        S.q2 = at(o0 + 2);
        S.c2 = w[(sb + 2) * n];
    }
    return S;
}

template <typename F>
__device__ __forceinline__ double apply_stencil(const Stencil &S, F f) {
    return S.c0 * f(S.q0) + S.c1 * f(S.q1) + S.c2 * f(S.q2);
}

// Launch shape shared by the per-point kernels below: blockIdx.x = row · nchunk + chunk, where a row is
// one (j, k) line along x and a chunk covers 256 consecutive x indices, so no lane divides by Nx.
struct RowChunk {
    int64_t i, j, k, r, t;
};

__device__ __forceinline__ RowChunk row_chunk(const GeomArgs &G, int64_t nchunk, int64_t item) {
    RowChunk c;
    c.r = item / nchunk;
    c.i = (item - c.r * nchunk) * 256 + threadIdx.x;
    c.j = c.r % G.n[1];
    c.k = c.r / G.n[1];
    c.t = c.r * G.n[0] + c.i;
    return c;
}

int64_t row_chunks(const GeomArgs &G) { return (G.n[0] + 255) / 256; }

// Block → work item for per-point kernels with stencils: the launch holds 8·⌈nblk/8⌉ blocks, which
// the hardware deals round-robin over the 8 XCDs; each XCD takes a contiguous range of items, so a
// row and its stencil neighbours are read through the same L2.
__device__ __forceinline__ int64_t xcd_block(int64_t nblk) {
    const int64_t chunk = (nblk + 7) / 8;
    return (int64_t)(blockIdx.x % 8) * chunk + blockIdx.x / 8;
}

static unsigned xcd_grid(int64_t nblk) { return (unsigned)(8 * ((nblk + 7) / 8)); }

// The same stencil with 32-bit node indices, for the strain of one slice (its points fit 31 bits).
struct Stencil32 {
    int q0, q1, q2;
    double c0, c1, c2;
};

// Branch-free: every kind loads three weights from valid slots and the unused ones are replaced by 0
// (nodes by t), so the loads of several points can be issued together. The weights and nodes it
// returns are those of the branching form above.
__device__ __forceinline__ Stencil32 stencil32(const GeomArgs &G, int axis, int kind, int t, int p, int stride) {
    const int n = (int)G.n[axis];
    const bool none = kind == STENCIL_NONE, three = !none && kind <= STENCIL_BACKWARD3;
    // per kind (4-bit fields, kind 0 lowest): first node offset + 2 (centred −1, backward 3-node −2,
    // backward 2-node −1, else 0) and first weight slot (none: 11, a valid slot whose weights are dropped)
    const int o0 = (int)((0x120212u >> (4 * kind)) & 15u) - 2;
    const int sb = (int)((0xB9630Bu >> (4 * kind)) & 15u);
    auto at = [&](int o) {   // t + (wrapped node − p)·stride
        const int q = p + o;
        return t + (o + (q < 0 ? n : 0) - (q >= n ? n : 0)) * stride;
    };
    const double *w = G.dcoef[axis] + p;  // slot s at w[s·n]
    const double c0 = w[sb * n], c1 = w[(sb + 1) * n], c2 = w[(three ? sb + 2 : sb) * n];
    // masks rather than conditional expressions, which the compiler turns into branches around loads
    const int mi = none ? 0 : -1, m3 = three ? -1 : 0;
    const uint64_t mw = none ? 0ull : ~0ull, mw3 = three ? ~0ull : 0ull;
    auto keep = [](double c, uint64_t m) { return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, c) & m); };
    Stencil32 S;
    S.q0 = t + ((at(o0) - t) & mi);
    S.q1 = t + ((at(o0 + 1) - t) & mi);
    S.q2 = t + ((at(o0 + 2) - t) & m3);
    S.c0 = keep(c0, mw);
    S.c1 = keep(c1, mw);
    S.c2 = keep(c2, mw3);
    return S;
}

template <typename F>
__device__ __forceinline__ double apply_stencil(const Stencil32 &S, F f) {
    return S.c0 * f(S.q0) + S.c1 * f(S.q1) + S.c2 * f(S.q2);
}

// Strain S̄ (with the spherical curvature terms, Diagnostics.jl:652-669), contraction
// (Diagnostics.jl:460-466, 683-698, 848-855) and −S̄:τ at one point t = (i, j, k) of slice b; the point's
// kinetic energy density ½ |ū|² A (Diagnostics.jl:1125-1134; the radial component is excluded on the
// sphere) goes to e. Inactive points give 0 for both. The loads at t (ū, τ) do not depend on the
// point's code and go out with it; only the neighbours' indices and weights follow from the code.
// Node indices are 32-bit within the slice; the sphere's row factors 1/(R cosφ) and tanφ/R come from
// the grid's table (one uniform load each instead of two FP64 divisions per point).
// S̄:τ in 2-D from the four derivatives of ū (frame components; on the sphere scaled by the row's metric
// factors and with the curvature terms, Diagnostics.jl:652-669) and τ's xx, xy, yy at the point.
template <int MODE>
__device__ __forceinline__ double strain_contract_2d(const GeomArgs &G, int j, double dudx, double dudy, double dvdx,
                                                     double dvdy, double ut, double vt, double txx, double txy,
                                                     double tyy) {
    double Sxx, Sxy, Syy;
    if (MODE == 1) {
        const int Nx = (int)G.n[0], Ny = (int)G.n[1];
        const double hx = G.trig[2 * Nx + 2 * Ny + j], tf = G.trig[2 * Nx + 3 * Ny + j];
        const double invR = 1.0 / G.R;   // uniform: once per wave
        dudx *= hx;
        dvdx *= hx;
        dudy *= invR;
        dvdy *= invR;
        Sxx = dudx - vt * tf;
        Syy = dvdy;
        Sxy = 0.5 * (dudy + dvdx + ut * tf);
    } else {
        Sxx = dudx;
        Syy = dvdy;
        Sxy = 0.5 * (dudy + dvdx);
    }
    return Sxx * txx + 2.0 * Sxy * txy + Syy * tyy;
}

template <typename T, int MODE>
__device__ __forceinline__ double strain_point(const GeomArgs &G, const double *__restrict__ U, const T *__restrict__ tb,
                                               int t, int i, int j, int k, double &e) {
    const int P = (int)G.npts, Nx = (int)G.n[0], Ny = (int)G.n[1];
    const double *V = U + P, *W = U + 2 * P;
    const unsigned code = G.scode[t];
    const double ut = U[t], vt = V[t];
    auto fu = [&](int q) { return U[q]; };
    auto fv = [&](int q) { return V[q]; };
    const Stencil32 sx = stencil32(G, 0, code & 7, t, i, 1);
    const Stencil32 sy = stencil32(G, 1, (code >> 3) & 7, t, j, Nx);
    double cc, v2;
    if (MODE == 2) {
        const double wt = W[t];
        auto fw = [&](int q) { return W[q]; };
        const Stencil32 sz = stencil32(G, 2, (code >> 6) & 7, t, k, Nx * Ny);
        const double Sxx = apply_stencil(sx, fu), Syy = apply_stencil(sy, fv), Szz = apply_stencil(sz, fw);
        const double Sxy = 0.5 * (apply_stencil(sy, fu) + apply_stencil(sx, fv));
        const double Sxz = 0.5 * (apply_stencil(sz, fu) + apply_stencil(sx, fw));
        const double Syz = 0.5 * (apply_stencil(sz, fv) + apply_stencil(sy, fw));
        cc = Sxx * (double)tb[t] + 2.0 * Sxy * (double)tb[P + t] + Syy * (double)tb[3 * P + t] +
             2.0 * Sxz * (double)tb[2 * P + t] + 2.0 * Syz * (double)tb[4 * P + t] + Szz * (double)tb[5 * P + t];
        v2 = ut * ut + vt * vt + wt * wt;
    } else {
        cc = strain_contract_2d<MODE>(G, j, apply_stencil(sx, fu), apply_stencil(sy, fu), apply_stencil(sx, fv),
                                      apply_stencil(sy, fv), ut, vt, (double)tb[t], (double)tb[P + t],
                                      (double)tb[2 * P + t]);
        v2 = ut * ut + vt * vt;
    }
    const bool act = (code & STENCIL_ACTIVE) != 0;
    // cell_area for this MODE, the same products in the same order, without its branch on the mode
    double area;
    if (MODE == 1) area = G.R * G.R * G.trig[2 * Nx + Ny + j] * G.width[0][i] * G.width[1][j];
    else if (MODE == 2) area = G.width[0][i] * G.width[1][j] * G.width[2][k];
    else area = G.width[0][i] * G.width[1][j];
    e = act ? 0.5 * v2 * area : 0.0;
    return act ? -cc : 0.0;
}

// Π = −ρ₀ S̄:τ and per-block partial energies. A block covers 256 consecutive x indices (lane c) of
// STRAIN_RJ consecutive rows j of one level k, 512 threads: thread (c, h) takes rows h, h + 2, ….
constexpr int STRAIN_RJ = 8;
constexpr int STRAIN_NT = 512;

// Per-point loads (3-D volumes, odd row lengths): each point gathers its stencil nodes and weights.
template <typename T, int MODE>
__global__ __launch_bounds__(STRAIN_NT) void k_strain_pi(GeomArgs G, const double *__restrict__ ubar,
                                                         const T *__restrict__ tau, T *__restrict__ Pi, double rho0,
                                                         double *partials, int nblk, int64_t pi_stride, int nchunk,
                                                         int nrb) {
    __shared__ double sh[STRAIN_NT / 64];
    const int64_t b = blockIdx.y;
    const int item = (int)xcd_block(nblk);
    if (item >= nblk) return;  // whole block: padding of the XCD-ordered launch
    const int rk = item / nchunk;   // level · nrb + row block
    const int i = (item - rk * nchunk) * 256 + (int)(threadIdx.x & 255);
    const int k = rk / nrb, j0 = (rk - k * nrb) * STRAIN_RJ;
    const int Nx = (int)G.n[0], Ny = (int)G.n[1];
    const double *U = ubar + b * 3 * G.npts;
    const T *tb = tau + b * 6 * G.npts;
    T *Pb = Pi + b * pi_stride;
    double e = 0.0;
    if (i < Nx)
        for (int q = (int)(threadIdx.x >> 8); q < STRAIN_RJ && j0 + q < Ny; q += 2) {
            const int j = j0 + q, t = i + Nx * (j + Ny * k);
            double eq;
            const double out = strain_point<T, MODE>(G, U, tb, t, i, j, k, eq);
            Pb[t] = (T)(rho0 * out);
            e += eq;
        }
    if (partials) {
        const double s = block_sum<STRAIN_NT>(e, sh);
        if (threadIdx.x == 0) partials[b * nblk + item] = s;
    }
}

// 2-D with an even row length: ū, v̄ of the block's rows j0 − 2 … j0 + STRAIN_RJ + 1 and columns
// i0 − 2 … i0 + 257 (wrapped on periodic axes, clamped otherwise: a stencil never reaches a clamped
// node) are staged in LDS by 16-byte loads with the y weights of its rows; a lane holds its column's
// centred x weights. A point reads from memory only its stencil code and τ (and the x weights of a
// non-centred kind); per-point gathers of nodes and weights had the texture addresser busy for most of
// the kernel. Same nodes, weights and arithmetic as k_strain_pi. Six waves per SIMD (three blocks per
// CU by LDS) for FP32 τ; FP64 τ needs the registers of four.
constexpr int STRAIN_SR = STRAIN_RJ + 4, STRAIN_SC = 260;
template <typename T, int MODE>
__global__ __launch_bounds__(STRAIN_NT) __attribute__((amdgpu_waves_per_eu(MODE != 2 && sizeof(T) == 4 ? 6 : 4, MODE != 2 && sizeof(T) == 4 ? 6 : 4))) void k_strain_pi_tile(GeomArgs G, const double *__restrict__ ubar,
                                                              const T *__restrict__ tau, T *__restrict__ Pi,
                                                              double rho0, double *partials, int nblk,
                                                              int64_t pi_stride, int nchunk, int nrb) {
    constexpr int NF = MODE == 2 ? 3 : 2, SFL = STRAIN_SR * STRAIN_SC;   // staged fields: ū, v̄ (, w̄)
    __shared__ __attribute__((aligned(16))) double sfl[NF * SFL];
    __shared__ double syw[STRAIN_RJ * DCOEF_SLOTS], szw[DCOEF_SLOTS];
    double *su = sfl, *sv = sfl + SFL;
    __shared__ double sh[STRAIN_NT / 64];
    const int64_t b = blockIdx.y;
    const int item = (int)xcd_block(nblk);
    if (item >= nblk) return;  // whole block: padding of the XCD-ordered launch
    const int rk = item / nchunk;
    const int i0 = (item - rk * nchunk) * 256;
    const int k = rk / nrb, j0 = (rk - k * nrb) * STRAIN_RJ;
    const int Nx = (int)G.n[0], Ny = (int)G.n[1], P = (int)G.npts, lev = Nx * Ny * k;
    const double *U = ubar + b * 3 * G.npts;
    const int tid = (int)threadIdx.x;
    // Every staging load is issued before the first LDS store (a load-store loop waited out each
    // load's latency in turn): a thread's index runs past the end as repeats of the last element,
    // which store the same value to the same place.
    constexpr int NPAIR = STRAIN_SR * (STRAIN_SC / 2), NPS = (NF * NPAIR + STRAIN_NT - 1) / STRAIN_NT;
    double2 xs[NPS];
#pragma unroll
    for (int n = 0; n < NPS; n++) {
        const int e = min(tid + n * STRAIN_NT, NF * NPAIR - 1);
        const int f = e / NPAIR, ee = e - f * NPAIR;
        const int r = ee / (STRAIN_SC / 2), pc = ee - r * (STRAIN_SC / 2);
        int jr = j0 - 2 + r, gc = i0 - 2 + 2 * pc;   // gc even, Nx even: a pair never straddles the wrap
        if (G.periodic[1]) jr = (int)wrapi(jr, Ny);
        else jr = min(max(jr, 0), Ny - 1);
        if (G.periodic[0]) gc = (int)wrapi(gc, Nx);
        else gc = min(max(gc, 0), Nx - 2);
        xs[n] = *reinterpret_cast<const double2 *>(U + f * P + lev + jr * Nx + gc);
    }
    // the thread's points: column c, rows h, h + 2, …; their codes and τ requested with the staging
    // (a column past the row's end or a row past the last repeats the last and is not stored)
    const int c = tid & 255, i = i0 + c, h = tid >> 8, ic = min(i, Nx - 1);
    const T *tb = tau + b * 6 * G.npts;
    // the column's centred x weights (slots 0-2); other kinds read theirs when they occur
    const double *wxc = G.dcoef[0] + ic;
    const double cxc0 = wxc[0], cxc1 = wxc[Nx], cxc2 = wxc[2 * Nx];
    constexpr int NQ = STRAIN_RJ / 2;
    unsigned codes[NQ];
    T txx[NQ], txy[NQ], tyy[NQ];   // held in the storage type until used
#pragma unroll
    for (int n = 0; n < NQ; n++) {
        const int t = lev + min(j0 + h + 2 * n, Ny - 1) * Nx + ic;
        codes[n] = G.scode[t];
        if (MODE != 2) {   // 3-D: τ's six components are read with the z nodes, after the barrier
            txx[n] = tb[t];
            txy[n] = tb[P + t];
            tyy[n] = tb[2 * P + t];
        }
    }
    const int qy = min(tid, STRAIN_RJ * DCOEF_SLOTS - 1) / DCOEF_SLOTS;
    const double wy = G.dcoef[1][(min(tid, STRAIN_RJ * DCOEF_SLOTS - 1) - qy * DCOEF_SLOTS) * Ny + min(j0 + qy, Ny - 1)];
#pragma unroll
    for (int n = 0; n < NPS; n++) {
        const int e = min(tid + n * STRAIN_NT, NF * NPAIR - 1);
        const int f = e / NPAIR, ee = e - f * NPAIR;
        const int r = ee / (STRAIN_SC / 2), pc = ee - r * (STRAIN_SC / 2);
        *reinterpret_cast<double2 *>(sfl + f * SFL + r * STRAIN_SC + 2 * pc) = xs[n];
    }
    if (MODE == 2 && tid < DCOEF_SLOTS) szw[tid] = G.dcoef[2][tid * G.n[2] + k];   // the level's z weights
    syw[min(tid, STRAIN_RJ * DCOEF_SLOTS - 1)] = wy;
#pragma unroll
    for (int n = 0; n < NQ; n++)   // keeps the compiler from sinking these loads below the barrier
        if (MODE != 2) asm volatile("" ::"v"(codes[n]), "v"(txx[n]), "v"(txy[n]), "v"(tyy[n]));
    __syncthreads();
    T *Pb = Pi + b * pi_stride;
    double e = 0.0;
    if (i < Nx) {
        const double wxi = G.width[0][i];
#pragma unroll
        for (int n = 0; n < NQ; n++) {
            const int q = h + 2 * n, j = min(j0 + q, Ny - 1), t = lev + j * Nx + i;
            const unsigned code = codes[n];
            const double *pu = su + (q + 2) * STRAIN_SC + c + 2, *pv = sv + (q + 2) * STRAIN_SC + c + 2;
            // the stencil of stencil32, as LDS offsets from the point and slots of the staged weights
            auto stencil = [&](int kind, int &o0, int &o1, int &o2, int &s0, int &s1, int &s2, uint64_t &m01,
                               uint64_t &m2) {
                const bool none = kind == STENCIL_NONE, three = !none && kind <= STENCIL_BACKWARD3;
                const int oo = (int)((0x120212u >> (4 * kind)) & 15u) - 2;
                const int sb = (int)((0xB9630Bu >> (4 * kind)) & 15u);
                const int mi = none ? 0 : -1, m3 = three ? -1 : 0;
                o0 = oo & mi;
                o1 = (oo + 1) & mi;
                o2 = (oo + 2) & m3;
                s0 = sb;
                s1 = sb + 1;
                s2 = three ? sb + 2 : sb;
                m01 = none ? 0ull : ~0ull;
                m2 = three ? ~0ull : 0ull;
            };
            auto keep = [](double w, uint64_t m) {
                return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, w) & m);
            };
            int x0, x1, x2, xs0, xs1, xs2, y0, y1, y2, ys0, ys1, ys2;
            uint64_t xm, xm2, ym, ym2;
            stencil(code & 7, x0, x1, x2, xs0, xs1, xs2, xm, xm2);
            stencil((code >> 3) & 7, y0, y1, y2, ys0, ys1, ys2, ym, ym2);
            double cx0 = cxc0, cx1 = cxc1, cx2 = cxc2;
            if ((code & 7) != STENCIL_CENTRED) {
                cx0 = keep(wxc[xs0 * Nx], xm);
                cx1 = keep(wxc[xs1 * Nx], xm);
                cx2 = keep(wxc[xs2 * Nx], xm2);
            }
            const double *yw = syw + q * DCOEF_SLOTS;
            const double cy0 = keep(yw[ys0], ym), cy1 = keep(yw[ys1], ym), cy2 = keep(yw[ys2], ym2);
            const double ut = pu[0], vt = pv[0];
            const double dudx = cx0 * pu[x0] + cx1 * pu[x1] + cx2 * pu[x2];
            const double dudy = cy0 * pu[y0 * STRAIN_SC] + cy1 * pu[y1 * STRAIN_SC] + cy2 * pu[y2 * STRAIN_SC];
            const double dvdx = cx0 * pv[x0] + cx1 * pv[x1] + cx2 * pv[x2];
            const double dvdy = cy0 * pv[y0 * STRAIN_SC] + cy1 * pv[y1 * STRAIN_SC] + cy2 * pv[y2 * STRAIN_SC];
            double cc, v2, area;
            if constexpr (MODE == 2) {
                // the z nodes of the point's column from memory (levels k + o, wrapped on a periodic axis),
                // the level's z weights from LDS; the arithmetic of strain_point's 3-D branch
                const double *pw = pu + 2 * SFL;
                const int Nz = (int)G.n[2], NxNy = Nx * Ny;
                int z0, z1, z2, zs0, zs1, zs2;
                uint64_t zm, zm2;
                stencil((code >> 6) & 7, z0, z1, z2, zs0, zs1, zs2, zm, zm2);
                auto lvl = [&](int o) {   // the point's index at level k + o
                    int kw = k + o;
                    if (G.periodic[2]) kw = (int)((unsigned)(kw + 2 * Nz) % (unsigned)Nz);
                    return i + Nx * j + NxNy * kw;
                };
                const int q0 = lvl(z0), q1 = lvl(z1), q2 = lvl(z2);
                const double cz0 = keep(szw[zs0], zm), cz1 = keep(szw[zs1], zm), cz2 = keep(szw[zs2], zm2);
                const double *Vp = U + P, *Wp = U + 2 * P;
                const double wt = pw[0];
                const double dwdx = cx0 * pw[x0] + cx1 * pw[x1] + cx2 * pw[x2];
                const double dwdy = cy0 * pw[y0 * STRAIN_SC] + cy1 * pw[y1 * STRAIN_SC] + cy2 * pw[y2 * STRAIN_SC];
                const double dudz = cz0 * U[q0] + cz1 * U[q1] + cz2 * U[q2];
                const double dvdz = cz0 * Vp[q0] + cz1 * Vp[q1] + cz2 * Vp[q2];
                const double dwdz = cz0 * Wp[q0] + cz1 * Wp[q1] + cz2 * Wp[q2];
                const double Sxx = dudx, Syy = dvdy, Szz = dwdz;
                const double Sxy = 0.5 * (dudy + dvdx);
                const double Sxz = 0.5 * (dudz + dwdx);
                const double Syz = 0.5 * (dvdz + dwdy);
                cc = Sxx * (double)tb[t] + 2.0 * Sxy * (double)tb[P + t] + Syy * (double)tb[3 * P + t] +
                     2.0 * Sxz * (double)tb[2 * P + t] + 2.0 * Syz * (double)tb[4 * P + t] + Szz * (double)tb[5 * P + t];
                v2 = ut * ut + vt * vt + wt * wt;
                area = wxi * G.width[1][j] * G.width[2][k];
            } else {
                cc = strain_contract_2d<MODE>(G, j, dudx, dudy, dvdx, dvdy, ut, vt, (double)txx[n], (double)txy[n],
                                              (double)tyy[n]);
                v2 = ut * ut + vt * vt;
                if (MODE == 1) area = G.R * G.R * G.trig[2 * Nx + Ny + j] * wxi * G.width[1][j];
                else area = wxi * G.width[1][j];
            }
            const bool act = (code & STENCIL_ACTIVE) != 0;
            if (j0 + q < Ny) {
                Pb[t] = (T)(rho0 * (act ? -cc : 0.0));
                e += act ? 0.5 * v2 * area : 0.0;
            }
        }
    }
    if (partials) {
        const double s = block_sum<STRAIN_NT>(e, sh);
        if (threadIdx.x == 0) partials[b * nblk + item] = s;
    }
}

int64_t strain_blocks(const GeomArgs &G) { return row_chunks(G) * ((G.n[1] + STRAIN_RJ - 1) / STRAIN_RJ) * G.n[2]; }

int64_t energy_blocks(const GeomArgs &G) { return row_chunks(G) * G.n[1] * G.n[2]; }

void launch_strain_pi(const GeomArgs &G, int dtype, const double *ubar, const void *tau, void *Pi, double rho0,
                      int64_t nbatch, double *partials, int64_t npartials, hipStream_t s, int64_t pi_stride) {
    if (G.npts * 6 >= ((int64_t)1 << 31)) fail(CGEF_ERR_DIMENSION, "the strain kernel indexes a slice in 32 bits");
    const int64_t nblk = strain_blocks(G);
    double *pp = (partials && nblk * nbatch <= npartials) ? partials : nullptr;
    dim3 grid(xcd_grid(nblk), (unsigned)nbatch);
    const int64_t ps = pi_stride > 0 ? pi_stride : G.npts;
    const int nch = (int)row_chunks(G), nrb = (int)((G.n[1] + STRAIN_RJ - 1) / STRAIN_RJ);
    // the LDS-staged tile kernel for 2-D slices with 16-byte aligned pairs of an even row length
    const bool tile = (G.mode == 2 || G.n[2] == 1) && G.n[0] % 2 == 0 && (reinterpret_cast<uintptr_t>(ubar) & 15) == 0;
#define CGEF_SP(KK, TT, MM) hipLaunchKernelGGL((KK<TT, MM>), grid, dim3(STRAIN_NT), 0, s, G, ubar, (const TT *)tau, (TT *)Pi, rho0, pp, (int)nblk, ps, nch, nrb)
    if (dtype == CGEF_F32) {
        if (G.mode == 0 && tile) CGEF_SP(k_strain_pi_tile, float, 0);
        else if (G.mode == 1 && tile) CGEF_SP(k_strain_pi_tile, float, 1);
        else if (G.mode == 2 && tile) CGEF_SP(k_strain_pi_tile, float, 2);
        else if (G.mode == 0) CGEF_SP(k_strain_pi, float, 0);
        else if (G.mode == 1) CGEF_SP(k_strain_pi, float, 1);
        else CGEF_SP(k_strain_pi, float, 2);
    } else {
        if (G.mode == 0 && tile) CGEF_SP(k_strain_pi_tile, double, 0);
        else if (G.mode == 1 && tile) CGEF_SP(k_strain_pi_tile, double, 1);
        else if (G.mode == 2 && tile) CGEF_SP(k_strain_pi_tile, double, 2);
        else if (G.mode == 0) CGEF_SP(k_strain_pi, double, 0);
        else if (G.mode == 1) CGEF_SP(k_strain_pi, double, 1);
        else CGEF_SP(k_strain_pi, double, 2);
    }
#undef CGEF_SP
}

__global__ void k_energy_final(const double *partials, int64_t nblk, double inv_area, double *E) {
    __shared__ double sh[4];
    const int64_t b = blockIdx.x;
    double acc = 0.0;
    for (int64_t q = threadIdx.x; q < nblk; q += 256) acc += partials[b * nblk + q];
    const double s = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) E[b] = s * inv_area;
}

void launch_energy_final(int64_t nbatch, int64_t nblocks, const double *partials, double active_area, double *E_dev,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_energy_final, dim3((unsigned)nbatch), dim3(256), 0, s, partials, nblocks, 1.0 / active_area, E_dev);
}

template <typename T>
__global__ __launch_bounds__(256) void k_energy_fields(GeomArgs G, const T *ub, const T *vb, const T *wb, double *partials,
                                                       int64_t nblk, int64_t nchunk) {
    __shared__ double sh[4];
    const int64_t b = blockIdx.y;
    const int64_t item = xcd_block(nblk);
    if (item >= nblk) return;
    const RowChunk c = row_chunk(G, nchunk, item);
    const int64_t P = G.npts, t = c.t;
    double e = 0.0;
    if (c.i < G.n[0] && G.mask[t]) {
        const double u = (double)ub[b * P + t], v = (double)vb[b * P + t];
        double v2 = u * u + v * v;
        if (wb) {
            const double w = (double)wb[b * P + t];
            v2 += w * w;
        }
        e = 0.5 * v2 * cell_area(G, c.i, c.j, c.k);
    }
    const double s = block_sum<256>(e, sh);
    if (threadIdx.x == 0) partials[b * nblk + item] = s;
}

void launch_energy_fields(const GeomArgs &G, int dtype, const void *ub, const void *vb, const void *wb, int64_t nbatch,
                          double *partials, int64_t npartials, hipStream_t s) {
    const int64_t nblk = energy_blocks(G);
    (void)npartials;
    dim3 grid(xcd_grid(nblk), (unsigned)nbatch);
    if (dtype == CGEF_F32)
        hipLaunchKernelGGL(k_energy_fields<float>, grid, dim3(256), 0, s, G, (const float *)ub, (const float *)vb,
                           (const float *)wb, partials, nblk, row_chunks(G));
    else
        hipLaunchKernelGGL(k_energy_fields<double>, grid, dim3(256), 0, s, G, (const double *)ub, (const double *)vb,
                           (const double *)wb, partials, nblk, row_chunks(G));
}

template <typename T>
__global__ __launch_bounds__(256) void k_derivative(GeomArgs G, int axis, const T *in, T *out, int64_t nchunk) {
    const int64_t b = blockIdx.y;
    const int64_t nblk = nchunk * G.n[1] * G.n[2];
    const int64_t item = xcd_block(nblk);
    if (item >= nblk) return;
    const RowChunk c = row_chunk(G, nchunk, item);
    const int64_t P = G.npts, t = c.t, Nx = G.n[0];
    if (c.i >= Nx) return;
    double v = 0.0;
    const unsigned code = G.scode[t];
    if (code & STENCIL_ACTIVE) {
        const T *f = in + b * P;
        const int64_t p = axis == 0 ? c.i : (axis == 1 ? c.j : c.k);
        const int64_t stride = axis == 0 ? 1 : (axis == 1 ? Nx : Nx * G.n[1]);
        v = apply_stencil(stencil_of(G, axis, (code >> (3 * axis)) & 7, t, p, stride),
                          [&](int64_t q) { return (double)f[q]; });
        if (G.mode == 1) {
            const double cp = G.trig[2 * Nx + G.n[1] + c.j];
            if (axis == 0) v = (fabs(cp) < sqrt(DBL_EPSILON)) ? 0.0 : v / (G.R * cp);
            else v = v / G.R;
        }
    }
    out[b * P + t] = (T)v;
}

void launch_derivative(const GeomArgs &G, int axis, int dtype, const void *in, void *out, int64_t nbatch, hipStream_t s) {
    const int64_t nch = row_chunks(G);
    dim3 grid(xcd_grid(nch * G.n[1] * G.n[2]), (unsigned)nbatch);
    if (dtype == CGEF_F32)
        hipLaunchKernelGGL(k_derivative<float>, grid, dim3(256), 0, s, G, axis, (const float *)in, (float *)out, nch);
    else
        hipLaunchKernelGGL(k_derivative<double>, grid, dim3(256), 0, s, G, axis, (const double *)in, (double *)out, nch);
}

// ----------------------------------------------------------- linear filter of K given fields ----

// Same run walk as k_pi_filter for plain fields (filter_apply!): the staged cells are f·active and
// a flag (0 active, 1 inactive, 2 outside a non-periodic x axis). FP64 accumulation.
template <typename T, int K, bool MASKED, bool EDGE_X>
__global__ __launch_bounds__(TI) void k_filter_lin(GeomArgs G, BandArgs B, const T *__restrict__ in, T *__restrict__ out,
                                                   int64_t nslices) {
    extern __shared__ unsigned char smem_raw[];
    const int L_max = TI + B.max_run - 1;
    T *seg = reinterpret_cast<T *>(smem_raw);
    unsigned char *flag = smem_raw + (size_t)K * L_max * sizeof(T);
    const int tid = threadIdx.x;
    const int64_t Nx = G.n[0], Ny = G.n[1], Nz = G.n[2];
    const int32_t r = B.row_order[blockIdx.y];
    const int64_t j = r % Ny, k = r / Ny;
    const int64_t s0 = (int64_t)blockIdx.z * K;
    const int64_t i0 = (int64_t)blockIdx.x * TI;
    const int64_t i = i0 + tid;
    const bool inb = i < Nx;
    const int64_t tpt = i + Nx * (int64_t)r;
    const bool tact = inb && (!MASKED || G.mask[tpt] != 0);
    const int band = B.band_of_row[r];
    const int rb = B.run_ptr[band], re = B.run_ptr[band + 1];
    double acc[K];
#pragma unroll
    for (int m = 0; m < K; m++) acc[m] = 0.0;
    double mo = 0.0, din = 0.0;
    for (int ru = rb; ru < re; ++ru) {
        const Run R = B.runs[ru];
        int64_t jj = j + R.dy, kk = k + R.dz;
        if (jj < 0 || jj >= Ny) {
            if (!G.periodic[1]) continue;
            jj = wrapi(jj, Ny);
        }
        if (kk < 0 || kk >= Nz) {
            if (!G.periodic[2]) continue;
            kk = wrapi(kk, Nz);
        }
        const int L = TI + R.hi - R.lo;
        const int64_t rowbase = Nx * (jj + Ny * kk);
        __syncthreads();
        for (int s = tid; s < L; s += TI) {
            int64_t gx = i0 + R.lo + s;
            unsigned char fl = 0;
            if (gx < 0 || gx >= Nx) {
                if (G.periodic[0]) gx = wrapi(gx, Nx);
                else fl = 2;
            }
            if (fl == 0 && MASKED && !G.mask[rowbase + gx]) fl = 1;
            flag[s] = fl;
#pragma unroll
            for (int m = 0; m < K; m++) {
                const int64_t sl = s0 + m;
                seg[m * L_max + s] = (fl == 0 && sl < nslices) ? in[sl * G.npts + rowbase + gx] : (T)0;
            }
        }
        __syncthreads();
        if (tact) {
            const double *wd = B.wd + R.woff;
            const int n = R.hi - R.lo + 1;
            for (int e = 0; e < n; ++e) {
                const double w = wd[e];
#pragma unroll
                for (int m = 0; m < K; m++) acc[m] = fma(w, (double)seg[m * L_max + tid + e], acc[m]);
                if (EDGE_X || MASKED) {
                    const unsigned char fl = flag[tid + e];
                    if (fl == 1) mo += w;
                    if (EDGE_X && fl != 2) din += w;
                }
            }
        }
    }
    if (!inb) return;
    const double D = EDGE_X ? din : B.D_row[r];
    const double den = (B.strategy == CGEF_DEFORMABLE) ? D - mo : D;
    const bool ok = tact && den > 1e-15;
#pragma unroll
    for (int m = 0; m < K; m++) {
        const int64_t sl = s0 + m;
        if (sl < nslices) out[sl * G.npts + tpt] = ok ? (T)(acc[m] / den) : (T)0;
    }
}

template <typename T, int K>
static void filter_lin_k(const GeomArgs &G, const BandArgs &B, bool masked, bool edge_x, const void *in, void *out,
                         int64_t nslices, hipStream_t s) {
    const size_t L_max = TI + B.max_run - 1;
    const size_t shmem = K * L_max * sizeof(T) + L_max;
    dim3 grid((unsigned)((G.n[0] + TI - 1) / TI), (unsigned)(G.n[1] * G.n[2]), (unsigned)((nslices + K - 1) / K));
    if (edge_x)
        hipLaunchKernelGGL((k_filter_lin<T, K, true, true>), grid, dim3(TI), shmem, s, G, B, (const T *)in, (T *)out, nslices);
    else if (masked)
        hipLaunchKernelGGL((k_filter_lin<T, K, true, false>), grid, dim3(TI), shmem, s, G, B, (const T *)in, (T *)out, nslices);
    else
        hipLaunchKernelGGL((k_filter_lin<T, K, false, false>), grid, dim3(TI), shmem, s, G, B, (const T *)in, (T *)out, nslices);
}

void launch_filter_lin(const GeomArgs &G, const BandArgs &B, int dtype, bool masked, bool edge_x, const void *in,
                       void *out, int64_t nslices, hipStream_t s) {
    // up to 4 slices share one neighbour walk (filter_apply_batch!); larger batches take more z-blocks
    const int K = nslices >= 4 ? 4 : (int)nslices;
    if (dtype == CGEF_F32) {
        if (K == 4) filter_lin_k<float, 4>(G, B, masked, edge_x, in, out, nslices, s);
        else if (K == 3) filter_lin_k<float, 3>(G, B, masked, edge_x, in, out, nslices, s);
        else if (K == 2) filter_lin_k<float, 2>(G, B, masked, edge_x, in, out, nslices, s);
        else filter_lin_k<float, 1>(G, B, masked, edge_x, in, out, nslices, s);
    } else {
        if (K == 4) filter_lin_k<double, 4>(G, B, masked, edge_x, in, out, nslices, s);
        else if (K == 3) filter_lin_k<double, 3>(G, B, masked, edge_x, in, out, nslices, s);
        else if (K == 2) filter_lin_k<double, 2>(G, B, masked, edge_x, in, out, nslices, s);
        else filter_lin_k<double, 1>(G, B, masked, edge_x, in, out, nslices, s);
    }
}

// ----------------------------------------------- plain FP64 direct sum (the KA kernel restated) ----

// KernelAbstractionsExt.jl:37-90 `_cgef_filter_kernel!`: one thread per output point, the CSR
// footprint (di, dj[, dk], w, ptr) read from global memory, ZeroFill/Deformable, ws/wn > 1e-15.
__global__ __launch_bounds__(256) void k_baseline(GeomArgs G, const int32_t *__restrict__ di, const int32_t *__restrict__ dj,
                                                  const int32_t *__restrict__ dk, const double *__restrict__ w,
                                                  const int64_t *__restrict__ ptr, const int32_t *__restrict__ band_of_row,
                                                  int strategy, const double *__restrict__ field, double *__restrict__ out) {
    const int64_t Nx = G.n[0], Ny = G.n[1], Nz = G.n[2];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = (int64_t)blockIdx.y * blockDim.y + threadIdx.y;
    const int64_t b = blockIdx.z;
    if (i >= Nx || r >= Ny * Nz) return;
    const int64_t j = r % Ny, k = r / Ny;
    const int64_t t = i + Nx * r;
    const double *f = field + b * G.npts;
    if (!G.mask[t]) {
        out[b * G.npts + t] = 0.0;
        return;
    }
    const int band = band_of_row[r];
    double ws = 0.0, wn = 0.0;
    for (int64_t e = ptr[band]; e < ptr[band + 1]; e++) {
        int64_t jj = j + dj[e];
        if (jj < 0 || jj >= Ny) {
            if (!G.periodic[1]) continue;
            jj = wrapi(jj, Ny);
        }
        int64_t kk = k + dk[e];
        if (kk < 0 || kk >= Nz) {
            if (!G.periodic[2]) continue;
            kk = wrapi(kk, Nz);
        }
        int64_t ii = i + di[e];
        if (ii < 0 || ii >= Nx) {
            if (!G.periodic[0]) continue;
            ii = wrapi(ii, Nx);
        }
        const int64_t sidx = ii + Nx * (jj + Ny * kk);
        const bool act = G.mask[sidx] != 0;
        const double wk = w[e];
        if (strategy == CGEF_ZEROFILL) {
            wn += wk;
            if (act) ws += wk * f[sidx];
        } else if (act) {
            wn += wk;
            ws += wk * f[sidx];
        }
    }
    out[b * G.npts + t] = wn > 1e-15 ? ws / wn : 0.0;
}

void launch_baseline_filter(const GeomArgs &G, const Plan &p, const double *in, double *out, int64_t nslices,
                            hipStream_t s) {
    dim3 block(64, 4);
    dim3 grid((unsigned)((G.n[0] + 63) / 64), (unsigned)((G.n[1] * G.n[2] + 3) / 4), (unsigned)nslices);
    hipLaunchKernelGGL(k_baseline, grid, block, 0, s, G, p.csr_di_d, p.csr_dj_d, p.csr_dk_d, p.w64_d, p.csr_ptr_d,
                       p.band_of_row_d, p.strategy, in, out);
}

// ------------------------------------------------------------- separable Gaussian (Cartesian) ----

// Filtering.jl:1210-1279 row pass (masked input) then column pass; 1510-1528 normalization.
template <typename T>
__global__ __launch_bounds__(256) void k_sep_row(GeomArgs G, const double *__restrict__ gx, int lx, const T *__restrict__ in,
                                                 double *__restrict__ rp) {
    const int64_t b = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= G.npts) return;
    const int64_t Nx = G.n[0];
    const int64_t i = t % Nx, rowbase = t - i;
    const T *f = in + b * G.npts;
    double acc = 0.0;
    for (int d = -lx; d <= lx; d++) {
        int64_t ii = i + d;
        if (ii < 0 || ii >= Nx) {
            if (!G.periodic[0]) continue;
            ii = wrapi(ii, Nx);
        }
        const int64_t q = rowbase + ii;
        acc += gx[d + lx] * (G.mask[q] ? (double)f[q] : 0.0);
    }
    rp[b * G.npts + t] = acc;
}

template <typename T>
__global__ __launch_bounds__(256) void k_sep_col(GeomArgs G, const double *__restrict__ gy, int ly, const double *__restrict__ rp,
                                                 const double *__restrict__ px, const double *__restrict__ py,
                                                 const double *__restrict__ invden, T *__restrict__ out) {
    const int64_t b = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= G.npts) return;
    const int64_t Nx = G.n[0], Ny = G.n[1];
    const int64_t i = t % Nx, j = t / Nx;
    const double *r = rp + b * G.npts;
    double acc = 0.0;
    for (int d = -ly; d <= ly; d++) {
        int64_t jj = j + d;
        if (jj < 0 || jj >= Ny) {
            if (!G.periodic[1]) continue;
            jj = wrapi(jj, Ny);
        }
        acc += gy[d + ly] * r[i + Nx * jj];
    }
    double v;
    if (invden) {
        v = acc * invden[t];
    } else {
        const double den = px[i] * py[j];
        v = den > 1e-15 ? acc / den : 0.0;
    }
    if (!G.mask[t]) v = 0.0;
    out[b * G.npts + t] = (T)v;
}

void launch_separable(const GeomArgs &G, const Plan &p, int dtype, const void *in, void *out, int64_t nslices,
                      double *rowpass, hipStream_t s) {
    dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)nslices);
    if (dtype == CGEF_F32) {
        hipLaunchKernelGGL(k_sep_row<float>, grid, dim3(256), 0, s, G, p.sep_g_d[0], p.sep_lim[0], (const float *)in, rowpass);
        hipLaunchKernelGGL(k_sep_col<float>, grid, dim3(256), 0, s, G, p.sep_g_d[1], p.sep_lim[1], rowpass, p.sep_prof_d[0],
                           p.sep_prof_d[1], p.sep_invden_d, (float *)out);
    } else {
        hipLaunchKernelGGL(k_sep_row<double>, grid, dim3(256), 0, s, G, p.sep_g_d[0], p.sep_lim[0], (const double *)in, rowpass);
        hipLaunchKernelGGL(k_sep_col<double>, grid, dim3(256), 0, s, G, p.sep_g_d[1], p.sep_lim[1], rowpass, p.sep_prof_d[0],
                           p.sep_prof_d[1], p.sep_invden_d, (double *)out);
    }
}

// Separable Π path: the row pass forms the five Cartesian planes (u, v, uu, uv, vv of the masked
// velocity) from (u, v) staged in LDS for a 256-cell row segment plus its halo; the column pass walks
// down TJ output rows per lane with the input rows streamed once (each read row feeds every output
// row within reach), finishes all five planes and hands each point to the shared epilogue (ZeroFill:
// product of the axis profiles; Deformable: 1 / filtered mask). Taps are summed in the same order
// as the plain two-pass filter.
template <typename T>
__global__ __launch_bounds__(256) void k_sep_pi_row(GeomArgs G, const double *__restrict__ gx, int lx,
                                                    const T *__restrict__ u, const T *__restrict__ v,
                                                    double *__restrict__ rp, int64_t nchunk) {
    extern __shared__ double srow[];   // u and v of the segment and its halo, then the 2·lx + 1 taps
    const int64_t b = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x / nchunk;
    const int64_t i0 = ((int64_t)blockIdx.x - j * nchunk) * 256;
    const int64_t Nx = G.n[0], P = G.npts, rowbase = j * Nx;
    const int L = 256 + 2 * lx;
    double *su = srow, *sv = srow + L, *sg = srow + 2 * L;
    for (int e = threadIdx.x; e <= 2 * lx; e += 256) sg[e] = gx[e];
    const T *ub = u + b * P, *vb = v + b * P;
    // two elements per lane per trip, both requested before either is stored (one trip for L ≤ 512);
    // the wrap is 32-bit and branch-free (ii + wadd ≥ 0), so no branch separates the loads
    const int nx = (int)Nx, wadd = (lx / nx + 1) * nx;
    const bool per = G.periodic[0] != 0;
    for (int s0 = 0; s0 < L; s0 += 512) {
        T xu[2], xv[2];
        bool ac[2];
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const int sidx = s0 + (int)threadIdx.x + 256 * m;
            const int ii = (int)i0 - lx + sidx;
            int iw = (int)((unsigned)(ii + wadd) % (unsigned)nx);
            asm volatile("" : "+v"(iw));   // computed unconditionally
            const bool inr = ii >= 0 && ii < nx;
            const bool in = sidx < L && (inr || per);
            const int64_t q = rowbase + (inr ? ii : (per ? iw : 0));   // a valid index either way
            const bool mk = G.mask[q] != 0;                                 // read unconditionally
            ac[m] = in && mk;
            xu[m] = ub[q];
            xv[m] = vb[q];
        }
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const int sidx = s0 + (int)threadIdx.x + 256 * m;
            if (sidx < L) {
                su[sidx] = ac[m] ? (double)xu[m] : 0.0;
                sv[sidx] = ac[m] ? (double)xv[m] : 0.0;
            }
        }
    }
    __syncthreads();
    const int64_t i = i0 + threadIdx.x;
    if (i >= Nx) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
    for (int d = 0; d <= 2 * lx; d++) {
        const double g = sg[d], x = su[threadIdx.x + d], y = sv[threadIdx.x + d];
        a0 += g * x;
        a1 += g * y;
        a2 += g * (x * x);
        a3 += g * (x * y);
        a4 += g * (y * y);
    }
    const int64_t t = rowbase + i;
    double *o = rp + b * 5 * P;
    o[t] = a0;
    o[P + t] = a1;
    o[2 * P + t] = a2;
    o[3 * P + t] = a3;
    o[4 * P + t] = a4;
}

// Column pass of the separable Π on 64-column × 64-row tiles (512 threads: lane = column, wave = an
// 8-row band). The five planes pass through LDS one at a time — the tile's 64 + 2·ly rows of that
// plane, staged with coalesced 512-byte row pieces, so each element is read from memory once per tile
// instead of once per 8-row band — and each lane sums its 8 outputs from LDS, the taps in a register
// window that shifts by one per row (taps outside the kernel are 0, and a row beyond a non-periodic
// edge reads as 0: adding +0 leaves a sum unchanged, so this equals skipping them).
constexpr int SEPC_W = 64, SEPC_H = 64, SEPC_TJ = 8;
constexpr int SEPC_AH = 32;   // rows per workgroup of the column pass with direct global → LDS loads
template <typename T>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_sep_pi_col(GeomArgs G, const double *__restrict__ gy, int ly,
                                                    const double *__restrict__ rp, const double *__restrict__ px,
                                                    const double *__restrict__ py, const double *__restrict__ invden,
                                                    double *__restrict__ ubar, T *__restrict__ tau, T *__restrict__ uo,
                                                    T *__restrict__ vo, T *__restrict__ wo, int64_t ntx) {
    extern __shared__ double smem[];   // the 2·ly + 1 taps, then one plane's 64 + 2·ly rows × 64 columns
    const int taps = 2 * ly + 1, R = SEPC_H + 2 * ly;
    double *sg = smem, *st = smem + taps;
    const int64_t b = blockIdx.y;
    const int64_t tx = (int64_t)blockIdx.x % ntx, ty = (int64_t)blockIdx.x / ntx;
    const int64_t Nx = G.n[0], Ny = G.n[1], P = G.npts;
    const int c = threadIdx.x & (SEPC_W - 1), band = threadIdx.x / SEPC_W;
    const int64_t i = tx * SEPC_W + c, j0 = ty * SEPC_H + band * SEPC_TJ;
    const double *r = rp + b * 5 * P;
    for (int e = threadIdx.x; e < taps; e += 512) sg[e] = gy[e];
    double F[SEPC_TJ][5];
#pragma unroll
    for (int q = 0; q < SEPC_TJ; q++)
#pragma unroll
        for (int f = 0; f < 5; f++) F[q][f] = 0.0;
#pragma unroll
    for (int f = 0; f < 5; f++) {
        if (f) __syncthreads();   // the previous plane's readers are done
        // lane: column c, rows band, band + 8, …: the row tests are wave-uniform branches
#pragma unroll 4
        for (int k = band; k < R; k += 512 / SEPC_W) {
            int64_t jw = ty * SEPC_H - ly + k;
            bool in = i < Nx;
            if (jw < 0 || jw >= Ny) {
                if (G.periodic[1]) jw = wrapi(jw, Ny);
                else in = false;
            }
            st[k * SEPC_W + c] = in ? r[f * P + jw * Nx + i] : 0.0;
        }
        __syncthreads();
        double g[SEPC_TJ];
#pragma unroll
        for (int q = 0; q < SEPC_TJ; q++) g[q] = 0.0;
        const double *col = st + band * SEPC_TJ * SEPC_W + c;
        for (int k = 0; k < taps + SEPC_TJ - 1; k++) {
#pragma unroll
            for (int q = SEPC_TJ - 1; q > 0; q--) g[q] = g[q - 1];
            g[0] = k < taps ? sg[k] : 0.0;
            const double x = col[k * SEPC_W];
#pragma unroll
            for (int q = 0; q < SEPC_TJ; q++) F[q][f] += g[q] * x;
        }
    }
    if (i >= Nx) return;
#pragma unroll
    for (int q = 0; q < SEPC_TJ; q++) {
        const int64_t j = j0 + q;
        if (j >= Ny) break;
        const int64_t t = i + Nx * j;
        bool ok;
        double inv;
        if (invden) {
            inv = invden[t];
            ok = G.mask[t] != 0 && inv > 0.0;
        } else {
            const double den = px[i] * py[j];
            ok = G.mask[t] != 0 && den > 1e-15;
            inv = ok ? 1.0 / den : 0.0;
        }
        pi_point_epilogue_inv<T>(G, t, i, j, F[q], ok, inv, b, ubar, tau, uo, vo, wo);
    }
}

// The same tiled column pass with the planes streamed into two LDS buffers by direct global → LDS
// loads (global_load_lds_dwordx4: lane l's 16 bytes land at the wave's LDS base + 16·l, no registers
// on the way), so plane f + 1 is in flight while plane f is summed. A wave's load covers two 64-column
// rows; a row beyond a non-periodic edge is written as zeros by its lanes instead. Needs an even Nx
// (16-byte aligned pairs); columns beyond Nx read the row's last pair (their outputs are dropped).
template <typename T, int TH>
__global__ __launch_bounds__(TH * 8) __attribute__((amdgpu_waves_per_eu(2, 4))) void k_sep_pi_col_async(
    GeomArgs G, const double *__restrict__ gy, int ly, const double *__restrict__ rp, const double *__restrict__ px,
    const double *__restrict__ py, const double *__restrict__ invden, double *__restrict__ ubar, T *__restrict__ tau,
    T *__restrict__ uo, T *__restrict__ vo, T *__restrict__ wo, int64_t ntx) {
    extern __shared__ __attribute__((aligned(16))) double smem[];   // taps, then two planes' 64 + 2·ly rows × 64
    constexpr int NT = TH * 8, NW = NT / 64;   // threads, waves: one 8-row band per wave
    const int taps = 2 * ly + 1, R = TH + 2 * ly, tpad = (taps + 1) & ~1;
    double *sg = smem, *buf0 = smem + tpad, *buf1 = buf0 + R * SEPC_W;
    const int64_t b = blockIdx.y;
    const int64_t tx = (int64_t)blockIdx.x % ntx, ty = (int64_t)blockIdx.x / ntx;
    const int64_t Nx = G.n[0], Ny = G.n[1], P = G.npts;
    const int c = threadIdx.x & (SEPC_W - 1), band = threadIdx.x / SEPC_W, wave = threadIdx.x >> 6;
    const int64_t i = tx * SEPC_W + c, j0 = ty * TH + band * SEPC_TJ;
    const double *r = rp + b * 5 * P;
    for (int e = threadIdx.x; e < taps; e += NT) sg[e] = gy[e];
    // lane l of a wave: row 2·pair + l / 32 of the pair, columns 2·(l % 32) and + 1
    const int lane = threadIdx.x & 63, half = lane >> 5, cc = 2 * (lane & 31);
    const int64_t ic = tx * SEPC_W + cc < Nx ? tx * SEPC_W + cc : Nx - 2;
    auto stage = [&](int f, double *buf) {
        for (int pr = wave; pr < R / 2; pr += NW) {
            const int k = 2 * pr + half;
            int64_t jw = ty * TH - ly + k;
            bool in = true;
            if (jw < 0 || jw >= Ny) {
                if (G.periodic[1]) jw = wrapi(jw, Ny);
                else in = false;
            }
            if (in)
                __builtin_amdgcn_global_load_lds((const void *)(r + f * P + jw * Nx + ic),
                                                 (void __attribute__((address_space(3))) *)(buf + 2 * pr * SEPC_W), 16,
                                                 0, 0);
            else
                *reinterpret_cast<double2 *>(buf + k * SEPC_W + cc) = make_double2(0.0, 0.0);
        }
    };
    double F[SEPC_TJ][5];
#pragma unroll
    for (int q = 0; q < SEPC_TJ; q++)
#pragma unroll
        for (int f = 0; f < 5; f++) F[q][f] = 0.0;
    stage(0, buf0);
#pragma unroll
    for (int f = 0; f < 5; f++) {
        __builtin_amdgcn_s_waitcnt(0);   // this wave's part of plane f has landed
        __syncthreads();                 // everyone's; and plane f − 1's readers are done with the other buffer
        if (f < 4) stage(f + 1, (f & 1) ? buf0 : buf1);
        double g[SEPC_TJ];
#pragma unroll
        for (int q = 0; q < SEPC_TJ; q++) g[q] = 0.0;
        const double *col = ((f & 1) ? buf1 : buf0) + band * SEPC_TJ * SEPC_W + c;
        for (int k = 0; k < taps + SEPC_TJ - 1; k++) {
#pragma unroll
            for (int q = SEPC_TJ - 1; q > 0; q--) g[q] = g[q - 1];
            g[0] = k < taps ? sg[k] : 0.0;
            const double x = col[k * SEPC_W];
#pragma unroll
            for (int q = 0; q < SEPC_TJ; q++) F[q][f] += g[q] * x;
        }
    }
    if (i >= Nx) return;
#pragma unroll
    for (int q = 0; q < SEPC_TJ; q++) {
        const int64_t j = j0 + q;
        if (j >= Ny) break;
        const int64_t t = i + Nx * j;
        bool ok;
        double inv;
        if (invden) {
            inv = invden[t];
            ok = G.mask[t] != 0 && inv > 0.0;
        } else {
            const double den = px[i] * py[j];
            ok = G.mask[t] != 0 && den > 1e-15;
            inv = ok ? 1.0 / den : 0.0;
        }
        pi_point_epilogue_inv<T>(G, t, i, j, F[q], ok, inv, b, ubar, tau, uo, vo, wo);
    }
}

// Column pass for grids with too few 64 × 64 tiles to fill the chip: one lane per column walking
// TJ output rows (each read row feeds every output row within reach), the planes read from memory.
template <typename T, int TJ>
__global__ __launch_bounds__(256) void k_sep_pi_col_lanes(GeomArgs G, const double *__restrict__ gy, int ly,
                                                          const double *__restrict__ rp, const double *__restrict__ px,
                                                          const double *__restrict__ py,
                                                          const double *__restrict__ invden, double *__restrict__ ubar,
                                                          T *__restrict__ tau, T *__restrict__ uo, T *__restrict__ vo,
                                                          T *__restrict__ wo, int64_t nchunk) {
    const int64_t b = blockIdx.y;
    const int64_t jt = (int64_t)blockIdx.x / nchunk;
    const int64_t i = ((int64_t)blockIdx.x - jt * nchunk) * 256 + threadIdx.x;
    const int64_t Nx = G.n[0], Ny = G.n[1], P = G.npts;
    if (i >= Nx) return;
    const int64_t j0 = jt * TJ;
    const double *r = rp + b * 5 * P;
    double F[TJ][5];
#pragma unroll
    for (int q = 0; q < TJ; q++)
#pragma unroll
        for (int f = 0; f < 5; f++) F[q][f] = 0.0;
    for (int64_t jj = j0 - ly; jj < j0 + TJ + ly; jj++) {
        int64_t jw = jj;
        if (jw < 0 || jw >= Ny) {
            if (!G.periodic[1]) continue;
            jw = wrapi(jw, Ny);
        }
        const int64_t qi = i + Nx * jw;
        double x[5];
#pragma unroll
        for (int f = 0; f < 5; f++) x[f] = r[f * P + qi];
#pragma unroll
        for (int q = 0; q < TJ; q++) {
            const int64_t d = jj - (j0 + q);
            if (d < -ly || d > ly) continue;
            const double g = gy[d + ly];
#pragma unroll
            for (int f = 0; f < 5; f++) F[q][f] += g * x[f];
        }
    }
#pragma unroll
    for (int q = 0; q < TJ; q++) {
        const int64_t j = j0 + q;
        if (j >= Ny) break;
        const int64_t t = i + Nx * j;
        bool ok;
        double inv;
        if (invden) {
            inv = invden[t];
            ok = G.mask[t] != 0 && inv > 0.0;
        } else {
            const double den = px[i] * py[j];
            ok = G.mask[t] != 0 && den > 1e-15;
            inv = ok ? 1.0 / den : 0.0;
        }
        pi_point_epilogue_inv<T>(G, t, i, j, F[q], ok, inv, b, ubar, tau, uo, vo, wo);
    }
}

void launch_separable_pi(const GeomArgs &G, const Plan &p, int dtype, const void *u, const void *v, int64_t nbatch,
                         double *rp, double *ubar, void *tau, void *uo, void *vo, void *wo, hipStream_t s) {
    const int64_t nchunk = (G.n[0] + 255) / 256;
    const int lx = p.sep_lim[0], ly = p.sep_lim[1];
    const size_t sh = (size_t)(2 * (256 + 2 * lx) + 2 * lx + 1) * sizeof(double);
    const size_t shc = (size_t)(2 * ly + 1 + (SEPC_H + 2 * ly) * SEPC_W) * sizeof(double);
    if (sh > 160 * 1024) fail(CGEF_ERR_UNSUPPORTED, "separable row half-width too large for the row pass");
    const int64_t ntx = (G.n[0] + SEPC_W - 1) / SEPC_W, nty = (G.n[1] + SEPC_H - 1) / SEPC_H;
    // the tiled column pass where its 64 × 64 tiles fill the chip and the tile fits in LDS; else a lane
    // per column, 8 output rows per lane where that still leaves ≥ 1024 workgroups, else one
    const bool tiled = ntx * nty * nbatch >= 1024 && shc <= 160 * 1024;
    // two plane buffers filled by direct global → LDS loads where they fit and Nx is even
    const size_t sha = (size_t)(((2 * ly + 2) & ~1) + 2 * (SEPC_AH + 2 * ly) * SEPC_W) * sizeof(double);
    const bool async = tiled && sha <= 160 * 1024 && G.n[0] % 2 == 0 && G.n[0] >= 2 &&
                       (reinterpret_cast<uintptr_t>(rp) & 15) == 0;
    const void *fr = dtype == CGEF_F32 ? (const void *)k_sep_pi_row<float> : (const void *)k_sep_pi_row<double>;
    const void *fc = dtype == CGEF_F32 ? (const void *)k_sep_pi_col<float> : (const void *)k_sep_pi_col<double>;
    if (sh > 64 * 1024) CGEF_HIP(hipFuncSetAttribute(fr, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    if (tiled && shc > 64 * 1024) CGEF_HIP(hipFuncSetAttribute(fc, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shc));
    const dim3 gr((unsigned)(nchunk * G.n[1]), (unsigned)nbatch);
    const dim3 gc((unsigned)(ntx * nty), (unsigned)nbatch);
    const bool tall = nchunk * ((G.n[1] + 7) / 8) * nbatch >= 1024;
    const int TJ = tall ? 8 : 1;
    const dim3 gl((unsigned)(nchunk * ((G.n[1] + TJ - 1) / TJ)), (unsigned)nbatch);
    auto col = [&](auto async_k, auto tiled_k, auto lanes8_k, auto lanes1_k, auto *tauT) {
        using TT = std::remove_pointer_t<decltype(tauT)>;
        if (async) {
            if (sha > 64 * 1024)
                CGEF_HIP(hipFuncSetAttribute((const void *)async_k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sha));
            const dim3 ga((unsigned)(ntx * ((G.n[1] + SEPC_AH - 1) / SEPC_AH)), (unsigned)nbatch);
            hipLaunchKernelGGL(async_k, ga, dim3(SEPC_AH * 8), sha, s, G, p.sep_g_d[1], ly, rp, p.sep_prof_d[0],
                               p.sep_prof_d[1], p.sep_invden_d, ubar, (TT *)tau, (TT *)uo, (TT *)vo, (TT *)wo, ntx);
        } else if (tiled)
            hipLaunchKernelGGL(tiled_k, gc, dim3(512), shc, s, G, p.sep_g_d[1], ly, rp, p.sep_prof_d[0],
                               p.sep_prof_d[1], p.sep_invden_d, ubar, (TT *)tau, (TT *)uo, (TT *)vo, (TT *)wo, ntx);
        else
            hipLaunchKernelGGL(tall ? lanes8_k : lanes1_k, gl, dim3(256), 0, s, G, p.sep_g_d[1], ly, rp,
                               p.sep_prof_d[0], p.sep_prof_d[1], p.sep_invden_d, ubar, (TT *)tau, (TT *)uo, (TT *)vo,
                               (TT *)wo, nchunk);
    };
    if (dtype == CGEF_F32) {
        hipLaunchKernelGGL(k_sep_pi_row<float>, gr, dim3(256), sh, s, G, p.sep_g_d[0], lx, (const float *)u,
                           (const float *)v, rp, nchunk);
        col(k_sep_pi_col_async<float, SEPC_AH>, k_sep_pi_col<float>, k_sep_pi_col_lanes<float, 8>, k_sep_pi_col_lanes<float, 1>,
            (float *)nullptr);
    } else {
        hipLaunchKernelGGL(k_sep_pi_row<double>, gr, dim3(256), sh, s, G, p.sep_g_d[0], lx, (const double *)u,
                           (const double *)v, rp, nchunk);
        col(k_sep_pi_col_async<double, SEPC_AH>, k_sep_pi_col<double>, k_sep_pi_col_lanes<double, 8>,
            k_sep_pi_col_lanes<double, 1>, (double *)nullptr);
    }
}

// --------------------------------------------------- unfused Π (baseline and separable engines) ----

// Diagnostics.jl:489-527 / 550-650 / 805-829 as separate elementwise steps in FP64: the frame
// velocity (rotated on the sphere, masked cells 0 there), its products, then τ and ū in the local
// frame from the filtered planes.
template <typename T>
__global__ __launch_bounds__(256) void k_unfused_prep(GeomArgs G, const void *u, const void *v, const void *w, double *vel) {
    const int64_t b = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= G.npts) return;
    const int64_t P = G.npts, Nx = G.n[0], Ny = G.n[1];
    const int64_t i = t % Nx, j = (t / Nx) % Ny;
    double U[3];
    frame_velocity<T>(G, u, v, w, b * P, t, i, j, U);
    if (G.mode == 1 && !G.mask[t]) U[0] = U[1] = U[2] = 0.0;
    double *o = vel + b * 3 * P;
    o[t] = U[0];
    o[P + t] = U[1];
    o[2 * P + t] = U[2];
}

void launch_unfused_prep(const GeomArgs &G, int dtype, const void *u, const void *v, const void *w, int64_t nbatch,
                         double *vel, hipStream_t s) {
    dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)nbatch);
    if (dtype == CGEF_F32) hipLaunchKernelGGL(k_unfused_prep<float>, grid, dim3(256), 0, s, G, u, v, w, vel);
    else hipLaunchKernelGGL(k_unfused_prep<double>, grid, dim3(256), 0, s, G, u, v, w, vel);
}

__global__ __launch_bounds__(256) void k_products(int64_t P, const double *vel, double *prod) {
    const int64_t b = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= P) return;
    const double *U = vel + b * 3 * P;
    const double a = U[t], c = U[P + t], d = U[2 * P + t];
    double *o = prod + b * 6 * P;
    o[t] = a * a;
    o[P + t] = a * c;
    o[2 * P + t] = a * d;
    o[3 * P + t] = c * c;
    o[4 * P + t] = c * d;
    o[5 * P + t] = d * d;
}

void launch_products(const GeomArgs &G, const double *vel, double *prod, int64_t nbatch, hipStream_t s) {
    dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)nbatch);
    hipLaunchKernelGGL(k_products, grid, dim3(256), 0, s, G.npts, vel, prod);
}

__global__ __launch_bounds__(256) void k_unfused_tau(GeomArgs G, const double *velf, const double *prodf, double *ubar,
                                                     double *tau) {
    const int64_t b = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t P = G.npts;
    if (t >= P) return;
    const double *Uf = velf + b * 3 * P;
    const double *F = prodf + b * 6 * P;
    double *ub = ubar + b * 3 * P;
    double *tb = tau + b * 6 * P;
    const double u0 = Uf[t], u1 = Uf[P + t], u2 = Uf[2 * P + t];
    const double tc[6] = {F[t] - u0 * u0, F[P + t] - u0 * u1, F[2 * P + t] - u0 * u2,
                          F[3 * P + t] - u1 * u1, F[4 * P + t] - u1 * u2, F[5 * P + t] - u2 * u2};
    if (G.mode == 1) {
        if (!G.mask[t]) {
            ub[t] = ub[P + t] = ub[2 * P + t] = 0.0;
            tb[t] = tb[P + t] = tb[2 * P + t] = 0.0;
            return;
        }
        const int64_t Nx = G.n[0], Ny = G.n[1];
        const int64_t i = t % Nx, j = (t / Nx) % Ny;
        const double sl = G.trig[i], cl = G.trig[Nx + i];
        const double sp = G.trig[2 * Nx + j], cp = G.trig[2 * Nx + Ny + j];
        const double ee[3] = {-sl, cl, 0.0}, en[3] = {-sp * cl, -sp * sl, cp}, er[3] = {cp * cl, cp * sl, sp};
        ub[t] = ee[0] * u0 + ee[1] * u1;
        ub[P + t] = en[0] * u0 + en[1] * u1 + en[2] * u2;
        ub[2 * P + t] = er[0] * u0 + er[1] * u1 + er[2] * u2;
        const double T3[3][3] = {{tc[0], tc[1], tc[2]}, {tc[1], tc[3], tc[4]}, {tc[2], tc[4], tc[5]}};
        double Te[3], Tn[3];
        for (int rr = 0; rr < 3; rr++) {
            Te[rr] = T3[rr][0] * ee[0] + T3[rr][1] * ee[1] + T3[rr][2] * ee[2];
            Tn[rr] = T3[rr][0] * en[0] + T3[rr][1] * en[1] + T3[rr][2] * en[2];
        }
        tb[t] = ee[0] * Te[0] + ee[1] * Te[1] + ee[2] * Te[2];
        tb[P + t] = ee[0] * Tn[0] + ee[1] * Tn[1] + ee[2] * Tn[2];
        tb[2 * P + t] = en[0] * Tn[0] + en[1] * Tn[1] + en[2] * Tn[2];
    } else if (G.mode == 0) {
        ub[t] = u0;
        ub[P + t] = u1;
        ub[2 * P + t] = 0.0;
        tb[t] = tc[0];
        tb[P + t] = tc[1];
        tb[2 * P + t] = tc[3];
    } else {
        ub[t] = u0;
        ub[P + t] = u1;
        ub[2 * P + t] = u2;
        for (int q = 0; q < 6; q++) tb[q * P + t] = tc[q];
    }
}

void launch_unfused_tau(const GeomArgs &G, const double *velf, const double *prodf, double *ubar, double *tau,
                        int64_t nbatch, hipStream_t s) {
    dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)nbatch);
    hipLaunchKernelGGL(k_unfused_tau, grid, dim3(256), 0, s, G, velf, prodf, ubar, tau);
}

template <typename T>
__global__ __launch_bounds__(256) void k_cast(const double *in, T *out, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) out[t] = (T)in[t];
}

void launch_cast(int dtype_out, const double *in, void *out, int64_t n, hipStream_t s) {
    dim3 grid((unsigned)((n + 255) / 256));
    if (dtype_out == CGEF_F32) hipLaunchKernelGGL(k_cast<float>, grid, dim3(256), 0, s, in, (float *)out, n);
    else hipLaunchKernelGGL(k_cast<double>, grid, dim3(256), 0, s, in, (double *)out, n);
}

}  // namespace cgef
