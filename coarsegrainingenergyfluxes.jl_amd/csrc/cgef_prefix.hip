// Prefix-sum top-hat engine (Filtering.jl:751-1141) on gfx950.
//
// A top-hat weight is constant inside its support and the cell measure of a rectilinear grid is
// separable (area = wx[i]·wy[j]), so the footprint sum over one source row is an interval sum of
// mask·f·wx along axis 1: O(1) from that row's prefix sum. Two kernels:
//   k_prefix_rows  — one workgroup per (slice, row): inclusive scan of mask·f·wx in FP64, in tiles of
//                    256 coalesced elements (wave scan with shuffles, then across the 4 waves);
//   k_prefix_apply — one lane per target point, one workgroup per 256 targets of a row: for each row
//                    of the band, numerator and denominator as prefix differences.
// On a periodic axis the prefix over the tripled axis is (k div N)·P[N] + P[k mod N], so only the
// N + 1 real-axis prefixes are stored.
#include <hip/hip_runtime.h>

#include "cgef_pi_point.hpp"

namespace cgef {

__device__ __forceinline__ double wave_incl_scan(double v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// P[s][row][0..N]: P[0] = 0, P[i+1] = Σ_{i' ≤ i} mask·f·wx (Filtering.jl:913-945).
template <typename T>
__global__ __launch_bounds__(256) void k_prefix_rows(PrefixArgs A, GeomArgs G, const T *__restrict__ in,
                                                     double *__restrict__ P) {
    __shared__ double wsum[4];
    const int64_t row = blockIdx.x, s = blockIdx.y;
    const int64_t N = A.nx;
    const int64_t moff = row * N;            // the mask is one plane for every slice
    const int64_t off = s * G.npts + moff;
    double *Pr = P + (s * A.ny + row) * (N + 1);
    if (threadIdx.x == 0) Pr[0] = 0.0;
    double carry = 0.0;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int64_t base = 0; base < N; base += 256) {
        const int64_t i = base + threadIdx.x;
        double v = 0.0;
        if (i < N && G.mask[moff + i]) v = (double)in[off + i] * A.wx[i];
        v = wave_incl_scan(v);
        if (lane == 63) wsum[wid] = v;
        __syncthreads();
        double pre = carry;
        for (int q = 0; q < wid; q++) pre += wsum[q];
        const double tile = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (i < N) Pr[i + 1] = pre + v;
        carry += tile;
        __syncthreads();
    }
}

// Prefix of the tripled axis at extended index k ∈ [0, ne].
__device__ __forceinline__ double pext(const double *Pr, int64_t k, int64_t N) {
    const int64_t r = k / N;
    return (double)r * Pr[N] + Pr[k - r * N];
}

// First extended index with xe[k] ≥ v (the reference's lower pointer: it skips xe < x − hw).
__device__ __forceinline__ int64_t first_ge(const PrefixArgs &A, double v) {
    int64_t lo = 0, hi = A.ne;
    if (A.uniform_x) {
        double gs = ceil((v - A.x0) / A.dx);
        gs = gs < 0.0 ? 0.0 : (gs > (double)A.ne ? (double)A.ne : gs);
        int64_t k = (int64_t)gs;
        while (k > 0 && A.xe[k - 1] >= v) k--;
        while (k < A.ne && A.xe[k] < v) k++;
        return k;
    }
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (A.xe[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Number of extended entries with xe[k] ≤ v (the reference's upper pointer: last k is this − 1).
__device__ __forceinline__ int64_t count_le(const PrefixArgs &A, double v) {
    int64_t lo = 0, hi = A.ne;
    if (A.uniform_x) {
        double gs = floor((v - A.x0) / A.dx) + 1.0;
        gs = gs < 0.0 ? 0.0 : (gs > (double)A.ne ? (double)A.ne : gs);
        int64_t k = (int64_t)gs;
        while (k > 0 && A.xe[k - 1] > v) k--;
        while (k < A.ne && A.xe[k] <= v) k++;
        return k;
    }
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (A.xe[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Filtering.jl:974-1078: out = Σ_b wy·(numerator interval) / Σ_b wy·(denominator interval) at active
// targets with a denominator above 1e-15, else 0.
template <typename T>
__global__ __launch_bounds__(256) void k_prefix_apply(PrefixArgs A, GeomArgs G, const double *__restrict__ P,
                                                      T *__restrict__ out, int64_t nchunk) {
    const int64_t s = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x / nchunk;
    const int64_t i = ((int64_t)blockIdx.x - j * nchunk) * 256 + threadIdx.x;
    const int64_t N = A.nx;
    if (i >= N) return;
    const int64_t t = j * N + i;
    const double *Ps = P + s * A.ny * (N + 1);
    const double xi = G.coord[0][i];
    double num = 0.0, den = 0.0;
    for (int b = 0; b < A.nb; b++) {
        const int64_t e = j * A.nb + b;
        const int jj = A.jj[e];
        if (jj < 0) continue;
        const double wyj = A.wy[e];
        const int wl = A.wl[e];
        const double *Pr = Ps + (int64_t)jj * (N + 1);
        const double *Dr = A.deform ? A.pd + (int64_t)jj * (N + 1) : A.pwx;
        if (wl == -2) {
            num += wyj * (Pr[N] - Pr[0]);
            den += wyj * (Dr[N] - Dr[0]);
        } else if (wl >= 0) {
            const int64_t lo = i - wl < 0 ? 0 : i - wl, hi = i + wl > N - 1 ? N - 1 : i + wl;
            num += wyj * (Pr[hi + 1] - Pr[lo]);
            den += wyj * (Dr[hi + 1] - Dr[lo]);
        } else {
            const double hw = A.hw[e];
            const int64_t lo = first_ge(A, xi - hw), hi1 = count_le(A, xi + hw);
            if (hi1 > lo) {
                num += wyj * (pext(Pr, hi1, N) - pext(Pr, lo, N));
                den += wyj * (pext(Dr, hi1, N) - pext(Dr, lo, N));
            }
        }
    }
    out[s * G.npts + t] = (T)((G.mask[t] && den > 1e-15) ? num / den : 0.0);
}

// Π path: the planes are formed from (u, v[, w]) inside the scan (one workgroup per (row, plane
// column), columns fastest so the planes of a row re-read its velocities from L2), and the apply
// finds each row's interval once for all NF planes, then runs the per-point epilogue.
__global__ __launch_bounds__(256) void k_prefix_pi_rows(PrefixArgs A, GeomArgs G, int dtype, const void *u, const void *v,
                                                        const void *w, int nf, int64_t ncol, double *__restrict__ P) {
    __shared__ double wsum[4];
    const int64_t row = (int64_t)blockIdx.x / ncol, col = (int64_t)blockIdx.x - row * ncol;
    const int64_t b = col / nf;
    const int f = (int)(col - b * nf);
    const int64_t N = A.nx, j = row % A.ny;
    const int64_t moff = row * N, off = b * G.npts;
    double *Pr = P + (col * A.ny + row) * (N + 1);
    if (threadIdx.x == 0) Pr[0] = 0.0;
    double carry = 0.0;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int64_t base = 0; base < N; base += 256) {
        const int64_t i = base + threadIdx.x;
        double val = 0.0;
        if (i < N && G.mask[moff + i]) {
            double q[3];
            frame_vel(G, dtype, u, v, w, off, moff + i, i, j, q);
            val = pi_plane_value(G.mode, f, q) * A.wx[i];
        }
        val = wave_incl_scan(val);
        if (lane == 63) wsum[wid] = val;
        __syncthreads();
        double pre = carry;
        for (int q = 0; q < wid; q++) pre += wsum[q];
        const double tile = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (i < N) Pr[i + 1] = pre + val;
        carry += tile;
        __syncthreads();
    }
}

template <typename T, int NF>
__global__ __launch_bounds__(256) void k_prefix_pi_apply(PrefixArgs A, GeomArgs G, const double *__restrict__ P,
                                                         int64_t nchunk, double *__restrict__ ubar, T *__restrict__ tau,
                                                         T *__restrict__ uo, T *__restrict__ vo, T *__restrict__ wo) {
    const int64_t b = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x / nchunk;
    const int64_t i = ((int64_t)blockIdx.x - j * nchunk) * 256 + threadIdx.x;
    const int64_t N = A.nx, NR = A.ny * (N + 1);
    if (i >= N) return;
    const int64_t t = j * N + i;
    const double *Pb = P + b * NF * NR;
    const double xi = G.coord[0][i];
    double F[NF], den = 0.0;
#pragma unroll
    for (int f = 0; f < NF; f++) F[f] = 0.0;
    for (int bb = 0; bb < A.nb; bb++) {
        const int64_t e = j * A.nb + bb;
        const int jj = A.jj[e];
        if (jj < 0) continue;
        const double wyj = A.wy[e];
        const int wl = A.wl[e];
        const double *Pr = Pb + (int64_t)jj * (N + 1);
        const double *Dr = A.deform ? A.pd + (int64_t)jj * (N + 1) : A.pwx;
        int64_t lo, hi1;  // interval [lo, hi1) of the extended axis
        if (wl == -2) {
            lo = 0;
            hi1 = N;
        } else if (wl >= 0) {
            lo = i - wl < 0 ? 0 : i - wl;
            hi1 = (i + wl > N - 1 ? N - 1 : i + wl) + 1;
        } else {
            const double hw = A.hw[e];
            lo = first_ge(A, xi - hw);
            hi1 = count_le(A, xi + hw);
            if (hi1 <= lo) continue;
        }
        den += wyj * (pext(Dr, hi1, N) - pext(Dr, lo, N));
#pragma unroll
        for (int f = 0; f < NF; f++) F[f] += wyj * (pext(Pr + f * NR, hi1, N) - pext(Pr + f * NR, lo, N));
    }
    pi_point_epilogue<T>(G, t, i, j % G.n[1], F, den, b, ubar, tau, uo, vo, wo);
}

void launch_prefix_pi(const GeomArgs &G, const Plan &p, int dtype, const void *u, const void *v, const void *w,
                      int64_t nbatch, double *Pbuf, double *ubar, void *tau, void *uo, void *vo, void *wo,
                      hipStream_t s) {
    const PrefixArgs &A = p.pre.args;
    const int nf = G.mode == 0 ? 5 : 9;
    const int64_t ncol = nf * nbatch;
    const int64_t nchunk = (A.nx + 255) / 256;
    hipLaunchKernelGGL(k_prefix_pi_rows, dim3((unsigned)(A.ny * ncol)), dim3(256), 0, s, A, G, dtype, u, v, w, nf, ncol,
                       Pbuf);
    const dim3 ga((unsigned)(nchunk * A.ny), (unsigned)nbatch);
#define CGEF_PPA(TT, NN)                                                                                           \
    hipLaunchKernelGGL((k_prefix_pi_apply<TT, NN>), ga, dim3(256), 0, s, A, G, Pbuf, nchunk, ubar, (TT *)tau, (TT *)uo, \
                       (TT *)vo, (TT *)wo)
    if (dtype == CGEF_F32) {
        if (nf == 5) CGEF_PPA(float, 5);
        else CGEF_PPA(float, 9);
    } else {
        if (nf == 5) CGEF_PPA(double, 5);
        else CGEF_PPA(double, 9);
    }
#undef CGEF_PPA
}

void launch_prefix(const GeomArgs &G, const Plan &p, int dtype, const void *in, void *out, int64_t nslices,
                   double *Pbuf, hipStream_t s) {
    const PrefixArgs &A = p.pre.args;
    const int64_t nchunk = (A.nx + 255) / 256;
    const dim3 gs((unsigned)A.ny, (unsigned)nslices), ga((unsigned)(nchunk * A.ny), (unsigned)nslices);
    if (dtype == CGEF_F32) {
        hipLaunchKernelGGL(k_prefix_rows<float>, gs, dim3(256), 0, s, A, G, (const float *)in, Pbuf);
        hipLaunchKernelGGL(k_prefix_apply<float>, ga, dim3(256), 0, s, A, G, Pbuf, (float *)out, nchunk);
    } else {
        hipLaunchKernelGGL(k_prefix_rows<double>, gs, dim3(256), 0, s, A, G, (const double *)in, Pbuf);
        hipLaunchKernelGGL(k_prefix_apply<double>, ga, dim3(256), 0, s, A, G, Pbuf, (double *)out, nchunk);
    }
}

}  // namespace cgef
