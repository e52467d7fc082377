// Ring engine: the banded footprint evaluated in Fourier space along the periodic x (longitude) axis.
//
// For a target row r and a source row rr the footprint restricted to (r, rr) is a circular
// correlation along x with weights w_{r,rr}(d) that are symmetric in d (|d| = its run half-width,
// the ring run −N/2..N/2−1 being symmetric modulo N). Its DFT Ŵ_{r,rr}(m) = Σ_d w(d) cos(2πmd/N) is
// real, so the whole filter becomes, per Fourier mode m, a banded real-weight contraction over rows:
//     OUT_r(m) = Σ_{(dy,dz)} Ŵ_{band(r),(dy,dz)}(m) · F_{r+(dy,dz)}(m).
// Same operator as the direct sum (same weights, truncation, ring rule, wrap and clipping), evaluated
// at O(N log N + N·rows·band) instead of O(N·taps). Everything here is FP64, so the one-pass form
// τ = F[qq] − q̄q̄ keeps ~13 significant digits even where τ is 10⁻³ of the resolved energy.
//
// Kernels: k_fft_rows (two real rows per complex FFT, mixed-radix Stockham in LDS), k_contract
// (source rows staged in LDS per column chunk, Ŵ read coalesced along rows), k_ifft_rows, and the
// prep/epilogue kernels that turn velocities into the filtered fields and τ, ū.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "cgef_ring.hpp"

namespace cgef {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// Twiddle tw[k] = exp(−2πik/N) (conjugated for the inverse).
__device__ __forceinline__ double2 twid(const double2 *tw, int64_t k, int n, bool inv) {
    double2 t = tw[k % n];
    return inv ? cconj(t) : t;
}

// In-LDS DFT of length R on v (R ≤ RING_MAX_RADIX), roots from the N-point table.
template <int R>
__device__ __forceinline__ void dft_small(double2 *v, const double2 *tw, int n, bool inv) {
    if (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (R == 4) {
        const double2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        const double2 c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        // −i·d forward, +i·d inverse
        const double2 dj = inv ? make_double2(-d.y, d.x) : make_double2(d.y, -d.x);
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(b, dj);
        v[3] = csub(b, dj);
    } else {
        double2 o[R];
        const int step = n / R;
#pragma unroll
        for (int a = 0; a < R; a++) {
            double2 s = v[0];
#pragma unroll
            for (int b = 1; b < R; b++) s = cadd(s, cmul(v[b], twid(tw, (int64_t)((a * b) % R) * step, n, inv)));
            o[a] = s;
        }
#pragma unroll
        for (int a = 0; a < R; a++) v[a] = o[a];
    }
}

template <int R>
__device__ void stockham_stage(const double2 *src, double2 *dst, int n, int Ns, const double2 *tw, bool inv,
                               int nseq) {
    const int nb = n / R;
    for (int t = threadIdx.x; t < nb * nseq; t += blockDim.x) {
        const int s = t / nb, j = t - s * nb;
        const double2 *x = src + s * n;
        double2 *y = dst + s * n;
        const int k = j % Ns;
        const int step = n / (Ns * R);
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const double2 a = x[j + r * nb];
            v[r] = r == 0 ? a : cmul(a, twid(tw, (int64_t)r * k * step, n, inv));
        }
        dft_small<R>(v, tw, n, inv);
        const int idx = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) y[idx + r * Ns] = v[r];
    }
}

// nseq complex sequences of length n in buf[0..), Stockham ping-pong with buf2; returns the buffer
// that holds the result.
__device__ double2 *fft_lds(double2 *buf, double2 *buf2, const RingFft &F, const double2 *tw, bool inv, int nseq) {
    double2 *src = buf, *dst = buf2;
    int Ns = 1;
    for (int st = 0; st < F.nstage; st++) {
        const int R = F.radix[st];
        __syncthreads();
        switch (R) {
            case 2: stockham_stage<2>(src, dst, F.n, Ns, tw, inv, nseq); break;
            case 3: stockham_stage<3>(src, dst, F.n, Ns, tw, inv, nseq); break;
            case 4: stockham_stage<4>(src, dst, F.n, Ns, tw, inv, nseq); break;
            case 5: stockham_stage<5>(src, dst, F.n, Ns, tw, inv, nseq); break;
            case 7: stockham_stage<7>(src, dst, F.n, Ns, tw, inv, nseq); break;
            case 8: stockham_stage<8>(src, dst, F.n, Ns, tw, inv, nseq); break;
            case 11: stockham_stage<11>(src, dst, F.n, Ns, tw, inv, nseq); break;
            case 13: stockham_stage<13>(src, dst, F.n, Ns, tw, inv, nseq); break;
            default: break;  // factorization is restricted to these radices on the host
        }
        Ns *= R;
        double2 *t = src;
        src = dst;
        dst = t;
    }
    __syncthreads();
    return src;
}

// Forward: real rows (r0, r0+1) of column `col` (xin[col][row][i]) → XF[m][col][row], m ≤ N/2.
__global__ __launch_bounds__(256) void k_fft_rows(RingFft F, const double2 *__restrict__ tw, const double *__restrict__ xin,
                                                  double2 *__restrict__ XF, int64_t nrows, int64_t ncol) {
    extern __shared__ double2 sm[];
    const int n = F.n;
    const int64_t col = blockIdx.y;
    const int64_t r0 = 2 * (int64_t)blockIdx.x, r1 = r0 + 1;
    const bool has1 = r1 < nrows;
    const double *x0 = xin + (col * nrows + r0) * n;
    const double *x1 = has1 ? x0 + n : nullptr;
    for (int i = threadIdx.x; i < n; i += blockDim.x) sm[i] = make_double2(x0[i], has1 ? x1[i] : 0.0);
    const double2 *Z = fft_lds(sm, sm + n, F, tw, false, 1);
    const int mh = n / 2 + 1;
    for (int m = threadIdx.x; m < mh; m += blockDim.x) {
        const double2 a = Z[m], b = cconj(Z[(n - m) % n]);
        const double2 X = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
        const double2 d = csub(a, b);
        const double2 Y = make_double2(0.5 * d.y, -0.5 * d.x);  // (a − b)/(2i)
        double2 *o = XF + ((int64_t)m * ncol + col) * nrows;
        o[r0] = X;
        if (has1) o[r1] = Y;
    }
}

// Inverse: OUT[m][col][row] (Hermitian halves) → yout[col][row][i], normalized by 1/N.
__global__ __launch_bounds__(256) void k_ifft_rows(RingFft F, const double2 *__restrict__ tw, const double2 *__restrict__ OUT,
                                                   double *__restrict__ yout, int64_t nrows, int64_t ncol) {
    extern __shared__ double2 sm[];
    const int n = F.n;
    const int64_t col = blockIdx.y;
    const int64_t r0 = 2 * (int64_t)blockIdx.x, r1 = r0 + 1;
    const bool has1 = r1 < nrows;
    const int mh = n / 2 + 1;
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        const bool lo = m < mh;
        const int mm = lo ? m : n - m;
        const double2 *o = OUT + ((int64_t)mm * ncol + col) * nrows;
        double2 X = o[r0];
        double2 Y = has1 ? o[r1] : make_double2(0.0, 0.0);
        if (!lo) {
            X = cconj(X);
            Y = cconj(Y);
        }
        sm[m] = make_double2(X.x - Y.y, X.y + Y.x);  // X + iY
    }
    const double2 *z = fft_lds(sm, sm + n, F, tw, true, 1);
    const double s = 1.0 / n;
    double *y0 = yout + (col * nrows + r0) * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        y0[i] = z[i].x * s;
        if (has1) y0[n + i] = z[i].y * s;
    }
}

// Banded contraction per Fourier mode m. Block: 256 target rows of one plane (j0..j0+255, level k)
// × a chunk of CC columns; source rows of each dz plane staged in LDS (zero outside a non-periodic
// axis), Ŵ read along rows (coalesced on the sphere, broadcast on translation-invariant grids).
template <int CC>
__global__ __launch_bounds__(256) void k_contract(RingArgs A, const double2 *__restrict__ XF, double2 *__restrict__ OUT,
                                                  int64_t ncol) {
    extern __shared__ double2 st[];
    const int64_t Ny = A.ny, Nz = A.nz;
    const int64_t c0 = (int64_t)blockIdx.x * CC;
    const int64_t nyb = (Ny + 255) / 256;
    const int64_t k = blockIdx.y / nyb, j0 = (blockIdx.y % nyb) * 256;
    const int64_t m = blockIdx.z;
    const int tid = threadIdx.x;
    const int64_t j = j0 + tid;
    const bool valid = j < Ny;
    const int64_t r = valid ? j + Ny * k : 0;
    const int band = valid ? A.band_of_row[r] : 0;
    const int span = 256 + A.dy_max - A.dy_min;
    const int ncc = (int)((ncol - c0) < CC ? (ncol - c0) : CC);
    double2 acc[CC];
#pragma unroll
    for (int c = 0; c < CC; c++) acc[c] = make_double2(0.0, 0.0);
    const int64_t mrow = (int64_t)m * A.noff;
    for (int g = 0; g < A.ndz; g++) {
        const int dz = A.dz_val[g];
        int64_t kk = k + dz;
        if (kk < 0 || kk >= Nz) {
            if (!A.pz) continue;
            kk = ((kk % Nz) + Nz) % Nz;
        }
        __syncthreads();
        for (int q = tid; q < span * CC; q += 256) {
            const int c = q / span, s = q - c * span;
            int64_t jj = j0 + A.dy_min + s;
            double2 v = make_double2(0.0, 0.0);
            bool in = true;
            if (jj < 0 || jj >= Ny) {
                if (A.py) jj = ((jj % Ny) + Ny) % Ny;
                else in = false;
            }
            if (in && c < ncc) v = XF[(m * ncol + c0 + c) * A.nrows + jj + Ny * kk];
            st[c * span + s] = v;
        }
        __syncthreads();
        if (valid) {
            for (int o = A.dz_off0[g]; o < A.dz_off0[g + 1]; o++) {
                const double w = A.what[(mrow + o) * A.nbands + band];
                if (w == 0.0) continue;
                const int s = tid + A.off_dy[o] - A.dy_min;
#pragma unroll
                for (int c = 0; c < CC; c++) {
                    const double2 v = st[c * span + s];
                    acc[c].x = fma(w, v.x, acc[c].x);
                    acc[c].y = fma(w, v.y, acc[c].y);
                }
            }
        }
    }
    if (!valid) return;
    for (int c = 0; c < ncc; c++) OUT[(m * ncol + c0 + c) * A.nrows + r] = acc[c];
}

// Ŵ for every (band, offset): the run's weights laid out circularly on a length-N row, FFT'd two
// rows at a time; the real part is kept (the weights are symmetric, so the spectrum is real).
__global__ __launch_bounds__(256) void k_build_what(RingFft F, const double2 *__restrict__ tw, RingArgs A,
                                                    const Run *__restrict__ runs, const int32_t *__restrict__ run_off,
                                                    const double *__restrict__ w64, int64_t nruns, double *__restrict__ what) {
    extern __shared__ double2 sm[];
    const int n = F.n;
    const int64_t q0 = 2 * (int64_t)blockIdx.x, q1 = q0 + 1;
    for (int i = threadIdx.x; i < n; i += blockDim.x) sm[i] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int p = 0; p < 2; p++) {
        const int64_t q = q0 + p;
        if (q >= nruns) break;
        const Run R = runs[q];
        const int len = R.hi - R.lo + 1;
        if (len <= n) {
            // each column at most once: parallel
            for (int e = threadIdx.x; e < len; e += blockDim.x) {
                const int idx = (((R.lo + e) % n) + n) % n;
                const double wv = w64[R.woff + e];
                if (p == 0) sm[idx].x = wv;
                else sm[idx].y = wv;
            }
        } else if (threadIdx.x == 0) {
            // a run longer than the ring (Cartesian rad > L) wraps more than once: accumulate
            for (int e = 0; e < len; e++) {
                const int idx = (((R.lo + e) % n) + n) % n;
                const double wv = w64[R.woff + e];
                if (p == 0) sm[idx].x += wv;
                else sm[idx].y += wv;
            }
        }
        __syncthreads();
    }
    const double2 *Z = fft_lds(sm, sm + n, F, tw, false, 1);
    const int mh = n / 2 + 1;
    for (int m = threadIdx.x; m < mh; m += blockDim.x) {
        const double2 a = Z[m], b = cconj(Z[(n - m) % n]);
        for (int p = 0; p < 2; p++) {
            const int64_t q = q0 + p;
            if (q >= nruns) break;
            // real parts of X = (a + b)/2 and Y = (a − b)/(2i)
            const double val = p == 0 ? 0.5 * (a.x + b.x) : 0.5 * (a.y - b.y);
            const int band = run_off[2 * q], o = run_off[2 * q + 1];
            what[((int64_t)m * A.noff + o) * A.nbands + band] = val;
        }
    }
}

// ---- prep: shifted masked components (planetary frame on the sphere), their products, the mask --

__device__ __forceinline__ void frame_vel(const GeomArgs &G, int dtype, const void *u, const void *v, const void *w,
                                          int64_t off, int64_t t, int64_t i, int64_t j, double U[3]) {
    const double uu = dtype == CGEF_F32 ? (double)((const float *)u)[off + t] : ((const double *)u)[off + t];
    const double vv = dtype == CGEF_F32 ? (double)((const float *)v)[off + t] : ((const double *)v)[off + t];
    if (G.mode == 1) {
        const double sl = G.trig[i], cl = G.trig[G.n[0] + i];
        const double sp = G.trig[2 * G.n[0] + j], cp = G.trig[2 * G.n[0] + G.n[1] + j];
        U[0] = -sl * uu - sp * cl * vv;
        U[1] = cl * uu - sp * sl * vv;
        U[2] = cp * vv;
    } else {
        U[0] = uu;
        U[1] = vv;
        U[2] = G.mode == 2 ? (dtype == CGEF_F32 ? (double)((const float *)w)[off + t] : ((const double *)w)[off + t]) : 0.0;
    }
}

// Field order per slice: q (NC), products (NP: xx xy [xz] yy [yz zz]), mask (if masked).
__global__ __launch_bounds__(256) void k_ring_prep(GeomArgs G, int dtype, const void *u, const void *v, const void *w,
                                                   const double *cmean, int nf, int masked, double *xin) {
    const int64_t b = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t P = G.npts;
    if (t >= P) return;
    const int64_t Nx = G.n[0], Ny = G.n[1];
    const int64_t i = t % Nx, j = (t / Nx) % Ny;
    const bool act = G.mask[t] != 0;
    double q[3] = {0.0, 0.0, 0.0};
    if (act) {
        double U[3];
        frame_vel(G, dtype, u, v, w, b * P, t, i, j, U);
        for (int a = 0; a < 3; a++) q[a] = U[a] - cmean[3 * b + a];
    }
    double *o = xin + b * nf * P;
    if (G.mode == 0) {
        o[t] = q[0];
        o[P + t] = q[1];
        o[2 * P + t] = q[0] * q[0];
        o[3 * P + t] = q[0] * q[1];
        o[4 * P + t] = q[1] * q[1];
        if (masked) o[5 * P + t] = act ? 1.0 : 0.0;
    } else {
        o[t] = q[0];
        o[P + t] = q[1];
        o[2 * P + t] = q[2];
        o[3 * P + t] = q[0] * q[0];
        o[4 * P + t] = q[0] * q[1];
        o[5 * P + t] = q[0] * q[2];
        o[6 * P + t] = q[1] * q[1];
        o[7 * P + t] = q[1] * q[2];
        o[8 * P + t] = q[2] * q[2];
        if (masked) o[9 * P + t] = act ? 1.0 : 0.0;
    }
}

// τ and ū from the filtered fields (numerators), the Galilean correction of the mean shift, and the
// local frame on the sphere (Diagnostics.jl:596-650).
template <typename T>
__global__ __launch_bounds__(256) void k_ring_epilogue(GeomArgs G, const double *__restrict__ yout, const double *__restrict__ D_row,
                                                       const double *__restrict__ cmean, int nf, int masked, int strategy,
                                                       double *__restrict__ ubar, T *__restrict__ tau) {
    const int64_t b = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t P = G.npts;
    if (t >= P) return;
    const int64_t Nx = G.n[0], Ny = G.n[1];
    const int64_t i = t % Nx, r = t / Nx, j = r % Ny;
    const double *y = yout + b * nf * P;
    const int NC = G.mode == 0 ? 2 : 3;
    const double D = D_row[r];
    const double msum = masked ? y[(nf - 1) * P + t] : D;     // Σ_active w
    const double den = strategy == CGEF_DEFORMABLE ? msum : D;
    const bool ok = G.mask[t] != 0 && den > 1e-15;
    double U[3] = {0.0, 0.0, 0.0}, tc[6] = {0, 0, 0, 0, 0, 0};
    if (ok) {
        const double *c = cmean + 3 * b;
        double qb[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < NC; a++) qb[a] = y[a * P + t] / den;
        double Pq[6] = {0, 0, 0, 0, 0, 0};
        if (NC == 2) {
            Pq[0] = y[2 * P + t] / den;
            Pq[1] = y[3 * P + t] / den;
            Pq[3] = y[4 * P + t] / den;
        } else {
            for (int q6 = 0; q6 < 6; q6++) Pq[q6] = y[(3 + q6) * P + t] / den;
        }
        const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
        const double om = (strategy == CGEF_ZEROFILL && masked) ? 1.0 - msum / D : 0.0;  // 1 − m
        const double mm = 1.0 - om;
        for (int a = 0; a < 3; a++) U[a] = qb[a] + c[a] * (strategy == CGEF_ZEROFILL ? mm : 1.0);
        for (int q6 = 0; q6 < 6; q6++) {
            const int a = ia[q6], bb = ib[q6];
            tc[q6] = Pq[q6] - qb[a] * qb[bb] + om * (c[a] * U[bb] + c[bb] * U[a]) - c[a] * c[bb] * mm * om;
        }
    }
    double *ub = ubar + b * 3 * P;
    T *tb = tau + b * 6 * P;
    if (G.mode == 1) {
        const double sl = G.trig[i], cl = G.trig[Nx + i];
        const double sp = G.trig[2 * Nx + j], cp = G.trig[2 * Nx + Ny + j];
        const double ee[3] = {-sl, cl, 0.0}, en[3] = {-sp * cl, -sp * sl, cp}, er[3] = {cp * cl, cp * sl, sp};
        ub[t] = ee[0] * U[0] + ee[1] * U[1];
        ub[P + t] = en[0] * U[0] + en[1] * U[1] + en[2] * U[2];
        ub[2 * P + t] = er[0] * U[0] + er[1] * U[1] + er[2] * U[2];
        const double T3[3][3] = {{tc[0], tc[1], tc[2]}, {tc[1], tc[3], tc[4]}, {tc[2], tc[4], tc[5]}};
        double Te[3], Tn[3];
        for (int rr = 0; rr < 3; rr++) {
            Te[rr] = T3[rr][0] * ee[0] + T3[rr][1] * ee[1] + T3[rr][2] * ee[2];
            Tn[rr] = T3[rr][0] * en[0] + T3[rr][1] * en[1] + T3[rr][2] * en[2];
        }
        tb[t] = (T)(ee[0] * Te[0] + ee[1] * Te[1] + ee[2] * Te[2]);
        tb[P + t] = (T)(ee[0] * Tn[0] + ee[1] * Tn[1] + ee[2] * Tn[2]);
        tb[2 * P + t] = (T)(en[0] * Tn[0] + en[1] * Tn[1] + en[2] * Tn[2]);
    } else if (G.mode == 0) {
        ub[t] = U[0];
        ub[P + t] = U[1];
        ub[2 * P + t] = 0.0;
        tb[t] = (T)tc[0];
        tb[P + t] = (T)tc[1];
        tb[2 * P + t] = (T)tc[3];
    } else {
        ub[t] = U[0];
        ub[P + t] = U[1];
        ub[2 * P + t] = U[2];
        for (int q6 = 0; q6 < 6; q6++) tb[q6 * P + t] = (T)tc[q6];
    }
}

// plain fields for filter_apply: x = f·active (FP64), plus the mask when a Deformable denominator is needed
__global__ __launch_bounds__(256) void k_ring_prep_fields(GeomArgs G, int dtype, const void *in, int64_t nsl, int withmask,
                                                          double *xin) {
    const int64_t s = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t P = G.npts;
    if (t >= P) return;
    const bool act = G.mask[t] != 0;
    if (s < nsl) {
        const double v = dtype == CGEF_F32 ? (double)((const float *)in)[s * P + t] : ((const double *)in)[s * P + t];
        xin[s * P + t] = act ? v : 0.0;
    } else if (withmask) {
        xin[s * P + t] = act ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(256) void k_ring_normalize(GeomArgs G, int dtype, const double *__restrict__ yout,
                                                        const double *__restrict__ D_row, int64_t nsl, int withmask,
                                                        int strategy, void *out) {
    const int64_t s = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t P = G.npts;
    if (t >= P) return;
    const int64_t r = t / G.n[0];
    const double D = D_row[r];
    const double den = (strategy == CGEF_DEFORMABLE && withmask) ? yout[nsl * P + t] : D;
    const bool ok = G.mask[t] != 0 && den > 1e-15;
    const double v = ok ? yout[s * P + t] / den : 0.0;
    if (dtype == CGEF_F32) ((float *)out)[s * P + t] = (float)v;
    else ((double *)out)[s * P + t] = v;
}

// ------------------------------------------------------------------------------------ launches ----

static size_t fft_shmem(const RingFft &F) { return (size_t)2 * F.n * sizeof(double2); }

static void set_shmem(const void *fn, size_t bytes) {
    if (bytes > 64 * 1024) CGEF_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

void ring_build_what(const RingPlan &R, const Run *runs_d, const int32_t *run_off_d, const double *w64_d, int64_t nruns,
                     hipStream_t s) {
    const size_t sh = fft_shmem(R.fft);
    set_shmem((const void *)k_build_what, sh);
    CGEF_HIP(hipMemsetAsync(R.what_d, 0, (size_t)R.mh * R.args.noff * R.args.nbands * sizeof(double), s));
    hipLaunchKernelGGL(k_build_what, dim3((unsigned)((nruns + 1) / 2)), dim3(256), sh, s, R.fft, R.tw_d, R.args, runs_d,
                       run_off_d, w64_d, nruns, R.what_d);
}

void ring_forward(const RingPlan &R, const double *xin, double2 *XF, int64_t ncol, hipStream_t s) {
    const size_t sh = fft_shmem(R.fft);
    set_shmem((const void *)k_fft_rows, sh);
    hipLaunchKernelGGL(k_fft_rows, dim3((unsigned)((R.args.nrows + 1) / 2), (unsigned)ncol), dim3(256), sh, s, R.fft,
                       R.tw_d, xin, XF, R.args.nrows, ncol);
}

void ring_inverse(const RingPlan &R, const double2 *OUT, double *yout, int64_t ncol, hipStream_t s) {
    const size_t sh = fft_shmem(R.fft);
    set_shmem((const void *)k_ifft_rows, sh);
    hipLaunchKernelGGL(k_ifft_rows, dim3((unsigned)((R.args.nrows + 1) / 2), (unsigned)ncol), dim3(256), sh, s, R.fft,
                       R.tw_d, OUT, yout, R.args.nrows, ncol);
}

void ring_contract(const RingPlan &R, const double2 *XF, double2 *OUT, int64_t ncol, hipStream_t s) {
    constexpr int CC = 10;
    const int span = 256 + R.args.dy_max - R.args.dy_min;
    const size_t sh = (size_t)span * CC * sizeof(double2);
    set_shmem((const void *)k_contract<CC>, sh);
    const int64_t nyb = (R.args.ny + 255) / 256;
    dim3 grid((unsigned)((ncol + CC - 1) / CC), (unsigned)(nyb * R.args.nz), (unsigned)R.mh);
    hipLaunchKernelGGL(k_contract<CC>, grid, dim3(256), sh, s, R.args, XF, OUT, ncol);
}

void ring_prep(const GeomArgs &G, int dtype, const void *u, const void *v, const void *w, const double *cmean, int nf,
               int masked, double *xin, int64_t nbatch, hipStream_t s) {
    dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)nbatch);
    hipLaunchKernelGGL(k_ring_prep, grid, dim3(256), 0, s, G, dtype, u, v, w, cmean, nf, masked, xin);
}

void ring_epilogue(const GeomArgs &G, int dtype, const double *yout, const double *D_row, const double *cmean, int nf,
                   int masked, int strategy, double *ubar, void *tau, int64_t nbatch, hipStream_t s) {
    dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)nbatch);
    if (dtype == CGEF_F32)
        hipLaunchKernelGGL(k_ring_epilogue<float>, grid, dim3(256), 0, s, G, yout, D_row, cmean, nf, masked, strategy, ubar,
                           (float *)tau);
    else
        hipLaunchKernelGGL(k_ring_epilogue<double>, grid, dim3(256), 0, s, G, yout, D_row, cmean, nf, masked, strategy, ubar,
                           (double *)tau);
}

void ring_prep_fields(const GeomArgs &G, int dtype, const void *in, int64_t nsl, int withmask, double *xin, hipStream_t s) {
    dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)(nsl + (withmask ? 1 : 0)));
    hipLaunchKernelGGL(k_ring_prep_fields, grid, dim3(256), 0, s, G, dtype, in, nsl, withmask, xin);
}

void ring_normalize(const GeomArgs &G, int dtype, const double *yout, const double *D_row, int64_t nsl, int withmask,
                    int strategy, void *out, hipStream_t s) {
    dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)nsl);
    hipLaunchKernelGGL(k_ring_normalize, grid, dim3(256), 0, s, G, dtype, yout, D_row, nsl, withmask, strategy, out);
}

}  // namespace cgef
