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
// Kernels: k_fft_rows_staged / k_ifft_rows_staged (two real rows per complex FFT, register-staged
// mixed-radix Stockham with radices up to 16, several row pairs per block for short rows; the LDS
// ping-pong k_fft_rows / k_ifft_rows for other lengths), k_contract (source rows staged in LDS per
// column chunk, Ŵ read coalesced along rows), the prep/epilogue kernels that turn velocities into the
// filtered fields and τ, ū, and the spectral engine's column transforms (k_spec_cols).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include <cfloat>
#include <type_traits>
#include <utility>

#include "cgef_pi_point.hpp"
#include "cgef_ring.hpp"

namespace cgef {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// Twiddle tw[k] = exp(−2πik/N) (conjugated for the inverse); callers guarantee 0 ≤ k < N.
__device__ __forceinline__ double2 twid(const double2 *tw, int k, bool inv) {
    const double2 t = tw[k];
    return inv ? cconj(t) : t;
}

// exp(−2πi·e/R) at compile time: the angle is reduced to |x| ≤ π/4 about the nearest quarter turn
// and cos, sin of x are summed from their Taylor series (terms below 1e-20 at |x| ≤ π/4).
struct Root {
    double c, s;
};
constexpr double kPi = 3.14159265358979323846;
constexpr Root taylor_cs(double x) {
    double c = 1.0, s = x, tc = 1.0, ts = x;
    const double x2 = x * x;
    for (int k = 1; k < 12; k++) {
        tc *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
        ts *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
        c += tc;
        s += ts;
    }
    return Root{c, s};
}
constexpr Root root_of_unity(int e, int R) {
    e %= R;
    const int q = (8 * e + R) / (2 * R);  // nearest quarter turn: round(4e/R)
    const double x = 2.0 * kPi * (double)(4 * e - q * R) / (4.0 * R);
    const Root r = taylor_cs(x);
    // angle = q·π/2 + x; exp(−i·angle) = (cos, −sin)
    double c = r.c, sn = r.s;
    switch (q & 3) {
        case 1: c = -r.s; sn = r.c; break;
        case 2: c = -r.c; sn = -r.s; break;
        case 3: c = r.s; sn = -r.c; break;
        default: break;
    }
    return Root{c, -sn};
}

// v · exp(∓2πi·e/R) with the root folded to constants (∓: forward / inverse)
template <int E, int R>
__device__ __forceinline__ double2 mul_root(double2 v, bool inv) {
    constexpr Root w = root_of_unity(E, R);
    const double ws = inv ? -w.s : w.s;
    return make_double2(v.x * w.c - v.y * ws, v.x * ws + v.y * w.c);
}

template <int A, int B>
__device__ __forceinline__ void dft_comp(double2 *v, bool inv);
template <int R>
__device__ __forceinline__ void dft_direct(double2 *v, const double2 *tw, int n, bool inv);

// In-register DFT of length R on v, roots from the N-point table (R divides N): closed forms for
// 2, 3, 4, 5, two-level products of those for the composite radices, the direct sum otherwise.
template <int R>
__device__ __forceinline__ void dft_small(double2 *v, const double2 *tw, int n, bool inv) {
    if constexpr (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const double2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        const double2 c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        // −i·d forward, +i·d inverse
        const double2 dj = inv ? make_double2(-d.y, d.x) : make_double2(d.y, -d.x);
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(b, dj);
        v[3] = csub(b, dj);
    } else if constexpr (R == 3) {
        // y1,2 = v0 − (v1+v2)/2 ∓ i·(√3/2)(v1 − v2) (forward; signs swap for the inverse)
        const double s3 = 0.86602540378443865;  // sin(2π/3)
        const double2 a = cadd(v[1], v[2]), b = csub(v[1], v[2]);
        const double2 m = make_double2(v[0].x - 0.5 * a.x, v[0].y - 0.5 * a.y);
        const double2 d = make_double2(s3 * b.x, s3 * b.y);
        const double2 dj = inv ? make_double2(-d.y, d.x) : make_double2(d.y, -d.x);  // ∓i·d
        v[0] = cadd(v[0], a);
        v[1] = cadd(m, dj);
        v[2] = csub(m, dj);
    } else if constexpr (R == 5) {
        // pairs a_k = v_k + v_{5−k}, b_k = v_k − v_{5−k}; cosines mix the a's, sines the b's
        const double c1 = 0.30901699437494742, c2 = -0.80901699437494742;  // cos(2π/5), cos(4π/5)
        const double s1 = 0.95105651629515357, s2 = 0.58778525229247313;   // sin(2π/5), sin(4π/5)
        const double2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        const double2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        const double2 m1 = make_double2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
        const double2 m2 = make_double2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
        const double2 n1 = make_double2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
        const double2 n2 = make_double2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
        const double2 j1 = inv ? make_double2(-n1.y, n1.x) : make_double2(n1.y, -n1.x);  // ∓i·n1
        const double2 j2 = inv ? make_double2(-n2.y, n2.x) : make_double2(n2.y, -n2.x);
        v[0] = make_double2(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
        v[1] = cadd(m1, j1);
        v[4] = csub(m1, j1);
        v[2] = cadd(m2, j2);
        v[3] = csub(m2, j2);
    } else if constexpr (R == 6) {
        dft_comp<3, 2>(v, inv);
    } else if constexpr (R == 8) {
        dft_comp<4, 2>(v, inv);
    } else if constexpr (R == 9) {
        dft_comp<3, 3>(v, inv);
    } else if constexpr (R == 10) {
        dft_comp<5, 2>(v, inv);
    } else if constexpr (R == 12) {
        dft_comp<4, 3>(v, inv);
    } else if constexpr (R == 15) {
        dft_comp<5, 3>(v, inv);
    } else if constexpr (R == 16) {
        dft_comp<4, 4>(v, inv);
    } else {
        dft_direct<R>(v, tw, n, inv);
    }
}

// The direct sum for the prime radices 7, 11, 13 (roots from the N-point table; these radices serve
// only the LDS ping-pong FFT).
template <int R>
__device__ __forceinline__ void dft_direct(double2 *v, const double2 *tw, int n, bool inv) {
    double2 o[R];
    const int step = n / R;
#pragma unroll
    for (int a = 0; a < R; a++) {
        double2 s = v[0];
#pragma unroll
        for (int b = 1; b < R; b++) s = cadd(s, cmul(v[b], twid(tw, ((a * b) % R) * step, inv)));
        o[a] = s;
    }
#pragma unroll
    for (int a = 0; a < R; a++) v[a] = o[a];
}

// R = A·B: with q = q2 + B·q1 and k = k1 + A·k2,
// X[k] = Σ_q2 ω_B^(q2·k2) · ω_R^(q2·k1) · DFT_A(v[q2 + B·q1])[k1].
template <int A, int R, int Q2, int... K>
__device__ __forceinline__ void twiddle_col(double2 *y, const double2 *t, bool inv, std::integer_sequence<int, K...>) {
    y[Q2 * A] = t[0];
    ((y[Q2 * A + K + 1] = mul_root<Q2 * (K + 1), R>(t[K + 1], inv)), ...);
}

template <int A, int B, int Q2 = 0>
__device__ __forceinline__ void comp_cols(double2 *v, double2 *y, bool inv) {
    if constexpr (Q2 < B) {
        double2 t[A];
#pragma unroll
        for (int q1 = 0; q1 < A; q1++) t[q1] = v[Q2 + B * q1];
        dft_small<A>(t, nullptr, 0, inv);
        twiddle_col<A, A * B, Q2>(y, t, inv, std::make_integer_sequence<int, A - 1>{});
        comp_cols<A, B, Q2 + 1>(v, y, inv);
    }
}

template <int A, int B>
__device__ __forceinline__ void dft_comp(double2 *v, bool inv) {
    constexpr int R = A * B;
    double2 y[R];  // y[q2·A + k1]
    comp_cols<A, B>(v, y, inv);
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) {
        double2 t[B];
#pragma unroll
        for (int q2 = 0; q2 < B; q2++) t[q2] = y[q2 * A + k1];
        dft_small<B>(t, nullptr, 0, inv);
#pragma unroll
        for (int k2 = 0; k2 < B; k2++) v[k1 + A * k2] = t[k2];
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
            v[r] = r == 0 ? a : cmul(a, twid(tw, r * k * step, inv));
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

// Work item of a row-pair kernel. Workgroups are dispatched round-robin over the 8 XCDs, each with its
// own L2; giving every XCD a contiguous range of (column, row pair) items makes neighbouring row pairs
// — whose spectra share cache lines of the [m][col][row] layout — write through the same L2.
__device__ __forceinline__ bool xcd_item(int64_t total, int64_t &item) {
    const int64_t chunk = (total + 7) / 8;
    item = (int64_t)(blockIdx.x % 8) * chunk + blockIdx.x / 8;
    return item < total;
}

// Offset of mode m of row `row` of column `col` in a spectra set [m][col][row].
__device__ __forceinline__ int64_t spec_at(int64_t m, int64_t col, int64_t row, int64_t nrows, int64_t ncol) {
    return (m * ncol + col) * nrows + row;
}

// The forward half-spectra of row pairs from the complex transforms z_sq of x_(r0+2sq) + i·x_(r0+2sq+1)
// (element e of z_sq at z[at(sq, e)]): output e is mode e/(2·nsq) of row r0 + e%(2·nsq), so that
// consecutive lanes write the block's consecutive rows of one mode.
template <typename At>
__device__ __forceinline__ void write_pair_spectra(const double2 *z, At at, int n, int nsq, double2 *XF, int64_t col,
                                                   int64_t r0, int64_t nrows, int64_t ncol) {
    const int mh = n / 2 + 1, rpb = 2 * nsq;
    for (int e = threadIdx.x; e < rpb * mh; e += blockDim.x) {
        const int m = e / rpb, ro = e - m * rpb, sq = ro >> 1;
        if (r0 + ro >= nrows) continue;
        const double2 a = z[at(sq, m)], b = cconj(z[at(sq, m == 0 ? 0 : n - m)]);
        // X = (a + b)/2 for row r0 + 2sq, Y = (a − b)/(2i) for row r0 + 2sq + 1
        XF[spec_at(m, col, r0 + ro, nrows, ncol)] = (ro & 1) == 0 ? make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y))
                                                                  : make_double2(0.5 * (a.y - b.y), -0.5 * (a.x - b.x));
    }
}

// Forward: real rows (r0, r0+1) of column `col` (xin[col][row][i]) → XF[m][col][row], m ≤ N/2.
__global__ __launch_bounds__(256) void k_fft_rows(RingFft F, const double2 *__restrict__ tw, const double *__restrict__ xin,
                                                  double2 *__restrict__ XF, int64_t nrows, int64_t ncol) {
    extern __shared__ double2 sm[];
    const int n = F.n;
    const int64_t npair = (nrows + 1) / 2;
    int64_t item;
    if (!xcd_item(npair * ncol, item)) return;
    const int64_t col = item / npair;
    const int64_t r0 = 2 * (item % npair), r1 = r0 + 1;
    const bool has1 = r1 < nrows;
    const double *x0 = xin + (col * nrows + r0) * n;
    const double *x1 = has1 ? x0 + n : nullptr;
    for (int i = threadIdx.x; i < n; i += blockDim.x) sm[i] = make_double2(x0[i], has1 ? x1[i] : 0.0);
    const double2 *Z = fft_lds(sm, sm + n, F, tw, false, 1);
    write_pair_spectra(Z, [](int, int e) { return e; }, n, 1, XF, col, r0, nrows, ncol);
}

// Inverse: OUT[m][col][row] (Hermitian halves) → yout[col][row][i], normalized by 1/N. With a mode cut
// (mcut non-null) row r reads only its modes m < mcut[r]; the others are zero.
__global__ __launch_bounds__(256) void k_ifft_rows(RingFft F, const double2 *__restrict__ tw, const double2 *__restrict__ OUT,
                                                   double *__restrict__ yout, int64_t nrows, int64_t ncol,
                                                   const int32_t *__restrict__ mcut) {
    extern __shared__ double2 sm[];
    const int n = F.n;
    const int64_t npair = (nrows + 1) / 2;
    int64_t item;
    if (!xcd_item(npair * ncol, item)) return;
    const int64_t col = item / npair;
    const int64_t r0 = 2 * (item % npair), r1 = r0 + 1;
    const bool has1 = r1 < nrows;
    const int mh = n / 2 + 1;
    const int mc0 = mcut ? mcut[r0] : mh, mc1 = mcut && has1 ? mcut[r1] : mh;
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        const bool lo = m < mh;
        const int mm = lo ? m : n - m;
        double2 X = mm < mc0 ? OUT[spec_at(mm, col, r0, nrows, ncol)] : make_double2(0.0, 0.0);
        double2 Y = has1 && mm < mc1 ? OUT[spec_at(mm, col, r1, nrows, ncol)] : make_double2(0.0, 0.0);
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

// ---- register-staged FFT: one sequence per 256-thread block in LDS; each stage reads its butterflies
// into registers, waits, then writes them back in Stockham order. The stages use radices up to 16
// (composite radices evaluated in registers), so a 1440-point transform takes 3 passes through LDS
// instead of 6. The first stage's writes step by R elements across lanes: an odd first radix keeps
// them on distinct banks; an even one (power-of-two lengths) gets a padded layout, element e at
// e + e/16 (F.lpsh = 4; 31 leaves the layout unpadded).

__device__ __forceinline__ int lp(int e, int sh) { return e + (e >> sh); }

// F.nsq sequences of F.n points at sm + sq·F.sstr are transformed together by the block's BLK threads
// (short rows: each stage then has a wave's worth of butterflies or more).
template <int R, int MB, int BLK = 256>
__device__ __forceinline__ void stage_staged(double2 *sm, int n, int Ns, int sh, int nsq, int sstr,
                                             const double2 *__restrict__ tws, const double2 *__restrict__ tw, bool inv) {
    // the stage parameters pass through an empty asm so that this radix's addresses are formed here:
    // hoisted out of the stage loop, every radix of the switch would keep its own set of registers.
    n = __builtin_amdgcn_readfirstlane(n);
    sh = __builtin_amdgcn_readfirstlane(sh);
    nsq = __builtin_amdgcn_readfirstlane(nsq);
    asm volatile("" : "+s"(n), "+s"(sh), "+s"(nsq));
    const int nb = n / R, tot = nb * nsq;
    double2 v[MB][R];
    int dst[MB], base[MB];
#pragma unroll
    for (int q = 0; q < MB; q++) {
        const int p = (int)threadIdx.x + q * BLK;
        if (p < tot) {
            const int sq = nsq == 1 ? 0 : p / nb;
            const int j = p - sq * nb;
            const int k = j % Ns;
            const double2 *t = tws + k * (R - 1);
            const double2 *x = sm + sq * sstr;
            v[q][0] = x[lp(j, sh)];
            if (Ns == 1) {  // first stage: every twiddle is 1
#pragma unroll
                for (int r = 1; r < R; r++) v[q][r] = x[lp(j + r * nb, sh)];
            } else {
                // one table load per butterfly: ω^k, its powers ω^(rk) = ω^(⌈r/2⌉k)·ω^(⌊r/2⌋k) (at most
                // four products deep for r ≤ 15), in place of R − 1 loads through the texture path
                double2 wp[R];
                wp[1] = inv ? cconj(t[0]) : t[0];
#pragma unroll
                for (int r = 2; r < R; r++) wp[r] = cmul(wp[(r + 1) / 2], wp[r / 2]);
#pragma unroll
                for (int r = 1; r < R; r++) v[q][r] = cmul(x[lp(j + r * nb, sh)], wp[r]);
            }
            dft_small<R>(v[q], tw, n, inv);
            dst[q] = (j - k) * R + k;
            base[q] = sq * sstr;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MB; q++) {
        const int p = (int)threadIdx.x + q * BLK;
        if (p < tot) {
#pragma unroll
            for (int r = 0; r < R; r++) sm[base[q] + lp(dst[q] + r * Ns, sh)] = v[q][r];
        }
    }
    __syncthreads();
}

// F.nsq sequences of length F.n at sm + sq·F.sstr (padded layout, transformed in place) with the
// staged factorization F.srad (every stage's butterflies fit the per-thread budget the host checked).
__device__ void fft_staged(double2 *sm, const RingFft &F, const double2 *tw, const double2 *tws, bool inv) {
    __syncthreads();
    int Ns = 1;
    for (int st = 0; st < F.snst; st++) {
        const int R = F.srad[st];
        const int mb = (F.n / R * F.nsq + 255) / 256;
        const double2 *t = tws + F.twoff[st];
        switch (R * 16 + mb) {
#define CGEF_ST(RR, MM) \
    case RR * 16 + MM: stage_staged<RR, MM>(sm, F.n, Ns, F.lpsh, F.nsq, F.sstr, t, tw, inv); break;
            CGEF_ST(2, 1) CGEF_ST(2, 2) CGEF_ST(2, 3) CGEF_ST(2, 4) CGEF_ST(2, 5) CGEF_ST(2, 6)
            CGEF_ST(3, 1) CGEF_ST(3, 2) CGEF_ST(3, 3) CGEF_ST(3, 4)
            CGEF_ST(4, 1) CGEF_ST(4, 2) CGEF_ST(4, 3)
            CGEF_ST(5, 1) CGEF_ST(5, 2)
            CGEF_ST(6, 1) CGEF_ST(6, 2)
            CGEF_ST(8, 1) CGEF_ST(8, 2)
            CGEF_ST(9, 1) CGEF_ST(10, 1) CGEF_ST(12, 1) CGEF_ST(15, 1) CGEF_ST(16, 1)
#undef CGEF_ST
            default: break;
        }
        Ns *= R;
    }
}

// The same with the factorization fixed at compile time (RingFft::fixed; every stage one butterfly per
// lane over the block's sequences, BLK lanes): only these stages' registers are allocated, where the switch
// above holds those of its largest radix (96 VGPRs against 67 for 12·12·10 and 30 for 4·4·4·4).
template <int... Rs>
struct Radices {};
template <int FIX>
struct FixedRadices;
template <>
struct FixedRadices<RING_FIXED_1440> {
    using type = Radices<12, 12, 10>;
};
template <>
struct FixedRadices<RING_FIXED_256> {
    using type = Radices<4, 4, 4, 4>;
};

template <int BLK, int Ns, int R, int... Rest>
__device__ __forceinline__ void fixed_stages(double2 *sm, const RingFft &F, const double2 *tw, const double2 *tws,
                                             bool inv, int st) {
    stage_staged<R, 1, BLK>(sm, F.n, Ns, F.lpsh, F.nsq, F.sstr, tws + F.twoff[st], tw, inv);
    if constexpr (sizeof...(Rest) > 0) fixed_stages<BLK, Ns * R, Rest...>(sm, F, tw, tws, inv, st + 1);
}

template <int BLK, int... Rs>
__device__ __forceinline__ void fft_fixed(Radices<Rs...>, double2 *sm, const RingFft &F, const double2 *tw,
                                          const double2 *tws, bool inv) {
    __syncthreads();
    fixed_stages<BLK, 1, Rs...>(sm, F, tw, tws, inv, 0);
}

template <int FIX, int BLK = 256>
__device__ __forceinline__ void fft_rows_any(double2 *sm, const RingFft &F, const double2 *tw, const double2 *tws, bool inv) {
    if constexpr (FIX == 0) fft_staged(sm, F, tw, tws, inv);
    else fft_fixed<BLK>(typename FixedRadices<FIX>::type{}, sm, F, tw, tws, inv);
}

// Forward, 2·F.nsq real rows per block: rows r0.. of column `col` → XF[m][col][row], m ≤ N/2.
// LPT·256 ≥ nsq·N: every load of the rows is issued before the first is waited for.
template <int LPT, int FIX>
__global__ __launch_bounds__(256) void k_fft_rows_staged(RingFft F, const double2 *__restrict__ tw,
                                                         const double2 *__restrict__ tws, const double *__restrict__ xin,
                                                         double2 *__restrict__ XF, int64_t nrows, int64_t ncol) {
    extern __shared__ double2 sm[];
    const int n = F.n, nsq = F.nsq, sh = F.lpsh, sstr = F.sstr;
    const int64_t rpb = 2 * nsq;
    const int64_t ngrp = (nrows + rpb - 1) / rpb;
    int64_t item;
    if (!xcd_item(ngrp * ncol, item)) return;
    const int64_t col = item / ngrp, r0 = (item % ngrp) * rpb;
    const double *x = xin + (col * nrows + r0) * n;
    double a_[LPT], b_[LPT];
#pragma unroll
    for (int q = 0; q < LPT; q++) {
        const int e = (int)threadIdx.x + q * 256;
        const int sq = e / n, i = e - sq * n;
        const bool in = e < nsq * n;
        a_[q] = in && r0 + 2 * sq < nrows ? x[(int64_t)(2 * sq) * n + i] : 0.0;
        b_[q] = in && r0 + 2 * sq + 1 < nrows ? x[(int64_t)(2 * sq + 1) * n + i] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < LPT; q++) {
        const int e = (int)threadIdx.x + q * 256;
        const int sq = e / n, i = e - sq * n;
        if (e < nsq * n) sm[sq * sstr + lp(i, sh)] = make_double2(a_[q], b_[q]);
    }
    fft_rows_any<FIX>(sm, F, tw, tws, false);
    write_pair_spectra(sm, [sh, sstr](int sq, int e) { return sq * sstr + lp(e, sh); }, n, nsq, XF, col, r0, nrows, ncol);
}

// Inverse, 2·F.nsq rows per block: OUT[m][col][row] (Hermitian halves) → yout[col][row][i] / N. Each
// stored value is read once, lane per (mode, row) so that the block's rows of one mode are adjacent;
// each lane pair exchanges its X, Y and writes X + iY at m and the conjugate mirror at N − m.
// LPT·256 ≥ 2·nsq·(N/2 + 1). With a mode cut (mcut non-null) row r reads only its modes m < mcut[r].
template <int LPT, int FIX>
__global__ __launch_bounds__(256) void k_ifft_rows_staged(RingFft F, const double2 *__restrict__ tw,
                                                          const double2 *__restrict__ tws, const double2 *__restrict__ OUT,
                                                          double *__restrict__ yout, int64_t nrows, int64_t ncol,
                                                          const int32_t *__restrict__ mcut) {
    extern __shared__ double2 sm[];
    const int n = F.n, nsq = F.nsq, sh = F.lpsh, sstr = F.sstr;
    const int rpb = 2 * nsq;
    const int64_t ngrp = (nrows + rpb - 1) / rpb;
    int64_t item;
    if (!xcd_item(ngrp * ncol, item)) return;
    const int64_t col = item / ngrp, r0 = (item % ngrp) * rpb;
    const int mh = n / 2 + 1;
    double2 v_[LPT];  // all gathers in flight together
#pragma unroll
    for (int q = 0; q < LPT; q++) {
        const int e = (int)threadIdx.x + q * 256;
        const int m = e / rpb, ro = e - m * rpb;
        const bool in = m < mh && r0 + ro < nrows && (!mcut || m < mcut[r0 + ro]);
        v_[q] = in ? OUT[spec_at(m, col, r0 + ro, nrows, ncol)] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < LPT; q++) {
        const int e = (int)threadIdx.x + q * 256;
        const int m = e / rpb, ro = e - m * rpb;
        // lanes with ro = 2sq, 2sq+1 hold X (row r0+2sq) and Y (row r0+2sq+1) of mode m
        const double2 o = make_double2(__shfl_xor(v_[q].x, 1), __shfl_xor(v_[q].y, 1));
        const bool even = (ro & 1) == 0;
        const double2 X = even ? v_[q] : o, Y = even ? o : v_[q];
        double2 *z = sm + (ro >> 1) * sstr;
        if (m < mh) {
            if (even) z[lp(m, sh)] = make_double2(X.x - Y.y, X.y + Y.x);  // X + iY
            else if (m > 0 && n - m >= mh)
                z[lp(n - m, sh)] = make_double2(X.x + Y.y, Y.x - X.y);  // conj(X) + i·conj(Y)
        }
    }
    fft_rows_any<FIX>(sm, F, tw, tws, true);
    const double sc = 1.0 / n;
    double *y = yout + (col * nrows + r0) * n;
    for (int e = threadIdx.x; e < nsq * n; e += 256) {
        const int sq = e / n, i = e - sq * n;
        const double2 z = sm[sq * sstr + lp(i, sh)];
        if (r0 + 2 * sq < nrows) y[(int64_t)(2 * sq) * n + i] = z.x * sc;
        if (r0 + 2 * sq + 1 < nrows) y[(int64_t)(2 * sq + 1) * n + i] = z.y * sc;
    }
}

// Banded contraction per Fourier mode m. Block: 256 target rows of one plane (j0..j0+255, level k)
// × a chunk of CC columns, one target row per lane. The source rows of each dz plane are staged in
// LDS (zero outside a non-periodic axis) and every staged value feeds CC columns' worth of FMAs per
// weight; Ŵ is read along rows (coalesced on the sphere, broadcast on translation-invariant grids),
// in batches of PB whose next batch is in flight while the current one is used — the first batch of
// a plane is issued before its staging, so the weight latency hides behind the staging loads. PB and
// the occupancy bound MINW (waves per SIMD the register allocation must allow) trade prefetch depth
// against resident blocks; ring_contract picks them per launch.

template <int CC, int PB, int MINW>
__global__ __launch_bounds__(256, MINW) void k_contract(RingArgs A, const double2 *__restrict__ XF, double2 *__restrict__ OUT,
                                                        int64_t ncol, int out_rowmajor, const int32_t *__restrict__ items,
                                                        int64_t nitems, const int32_t *__restrict__ mcut) {
    extern __shared__ double2 st[];
    const int64_t Ny = A.ny, Nz = A.nz;
    const int64_t nyb = (Ny + 255) / 256;
    const int64_t ncc = (ncol + CC - 1) / CC;
    // items (m, row block, column chunk), column chunk fastest, in contiguous ranges per XCD: the
    // chunks of one (m, row block) read the same Ŵ slab and so should share an L2. With a mode cut the
    // (m, row block) pairs come from its work list (ascending m), the others are not launched.
    int64_t item;
    if (!xcd_item((items ? nitems : (int64_t)A.mh * nyb * Nz) * ncc, item)) return;
    const int64_t c0 = (item % ncc) * CC;
    const int64_t pr = items ? (int64_t)items[item / ncc] : item / ncc;
    const int64_t rbk = pr % (nyb * Nz);
    const int64_t m = pr / (nyb * Nz);
    const int64_t k = rbk / nyb, j0 = (rbk % nyb) * 256;
    const int tid = threadIdx.x;
    const int64_t j = j0 + tid;
    const bool valid = j < Ny;
    // rows whose cut lies at or below m take part in the staging only: their mode m is never read
    const bool act = valid && (!mcut || m < mcut[j + Ny * k]);
    const int band = valid ? A.band_of_row[j + Ny * k] : 0;
    const int span = 256 + A.dy_max - A.dy_min;
    const int nvc = (int)((ncol - c0) < CC ? (ncol - c0) : CC);  // valid columns in this chunk
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
        const int o0 = A.dz_off0[g], ng = A.dz_off0[g + 1] - o0;
        const int e0 = A.off_dy[o0] - A.dy_min;  // staged index of offset 0 for lane 0
        const double *wrow = A.what + (mrow + o0) * A.nbands + band;
        double wn[PB];
#pragma unroll
        for (int u = 0; u < PB; u++) wn[u] = (act && u < ng) ? wrow[(int64_t)u * A.nbands] : 0.0;
        __syncthreads();
        // staging: a lane takes source-row offsets s = tid, tid + 256, … and loads the CC columns of
        // each together (CC independent loads in flight, no division by the span)
        for (int s = tid; s < span; s += 256) {
            int64_t jj = j0 + A.dy_min + s;
            bool in = true;
            if (jj < 0 || jj >= Ny) {
                if (A.py) jj = ((jj % Ny) + Ny) % Ny;
                else in = false;
            }
            const double2 *xs = XF + (m * ncol + c0) * A.nrows + jj + Ny * kk;
            double2 v[CC];
#pragma unroll
            for (int c = 0; c < CC; c++)
                v[c] = (in && c < nvc) ? xs[(int64_t)c * A.nrows] : make_double2(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < CC; c++) st[c * span + s] = v[c];
        }
        __syncthreads();
        if (!act) continue;
        const double2 *src = st + tid + e0;
        for (int p0 = 0; p0 < ng; p0 += PB) {
            double w[PB];
#pragma unroll
            for (int u = 0; u < PB; u++) w[u] = wn[u];
            if (p0 + PB < ng) {
#pragma unroll
                for (int u = 0; u < PB; u++) wn[u] = (p0 + PB + u < ng) ? wrow[(int64_t)(p0 + PB + u) * A.nbands] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < PB; u++) {
                if (p0 + u >= ng) break;
#pragma unroll
                for (int c = 0; c < CC; c++) {
                    const double2 v = src[c * span + p0 + u];
                    acc[c].x = fma(w[u], v.x, acc[c].x);
                    acc[c].y = fma(w[u], v.y, acc[c].y);
                }
            }
        }
    }
    if (out_rowmajor) {
        // OUT[m][row][col]: the block's 256 rows × CC columns pass through LDS and are stored row by
        // row, CC adjacent columns per row, so a wave's stores cover about 13 rows instead of 64
        __syncthreads();   // every lane has finished reading the staged spectra
#pragma unroll
        for (int c = 0; c < CC; c++) st[tid * CC + c] = acc[c];
        __syncthreads();
        const int nrv = (int)((Ny - j0) < 256 ? (Ny - j0) : 256);   // valid rows of this block
        for (int e = tid; e < nrv * CC; e += 256) {
            const int rr = e / CC, c = e - rr * CC;
            const int64_t row = j0 + rr + Ny * k;
            if (c < nvc && (!mcut || m < mcut[row])) OUT[(m * A.nrows + row) * ncol + c0 + c] = st[e];
        }
        return;
    }
    if (!act) return;
    // fully unrolled with a predicate: a loop bounded by nvc would index acc dynamically and move the
    // accumulators to scratch memory
    const int64_t r = j + Ny * k;
#pragma unroll
    for (int c = 0; c < CC; c++)
        if (c < nvc) OUT[(m * ncol + c0 + c) * A.nrows + r] = acc[c];
}

// ---- the contraction on the matrix cores (2-D grids, wide footprints) ------------------------------
//
// Per Fourier mode m the contraction is a banded real matrix times the spectra of the plane columns:
//     OUT[c][r] = Σ_s A[r][s] · X[c][s],   A[r][s] = Ŵ[m][s − r − dy_min][band(r)] inside the band.
// With the real and imaginary parts of the columns as separate real columns it maps onto
// v_mfma_f64_16x16x4f64: D(16 target rows × 16 real columns) += A(16 × 4 source rows) · B(4 × 16).
// A workgroup (4 waves) takes 64 consecutive target rows of one mode (one 16-row tile per wave) and
// NT·16 real columns; the source rows its tiles reach are staged through LDS KC rows at a time
// ([row][real column], shared by the four waves), each wave running the 4-row steps of its tile that
// fall in the staged rows: one A fragment (one coalesced 8-byte load per lane from the lane-ordered
// weight table) serves NT MFMAs, each with a B fragment read from LDS. Operand layouts (probed on
// gfx950): lane l holds A[l%16][l/16] and B[l/16][l%16]; the accumulator's value v is D[l/16 + 4v][l%16].
// The staging is double-buffered: the next pass's spectra and weights are requested into registers
// before this pass's MFMAs are issued and reach the other LDS buffer after them, so one barrier per
// pass separates the buffers and the loads' latency runs under the matrix unit's work.
// Each staged spectrum value feeds up to 16 products in the matrix unit, where the VALU contraction
// reads it from LDS once per FMA pair; the tile's band edges add (15 + offsets)/offsets of zero work.
typedef double mfma_f64x4 __attribute__((ext_vector_type(4)));
constexpr int MFMA_KC = 32;                   // source rows staged per pass (8 four-row steps)

template <int NT>
__global__ __launch_bounds__(256, 2) void k_contract_mfma(RingArgs A, const double *__restrict__ wt, int nsteps, int wtiles,
                                                       const double2 *__restrict__ XF, double2 *__restrict__ OUT,
                                                       int64_t ncol, int out_rowmajor, const int32_t *__restrict__ items,
                                                       int64_t nitems) {
    constexpr int LDW = NT * 16 + 2;          // row stride in doubles (≡ 4 banks mod 32: 16-byte stores spread)
    constexpr int BUF = MFMA_KC * LDW;        // one staging buffer
    static_assert(MFMA_KC == 32, "staging: a thread owns one row (tid % 32) and NT columns (tid / 32 + 8i)");
    extern __shared__ double lds[];
    const int64_t Ny = A.ny;
    const int64_t nrb = (Ny + 63) / 64;
    const int64_t ncc = (2 * ncol + NT * 16 - 1) / (NT * 16);
    // items (m, row block, column chunk), column chunk fastest, contiguous ranges per XCD: the row
    // blocks of one mode re-read overlapping source rows through the same L2. With a mode cut the
    // (m, row block) pairs come from its work list.
    int64_t item;
    if (!xcd_item((items ? nitems : (int64_t)A.mh * nrb) * ncc, item)) return;
    const int64_t cc = item % ncc;
    const int64_t pr = items ? (int64_t)items[item / ncc] : item / ncc;
    const int64_t rb = pr % nrb;
    const int64_t m = pr / nrb;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t R0 = rb * 64;
    const int ntile = (int)((Ny - R0) < 64 ? (Ny - R0 + 15) / 16 : 4);   // tiles in this block
    const int64_t c0 = cc * (NT * 8);                                    // first complex column
    const int nvc = (int)((ncol - c0) < NT * 8 ? (ncol - c0) : NT * 8);
    const bool active = wave < ntile;
    const int64_t tile = R0 / 16 + wave;
    const double *wp = wt + ((m * wtiles + (wtiles == 1 ? 0 : tile)) * nsteps) * 64 + lane;
    const int krows = 16 * (ntile - 1) + 4 * nsteps;                     // block source rows used
    const int64_t S0 = R0 + A.dy_min;                                    // block source row 0
    // on a non-periodic axis the source rows outside 0..Ny−1 contribute nothing: skip the 4-row steps
    // (and staged passes) that lie wholly outside (the rows beyond a pole at wide footprints)
    const int brlo = A.py ? 0 : (int)(S0 < 0 ? (-S0 < krows ? -S0 : krows) : 0);
    const int brhi = A.py ? krows : (int)(Ny - S0 < krows ? (Ny - S0 > 0 ? Ny - S0 : 0) : krows);
    const int brlo4 = brlo & ~3;
    const double2 *xs = XF + (m * ncol + c0) * A.nrows;
    const int srow = tid & 31, scol = tid >> 5;
    // this wave's 4-row steps within the staged rows [base, base + KC) of the block's source rows
    auto steps_of = [&](int base, int &lo, int &hi) {
        lo = base > 16 * wave ? base : 16 * wave;
        lo = lo > brlo4 ? lo : brlo4;
        hi = base + MFMA_KC;
        hi = hi < 16 * wave + 4 * nsteps ? hi : 16 * wave + 4 * nsteps;
        hi = hi < brhi ? hi : brhi;
    };
    // the pass's requests: row S0 + base + tid%32 of columns tid/32 + 8i (32 consecutive rows of a column
    // per half wave: 512 contiguous bytes; zero outside a non-periodic axis and beyond the valid
    // columns), and this wave's A fragments
    auto request = [&](int base, double2 (&v)[NT], double (&a)[MFMA_KC / 4]) {
        int64_t s = S0 + base + srow;
        bool in = true;
        if (s < 0 || s >= Ny) {
            if (A.py) s = ((s % Ny) + Ny) % Ny;
            else in = false;
        }
        const double2 *xr = xs + s;
#pragma unroll
        for (int i = 0; i < NT; i++) {
            const int col = scol + 8 * i;
            v[i] = (in && col < nvc) ? xr[(int64_t)col * A.nrows] : make_double2(0.0, 0.0);
        }
        int lo, hi;
        steps_of(base, lo, hi);
#pragma unroll
        for (int q = 0; q < MFMA_KC / 4; q++) {
            const int br = lo + 4 * q;
            a[q] = (active && br < hi) ? wp[(int64_t)((br - 16 * wave) >> 2) * 64] : 0.0;
        }
    };
    auto deposit = [&](double *buf, const double2 (&v)[NT]) {
#pragma unroll
        for (int i = 0; i < NT; i++) *reinterpret_cast<double2 *>(buf + srow * LDW + 2 * (scol + 8 * i)) = v[i];
    };
    mfma_f64x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = mfma_f64x4{0.0, 0.0, 0.0, 0.0};
    const int pfirst = brlo / MFMA_KC * MFMA_KC;
    double2 v[NT];
    double a[MFMA_KC / 4], an[MFMA_KC / 4];
    if (pfirst < brhi) {
        request(pfirst, v, a);
        deposit(lds, v);
    }
    __syncthreads();
    int cur = 0;
    for (int base = pfirst; base < brhi; base += MFMA_KC) {
        const bool more = base + MFMA_KC < brhi;
        if (more) request(base + MFMA_KC, v, an);
        if (active) {
            int lo, hi;
            steps_of(base, lo, hi);
            const double *sb = lds + cur * BUF;
#pragma unroll
            for (int q = 0; q < MFMA_KC / 4; q++) {
                const int br = lo + 4 * q;
                if (br >= hi) break;
                const double *brow = sb + (br - base + (lane >> 4)) * LDW + (lane & 15);
#pragma unroll
                for (int t = 0; t < NT; t++)
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], brow[t * 16], acc[t], 0, 0, 0);
            }
        }
        if (more) deposit(lds + (cur ^ 1) * BUF, v);   // the other buffer: its last reads ended with the previous pass
        __syncthreads();
        cur ^= 1;
#pragma unroll
        for (int q = 0; q < MFMA_KC / 4; q++) a[q] = an[q];
    }
    if (!active) return;
    // lane holds D[lane/16 + 4v][lane%16] of each column tile: target row, real column (re/im of a
    // complex column); in the [m][row][col] layout a row's 16 real columns are 128 contiguous bytes
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int j = t * 16 + (lane & 15);
        const int c = j >> 1;
        if (c >= nvc) continue;
#pragma unroll
        for (int v4 = 0; v4 < 4; v4++) {
            const int64_t r = tile * 16 + (lane >> 4) + 4 * v4;
            if (r >= Ny) continue;
            double *o = out_rowmajor ? reinterpret_cast<double *>(OUT + (m * A.nrows + r) * ncol + c0 + c)
                                     : reinterpret_cast<double *>(OUT + (m * ncol + c0 + c) * A.nrows + r);
            o[j & 1] = acc[t][v4];
        }
    }
}

// The A-operand table of k_contract_mfma from Ŵ: entry (m, tile, step, lane) = Ŵ[m][o][band(r)] with
// r = 16·tile + lane%16 and o = 4·step + lane/16 − lane%16, zero where o is outside 0..noff−1 or r ≥ Ny.
__global__ __launch_bounds__(256) void k_build_wtiles(RingArgs A, int nsteps, int wtiles, double *__restrict__ wt) {
    const int64_t total = (int64_t)A.mh * wtiles * nsteps * 64;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int lane = (int)(e % 64);
        const int64_t step = (e / 64) % nsteps;
        const int64_t tile = (e / (64 * nsteps)) % wtiles;
        const int64_t m = e / (64 * nsteps * wtiles);
        const int i = lane & 15, k = lane >> 4;
        const int64_t r = tile * 16 + i;
        const int64_t o = 4 * step + k - i;
        double v = 0.0;
        if (o >= 0 && o < A.noff && r < A.ny) v = A.what[(m * A.noff + o) * A.nbands + A.band_of_row[r]];
        wt[e] = v;
    }
}

// The mode cut (RING_MODE_TOL, RING_MODE_SLOPE_TOL), first step: S[m][band] = Σ_offsets |Ŵ[m][offset][band]|,
// lanes over consecutive bands (coalesced along the [m][offset][band] table).
__global__ __launch_bounds__(256) void k_mode_sums(RingArgs A, double *__restrict__ S) {
    const int64_t total = (int64_t)A.mh * A.nbands;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t m = e / A.nbands, b = e - m * A.nbands;
        const double *w = A.what + m * A.noff * A.nbands + b;
        double acc = 0.0;
        for (int o = 0; o < A.noff; o++) acc += fabs(w[(int64_t)o * A.nbands]);
        S[e] = acc;
    }
}

// Second step, a lane per band: the in-range weight sum S0 = Σ_offsets Ŵ[0][offset][band] and, from the
// highest mode down, the value tail Σ c_m·S[m] and the slope tail Σ c_m·S[m]·max(1, 4ℓ·sin(πm/N)/Δx)
// (c_m = 1 for modes 0 and N/2 of the full spectrum, else 2); mband = the smallest M at which both
// tails are within their bounds (every mode when S0 is not positive). dxb: the band's smallest
// x spacing.
__global__ __launch_bounds__(256) void k_mode_cut(RingArgs A, int n, double tol, double slope_tol, double ell,
                                                  const double *__restrict__ dxb, const double *__restrict__ S,
                                                  int32_t *__restrict__ mband) {
    const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (b >= A.nbands) return;
    double s0 = 0.0;
    for (int o = 0; o < A.noff; o++) s0 += A.what[(int64_t)o * A.nbands + b];
    int M = (int)A.mh;
    if (s0 > 0.0) {
        const double g = 4.0 * ell / dxb[b];
        double tv = 0.0, ts = 0.0;
        int mv = -1, ms = -1;
        for (int64_t m = A.mh - 1; m >= 0 && (mv < 0 || ms < 0); m--) {
            const double c = (m == 0 || 2 * m == n ? 1.0 : 2.0) * S[m * A.nbands + b];
            tv += c;
            ts += c * fmax(1.0, g * sin(M_PI * (double)m / (double)n));
            if (mv < 0 && tv > tol * s0) mv = (int)m + 1;
            if (ms < 0 && ts > slope_tol * s0) ms = (int)m + 1;
        }
        M = mv > ms ? mv : ms;
    }
    mband[b] = M;
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

// Field order per slice: q (NC), products (NP: xx xy [xz] yy [yz zz]), mask (if masked).
// VP consecutive points per lane (2 when the plane length is even: 16-byte stores of both points).
template <int VP>
__global__ __launch_bounds__(256) void k_ring_prep(GeomArgs G, int dtype, const void *u, const void *v, const void *w,
                                                   int nf, int withmask, double *xin) {
    const int64_t b = blockIdx.y;
    const int64_t t0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * VP;
    const int64_t P = G.npts;
    if (t0 >= P) return;  // with VP = 2, P is even, so t0 + 1 < P as well
    const int64_t Nx = G.n[0], Ny = G.n[1];
    double q[VP][3];
    double act[VP];
#pragma unroll
    for (int p = 0; p < VP; p++) {
        const int64_t t = t0 + p;
        const int64_t i = t % Nx, j = (t / Nx) % Ny;
        act[p] = G.mask[t] != 0 ? 1.0 : 0.0;
        q[p][0] = q[p][1] = q[p][2] = 0.0;
        if (act[p] != 0.0) frame_vel(G, dtype, u, v, w, b * P, t, i, j, q[p]);
    }
    double *o = xin + b * nf * P + t0;
    const int npl = G.mode == 0 ? 5 : 9;
#pragma unroll
    for (int f = 0; f < 9; f++) {
        if (f < npl) {
            if (VP == 2)
                *(double2 *)(o + f * P) = make_double2(pi_plane_value(G.mode, f, q[0]), pi_plane_value(G.mode, f, q[VP - 1]));
            else
                o[f * P] = pi_plane_value(G.mode, f, q[0]);
        }
    }
    if (withmask) {
        if (VP == 2) *(double2 *)(o + npl * P) = make_double2(act[0], act[VP - 1]);
        else o[npl * P] = act[0];
    }
}

// τ = F[q_a q_b]/den − ū_a ū_b and ū = F[q]/den from the filtered planes, then the local frame on the
// sphere (Diagnostics.jl:596-650). FP64 throughout, so no mean shift is needed against cancellation.
template <typename T>
__global__ __launch_bounds__(256) void k_ring_epilogue(GeomArgs G, const double *__restrict__ yout, const double *__restrict__ D_row,
                                                       int nf, int withmask, int strategy, double *__restrict__ ubar,
                                                       T *__restrict__ tau, T *__restrict__ uo, T *__restrict__ vo,
                                                       T *__restrict__ wo) {
    const int64_t b = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t P = G.npts;
    if (t >= P) return;
    const int64_t Nx = G.n[0], Ny = G.n[1];
    const int64_t i = t % Nx, r = t / Nx, j = r % Ny;
    const double *y = yout + b * nf * P;
    // ZeroFill divides by the in-range weight sum of the row; Deformable by the filtered mask
    // (Filtering.jl:3180-3240)
    const double den = (strategy == CGEF_DEFORMABLE && withmask) ? y[(nf - 1) * P + t] : D_row[r];
    double F[9];
    const int np = G.mode == 0 ? 5 : 9;
    for (int f = 0; f < np; f++) F[f] = y[f * P + t];
    pi_point_epilogue<T>(G, t, i, j, F, den, b, ubar, tau, uo, vo, wo);
}

// The inverse row transforms fused with the epilogue (lengths with a compiled-in factorization): a
// block takes NSQ consecutive rows of one snapshot and inverse-transforms that snapshot's planes of
// those rows two at a time (z = X_a + i·X_b gives both real rows), NPR plane pairs per round as
// separate sequences of one transform, folding each transformed plane into per-point state as soon
// as it leaves LDS — the numerators of ū and, on the sphere, of the three local-frame components of τ
// (e_xᵀ F[q qᵀ] e_y is linear in the product planes) — so the filtered planes never return to HBM.
// The epilogue then forms ū = F[q]/den and τ = F[q_a q_b]/den − ū_a ū_b (Diagnostics.jl:496-527,
// 596-650) and writes what k_ring_epilogue writes. Plane order per slice: the NC components, the
// products (xx xy [xz] yy [yz zz]), then the mask when WM (Deformable). 1440-point rows: one row per
// block, three pairs per round (2 rounds for 9 or 10 planes), 512 threads (the compiled-in stages use
// at most 3·144 butterflies, one per thread; every thread shares the per-point work); 256-point rows:
// four rows per block, one pair per round, 256 threads.
template <int FIX>
struct IfeShape;
template <>
struct IfeShape<RING_FIXED_1440> {
    static constexpr int N = 1440, NSQ = 1, NPR = 3, BLK = 512;
};
template <>
struct IfeShape<RING_FIXED_256> {
    static constexpr int N = 256, NSQ = 4, NPR = 1, BLK = 256;
};

template <typename T, int FIX, int MODE, int WM>
__global__ __launch_bounds__(IfeShape<FIX>::BLK) void k_ifft_epilogue(RingFft F, const double2 *__restrict__ tw,
                                                                      const double2 *__restrict__ tws,
                                                                      const double2 *__restrict__ OUT, GeomArgs G,
                                                                      const double *__restrict__ D_row, int64_t nbatch,
                                                                      double *__restrict__ ubar, T *__restrict__ tau,
                                                                      T *__restrict__ uo, T *__restrict__ vo,
                                                                      T *__restrict__ wo, const int32_t *__restrict__ mcut) {
    using S = IfeShape<FIX>;
    constexpr int N = S::N, NSQ = S::NSQ, NPR = S::NPR, BLK = S::BLK;
    constexpr int MH = N / 2 + 1;
    constexpr int PPT = (NSQ * N + BLK - 1) / BLK;           // points per lane
    constexpr int NC = MODE == 0 ? 2 : 3;
    constexpr int NP = MODE == 0 ? 3 : 6;
    constexpr int NF = NC + NP + WM;
    constexpr int NPAIR = (NF + 1) / 2;
    constexpr int NROUND = (NPAIR + NPR - 1) / NPR;
    constexpr int NS = MODE == 1 ? 6 : NC + NP;              // state per point (sphere: ū numerators, 3 local τ sums)
    extern __shared__ double2 sm[];
    const int sh = F.lpsh, sstr = F.sstr;
    const int64_t nrows = G.n[1] * G.n[2], P = G.npts, Ny = G.n[1];
    const int64_t ngrp = (nrows + NSQ - 1) / NSQ;
    int64_t item;
    if (!xcd_item(ngrp * nbatch, item)) return;
    const int64_t b = item % nbatch, r0 = (item / nbatch) * NSQ;   // snapshot fastest: rows' planes share lines
    const int64_t ncolt = nbatch * NF;                       // columns of the [m][row][col] spectra
    const int mstride = (int)(ncolt * nrows);                // one mode (< 2^31)
    const double2 *spec = OUT + r0 * ncolt + b * NF;         // row r0, plane 0, mode 0
    constexpr double sc = 1.0 / N;
    double sp_[NSQ], cp_[NSQ];                               // the rows' sinφ, cosφ (sphere)
#pragma unroll
    for (int q = 0; q < NSQ; q++) {
        const int64_t j = (r0 + q) % Ny;
        sp_[q] = MODE == 1 ? G.trig[2 * G.n[0] + j] : 0.0;
        cp_[q] = MODE == 1 ? G.trig[2 * G.n[0] + Ny + j] : 0.0;
    }
    double st[PPT][NS], dn[PPT];
#pragma unroll
    for (int p = 0; p < PPT; p++) {
        dn[p] = 0.0;
#pragma unroll
        for (int q = 0; q < NS; q++) st[p][q] = 0.0;
    }
#pragma unroll
    for (int rd = 0; rd < NROUND; rd++) {
        const int npr = NPAIR - rd * NPR < NPR ? NPAIR - rd * NPR : NPR;   // pairs this round
        const int nseq = npr * NSQ;                                        // sequence = pair-major, then row
        // compose z = X_a + i·X_b per sequence and its conjugate mirror; lane per (mode, sequence)
        for (int e = threadIdx.x; e < nseq * MH; e += BLK) {
            const int m = e / nseq, sq = e - m * nseq;
            const int k = sq / NSQ, rw = sq - k * NSQ;
            const int fa = 2 * (rd * NPR + k), fb = fa + 1;
            double2 X = make_double2(0.0, 0.0), Y = X;
            if (r0 + rw < nrows && (!mcut || m < mcut[r0 + rw])) {
                const double2 *x = spec + (int64_t)m * mstride + rw * ncolt;   // planes fa, fb adjacent
                X = x[fa];
                if (fb < NF) Y = x[fb];
            }
            double2 *z = sm + sq * sstr;
            z[lp(m, sh)] = make_double2(X.x - Y.y, X.y + Y.x);
            if (m > 0 && N - m >= MH) z[lp(N - m, sh)] = make_double2(X.x + Y.y, Y.x - X.y);
        }
        RingFft Fr = F;
        Fr.nsq = nseq;
        fft_rows_any<FIX, BLK>(sm, Fr, tw, tws, true);
#pragma unroll
        for (int p = 0; p < PPT; p++) {
            const int e = (int)threadIdx.x + p * BLK;
            if (e >= NSQ * N) continue;
            const int rw = e / N, i = e - rw * N;
            double ee[3] = {0.0, 0.0, 0.0}, en[3] = {0.0, 0.0, 0.0};
            if (MODE == 1 && (rd * NPR + NPR) * 2 > NC) {
                // e_east, e_north at this point (a round holding a product plane)
                const double sl = G.trig[i], cl = G.trig[N + i];
                ee[0] = -sl;
                ee[1] = cl;
                en[0] = -sp_[rw] * cl;
                en[1] = -sp_[rw] * sl;
                en[2] = cp_[rw];
            }
#pragma unroll
            for (int k = 0; k < NPR; k++) {
                const int pr = rd * NPR + k;
                if (pr >= NPAIR) continue;
                const double2 z = sm[(k * NSQ + rw) * sstr + lp(i, sh)];
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int f = 2 * pr + h;
                    if (f >= NF) continue;
                    const double x = (h == 0 ? z.x : z.y) * sc;
                    if (f < NC) {
                        st[p][f] = x;
                    } else if (f < NC + NP) {
                        if constexpr (MODE == 1) {
                            // planetary product plane (a, c): its share of e_eᵀFe_e, e_eᵀFe_n, e_nᵀFe_n
                            const int k6 = f - NC;
                            const int a = k6 < 3 ? 0 : (k6 < 5 ? 1 : 2);
                            const int c = k6 < 3 ? k6 : (k6 < 5 ? k6 - 2 : 2);
                            const double wee = a == c ? ee[a] * ee[a] : 2.0 * ee[a] * ee[c];
                            const double wen = a == c ? ee[a] * en[a] : ee[a] * en[c] + ee[c] * en[a];
                            const double wnn = a == c ? en[a] * en[a] : 2.0 * en[a] * en[c];
                            st[p][3] += wee * x;
                            st[p][4] += wen * x;
                            st[p][5] += wnn * x;
                        } else {
                            st[p][f] = x;
                        }
                    } else {
                        dn[p] = x;   // the filtered mask: Deformable's denominator
                    }
                }
            }
        }
        __syncthreads();   // every lane has read the transforms before the next round is composed
    }
#pragma unroll
    for (int p = 0; p < PPT; p++) {
        const int e = (int)threadIdx.x + p * BLK;
        if (e >= NSQ * N) continue;
        const int rw = e / N, i = e - rw * N;
        const int64_t row = r0 + rw;
        if (row >= nrows) continue;
        const int64_t t = row * N + i;
        const double den = WM ? dn[p] : D_row[row];
        const bool ok = G.mask[t] != 0 && den > 1e-15;
        const double inv = ok ? 1.0 / den : 0.0;
        double *ub = ubar + b * 3 * P;
        T *tb = tau + b * 6 * P;
        double o0, o1, o2;
        if constexpr (MODE == 1) {
            const double sl = G.trig[i], cl = G.trig[N + i];
            const double sp = sp_[rw], cp = cp_[rw];
            const double U0 = st[p][0] * inv, U1 = st[p][1] * inv, U2 = st[p][2] * inv;
            o0 = -sl * U0 + cl * U1;
            o1 = -sp * cl * U0 - sp * sl * U1 + cp * U2;
            o2 = cp * cl * U0 + cp * sl * U1 + sp * U2;
            tb[t] = (T)(st[p][3] * inv - o0 * o0);
            tb[P + t] = (T)(st[p][4] * inv - o0 * o1);
            tb[2 * P + t] = (T)(st[p][5] * inv - o1 * o1);
        } else if constexpr (MODE == 0) {
            o0 = st[p][0] * inv;
            o1 = st[p][1] * inv;
            o2 = 0.0;
            tb[t] = (T)(st[p][2] * inv - o0 * o0);
            tb[P + t] = (T)(st[p][3] * inv - o0 * o1);
            tb[2 * P + t] = (T)(st[p][4] * inv - o1 * o1);
        } else {
            o0 = st[p][0] * inv;
            o1 = st[p][1] * inv;
            o2 = st[p][2] * inv;
            const double U[3] = {o0, o1, o2};
#pragma unroll
            for (int q6 = 0; q6 < 6; q6++) {
                const int a = q6 < 3 ? 0 : (q6 < 5 ? 1 : 2);
                const int c = q6 < 3 ? q6 : (q6 < 5 ? q6 - 2 : 2);
                tb[q6 * P + t] = (T)(st[p][3 + q6] * inv - U[a] * U[c]);
            }
        }
        ub[t] = o0;
        ub[P + t] = o1;
        if (MODE == 2) ub[2 * P + t] = o2;
        if (uo) uo[b * P + t] = (T)o0;
        if (vo) vo[b * P + t] = (T)o1;
        if (wo) wo[b * P + t] = (T)o2;
    }
}

// The sphere's fused inverse transform + epilogue with the rotation to the local frame done in Fourier
// space (1440-point rows). The east/north frame weights are trigonometric
// polynomials in λ of degree ≤ 2: with λ_k = λ₀ + 2πk/N, multiplying a row by cos λ or sin λ mixes
// modes m ∓ 1 with phase w = e^{iλ₀}, by cos 2λ or sin 2λ modes m ∓ 2 with phase w². So per mode the
// block forms, from the nine planetary planes, the spectra of the five local numerators
//     E  = −s·U0 + c·U1,                       N   = −sinφ (c·U0 + s·U1) + cosφ·U2,
//     Tee = D1 − Q,   Ten = sinφ·R + cosφ (c·Pyz − s·Pxz),
//     Tnn = sin²φ (D1 + Q) − 2 sinφ cosφ (c·Pxz + s·Pyz) + cos²φ·Pzz,
// with c, s = cos λ, sin λ, D1 = (Pxx + Pyy)/2, D2 = (Pxx − Pyy)/2, Q = cos 2λ·D2 + sin 2λ·Pxy and
// R = sin 2λ·D2 − cos 2λ·Pxy (e_east = (−s, c, 0), e_north = (−sinφ c, −sinφ s, cosφ); Diagnostics.jl:
// 596-650) — and inverse-transforms only those (E + iN, Tee + iTen, Tnn + i·mask): three sequences in
// one round instead of five pairs in two. Modes beyond 0..N/2 come from X[−m] = X[N−m] = conj X[m].
// The products and the components are staged through the LDS that the transforms then use. RAD (no
// mask plane): the radial component cosφ (c·U0 + s·U1) + sinφ·U2 rides in the third sequence's
// imaginary part and is exported through wo.
template <typename T, int WM, int RAD>
__global__ __launch_bounds__(512) void k_ifft_epilogue_sphere(RingFft F, const double2 *__restrict__ tw,
                                                             const double2 *__restrict__ tws,
                                                             const double2 *__restrict__ OUT, GeomArgs G,
                                                             const double *__restrict__ D_row, int64_t nbatch,
                                                             double *__restrict__ ubar, T *__restrict__ tau,
                                                             T *__restrict__ uo, T *__restrict__ vo,
                                                             T *__restrict__ wo, const int32_t *__restrict__ mcut) {
    static_assert(!(WM && RAD), "the mask plane and the radial component share one sequence slot");
    constexpr int N = 1440, MH = N / 2 + 1, BLK = 512;
    constexpr int PPT = (N + BLK - 1) / BLK;                 // points per lane
    constexpr int MPT = (MH + BLK - 1) / BLK;                // modes per lane
    constexpr int NF = 9 + WM;
    constexpr double sc = 1.0 / N;
    extern __shared__ double2 sm[];
    const int sh = F.lpsh, sstr = F.sstr;
    const int64_t nrows = G.n[1] * G.n[2], P = G.npts, Ny = G.n[1];
    int64_t item;
    if (!xcd_item(nrows * nbatch, item)) return;
    // snapshot fastest: the blocks of one row's snapshots run together on one XCD and share the cache
    // lines their adjacent 144-byte plane groups straddle
    const int64_t row = item / nbatch, b = item % nbatch;
    const int mstride = (int)(nbatch * NF * nrows);          // one mode of the [m][row][col] spectra (< 2^31)
    const double2 *spec = OUT + row * (nbatch * NF) + b * NF; // plane f, mode m: spec[m·mstride + f]
    const int64_t j = row % Ny;
    const double sp = G.trig[2 * N + j], cp = G.trig[2 * N + Ny + j];
    const double s0 = G.trig[0], c0 = G.trig[N];             // w = e^{iλ₀}
    const double drow = WM ? 0.0 : D_row[row];               // what the epilogue reads, requested early
    const int mc = mcut ? mcut[row] : MH;                    // modes kept by the cut (the rest are zero)
    bool act[PPT];
#pragma unroll
    for (int p = 0; p < PPT; p++) {
        const int i = (int)threadIdx.x + p * BLK;
        act[p] = i < N && G.mask[row * N + i] != 0;
    }
    const double2 w1 = make_double2(c0, s0), w2 = make_double2(c0 * c0 - s0 * s0, 2.0 * s0 * c0);
    double2 *pl = sm;                                        // staged spectra [plane][MH]
    auto at = [&](int k, int m) {                            // plane k at any mode in −2..MH+1
        if (m < 0) return cconj(pl[k * MH - m]);
        if (m >= MH) return cconj(pl[k * MH + N - m]);
        return pl[k * MH + m];
    };
    auto mi = [](double2 a) { return make_double2(-a.y, a.x); };                   // i·a
    auto half = [](double2 a) { return make_double2(0.5 * a.x, 0.5 * a.y); };
    // the products xx xy xz yy yz zz (planes 3..8: 96 contiguous bytes per mode), every load of the
    // lane issued before the first LDS store
    {
        constexpr int NA = (6 * MH + BLK - 1) / BLK;
        double2 v[NA];
#pragma unroll
        for (int q = 0; q < NA; q++) {
            const int e = (int)threadIdx.x + q * BLK;
            const int m = e / 6, k = e - 6 * m;
            v[q] = e < 6 * MH && m < mc ? spec[(int64_t)m * mstride + 3 + k] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < NA; q++) {
            const int e = (int)threadIdx.x + q * BLK;
            const int m = e / 6, k = e - 6 * m;
            if (e < 6 * MH) pl[k * MH + m] = v[q];
        }
    }
    __syncthreads();
    // the components U0 U1 U2 (planes 0..2) and, for Deformable, the mask (plane 9): requested now, so
    // that their latency runs under the products' mixing; they reach LDS once the products are read
    constexpr int NC4 = 3 + WM, NB = (NC4 * MH + BLK - 1) / BLK;
    double2 cv[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int e = (int)threadIdx.x + q * BLK;
        const int m = e / NC4, k = e - NC4 * m;
        cv[q] = e < NC4 * MH && m < mc ? spec[(int64_t)m * mstride + (k < 3 ? k : 9)] : make_double2(0.0, 0.0);
    }
    double2 Tee[MPT], Ten[MPT], Tnn[MPT];
#pragma unroll
    for (int q = 0; q < MPT; q++) {
        const int m = (int)threadIdx.x + q * BLK;
        if (m >= MH) continue;
        // G = D2 − i·Pxy at m − 2, H = D2 + i·Pxy at m + 2: Q = (w²G + w̄²H)/2, R = (w²G − w̄²H)/(2i)
        const double2 d2l = half(csub(at(0, m - 2), at(3, m - 2))), d2h = half(csub(at(0, m + 2), at(3, m + 2)));
        const double2 gl = cmul(w2, csub(d2l, mi(at(1, m - 2))));
        const double2 gh = cmul(cconj(w2), cadd(d2h, mi(at(1, m + 2))));
        const double2 Q = half(cadd(gl, gh));
        const double2 dq = half(csub(gl, gh));
        const double2 R = make_double2(dq.y, -dq.x);                             // dq / i
        const double2 D1 = half(cadd(at(0, m), at(3, m)));
        // c·Pxz + s·Pyz and c·Pyz − s·Pxz: (w·(X − iY)[m−1] + w̄·(X + iY)[m+1]) / 2 with (X, Y) = (Pxz, Pyz), (Pyz, −Pxz)
        const double2 xl = at(2, m - 1), yl = at(4, m - 1), xh = at(2, m + 1), yh = at(4, m + 1);
        const double2 A1 = half(cadd(cmul(w1, csub(xl, mi(yl))), cmul(cconj(w1), cadd(xh, mi(yh)))));
        const double2 B1 = half(cadd(cmul(w1, cadd(yl, mi(xl))), cmul(cconj(w1), csub(yh, mi(xh)))));
        const double2 pzz = at(5, m);
        Tee[q] = csub(D1, Q);
        Ten[q] = make_double2(sp * R.x + cp * B1.x, sp * R.y + cp * B1.y);
        Tnn[q] = make_double2(sp * sp * (D1.x + Q.x) - 2.0 * sp * cp * A1.x + cp * cp * pzz.x,
                              sp * sp * (D1.y + Q.y) - 2.0 * sp * cp * A1.y + cp * cp * pzz.y);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int e = (int)threadIdx.x + q * BLK;
        const int m = e / NC4, k = e - NC4 * m;
        if (e < NC4 * MH) pl[k * MH + m] = cv[q];
    }
    __syncthreads();
    double2 Ev[MPT], Nv[MPT], Mv[MPT];
#pragma unroll
    for (int q = 0; q < MPT; q++) {
        const int m = (int)threadIdx.x + q * BLK;
        if (m >= MH) continue;
        const double2 u0l = at(0, m - 1), u1l = at(1, m - 1), u0h = at(0, m + 1), u1h = at(1, m + 1);
        // −s·U0 + c·U1 = (w·(U1 + iU0)[m−1] + w̄·(U1 − iU0)[m+1]) / 2;  c·U0 + s·U1 likewise with (U0, −U1)
        Ev[q] = half(cadd(cmul(w1, cadd(u1l, mi(u0l))), cmul(cconj(w1), csub(u1h, mi(u0h)))));
        const double2 cs = half(cadd(cmul(w1, csub(u0l, mi(u1l))), cmul(cconj(w1), cadd(u0h, mi(u1h)))));
        const double2 u2 = at(2, m);
        Nv[q] = make_double2(cp * u2.x - sp * cs.x, cp * u2.y - sp * cs.y);
        if (WM) Mv[q] = pl[3 * MH + m];
        else if (RAD) Mv[q] = make_double2(cp * cs.x + sp * u2.x, cp * cs.y + sp * u2.y);
        else Mv[q] = make_double2(0.0, 0.0);
    }
    __syncthreads();   // the staged spectra are read: the sequences may overwrite them
#pragma unroll
    for (int q = 0; q < MPT; q++) {
        const int m = (int)threadIdx.x + q * BLK;
        if (m >= MH) continue;
        const double2 X[3] = {Ev[q], Tee[q], Tnn[q]}, Y[3] = {Nv[q], Ten[q], Mv[q]};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double2 *z = sm + k * sstr;
            z[lp(m, sh)] = make_double2(X[k].x - Y[k].y, X[k].y + Y[k].x);
            if (m > 0 && N - m >= MH) z[lp(N - m, sh)] = make_double2(X[k].x + Y[k].y, Y[k].x - X[k].y);
        }
    }
    RingFft Fr = F;
    Fr.nsq = 3;
    fft_rows_any<RING_FIXED_1440, BLK>(sm, Fr, tw, tws, true);
    double *ub = ubar + b * 3 * P;
    T *tb = tau + b * 6 * P;
#pragma unroll
    for (int p = 0; p < PPT; p++) {
        const int i = (int)threadIdx.x + p * BLK;
        if (i >= N) continue;
        const double2 z0 = sm[lp(i, sh)], z1 = sm[sstr + lp(i, sh)], z2 = sm[2 * sstr + lp(i, sh)];
        const int64_t t = row * N + i;
        const double den = WM ? z2.y * sc : drow;
        const bool ok = act[p] && den > 1e-15;
        const double inv = ok ? sc / den : 0.0;
        const double o0 = z0.x * inv, o1 = z0.y * inv;
        tb[t] = (T)(z1.x * inv - o0 * o0);
        tb[P + t] = (T)(z1.y * inv - o0 * o1);
        tb[2 * P + t] = (T)(z2.x * inv - o1 * o1);
        ub[t] = o0;
        ub[P + t] = o1;
        if (uo) uo[b * P + t] = (T)o0;
        if (vo) vo[b * P + t] = (T)o1;
        if (RAD) wo[b * P + t] = (T)(z2.y * inv);
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

// ---- spectral engine (FFTWExt.jl:121-208): the row spectra XF[m][col][row] of a doubly periodic
// 2-D grid, per (m, col) a contiguous column of Ny values, are transformed along y, multiplied by the
// kernel's transfer function Ĝ(|k|) and transformed back, so the row inverse yields
// irfft2(Ĝ · rfft2(mask·f)).

// Kernels.jl spectral_transfer: Gaussian exp(−k²ℓ²/(4α)); top-hat 2·J₁(x)/x with x = kℓ/2.
__device__ __forceinline__ double spectral_transfer(int kind, double alpha, double k, double ell) {
    if (kind == CGEF_GAUSSIAN) return exp(-k * k * ell * ell / (4.0 * alpha));
    const double x = 0.5 * k * ell;
    return x == 0.0 ? 1.0 : 2.0 * j1(x) / x;
}

// MODE 0: transform along y, × Ĝ/Ny, transform back (one scale); MODE 1: the forward transform only
// (the analysis a sweep over scales shares, FFTWExt.jl:100-110); MODE 2: × Ĝ/Ny and back (one
// scale's synthesis from the shared analysis, FFTWExt.jl:112-120).
// LPT > 0: the register-staged transform (padded LDS, LPT·256 ≥ Ny); LPT = 0: the LDS ping-pong.
template <int LPT, int MODE>
__global__ __launch_bounds__(256) void k_spec_cols(RingFft F, const double2 *__restrict__ tw,
                                                   const double2 *__restrict__ tws, int kind, double alpha, double ell,
                                                   double dkx, double dky, const double2 *__restrict__ XF,
                                                   double2 *__restrict__ OUT, int64_t ncol, int64_t mh) {
    extern __shared__ double2 sm[];
    constexpr bool STAGED = LPT > 0;
    const int n = F.n;
    int64_t item;
    if (!xcd_item(mh * ncol, item)) return;
    const int64_t m = item / ncol;
    const double2 *x = XF + item * n;
    const int sh = STAGED ? F.lpsh : 31;
    auto at = [sh](int e) { return lp(e, sh); };
    if constexpr (STAGED) {  // all loads in flight together
        double2 v_[LPT];
#pragma unroll
        for (int q = 0; q < LPT; q++) {
            const int e = (int)threadIdx.x + q * 256;
            v_[q] = e < n ? x[e] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < LPT; q++) {
            const int e = (int)threadIdx.x + q * 256;
            if (e < n) sm[lp(e, sh)] = v_[q];
        }
    } else {
        for (int e = threadIdx.x; e < n; e += 256) sm[e] = x[e];
    }
    double2 *z = sm;
    if (MODE != 2) {
        if (STAGED) fft_staged(sm, F, tw, tws, false);
        else z = fft_lds(sm, sm + n, F, tw, false, 1);
    }
    double2 *r = z;
    if (MODE != 1) {
        if (MODE == 2) __syncthreads();
        // angular wavenumbers 2π·m/Lx (rfftfreq) and 2π·n/Ly with n folded to −Ny/2..(Ny−1)/2 (fftfreq)
        const double kx = dkx * (double)m, sc = 1.0 / (double)n;
        for (int e = threadIdx.x; e < n; e += 256) {
            const int nn = e <= (n - 1) / 2 ? e : e - n;
            const double ky = dky * (double)nn;
            const double g = spectral_transfer(kind, alpha, sqrt(kx * kx + ky * ky), ell) * sc;
            const double2 c = z[at(e)];
            z[at(e)] = make_double2(c.x * g, c.y * g);
        }
        if (STAGED) fft_staged(z, F, tw, tws, true);
        else r = fft_lds(z, z == sm ? sm + n : sm, F, tw, true, 1);
    }
    if (MODE == 1) __syncthreads();
    double2 *o = OUT + item * n;
    for (int e = threadIdx.x; e < n; e += 256) o[e] = r[at(e)];
}

// FFTWExt.jl:162-176: the filtered mask inverted where |filter(mask)| ≥ 0.01, else 0
__global__ __launch_bounds__(256) void k_spec_invert(int64_t P, const double *__restrict__ r, double *__restrict__ ir) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= P) return;
    const double v = r[t];
    ir[t] = fabs(v) >= 0.01 ? 1.0 / v : 0.0;
}

// filter_apply: the transform's output (× 1/filter(mask) for Deformable on a masked grid), with no
// masking of the output (FFTWExt.jl:193-208)
__global__ __launch_bounds__(256) void k_spec_normalize(GeomArgs G, int dtype, const double *__restrict__ yout,
                                                        const double *__restrict__ invrenorm, void *out) {
    const int64_t s = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t P = G.npts;
    if (t >= P) return;
    const double v = invrenorm ? yout[s * P + t] * invrenorm[t] : yout[s * P + t];
    if (dtype == CGEF_F32) ((float *)out)[s * P + t] = (float)v;
    else ((double *)out)[s * P + t] = v;
}

// Π from the five filtered planes of a 2-D Cartesian slice: every point carries the transform's value
// (the strain reads masked neighbours as the filter left them); Π itself is masked by the strain kernel.
template <typename T>
__global__ __launch_bounds__(256) void k_spec_epilogue(GeomArgs G, const double *__restrict__ yout,
                                                       const double *__restrict__ invrenorm, double *__restrict__ ubar,
                                                       T *__restrict__ tau, T *__restrict__ uo, T *__restrict__ vo,
                                                       T *__restrict__ wo) {
    const int64_t b = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t P = G.npts;
    if (t >= P) return;
    const int64_t Nx = G.n[0];
    const int64_t i = t % Nx, j = t / Nx;
    const double *y = yout + b * 5 * P;
    double F[9];
#pragma unroll
    for (int f = 0; f < 5; f++) F[f] = y[f * P + t];
    pi_point_epilogue_inv<T>(G, t, i, j, F, true, invrenorm ? invrenorm[t] : 1.0, b, ubar, tau, uo, vo, wo);
}

// ------------------------------------------------------------------------------------ launches ----

static size_t fft_shmem(const RingFft &F) { return (size_t)2 * F.n * sizeof(double2); }
static size_t staged_shmem(const RingFft &F) { return (size_t)F.nsq * F.sstr * sizeof(double2); }

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

// The staged row transform of a launch: the plan's row pairs per block, halved while the launch would
// have fewer than 4 blocks per CU (small grids keep their parallelism).
// The compile-time factorization serves a launch when every stage is one butterfly per lane.
static bool fixed_fits(const RingFft &F) {
    if (F.fixed == 0) return false;
    for (int st = 0; st < F.snst; st++)
        if (F.n / F.srad[st] * F.nsq > 256) return false;
    return true;
}

static RingFft launch_rows_fft(const RingPlan &R, int64_t ncol) {
    RingFft F = R.fft;
    while (F.nsq > 1 && (R.args.nrows + 2 * F.nsq - 1) / (2 * F.nsq) * ncol < 4 * 256) F.nsq /= 2;
    return F;
}

void ring_forward(const RingPlan &R, const double *xin, double2 *XF, int64_t ncol, hipStream_t s) {
    if (R.fft.nseq > 0) {
        const RingFft F = launch_rows_fft(R, ncol);
        const int64_t items = (R.args.nrows + 2 * F.nsq - 1) / (2 * F.nsq) * ncol;
        const dim3 grid((unsigned)(8 * ((items + 7) / 8)));
        const size_t sh = staged_shmem(F);
        if (fixed_fits(F) && F.fixed == RING_FIXED_1440) {
            hipLaunchKernelGGL((k_fft_rows_staged<8, RING_FIXED_1440>), grid, dim3(256), sh, s, F, R.tw_d, R.tws_d, xin, XF,
                               R.args.nrows, ncol);
        } else if (fixed_fits(F) && F.fixed == RING_FIXED_256) {  // ≤ 4 row pairs: 1024 values per block
            hipLaunchKernelGGL((k_fft_rows_staged<4, RING_FIXED_256>), grid, dim3(256), sh, s, F, R.tw_d, R.tws_d, xin, XF,
                               R.args.nrows, ncol);
        } else if (F.n <= 2048) {
            hipLaunchKernelGGL((k_fft_rows_staged<8, 0>), grid, dim3(256), sh, s, F, R.tw_d, R.tws_d, xin, XF, R.args.nrows,
                               ncol);
        } else {
            set_shmem((const void *)k_fft_rows_staged<16, 0>, sh);
            hipLaunchKernelGGL((k_fft_rows_staged<16, 0>), grid, dim3(256), sh, s, F, R.tw_d, R.tws_d, xin, XF, R.args.nrows,
                               ncol);
        }
        return;
    }
    const int64_t items = (R.args.nrows + 1) / 2 * ncol;
    const dim3 grid((unsigned)(8 * ((items + 7) / 8)));
    const size_t sh = fft_shmem(R.fft);
    set_shmem((const void *)k_fft_rows, sh);
    hipLaunchKernelGGL(k_fft_rows, grid, dim3(256), sh, s, R.fft, R.tw_d, xin, XF, R.args.nrows, ncol);
}

void ring_inverse(const RingPlan &R, const double2 *OUT, double *yout, int64_t ncol, hipStream_t s, const int32_t *mcut) {
    if (R.fft.nseq > 0) {
        const RingFft F = launch_rows_fft(R, ncol);
        const int64_t items = (R.args.nrows + 2 * F.nsq - 1) / (2 * F.nsq) * ncol;
        const dim3 grid((unsigned)(8 * ((items + 7) / 8)));
        const size_t sh = staged_shmem(F);
        if (fixed_fits(F) && F.fixed == RING_FIXED_1440) {
            hipLaunchKernelGGL((k_ifft_rows_staged<9, RING_FIXED_1440>), grid, dim3(256), sh, s, F, R.tw_d, R.tws_d, OUT,
                               yout, R.args.nrows, ncol, mcut);
        } else if (fixed_fits(F) && F.fixed == RING_FIXED_256) {  // ≤ 4 row pairs: 8 × 129 stored modes
            hipLaunchKernelGGL((k_ifft_rows_staged<5, RING_FIXED_256>), grid, dim3(256), sh, s, F, R.tw_d, R.tws_d, OUT,
                               yout, R.args.nrows, ncol, mcut);
        } else if (F.n <= 2048) {
            hipLaunchKernelGGL((k_ifft_rows_staged<9, 0>), grid, dim3(256), sh, s, F, R.tw_d, R.tws_d, OUT, yout,
                               R.args.nrows, ncol, mcut);
        } else {
            set_shmem((const void *)k_ifft_rows_staged<17, 0>, sh);
            hipLaunchKernelGGL((k_ifft_rows_staged<17, 0>), grid, dim3(256), sh, s, F, R.tw_d, R.tws_d, OUT, yout,
                               R.args.nrows, ncol, mcut);
        }
        return;
    }
    const int64_t items = (R.args.nrows + 1) / 2 * ncol;
    const dim3 grid((unsigned)(8 * ((items + 7) / 8)));
    const size_t sh = fft_shmem(R.fft);
    set_shmem((const void *)k_ifft_rows, sh);
    hipLaunchKernelGGL(k_ifft_rows, grid, dim3(256), sh, s, R.fft, R.tw_d, OUT, yout, R.args.nrows, ncol, mcut);
}

// Weight batch depth and occupancy bound of the contraction, from the blocks per CU its LDS slab
// allows (160 KiB per CU), measured on MI355X: with at most 5 resident blocks the deepest prefetch
// (PB 8, 94 VGPRs); with room for 7 and many blocks to schedule (many columns, or several dz groups
// of few offsets each) PB 4 at 72 VGPRs; otherwise PB 6 held to 80 VGPRs for 6 waves per SIMD.
// The contraction comes in three variants (weight batch PB, waves per SIMD held): 0 = <8, 1>,
// 1 = <4, 1>, 2 = <6, 6>. Which is fastest depends on the footprint's row span (LDS slab, hence
// blocks per CU), the number of vertical offset groups and the column count. The first call of a
// plan for a column count times all three on that call's operands (each overwrites OUT with the same
// sums; the chosen one runs last), each by the median of three launches, and keeps the fastest; later calls, and calls under stream capture,
// use the kept choice or, before any measurement, the occupancy rule below. CGEF_TUNE=0 turns the
// measurement off; CGEF_CONTRACT_VARIANT=0|1|2 forces a variant (testing).
static int contract_rule(const RingPlan &R, int64_t ncol, size_t lds_blocks) {
    if (lds_blocks <= 5) return 0;
    if (lds_blocks >= 7 && (R.args.ndz > 1 || ncol > 16)) return 1;
    return 2;
}

// Column tiles (16 real columns each) per workgroup of the matrix-core contraction for a launch of
// ncol complex columns: the instance that wastes the fewest padded columns, the wider on a tie.
static int mfma_nt(int64_t ncol) {
    const int64_t need = (2 * ncol + 15) / 16;
    const int cand[5] = {9, 5, 4, 3, 2};
    int best = 9;
    int64_t best_cols = INT64_MAX;
    for (int nt : cand) {
        const int64_t cols = (need + nt - 1) / nt * nt;
        if (cols < best_cols) {
            best_cols = cols;
            best = nt;
        }
    }
    return best;
}

static void launch_contract_mfma(const RingPlan &R, const double2 *XF, double2 *OUT, int64_t ncol, int orm,
                                 bool cut, hipStream_t s) {
    const int nt = mfma_nt(ncol);
    const int64_t nrb = (R.args.ny + 63) / 64;
    const int64_t ncc = (2 * ncol + nt * 16 - 1) / (nt * 16);
    const int32_t *wl = cut ? R.mitems_d : nullptr;
    const int64_t nwl = cut ? R.nmitems : 0;
    const int64_t items = (cut ? nwl : R.mh * nrb) * ncc;
    const dim3 grid((unsigned)(8 * ((items + 7) / 8)));
    const size_t sh = 2 * (size_t)MFMA_KC * (nt * 16 + 2) * sizeof(double);   // two staging buffers
    auto go = [&](auto kern) {
        set_shmem((const void *)kern, sh);
        hipLaunchKernelGGL(kern, grid, dim3(256), sh, s, R.args, R.wt_d, R.wt_steps, R.wt_tiles, XF, OUT, ncol, orm,
                           wl, nwl);
    };
    if (nt == 9) go(k_contract_mfma<9>);
    else if (nt == 5) go(k_contract_mfma<5>);
    else if (nt == 4) go(k_contract_mfma<4>);
    else if (nt == 3) go(k_contract_mfma<3>);
    else go(k_contract_mfma<2>);
}

void ring_mode_band_cut(const RingPlan &R, double ell, const double *dxb, double *S, int32_t *mband, hipStream_t s) {
    const RingArgs &A = R.args;
    const int64_t blocks = std::min<int64_t>((A.mh * A.nbands + 255) / 256, 8 * 1024);
    hipLaunchKernelGGL(k_mode_sums, dim3((unsigned)blocks), dim3(256), 0, s, A, S);
    hipLaunchKernelGGL(k_mode_cut, dim3((unsigned)((A.nbands + 255) / 256)), dim3(256), 0, s, A, R.fft.n, RING_MODE_TOL,
                       RING_MODE_SLOPE_TOL, ell, dxb, (const double *)S, mband);
}

void ring_build_wtiles(RingPlan &R, hipStream_t s) {
    const int64_t total = R.mh * R.wt_tiles * R.wt_steps * 64;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 8 * 1024);
    hipLaunchKernelGGL(k_build_wtiles, dim3((unsigned)blocks), dim3(256), 0, s, R.args, R.wt_steps, R.wt_tiles, R.wt_d);
}

const int32_t *ring_mode_cut(const RingPlan &R, int dtype) {
    if (dtype != CGEF_F32 || !R.mcut_d) return nullptr;
    const char *e = std::getenv("CGEF_MODE_CUT");
    return e && e[0] == '0' ? nullptr : R.mcut_d;
}

void ring_contract(const RingPlan &R, const double2 *XF, double2 *OUT, int64_t ncol, hipStream_t s, bool out_rowmajor,
                   const int32_t *mcut) {
    constexpr int CC = RING_CONTRACT_CC;
    const int span = 256 + R.args.dy_max - R.args.dy_min;
    const size_t sh = (size_t)CC * span * sizeof(double2);
    const int64_t nyb = (R.args.ny + 255) / 256;
    const bool cut = mcut != nullptr;
    const int32_t *wl = cut ? R.citems_d : nullptr;
    const int64_t nwl = cut ? R.ncitems : 0;
    const int64_t items = (ncol + CC - 1) / CC * (cut ? nwl : nyb * R.args.nz * R.mh);
    const dim3 grid((unsigned)(8 * ((items + 7) / 8)));
    const size_t lds_blocks = LDS_BYTES_PER_CU / sh;
    const bool has_mfma = R.wt_d != nullptr;
    const int nvar = has_mfma ? 4 : 3;
    const int orm = out_rowmajor ? 1 : 0;
    const int64_t key = 4 * ncol + 2 * (cut ? 1 : 0) + orm;   // the kept variant per column count, cut and layout
    auto launch = [&](int variant) {
        auto go = [&](auto kern) {
            set_shmem((const void *)kern, sh);
            hipLaunchKernelGGL(kern, grid, dim3(256), sh, s, R.args, XF, OUT, ncol, orm, wl, nwl, mcut);
        };
        if (variant == 3) launch_contract_mfma(R, XF, OUT, ncol, orm, cut, s);
        else if (variant == 0) go(k_contract<CC, 8, 1>);
        else if (variant == 1) go(k_contract<CC, 4, 1>);
        else go(k_contract<CC, 6, 6>);
    };
    if (const char *f = std::getenv("CGEF_CONTRACT_VARIANT")) {
        const int v = std::atoi(f);
        if (v >= 0 && v < nvar) {
            launch(v);
            return;
        }
    }
    int variant = -1;
    {
        std::lock_guard<std::mutex> lk(R.tune->mu);
        for (const auto &c : R.tune->choice)
            if (c.first == key) variant = c.second;
    }
    if (variant >= 0) {
        launch(variant);
        return;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    CGEF_HIP(hipStreamIsCapturing(s, &cap));
    const char *tune = std::getenv("CGEF_TUNE");
    if (cap != hipStreamCaptureStatusNone || (tune && tune[0] == '0')) {
        launch(contract_rule(R, ncol, lds_blocks));
        return;
    }
    // each variant: one untimed launch (code load, LDS limit), then the median of three timed ones
    constexpr int kReps = 3;
    hipEvent_t ev[kReps + 1];
    for (auto &e : ev) CGEF_HIP(hipEventCreate(&e));
    float best_ms = 0.0f;
    int best = contract_rule(R, ncol, lds_blocks);
    for (int v = 0; v < nvar; ++v) {
        launch(v);
        CGEF_HIP(hipEventRecord(ev[0], s));
        for (int r = 0; r < kReps; ++r) {
            launch(v);
            CGEF_HIP(hipEventRecord(ev[r + 1], s));
        }
        CGEF_HIP(hipEventSynchronize(ev[kReps]));
        float t[kReps];
        for (int r = 0; r < kReps; ++r) CGEF_HIP(hipEventElapsedTime(&t[r], ev[r], ev[r + 1]));
        std::sort(t, t + kReps);
        const float ms = t[kReps / 2];
        if (v == 0 || ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    launch(best);
    std::lock_guard<std::mutex> lk(R.tune->mu);
    R.tune->choice.emplace_back(key, best);
}

void ring_prep(const GeomArgs &G, int dtype, const void *u, const void *v, const void *w, int nf, int withmask,
               double *xin, int64_t nbatch, hipStream_t s) {
    if (G.npts % 2 == 0) {
        dim3 grid((unsigned)((G.npts / 2 + 255) / 256), (unsigned)nbatch);
        hipLaunchKernelGGL(k_ring_prep<2>, grid, dim3(256), 0, s, G, dtype, u, v, w, nf, withmask, xin);
    } else {
        dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)nbatch);
        hipLaunchKernelGGL(k_ring_prep<1>, grid, dim3(256), 0, s, G, dtype, u, v, w, nf, withmask, xin);
    }
}

// The fused inverse + epilogue where the row length has a compiled-in factorization that one block's
// sequences fit (1440 points: one row per block, 6 points per lane; 256 points: 4 rows, 4 per lane);
// false when the two separate kernels must serve.
bool ring_inverse_epilogue_serves(const RingPlan &R, const GeomArgs &G, int nf, int withmask, int strategy,
                                  int64_t nbatch) {
    const RingFft &F = R.fft;
    if (F.nseq == 0 || (F.fixed != RING_FIXED_1440 && F.fixed != RING_FIXED_256)) return false;
    if (withmask && strategy != CGEF_DEFORMABLE) return false;
    const int mode = G.mode;
    if (nf != (mode == 0 ? 5 : 9) + withmask) return false;
    if (F.fixed == RING_FIXED_1440 && mode != 1) return false;   // 1440 points: the 0.25° longitude
    const int64_t nrows = G.n[1] * G.n[2];
    if (nbatch * nf * nrows >= ((int64_t)1 << 31) / R.mh) return false;   // 32-bit mode offsets
    const bool big = F.fixed == RING_FIXED_1440;
    const int nsq = big ? IfeShape<RING_FIXED_1440>::NSQ : IfeShape<RING_FIXED_256>::NSQ;
    const int npr = big ? IfeShape<RING_FIXED_1440>::NPR : IfeShape<RING_FIXED_256>::NPR;
    const int blk = big ? IfeShape<RING_FIXED_1440>::BLK : IfeShape<RING_FIXED_256>::BLK;
    for (int st = 0; st < F.snst; st++)   // one butterfly per lane in every stage, all of a round's sequences
        if (F.n / F.srad[st] * nsq * npr > blk) return false;
    return true;
}

void ring_inverse_epilogue(const RingPlan &R, const GeomArgs &G, int dtype, const double2 *OUT, const double *D_row,
                           int nf, int withmask, int strategy, double *ubar, void *tau, void *uo, void *vo, void *wo,
                           int64_t nbatch, bool spectral_rotation, hipStream_t s, const int32_t *mcut) {
    if (!ring_inverse_epilogue_serves(R, G, nf, withmask, strategy, nbatch))
        fail(CGEF_ERR_UNSUPPORTED, "the fused inverse transform + epilogue does not serve this plan");
    const RingFft &F = R.fft;
    const int mode = G.mode;
    const int64_t nrows = G.n[1] * G.n[2];
    const bool big = F.fixed == RING_FIXED_1440;
    if (spectral_rotation && big && !(wo && withmask)) {
        // the sphere's local frame formed in Fourier space: three sequences per row in one round
        bool fits = (size_t)6 * R.mh <= (size_t)3 * F.sstr;
        for (int st = 0; st < F.snst; st++) fits = fits && F.n / F.srad[st] * 3 <= 512;
        if (fits) {
            const int64_t items = nrows * nbatch;
            const dim3 grid((unsigned)(8 * ((items + 7) / 8)));
            const size_t sh = (size_t)3 * F.sstr * sizeof(double2);
            auto go = [&](auto kern, auto *tp) {
                using TT = std::remove_pointer_t<decltype(tp)>;
                set_shmem((const void *)kern, sh);
                hipLaunchKernelGGL(kern, grid, dim3(512), sh, s, F, R.tw_d, R.tws_d, OUT, G, D_row, nbatch, ubar,
                                   (TT *)tau, (TT *)uo, (TT *)vo, (TT *)wo, mcut);
            };
            if (dtype == CGEF_F32) {
                if (withmask) go(k_ifft_epilogue_sphere<float, 1, 0>, (float *)nullptr);
                else if (wo) go(k_ifft_epilogue_sphere<float, 0, 1>, (float *)nullptr);
                else go(k_ifft_epilogue_sphere<float, 0, 0>, (float *)nullptr);
            } else {
                if (withmask) go(k_ifft_epilogue_sphere<double, 1, 0>, (double *)nullptr);
                else if (wo) go(k_ifft_epilogue_sphere<double, 0, 1>, (double *)nullptr);
                else go(k_ifft_epilogue_sphere<double, 0, 0>, (double *)nullptr);
            }
            return;
        }
    }
    const int nsq = big ? IfeShape<RING_FIXED_1440>::NSQ : IfeShape<RING_FIXED_256>::NSQ;
    const int npr = big ? IfeShape<RING_FIXED_1440>::NPR : IfeShape<RING_FIXED_256>::NPR;
    const int blk = big ? IfeShape<RING_FIXED_1440>::BLK : IfeShape<RING_FIXED_256>::BLK;
    const int64_t items = (nrows + nsq - 1) / nsq * nbatch;
    const dim3 grid((unsigned)(8 * ((items + 7) / 8)));
    const size_t sh = (size_t)nsq * npr * F.sstr * sizeof(double2);
    auto go = [&](auto kern, auto *tp) {
        using TT = std::remove_pointer_t<decltype(tp)>;
        set_shmem((const void *)kern, sh);
        hipLaunchKernelGGL(kern, grid, dim3(blk), sh, s, F, R.tw_d, R.tws_d, OUT, G, D_row, nbatch, ubar, (TT *)tau,
                           (TT *)uo, (TT *)vo, (TT *)wo, mcut);
    };
#define CGEF_IFE(TT, FIXV, MODEV)                                                                    \
    if (withmask) go(k_ifft_epilogue<TT, FIXV, MODEV, 1>, (TT *)nullptr);                           \
    else go(k_ifft_epilogue<TT, FIXV, MODEV, 0>, (TT *)nullptr);
    if (big) {
        if (dtype == CGEF_F32) { CGEF_IFE(float, RING_FIXED_1440, 1) } else { CGEF_IFE(double, RING_FIXED_1440, 1) }
    } else if (dtype == CGEF_F32) {
        if (mode == 0) { CGEF_IFE(float, RING_FIXED_256, 0) }
        else if (mode == 1) { CGEF_IFE(float, RING_FIXED_256, 1) }
        else { CGEF_IFE(float, RING_FIXED_256, 2) }
    } else {
        if (mode == 0) { CGEF_IFE(double, RING_FIXED_256, 0) }
        else if (mode == 1) { CGEF_IFE(double, RING_FIXED_256, 1) }
        else { CGEF_IFE(double, RING_FIXED_256, 2) }
    }
#undef CGEF_IFE
}

void ring_epilogue(const GeomArgs &G, int dtype, const double *yout, const double *D_row, int nf, int withmask,
                   int strategy, double *ubar, void *tau, void *uo, void *vo, void *wo, int64_t nbatch, hipStream_t s) {
    dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)nbatch);
    if (dtype == CGEF_F32)
        hipLaunchKernelGGL(k_ring_epilogue<float>, grid, dim3(256), 0, s, G, yout, D_row, nf, withmask, strategy, ubar,
                           (float *)tau, (float *)uo, (float *)vo, (float *)wo);
    else
        hipLaunchKernelGGL(k_ring_epilogue<double>, grid, dim3(256), 0, s, G, yout, D_row, nf, withmask, strategy, ubar,
                           (double *)tau, (double *)uo, (double *)vo, (double *)wo);
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

template <int MODE>
static void spec_cols(const RingPlan &R, const double2 *XF, double2 *OUT, int64_t ncol, hipStream_t s) {
    const int64_t items = R.mh * ncol;
    const dim3 grid((unsigned)(8 * ((items + 7) / 8)));
    if (R.fy.nseq > 0) {
        const size_t sh = staged_shmem(R.fy);
        if (R.fy.n <= 2048) {
            hipLaunchKernelGGL((k_spec_cols<8, MODE>), grid, dim3(256), sh, s, R.fy, R.twy_d, R.twsy_d, R.kind, R.alpha,
                               R.ell, R.dkx, R.dky, XF, OUT, ncol, R.mh);
        } else {
            set_shmem((const void *)k_spec_cols<16, MODE>, sh);
            hipLaunchKernelGGL((k_spec_cols<16, MODE>), grid, dim3(256), sh, s, R.fy, R.twy_d, R.twsy_d, R.kind, R.alpha,
                               R.ell, R.dkx, R.dky, XF, OUT, ncol, R.mh);
        }
    } else {
        const size_t sh = fft_shmem(R.fy);
        set_shmem((const void *)k_spec_cols<0, MODE>, sh);
        hipLaunchKernelGGL((k_spec_cols<0, MODE>), grid, dim3(256), sh, s, R.fy, R.twy_d, R.twsy_d, R.kind, R.alpha,
                           R.ell, R.dkx, R.dky, XF, OUT, ncol, R.mh);
    }
}

void ring_spectral_cols(const RingPlan &R, const double2 *XF, double2 *OUT, int64_t ncol, hipStream_t s) {
    spec_cols<0>(R, XF, OUT, ncol, s);
}
void ring_spectral_analyze_cols(const RingPlan &R, const double2 *XF, double2 *OUT, int64_t ncol, hipStream_t s) {
    spec_cols<1>(R, XF, OUT, ncol, s);
}
void ring_spectral_synth_cols(const RingPlan &R, const double2 *XF, double2 *OUT, int64_t ncol, hipStream_t s) {
    spec_cols<2>(R, XF, OUT, ncol, s);
}

void ring_spectral_invert(const GeomArgs &G, const double *renorm, double *invrenorm, hipStream_t s) {
    hipLaunchKernelGGL(k_spec_invert, dim3((unsigned)((G.npts + 255) / 256)), dim3(256), 0, s, G.npts, renorm, invrenorm);
}

void ring_spectral_normalize(const GeomArgs &G, int dtype, const double *yout, const double *invrenorm, int64_t nsl,
                             void *out, hipStream_t s) {
    dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)nsl);
    hipLaunchKernelGGL(k_spec_normalize, grid, dim3(256), 0, s, G, dtype, yout, invrenorm, out);
}

void ring_spectral_epilogue(const GeomArgs &G, int dtype, const double *yout, const double *invrenorm, double *ubar,
                            void *tau, void *uo, void *vo, void *wo, int64_t nbatch, hipStream_t s) {
    dim3 grid((unsigned)((G.npts + 255) / 256), (unsigned)nbatch);
    if (dtype == CGEF_F32)
        hipLaunchKernelGGL(k_spec_epilogue<float>, grid, dim3(256), 0, s, G, yout, invrenorm, ubar, (float *)tau,
                           (float *)uo, (float *)vo, (float *)wo);
    else
        hipLaunchKernelGGL(k_spec_epilogue<double>, grid, dim3(256), 0, s, G, yout, invrenorm, ubar, (double *)tau,
                           (double *)uo, (double *)vo, (double *)wo);
}

}  // namespace cgef
