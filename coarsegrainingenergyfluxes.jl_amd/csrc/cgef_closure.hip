// Learned closure for τ_ℓ: a 4-layer 5×5 convolutional network, trained and run on the device with
// hand-written kernels (no reference counterpart: the reference computes τ only by the exact filter).
//
// Storage: activations and weights BF16, accumulation FP32, optimizer state and master weights FP32.
// Tensors are (n, channels, rows, cols) with cols fastest. A hidden layer's epilogue stores its
// pre-activation Z and its activation A = GELU(Z), both BF16: GELU is evaluated once per element,
// the next layer and the weight gradient read A, and the data gradient reads Z for GELU'.
//
//   forward        Z_l = conv5_valid(A_{l-1}, W_l) + b_l, A_l = GELU(Z_l)   k_conv5(_mfma), pad 0
//   backward data  dZ_{l-1} = conv5_full(dZ_l, flip(W_l)ᵀ) · GELU'(Z_{l-1})  k_conv5, pad 4
//   backward W     dW_l[o][i][k] = Σ_{n,p} dZ_l[n][o][p] · A_{l-1}[n][i][p+k]  k_conv5_wgrad
//
// Weights reach the convolution in two BF16 layouts written from the FP32 master copy:
// forward [i][k][o] and backward-data [o][24−k][i] (output channel fastest in both, padded to a
// multiple of the output-channel block), so each tap's weights for a block of output channels are
// one contiguous wave-uniform read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <type_traits>

#include "cgef_internal.hpp"

namespace cgef {
namespace {

constexpr int CT = 16;       // output tile edge (16×16 outputs per 256-thread block)
constexpr int CS = CT + 4;   // staged input tile edge (5×5 stencil halo)
constexpr int ICC = 8;       // input channels staged per LDS pass in k_conv5
constexpr int WG = 16;       // output / input channels per k_conv5_wgrad block (16 × 16 lanes)
constexpr int ASTR = CS * (CS + 1) + 1;  // LDS channel stride of the wgrad input tile (odd: 16 channels on distinct banks)
constexpr int DSTR = CT * CT + 1;        // LDS channel stride of the wgrad dZ tile

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// FP32 → BF16, round to nearest even: one v_cvt_pk_bf16_f32 per pair.
__device__ __forceinline__ uint32_t pack_bf2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t{a, b}), bf16x2_t));
}
__device__ __forceinline__ uint16_t f2bf(float f) { return uint16_t(pack_bf2(f, 0.0f)); }

// erf for the GELU: Abramowitz & Stegun 7.1.26 on |x| (absolute error at most 1.5e-7, below the
// BF16 activations' resolution and the FP32 gradients' tolerance), extended as an odd function;
// also returns exp(−x²), which the GELU derivative reuses.
__device__ __forceinline__ float erf_as(float x, float &ex2) {
    const float ax = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * ax);
    const float poly =
        t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
    ex2 = __expf(-ax * ax);
    return copysignf(1.0f - poly * ex2, x);
}

// GELU(x) = x·Φ(x) with the normal CDF Φ(x) = (1 + erf(x/√2)) / 2, and its derivative Φ(x) + x·φ(x).
__device__ __forceinline__ float gelu(float x) {
    float e;
    return 0.5f * x * (1.0f + erf_as(x * 0.7071067811865476f, e));
}
__device__ __forceinline__ float gelu_grad(float x) {
    float e;   // exp(−x²/2)
    const float cdf = 0.5f * (1.0f + erf_as(x * 0.7071067811865476f, e));
    return cdf + x * 0.3989422804014327f * e;
}

template <int T>
__device__ __forceinline__ float load_as_float(const void *p, int64_t i) {
    if constexpr (T == CGEF_BF16) return bf2f(static_cast<const uint16_t *>(p)[i]);
    else if constexpr (T == CGEF_F32) return static_cast<const float *>(p)[i];
    else return float(static_cast<const double *>(p)[i]);
}

// Work item of a workgroup in a 1-D grid whose length is a multiple of 8. Workgroups are dispatched
// round-robin over the 8 XCDs, each with its own L2; XCD k takes the k-th contiguous range of the
// items, so items that are close in the item order (and read the same input tiles, halo rows or
// gradient tiles) share one L2.
__device__ __forceinline__ bool xcd_item(int64_t total, int64_t &item) {
    const int64_t chunk = gridDim.x / 8;
    item = int64_t(blockIdx.x % 8) * chunk + blockIdx.x / 8;
    return item < total;
}


// One 16×16 output tile × OCB output channels per block; lane = output pixel, OCB FP32 accumulators.
// Input channels pass through LDS ICC at a time (zero outside the input: pad 4 gives the full
// convolution used by the data gradient). OUT: 0 BF16 Z + bias, 1 FP32 + bias, 2 FP32 × GELU'(zg).
template <int OCB, int IN_T, int IN_GELU, int OUT>
__global__ __launch_bounds__(256) void k_conv5(const void *__restrict__ in, const uint16_t *__restrict__ w,
                                               const float *__restrict__ bias, const uint16_t *__restrict__ zg,
                                               void *__restrict__ out, uint16_t *__restrict__ act, int64_t cin,
                                               int64_t hin, int64_t win,
                                               int64_t cout, int64_t ocs, int64_t hout, int64_t wout, int pad,
                                               int tiles_x, int64_t ntiles, int nob, int64_t nimg) {
    constexpr int KUNROLL = OCB >= 64 ? 1 : 25;
    __shared__ float tile[ICC][CS][CS + 1];
    const int t = threadIdx.x, tx = t & (CT - 1), ty = t >> 4;
    int64_t item;   // (output-channel block, tile, image), channel block fastest: the blocks of one tile together
    if (!xcd_item(nob * ntiles * nimg, item)) return;
    const int64_t ob = item % nob, tl = (item / nob) % ntiles, img = item / (nob * ntiles);
    const int64_t ox0 = (tl % tiles_x) * CT, oy0 = (tl / tiles_x) * CT;
    const int64_t iy0 = oy0 - pad, ix0 = ox0 - pad;
    float acc[OCB];
#pragma unroll
    for (int o = 0; o < OCB; ++o) acc[o] = 0.0f;

    for (int64_t c0 = 0; c0 < cin; c0 += ICC) {
        const int nc = int(cin - c0 < ICC ? cin - c0 : ICC);
        __syncthreads();
        {
            const int64_t plane = hin * win;
            for (int r = t; r < CS * CS; r += 256) {
                const int yy = r / CS, xx = r - yy * CS;
                const int64_t iy = iy0 + yy, ix = ix0 + xx;
                const bool ok = iy >= 0 && iy < hin && ix >= 0 && ix < win;
                const int64_t off = ok ? (img * cin + c0) * plane + iy * win + ix : 0;
                float v[ICC];
#pragma unroll
                for (int c = 0; c < ICC; ++c) {
                    v[c] = (ok && c < nc) ? load_as_float<IN_T>(in, off + c * plane) : 0.0f;
                    if constexpr (IN_GELU) v[c] = gelu(v[c]);
                }
#pragma unroll
                for (int c = 0; c < ICC; ++c) tile[c][yy][xx] = v[c];
            }
        }
        __syncthreads();
        for (int c = 0; c < nc; ++c) {
            const uint32_t *wc = reinterpret_cast<const uint32_t *>(w + ((c0 + c) * 25) * ocs + ob * OCB);
            // a tap's OCB weights are wave-uniform scalar loads; with 64 of them per tap the taps are
            // not unrolled, so the scalar registers hold one tap's weights at a time
#pragma unroll KUNROLL
            for (int k = 0; k < 25; ++k) {
                const float x = tile[c][ty + k / 5][tx + k % 5];
                const uint32_t *wk = wc + k * (ocs >> 1);
#pragma unroll
                for (int o = 0; o < OCB; o += 2) {
                    const uint32_t pr = wk[o >> 1];
                    acc[o] = fmaf(x, __uint_as_float(pr << 16), acc[o]);
                    acc[o + 1] = fmaf(x, __uint_as_float(pr & 0xFFFF0000u), acc[o + 1]);
                }
            }
        }
    }

    const int64_t oy = oy0 + ty, ox = ox0 + tx;
    if (oy >= hout || ox >= wout) return;
#pragma unroll
    for (int o = 0; o < OCB; ++o) {
        const int64_t co = ob * OCB + o;
        if (co >= cout) break;
        const int64_t idx = ((img * cout + co) * hout + oy) * wout + ox;
        if constexpr (OUT == 0) {
            const uint16_t z = f2bf(acc[o] + bias[co]);
            static_cast<uint16_t *>(out)[idx] = z;
            if (act) act[idx] = f2bf(gelu(bf2f(z)));
        } else if constexpr (OUT == 1) {
            static_cast<float *>(out)[idx] = acc[o] + bias[co];
        } else {
            static_cast<float *>(out)[idx] = acc[o] * gelu_grad(bf2f(zg[idx]));
        }
    }
}

// Layers with at most 4 output channels (the 64 → 3 output layer): one 8×8 output tile per block, the
// input channels split over the four waves (wave w takes channels ≡ w mod 4 of each staged group of
// 16), lane = output pixel, 4 FP32 accumulators; the waves' partial sums are added in LDS in wave
// order (deterministic) and wave 0 writes the outputs as k_conv5 does. Four times the blocks of a
// 16×16 tile and a quarter of the channel loop per lane.
constexpr int NT8 = 8;                  // output tile edge
constexpr int NS8 = NT8 + 4;            // staged input tile edge
constexpr int NCG = 16;                 // input channels staged per pass (4 per wave)
template <int IN_T, int IN_GELU, int OUT>
__global__ __launch_bounds__(256) void k_conv5_narrow(const void *__restrict__ in, const uint16_t *__restrict__ w,
                                                      const float *__restrict__ bias, const uint16_t *__restrict__ zg,
                                                      void *__restrict__ out, uint16_t *__restrict__ act, int64_t cin,
                                                      int64_t hin, int64_t win, int64_t cout, int64_t ocs, int64_t hout,
                                                      int64_t wout, int pad, int tiles_x, int64_t ntiles,
                                                      int64_t nimg) {
    __shared__ float tile[NCG][NS8][NS8 + 1];
    __shared__ float red[4][64][4];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63, tx = lane & (NT8 - 1), ty = lane >> 3;
    int64_t item;   // (tile, image), tile fastest
    if (!xcd_item(ntiles * nimg, item)) return;
    const int64_t tl = item % ntiles, img = item / ntiles;
    const int64_t ox0 = (tl % tiles_x) * NT8, oy0 = (tl / tiles_x) * NT8;
    const int64_t iy0 = oy0 - pad, ix0 = ox0 - pad;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const int64_t plane = hin * win;
    for (int64_t c0 = 0; c0 < cin; c0 += NCG) {
        const int nc = int(cin - c0 < NCG ? cin - c0 : NCG);
        __syncthreads();
        for (int e = t; e < NCG * NS8 * NS8; e += 256) {
            const int c = e / (NS8 * NS8), r = e - c * (NS8 * NS8), yy = r / NS8, xx = r - yy * NS8;
            const int64_t iy = iy0 + yy, ix = ix0 + xx;
            float v = 0.0f;
            if (c < nc && iy >= 0 && iy < hin && ix >= 0 && ix < win) {
                v = load_as_float<IN_T>(in, (img * cin + c0 + c) * plane + iy * win + ix);
                if constexpr (IN_GELU) v = gelu(v);
            }
            tile[c][yy][xx] = v;
        }
        __syncthreads();
        for (int cc = wave; cc < nc; cc += 4) {
            const uint32_t *wc = reinterpret_cast<const uint32_t *>(w + ((c0 + cc) * 25) * ocs);
#pragma unroll
            for (int k = 0; k < 25; ++k) {
                const float x = tile[cc][ty + k / 5][tx + k % 5];
                const uint32_t *wk = wc + k * (ocs >> 1);
                const uint32_t p0 = wk[0], p1 = wk[1];
                acc[0] = fmaf(x, __uint_as_float(p0 << 16), acc[0]);
                acc[1] = fmaf(x, __uint_as_float(p0 & 0xFFFF0000u), acc[1]);
                acc[2] = fmaf(x, __uint_as_float(p1 << 16), acc[2]);
                acc[3] = fmaf(x, __uint_as_float(p1 & 0xFFFF0000u), acc[3]);
            }
        }
    }
#pragma unroll
    for (int o = 0; o < 4; ++o) red[wave][lane][o] = acc[o];
    __syncthreads();
    if (wave != 0) return;
    const int64_t oy = oy0 + ty, ox = ox0 + tx;
    if (oy >= hout || ox >= wout) return;
#pragma unroll
    for (int o = 0; o < 4; ++o) {
        const int64_t co = o;
        if (co >= cout) break;
        const float sum = ((red[0][lane][o] + red[1][lane][o]) + red[2][lane][o]) + red[3][lane][o];
        const int64_t idx = ((img * cout + co) * hout + oy) * wout + ox;
        if constexpr (OUT == 0) {
            const uint16_t z = f2bf(sum + bias[co]);
            static_cast<uint16_t *>(out)[idx] = z;
            if (act) act[idx] = f2bf(gelu(bf2f(z)));
        } else if constexpr (OUT == 1) {
            static_cast<float *>(out)[idx] = sum + bias[co];
        } else {
            static_cast<float *>(out)[idx] = sum * gelu_grad(bf2f(zg[idx]));
        }
    }
}

// The 64 → 64-channel layers on the matrix cores: an implicit GEMM, output pixels × output channels
// over K = 25 taps × 64 input channels, with v_mfma_f32_16x16x32_bf16. The block's 20×20 input tile
// is staged in LDS channel-fastest ([pixel][channel], pixel stride MPS), so the 8 consecutive K values
// a lane supplies (8 channels of one tap at one pixel) are one 16-byte LDS read; weights are laid out
// [tap][out][in] for the same reason, and each tap's 64 × 64 block passes through a double-buffered
// LDS slab (the next tap's block is loaded into registers while the current one is multiplied). Wave w computes output rows 4w..4w+3 of the 16×16 tile (four
// 16-pixel M tiles) × all 64 output channels (four N tiles): 16 accumulators of 4 floats per lane.
// Lane layout of the 16×16 MFMA: A row = lane % 16, B column = lane % 16, both supplying the K slice
// of lane group lane / 16 (the same slice for A and B, so the product does not depend on which K
// values a group holds); result rows 4·(lane / 16) + i, column lane % 16.
constexpr int MCH = 64;
constexpr int MPS = MCH + 8;

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

template <int IN_T, int IN_GELU, int OUT, int CP, int NT, int EXACT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_conv5_mfma(const void *__restrict__ in, const uint16_t *__restrict__ wm,
                                                    const float *__restrict__ bias, const uint16_t *__restrict__ zg,
                                                    void *__restrict__ out, uint16_t *__restrict__ act, int64_t cin_arg,
                                                    int64_t cout_arg, int64_t hin, int64_t win, int64_t hout,
                                                    int64_t wout, int pad, int tiles_x, int64_t ntiles, int64_t nimg) {
    static_assert(CP == 32 || CP == 64, "input channels padded to one or two 32-wide K slices per tap");
    constexpr int KS = CP + 8;     // LDS pixel stride of the input tile and row stride of a weight slab
    constexpr int NO = 16 * NT;    // output channels, padded to whole 16-wide N tiles
    // EXACT: the channel counts are the padded ones (the 64 → 64 layers), known at compile time
    const int64_t cin = EXACT ? CP : cin_arg, cout = EXACT ? NO : cout_arg;
    constexpr int H = CP / 32;     // K slices per tap
    __shared__ __attribute__((aligned(16))) uint16_t xs[CS * CS * KS];
    __shared__ __attribute__((aligned(16))) uint16_t wsm[2][NO * KS];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63, g = lane >> 4, r16 = lane & 15;
    // Persistent over (image, tile) items. Workgroups go round-robin to the 8 XCDs (gridDim.x is a
    // multiple of 8): XCD k owns the k-th contiguous range of items and its workgroups step through
    // it together, so tiles that share input halo rows and output cache lines (the 16-pixel rows of
    // neighbouring tiles) meet in one L2.
    const int64_t total = ntiles * nimg, chunk = (total + 7) / 8, nb = gridDim.x / 8;
    const int64_t lo = int64_t(blockIdx.x % 8) * chunk, hi = lo + chunk < total ? lo + chunk : total;
    int64_t item = lo + blockIdx.x / 8;
    if (item >= hi) return;
    // a tap's NO × CP weights are NW16 16-byte pieces; lane t moves pieces t and t + 256 (where they
    // exist): piece c is output row c / (CP / 8), input channels 8·(c % (CP / 8))…
    constexpr int NW16 = NO * CP / 8;
    const bool w0ok = t < NW16, w1ok = t + 256 < NW16;
    const uint4 *wsrc0 = reinterpret_cast<const uint4 *>(wm) + t;
    const int wo0 = (t / (CP / 8)) * KS + (t % (CP / 8)) * 8;
    const int wo1 = ((t + 256) / (CP / 8)) * KS + ((t + 256) % (CP / 8)) * 8;
    uint4 wr0 = uint4{0u, 0u, 0u, 0u}, wr1 = uint4{0u, 0u, 0u, 0u};
    auto store_w = [&](int slab) {
        if (w0ok) *reinterpret_cast<uint4 *>(&wsm[slab][wo0]) = wr0;
        if (w1ok) *reinterpret_cast<uint4 *>(&wsm[slab][wo1]) = wr1;
    };
    float bv[NT];   // the bias of the lane's output channels
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) bv[nt] = OUT == 2 || nt * 16 + r16 >= cout ? 0.0f : bias[nt * 16 + r16];

    // Input staging, channel-fastest, channels cin … CP − 1 zero. With the input 4-element aligned
    // (pad is 0 or 4, so every tile origin is too, and each 4-element group of a row lies wholly
    // inside or wholly outside the input), work items are (channel pair, tile row, 4-pixel group),
    // group fastest, so consecutive lanes read consecutive 8-byte (BF16) or 16-byte (FP32) pieces of
    // one row; an item's two channels go to LDS as one 4-byte store per pixel. The first NPF items of
    // the next tile are requested into registers during the current tile's last taps; the rest at
    // the tile's start. Otherwise a lane takes (channel pair, row) items and loads element by element.
    using Raw = typename std::conditional<IN_T == CGEF_BF16, uint2, float4>::type;
    constexpr int NG = CS / 4, NITEM = (CP / 2) * CS * NG, NI = (NITEM + 255) / 256;
    // prefetched items: as many as the register budget of two waves per SIMD leaves
    constexpr int NPF = IN_GELU ? 0 : (IN_T == CGEF_BF16 ? (OUT == 2 ? 5 : NI) : (OUT == 2 ? 3 : 5));
    constexpr int NB = IN_GELU ? 1 : (IN_T == CGEF_BF16 ? (NI - NPF > 0 ? NI - NPF : 1) : 5);   // batch of the items left
    const int64_t plane = hin * win;
    const bool vec_ok = (reinterpret_cast<uintptr_t>(in) & (IN_T == CGEF_BF16 ? 7 : 15)) == 0 && (win & 3) == 0 &&
                        (pad & 3) == 0;
    auto store_px = [&](int yy, int xx, int cp, float v0, float v1) {
        if constexpr (IN_GELU) v0 = gelu(v0), v1 = gelu(v1);
        reinterpret_cast<uint32_t *>(xs + (yy * CS + xx) * KS)[cp] = pack_bf2(v0, v1);
    };
    // lane index for the item arithmetic, made opaque per tile (below) so that the compiler does not
    // hoist every item's offsets out of the tile loop into registers
    int tq = t;
    auto load_item = [&](int i, int64_t img, int64_t iy0, int64_t ix0, Raw(&r)[2]) {
        const int idx = tq + 256 * i;
        const int q = idx % NG, rest = idx / NG, cp = rest % (CP / 2), yy = rest / (CP / 2);
        const int64_t iy = iy0 + yy, ix = ix0 + 4 * q;
        const bool ok = i < NI && idx < NITEM && iy >= 0 && iy < hin && ix >= 0 && ix + 4 <= win;
        const Raw *src = reinterpret_cast<const Raw *>(static_cast<const char *>(in) +
                                                       ((img * cin + 2 * cp) * plane + iy * win + ix) *
                                                           (IN_T == CGEF_BF16 ? 2 : 4));
        r[0] = ok && 2 * cp < cin ? src[0] : Raw{};
        r[1] = ok && 2 * cp + 1 < cin ? src[plane / 4] : Raw{};
    };
    auto store_item = [&](int i, const Raw(&r)[2]) {
        const int idx = tq + 256 * i;
        if (i >= NI || idx >= NITEM) return;
        const int q = idx % NG, rest = idx / NG, cp = rest % (CP / 2), yy = rest / (CP / 2);
        if constexpr (IN_T == CGEF_BF16 && !IN_GELU) {   // BF16 in, BF16 out: the halves regrouped
            const uint2 w0 = r[0], w1 = r[1];
            uint32_t *d = reinterpret_cast<uint32_t *>(xs + (yy * CS + 4 * q) * KS) + cp;
            d[0] = (w0.x & 0xFFFFu) | (w1.x << 16);
            d[KS / 2] = (w0.x >> 16) | (w1.x & 0xFFFF0000u);
            d[KS] = (w0.y & 0xFFFFu) | (w1.y << 16);
            d[3 * KS / 2] = (w0.y >> 16) | (w1.y & 0xFFFF0000u);
        } else if constexpr (IN_T == CGEF_BF16) {
            const uint2 w0 = r[0], w1 = r[1];
            store_px(yy, 4 * q, cp, __uint_as_float(w0.x << 16), __uint_as_float(w1.x << 16));
            store_px(yy, 4 * q + 1, cp, __uint_as_float(w0.x & 0xFFFF0000u), __uint_as_float(w1.x & 0xFFFF0000u));
            store_px(yy, 4 * q + 2, cp, __uint_as_float(w0.y << 16), __uint_as_float(w1.y << 16));
            store_px(yy, 4 * q + 3, cp, __uint_as_float(w0.y & 0xFFFF0000u), __uint_as_float(w1.y & 0xFFFF0000u));
        } else {
            const float4 f0 = r[0], f1 = r[1];
            store_px(yy, 4 * q, cp, f0.x, f1.x);
            store_px(yy, 4 * q + 1, cp, f0.y, f1.y);
            store_px(yy, 4 * q + 2, cp, f0.z, f1.z);
            store_px(yy, 4 * q + 3, cp, f0.w, f1.w);
        }
    };
    auto item_origin = [&](int64_t it, int64_t &img, int64_t &oy0, int64_t &ox0) {
        img = it / ntiles;
        const int64_t tile = it - img * ntiles;
        ox0 = (tile % tiles_x) * CT, oy0 = (tile / tiles_x) * CT;
    };
    Raw pf[NPF > 0 ? NPF : 1][2];
    int64_t img, oy0, ox0;
    item_origin(item, img, oy0, ox0);
    if (vec_ok) {
#pragma unroll
        for (int i = 0; i < NPF; ++i) load_item(i, img, oy0 - pad, ox0 - pad, pf[i]);
    }

    const uint16_t *xw = xs + (wave * 4 * CS + r16) * KS + g * 8;
    const uint16_t *wf = &wsm[0][r16 * KS + g * 8];
    auto frag_a = [&](bf16x8_t(&a)[4], int tap, int half) {
        const int ky = tap / 5, kx = tap - 5 * ky;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
            a[mt] = *reinterpret_cast<const bf16x8_t *>(xw + ((mt + ky) * CS + kx) * KS + half * 32);
    };
    auto frag_b = [&](bf16x8_t(&b)[NT], int tap, int half) {
        const uint16_t *wt = wf + (tap & 1) * (NO * KS);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) b[nt] = *reinterpret_cast<const bf16x8_t *>(wt + nt * 16 * KS + half * 32);
    };
    f32x4_t acc[4][NT];
    auto products = [&](const bf16x8_t(&a)[4], const bf16x8_t(&b)[NT]) {
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
    };

    for (;;) {
        const int64_t iy0 = oy0 - pad, ix0 = ox0 - pad;
        // Opaque copy of the weight pointer per tile: otherwise the per-tap weight addresses, being
        // the same for every tile, are hoisted out of this loop and held in ~100 registers.
        const uint4 *wsrc = wsrc0;
        asm volatile("" : "+v"(wsrc));
        tq = t;
        asm volatile("" : "+v"(tq));
        __syncthreads();   // the previous tile's products have read xs and both weight slabs
        if (vec_ok) {
#pragma unroll
            for (int i = 0; i < NPF; ++i) store_item(i, pf[i]);
#pragma unroll 1
            for (int i0 = NPF; i0 < NI; i0 += NB) {
                Raw r[NB > 0 ? NB : 1][2];
#pragma unroll
                for (int b = 0; b < NB; ++b) load_item(i0 + b, img, iy0, ix0, r[b]);
#pragma unroll
                for (int b = 0; b < NB; ++b) store_item(i0 + b, r[b]);
            }
        } else {
            const int cp = t % (CP / 2);
            for (int yy = t / (CP / 2); yy < CS; yy += 256 / (CP / 2)) {
                const int64_t iy = iy0 + yy;
                const bool row_in = iy >= 0 && iy < hin;
                const int64_t base = (img * cin + 2 * cp) * plane + iy * win + ix0;
                const bool c0 = 2 * cp < cin, c1 = 2 * cp + 1 < cin;
#pragma unroll 4
                for (int xx = 0; xx < CS; ++xx) {
                    const int64_t ix = ix0 + xx;
                    const bool ok = row_in && ix >= 0 && ix < win;
                    store_px(yy, xx, cp, ok && c0 ? load_as_float<IN_T>(in, base + xx) : 0.0f,
                             ok && c1 ? load_as_float<IN_T>(in, base + plane + xx) : 0.0f);
                }
            }
        }
        if (w0ok) wr0 = wsrc[0];
        if (w1ok) wr1 = wsrc[256];
        store_w(0);
        if (w0ok) wr0 = wsrc[NW16];   // tap 1's weights, stored to LDS during tap 0
        if (w1ok) wr1 = wsrc[NW16 + 256];
        __syncthreads();
        const int64_t next = item + nb;
        int64_t nimg_, noy0, nox0;
        item_origin(next < hi ? next : item, nimg_, noy0, nox0);

#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
        // Software-pipelined over the (tap, K slice) steps: the fragments of the next step are
        // requested from LDS before the current step's products. Tap t + 1's weights are stored to
        // LDS slab (t + 1) & 1 after tap t's first products — that slab was last read before the
        // barrier that ended tap t − 1 — and the barrier ending tap t publishes them; only the first
        // weight fragments of each tap are requested after a barrier. Tap t + 2's weights are
        // requested after that barrier; the next tile's input after tap 24's weights are stored.
        bf16x8_t a0[4], b0[NT], a1[4], b1[NT];
        frag_a(a0, 0, 0);
        frag_b(b0, 0, 0);
        for (int tap = 0; tap < 25; ++tap) {
            const bool more = tap + 1 < 25;
            if constexpr (H == 2) {
                frag_a(a1, tap, 1);
                frag_b(b1, tap, 1);
            } else if (more) {
                frag_a(a1, tap + 1, 0);
            }
            __builtin_amdgcn_sched_barrier(0);   // the requests above stay ahead of the products below
            products(a0, b0);
            __builtin_amdgcn_sched_barrier(0);   // the weight stores wait for their loads: after the products
            if (more) {
                store_w((tap + 1) & 1);
                if constexpr (H == 2) frag_a(a0, tap + 1, 0);
            }
            if (tap == 23 && vec_ok && next < hi) {
                tq = t;   // opaque again: the item addresses are formed here, not held through the taps
                asm volatile("" : "+v"(tq));
#pragma unroll
                for (int i = 0; i < NPF; ++i) load_item(i, nimg_, noy0 - pad, nox0 - pad, pf[i]);
            }
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (H == 2) products(a1, b1);
            if (more) {
                __syncthreads();
                frag_b(b0, tap + 1, 0);
                if constexpr (H == 1) {
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt) a0[mt] = a1[mt];
                }
                if (tap + 2 < 25) {
                    if (w0ok) wr0 = wsrc[(tap + 2) * NW16];
                    if (w1ok) wr1 = wsrc[(tap + 2) * NW16 + 256];
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }

        // Epilogue: lane (g, r16) holds pixels 4g … 4g + 3 of output row py for channel nt·16 + r16, i.e.
        // four consecutive outputs: one 8-byte (BF16) or 16-byte (FP32) store each when the row length
        // keeps them aligned and inside the output, element by element otherwise.
        const bool vec_out = (wout & 3) == 0 && ((reinterpret_cast<uintptr_t>(out) | reinterpret_cast<uintptr_t>(act) |
                                                  reinterpret_cast<uintptr_t>(zg)) & 15) == 0;
        const int64_t px0 = ox0 + 4 * g;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int64_t py = oy0 + wave * 4 + mt;
            if (py >= hout) continue;
            float gz[NT][4];   // OUT 2: GELU'(Z) of the four pixels, requested for all channels first
            if constexpr (OUT == 2) {
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const int64_t co = nt * 16 + r16;
                    const int64_t idx0 = ((img * cout + co) * hout + py) * wout + px0;
                    if (co >= cout) {
#pragma unroll
                        for (int i = 0; i < 4; ++i) gz[nt][i] = 0.0f;
                    } else if (vec_out && px0 + 4 <= wout) {
                        const uint2 z4 = *reinterpret_cast<const uint2 *>(zg + idx0);
                        gz[nt][0] = __uint_as_float(z4.x << 16), gz[nt][1] = __uint_as_float(z4.x & 0xFFFF0000u);
                        gz[nt][2] = __uint_as_float(z4.y << 16), gz[nt][3] = __uint_as_float(z4.y & 0xFFFF0000u);
                    } else {
#pragma unroll
                        for (int i = 0; i < 4; ++i) gz[nt][i] = px0 + i < wout ? bf2f(zg[idx0 + i]) : 0.0f;
                    }
                }
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) gz[nt][i] = gelu_grad(gz[nt][i]);
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int64_t co = nt * 16 + r16;
                if (co >= cout) continue;
                const int64_t idx0 = ((img * cout + co) * hout + py) * wout + px0;
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if constexpr (OUT == 2) v[i] = acc[mt][nt][i] * gz[nt][i];
                    else v[i] = acc[mt][nt][i] + bv[nt];
                }
                if (vec_out && px0 + 4 <= wout) {
                    if constexpr (OUT == 0) {
                        uint16_t z[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i) z[i] = f2bf(v[i]);
                        *reinterpret_cast<uint2 *>(static_cast<uint16_t *>(out) + idx0) =
                            uint2{uint32_t(z[0]) | (uint32_t(z[1]) << 16), uint32_t(z[2]) | (uint32_t(z[3]) << 16)};
                        if (act)
                            *reinterpret_cast<uint2 *>(act + idx0) =
                                uint2{pack_bf2(gelu(bf2f(z[0])), gelu(bf2f(z[1]))), pack_bf2(gelu(bf2f(z[2])), gelu(bf2f(z[3])))};
                    } else {
                        *reinterpret_cast<float4 *>(static_cast<float *>(out) + idx0) = float4{v[0], v[1], v[2], v[3]};
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        if (px0 + i >= wout) continue;
                        if constexpr (OUT == 0) {
                            const uint16_t z = f2bf(v[i]);
                            static_cast<uint16_t *>(out)[idx0 + i] = z;
                            if (act) act[idx0 + i] = f2bf(gelu(bf2f(z)));
                        } else static_cast<float *>(out)[idx0 + i] = v[i];
                    }
                }
                __builtin_amdgcn_sched_barrier(0);   // one (row, channel group) at a time: bounded temporaries
            }
        }
        item = next;
        if (item >= hi) break;
        img = nimg_, oy0 = noy0, ox0 = nox0;
    }
}

// BF16 matrix-core layouts of a 64 × 64 × 25 master weight tensor: forward wmf[k][o][i] and
// backward-data wmb[24 − k][i][o].
__global__ void k_pack_mfma(const float *__restrict__ p, int64_t cout, int64_t cin, uint16_t *wmf, int fo, int fi,
                            uint16_t *wmb, int bo, int bi) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (wmf && idx < 25 * fo * fi) {   // [k][o][i], zero beyond cout × cin
        const int64_t k = idx / (fo * fi), o = (idx / fi) % fo, i = idx % fi;
        wmf[idx] = o < cout && i < cin ? f2bf(p[(o * cin + i) * 25 + k]) : 0;
    }
    if (wmb && idx < 25 * bo * bi) {   // [24 − k][i][o]: the data gradient's input is dZ (cout channels)
        const int64_t kk = idx / (bo * bi), i = (idx / bi) % bo, o = idx % bi;
        wmb[idx] = o < cout && i < cin ? f2bf(p[(o * cin + i) * 25 + (24 - kk)]) : 0;
    }
}

// Weight and bias gradients of a valid 5×5 convolution. Block = OG output × IG input channels (one
// (o, i) pair per lane, its 25 taps in registers; OG·IG = 256: 16 × 16, 64 × 4 for an input of at
// most 4 channels so that no lane idles, 8 × 32 for at most 8 output channels) over a range of 16×16 output tiles of one image; each tile
// stages dZ (OG channels × 256 pixels) and the input (IG channels × 20×20) in LDS, and a lane walks
// each output row with a 5×5 register window of the input, 6 LDS reads per 25 FMAs. The block's sums
// go to its own slot of `part` (deterministic; k_wgrad_reduce adds the slots).
template <int OG, int IG, int IN_T, int IN_GELU>
__global__ __launch_bounds__(256) void k_conv5_wgrad(const float *__restrict__ dz, const void *__restrict__ in,
                                                     float *__restrict__ part, int64_t cin, int64_t cout,
                                                     int64_t hout, int64_t wout, int tiles_x, int ntiles,
                                                     int tiles_per_block, int nig, int nog, int64_t nimg, int nsplit) {
    static_assert(OG * IG == 256, "one (o, i) pair per lane");
    __shared__ float ds[OG * DSTR];
    __shared__ float as[IG * ASTR];
    const int t = threadIdx.x, ol = t / IG, il = t % IG;
    int64_t item;   // (channel-pair block, image, tile range), channel blocks fastest
    const int64_t nblk = int64_t(nog) * nig;
    if (!xcd_item(nblk * nimg * nsplit, item)) return;
    const int64_t og = (item % nblk) / nig, ig = (item % nblk) % nig;
    const int64_t img = (item / nblk) % nimg;
    const int split = int(item / (nblk * nimg));
    const int64_t hin = hout + 4, win = wout + 4;
    float acc[25];
#pragma unroll
    for (int k = 0; k < 25; ++k) acc[k] = 0.0f;
    float dsum = 0.0f;

    const int tile_lo = split * tiles_per_block;
    const int tile_hi = min(ntiles, tile_lo + tiles_per_block);
    for (int tl = tile_lo; tl < tile_hi; ++tl) {
        const int64_t ox0 = int64_t(tl % tiles_x) * CT, oy0 = int64_t(tl / tiles_x) * CT;
        __syncthreads();
        {
            // dZ: lane t owns output pixel t of the tile, 16 channel planes
            const int64_t oplane = hout * wout;
            {
                const int64_t oy = oy0 + (t >> 4), ox = ox0 + (t & 15);
                const bool ok = oy < hout && ox < wout;
                const float *src = dz + (img * cout + og * OG) * oplane + (ok ? oy * wout + ox : 0);
                // 16 planes at a time (their loads in flight together), so 64 output channels do not hold
                // 64 registers
                constexpr int CH = OG < 16 ? OG : 16;
#pragma unroll 1
                for (int c0 = 0; c0 < OG; c0 += CH) {
                    float v[CH];
#pragma unroll
                    for (int c = 0; c < CH; ++c)
                        v[c] = (ok && og * OG + c0 + c < cout) ? src[(c0 + c) * oplane] : 0.0f;
#pragma unroll
                    for (int c = 0; c < CH; ++c) ds[(c0 + c) * DSTR + t] = v[c];
                }
            }
            // input: lane owns pixels r = t, t + 256 of the 20×20 tile, IG channel planes
            const int64_t iplane = hin * win;
            for (int r = t; r < CS * CS; r += 256) {
                const int yy = r / CS, xx = r - yy * CS;
                const int64_t iy = oy0 + yy, ix = ox0 + xx;
                const bool ok = iy < hin && ix < win;
                const int64_t off = ok ? (img * cin + ig * IG) * iplane + iy * win + ix : 0;
                float v[IG];
#pragma unroll
                for (int c = 0; c < IG; ++c) {
                    v[c] = (ok && ig * IG + c < cin) ? load_as_float<IN_T>(in, off + c * iplane) : 0.0f;
                    if constexpr (IN_GELU) v[c] = gelu(v[c]);
                }
#pragma unroll
                for (int c = 0; c < IG; ++c) as[c * ASTR + yy * (CS + 1) + xx] = v[c];
            }
        }
        __syncthreads();
        const float *dsl = ds + ol * DSTR;
        const float *asl = as + il * ASTR;
        for (int py = 0; py < CT; ++py) {
            float wnd[5][5];
#pragma unroll
            for (int ky = 0; ky < 5; ++ky)
#pragma unroll
                for (int kx = 0; kx < 4; ++kx) wnd[ky][kx] = asl[(py + ky) * (CS + 1) + kx];
#pragma unroll
            for (int px = 0; px < CT; ++px) {
                const float d = dsl[py * CT + px];
                dsum += d;
#pragma unroll
                for (int ky = 0; ky < 5; ++ky) wnd[ky][4] = asl[(py + ky) * (CS + 1) + px + 4];
#pragma unroll
                for (int ky = 0; ky < 5; ++ky)
#pragma unroll
                    for (int kx = 0; kx < 5; ++kx) acc[ky * 5 + kx] = fmaf(d, wnd[ky][kx], acc[ky * 5 + kx]);
#pragma unroll
                for (int ky = 0; ky < 5; ++ky)
#pragma unroll
                    for (int kx = 0; kx < 4; ++kx) wnd[ky][kx] = wnd[ky][kx + 1];
            }
        }
    }

    // slot layout: [slot][cout·cin·25 weights | cout biases]
    const int64_t nw = cout * cin * 25;
    float *slot = part + (img * nsplit + split) * (nw + cout);
    const int64_t co = og * OG + ol, ci = ig * IG + il;
    if (co < cout && ci < cin) {
#pragma unroll
        for (int k = 0; k < 25; ++k) slot[(co * cin + ci) * 25 + k] = acc[k];
    }
    if (co < cout && il == 0 && ig == 0) slot[nw + co] = dsum;
}

// Weight and bias gradients of a 64 → 64-channel layer on the matrix cores. Per tap (ky, kx) the
// gradient is a 64 × 64 GEMM over pixels: dW[o][i] = Σ_p dZ[o][p] · A[i][p + (ky, kx)]. A block takes
// one kernel row ky (5 taps) and a range of 16 × 8-pixel output tiles of one image; wave w owns input
// channels 16w…16w+15 for all 64 output channels (4 M tiles) and the 5 taps: 20 accumulator tiles.
// K runs over the tile's pixels, 32 per MFMA: lane group g supplies 8 consecutive pixels of row
// g / 2, columns 8·(g % 2)…+7. dZ (staged BF16, [o][pixel]) gives aligned 16-byte fragments. The
// input ([i][row][col], staged with the GELU applied, rows 16-byte aligned) is read as the two aligned
// 16-byte halves of columns 8·(g % 2)…+15 once per step; the five taps' shifted fragments are
// selected from those eight dwords (odd shifts joined across dword pairs with v_alignbyte), so a step
// costs the lane 2 LDS reads for its 5 B fragments instead of 40 two-byte reads. Bias gradients come
// from ky = 0 blocks (wave 0 sums its dZ fragments). Sums go to the block's slot of `part`, tap-major
// ([tap][o][i], so a store instruction writes 16 consecutive floats per lane group) — k_wgrad_reduce
// turns them into [o][i][tap].
constexpr int WTY = 8;                 // output rows per wgrad tile (tile = WTY × 16 pixels)
constexpr int DZS = WTY * CT + 8;      // LDS stride of a dZ channel (BF16)
constexpr int AXP = 24;                // LDS row pitch of the staged input (BF16, 48 B; 20 used)
constexpr int AXC = WTY * AXP + 8;     // LDS stride of an input channel (400 B: the 16 channels a wave's
                                       // lanes read start on distinct 16-byte bank groups)
static_assert((MCH * WTY * 2) % 256 == 0 && (MCH * WTY * 5) % 256 == 0, "wgrad staging trip counts");
static_assert((AXP * 2) % 16 == 0 && (AXC * 2) % 16 == 0, "staged input rows 16-byte aligned");

template <int IN_T, int IN_GELU>
__global__ __launch_bounds__(256) void k_conv5_wgrad_mfma(const float *__restrict__ dz, const void *__restrict__ in,
                                                          float *__restrict__ part, int64_t hout, int64_t wout,
                                                          int tiles_x, int ntiles, int tiles_per_block, int64_t nimg,
                                                          int nsplit) {
    __shared__ __attribute__((aligned(16))) uint16_t dzs[MCH * DZS];
    __shared__ __attribute__((aligned(16))) uint16_t axs[MCH * AXC];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63, g = lane >> 4, r16 = lane & 15;
    // (kernel row, image, tile range) items, kernel row fastest: the five kernel-row blocks that stage
    // the same dZ tiles run together on one XCD and share them through its L2
    int64_t item;
    if (!xcd_item(5 * nimg * nsplit, item)) return;
    const int ky = int(item % 5);
    const int64_t img = (item / 5) % nimg;
    const int split = int(item / (5 * nimg));
    const int64_t hin = hout + 4, win = wout + 4;
    f32x4_t acc[5][4];
#pragma unroll
    for (int kx = 0; kx < 5; ++kx)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) acc[kx][mt] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const int tile_lo = split * tiles_per_block;
    const int tile_hi = min(ntiles, tile_lo + tiles_per_block);
    for (int tl = tile_lo; tl < tile_hi; ++tl) {
        const int64_t ox0 = int64_t(tl % tiles_x) * CT, oy0 = int64_t(tl / tiles_x) * WTY;
        __syncthreads();
        {
            // dZ: a lane owns (channel, row, 8-column half) segments and writes each as one 16-byte store
            const int64_t oplane = hout * wout;
#pragma unroll
            for (int it = 0; it < MCH * WTY * 2 / 256; ++it) {
                const int sgi = t + 256 * it;
                const int c = sgi / (WTY * 2), rr = (sgi / 2) % WTY, hf = sgi & 1;
                const int64_t oy = oy0 + rr, oxs = ox0 + 8 * hf;
                const float *src = dz + (img * MCH + c) * oplane + oy * wout + oxs;
                float v[8];
                if (oy < hout && oxs + 8 <= wout && (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
                    // the segment whole and 16-byte aligned: two vector loads
                    const float4 p0 = reinterpret_cast<const float4 *>(src)[0], p1 = reinterpret_cast<const float4 *>(src)[1];
                    v[0] = p0.x, v[1] = p0.y, v[2] = p0.z, v[3] = p0.w, v[4] = p1.x, v[5] = p1.y, v[6] = p1.z, v[7] = p1.w;
                } else {
#pragma unroll
                    for (int j = 0; j < 8; ++j) v[j] = (oy < hout && oxs + j < wout) ? src[j] : 0.0f;
                }
                *reinterpret_cast<uint4 *>(&dzs[c * DZS + rr * CT + 8 * hf]) =
                    uint4{pack_bf2(v[0], v[1]), pack_bf2(v[2], v[3]), pack_bf2(v[4], v[5]), pack_bf2(v[6], v[7])};
            }
            // input rows oy0 + ky …: (channel, row, 4-column quarter) segments, 8-byte stores
            const int64_t iplane = hin * win;
#pragma unroll 5
            for (int it = 0; it < MCH * WTY * 5 / 256; ++it) {
                const int sgi = t + 256 * it;
                const int c = sgi / (WTY * 5), rem = sgi - c * (WTY * 5), yy = rem / 5, q = rem - yy * 5;
                const int64_t iy = oy0 + yy + ky, ixs = ox0 + 4 * q;
                const int64_t base = (img * MCH + c) * iplane + iy * win + ixs;
                float v[4];
                if (IN_T == CGEF_BF16 && iy < hin && ixs + 4 <= win && (base & 3) == 0 &&
                    (reinterpret_cast<uintptr_t>(in) & 7) == 0) {
                    // four BF16 values, 8-byte aligned: one load
                    const uint2 q4 = *reinterpret_cast<const uint2 *>(static_cast<const uint16_t *>(in) + base);
                    v[0] = __uint_as_float(q4.x << 16), v[1] = __uint_as_float(q4.x & 0xFFFF0000u);
                    v[2] = __uint_as_float(q4.y << 16), v[3] = __uint_as_float(q4.y & 0xFFFF0000u);
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] = (iy < hin && ixs + j < win) ? load_as_float<IN_T>(in, base + j) : 0.0f;
                }
                if constexpr (IN_GELU) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] = gelu(v[j]);
                }
                *reinterpret_cast<uint2 *>(&axs[c * AXC + yy * AXP + 4 * q]) =
                    uint2{pack_bf2(v[0], v[1]), pack_bf2(v[2], v[3])};
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < WTY / 2; ++ks) {
            const int row = 2 * ks + (g >> 1), col = 8 * (g & 1);
            bf16x8_t a[4];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
                a[mt] = *reinterpret_cast<const bf16x8_t *>(&dzs[(mt * 16 + r16) * DZS + row * CT + col]);
            if (wave == 0 && ky == 0) {
#pragma unroll
                for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                    for (int j = 0; j < 8; ++j) bsum[mt] += float(a[mt][j]);
            }
            // columns col…col+15 of the lane's input row (only col…col+11 are ever selected)
            const uint16_t *arow = &axs[(wave * 16 + r16) * AXC + row * AXP + col];
            const uint4 lo = *reinterpret_cast<const uint4 *>(arow), hi = *reinterpret_cast<const uint4 *>(arow + 8);
            const uint32_t dw[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
            for (int kx = 0; kx < 5; ++kx) {
                // elements kx…kx+7: dwords kx/2… when kx is even, else the upper half of dword (kx−1)/2 + j
                // joined with the lower half of the next
                uint32_t f[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    f[j] = (kx & 1) ? __builtin_amdgcn_alignbyte(dw[(kx + 1) / 2 + j], dw[(kx - 1) / 2 + j], 2)
                                    : dw[kx / 2 + j];
                const bf16x8_t b = __builtin_bit_cast(bf16x8_t, uint4{f[0], f[1], f[2], f[3]});
#pragma unroll
                for (int mt = 0; mt < 4; ++mt)
                    acc[kx][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], b, acc[kx][mt], 0, 0, 0);
            }
        }
    }

    // D[m = 4g + i][n = r16]: output channel 16·mt + 4g + i, input channel 16·wave + r16; slot layout
    // [tap][o][i]
    const int64_t nw = int64_t(MCH) * MCH * 25;
    float *slot = part + (img * nsplit + split) * (nw + MCH);
    const int ci = wave * 16 + r16;
#pragma unroll
    for (int kx = 0; kx < 5; ++kx)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int co = mt * 16 + 4 * g + i;
                slot[(int64_t(ky * 5 + kx) * MCH + co) * MCH + ci] = acc[kx][mt][i];
            }
    if (wave == 0 && ky == 0) {
        // lanes r16 of the four groups hold partial sums of output channel 16·mt + r16
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            float v = bsum[mt];
            v += __shfl_xor(v, 16, 64);
            v += __shfl_xor(v, 32, 64);
            if (g == 0) slot[nw + mt * 16 + r16] = v;
        }
    }
}

// Adds the per-block slots: RDO outputs × RDP slot phases per block (phase p sums slots p, p + RDP, …),
// the phases combined in LDS in phase order (deterministic). The slots hold [o][i][tap] (VALU kernel)
// or, with tapmajor, [tap][o][i] (matrix-core kernel: cin input channels); dw is always [o][i][tap].
// 64 × 4 where the outputs alone give many blocks (coalesced 256-byte rows), 16 × 16 where they do not.
template <int RDO, int RDP>
__global__ __launch_bounds__(256) void k_wgrad_reduce(const float *__restrict__ part, int64_t nslots, int64_t nw,
                                                      int64_t cout, float *__restrict__ dw, float *__restrict__ db,
                                                      int tapmajor, int64_t cin) {
    __shared__ float red[RDP][RDO];
    const int o = threadIdx.x % RDO, ph = threadIdx.x / RDO;
    const int64_t i = int64_t(blockIdx.x) * RDO + o;
    const int64_t len = nw + cout;
    float s = 0.0f;
    if (i < len)
        for (int64_t p = ph; p < nslots; p += RDP) s += part[p * len + i];
    red[ph][o] = s;
    __syncthreads();
    if (ph == 0 && i < len) {
        s = 0.0f;
#pragma unroll
        for (int q = 0; q < RDP; ++q) s += red[q][o];
        if (i < nw) {
            if (tapmajor) {
                const int64_t tap = i / (cout * cin), oi = i - tap * (cout * cin);
                dw[oi * 25 + tap] = s;
            } else {
                dw[i] = s;
            }
        } else {
            db[i - nw] = s;
        }
    }
}

// BF16 copies of a master weight tensor W[o][i][k] (k = 5·ky + kx): forward [i][k][o] (stride ocs_f
// per (i, k)), backward-data [o][24 − k][i] (stride ocs_b per (o, k)).
__device__ __forceinline__ void write_layouts(float p, int64_t idx, int64_t cin, uint16_t *wf, int64_t ocs_f,
                                              uint16_t *wb, int64_t ocs_b) {
    const int64_t k = idx % 25, oi = idx / 25, i = oi % cin, o = oi / cin;
    const uint16_t h = f2bf(p);
    if (wf) wf[(i * 25 + k) * ocs_f + o] = h;
    if (wb) wb[(o * 25 + (24 - k)) * ocs_b + i] = h;
}

__global__ void k_pack_weights(const float *__restrict__ p, int64_t count, int64_t cin, uint16_t *wf, int64_t ocs_f,
                               uint16_t *wb, int64_t ocs_b) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx < count) write_layouts(p[idx], idx, cin, wf, ocs_f, wb, ocs_b);
}

// Adam (Kingma & Ba 2015, Algorithm 1) with bias correction; bc1 = 1 − β₁ᵗ, bc2 = 1 − β₂ᵗ.
__global__ void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, int64_t count, float lr, float b1, float b2, float eps,
                       float bc1, float bc2, int64_t cin, uint16_t *wf, int64_t ocs_f, uint16_t *wb,
                       int64_t ocs_b) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const float gi = g[idx];
    const float mi = b1 * m[idx] + (1.0f - b1) * gi;
    const float vi = b2 * v[idx] + (1.0f - b2) * gi * gi;
    m[idx] = mi;
    v[idx] = vi;
    const float pi = p[idx] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[idx] = pi;
    if (wf || wb) write_layouts(pi, idx, cin, wf, ocs_f, wb, ocs_b);
}

// Mean-squared error: dpred = 2(pred − target)/count; Σ(pred − target)² added to *loss_sum.
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ pred, const float *__restrict__ target,
                                             int64_t count, float inv_count, float *__restrict__ dpred,
                                             float *__restrict__ loss_sum) {
    __shared__ float red[4];
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float e = 0.0f;
    if (idx < count) {
        e = pred[idx] - target[idx];
        dpred[idx] = 2.0f * e * inv_count;
    }
    float s = e * e;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0 && loss_sum) atomicAdd(loss_sum, red[0] + red[1] + red[2] + red[3]);
}

// Loss with a Π term: L = mean_τ (pred − target)² + λ mean_p (Π(pred) − Π(target))², where at each
// pixel Π(τ) = −(a0 τxx + a1 τxy + a2 τyy) with a0 = g0·c0, a1 = (g1 + g2)·c1, a2 = g3·c2 from the
// normalized gradients g, so that Π is in normalized units:
// coef = (1/s_g0, 1/s_g1, 1/s_g2, 1/s_g3, c0, c1, c2): g_phys = g·coef[k], c = τ scale × Π scale.
// One lane per pixel (all three τ channels); dpred overwritten; loss[0] += Σ τ part, loss[1] += Σ Π part.
__global__ __launch_bounds__(256) void k_tau_pi_loss(const float *__restrict__ pred, const float *__restrict__ target,
                                                     const float *__restrict__ feat, const float *__restrict__ coef,
                                                     float lam, int64_t npix, int64_t total, float inv_tau,
                                                     float inv_pix, float *__restrict__ dpred,
                                                     float *__restrict__ loss) {
    __shared__ float red[2][4];
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float lt = 0.0f, lp = 0.0f;
    if (idx < total) {
        const int64_t img = idx / npix, p = idx - img * npix;
        const float *pr = pred + img * 3 * npix + p;
        const float *tg = target + img * 3 * npix + p;
        const float *g = feat + img * 4 * npix + p;
        const float a0 = g[0] * coef[0] * coef[4];
        const float a1 = (g[npix] * coef[1] + g[2 * npix] * coef[2]) * coef[5];
        const float a2 = g[3 * npix] * coef[3] * coef[6];
        const float e0 = pr[0] - tg[0], e1 = pr[npix] - tg[npix], e2 = pr[2 * npix] - tg[2 * npix];
        const float dpi = -(a0 * e0 + a1 * e1 + a2 * e2);
        lt = e0 * e0 + e1 * e1 + e2 * e2;
        lp = dpi * dpi;
        const float kp = -2.0f * lam * dpi * inv_pix, kt = 2.0f * inv_tau;
        float *dp = dpred + img * 3 * npix + p;
        dp[0] = kt * e0 + kp * a0;
        dp[npix] = kt * e1 + kp * a1;
        dp[2 * npix] = kt * e2 + kp * a2;
    }
    for (int off = 32; off > 0; off >>= 1) {
        lt += __shfl_down(lt, off, 64);
        lp += __shfl_down(lp, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = lt;
        red[1][threadIdx.x >> 6] = lp;
    }
    __syncthreads();
    if (threadIdx.x == 0 && loss) {
        atomicAdd(loss, red[0][0] + red[0][1] + red[0][2] + red[0][3]);
        atomicAdd(loss + 1, red[1][0] + red[1][1] + red[1][2] + red[1][3]);
    }
}

// Periodic crop: dst[i][c][y][x] = scale[c] · src[img_i][c][(y0_i + y) mod H][(x0_i + x) mod W].
// An origin whose image index is outside [0, nsrc) yields zeros.
template <int ST, int DT>
__global__ void k_crop(const void *__restrict__ src, int64_t nsrc, int64_t c, int64_t H, int64_t W,
                       const int32_t *__restrict__ org, int64_t h, int64_t w, const float *__restrict__ scale,
                       void *__restrict__ dst, int64_t total) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int64_t x = idx % w, y = (idx / w) % h, ch = (idx / (w * h)) % c, i = idx / (w * h * c);
    const int64_t img = org[3 * i];
    int64_t sy = (int64_t(org[3 * i + 1]) + y) % H, sx = (int64_t(org[3 * i + 2]) + x) % W;
    if (sy < 0) sy += H;
    if (sx < 0) sx += W;
    float val = 0.0f;
    if (img >= 0 && img < nsrc) val = load_as_float<ST>(src, ((img * c + ch) * H + sy) * W + sx);
    if (scale) val *= scale[ch];
    if constexpr (DT == CGEF_BF16) static_cast<uint16_t *>(dst)[idx] = f2bf(val);
    else static_cast<float *>(dst)[idx] = val;
}

// The five planes the exact τ needs, per slice: u, v, u², u·v, v² (FP64).
template <typename T>
__global__ void k_products(const T *__restrict__ u, const T *__restrict__ v, int64_t npts, int64_t total,
                           double *__restrict__ out) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int64_t s = idx / npts, p = idx - s * npts;
    const double a = double(u[idx]), b = double(v[idx]);
    double *o = out + s * 5 * npts + p;
    o[0] = a;
    o[npts] = b;
    o[2 * npts] = a * a;
    o[3 * npts] = a * b;
    o[4 * npts] = b * b;
}

// τ = F[q_a q_b] − q̄_a q̄_b (FP64, stored FP32; skipped when tau is null) and the four filtered-velocity
// gradients as FP32.
__global__ void k_targets(const double *__restrict__ filt, const double *__restrict__ grad, int64_t npts,
                          int64_t total, float *__restrict__ tau, float *__restrict__ feat) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int64_t s = idx / npts, p = idx - s * npts;
    if (tau) {
        const double *f = filt + s * 5 * npts + p;
        const double ub = f[0], vb = f[npts];
        float *tt = tau + s * 3 * npts + p;
        tt[0] = float(f[2 * npts] - ub * ub);
        tt[npts] = float(f[3 * npts] - ub * vb);
        tt[2 * npts] = float(f[4 * npts] - vb * vb);
    }
    const double *g = grad + s * 4 * npts + p;
    float *fo = feat + s * 4 * npts + p;
#pragma unroll
    for (int c = 0; c < 4; ++c) fo[c * npts] = float(g[c * npts]);
}

// Π = −(S̄xx τxx + 2 S̄xy τxy + S̄yy τyy) from the gradients (∂xū, ∂yū, ∂xv̄, ∂yv̄) and the stress
// τ = tau·tau_scale + σ² G Gᵀ (G the velocity-gradient matrix; σ² = 0: τ as given).
__global__ void k_closure_pi(const float *__restrict__ feat, const float *__restrict__ tau,
                             const float *__restrict__ tau_scale, float sigma2, int64_t npts, int64_t total,
                             float *__restrict__ Pi) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int64_t s = idx / npts, p = idx - s * npts;
    const float *g = feat + s * 4 * npts + p;
    const float *tt = tau + s * 3 * npts + p;
    const float s0 = tau_scale ? tau_scale[0] : 1.0f, s1 = tau_scale ? tau_scale[1] : 1.0f,
                s2 = tau_scale ? tau_scale[2] : 1.0f;
    const float ux = g[0], uy = g[npts], vx = g[2 * npts], vy = g[3 * npts];
    const float txx = tt[0] * s0 + sigma2 * (ux * ux + uy * uy);
    const float txy = tt[npts] * s1 + sigma2 * (ux * vx + uy * vy);
    const float tyy = tt[2 * npts] * s2 + sigma2 * (vx * vx + vy * vy);
    Pi[idx] = -(ux * txx + (uy + vx) * txy + vy * tyy);
}

// Gradient (nonlinear) model τ_ab = σ² Σ_k ∂_k ū_a ∂_k ū_b, the closure without learning, added to
// `base` when given (base = τ and σ² < 0 gives the residual the network learns).
__global__ void k_gradient_model(const float *__restrict__ feat, float sigma2, const float *__restrict__ base,
                                 int64_t npts, int64_t total, float *__restrict__ tau) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int64_t s = idx / npts, p = idx - s * npts;
    const float *g = feat + s * 4 * npts + p;
    const float ux = g[0], uy = g[npts], vx = g[2 * npts], vy = g[3 * npts];
    float *tt = tau + s * 3 * npts + p;
    const float *bb = base ? base + s * 3 * npts + p : nullptr;
    tt[0] = (bb ? bb[0] : 0.0f) + sigma2 * (ux * ux + uy * uy);
    tt[npts] = (bb ? bb[npts] : 0.0f) + sigma2 * (ux * vx + uy * vy);
    tt[2 * npts] = (bb ? bb[2 * npts] : 0.0f) + sigma2 * (vx * vx + vy * vy);
}

template <class F>
cgef_status guarded(F &&f) {
    try {
        f();
        return CGEF_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.status;
    } catch (const std::exception &e) {
        set_error(e.what());
        return CGEF_ERR_ARGUMENT;
    }
}

void need(bool ok, cgef_status s, const char *msg) {
    if (!ok) fail(s, msg);
}

unsigned blocks_for(int64_t n, int b = 256) { return unsigned((n + b - 1) / b); }

void check_launch() { CGEF_HIP(hipGetLastError()); }

// Matrix-core layer shapes: input channels padded to one or two 32-wide K slices per tap, output
// channels to one or four 16-wide N tiles.
int mfma_cin_pad(int64_t cin) { return cin <= 32 ? 32 : 64; }
int mfma_cout_pad(int64_t cout) { return cout <= 16 ? 16 : 64; }

// Workgroups of 256 threads of kernel fn that one CU holds at once (at least 1).
int resident_blocks(const void *fn) {
    int nb = 0;
    CGEF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, 0));
    return nb > 0 ? nb : 1;
}

// Compute units of the current device (cached per device).
int device_cus() {
    static int cus[64] = {};
    int dev = 0;
    CGEF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return 256;
    if (!cus[dev]) {
        int c = 0;
        CGEF_HIP(hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev));
        cus[dev] = c > 0 ? c : 256;
    }
    return cus[dev];
}

template <int OCB, int IN_T, int IN_GELU, int OUT>
void launch_conv(dim3 g, hipStream_t st, const void *in, const uint16_t *w, const float *bias, const uint16_t *zg,
                 void *out, uint16_t *act, int64_t cin, int64_t hin, int64_t win, int64_t cout, int64_t ocs,
                 int64_t hout, int64_t wout, int pad, int tiles_x) {
    // g carries (tiles, output-channel blocks, images); the kernel takes them as 1-D XCD-ranged items
    const int64_t items = int64_t(g.x) * g.y * g.z;
    hipLaunchKernelGGL((k_conv5<OCB, IN_T, IN_GELU, OUT>), dim3(unsigned((items + 7) / 8 * 8)), dim3(256), 0, st, in,
                       w, bias, zg, out, act, cin, hin, win, cout, ocs, hout, wout, pad, tiles_x, int64_t(g.x),
                       int(g.y), int64_t(g.z));
}

template <int OCB, int IN_T, int IN_GELU>
void conv_out(int out_kind, dim3 g, hipStream_t st, const void *in, const uint16_t *w, const float *bias,
              const uint16_t *zg, void *out, uint16_t *act, int64_t cin, int64_t hin, int64_t win, int64_t cout, int64_t ocs,
              int64_t hout, int64_t wout, int pad, int tiles_x) {
    if (out_kind == 0)
        launch_conv<OCB, IN_T, IN_GELU, 0>(g, st, in, w, bias, zg, out, act, cin, hin, win, cout, ocs, hout, wout, pad,
                                           tiles_x);
    else if (out_kind == 1)
        launch_conv<OCB, IN_T, IN_GELU, 1>(g, st, in, w, bias, zg, out, act, cin, hin, win, cout, ocs, hout, wout, pad,
                                           tiles_x);
    else
        launch_conv<OCB, IN_T, IN_GELU, 2>(g, st, in, w, bias, zg, out, act, cin, hin, win, cout, ocs, hout, wout, pad,
                                           tiles_x);
}

template <int OCB>
void conv_in(int in_t, int in_gelu, int out_kind, dim3 g, hipStream_t st, const void *in, const uint16_t *w,
             const float *bias, const uint16_t *zg, void *out, uint16_t *act, int64_t cin, int64_t hin, int64_t win, int64_t cout,
             int64_t ocs, int64_t hout, int64_t wout, int pad, int tiles_x) {
    if (in_t == CGEF_BF16 && in_gelu)
        conv_out<OCB, CGEF_BF16, 1>(out_kind, g, st, in, w, bias, zg, out, act, cin, hin, win, cout, ocs, hout, wout, pad,
                                    tiles_x);
    else if (in_t == CGEF_BF16)
        conv_out<OCB, CGEF_BF16, 0>(out_kind, g, st, in, w, bias, zg, out, act, cin, hin, win, cout, ocs, hout, wout, pad,
                                    tiles_x);
    else
        conv_out<OCB, CGEF_F32, 0>(out_kind, g, st, in, w, bias, zg, out, act, cin, hin, win, cout, ocs, hout, wout, pad,
                                   tiles_x);
}

// Split of the output tiles over blocks for the weight gradient: about 2048 blocks in flight (VALU:
// 16×16 tiles, channel groups of 16; matrix cores: 8×16 tiles, one block per kernel row, ~1024).
void wgrad_split(bool mfma, int64_t n, int64_t cin, int64_t cout, int64_t hout, int64_t wout, int &tiles_x,
                 int &ntiles, int &tpb, int &nsplit, int &nig, int &nog) {
    tiles_x = int((wout + CT - 1) / CT);
    ntiles = tiles_x * int((hout + (mfma ? WTY : CT) - 1) / (mfma ? WTY : CT));
    // k_conv5_wgrad's block shape (OG × IG): 64 × 4 for a ≤ 4-channel input, 8 × 32 for ≤ 8 output
    // channels (the 64 → 3 layer: 3 of 8 lane rows busy instead of 3 of 16), else 16 × 16
    const int ig = cin <= 4 ? 4 : (cout <= 8 ? 32 : WG), og = cin <= 4 ? 64 : (cout <= 8 ? 8 : WG);
    nig = int((cin + ig - 1) / ig);
    nog = int((cout + og - 1) / og);
    const int64_t base = mfma ? 5 * n : n * nig * nog;
    int64_t want = std::max<int64_t>(1, (mfma ? 1024 : 2048) / std::max<int64_t>(base, 1));
    want = std::min<int64_t>(want, ntiles);
    tpb = int((ntiles + want - 1) / want);
    nsplit = (ntiles + tpb - 1) / tpb;
}

}  // namespace
}  // namespace cgef

using namespace cgef;

extern "C" {

cgef_status cgef_conv5(const void *in, int32_t in_dtype, int32_t in_gelu, int64_t n, int64_t cin, int64_t hin,
                       int64_t win, const uint16_t *w, int64_t ocs, const float *bias, int64_t cout, int32_t pad,
                       void *out, int32_t out_dtype, const uint16_t *zgrad, uint16_t *act, void *stream) {
    return guarded([&] {
        need(in && w && out, CGEF_ERR_ARGUMENT, "cgef_conv5: null pointer");
        need(in_dtype == CGEF_BF16 || (in_dtype == CGEF_F32 && !in_gelu), CGEF_ERR_ARGUMENT,
             "cgef_conv5: input must be BF16 (optionally GELU-applied) or FP32");
        need(pad == 0 || pad == 4, CGEF_ERR_ARGUMENT, "cgef_conv5: pad must be 0 (valid) or 4 (full)");
        need(n >= 1 && n <= 65535 && cin >= 1 && cout >= 1, CGEF_ERR_DIMENSION, "cgef_conv5: bad n / channels");
        const int64_t hout = hin + 2 * pad - 4, wout = win + 2 * pad - 4;
        need(hin >= 1 && win >= 1 && hout >= 1 && wout >= 1, CGEF_ERR_DIMENSION, "cgef_conv5: input smaller than 5x5");
        int out_kind;
        if (zgrad) {
            need(out_dtype == CGEF_F32, CGEF_ERR_ARGUMENT, "cgef_conv5: GELU'-scaled output is FP32");
            out_kind = 2;
        } else {
            need(bias != nullptr, CGEF_ERR_ARGUMENT, "cgef_conv5: bias required without zgrad");
            need(out_dtype == CGEF_BF16 || out_dtype == CGEF_F32, CGEF_ERR_ARGUMENT, "cgef_conv5: output BF16 or FP32");
            out_kind = out_dtype == CGEF_BF16 ? 0 : 1;
        }
        need(!act || out_kind == 0, CGEF_ERR_ARGUMENT, "cgef_conv5: the activation is written with a BF16 output");
        // output channels per block: 64 lanes' worth of accumulators for wide layers; 16 where the input has
        // few channels (the 4 → 64 layer, the data gradient through the 64 → 3 layer), which gives four
        // times the blocks for a short channel loop
        const int ocb = cout <= 4 ? 4 : (cout <= 16 || cin <= 16 ? 16 : 64);
        need(ocs % ocb == 0 && ocs >= cout, CGEF_ERR_DIMENSION,
             "cgef_conv5: weight layout stride must cover cout in whole output-channel blocks (4, 16 or 64)");
        const int tiles_x = int((wout + CT - 1) / CT);
        const int64_t ntiles = int64_t(tiles_x) * ((hout + CT - 1) / CT);
        need(ntiles * ((cout + ocb - 1) / ocb) * n <= 0x7ffffff0, CGEF_ERR_DIMENSION, "cgef_conv5: too many tiles");
        dim3 g(unsigned(ntiles), unsigned((cout + ocb - 1) / ocb), unsigned(n));
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (ocb == 4 && ntiles * n < 2048) {
            // few 16×16 tiles (a training batch): 8×8 tiles, the input channels split over the waves
            // (k_conv5_narrow); large fields keep k_conv5's 16×16 tiles
            const int tx8 = int((wout + NT8 - 1) / NT8);
            const int64_t nt8 = int64_t(tx8) * ((hout + NT8 - 1) / NT8);
            need(nt8 <= 0x7fffffff, CGEF_ERR_DIMENSION, "cgef_conv5: too many tiles");
            need(nt8 * n <= 0x7ffffff0, CGEF_ERR_DIMENSION, "cgef_conv5: too many tiles");
            const dim3 g8(unsigned((nt8 * n + 7) / 8 * 8));
            auto go = [&](auto kern) {
                hipLaunchKernelGGL(kern, g8, dim3(256), 0, st, in, w, bias, zgrad, out, act, cin, hin, win, cout, ocs, hout,
                                   wout, pad, tx8, nt8, n);
            };
#define CGEF_NARROW(IT, IG)                                                  \
    do {                                                                     \
        if (out_kind == 0) go(k_conv5_narrow<IT, IG, 0>);                    \
        else if (out_kind == 1) go(k_conv5_narrow<IT, IG, 1>);               \
        else go(k_conv5_narrow<IT, IG, 2>);                                  \
    } while (0)
            if (in_dtype == CGEF_BF16 && in_gelu) CGEF_NARROW(CGEF_BF16, 1);
            else if (in_dtype == CGEF_BF16) CGEF_NARROW(CGEF_BF16, 0);
            else CGEF_NARROW(CGEF_F32, 0);
#undef CGEF_NARROW
        } else if (ocb == 4)
            conv_in<4>(in_dtype, in_gelu, out_kind, g, st, in, w, bias, zgrad, out, act, cin, hin, win, cout, ocs, hout,
                       wout, pad, tiles_x);
        else if (ocb == 16)
            conv_in<16>(in_dtype, in_gelu, out_kind, g, st, in, w, bias, zgrad, out, act, cin, hin, win, cout, ocs, hout,
                        wout, pad, tiles_x);
        else
            conv_in<64>(in_dtype, in_gelu, out_kind, g, st, in, w, bias, zgrad, out, act, cin, hin, win, cout, ocs, hout,
                        wout, pad, tiles_x);
        check_launch();
    });
}

cgef_status cgef_conv5_mfma_ex(const void *in, int32_t in_dtype, int32_t in_gelu, int64_t n, int64_t cin, int64_t hin,
                               int64_t win, const uint16_t *wm, const float *bias, int64_t cout, int32_t pad, void *out,
                               int32_t out_dtype, const uint16_t *zgrad, uint16_t *act, void *stream) {
    return guarded([&] {
        need(in && wm && out, CGEF_ERR_ARGUMENT, "cgef_conv5_mfma: null pointer");
        need(in_dtype == CGEF_BF16 || (in_dtype == CGEF_F32 && !in_gelu), CGEF_ERR_ARGUMENT,
             "cgef_conv5_mfma: input must be BF16 (optionally GELU-applied) or FP32");
        need(pad == 0 || pad == 4, CGEF_ERR_ARGUMENT, "cgef_conv5_mfma: pad must be 0 (valid) or 4 (full)");
        need(n >= 1 && n <= 65535, CGEF_ERR_DIMENSION, "cgef_conv5_mfma: bad n");
        need(cin >= 1 && cin <= MCH && cout >= 1 && cout <= MCH, CGEF_ERR_DIMENSION,
             "cgef_conv5_mfma: 1 to 64 input and output channels");
        const int cp = mfma_cin_pad(cin), no = mfma_cout_pad(cout);
        need(!in_gelu || (cp == 64 && no == 64), CGEF_ERR_UNSUPPORTED,
             "cgef_conv5_mfma: GELU on load only for 33-64 input and output channels");
        const int64_t hout = hin + 2 * pad - 4, wout = win + 2 * pad - 4;
        need(hin >= 1 && win >= 1 && hout >= 1 && wout >= 1, CGEF_ERR_DIMENSION,
             "cgef_conv5_mfma: input smaller than 5x5");
        int out_kind;
        if (zgrad) {
            need(out_dtype == CGEF_F32, CGEF_ERR_ARGUMENT, "cgef_conv5_mfma: GELU'-scaled output is FP32");
            out_kind = 2;
        } else {
            need(bias != nullptr, CGEF_ERR_ARGUMENT, "cgef_conv5_mfma: bias required without zgrad");
            need(out_dtype == CGEF_BF16 || out_dtype == CGEF_F32, CGEF_ERR_ARGUMENT,
                 "cgef_conv5_mfma: output BF16 or FP32");
            out_kind = out_dtype == CGEF_BF16 ? 0 : 1;
        }
        need(!act || out_kind == 0, CGEF_ERR_ARGUMENT, "cgef_conv5_mfma: the activation is written with a BF16 output");
        const int tiles_x = int((wout + CT - 1) / CT);
        const int64_t ntiles = int64_t(tiles_x) * ((hout + CT - 1) / CT);
        const int64_t items = ntiles * n;
        hipStream_t st = static_cast<hipStream_t>(stream);
        // persistent workgroups: as many per CU as the kernel's LDS and registers allow, in 8 equal XCD groups
        auto launch = [&](const void *fn, auto kern) {
            const int per_cu = resident_blocks(fn);
            const int64_t per_xcd = std::min<int64_t>((items + 7) / 8, int64_t(per_cu) * std::max(1, device_cus() / 8));
            hipLaunchKernelGGL(kern, dim3(unsigned(8 * per_xcd)), dim3(256), 0, st, in, wm, bias, zgrad, out, act, cin,
                               cout, hin, win, hout, wout, pad, tiles_x, ntiles, n);
        };
#define CGEF_MFMA_K(IT, IG, OK, CPV, NTV)                                                                            \
    do {                                                                                                             \
        if (CPV == 64 && NTV == 4 && cin == 64 && cout == 64)                                                        \
            launch((const void *)k_conv5_mfma<IT, IG, OK, CPV, NTV, (CPV == 64 && NTV == 4)>,                       \
                   k_conv5_mfma<IT, IG, OK, CPV, NTV, (CPV == 64 && NTV == 4)>);                                     \
        else                                                                                                         \
            launch((const void *)k_conv5_mfma<IT, IG, OK, CPV, NTV, 0>, k_conv5_mfma<IT, IG, OK, CPV, NTV, 0>);     \
    } while (0)
#define CGEF_MFMA_OUT(IT, IG, CPV, NTV)                                              \
    do {                                                                             \
        if (out_kind == 0) CGEF_MFMA_K(IT, IG, 0, CPV, NTV);                         \
        else if (out_kind == 1) CGEF_MFMA_K(IT, IG, 1, CPV, NTV);                    \
        else CGEF_MFMA_K(IT, IG, 2, CPV, NTV);                                       \
    } while (0)
#define CGEF_MFMA_IN(CPV, NTV)                                                       \
    do {                                                                             \
        if (in_dtype == CGEF_BF16) CGEF_MFMA_OUT(CGEF_BF16, 0, CPV, NTV);            \
        else CGEF_MFMA_OUT(CGEF_F32, 0, CPV, NTV);                                   \
    } while (0)
        if (cp == 64 && no == 64) {
            if (in_dtype == CGEF_BF16 && in_gelu) CGEF_MFMA_OUT(CGEF_BF16, 1, 64, 4);
            else CGEF_MFMA_IN(64, 4);
        } else if (cp == 64) CGEF_MFMA_IN(64, 1);
        else if (no == 64) CGEF_MFMA_IN(32, 4);
        else CGEF_MFMA_IN(32, 1);
#undef CGEF_MFMA_IN
#undef CGEF_MFMA_OUT
#undef CGEF_MFMA_K
        check_launch();
    });
}

cgef_status cgef_conv5_mfma(const void *in, int32_t in_dtype, int32_t in_gelu, int64_t n, int64_t hin, int64_t win,
                            const uint16_t *wm, const float *bias, int32_t pad, void *out, int32_t out_dtype,
                            const uint16_t *zgrad, uint16_t *act, void *stream) {
    return cgef_conv5_mfma_ex(in, in_dtype, in_gelu, n, MCH, hin, win, wm, bias, MCH, pad, out, out_dtype, zgrad, act,
                              stream);
}

cgef_status cgef_conv5_pack_mfma_ex(const float *w, int64_t cout, int64_t cin, uint16_t *wmf, uint16_t *wmb,
                                    void *stream) {
    return guarded([&] {
        need(w && (wmf || wmb), CGEF_ERR_ARGUMENT, "cgef_conv5_pack_mfma: null pointer");
        need(cin >= 1 && cin <= MCH && cout >= 1 && cout <= MCH, CGEF_ERR_DIMENSION,
             "cgef_conv5_pack_mfma: 1 to 64 input and output channels");
        const int fo = mfma_cout_pad(cout), fi = mfma_cin_pad(cin), bo = mfma_cout_pad(cin), bi = mfma_cin_pad(cout);
        const int64_t count = 25 * std::max(fo * fi, bo * bi);
        hipLaunchKernelGGL(k_pack_mfma, dim3(blocks_for(count)), dim3(256), 0, static_cast<hipStream_t>(stream), w, cout,
                           cin, wmf, fo, fi, wmb, bo, bi);
        check_launch();
    });
}

cgef_status cgef_conv5_pack_mfma(const float *w, uint16_t *wmf, uint16_t *wmb, void *stream) {
    return cgef_conv5_pack_mfma_ex(w, MCH, MCH, wmf, wmb, stream);
}

cgef_status cgef_conv5_wgrad_scratch(int64_t n, int64_t cin, int64_t cout, int64_t hout, int64_t wout, int32_t mfma,
                                     int64_t *floats) {
    return guarded([&] {
        need(floats != nullptr, CGEF_ERR_ARGUMENT, "null output");
        need(n >= 1 && cin >= 1 && cout >= 1 && hout >= 1 && wout >= 1, CGEF_ERR_DIMENSION, "bad sizes");
        need(!mfma || (cin == MCH && cout == MCH), CGEF_ERR_DIMENSION, "the matrix-core path takes 64 -> 64 layers");
        int tiles_x, ntiles, tpb, nsplit, nig, nog;
        wgrad_split(mfma != 0, n, cin, cout, hout, wout, tiles_x, ntiles, tpb, nsplit, nig, nog);
        *floats = n * nsplit * (cout * cin * 25 + cout);
    });
}

cgef_status cgef_conv5_wgrad(const float *dz, const void *in, int32_t in_dtype, int32_t in_gelu, int64_t n,
                             int64_t cin, int64_t cout, int64_t hout, int64_t wout, int32_t mfma, float *dw, float *db,
                             float *scratch, int64_t scratch_floats, void *stream) {
    return guarded([&] {
        need(dz && in && dw && db && scratch, CGEF_ERR_ARGUMENT, "cgef_conv5_wgrad: null pointer");
        need(in_dtype == CGEF_BF16 || (in_dtype == CGEF_F32 && !in_gelu), CGEF_ERR_ARGUMENT,
             "cgef_conv5_wgrad: input must be BF16 (optionally GELU-applied) or FP32");
        need(n >= 1 && n <= 65535 && cin >= 1 && cout >= 1 && hout >= 1 && wout >= 1, CGEF_ERR_DIMENSION,
             "cgef_conv5_wgrad: bad sizes");
        need(!mfma || (cin == MCH && cout == MCH), CGEF_ERR_DIMENSION,
             "cgef_conv5_wgrad: the matrix-core path takes 64 -> 64 layers");
        int tiles_x, ntiles, tpb, nsplit, nig, nog;
        wgrad_split(mfma != 0, n, cin, cout, hout, wout, tiles_x, ntiles, tpb, nsplit, nig, nog);
        const int64_t nw = cout * cin * 25, nslots = n * nsplit;
        need(scratch_floats >= nslots * (nw + cout), CGEF_ERR_DIMENSION,
             "cgef_conv5_wgrad: scratch smaller than cgef_conv5_wgrad_scratch");
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (mfma) {
            const int64_t nitems = 5 * n * nsplit;
            need(nitems <= 0x7ffffff0, CGEF_ERR_DIMENSION, "cgef_conv5_wgrad: too many blocks");
            dim3 g(unsigned((nitems + 7) / 8 * 8));
            if (in_dtype == CGEF_BF16 && in_gelu)
                hipLaunchKernelGGL((k_conv5_wgrad_mfma<CGEF_BF16, 1>), g, dim3(256), 0, st, dz, in, scratch, hout, wout,
                                   tiles_x, ntiles, tpb, n, nsplit);
            else if (in_dtype == CGEF_BF16)
                hipLaunchKernelGGL((k_conv5_wgrad_mfma<CGEF_BF16, 0>), g, dim3(256), 0, st, dz, in, scratch, hout, wout,
                                   tiles_x, ntiles, tpb, n, nsplit);
            else
                hipLaunchKernelGGL((k_conv5_wgrad_mfma<CGEF_F32, 0>), g, dim3(256), 0, st, dz, in, scratch, hout, wout,
                                   tiles_x, ntiles, tpb, n, nsplit);
        } else {
            const int64_t nitems = int64_t(nog) * nig * n * nsplit;
            need(nitems <= 0x7ffffff0, CGEF_ERR_DIMENSION, "cgef_conv5_wgrad: too many blocks");
            dim3 g(unsigned((nitems + 7) / 8 * 8));
            auto go = [&](auto kern) {
                hipLaunchKernelGGL(kern, g, dim3(256), 0, st, dz, in, scratch, cin, cout, hout, wout, tiles_x, ntiles, tpb,
                                   nig, nog, n, nsplit);
            };
#define CGEF_WG(OGV, IGV)                                                                        \
    do {                                                                                         \
        if (in_dtype == CGEF_BF16 && in_gelu) go(k_conv5_wgrad<OGV, IGV, CGEF_BF16, 1>);         \
        else if (in_dtype == CGEF_BF16) go(k_conv5_wgrad<OGV, IGV, CGEF_BF16, 0>);               \
        else go(k_conv5_wgrad<OGV, IGV, CGEF_F32, 0>);                                           \
    } while (0)
            if (cin <= 4) CGEF_WG(64, 4);
            else if (cout <= 8) CGEF_WG(8, 32);
            else CGEF_WG(WG, WG);
#undef CGEF_WG
        }
        check_launch();
        const int64_t len = nw + cout;
        if ((len + 63) / 64 >= 1024)
            hipLaunchKernelGGL((k_wgrad_reduce<64, 4>), dim3(unsigned((len + 63) / 64)), dim3(256), 0, st, scratch, nslots,
                               nw, cout, dw, db, mfma ? 1 : 0, cin);
        else
            hipLaunchKernelGGL((k_wgrad_reduce<16, 16>), dim3(unsigned((len + 15) / 16)), dim3(256), 0, st, scratch,
                               nslots, nw, cout, dw, db, mfma ? 1 : 0, cin);
        check_launch();
    });
}

cgef_status cgef_conv5_pack_weights(const float *w, int64_t cout, int64_t cin, uint16_t *wf, int64_t ocs_f,
                                    uint16_t *wb, int64_t ocs_b, void *stream) {
    return guarded([&] {
        need(w && (wf || wb), CGEF_ERR_ARGUMENT, "cgef_conv5_pack_weights: null pointer");
        need(cout >= 1 && cin >= 1 && (!wf || ocs_f >= cout) && (!wb || ocs_b >= cin), CGEF_ERR_DIMENSION,
             "cgef_conv5_pack_weights: layout strides must cover the channels");
        const int64_t count = cout * cin * 25;
        hipLaunchKernelGGL(k_pack_weights, dim3(blocks_for(count)), dim3(256), 0, static_cast<hipStream_t>(stream), w,
                           count, cin, wf, ocs_f, wb, ocs_b);
        check_launch();
    });
}

cgef_status cgef_adam_update(float *param, float *m, float *v, const float *grad, int64_t count, double lr,
                             double beta1, double beta2, double eps, int64_t step, int64_t cout, int64_t cin,
                             uint16_t *wf, int64_t ocs_f, uint16_t *wb, int64_t ocs_b, void *stream) {
    return guarded([&] {
        need(param && m && v && grad, CGEF_ERR_ARGUMENT, "cgef_adam_update: null pointer");
        need(step >= 1 && count >= 1, CGEF_ERR_ARGUMENT, "cgef_adam_update: step counts from 1");
        if (wf || wb)
            need(count == cout * cin * 25 && (!wf || ocs_f >= cout) && (!wb || ocs_b >= cin), CGEF_ERR_DIMENSION,
                 "cgef_adam_update: BF16 layouts need count = cout·cin·25 and strides covering the channels");
        const double bc1 = 1.0 - std::pow(beta1, double(step)), bc2 = 1.0 - std::pow(beta2, double(step));
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(count)), dim3(256), 0, static_cast<hipStream_t>(stream), param, m, v,
                           grad, count, float(lr), float(beta1), float(beta2), float(eps), float(bc1), float(bc2),
                           cin > 0 ? cin : 1, wf, ocs_f, wb, ocs_b);
        check_launch();
    });
}

cgef_status cgef_mse_grad(const float *pred, const float *target, int64_t count, float *dpred, float *loss_sum,
                          void *stream) {
    return guarded([&] {
        need(pred && target && dpred && count >= 1, CGEF_ERR_ARGUMENT, "cgef_mse_grad: bad arguments");
        hipLaunchKernelGGL(k_mse, dim3(blocks_for(count)), dim3(256), 0, static_cast<hipStream_t>(stream), pred, target,
                           count, float(1.0 / double(count)), dpred, loss_sum);
        check_launch();
    });
}

cgef_status cgef_closure_loss_grad(const float *pred, const float *target, const float *feat, const float *coef,
                                   double lambda_pi, int64_t n, int64_t npix, float *dpred, float *loss, void *stream) {
    return guarded([&] {
        need(pred && target && feat && coef && dpred && n >= 1 && npix >= 1, CGEF_ERR_ARGUMENT,
             "cgef_closure_loss_grad: bad arguments");
        const int64_t total = n * npix;
        hipLaunchKernelGGL(k_tau_pi_loss, dim3(blocks_for(total)), dim3(256), 0, static_cast<hipStream_t>(stream), pred,
                           target, feat, coef, float(lambda_pi), npix, total, float(1.0 / double(3 * total)),
                           float(1.0 / double(total)), dpred, loss);
        check_launch();
    });
}

cgef_status cgef_closure_crop(const void *src, int32_t src_dtype, int64_t nsrc, int64_t c, int64_t H, int64_t W,
                              const int32_t *origins, int64_t n, int64_t h, int64_t w, const float *chan_scale,
                              void *dst, int32_t dst_dtype, void *stream) {
    return guarded([&] {
        need(src && origins && dst, CGEF_ERR_ARGUMENT, "cgef_closure_crop: null pointer");
        need(src_dtype == CGEF_F32 && (dst_dtype == CGEF_F32 || dst_dtype == CGEF_BF16), CGEF_ERR_ARGUMENT,
             "cgef_closure_crop: FP32 source, FP32 or BF16 destination");
        need(nsrc >= 1 && c >= 1 && H >= 1 && W >= 1 && n >= 1 && h >= 1 && w >= 1, CGEF_ERR_DIMENSION,
             "cgef_closure_crop: bad sizes");
        const int64_t total = n * c * h * w;
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (dst_dtype == CGEF_BF16)
            hipLaunchKernelGGL((k_crop<CGEF_F32, CGEF_BF16>), dim3(blocks_for(total)), dim3(256), 0, st, src, nsrc, c, H, W,
                               origins, h, w, chan_scale, dst, total);
        else
            hipLaunchKernelGGL((k_crop<CGEF_F32, CGEF_F32>), dim3(blocks_for(total)), dim3(256), 0, st, src, nsrc, c, H, W,
                               origins, h, w, chan_scale, dst, total);
        check_launch();
    });
}

cgef_status cgef_closure_products(const void *u, const void *v, int32_t dtype, int64_t npts, int64_t nb, double *out5,
                                  void *stream) {
    return guarded([&] {
        need(u && v && out5 && npts >= 1 && nb >= 1, CGEF_ERR_ARGUMENT, "cgef_closure_products: bad arguments");
        need(dtype == CGEF_F32 || dtype == CGEF_F64, CGEF_ERR_ARGUMENT, "cgef_closure_products: FP32 or FP64 fields");
        const int64_t total = npts * nb;
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (dtype == CGEF_F32)
            hipLaunchKernelGGL(k_products<float>, dim3(blocks_for(total)), dim3(256), 0, st,
                               static_cast<const float *>(u), static_cast<const float *>(v), npts, total, out5);
        else
            hipLaunchKernelGGL(k_products<double>, dim3(blocks_for(total)), dim3(256), 0, st,
                               static_cast<const double *>(u), static_cast<const double *>(v), npts, total, out5);
        check_launch();
    });
}

cgef_status cgef_closure_targets(const double *filt5, const double *grad4, int64_t npts, int64_t nb, float *tau3,
                                 float *feat4, void *stream) {
    return guarded([&] {
        need(grad4 && feat4 && (!tau3 || filt5) && npts >= 1 && nb >= 1, CGEF_ERR_ARGUMENT,
             "cgef_closure_targets: bad arguments");
        const int64_t total = npts * nb;
        hipLaunchKernelGGL(k_targets, dim3(blocks_for(total)), dim3(256), 0, static_cast<hipStream_t>(stream), filt5,
                           grad4, npts, total, tau3, feat4);
        check_launch();
    });
}

cgef_status cgef_closure_pi(const float *feat4, const float *tau3, const float *tau_scale, double sigma2, int64_t npts,
                            int64_t nb, float *Pi, void *stream) {
    return guarded([&] {
        need(feat4 && tau3 && Pi && npts >= 1 && nb >= 1, CGEF_ERR_ARGUMENT, "cgef_closure_pi: bad arguments");
        const int64_t total = npts * nb;
        hipLaunchKernelGGL(k_closure_pi, dim3(blocks_for(total)), dim3(256), 0, static_cast<hipStream_t>(stream), feat4,
                           tau3, tau_scale, float(sigma2), npts, total, Pi);
        check_launch();
    });
}

cgef_status cgef_closure_gradient_model(const float *feat4, double sigma2, const float *base3, int64_t npts,
                                        int64_t nb, float *tau3, void *stream) {
    return guarded([&] {
        need(feat4 && tau3 && npts >= 1 && nb >= 1, CGEF_ERR_ARGUMENT, "cgef_closure_gradient_model: bad arguments");
        const int64_t total = npts * nb;
        hipLaunchKernelGGL(k_gradient_model, dim3(blocks_for(total)), dim3(256), 0, static_cast<hipStream_t>(stream),
                           feat4, float(sigma2), base3, npts, total, tau3);
        check_launch();
    });
}

}  // extern "C"
