// Learned warm start for the Poisson multigrid: a 4-level convolutional
// U-Net (32 base channels) mapping a charge density to a potential, with
// BF16 weights and activations and FP32 accumulation on the MFMA units.
//
// No counterpart in the reference (its Poisson start is Belos CG + ML AMG,
// PoissonSolver.cpp:85-201); the network only supplies the initial guess and
// the multigrid still iterates to its tolerance.
//
// Layout: activations NHWC (channels fastest), weights [Cout][3][3][Cin].
// The 3x3 convolutions (stride 1, zero padding 1) with 32 or more channels on
// both sides are a persistent implicit GEMM on v_mfma_f32_16x16x32_bf16 with
// LDS-staged input tiles and weights (k_conv32p); K runs over (tap, input
// channel).  The same kernel, with transposed weights and mirrored taps, gives
// the input gradient.  The one-channel input and output layers are load-bound
// and run on plain vector kernels.  The weight gradient is a GEMM over pixels
// on the same MFMA, split across workgroups with FP32 atomics (k_wgrad32,
// k_wgrad_one).  k_conv and k_conv_wgrad (16x16x16) remain for channel counts
// the 32-channel kernels do not take.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "csrst_hip.h"
#include "hip_util.h"
#include "poisson_mg.h"

using hu::DBuf;
using hu::Error;

namespace unet {

typedef short sh4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

constexpr int kB = 256;

__device__ __forceinline__ float bf2f(bf16 v) { return __uint_as_float((uint32_t)v << 16); }
// round to nearest even
__host__ __device__ __forceinline__ bf16 f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}

// MFMA 16x16x16 (BF16 in, FP32 accumulate).  Lane l holds A[l%16][4(l/16)..+3],
// B[4(l/16)..+3][l%16] and D[4(l/16)..+3][l%16].
__device__ __forceinline__ f4 mfma(sh4 a, sh4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

__device__ __forceinline__ sh4 load4(const bf16 *p)
{
    const uint2 v = *reinterpret_cast<const uint2 *>(p);
    sh4 r;
    r[0] = (short)(v.x & 0xFFFF); r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xFFFF); r[3] = (short)(v.y >> 16);
    return r;
}

// eight bf16 values as floats (one 16-byte load or store)
struct V8 { float v[8]; };

__device__ __forceinline__ V8 ld8(const bf16 *p)
{
    const uint4 u = *reinterpret_cast<const uint4 *>(p);
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
    V8 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        r.v[2 * i] = __uint_as_float(w[i] << 16);
        r.v[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
    }
    return r;
}

__device__ __forceinline__ void st8(bf16 *p, const V8 &r)
{
    uint32_t w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = (uint32_t)f2bf(r.v[2 * i]) | ((uint32_t)f2bf(r.v[2 * i + 1]) << 16);
    *reinterpret_cast<uint4 *>(p) = make_uint4(w[0], w[1], w[2], w[3]);
}

struct ConvArgs {
    const bf16 *x;       // [P][Cin]
    const bf16 *w;       // [CoutP][9][Cin] (rows past Cout are zero)
    const float *bias;   // [Cout] or null
    bf16 *y;             // [P][Cout] (bf16 out) or null
    float *yf;           // [P][Cout] (fp32 out) or null
    int B, H, W, Cin, Cout, CoutP;
    int flip;            // 1: taps mirrored (input gradient)
    int relu;
};

// Implicit-GEMM 3x3 convolution.  grid = (ceil(P / 256), CoutP / (16 MT));
// each of the 4 waves covers 64 consecutive pixels of the block's 256 and MT
// tiles of 16 output channels, so each input load feeds MT MFMAs.
template <int MT>
__global__ void __launch_bounds__(kB) k_conv(ConvArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int P = a.B * a.H * a.W;
    const int co0 = blockIdx.y * 16 * MT;
    const int pw = blockIdx.x * 256 + wave * 64;
    if (pw >= P) return;
    const int li = lane & 15, kg = lane >> 4;
    f4 acc[MT][4];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[m][t] = f4{0.f, 0.f, 0.f, 0.f};
    // pixel coordinates of this lane's B columns
    int pb[4], py[4], px[4];
    bool pv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int p = pw + 16 * t + li;
        pv[t] = p < P;
        const int pp = pv[t] ? p : 0;
        px[t] = pp % a.W;
        py[t] = (pp / a.W) % a.H;
        pb[t] = pp / (a.W * a.H);
    }
    const int Cin = a.Cin;
    if ((Cin & 15) == 0) {
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = (a.flip ? 1 - tap / 3 : tap / 3 - 1), dx = (a.flip ? 1 - tap % 3 : tap % 3 - 1);
            int off[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int ys = py[t] + dy, xs = px[t] + dx;
                const bool ok = pv[t] && ys >= 0 && ys < a.H && xs >= 0 && xs < a.W;
                off[t] = ok ? ((pb[t] * a.H + ys) * a.W + xs) * Cin : -1;
            }
            const bf16 *wr = a.w + ((size_t)(co0 + li) * 9 + tap) * Cin;
            const size_t mstride = (size_t)16 * 9 * Cin;   // next 16 output channels
            for (int c0 = 0; c0 < Cin; c0 += 16) {
                sh4 av[MT], bv[4];
#pragma unroll
                for (int m = 0; m < MT; ++m) av[m] = load4(wr + m * mstride + c0 + 4 * kg);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    bv[t] = sh4{0, 0, 0, 0};
                    if (off[t] >= 0) bv[t] = load4(a.x + off[t] + c0 + 4 * kg);
                }
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                    for (int t = 0; t < 4; ++t) acc[m][t] = mfma(av[m], bv[t], acc[m][t]);
            }
        }
    } else if (MT == 1) {
        // small Cin: K = 9 Cin flattened as k = tap * Cin + ci, padded to 16
        const int K = 9 * Cin;
        for (int k0 = 0; k0 < K; k0 += 16) {
            sh4 av;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = k0 + 4 * kg + r;
                av[r] = k < K ? (short)a.w[(size_t)(co0 + li) * 9 * Cin + k] : (short)0;
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                sh4 bv;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int k = k0 + 4 * kg + r;
                    short v = 0;
                    if (k < K && pv[t]) {
                        const int tap = k / Cin, ci = k % Cin;
                        const int dy = (a.flip ? 1 - tap / 3 : tap / 3 - 1), dx = (a.flip ? 1 - tap % 3 : tap % 3 - 1);
                        const int ys = py[t] + dy, xs = px[t] + dx;
                        if (ys >= 0 && ys < a.H && xs >= 0 && xs < a.W)
                            v = (short)a.x[((size_t)(pb[t] * a.H + ys) * a.W + xs) * Cin + ci];
                    }
                    bv[r] = v;
                }
                acc[0][t] = mfma(av, bv, acc[0][t]);
            }
        }
    }
    // epilogue: rows co0 + 16 m + 4 kg + r, column = pixel
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (!pv[t]) continue;
            const size_t p = (size_t)pw + 16 * t + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = co0 + 16 * m + 4 * kg + r;
                if (co >= a.Cout) continue;
                float v = acc[m][t][r] + (a.bias ? a.bias[co] : 0.f);
                if (a.relu) v = v > 0.f ? v : 0.f;
                if (a.y) a.y[p * a.Cout + co] = f2bf(v);
                if (a.yf) a.yf[p * a.Cout + co] = v;
            }
        }
}

// ---------------------------------------------------------------------------
// The 3x3 convolution for Cin a multiple of 32 and Cout of 16 MF on the
// gfx950 v_mfma_f32_16x16x32_bf16 (K = 32 input channels of one tap per MFMA),
// persistent: a workgroup of 8 waves keeps one block of 16 MF output channels
// and walks its work items (a 16 x 32 pixel tile, a chunk of 32 input
// channels) tile by tile.  Per item the (16+2) x (32+2) input pixels with their
// halo and the 16 MF x 9 x 32 weights sit in LDS and each wave runs 9 taps x
// MF x 4 MFMAs on tile rows 2w and 2w+1.  The global loads of the next DEPTH
// items are in flight in registers while the current one computes (DEPTH 2
// where one item's matrix work is too short to cover the memory latency);
// with Cin = 32 the weights are staged once for all tiles.  An LDS pixel (or
// weight row) holds 32 channels in 64 B padded to 80 B, so the 16-byte reads
// of eight consecutive lanes fall on distinct banks.  Lane l holds
// A[l%16][8(l/16)..+7] and B[8(l/16)..+7][l%16] (one 16-byte LDS read each)
// and D[4(l/16)..+3][l%16].
constexpr int kTW = 32;                       // output tile: pixels per row
constexpr int kIW = kTW + 2;                  // input tile row with the halo
constexpr int kPS = 40;                       // LDS stride of 32 channels, in bf16
constexpr int kPH = 16;                       // output rows per tile
constexpr int kPIH = kPH + 2;                 // input rows with the halo
constexpr int kPB = 512;                      // threads: 8 waves
constexpr int kXSeg = kPIH * kIW * 4;         // 16-byte segments of one input tile chunk
constexpr int kXPer = (kXSeg + kPB - 1) / kPB;
constexpr int kTH = 8, kIH = kTH + 2;         // 8 x 32 pixel tiles of the one-channel layers

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f4 mfma32(bf8 a, bf8 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

inline int conv32p_tiles(const ConvArgs &a)
{
    return a.B * ((a.H + kPH - 1) / kPH) * ((a.W + kTW - 1) / kTW);
}

template <int MF, int DEPTH>
__global__ void __launch_bounds__(kPB, 1) k_conv32p(ConvArgs a)
{
    constexpr int kWSeg = 9 * 16 * MF * 4;    // 16-byte segments of one weight chunk
    constexpr int kWPer = (kWSeg + kPB - 1) / kPB;
    __shared__ __attribute__((aligned(16))) bf16 sX[kPIH * kIW * kPS];
    __shared__ __attribute__((aligned(16))) bf16 sW[9 * 16 * MF * kPS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, kg = lane >> 4;
    const int nCo = a.CoutP / (16 * MF);
    const int tilesX = (a.W + kTW - 1) / kTW, tilesY = (a.H + kPH - 1) / kPH;
    const int nTiles = a.B * tilesX * tilesY;
    const int cb = blockIdx.x % nCo, g = blockIdx.x / nCo, G = gridDim.x / nCo;
    const int co0 = cb * 16 * MF;
    const int Cin = a.Cin, nChunks = Cin / 32;
    if (g >= nTiles) return;
    const int nItems = ((nTiles - g + G - 1) / G) * nChunks;   // tiles g, g + G, ... times chunks
    const bool wOnce = nChunks == 1;
    uint4 rxA[kXPer], rwA[kWPer], rxB[DEPTH == 2 ? kXPer : 1], rwB[DEPTH == 2 ? kWPer : 1];
    f4 acc[MF][4];
#pragma unroll
    for (int m = 0; m < MF; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[m][t] = f4{0.f, 0.f, 0.f, 0.f};
    // this lane's biases in registers: a global load in the epilogue would
    // make its wait also wait for the prefetch in flight
    float bias[MF][4];
#pragma unroll
    for (int m = 0; m < MF; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) bias[m][r] = a.bias ? a.bias[co0 + 16 * m + 4 * kg + r] : 0.f;
    // item -> (image, tile origin, first input channel)
#define CONV32P_ORIGIN(item)                                       \
    const int tile_ = g + ((item) / nChunks) * G;                  \
    const int c0 = ((item) % nChunks) * 32;                        \
    const int x0 = (tile_ % tilesX) * kTW;                         \
    const int y0 = ((tile_ / tilesX) % tilesY) * kPH;              \
    const int bimg = (tile_ / tilesX) / tilesY;
    // global loads of an item into rx (and rw)
#define CONV32P_FETCH(item, withW, rx, rw)                                                                           \
    {                                                                                                                \
        CONV32P_ORIGIN(item)                                                                                         \
        _Pragma("unroll") for (int j = 0; j < kXPer; ++j)                                                            \
        {                                                                                                            \
            const int s = tid + kPB * j;                                                                             \
            const int pix = s >> 2, seg = s & 3;                                                                     \
            const int iy = pix / kIW, ix = pix - iy * kIW;                                                           \
            const int ys = y0 + iy - 1, xs = x0 + ix - 1;                                                            \
            const bool in = s < kXSeg && ys >= 0 && ys < a.H && xs >= 0 && xs < a.W;                                 \
            rx[j] = in ? *reinterpret_cast<const uint4 *>(a.x + ((size_t)(bimg * a.H + ys) * a.W + xs) * Cin + c0 +  \
                                                          8 * seg)                                                   \
                       : make_uint4(0u, 0u, 0u, 0u);                                                                 \
        }                                                                                                            \
        if (withW) {                                                                                                 \
            _Pragma("unroll") for (int j = 0; j < kWPer; ++j)                                                        \
            {                                                                                                        \
                const int s = tid + kPB * j;                                                                         \
                const int seg = s & 3, r = s >> 2;              /* LDS row r = tap * 16 MF + col */                \
                const int tap = r / (16 * MF), col = r - tap * (16 * MF);                                            \
                rw[j] = s < kWSeg ? *reinterpret_cast<const uint4 *>(a.w + ((size_t)(co0 + col) * 9 + tap) * Cin +  \
                                                                     c0 + 8 * seg)                                   \
                                  : make_uint4(0u, 0u, 0u, 0u);                                                      \
            }                                                                                                        \
        }                                                                                                            \
    }
    // one item: its registers into LDS, the loads of item + DEPTH issued, the
    // MFMAs, and after the tile's last chunk bias, ReLU, bf16 store
#define CONV32P_STEP(item, rx, rw)                                                                                   \
    {                                                                                                                \
        const bool loadW = !wOnce || (item) == 0;                                                                    \
        __syncthreads();                                                                                             \
        _Pragma("unroll") for (int j = 0; j < kXPer; ++j)                                                            \
        {                                                                                                            \
            const int s = tid + kPB * j;                                                                             \
            if (s < kXSeg) *reinterpret_cast<uint4 *>(&sX[(s >> 2) * kPS + 8 * (s & 3)]) = rx[j];                    \
        }                                                                                                            \
        if (loadW)                                                                                                   \
            _Pragma("unroll") for (int j = 0; j < kWPer; ++j)                                                        \
            {                                                                                                        \
                const int s = tid + kPB * j;                                                                         \
                const int seg = s & 3;                                                                               \
                if (s < kWSeg) *reinterpret_cast<uint4 *>(&sW[(s >> 2) * kPS + 8 * seg]) = rw[j];                     \
            }                                                                                                        \
        __syncthreads();                                                                                             \
        if ((item) + DEPTH < nItems) CONV32P_FETCH((item) + DEPTH, !wOnce, rx, rw)                                   \
        _Pragma("unroll") for (int tap = 0; tap < 9; ++tap)                                                          \
        {                                                                                                            \
            const int dy = a.flip ? 1 - tap / 3 : tap / 3 - 1, dx = a.flip ? 1 - tap % 3 : tap % 3 - 1;              \
            bf8 av[MF], bv[4];                                                                                       \
            _Pragma("unroll") for (int m = 0; m < MF; ++m) av[m] =                                                   \
                *reinterpret_cast<const bf8 *>(&sW[(tap * 16 * MF + 16 * m + li) * kPS + 8 * kg]);                   \
            _Pragma("unroll") for (int t = 0; t < 4; ++t)                                                            \
            {                                                                                                        \
                const int ry = 2 * wave + (t >> 1) + 1 + dy, rxx = 16 * (t & 1) + li + 1 + dx;                      \
                bv[t] = *reinterpret_cast<const bf8 *>(&sX[(ry * kIW + rxx) * kPS + 8 * kg]);                        \
            }                                                                                                        \
            _Pragma("unroll") for (int m = 0; m < MF; ++m)                                                           \
                _Pragma("unroll") for (int t = 0; t < 4; ++t) acc[m][t] = mfma32(av[m], bv[t], acc[m][t]);           \
        }                                                                                                            \
        if ((item) % nChunks == nChunks - 1) {                                                                       \
            CONV32P_ORIGIN(item)                                                                                     \
            (void)c0;                                                                                                \
            _Pragma("unroll") for (int t = 0; t < 4; ++t)                                                            \
            {                                                                                                        \
                const int yy = y0 + 2 * wave + (t >> 1), xx = x0 + 16 * (t & 1) + li;                               \
                const bool ok = yy < a.H && xx < a.W;                                                                \
                const size_t p = ((size_t)bimg * a.H + yy) * a.W + xx;                                               \
                _Pragma("unroll") for (int m = 0; m < MF; ++m)                                                       \
                {                                                                                                    \
                    const int co = co0 + 16 * m + 4 * kg;                                                            \
                    float v[4];                                                                                      \
                    _Pragma("unroll") for (int r = 0; r < 4; ++r)                                                    \
                    {                                                                                                \
                        v[r] = acc[m][t][r] + bias[m][r];                                                            \
                        if (a.relu) v[r] = v[r] > 0.f ? v[r] : 0.f;                                                  \
                    }                                                                                                \
                    acc[m][t] = f4{0.f, 0.f, 0.f, 0.f};                                                              \
                    if (ok && a.y) {                                                                                 \
                        uint2 o;                                                                                     \
                        o.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);                                   \
                        o.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);                                   \
                        *reinterpret_cast<uint2 *>(a.y + p * a.Cout + co) = o;                                       \
                    }                                                                                                \
                    if (ok && a.yf)                                                                                  \
                        _Pragma("unroll") for (int r = 0; r < 4; ++r) a.yf[p * a.Cout + co + r] = v[r];              \
                }                                                                                                    \
            }                                                                                                        \
        }                                                                                                            \
    }
    CONV32P_FETCH(0, true, rxA, rwA)
    if (DEPTH == 2 && nItems > 1) CONV32P_FETCH(1, !wOnce, rxB, rwB)
    for (int item = 0; item < nItems; item += DEPTH) {
        CONV32P_STEP(item, rxA, rwA)
        if (DEPTH == 2 && item + 1 < nItems) CONV32P_STEP(item + 1, rxB, rwB)
    }
#undef CONV32P_STEP
#undef CONV32P_FETCH
#undef CONV32P_ORIGIN
}

// 3x3 convolution from one input channel to Cout (a multiple of 8, at most
// 256) channels: one thread per pixel; the weights as floats in LDS, [tap][co],
// read by every lane at one address.  Load-bound: 2 B in and 2 Cout B out per
// pixel.
__global__ void __launch_bounds__(kB) k_conv_c1(ConvArgs a)
{
    __shared__ float sw[9 * 256];
    for (int k = threadIdx.x; k < a.Cout * 9; k += kB) {
        const int tap = k / a.Cout, co = k - tap * a.Cout;
        sw[k] = bf2f(a.w[co * 9 + tap]);
    }
    __syncthreads();
    const int p = blockIdx.x * kB + threadIdx.x;        // pixels < 2^31 (checked by the caller)
    if (p >= a.B * a.H * a.W) return;
    const int xx = p % a.W, yy = (p / a.W) % a.H, b = p / (a.W * a.H);
    float xin[9];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = a.flip ? 1 - tap / 3 : tap / 3 - 1, dx = a.flip ? 1 - tap % 3 : tap % 3 - 1;
        const int ys = yy + dy, xs = xx + dx;
        xin[tap] = (ys >= 0 && ys < a.H && xs >= 0 && xs < a.W) ? bf2f(a.x[((size_t)b * a.H + ys) * a.W + xs]) : 0.f;
    }
    for (int g = 0; g < a.Cout / 8; ++g) {
        V8 o;
#pragma unroll
        for (int i = 0; i < 8; ++i) o.v[i] = a.bias ? a.bias[8 * g + i] : 0.f;
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const float4 w0 = *reinterpret_cast<const float4 *>(&sw[tap * a.Cout + 8 * g]);
            const float4 w1 = *reinterpret_cast<const float4 *>(&sw[tap * a.Cout + 8 * g + 4]);
            o.v[0] += w0.x * xin[tap]; o.v[1] += w0.y * xin[tap]; o.v[2] += w0.z * xin[tap]; o.v[3] += w0.w * xin[tap];
            o.v[4] += w1.x * xin[tap]; o.v[5] += w1.y * xin[tap]; o.v[6] += w1.z * xin[tap]; o.v[7] += w1.w * xin[tap];
        }
        if (a.relu)
#pragma unroll
            for (int i = 0; i < 8; ++i) o.v[i] = o.v[i] < 0.f ? 0.f : o.v[i];
        st8(a.y + (size_t)p * a.Cout + 8 * g, o);
    }
}

// 3x3 convolution from Cin (a multiple of 32) channels to one: an 8 x 32
// pixel tile per block of 256 threads, one pixel per thread; per chunk of 32
// channels the tile and its halo are staged in LDS as in k_conv32, the 9 x 32
// weights as floats.  Load-bound: 2 Cin B in and 2 (+4) B out per pixel.
__global__ void __launch_bounds__(kB) k_conv_to1(ConvArgs a)
{
    __shared__ __attribute__((aligned(16))) bf16 sX[kIH * kIW * kPS];
    __shared__ float sw[9 * 32];
    const int tilesX = (a.W + kTW - 1) / kTW, tilesY = (a.H + kTH - 1) / kTH;
    int tile = blockIdx.x;
    const int tx = tile % tilesX;
    tile /= tilesX;
    const int ty = tile % tilesY;
    const int bimg = tile / tilesY;
    const int x0 = tx * kTW, y0 = ty * kTH;
    const int ry0 = threadIdx.x / kTW, rx0 = threadIdx.x % kTW;
    const int Cin = a.Cin;
    float acc = 0.f;
    for (int c0 = 0; c0 < Cin; c0 += 32) {
        __syncthreads();
        for (int s = threadIdx.x; s < kIH * kIW * 4; s += kB) {
            const int pix = s >> 2, seg = s & 3;
            const int iy = pix / kIW, ix = pix - iy * kIW;
            const int ys = y0 + iy - 1, xs = x0 + ix - 1;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (ys >= 0 && ys < a.H && xs >= 0 && xs < a.W)
                v = *reinterpret_cast<const uint4 *>(a.x + ((size_t)(bimg * a.H + ys) * a.W + xs) * Cin + c0 + 8 * seg);
            *reinterpret_cast<uint4 *>(&sX[pix * kPS + 8 * seg]) = v;
        }
        for (int k = threadIdx.x; k < 9 * 32; k += kB) sw[k] = bf2f(a.w[(k >> 5) * Cin + c0 + (k & 31)]);
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = a.flip ? 1 - tap / 3 : tap / 3 - 1, dx = a.flip ? 1 - tap % 3 : tap % 3 - 1;
            const bf16 *q = &sX[((ry0 + 1 + dy) * kIW + rx0 + 1 + dx) * kPS];
#pragma unroll
            for (int seg = 0; seg < 4; ++seg) {
                const V8 v = ld8(q + 8 * seg);
#pragma unroll
                for (int i = 0; i < 8; ++i) acc += sw[tap * 32 + 8 * seg + i] * v.v[i];
            }
        }
    }
    const int yy = y0 + ry0, xx = x0 + rx0;
    if (yy >= a.H || xx >= a.W) return;
    const size_t p = ((size_t)bimg * a.H + yy) * a.W + xx;
    float v = acc + (a.bias ? a.bias[0] : 0.f);
    if (a.relu && v < 0.f) v = 0.f;
    if (a.y) a.y[p] = f2bf(v);
    if (a.yf) a.yf[p] = v;
}

// k_conv32p where the channels allow it (Cin a multiple of 32, Cout of 32);
// the one-channel input and output layers on their load-bound kernels;
// otherwise k_conv with the widest channel blocking that divides CoutP
// (small Cin takes the flattened path, one tile)
void launch_conv(const ConvArgs &a, hipStream_t s)
{
    if (a.Cin % 32 == 0 && a.Cout % 32 == 0 && a.CoutP == a.Cout) {
        // one persistent workgroup per CU and occupancy slot
        static int slots4 = 0, slots2 = 0;
        if (!slots4) {
            int dev = 0, cus = 0, o4 = 0, o2 = 0;
            HU_CHECK(hipGetDevice(&dev));
            HU_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            HU_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o4, k_conv32p<4, 1>, kPB, 0));
            HU_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o2, k_conv32p<2, 2>, kPB, 0));
            slots2 = cus * std::max(o2, 1);
            slots4 = cus * std::max(o4, 1);
        }
        const int mf = a.Cout % 64 == 0 ? 4 : 2, nCo = a.Cout / (16 * mf);
        const int G = std::max(1, std::min(conv32p_tiles(a), (mf == 4 ? slots4 : slots2) / nCo));
        if (mf == 4) k_conv32p<4, 1><<<G * nCo, kPB, 0, s>>>(a);
        else k_conv32p<2, 2><<<G * nCo, kPB, 0, s>>>(a);
        return;
    }
    if (a.Cin == 1 && a.Cout % 8 == 0 && a.Cout <= 256 && a.y && (double)a.B * a.H * a.W < 2e9) {
        const int n = a.B * a.H * a.W;
        k_conv_c1<<<(n + kB - 1) / kB, kB, 0, s>>>(a);
        return;
    }
    if (a.Cout == 1 && a.Cin % 32 == 0) {
        k_conv_to1<<<a.B * ((a.H + kTH - 1) / kTH) * ((a.W + kTW - 1) / kTW), kB, 0, s>>>(a);
        return;
    }
    const int P = a.B * a.H * a.W;
    const int mt = (a.Cin & 15) ? 1 : (a.CoutP % 64 == 0) ? 4 : (a.CoutP % 32 == 0) ? 2 : 1;
    const dim3 grid((P + 255) / 256, a.CoutP / (16 * mt));
    if (mt == 4) k_conv<4><<<grid, kB, 0, s>>>(a);
    else if (mt == 2) k_conv<2><<<grid, kB, 0, s>>>(a);
    else k_conv<1><<<grid, kB, 0, s>>>(a);
}

// Weight gradient dW[co][tap][ci] += sum_p dY[p][co] X[p + s(tap)][ci]:
// M = Cout, N = 9 Cin (n = tap * Cin + ci), K = pixels, split over blockIdx.z.
// grid = (ceil(9 Cin / 64), CoutP / 16, splits); 64 threads per block.
__global__ void __launch_bounds__(64) k_conv_wgrad(const bf16 *dy, const bf16 *x, float *dw, int B, int H, int W,
                                                  int Cin, int Cout, int chunk)
{
    const int lane = threadIdx.x, li = lane & 15, kg = lane >> 4;
    const int P = B * H * W, N = 9 * Cin;
    const int co0 = blockIdx.y * 16;
    const int n0 = blockIdx.x * 64;
    const int k0 = blockIdx.z * chunk, k1 = min(P, k0 + chunk);
    f4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    int tapn[4], cin[4];
    bool nv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int n = n0 + 16 * t + li;
        nv[t] = n < N;
        tapn[t] = nv[t] ? n / Cin : 0;
        cin[t] = nv[t] ? n % Cin : 0;
    }
    const int co = co0 + li;
    for (int kb = k0; kb < k1; kb += 16) {
        sh4 av;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int p = kb + 4 * kg + r;
            av[r] = (p < k1 && co < Cout) ? (short)dy[(size_t)p * Cout + co] : (short)0;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            sh4 bv;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = kb + 4 * kg + r;
                short v = 0;
                if (p < k1 && nv[t]) {
                    const int xx = p % W, yy = (p / W) % H, bb = p / (W * H);
                    const int ys = yy + tapn[t] / 3 - 1, xs = xx + tapn[t] % 3 - 1;
                    if (ys >= 0 && ys < H && xs >= 0 && xs < W)
                        v = (short)x[((size_t)(bb * H + ys) * W + xs) * Cin + cin[t]];
                }
                bv[r] = v;
            }
            acc[t] = mfma(av, bv, acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        if (!nv[t]) continue;
        const int n = n0 + 16 * t + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = co0 + 4 * kg + r;
            if (c < Cout) atomicAdd(&dw[(size_t)c * N + n], acc[t][r]);
        }
    }
}

// Weight gradient for Cin and Cout multiples of 32 on v_mfma_f32_16x16x32_bf16:
// dW[co][tap][ci] = sum_p dY[p][co] X[p + s(tap)][ci], K = pixels.  A
// workgroup (4 waves) owns a block of 16 CF WCO output channels x 64 / WCO
// input channels, all 9 taps, and sums over the 8 x 32 pixel tiles
// blockIdx / nBlk, + ksplit, ...; wave w takes output-channel fragments
// (w % WCO) CF .. + CF - 1 and input-channel fragment w / WCO.  Per tile dY
// and the input with its halo are staged in LDS transposed (pixels fastest,
// 4 pixels per 8-byte write), so a lane's 8 consecutive pixels of a row are
// one 16-byte read.  Per K step (one tile row) and tap row dy, two aligned
// reads of the input row give the three dx shifts by funnel shifts of the
// lane's registers.  Row pitches are padded so that 16 lanes' 16-byte reads
// fall on distinct banks.  One atomic per weight and workgroup at the end.
constexpr int kWgTH = 8, kWgTW = 32;          // pixel tile
constexpr int kDyPitch = kWgTH * kWgTW + 8;   // bf16 per output channel in LDS
constexpr int kXCols = 40;                    // input columns x0 - 1 .. x0 + 38
constexpr int kXPitch = (kWgTH + 2) * kXCols + 8;   // bf16 per input channel in LDS

typedef uint32_t u4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf8 as_bf8(u4v v) { return __builtin_bit_cast(bf8, v); }

template <int CF, int WCO>
__global__ void __launch_bounds__(kB, 2) k_wgrad32(const bf16 *dy, const bf16 *x, float *dw, int B, int H, int W,
                                                   int Cin, int Cout, int ksplit)
{
    constexpr int CO_BLK = 16 * CF * WCO, CI_BLK = 16 * (4 / WCO);
    __shared__ __attribute__((aligned(16))) bf16 sD[CO_BLK * kDyPitch];
    __shared__ __attribute__((aligned(16))) bf16 sXt[CI_BLK * kXPitch];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, kg = lane >> 4;
    const int nCoB = Cout / CO_BLK, nCiB = Cin / CI_BLK, nBlk = nCoB * nCiB;
    const int blk = blockIdx.x % nBlk, kz = blockIdx.x / nBlk;
    const int co0 = (blk % nCoB) * CO_BLK, ci0 = (blk / nCoB) * CI_BLK;
    const int cf0 = (wave % WCO) * CF, nf = wave / WCO;
    const int tilesX = (W + kWgTW - 1) / kWgTW, tilesY = (H + kWgTH - 1) / kWgTH, nTiles = B * tilesX * tilesY;
    f4 acc[CF][9];
#pragma unroll
    for (int m = 0; m < CF; ++m)
#pragma unroll
        for (int t = 0; t < 9; ++t) acc[m][t] = f4{0.f, 0.f, 0.f, 0.f};
    for (int tile = kz; tile < nTiles; tile += ksplit) {
        const int tx = tile % tilesX, rest = tile / tilesX;
        const int x0 = tx * kWgTW, y0 = (rest % tilesY) * kWgTH, bimg = rest / tilesY;
        __syncthreads();
        // dY: units (row, 4-pixel group, 8-channel group)
        for (int u = tid; u < kWgTH * (kWgTW / 4) * (CO_BLK / 8); u += kB) {
            const int cg = u % (CO_BLK / 8), r2 = u / (CO_BLK / 8);
            const int x4 = r2 % (kWgTW / 4), r = r2 / (kWgTW / 4);
            const int yy = y0 + r;
            uint4 v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int xx = x0 + 4 * x4 + q;
                v[q] = (yy < H && xx < W)
                           ? *reinterpret_cast<const uint4 *>(dy + (((size_t)bimg * H + yy) * W + xx) * Cout + co0 + 8 * cg)
                           : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                auto h = [&](const uint4 &w) -> uint32_t {
                    const uint32_t d = e < 2 ? w.x : e < 4 ? w.y : e < 6 ? w.z : w.w;
                    return (e & 1) ? (d >> 16) : (d & 0xFFFFu);
                };
                const uint2 o = make_uint2(h(v[0]) | (h(v[1]) << 16), h(v[2]) | (h(v[3]) << 16));
                *reinterpret_cast<uint2 *>(&sD[(8 * cg + e) * kDyPitch + r * kWgTW + 4 * x4]) = o;
            }
        }
        // input with its halo: units (row, 4-column group, 8-channel group)
        for (int u = tid; u < (kWgTH + 2) * (kXCols / 4) * (CI_BLK / 8); u += kB) {
            const int ig = u % (CI_BLK / 8), r2 = u / (CI_BLK / 8);
            const int c4 = r2 % (kXCols / 4), r = r2 / (kXCols / 4);
            const int ys = y0 + r - 1;
            uint4 v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int xs = x0 - 1 + 4 * c4 + q;
                v[q] = (ys >= 0 && ys < H && xs >= 0 && xs < W)
                           ? *reinterpret_cast<const uint4 *>(x + (((size_t)bimg * H + ys) * W + xs) * Cin + ci0 + 8 * ig)
                           : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                auto h = [&](const uint4 &w) -> uint32_t {
                    const uint32_t d = e < 2 ? w.x : e < 4 ? w.y : e < 6 ? w.z : w.w;
                    return (e & 1) ? (d >> 16) : (d & 0xFFFFu);
                };
                const uint2 o = make_uint2(h(v[0]) | (h(v[1]) << 16), h(v[2]) | (h(v[3]) << 16));
                *reinterpret_cast<uint2 *>(&sXt[(8 * ig + e) * kXPitch + r * kXCols + 4 * c4]) = o;
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int r = 0; r < kWgTH; ++r) {
            bf8 av[CF];
#pragma unroll
            for (int m = 0; m < CF; ++m)
                av[m] = *reinterpret_cast<const bf8 *>(&sD[(16 * (cf0 + m) + li) * kDyPitch + r * kWgTW + 8 * kg]);
#pragma unroll
            for (int dyi = 0; dyi < 3; ++dyi) {
                const bf16 *row = &sXt[(16 * nf + li) * kXPitch + (r + dyi) * kXCols + 8 * kg];
                const u4v f0 = *reinterpret_cast<const u4v *>(row), f1 = *reinterpret_cast<const u4v *>(row + 8);
                u4v s1, s2;   // shifted by one and by two columns
                s1[0] = __builtin_amdgcn_alignbit(f0[1], f0[0], 16);
                s1[1] = __builtin_amdgcn_alignbit(f0[2], f0[1], 16);
                s1[2] = __builtin_amdgcn_alignbit(f0[3], f0[2], 16);
                s1[3] = __builtin_amdgcn_alignbit(f1[0], f0[3], 16);
                s2[0] = f0[1]; s2[1] = f0[2]; s2[2] = f0[3]; s2[3] = f1[0];
                const bf8 bv[3] = {as_bf8(f0), as_bf8(s1), as_bf8(s2)};
#pragma unroll
                for (int dxi = 0; dxi < 3; ++dxi)
#pragma unroll
                    for (int m = 0; m < CF; ++m)
                        acc[m][3 * dyi + dxi] = mfma32(av[m], bv[dxi], acc[m][3 * dyi + dxi]);
            }
        }
    }
    // D[4 kg + rr][li]: output channel co0 + 16 (cf0 + m) + 4 kg + rr, input channel ci0 + 16 nf + li
    const int ci = ci0 + 16 * nf + li;
#pragma unroll
    for (int m = 0; m < CF; ++m)
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int co = co0 + 16 * (cf0 + m) + 4 * kg + rr;
                atomicAdd(&dw[((size_t)co * 9 + t) * Cin + ci], acc[m][t][rr]);
            }
}

// Weight gradient of the one-channel layers on the same MFMA path: OUT1 is the
// output layer (Cin = 32, Cout = 1), otherwise the input layer (Cin = 1,
// Cout = 32).  The one channel fills row (or column) 0 of its 16-wide MFMA
// operand and the other lanes hold zeros.  Wave w takes 16-channel fragment
// w & 1 of the 32-channel side and tile rows (w >> 1) + 2 j.  Staging and
// shifts as in k_wgrad32; pixel tiles blockIdx.x, + gridDim.x, ...
template <bool OUT1>
__global__ void __launch_bounds__(kB) k_wgrad_one(const bf16 *dy, const bf16 *x, float *dw, int B, int H, int W)
{
    constexpr int NCO = OUT1 ? 1 : 32, NCI = OUT1 ? 32 : 1;
    __shared__ __attribute__((aligned(16))) bf16 sD[NCO * kDyPitch];
    __shared__ __attribute__((aligned(16))) bf16 sXt[NCI * kXPitch];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, kg = lane >> 4, fr = wave & 1;
    const int tilesX = (W + kWgTW - 1) / kWgTW, tilesY = (H + kWgTH - 1) / kWgTH, nTiles = B * tilesX * tilesY;
    f4 acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    const u4v zero = {0u, 0u, 0u, 0u};
    for (int tile = blockIdx.x; tile < nTiles; tile += gridDim.x) {
        const int tx = tile % tilesX, rest = tile / tilesX;
        const int x0 = tx * kWgTW, y0 = (rest % tilesY) * kWgTH, bimg = rest / tilesY;
        __syncthreads();
        // dY: units (row, 4-pixel group, 8-channel group); one channel: 4 scalars
        for (int u = tid; u < kWgTH * (kWgTW / 4) * ((NCO + 7) / 8); u += kB) {
            const int cg = u % ((NCO + 7) / 8), r2 = u / ((NCO + 7) / 8);
            const int x4 = r2 % (kWgTW / 4), r = r2 / (kWgTW / 4);
            const int yy = y0 + r;
            if (NCO == 1) {
                uint32_t h[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int xx = x0 + 4 * x4 + q;
                    h[q] = (yy < H && xx < W) ? (uint32_t)dy[((size_t)bimg * H + yy) * W + xx] : 0u;
                }
                *reinterpret_cast<uint2 *>(&sD[r * kWgTW + 4 * x4]) = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
            } else {
                uint4 v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int xx = x0 + 4 * x4 + q;
                    v[q] = (yy < H && xx < W)
                               ? *reinterpret_cast<const uint4 *>(dy + (((size_t)bimg * H + yy) * W + xx) * NCO + 8 * cg)
                               : make_uint4(0u, 0u, 0u, 0u);
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    auto h = [&](const uint4 &w) -> uint32_t {
                        const uint32_t d = e < 2 ? w.x : e < 4 ? w.y : e < 6 ? w.z : w.w;
                        return (e & 1) ? (d >> 16) : (d & 0xFFFFu);
                    };
                    *reinterpret_cast<uint2 *>(&sD[(8 * cg + e) * kDyPitch + r * kWgTW + 4 * x4]) =
                        make_uint2(h(v[0]) | (h(v[1]) << 16), h(v[2]) | (h(v[3]) << 16));
                }
            }
        }
        // input with its halo: units (row, 4-column group, 8-channel group)
        for (int u = tid; u < (kWgTH + 2) * (kXCols / 4) * ((NCI + 7) / 8); u += kB) {
            const int ig = u % ((NCI + 7) / 8), r2 = u / ((NCI + 7) / 8);
            const int c4 = r2 % (kXCols / 4), r = r2 / (kXCols / 4);
            const int ys = y0 + r - 1;
            if (NCI == 1) {
                uint32_t h[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int xs = x0 - 1 + 4 * c4 + q;
                    h[q] = (ys >= 0 && ys < H && xs >= 0 && xs < W) ? (uint32_t)x[((size_t)bimg * H + ys) * W + xs] : 0u;
                }
                *reinterpret_cast<uint2 *>(&sXt[r * kXCols + 4 * c4]) = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
            } else {
                uint4 v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int xs = x0 - 1 + 4 * c4 + q;
                    v[q] = (ys >= 0 && ys < H && xs >= 0 && xs < W)
                               ? *reinterpret_cast<const uint4 *>(x + (((size_t)bimg * H + ys) * W + xs) * NCI + 8 * ig)
                               : make_uint4(0u, 0u, 0u, 0u);
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    auto h = [&](const uint4 &w) -> uint32_t {
                        const uint32_t d = e < 2 ? w.x : e < 4 ? w.y : e < 6 ? w.z : w.w;
                        return (e & 1) ? (d >> 16) : (d & 0xFFFFu);
                    };
                    *reinterpret_cast<uint2 *>(&sXt[(8 * ig + e) * kXPitch + r * kXCols + 4 * c4]) =
                        make_uint2(h(v[0]) | (h(v[1]) << 16), h(v[2]) | (h(v[3]) << 16));
                }
            }
        }
        __syncthreads();
        // the one-channel side's operand: lane li = 0 only
        const bool one = li == 0;
#pragma unroll 1
        for (int j = 0; j < kWgTH / 2; ++j) {
            const int r = (wave >> 1) + 2 * j;
            bf8 av;
            if (OUT1) av = one ? *reinterpret_cast<const bf8 *>(&sD[r * kWgTW + 8 * kg]) : as_bf8(zero);
            else av = *reinterpret_cast<const bf8 *>(&sD[(16 * fr + li) * kDyPitch + r * kWgTW + 8 * kg]);
#pragma unroll
            for (int dyi = 0; dyi < 3; ++dyi) {
                u4v f0 = zero, f1 = zero;
                if (!OUT1 ? one : true) {
                    const bf16 *row = &sXt[(OUT1 ? (16 * fr + li) * kXPitch : 0) + (r + dyi) * kXCols + 8 * kg];
                    f0 = *reinterpret_cast<const u4v *>(row);
                    f1 = *reinterpret_cast<const u4v *>(row + 8);
                }
                u4v s1, s2;
                s1[0] = __builtin_amdgcn_alignbit(f0[1], f0[0], 16);
                s1[1] = __builtin_amdgcn_alignbit(f0[2], f0[1], 16);
                s1[2] = __builtin_amdgcn_alignbit(f0[3], f0[2], 16);
                s1[3] = __builtin_amdgcn_alignbit(f1[0], f0[3], 16);
                s2[0] = f0[1]; s2[1] = f0[2]; s2[2] = f0[3]; s2[3] = f1[0];
                acc[3 * dyi + 0] = mfma32(av, as_bf8(f0), acc[3 * dyi + 0]);
                acc[3 * dyi + 1] = mfma32(av, as_bf8(s1), acc[3 * dyi + 1]);
                acc[3 * dyi + 2] = mfma32(av, as_bf8(s2), acc[3 * dyi + 2]);
            }
        }
    }
    // D[4 kg + rr][li]: row = output channel (OUT1: only row 0), column = input channel (else: only column 0)
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        if (OUT1) {
            if (kg == 0) atomicAdd(&dw[t * 32 + 16 * fr + li], acc[t][0]);
        } else if (li == 0) {
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) atomicAdd(&dw[(16 * fr + 4 * kg + rr) * 9 + t], acc[t][rr]);
        }
    }
}

// Elementwise kernels take 8 channels (one 16-byte vector) per thread: every
// activation they touch has a channel count that is a multiple of 8.

// bias gradient: db[co] += sum_p dY[p][co].  Thread t of a block reads the 8
// channels 8 (t % (C/8)) of pixels t / (C/8) + k 256 / (C/8); the partial
// sums of one channel group meet in LDS, one atomic per channel and block.
// C a multiple of 8 dividing 2048.
__global__ void __launch_bounds__(kB) k_bias_grad8(const bf16 *dy, float *db, int P, int C)
{
    __shared__ float sh[kB][9];
    const int groups = C / 8, cg = threadIdx.x % groups, row = threadIdx.x / groups, rows = kB / groups;
    V8 s;
#pragma unroll
    for (int i = 0; i < 8; ++i) s.v[i] = 0.f;
    for (int p = blockIdx.x * rows + row; p < P; p += gridDim.x * rows) {
        const V8 v = ld8(dy + (size_t)p * C + 8 * cg);
#pragma unroll
        for (int i = 0; i < 8; ++i) s.v[i] += v.v[i];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) sh[threadIdx.x][i] = s.v[i];
    __syncthreads();
    if ((int)threadIdx.x < C) {
        const int g = threadIdx.x / 8, i = threadIdx.x % 8;
        float t = 0.f;
        for (int r = 0; r < rows; ++r) t += sh[r * groups + g][i];
        atomicAdd(&db[threadIdx.x], t);
    }
}

// bias gradient for any C (the 1-channel output layer): one block per channel
__global__ void k_bias_grad(const bf16 *dy, float *db, int P, int C)
{
    __shared__ float sh[kB];
    const int c = blockIdx.y;
    float s = 0.f;
    for (int p = blockIdx.x * kB + threadIdx.x; p < P; p += gridDim.x * kB) s += bf2f(dy[(size_t)p * C + c]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&db[c], sh[0]);
}

// dPre = dY * (Y > 0); n8 = elements / 8
__global__ void k_relu_grad(const bf16 *dy, const bf16 *y, bf16 *out, size_t n8)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    if (k >= n8) return;
    const uint4 g = reinterpret_cast<const uint4 *>(dy)[k], a = reinterpret_cast<const uint4 *>(y)[k];
    const uint32_t gw[4] = {g.x, g.y, g.z, g.w}, aw[4] = {a.x, a.y, a.z, a.w};
    uint32_t o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t lo = __uint_as_float(aw[i] << 16) > 0.f ? 0x0000FFFFu : 0u;
        const uint32_t hi = __uint_as_float(aw[i] & 0xFFFF0000u) > 0.f ? 0xFFFF0000u : 0u;
        o[i] = gw[i] & (lo | hi);
    }
    reinterpret_cast<uint4 *>(out)[k] = make_uint4(o[0], o[1], o[2], o[3]);
}

// 2x2 average pool (NHWC); H, W even; one thread per output pixel and 8 channels
__global__ void k_pool(const bf16 *x, bf16 *y, int B, int H, int W, int C)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    if (k >= (size_t)B * Ho * Wo * C8) return;
    const int c = (int)(k % C8) * 8;
    size_t r = k / C8;
    const int xo = (int)(r % Wo); r /= Wo;
    const int yo = (int)(r % Ho);
    const int b = (int)(r / Ho);
    const bf16 *q = x + (((size_t)b * H + 2 * yo) * W + 2 * xo) * C + c;
    const V8 a = ld8(q), bb = ld8(q + C), cc = ld8(q + (size_t)W * C), d = ld8(q + (size_t)W * C + C);
    V8 o;
#pragma unroll
    for (int i = 0; i < 8; ++i) o.v[i] = 0.25f * (a.v[i] + bb.v[i] + cc.v[i] + d.v[i]);
    st8(y + k * 8, o);
}

__global__ void k_pool_grad(const bf16 *dy, bf16 *dx, int B, int H, int W, int C)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    const int C8 = C / 8;
    if (k >= (size_t)B * H * W * C8) return;
    const int c = (int)(k % C8) * 8;
    size_t r = k / C8;
    const int xx = (int)(r % W); r /= W;
    const int yy = (int)(r % H);
    const int b = (int)(r / H);
    V8 o = ld8(dy + (((size_t)b * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c);
#pragma unroll
    for (int i = 0; i < 8; ++i) o.v[i] *= 0.25f;
    st8(dx + k * 8, o);
}

// nearest x2 upsample plus skip: y = up(x) + s
__global__ void k_up_add(const bf16 *x, const bf16 *s, bf16 *y, int B, int H, int W, int C)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    const int C8 = C / 8;
    if (k >= (size_t)B * H * W * C8) return;
    const int c = (int)(k % C8) * 8;
    size_t r = k / C8;
    const int xx = (int)(r % W); r /= W;
    const int yy = (int)(r % H);
    const int b = (int)(r / H);
    V8 o = ld8(x + (((size_t)b * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c);
    const V8 t = ld8(s + k * 8);
#pragma unroll
    for (int i = 0; i < 8; ++i) o.v[i] += t.v[i];
    st8(y + k * 8, o);
}

// gradient of up(x): dx[y][x] = sum of the 2x2 block of dy
__global__ void k_up_grad(const bf16 *dy, bf16 *dx, int B, int Hs, int Ws, int C)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    const int C8 = C / 8;
    if (k >= (size_t)B * Hs * Ws * C8) return;
    const int c = (int)(k % C8) * 8;
    size_t r = k / C8;
    const int xx = (int)(r % Ws); r /= Ws;
    const int yy = (int)(r % Hs);
    const int b = (int)(r / Hs);
    const int W = 2 * Ws;
    const bf16 *q = dy + (((size_t)b * 2 * Hs + 2 * yy) * W + 2 * xx) * C + c;
    const V8 a = ld8(q), bb = ld8(q + C), cc = ld8(q + (size_t)W * C), d = ld8(q + (size_t)W * C + C);
    V8 o;
#pragma unroll
    for (int i = 0; i < 8; ++i) o.v[i] = a.v[i] + bb.v[i] + cc.v[i] + d.v[i];
    st8(dx + k * 8, o);
}

// y = a + b; n8 = elements / 8
__global__ void k_add_bf(const bf16 *a, const bf16 *b, bf16 *y, size_t n8)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    if (k >= n8) return;
    V8 o = ld8(a + 8 * k);
    const V8 t = ld8(b + 8 * k);
#pragma unroll
    for (int i = 0; i < 8; ++i) o.v[i] += t.v[i];
    st8(y + 8 * k, o);
}

__global__ void k_to_bf(const float *x, bf16 *y, size_t n, float scale)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    if (k < n) y[k] = f2bf(x[k] * scale);
}

// MSE loss of the fp32 output against the target; gradient 2 (o - t) / n in bf16
// (grid-stride over at most kMseBlocks blocks: one atomic per block)
constexpr int kMseBlocks = 1024;
__global__ void k_mse(const float *o, const float *t, bf16 *g, float *loss, size_t n)
{
    __shared__ float sh[kB];
    float acc = 0.f;
    for (size_t k = (size_t)blockIdx.x * kB + threadIdx.x; k < n; k += (size_t)gridDim.x * kB) {
        const float d = o[k] - t[k];
        g[k] = f2bf(2.f * d / (float)n);
        acc += d * d;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, sh[0] / (float)n);
}

// Adam on the fp32 master weights; refresh the bf16 copies (padded rows stay 0)
__global__ void k_adam(float *w, float *m, float *v, const float *g, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    if (k >= n) return;
    const float gg = g[k];
    const float mm = b1 * m[k] + (1.f - b1) * gg;
    const float vv = b2 * v[k] + (1.f - b2) * gg * gg;
    m[k] = mm;
    v[k] = vv;
    w[k] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// bf16 copies of one layer: wf [Cout][9][Cin] -> wb [CoutP][9][Cin] and the
// transposed wt [CinP][9][Cout] (wt[ci][tap][co] = w[co][tap][ci]) used by the
// input gradient; padded rows are zero.
__global__ void k_wcopy(const float *wf, bf16 *wb, bf16 *wt, int Cin, int Cout, int CoutP, int CinP)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    const size_t nb = (size_t)CoutP * 9 * Cin, nt = (size_t)CinP * 9 * Cout;
    if (k < nb) {
        const int co = (int)(k / (9 * Cin));
        wb[k] = co < Cout ? f2bf(wf[k]) : (bf16)0;
    }
    if (k < nt) {
        const int ci = (int)(k / (9 * Cout)), rem = (int)(k % (9 * Cout)), tap = rem / Cout, co = rem % Cout;
        wt[k] = ci < Cin ? f2bf(wf[((size_t)co * 9 + tap) * Cin + ci]) : (bf16)0;
    }
}

int nb(size_t n) { return (int)((n + kB - 1) / kB); }
int pad16(int c) { return (c + 15) / 16 * 16; }

struct Layer {
    int cin, cout, relu;
    size_t woff, boff;       // offsets into the fp32 parameter vector
};

}  // namespace unet

using namespace unet;

struct csrst_warmstart {
    std::string err;
    hipStream_t s = nullptr;
    int H = 0, W = 0, base = 32, maxBatch = 0;
    std::vector<Layer> L;          // 15 convolutions
    size_t nparam = 0;
    DBuf<float> w, m, v, g;        // master weights, Adam moments, gradients
    DBuf<bf16> wb, wt;             // bf16 copies per layer (offsets below)
    std::vector<size_t> wbOff, wtOff;
    DBuf<bf16> act;                // activations and gradients arena
    DBuf<float> outf, tgt, scratchF, lossBuf;
    int steps = 0;
    // activation offsets
    struct Buf { size_t off; int h, w, c; };
    std::vector<Buf> A;            // forward activations
    std::vector<Buf> Gd;           // gradient buffers (same shapes)
    int batch = 0;
    // the multigrid of the grid H x W and the buffers of the network-in-the-loop
    // solve and of its training pairs
    struct Loop {
        pmg::Solver mg;
        bool ready = false;
        double hx = 0, hy = 0;
        std::vector<uint8_t> cls;      // host copy of the node classes of the hierarchy
        size_t unknowns = 0;
        DBuf<uint8_t> dcls;
        DBuf<double> b, phi, pstar, part;
        DBuf<double> pprev, rprev;     // the previous iterate and its residual (two-direction step)
        DBuf<bf16> pairIn;             // [maxBatch][H][W] r / rms(r)
        DBuf<float> pairTgt;           // [maxBatch][H][W] e / rms(e)
    } loop;

    bf16 *ptr(const Buf &b) { return act.p + b.off; }
};

namespace {

// conv indices in order: 0 c00 1->32, 1 c01, 2 c10 32->64, 3 c11, 4 c20 64->128,
// 5 c21, 6 c30 128->256, 7 c31, 8 c22 256->128, 9 c23, 10 c12 128->64, 11 c13,
// 12 c02 64->32, 13 c03, 14 out 32->1
void make_layers(csrst_warmstart *u)
{
    const int b = u->base;
    const int io[15][3] = {{1, b, 1},          {b, b, 1},          {b, 2 * b, 1},     {2 * b, 2 * b, 1},
                           {2 * b, 4 * b, 1},  {4 * b, 4 * b, 1},  {4 * b, 8 * b, 1}, {8 * b, 8 * b, 1},
                           {8 * b, 4 * b, 1},  {4 * b, 4 * b, 1},  {4 * b, 2 * b, 1}, {2 * b, 2 * b, 1},
                           {2 * b, b, 1},      {b, b, 1},          {b, 1, 0}};
    size_t off = 0;
    u->L.clear();
    for (auto &e : io) {
        Layer l;
        l.cin = e[0]; l.cout = e[1]; l.relu = e[2];
        l.woff = off; off += (size_t)l.cout * 9 * l.cin;
        l.boff = off; off += l.cout;
        u->L.push_back(l);
    }
    u->nparam = off;
}

void conv(csrst_warmstart *u, int li, const bf16 *x, bf16 *y, float *yf, int B, int H, int W)
{
    const Layer &l = u->L[li];
    ConvArgs a;
    a.x = x; a.w = u->wb.p + u->wbOff[li]; a.bias = u->w.p + l.boff; a.y = y; a.yf = yf;
    a.B = B; a.H = H; a.W = W; a.Cin = l.cin; a.Cout = l.cout; a.CoutP = pad16(l.cout); a.flip = 0; a.relu = l.relu;
    launch_conv(a, u->s);
}

// input gradient of conv li: dx[P][cin] from dpre[P][cout]
void conv_dgrad(csrst_warmstart *u, int li, const bf16 *dpre, bf16 *dx, int B, int H, int W)
{
    const Layer &l = u->L[li];
    ConvArgs a;
    a.x = dpre; a.w = u->wt.p + u->wtOff[li]; a.bias = nullptr; a.y = dx; a.yf = nullptr;
    a.B = B; a.H = H; a.W = W; a.Cin = l.cout; a.Cout = l.cin; a.CoutP = pad16(l.cin); a.flip = 1; a.relu = 0;
    launch_conv(a, u->s);
}

void conv_wgrad(csrst_warmstart *u, int li, const bf16 *dpre, const bf16 *x, int B, int H, int W)
{
    const Layer &l = u->L[li];
    const int P = B * H * W;
    const int chunk = 1024;
    dim3 grid((9 * l.cin + 63) / 64, pad16(l.cout) / 16, (P + chunk - 1) / chunk);
    if (l.cin % 32 == 0 && l.cout % 32 == 0) {
        const int tiles = B * ((H + kWgTH - 1) / kWgTH) * ((W + kWgTW - 1) / kWgTW);
        auto go = [&](auto kern, int coBlk, int ciBlk) {
            const int nBlk = (l.cout / coBlk) * (l.cin / ciBlk);
            const int ksplit = std::max(1, std::min(tiles, 1024 / nBlk));
            kern<<<nBlk * ksplit, kB, 0, u->s>>>(dpre, x, u->g.p + l.woff, B, H, W, l.cin, l.cout, ksplit);
        };
        if (l.cin % 64 == 0) go(k_wgrad32<2, 1>, 32, 64);
        else if (l.cout % 64 == 0) go(k_wgrad32<2, 2>, 64, 32);
        else go(k_wgrad32<1, 2>, 32, 32);
    } else if ((l.cin == 1 && l.cout == 32) || (l.cin == 32 && l.cout == 1)) {
        const int tiles = B * ((H + kWgTH - 1) / kWgTH) * ((W + kWgTW - 1) / kWgTW);
        if (l.cout == 1) k_wgrad_one<true><<<std::min(tiles, 1024), kB, 0, u->s>>>(dpre, x, u->g.p + l.woff, B, H, W);
        else k_wgrad_one<false><<<std::min(tiles, 1024), kB, 0, u->s>>>(dpre, x, u->g.p + l.woff, B, H, W);
    } else
        k_conv_wgrad<<<grid, 64, 0, u->s>>>(dpre, x, u->g.p + l.woff, B, H, W, l.cin, l.cout, chunk);
    if (l.cout % 8 == 0 && kB % (l.cout / 8) == 0) {
        const int rows = kB / (l.cout / 8);
        k_bias_grad8<<<std::min(1024, (P + rows - 1) / rows), kB, 0, u->s>>>(dpre, u->g.p + l.boff, P, l.cout);
    } else {
        dim3 bg(std::min(64, (P + kB - 1) / kB), l.cout);
        k_bias_grad<<<bg, kB, 0, u->s>>>(dpre, u->g.p + l.boff, P, l.cout);
    }
}

void refresh_bf16(csrst_warmstart *u)
{
    for (size_t i = 0; i < u->L.size(); ++i) {
        const Layer &l = u->L[i];
        const size_t n = std::max((size_t)pad16(l.cout) * 9 * l.cin, (size_t)pad16(l.cin) * 9 * l.cout);
        k_wcopy<<<nb(n), kB, 0, u->s>>>(u->w.p + l.woff, u->wb.p + u->wbOff[i], u->wt.p + u->wtOff[i], l.cin, l.cout,
                                         pad16(l.cout), pad16(l.cin));
    }
    HU_LAUNCH_CHECK();
}

// activation plan for batch B: forward buffers A[0..] and gradient buffers
void plan(csrst_warmstart *u, int B)
{
    if (u->batch == B) return;
    const int H = u->H, W = u->W, b = u->base;
    u->A.clear();
    size_t off = 0;
    auto add = [&](std::vector<csrst_warmstart::Buf> &v, int h, int w, int c) {
        v.push_back({off, h, w, c});
        off += (size_t)B * h * w * c;
        off = (off + 63) / 64 * 64;
    };
    // forward: 0 in, 1 c00, 2 s0(c01), 3 p0, 4 c10, 5 s1(c11), 6 p1, 7 c20, 8 s2(c21), 9 p2, 10 c30, 11 c31,
    // 12 c22, 13 u2(up+s2), 14 c23, 15 c12, 16 u1, 17 c13, 18 c02, 19 u0, 20 c03, 21 out(bf16)
    const int sh[22][3] = {{H, W, 1},          {H, W, b},          {H, W, b},          {H / 2, W / 2, b},
                           {H / 2, W / 2, 2 * b}, {H / 2, W / 2, 2 * b}, {H / 4, W / 4, 2 * b}, {H / 4, W / 4, 4 * b},
                           {H / 4, W / 4, 4 * b}, {H / 8, W / 8, 4 * b}, {H / 8, W / 8, 8 * b}, {H / 8, W / 8, 8 * b},
                           {H / 8, W / 8, 4 * b}, {H / 4, W / 4, 4 * b}, {H / 4, W / 4, 4 * b}, {H / 4, W / 4, 2 * b},
                           {H / 2, W / 2, 2 * b}, {H / 2, W / 2, 2 * b}, {H / 2, W / 2, b},     {H, W, b},
                           {H, W, b},          {H, W, 1}};
    for (auto &e : sh) add(u->A, e[0], e[1], e[2]);
    u->Gd.clear();
    for (auto &e : sh) add(u->Gd, e[0], e[1], e[2]);
    // two scratch gradient buffers of the largest activation
    size_t big = 0;
    for (auto &e : sh) big = std::max(big, (size_t)B * e[0] * e[1] * e[2]);
    u->Gd.push_back({off, 0, 0, 0}); off += (big + 63) / 64 * 64;
    u->Gd.push_back({off, 0, 0, 0}); off += (big + 63) / 64 * 64;
    // grow only: a smaller batch reuses the arena with its own offsets
    if (off > u->act.n) u->act.alloc(off);
    if ((size_t)B * H * W > u->outf.n) u->outf.alloc((size_t)B * H * W);
    if ((size_t)B * H * W > u->tgt.n) u->tgt.alloc((size_t)B * H * W);
    u->batch = B;
}

void forward(csrst_warmstart *u, int B)
{
    auto &A = u->A;
    const int H = u->H, W = u->W;
    auto P = [&](int i) { return u->ptr(A[i]); };
    auto pool = [&](int src, int dst) {
        const auto &d = A[dst];
        k_pool<<<nb((size_t)B * d.h * d.w * d.c / 8), kB, 0, u->s>>>(P(src), P(dst), B, A[src].h, A[src].w, d.c);
    };
    auto upadd = [&](int src, int skip, int dst) {
        const auto &d = A[dst];
        k_up_add<<<nb((size_t)B * d.h * d.w * d.c / 8), kB, 0, u->s>>>(P(src), P(skip), P(dst), B, d.h, d.w, d.c);
    };
    conv(u, 0, P(0), P(1), nullptr, B, H, W);
    conv(u, 1, P(1), P(2), nullptr, B, H, W);
    pool(2, 3);
    conv(u, 2, P(3), P(4), nullptr, B, H / 2, W / 2);
    conv(u, 3, P(4), P(5), nullptr, B, H / 2, W / 2);
    pool(5, 6);
    conv(u, 4, P(6), P(7), nullptr, B, H / 4, W / 4);
    conv(u, 5, P(7), P(8), nullptr, B, H / 4, W / 4);
    pool(8, 9);
    conv(u, 6, P(9), P(10), nullptr, B, H / 8, W / 8);
    conv(u, 7, P(10), P(11), nullptr, B, H / 8, W / 8);
    conv(u, 8, P(11), P(12), nullptr, B, H / 8, W / 8);
    upadd(12, 8, 13);
    conv(u, 9, P(13), P(14), nullptr, B, H / 4, W / 4);
    conv(u, 10, P(14), P(15), nullptr, B, H / 4, W / 4);
    upadd(15, 5, 16);
    conv(u, 11, P(16), P(17), nullptr, B, H / 2, W / 2);
    conv(u, 12, P(17), P(18), nullptr, B, H / 2, W / 2);
    upadd(18, 2, 19);
    conv(u, 13, P(19), P(20), nullptr, B, H, W);
    conv(u, 14, P(20), P(21), u->outf.p, B, H, W);
    HU_LAUNCH_CHECK();
}

// backward from the bf16 output gradient in Gd[21]; fills u->g
void backward(csrst_warmstart *u, int B)
{
    auto &A = u->A;
    auto &G = u->Gd;
    const int H = u->H, W = u->W;
    auto P = [&](int i) { return u->ptr(A[i]); };
    auto D = [&](int i) { return u->ptr(G[i]); };
    bf16 *pre = u->ptr(G[22]);
    HU_CHECK(hipMemsetAsync(u->g.p, 0, u->nparam * sizeof(float), u->s));
    // conv li with input activation xin and output activation yout; dy given in D(yout)
    auto back = [&](int li, int xin, int yout, int h, int w, bool needDx) {
        const Layer &l = u->L[li];
        const size_t n = (size_t)B * h * w * l.cout;
        const bf16 *dpre = D(yout);
        if (l.relu) {
            k_relu_grad<<<nb(n / 8), kB, 0, u->s>>>(D(yout), P(yout), pre, n / 8);
            dpre = pre;
        }
        conv_wgrad(u, li, dpre, P(xin), B, h, w);
        if (needDx) conv_dgrad(u, li, dpre, D(xin), B, h, w);
    };
    auto addgrad = [&](int dst, const bf16 *src) {   // D(dst) += src
        const auto &d = A[dst];
        const size_t n = (size_t)B * d.h * d.w * d.c;
        k_add_bf<<<nb(n / 8), kB, 0, u->s>>>(D(dst), src, D(dst), n / 8);
    };
    auto upgrad = [&](int dstLow, int srcHigh) {      // D(dstLow) = up^T D(srcHigh)
        const auto &d = A[dstLow];
        k_up_grad<<<nb((size_t)B * d.h * d.w * d.c / 8), kB, 0, u->s>>>(D(srcHigh), D(dstLow), B, d.h, d.w, d.c);
    };
    auto poolgrad = [&](bf16 *dst, int srcLow, int hiIdx) {   // dst = pool^T D(srcLow)
        const auto &d = A[hiIdx];
        k_pool_grad<<<nb((size_t)B * d.h * d.w * d.c / 8), kB, 0, u->s>>>(D(srcLow), dst, B, d.h, d.w, d.c);
    };
    bf16 *tmp = u->ptr(G[23]);
    back(14, 20, 21, H, W, true);      // out -> D20
    back(13, 19, 20, H, W, true);      // -> D19 (= d(up(c02)) and d(s0))
    upgrad(18, 19);                    // D18 = up^T D19
    back(12, 17, 18, H / 2, W / 2, true);
    back(11, 16, 17, H / 2, W / 2, true);
    upgrad(15, 16);
    back(10, 14, 15, H / 4, W / 4, true);
    back(9, 13, 14, H / 4, W / 4, true);
    upgrad(12, 13);
    back(8, 11, 12, H / 8, W / 8, true);
    back(7, 10, 11, H / 8, W / 8, true);
    back(6, 9, 10, H / 8, W / 8, true);
    // D8 = pool^T D9 + D13 (skip)
    poolgrad(D(8), 9, 8);
    addgrad(8, D(13));
    back(5, 7, 8, H / 4, W / 4, true);
    back(4, 6, 7, H / 4, W / 4, true);
    poolgrad(D(5), 6, 5);
    addgrad(5, D(16));
    back(3, 4, 5, H / 2, W / 2, true);
    back(2, 3, 4, H / 2, W / 2, true);
    poolgrad(D(2), 3, 2);
    addgrad(2, D(19));
    back(1, 1, 2, H, W, true);
    back(0, 0, 1, H, W, false);
    (void)tmp;
    HU_LAUNCH_CHECK();
}

void upload_input(csrst_warmstart *u, const float *rho, int B, float scale)
{
    const size_t n = (size_t)B * u->H * u->W;
    u->scratchF.alloc(std::max(u->scratchF.n, n));
    HU_CHECK(hipMemcpyAsync(u->scratchF.p, rho, n * sizeof(float), hipMemcpyHostToDevice, u->s));
    k_to_bf<<<nb(n), kB, 0, u->s>>>(u->scratchF.p, u->ptr(u->A[0]), n, scale);
    HU_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// The network between V-cycles.  At an iterate phi_k of the multigrid the
// residual r_k = b - A phi_k, scaled to unit RMS over the unknowns, goes
// through the network; its output c is taken as the direction of the error
// e_k = A^-1 r_k and the iterate moves by the step that minimises the energy
// norm of the error along it: phi += alpha c, alpha = (c . r) / (c . A c).  A
// V-cycle follows.  Training pairs are (r_k / rms(r_k), e_k / rms(e_k)) over
// the iterates of the solver itself, with e_k from a solve to round-off.

constexpr int kRed = 256;      // blocks of the fixed-order reductions

// r * inv as bf16 network input, zero off the unknowns
__global__ void k_res_in(const double *res, const uint8_t *act, bf16 *in, size_t n, double inv)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    if (k < n) in[k] = act[k] ? f2bf((float)(res[k] * inv)) : (bf16)0;
}

// (pstar - phi) * inv as float target, zero off the unknowns
__global__ void k_err_tgt(const double *pstar, const double *phi, const uint8_t *act, float *t, size_t n, double inv)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    if (k < n) t[k] = act[k] ? (float)((pstar[k] - phi[k]) * inv) : 0.f;
}

__device__ __forceinline__ void block_sum2(double &a, double &b, double *part)
{
    __shared__ double sa[kB], sb[kB];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int o = kB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sa[threadIdx.x] += sa[threadIdx.x + o];
            sb[threadIdx.x] += sb[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sa[0];
        part[2 * blockIdx.x + 1] = sb[0];
    }
}

// partial sums of c . r and c . A c over the unknowns (A the level's 5-point
// operator with the unknowns' couplings only, as the multigrid applies it)
__global__ void __launch_bounds__(kB) k_step_dots(pmg::Level L, const float *c, double *part)
{
    double cr = 0.0, cac = 0.0;
    const size_t n = (size_t)L.nx * L.ny;
    for (size_t k = (size_t)blockIdx.x * kB + threadIdx.x; k < n; k += (size_t)gridDim.x * kB) {
        if (!L.act[k]) continue;
        const int i = (int)(k % L.nx), j = (int)(k / L.nx);
        const double ck = c[k];
        double sx = 0.0, sy = 0.0;
        if (i > 0 && L.act[k - 1]) sx += c[k - 1];
        if (i + 1 < L.nx && L.act[k + 1]) sx += c[k + 1];
        if (j > 0 && L.act[k - L.nx]) sy += c[k - L.nx];
        if (j + 1 < L.ny && L.act[k + L.nx]) sy += c[k + L.nx];
        cr += ck * L.res[k];
        cac += ck * (2.0 * (L.cx + L.cy) * ck - L.cx * sx - L.cy * sy);
    }
    block_sum2(cr, cac, part);
}

// partial sums of (pstar - phi)^2 over the unknowns (second slot unused)
__global__ void __launch_bounds__(kB) k_err_sumsq(const double *pstar, const double *phi, const uint8_t *act, size_t n,
                                                 double *part)
{
    double s = 0.0, z = 0.0;
    for (size_t k = (size_t)blockIdx.x * kB + threadIdx.x; k < n; k += (size_t)gridDim.x * kB)
        if (act[k]) {
            const double d = pstar[k] - phi[k];
            s += d * d;
        }
    block_sum2(s, z, part);
}

// The two-direction step: partial sums over the unknowns of c.Ac, c.Ad, d.Ad,
// c.r, d.r with d = phi - pprev (the previous iteration's change) and
// A d = rprev - r; c may be null (d alone).  Five doubles per block.
__global__ void __launch_bounds__(kB) k_step_dots5(pmg::Level L, const float *c, const double *pprev,
                                                   const double *rprev, double *part)
{
    __shared__ double sh[5][kB];
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const size_t n = (size_t)L.nx * L.ny;
    for (size_t k = (size_t)blockIdx.x * kB + threadIdx.x; k < n; k += (size_t)gridDim.x * kB) {
        if (!L.act[k]) continue;
        const double r = L.res[k], d = L.phi[k] - pprev[k], ad = rprev[k] - r;
        double ck = 0.0, ack = 0.0;
        if (c) {
            const int i = (int)(k % L.nx), j = (int)(k / L.nx);
            ck = c[k];
            double sx = 0.0, sy = 0.0;
            if (i > 0 && L.act[k - 1]) sx += c[k - 1];
            if (i + 1 < L.nx && L.act[k + 1]) sx += c[k + 1];
            if (j > 0 && L.act[k - L.nx]) sy += c[k - L.nx];
            if (j + 1 < L.ny && L.act[k + L.nx]) sy += c[k + L.nx];
            ack = 2.0 * (L.cx + L.cy) * ck - L.cx * sx - L.cy * sy;
        }
        s[0] += ck * ack;
        s[1] += ck * ad;
        s[2] += d * ad;
        s[3] += ck * r;
        s[4] += d * r;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int o = kB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
#pragma unroll
            for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 5) part[5 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_sum5_final(double *part, int nb)
{
    if (threadIdx.x >= 5 || blockIdx.x != 0) return;
    double a = 0.0;
    for (int k = 0; k < nb; ++k) a += part[5 * k + threadIdx.x];
    part[5 * nb + threadIdx.x] = a;
}

// phi += alpha c + beta (phi - pprev) on the unknowns; pprev, rprev take the
// iterate and residual before the step
__global__ void k_update2(pmg::Level L, const float *c, double alpha, double beta, double *pprev, double *rprev)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    if (k >= (size_t)L.nx * L.ny) return;
    if (!L.act[k]) {
        pprev[k] = 0.0;
        rprev[k] = 0.0;
        return;
    }
    const double ph = L.phi[k], d = ph - pprev[k];
    pprev[k] = ph;
    rprev[k] = L.res[k];
    L.phi[k] = ph + (c ? alpha * (double)c[k] : 0.0) + beta * d;
}

__global__ void k_sum2_final(double *part, int nb)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0, b = 0.0;
    for (int k = 0; k < nb; ++k) {
        a += part[2 * k];
        b += part[2 * k + 1];
    }
    part[2 * nb] = a;
    part[2 * nb + 1] = b;
}

__global__ void k_axpy_act(pmg::Level L, const float *c, double alpha)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    if (k < (size_t)L.nx * L.ny && L.act[k]) L.phi[k] += alpha * (double)c[k];
}

__global__ void k_copy_d(double *dst, const double *src, size_t n)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

__global__ void k_zero_d(double *p, size_t n)
{
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    if (k < n) p[k] = 0.0;
}

int ws_fail(csrst_warmstart *u, const std::exception &e)
{
    const Error *he = dynamic_cast<const Error *>(&e);
    if (u) u->err = e.what();
    hu::global_error() = e.what();
    return he ? he->code : CSRST_EINVAL;
}


// The hierarchy of the grid (hx, hy, cls), set up once and kept while the
// caller passes the same grid.
void loop_grid(csrst_warmstart *u, double hx, double hy, const uint8_t *cls)
{
    auto &Lp = u->loop;
    const size_t n = (size_t)u->H * u->W;
    if (!cls || !(hx > 0) || !(hy > 0)) throw Error(CSRST_EINVAL, "bad grid for the warm-start solve");
    if (Lp.ready && Lp.hx == hx && Lp.hy == hy && std::equal(Lp.cls.begin(), Lp.cls.end(), cls)) return;
    size_t unk = 0;
    for (size_t k = 0; k < n; ++k) {
        if (cls[k] > pmg::kKnown) throw Error(CSRST_EINVAL, hu::fmt("node class %d at %zu", cls[k], k));
        unk += cls[k] == pmg::kUnknown;
    }
    if (!unk) throw Error(CSRST_EINVAL, "the grid has no unknown nodes");
    Lp.ready = false;
    Lp.cls.assign(cls, cls + n);
    Lp.dcls.alloc(n);
    Lp.b.alloc(n); Lp.phi.alloc(n); Lp.pstar.alloc(n); Lp.part.alloc(5 * kRed + 5);
    Lp.pprev.alloc(n); Lp.rprev.alloc(n);
    HU_CHECK(hipMemcpyAsync(Lp.dcls.p, cls, n, hipMemcpyHostToDevice, u->s));
    Lp.mg.setup(u->W, u->H, hx, hy, Lp.dcls.p, u->s);
    Lp.hx = hx; Lp.hy = hy; Lp.unknowns = unk;
    Lp.ready = true;
}

// the two sums of the last k_step_dots / k_err_sumsq launch on the host
void loop_sums(csrst_warmstart *u, double &a, double &b)
{
    auto &Lp = u->loop;
    k_sum2_final<<<1, 1, 0, u->s>>>(Lp.part.p, kRed);
    HU_LAUNCH_CHECK();
    double h[2];
    HU_CHECK(hipMemcpyAsync(h, Lp.part.p + 2 * kRed, sizeof h, hipMemcpyDeviceToHost, u->s));
    HU_CHECK(hipStreamSynchronize(u->s));
    a = h[0];
    b = h[1];
}

// The network's step at the current iterate; rn = ||r|| from resnorm().
void loop_correct(csrst_warmstart *u, double rn)
{
    auto &Lp = u->loop;
    const pmg::Level &L0 = Lp.mg.level0();
    const size_t n = (size_t)u->H * u->W;
    const double rms = rn / std::sqrt((double)Lp.unknowns);
    if (!(rms > 0)) return;
    plan(u, 1);
    k_res_in<<<nb(n), kB, 0, u->s>>>(L0.res, L0.act, u->ptr(u->A[0]), n, 1.0 / rms);
    forward(u, 1);
    k_step_dots<<<kRed, kB, 0, u->s>>>(L0, u->outf.p, Lp.part.p);
    double cr = 0, cac = 0;
    loop_sums(u, cr, cac);
    if (cac > 0 && std::isfinite(cr / cac)) {
        k_axpy_act<<<nb(n), kB, 0, u->s>>>(L0, u->outf.p, cr / cac);
        HU_LAUNCH_CHECK();
    }
}

// The step over the network's direction (useNet) and the previous iteration's
// change: the energy-norm minimiser over span{c, d}, a 2 x 2 system (one
// direction when the other is absent or degenerate).  first: no previous
// iterate yet (d = 0).
void loop_step2(csrst_warmstart *u, double rn, bool useNet, bool first)
{
    auto &Lp = u->loop;
    const pmg::Level &L0 = Lp.mg.level0();
    const size_t n = (size_t)u->H * u->W;
    const double rms = rn / std::sqrt((double)Lp.unknowns);
    if (!(rms > 0)) return;
    if (first) {
        k_copy_d<<<nb(n), kB, 0, u->s>>>(Lp.pprev.p, L0.phi, n);
        k_copy_d<<<nb(n), kB, 0, u->s>>>(Lp.rprev.p, L0.res, n);
    }
    const float *c = nullptr;
    if (useNet) {
        plan(u, 1);
        k_res_in<<<nb(n), kB, 0, u->s>>>(L0.res, L0.act, u->ptr(u->A[0]), n, 1.0 / rms);
        forward(u, 1);
        c = u->outf.p;
    }
    k_step_dots5<<<kRed, kB, 0, u->s>>>(L0, c, Lp.pprev.p, Lp.rprev.p, Lp.part.p);
    k_sum5_final<<<1, 64, 0, u->s>>>(Lp.part.p, kRed);
    HU_LAUNCH_CHECK();
    double h[5];
    HU_CHECK(hipMemcpyAsync(h, Lp.part.p + 5 * kRed, sizeof h, hipMemcpyDeviceToHost, u->s));
    HU_CHECK(hipStreamSynchronize(u->s));
    const double cac = h[0], cad = h[1], dad = h[2], cr = h[3], dr = h[4];
    double alpha = 0.0, beta = 0.0;
    const bool hc = c && cac > 0, hd = dad > 0;
    const double det = cac * dad - cad * cad;
    if (hc && hd && det > 1e-12 * cac * dad) {
        alpha = (cr * dad - cad * dr) / det;
        beta = (cac * dr - cad * cr) / det;
    } else if (hc) {
        alpha = cr / cac;
    } else if (hd) {
        beta = dr / dad;
    }
    if (!std::isfinite(alpha) || !std::isfinite(beta)) alpha = beta = 0.0;
    k_update2<<<nb(n), kB, 0, u->s>>>(L0, c, alpha, beta, Lp.pprev.p, Lp.rprev.p);
    HU_LAUNCH_CHECK();
}

// One Adam step on the first B staged pairs; returns the loss before it.
float loop_train(csrst_warmstart *u, int B, float lr)
{
    auto &Lp = u->loop;
    const size_t np = (size_t)B * u->H * u->W;
    plan(u, B);
    HU_CHECK(hipMemcpyAsync(u->ptr(u->A[0]), Lp.pairIn.p, np * sizeof(bf16), hipMemcpyDeviceToDevice, u->s));
    HU_CHECK(hipMemcpyAsync(u->tgt.p, Lp.pairTgt.p, np * sizeof(float), hipMemcpyDeviceToDevice, u->s));
    forward(u, B);
    HU_CHECK(hipMemsetAsync(u->lossBuf.p, 0, sizeof(float), u->s));
    k_mse<<<std::min(nb(np), kMseBlocks), kB, 0, u->s>>>(u->outf.p, u->tgt.p, u->ptr(u->Gd[21]), u->lossBuf.p, np);
    backward(u, B);
    ++u->steps;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)u->steps), c2 = 1.f - std::pow(b2, (float)u->steps);
    k_adam<<<nb(u->nparam), kB, 0, u->s>>>(u->w.p, u->m.p, u->v.p, u->g.p, u->nparam, lr, b1, b2, 1e-8f, c1, c2);
    refresh_bf16(u);
    float loss = 0.f;
    HU_CHECK(hipMemcpyAsync(&loss, u->lossBuf.p, sizeof(float), hipMemcpyDeviceToHost, u->s));
    HU_CHECK(hipStreamSynchronize(u->s));
    return loss;
}

}  // namespace

extern "C" {

int csrst_warmstart_create(int H, int W, int maxBatch, csrst_warmstart **out)
{
    std::unique_ptr<csrst_warmstart> u(new csrst_warmstart());
    try {
        if (!out || H < 16 || W < 16 || H % 8 || W % 8 || maxBatch < 1)
            throw Error(CSRST_EINVAL, "grid must be at least 16 x 16 and divisible by 8; batch >= 1");
        u->H = H; u->W = W; u->maxBatch = maxBatch;
        HU_CHECK(hipStreamCreateWithFlags(&u->s, hipStreamNonBlocking));
        make_layers(u.get());
        u->w.alloc(u->nparam); u->m.alloc(u->nparam); u->v.alloc(u->nparam); u->g.alloc(u->nparam);
        u->m.zero(u->s); u->v.zero(u->s); u->g.zero(u->s);
        size_t ob = 0, ot = 0;
        for (const Layer &l : u->L) {
            u->wbOff.push_back(ob); ob += (size_t)pad16(l.cout) * 9 * l.cin;
            u->wtOff.push_back(ot); ot += (size_t)pad16(l.cin) * 9 * l.cout;
        }
        u->wb.alloc(ob); u->wt.alloc(ot);
        u->lossBuf.alloc(1);
        // He initialisation from a fixed LCG (weights), zero biases
        std::vector<float> h(u->nparam, 0.f);
        uint64_t st = 0x9E3779B97F4A7C15ull;
        auto uni = [&]() {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            return ((st >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        };
        for (const Layer &l : u->L) {
            const double sd = std::sqrt(2.0 / (9.0 * l.cin));
            for (size_t k = 0; k < (size_t)l.cout * 9 * l.cin; ++k) {
                const double a = uni(), b = uni();
                h[l.woff + k] = (float)(sd * std::sqrt(-2 * std::log(a)) * std::cos(6.283185307179586 * b));
            }
        }
        HU_CHECK(hipMemcpyAsync(u->w.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, u->s));
        refresh_bf16(u.get());
        HU_CHECK(hipStreamSynchronize(u->s));
        *out = u.release();
        return CSRST_OK;
    } catch (const std::exception &e) {
        if (u && u->s) { (void)hipStreamSynchronize(u->s); (void)hipStreamDestroy(u->s); u->s = nullptr; }
        return ws_fail(nullptr, e);
    }
}

int csrst_warmstart_destroy(csrst_warmstart *u)
{
    if (!u) return CSRST_OK;
    if (u->s) { (void)hipStreamSynchronize(u->s); (void)hipStreamDestroy(u->s); }
    delete u;
    return CSRST_OK;
}

int csrst_warmstart_num_params(csrst_warmstart *u, int *n)
{
    if (!u || !n) return CSRST_EINVAL;
    *n = (int)u->nparam;
    return CSRST_OK;
}

int csrst_warmstart_get_weights(csrst_warmstart *u, float *w, int n)
{
    try {
        if (!u || !w || n != (int)u->nparam) throw Error(CSRST_EINVAL, "weight count mismatch");
        HU_CHECK(hipMemcpyAsync(w, u->w.p, u->nparam * sizeof(float), hipMemcpyDeviceToHost, u->s));
        HU_CHECK(hipStreamSynchronize(u->s));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return ws_fail(u, e);
    }
}

int csrst_warmstart_set_weights(csrst_warmstart *u, const float *w, int n)
{
    try {
        if (!u || !w || n != (int)u->nparam) throw Error(CSRST_EINVAL, "weight count mismatch");
        HU_CHECK(hipMemcpyAsync(u->w.p, w, u->nparam * sizeof(float), hipMemcpyHostToDevice, u->s));
        u->m.zero(u->s); u->v.zero(u->s);
        u->steps = 0;
        refresh_bf16(u);
        HU_CHECK(hipStreamSynchronize(u->s));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return ws_fail(u, e);
    }
}

// Inference: rho [B][H][W] (host, float) scaled by `scale`, output [B][H][W].
int csrst_warmstart_infer(csrst_warmstart *u, const float *rho, float *out, int B, float scale)
{
    try {
        if (!u || !rho || !out || B < 1 || B > u->maxBatch) throw Error(CSRST_EINVAL, "bad inference arguments");
        plan(u, B);
        upload_input(u, rho, B, scale);
        forward(u, B);
        HU_CHECK(hipMemcpyAsync(out, u->outf.p, (size_t)B * u->H * u->W * sizeof(float), hipMemcpyDeviceToHost, u->s));
        HU_CHECK(hipStreamSynchronize(u->s));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return ws_fail(u, e);
    }
}

// Loss and parameter gradients of one batch (no update): grads [nparam].
int csrst_warmstart_gradients(csrst_warmstart *u, const float *rho, const float *phi, int B, float scale, float *loss,
                              float *grads, int n)
{
    try {
        if (!u || !rho || !phi || B < 1 || B > u->maxBatch) throw Error(CSRST_EINVAL, "bad gradient arguments");
        if (grads && n != (int)u->nparam) throw Error(CSRST_EINVAL, "gradient count mismatch");
        plan(u, B);
        upload_input(u, rho, B, scale);
        const size_t np = (size_t)B * u->H * u->W;
        HU_CHECK(hipMemcpyAsync(u->tgt.p, phi, np * sizeof(float), hipMemcpyHostToDevice, u->s));
        forward(u, B);
        HU_CHECK(hipMemsetAsync(u->lossBuf.p, 0, sizeof(float), u->s));
        k_mse<<<std::min(nb(np), kMseBlocks), kB, 0, u->s>>>(u->outf.p, u->tgt.p, u->ptr(u->Gd[21]), u->lossBuf.p, np);
        backward(u, B);
        if (loss) HU_CHECK(hipMemcpyAsync(loss, u->lossBuf.p, sizeof(float), hipMemcpyDeviceToHost, u->s));
        if (grads) HU_CHECK(hipMemcpyAsync(grads, u->g.p, u->nparam * sizeof(float), hipMemcpyDeviceToHost, u->s));
        HU_CHECK(hipStreamSynchronize(u->s));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return ws_fail(u, e);
    }
}

// One Adam step on a batch; loss before the update.
int csrst_warmstart_train_step(csrst_warmstart *u, const float *rho, const float *phi, int B, float scale, float lr,
                               float *loss)
{
    try {
        const int rc = csrst_warmstart_gradients(u, rho, phi, B, scale, loss, nullptr, 0);
        if (rc) return rc;
        ++u->steps;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)u->steps), c2 = 1.f - std::pow(b2, (float)u->steps);
        k_adam<<<nb(u->nparam), kB, 0, u->s>>>(u->w.p, u->m.p, u->v.p, u->g.p, u->nparam, lr, b1, b2, 1e-8f, c1, c2);
        refresh_bf16(u);
        HU_CHECK(hipStreamSynchronize(u->s));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return ws_fail(u, e);
    }
}


// Poisson solve on the network's H x W node grid (nx = W, ny = H) with the
// multigrid of csrst_poisson_solve.  mode 0: plain V-cycles from phi; 1: the
// network's step before every V-cycle; 2: before the first only (a one-shot
// start); 3: before every V-cycle the energy-norm minimiser over the
// network's direction and the previous iteration's change; 4: over the
// previous change alone (the same acceleration without the network).  phi: in the guess, out the solution; cycles and the final relative
// residual as csrst_poisson_solve reports them.
int csrst_warmstart_solve(csrst_warmstart *u, double hx, double hy, const uint8_t *cls, const double *b, double *phi,
                          double tol, int maxit, int mode, int *cycles, double *resid)
{
    try {
        if (!u || !b || !phi || maxit < 0 || !(tol >= 0) || mode < 0 || mode > 4)
            throw Error(CSRST_EINVAL, "bad warm-start solve arguments");
        loop_grid(u, hx, hy, cls);
        auto &Lp = u->loop;
        const size_t n = (size_t)u->H * u->W;
        HU_CHECK(hipMemcpyAsync(Lp.b.p, b, n * sizeof(double), hipMemcpyHostToDevice, u->s));
        HU_CHECK(hipMemcpyAsync(Lp.phi.p, phi, n * sizeof(double), hipMemcpyHostToDevice, u->s));
        const double bn = Lp.mg.begin(Lp.b.p, Lp.phi.p);
        double r = Lp.mg.resnorm();
        const double target = tol * (bn > 0 ? bn : 1.0);
        int cyc = 0, flat = 0;
        while (r > target && cyc < maxit && flat < 3) {
            if (mode == 1 || (mode == 2 && cyc == 0)) loop_correct(u, r);
            else if (mode >= 3) loop_step2(u, r, mode == 3, cyc == 0);
            Lp.mg.cycle();
            ++cyc;
            const double rn = Lp.mg.resnorm();
            if (!std::isfinite(rn)) throw Error(CSRST_EHIP, "multigrid diverged");
            flat = rn > 0.95 * r ? flat + 1 : 0;
            r = rn;
        }
        Lp.mg.store(Lp.b.p, Lp.phi.p);
        HU_CHECK(hipMemcpyAsync(phi, Lp.phi.p, n * sizeof(double), hipMemcpyDeviceToHost, u->s));
        HU_CHECK(hipStreamSynchronize(u->s));
        if (cycles) *cycles = cyc;
        if (resid) *resid = bn > 0 ? r / bn : r;
        return CSRST_OK;
    } catch (const std::exception &e) {
        return ws_fail(u, e);
    }
}

// Training on the solver's own iterates.  For each of the nrhs right-hand
// sides b[nrhs][H][W]: a plain solve to a relative residual of 1e-13 gives
// phi*; then from phi = 0, at iterates k = 0 .. kmax-1 (until the residual
// reaches 1e-12), the pair (r_k / rms(r_k), (phi* - phi_k) / rms) is staged
// and the iterate advances by a V-cycle, preceded by the network's step when
// inloop is 1 (pairs from the learned loop itself).  Every maxBatch staged
// pairs make one Adam step of rate lr.  loss: mean loss of those steps;
// pairs: pairs used.
int csrst_warmstart_train_iterates(csrst_warmstart *u, double hx, double hy, const uint8_t *cls, const double *b,
                                   int nrhs, int kmax, int inloop, float lr, float *loss, int *pairs)
{
    try {
        if (!u || !b || nrhs < 1 || kmax < 1 || !(lr > 0)) throw Error(CSRST_EINVAL, "bad training arguments");
        loop_grid(u, hx, hy, cls);
        auto &Lp = u->loop;
        const size_t n = (size_t)u->H * u->W;
        const int MB = u->maxBatch;
        Lp.pairIn.alloc(std::max(Lp.pairIn.n, (size_t)MB * n));
        Lp.pairTgt.alloc(std::max(Lp.pairTgt.n, (size_t)MB * n));
        const pmg::Level &L0 = Lp.mg.level0();
        int slot = 0, used = 0, nsteps = 0;
        double lsum = 0.0;
        for (int i = 0; i < nrhs; ++i) {
            HU_CHECK(hipMemcpyAsync(Lp.b.p, b + (size_t)i * n, n * sizeof(double), hipMemcpyHostToDevice, u->s));
            k_zero_d<<<nb(n), kB, 0, u->s>>>(Lp.phi.p, n);
            HU_LAUNCH_CHECK();
            const double bn = Lp.mg.begin(Lp.b.p, Lp.phi.p);
            if (!(bn > 0)) continue;
            double r = Lp.mg.resnorm();
            for (int c = 0, flat = 0; r > 1e-13 * bn && c < 200 && flat < 3; ++c) {
                Lp.mg.cycle();
                const double rn = Lp.mg.resnorm();
                flat = rn > 0.95 * r ? flat + 1 : 0;
                r = rn;
            }
            k_copy_d<<<nb(n), kB, 0, u->s>>>(Lp.pstar.p, L0.phi, n);
            Lp.mg.begin(Lp.b.p, Lp.phi.p);
            r = Lp.mg.resnorm();
            for (int k = 0; k < kmax && r > 1e-12 * bn; ++k) {
                k_err_sumsq<<<kRed, kB, 0, u->s>>>(Lp.pstar.p, L0.phi, L0.act, n, Lp.part.p);
                double e2 = 0, z = 0;
                loop_sums(u, e2, z);
                const double rr = r / std::sqrt((double)Lp.unknowns), re = std::sqrt(e2 / (double)Lp.unknowns);
                if (rr > 0 && re > 0) {
                    k_res_in<<<nb(n), kB, 0, u->s>>>(L0.res, L0.act, Lp.pairIn.p + (size_t)slot * n, n, 1.0 / rr);
                    k_err_tgt<<<nb(n), kB, 0, u->s>>>(Lp.pstar.p, L0.phi, L0.act, Lp.pairTgt.p + (size_t)slot * n, n,
                                                       1.0 / re);
                    HU_LAUNCH_CHECK();
                    if (++slot == MB) {
                        lsum += loop_train(u, slot, lr);
                        ++nsteps;
                        used += slot;
                        slot = 0;
                    }
                }
                if (inloop) loop_correct(u, r);
                Lp.mg.cycle();
                r = Lp.mg.resnorm();
            }
        }
        if (slot) {
            lsum += loop_train(u, slot, lr);
            ++nsteps;
            used += slot;
        }
        HU_CHECK(hipStreamSynchronize(u->s));
        if (loss) *loss = nsteps ? (float)(lsum / nsteps) : 0.f;
        if (pairs) *pairs = used;
        return CSRST_OK;
    } catch (const std::exception &e) {
        return ws_fail(u, e);
    }
}

}  // extern "C"
