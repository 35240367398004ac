// Beam statistics of the bunch on the GPU: the moments of
// PartBunch::calcMoments (PartBunch.cpp:659-756) and the derived parameters
// of PartBunch::calcBeamParameters (PartBunch.cpp:584-657), which the
// statistics file of DataSinkCommand (DataSink.cpp:91-148) prints.
//
// Two passes over the particles, both double-double per block with a fixed
// reduction tree, summed on the host in long double:
//   1. sums of x, px, y, py and the bounding box;
//   2. with the centroid and the mean momentum's direction known, the ten
//      second moments of (x, px, y, py) centred and rotated into the bunch
//      frame (x along the mean momentum), and the sum of gamma.
// A particle source is a struct with
//   __device__ void get(int k, double &x, double &y, double &px, double &py) const.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "hip_util.h"

namespace bs {

constexpr int kBlocks = 240, kThreads = 256;
constexpr int kRaw = 20;      // raw statistics, layout below
constexpr int kParams = 20;   // derived parameters, layout below

struct DD {
    double hi, lo;
};
__device__ inline DD add(DD a, DD b)
{
    double s = a.hi + b.hi;
    const double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (b.hi - bb);
    e += a.lo + b.lo;
    const double hi = s + e;
    return DD{hi, e - (hi - s)};
}

// block tree sum of M double-double values per thread into part[blockIdx.x * M + m]
template <int M>
__device__ void block_sum(DD (&v)[M], DD *part)
{
    __shared__ DD sh[M][kThreads];
    for (int m = 0; m < M; ++m) sh[m][threadIdx.x] = v[m];
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int m = 0; m < M; ++m) sh[m][threadIdx.x] = add(sh[m][threadIdx.x], sh[m][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int m = 0; m < M; ++m) part[blockIdx.x * M + m] = sh[m][0];
}

template <class Src>
__global__ void __launch_bounds__(kThreads) k_pass1(Src src, int n, DD *part, double *lim)
{
    DD v[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    double lo[2] = {DBL_MAX, DBL_MAX}, hi[2] = {-DBL_MAX, -DBL_MAX};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        double x, y, px, py;
        src.get(k, x, y, px, py);
        v[0] = add(v[0], DD{x, 0});
        v[1] = add(v[1], DD{px, 0});
        v[2] = add(v[2], DD{y, 0});
        v[3] = add(v[3], DD{py, 0});
        lo[0] = fmin(lo[0], x); lo[1] = fmin(lo[1], y);
        hi[0] = fmax(hi[0], x); hi[1] = fmax(hi[1], y);
    }
    block_sum<4>(v, part);
    __shared__ double sl[4][kThreads];
    sl[0][threadIdx.x] = lo[0]; sl[1][threadIdx.x] = lo[1];
    sl[2][threadIdx.x] = hi[0]; sl[3][threadIdx.x] = hi[1];
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int m = 0; m < 2; ++m) sl[m][threadIdx.x] = fmin(sl[m][threadIdx.x], sl[m][threadIdx.x + s]);
            for (int m = 2; m < 4; ++m) sl[m][threadIdx.x] = fmax(sl[m][threadIdx.x], sl[m][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int m = 0; m < 4; ++m) lim[blockIdx.x * 4 + m] = sl[m][0];
}

// PartBunch.cpp:732-745: part = (x', px', y', py') with positions centred on
// (ax, ay) and both vectors rotated by the mean momentum's angle (c, s)
template <class Src>
__global__ void __launch_bounds__(kThreads) k_pass2(Src src, int n, double ax, double ay, double c, double s,
                                                    DD *part)
{
    DD v[11];
    for (int m = 0; m < 11; ++m) v[m] = DD{0, 0};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        double x, y, px, py;
        src.get(k, x, y, px, py);
        double p[4];
        p[1] = c * px + s * py;
        p[3] = -s * px + c * py;
        const double dx = x - ax, dy = y - ay;
        p[0] = c * dx + s * dy;
        p[2] = -s * dx + c * dy;
        int m = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j <= i; ++j, ++m) v[m] = add(v[m], DD{p[i] * p[j], 0});
        v[10] = add(v[10], DD{sqrt(1.0 + (px * px + py * py)), 0});
    }
    block_sum<11>(v, part);
}

// Raw statistics: [0] N, [1..4] sums of x, px, y, py, [5..6] min x, y,
// [7..8] max x, y, [9..18] second moments (i, j <= i) in the order
// (0,0) (1,0) (1,1) (2,0) (2,1) (2,2) (3,0) (3,1) (3,2) (3,3) of
// (x', px', y', py'), [19] sum of gamma.  part: device, kBlocks * 11 DD;
// lim: device, kBlocks * 4 doubles.  Synchronises the stream.
template <class Src>
void raw_stats(const Src &src, int n, hipStream_t st, DD *part, double *lim, double raw[kRaw])
{
    for (int i = 0; i < kRaw; ++i) raw[i] = 0;
    raw[0] = n;
    if (n <= 0) return;
    std::vector<DD> hp(kBlocks * 11);
    std::vector<double> hl(kBlocks * 4);
    k_pass1<Src><<<kBlocks, kThreads, 0, st>>>(src, n, part, lim);
    HU_LAUNCH_CHECK();
    HU_CHECK(hipMemcpyAsync(hp.data(), part, kBlocks * 4 * sizeof(DD), hipMemcpyDeviceToHost, st));
    HU_CHECK(hipMemcpyAsync(hl.data(), lim, kBlocks * 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    HU_CHECK(hipStreamSynchronize(st));
    long double sum[4] = {0, 0, 0, 0};
    double lo[2] = {DBL_MAX, DBL_MAX}, hi[2] = {-DBL_MAX, -DBL_MAX};
    for (int b = 0; b < kBlocks; ++b) {
        for (int m = 0; m < 4; ++m) sum[m] += (long double)hp[b * 4 + m].hi + (long double)hp[b * 4 + m].lo;
        lo[0] = std::min(lo[0], hl[b * 4]); lo[1] = std::min(lo[1], hl[b * 4 + 1]);
        hi[0] = std::max(hi[0], hl[b * 4 + 2]); hi[1] = std::max(hi[1], hl[b * 4 + 3]);
    }
    for (int m = 0; m < 4; ++m) raw[1 + m] = (double)sum[m];
    raw[5] = lo[0]; raw[6] = lo[1]; raw[7] = hi[0]; raw[8] = hi[1];
    // PartBunch.cpp:724-730
    const double ax = (double)(sum[0] / n), ay = (double)(sum[2] / n);
    const double pxm = (double)(sum[1] / n), pym = (double)(sum[3] / n);
    const double am = std::sqrt(pxm * pxm + pym * pym);
    const double c = am > 0 ? pxm / am : 1.0, s = am > 0 ? pym / am : 0.0;   // a bunch at rest: no rotation
    k_pass2<Src><<<kBlocks, kThreads, 0, st>>>(src, n, ax, ay, c, s, part);
    HU_LAUNCH_CHECK();
    HU_CHECK(hipMemcpyAsync(hp.data(), part, kBlocks * 11 * sizeof(DD), hipMemcpyDeviceToHost, st));
    HU_CHECK(hipStreamSynchronize(st));
    long double mom[11];
    for (int m = 0; m < 11; ++m) mom[m] = 0;
    for (int b = 0; b < kBlocks; ++b)
        for (int m = 0; m < 11; ++m) mom[m] += (long double)hp[b * 11 + m].hi + (long double)hp[b * 11 + m].lo;
    for (int m = 0; m < 10; ++m) raw[9 + m] = (double)mom[m];
    raw[19] = (double)mom[10];
}

// Derived parameters (PartBunch.cpp:611-654), the PartBunchState fields:
// [0] Ekin (MeV), [1..2] rrms, [3..4] prms, [5..6] normalised emittance,
// [7..8] rmean, [9..10] pmean, [11] Dy, [12] DDy, [13] dE (MeV),
// [14..15] rmin, [16..17] rmax, [18..19] rprms.
inline void parameters(const double raw[kRaw], double out[kParams])
{
    for (int i = 0; i < kParams; ++i) out[i] = 0;
    const double N = raw[0];
    if (N == 0) return;
    const double m0 = 0.51099906;   // Physics::m_e (GeV) * 1000, PartBunch.cpp:588
    auto mom = [&](int i, int j) { return raw[9 + (i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i)]; };
    double rmean[2], pmean[2];
    for (int i = 0; i < 2; ++i) { rmean[i] = raw[1 + 2 * i] / N; pmean[i] = raw[2 + 2 * i] / N; }
    const double cen1 = pmean[0] * pmean[0] + pmean[1] * pmean[1];
    double rsq[2], psq[2], rp[2];
    for (int i = 0; i < 2; ++i) {
        rsq[i] = mom(2 * i, 2 * i);
        psq[i] = mom(2 * i + 1, 2 * i + 1) - N * (i == 0 ? cen1 : 0.0);
        if (psq[i] < 0) psq[i] = 0;
        rp[i] = mom(2 * i, 2 * i + 1);
    }
    for (int i = 0; i < 2; ++i) {
        const double eps2 = (rsq[i] * psq[i] - rp[i] * rp[i]) / (N * N);
        const double rpm = rp[i] / N;
        out[1 + i] = std::sqrt(rsq[i] / N);
        out[3 + i] = std::sqrt(psq[i] / N);
        out[5 + i] = std::sqrt(std::max(eps2, 0.0));
        const double tmp = out[1 + i] * out[3 + i];
        out[18 + i] = rpm * (tmp == 0 ? 0.0 : 1.0 / tmp);
        out[7 + i] = rmean[i];
        out[9 + i] = pmean[i];
    }
    out[11] = mom(2, 1) / N;
    out[12] = mom(3, 1) / N;
    const double gamma = raw[19] / N;
    out[0] = (gamma - 1.0) * m0;
    const double ek = out[0];
    out[13] = out[3] * std::sqrt(ek * (ek + 2. * m0) / (1. + ek * (ek + 2. * m0) / (m0 * m0)));
    out[14] = raw[5]; out[15] = raw[6]; out[16] = raw[7]; out[17] = raw[8];
}

}  // namespace bs
