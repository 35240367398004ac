// FP64 geometric multigrid for the Poisson initialisation (PoissonSolver.cpp:85-201
// replaces Belos block-CG + ML smoothed aggregation by line-relaxation
// multigrid on the same cutoffStencil operator, PoissonSolver.cpp:678-766).
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <vector>

#include "csrst_hip.h"
#include "hip_util.h"
#include "poisson_mg.h"

using hu::DBuf;
using hu::Error;

namespace pmg {
namespace {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 256;

int nblk(size_t n) { return (int)((n + kBlock - 1) / kBlock); }

__global__ void k_fill(double *p, size_t n, double v)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

__global__ void k_copy(double *dst, const double *src, size_t n)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

// Level-0 unknown mask from the node classes.
__global__ void k_act0(const uint8_t *cls, uint8_t *act, size_t n)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) act[k] = cls[k] == kUnknown ? 1 : 0;
}

// Coarse node (I, j) is an unknown iff fine node (2I, j) is.
__global__ void k_act_coarse(Level f, Level c)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)c.nx * c.ny) return;
    const int I = (int)(k % c.nx), j = (int)(k / c.nx);
    c.act[k] = f.act[(size_t)j * f.nx + 2 * I];
}

__global__ void k_colact(Level L)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.nx) return;
    uint8_t any = 0;
    for (int j = 0; j < L.ny; ++j) any |= L.act[(size_t)j * L.nx + i];
    L.colact[i] = any;
}

// Right-hand side of the unknown rows: b plus the known neighbours' values
// (cutoffStencil keeps the coupling to a flagged neighbour that is itself an
// identity row, PoissonSolver.cpp:700-741).
__global__ void k_prepare(const uint8_t *cls, const double *b, Level L)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)L.nx * L.ny) return;
    const int i = (int)(k % L.nx), j = (int)(k / L.nx);
    double r = 0.0;
    if (cls[k] == kUnknown) {
        r = b[k];
        if (i > 0 && cls[k - 1] == kKnown) r += L.cx * b[k - 1];
        if (i + 1 < L.nx && cls[k + 1] == kKnown) r += L.cx * b[k + 1];
        if (j > 0 && cls[k - L.nx] == kKnown) r += L.cy * b[k - L.nx];
        if (j + 1 < L.ny && cls[k + L.nx] == kKnown) r += L.cy * b[k + L.nx];
    }
    L.rhs[k] = r;
}

// Initial guess restricted to the unknowns.
__global__ void k_load_guess(const double *phi, Level L)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < (size_t)L.nx * L.ny) L.phi[k] = L.act[k] ? phi[k] : 0.0;
}

// Solution out: unknowns from the level, identity rows phi = b.
__global__ void k_store(const uint8_t *cls, const double *b, const double *src, double *phi, size_t n)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) phi[k] = cls[k] == kUnknown ? src[k] : b[k];
}

__device__ __forceinline__ double apply(const Level &L, size_t k, int i, int j)
{
    double s = 0.0;
    if (i > 0 && L.act[k - 1]) s += L.phi[k - 1];
    if (i + 1 < L.nx && L.act[k + 1]) s += L.phi[k + 1];
    double t = 0.0;
    if (j > 0 && L.act[k - L.nx]) t += L.phi[k - L.nx];
    if (j + 1 < L.ny && L.act[k + L.nx]) t += L.phi[k + L.nx];
    return 2.0 * (L.cx + L.cy) * L.phi[k] - L.cx * s - L.cy * t;
}

__global__ void k_residual(Level L)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)L.nx * L.ny) return;
    const int i = (int)(k % L.nx), j = (int)(k / L.nx);
    L.res[k] = L.act[k] ? L.rhs[k] - apply(L, k, i, j) : 0.0;
}

// Sum of squares, fixed order: block partials, then one thread.
__global__ void k_sumsq_partial(const double *v, size_t n, double *partial)
{
    __shared__ double sh[kBlock];
    double s = 0.0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        s += v[k] * v[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_sum_final(double *partial, int nb)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partial[b];
    partial[nb] = s;
}

// Exact y-line solve of the columns i = 2 blockIdx.x + color by parallel
// cyclic reduction.  Row j of column i, normalised by the diagonal:
//   a_j phi_{j-1} + phi_j + c_j phi_{j+1} = d_j,
// with the x neighbours taken from the current iterate.  Each reduction step
// eliminates the couplings at distance s using rows j -/+ s; after
// ceil(log2 ny) steps every row is decoupled and phi_j = d_j.
template <int R, int NT>
__global__ __launch_bounds__(NT) void k_line_pcr(Level L, int color)
{
    extern __shared__ double sm[];
    const int ny = L.ny;
    double *A = sm, *Cc = sm + ny, *D = sm + 2 * ny;
    // Consecutive workgroup ids run on the 8 XCDs in turn.  Each XCD gets a
    // contiguous range of columns, so the columns that share a cache line
    // (16 doubles of a row) are read through one L2.
    int bid = blockIdx.x;
    if (gridDim.x >= 64 && (gridDim.x & 7) == 0) bid = (bid & 7) * (gridDim.x >> 3) + (bid >> 3);
    const int i = 2 * bid + color;
    if (i >= L.nx || !L.colact[i]) return;           // uniform over the workgroup
    const double diag = 2.0 * (L.cx + L.cy);
    const double inv = 1.0 / diag;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = threadIdx.x + r * NT;
        if (j >= ny) break;
        const size_t k = (size_t)j * L.nx + i;
        double a = 0.0, c = 0.0, d = 0.0;
        if (L.act[k]) {
            double s = 0.0;
            if (i > 0 && L.act[k - 1]) s += L.phi[k - 1];
            if (i + 1 < L.nx && L.act[k + 1]) s += L.phi[k + 1];
            d = (L.rhs[k] + L.cx * s) * inv;
            a = (j > 0 && L.act[k - L.nx]) ? -L.cy * inv : 0.0;
            c = (j + 1 < ny && L.act[k + L.nx]) ? -L.cy * inv : 0.0;
        }
        A[j] = a; Cc[j] = c; D[j] = d;
    }
    __syncthreads();
    for (int s = 1; s < ny; s <<= 1) {
        double na[R], nc[R], nd[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = threadIdx.x + r * NT;
            na[r] = 0.0; nc[r] = 0.0; nd[r] = 0.0;
            if (j < ny) {
                const double a = A[j], c = Cc[j], d = D[j];
                double am = 0.0, cm = 0.0, dm = 0.0, ap = 0.0, cp = 0.0, dp = 0.0;
                if (j - s >= 0) { am = A[j - s]; cm = Cc[j - s]; dm = D[j - s]; }
                if (j + s < ny) { ap = A[j + s]; cp = Cc[j + s]; dp = D[j + s]; }
                const double ib = 1.0 / (1.0 - a * cm - c * ap);
                na[r] = -(a * am) * ib;
                nc[r] = -(c * cp) * ib;
                nd[r] = (d - a * dm - c * dp) * ib;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = threadIdx.x + r * NT;
            if (j < ny) { A[j] = na[r]; Cc[j] = nc[r]; D[j] = nd[r]; }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = threadIdx.x + r * NT;
        if (j >= ny) break;
        const size_t k = (size_t)j * L.nx + i;
        if (L.act[k]) L.phi[k] = D[j];
    }
}

// The same y-line solve for columns longer than the LDS holds: one thread per
// column, Thomas elimination (the forward coefficients go to the level's
// residual array, free during smoothing).
__global__ void k_line_thomas(Level L, int color)
{
    const int i = 2 * (blockIdx.x * blockDim.x + threadIdx.x) + color;
    if (i >= L.nx || !L.colact[i]) return;
    const int ny = L.ny;
    const double diag = 2.0 * (L.cx + L.cy);
    double cp = 0.0, dp = 0.0;
    for (int j = 0; j < ny; ++j) {
        const size_t k = (size_t)j * L.nx + i;
        double a = 0.0, c = 0.0, bd = 1.0, d = 0.0;
        if (L.act[k]) {
            double s = 0.0;
            if (i > 0 && L.act[k - 1]) s += L.phi[k - 1];
            if (i + 1 < L.nx && L.act[k + 1]) s += L.phi[k + 1];
            d = L.rhs[k] + L.cx * s;
            a = (j > 0 && L.act[k - L.nx]) ? -L.cy : 0.0;
            c = (j + 1 < ny && L.act[k + L.nx]) ? -L.cy : 0.0;
            bd = diag;
        }
        const double m = bd - a * cp;
        cp = c / m;
        dp = (d - a * dp) / m;
        L.res[k] = cp;
        L.phi[k] = dp;
    }
    double x = 0.0;
    for (int j = ny - 1; j >= 0; --j) {
        const size_t k = (size_t)j * L.nx + i;
        x = L.phi[k] - L.res[k] * x;
        L.phi[k] = L.act[k] ? x : 0.0;
    }
}

// Full weighting in x onto the coarse rhs; coarse iterate cleared.
__global__ void k_restrict(Level f, Level c)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)c.nx * c.ny) return;
    const int I = (int)(k % c.nx), j = (int)(k / c.nx);
    double v = 0.0;
    if (c.act[k]) {
        const size_t fk = (size_t)j * f.nx + 2 * I;
        v = 0.5 * f.res[fk];
        if (2 * I - 1 >= 0) v += 0.25 * f.res[fk - 1];
        if (2 * I + 1 < f.nx) v += 0.25 * f.res[fk + 1];
    }
    c.rhs[k] = v;
    c.phi[k] = 0.0;
}

// Linear interpolation in x of the coarse correction, added to the unknowns.
__global__ void k_prolong(Level c, Level f)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)f.nx * f.ny) return;
    if (!f.act[k]) return;
    const int i = (int)(k % f.nx), j = (int)(k / f.nx);
    const size_t row = (size_t)j * c.nx;
    double e;
    if ((i & 1) == 0) {
        e = c.phi[row + i / 2];
    } else {
        const int I0 = (i - 1) / 2, I1 = I0 + 1;
        e = 0.5 * (c.phi[row + I0] + (I1 < c.nx ? c.phi[row + I1] : 0.0));
    }
    f.phi[k] += e;
}

template <int R, int NT>
void launch_pcr(const Level &L, int color, hipStream_t s)
{
    const int nb = (L.nx + 1 - color) / 2;
    if (nb <= 0) return;
    k_line_pcr<R, NT><<<nb, NT, 3 * L.ny * sizeof(double), s>>>(L, color);
}

}  // namespace

void Solver::setup(int nx, int ny, double hx, double hy, const uint8_t *cls, hipStream_t s)
{
    if (nx < 1 || ny < 1) throw Error(CSRST_EINVAL, "empty Poisson grid");
    if (!(hx > 0) || !(hy > 0)) throw Error(CSRST_EINVAL, "Poisson spacings must be positive");
    nx_ = nx; ny_ = ny; s_ = s; cls_ = cls;
    lv_.clear();
    int n = nx;
    double cx = 1.0 / (hx * hx);
    const double cy = 1.0 / (hy * hy);
    size_t nodes = 0, cols = 0;
    while (true) {
        Level L;
        L.nx = n; L.ny = ny; L.cx = cx; L.cy = cy;
        lv_.push_back(L);
        nodes += (size_t)n * ny;
        cols += n;
        if (n <= 2) break;
        n = (n + 1) / 2;
        cx *= 0.25;
    }
    bytes_.alloc(nodes + cols);
    reals_.alloc(3 * nodes);
    partial_.alloc(kRedBlocks + 1);
    size_t ob = 0, orr = 0;
    for (auto &L : lv_) {
        const size_t m = (size_t)L.nx * L.ny;
        L.act = bytes_.p + ob; ob += m;
        L.colact = bytes_.p + ob; ob += L.nx;
        L.phi = reals_.p + orr; orr += m;
        L.rhs = reals_.p + orr; orr += m;
        L.res = reals_.p + orr; orr += m;
    }
    const size_t m0 = (size_t)nx * ny;
    k_act0<<<nblk(m0), kBlock, 0, s_>>>(cls, lv_[0].act, m0);
    for (size_t l = 1; l < lv_.size(); ++l)
        k_act_coarse<<<nblk((size_t)lv_[l].nx * ny), kBlock, 0, s_>>>(lv_[l - 1], lv_[l]);
    for (auto &L : lv_) k_colact<<<nblk(L.nx), kBlock, 0, s_>>>(L);
    HU_LAUNCH_CHECK();
}

void Solver::smooth(int l, int sweeps)
{
    const Level &L = lv_[l];
    for (int k = 0; k < sweeps; ++k)
        for (int color = 0; color < 2; ++color) {
            if (L.ny <= 256) launch_pcr<1, 256>(L, color, s_);
            else if (L.ny <= 512) launch_pcr<1, 512>(L, color, s_);
            else if (L.ny <= 1024) launch_pcr<1, 1024>(L, color, s_);
            else if (L.ny <= kMaxNy) launch_pcr<2, 1024>(L, color, s_);
            else {
                const int ncol = (L.nx + 1 - color) / 2;
                if (ncol > 0) k_line_thomas<<<(ncol + 63) / 64, 64, 0, s_>>>(L, color);
            }
        }
    HU_LAUNCH_CHECK();
}

double Solver::norm_res(int l)
{
    const Level &L = lv_[l];
    const size_t m = (size_t)L.nx * L.ny;
    k_residual<<<nblk(m), kBlock, 0, s_>>>(L);
    k_sumsq_partial<<<kRedBlocks, kBlock, 0, s_>>>(L.res, m, partial_.p);
    k_sum_final<<<1, 1, 0, s_>>>(partial_.p, kRedBlocks);
    HU_LAUNCH_CHECK();
    double v = 0.0;
    HU_CHECK(hipMemcpyAsync(&v, partial_.p + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, s_));
    HU_CHECK(hipStreamSynchronize(s_));
    return std::sqrt(v);
}

void Solver::vcycle(int l)
{
    if (l + 1 == (int)lv_.size()) {
        smooth(l, coarse_sweeps);
        return;
    }
    const Level &f = lv_[l], &c = lv_[l + 1];
    smooth(l, nu1);
    k_residual<<<nblk((size_t)f.nx * f.ny), kBlock, 0, s_>>>(f);
    k_restrict<<<nblk((size_t)c.nx * c.ny), kBlock, 0, s_>>>(f, c);
    HU_LAUNCH_CHECK();
    vcycle(l + 1);
    k_prolong<<<nblk((size_t)f.nx * f.ny), kBlock, 0, s_>>>(c, f);
    HU_LAUNCH_CHECK();
    smooth(l, nu2);
}

void Solver::prepare(const double *b)
{
    const Level &L = lv_[0];
    const size_t m = (size_t)L.nx * L.ny;
    k_prepare<<<nblk(m), kBlock, 0, s_>>>(cls_, b, L);
    k_sumsq_partial<<<kRedBlocks, kBlock, 0, s_>>>(L.rhs, m, partial_.p);
    k_sum_final<<<1, 1, 0, s_>>>(partial_.p, kRedBlocks);
    HU_LAUNCH_CHECK();
    double v = 0.0;
    HU_CHECK(hipMemcpyAsync(&v, partial_.p + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, s_));
    HU_CHECK(hipStreamSynchronize(s_));
    bnorm_ = std::sqrt(v);
}

double Solver::residual(const double *b, const double *phi)
{
    prepare(b);
    const Level &L = lv_[0];
    const size_t m = (size_t)L.nx * L.ny;
    k_load_guess<<<nblk(m), kBlock, 0, s_>>>(phi, L);
    HU_LAUNCH_CHECK();
    return norm_res(0);
}

double Solver::begin(const double *b, const double *phi0)
{
    prepare(b);
    const Level &L = lv_[0];
    k_load_guess<<<nblk((size_t)L.nx * L.ny), kBlock, 0, s_>>>(phi0, L);
    HU_LAUNCH_CHECK();
    return bnorm_;
}

void Solver::store(const double *b, double *phi)
{
    const size_t m = (size_t)nx_ * ny_;
    k_store<<<nblk(m), kBlock, 0, s_>>>(cls_, b, lv_[0].phi, phi, m);
    HU_LAUNCH_CHECK();
}

Stats Solver::solve(const double *b, double *phi, double tol, int maxit)
{
    Stats st;
    prepare(b);
    const Level &L = lv_[0];
    const size_t m = (size_t)L.nx * L.ny;
    k_load_guess<<<nblk(m), kBlock, 0, s_>>>(phi, L);
    HU_LAUNCH_CHECK();
    st.bnorm = bnorm_;
    double r = norm_res(0);
    st.resid0 = r;
    const double target = tol * (bnorm_ > 0 ? bnorm_ : 1.0);
    int flat = 0;   // cycles in a row that gained less than 5%: the round-off floor
    while (r > target && st.cycles < maxit && flat < 3) {
        vcycle(0);
        ++st.cycles;
        const double rn = norm_res(0);
        if (!std::isfinite(rn)) throw Error(CSRST_EHIP, "multigrid diverged");
        flat = rn > 0.95 * r ? flat + 1 : 0;
        r = rn;
    }
    st.resid = r;
    k_store<<<nblk(m), kBlock, 0, s_>>>(cls_, b, L.phi, phi, m);
    HU_LAUNCH_CHECK();
    return st;
}

}  // namespace pmg

// ---------------------------------------------------------------------------
// Standalone solve on host arrays.

extern "C" int csrst_poisson_solve(int nx, int ny, double hx, double hy, const uint8_t *cls, const double *b,
                                   double *phi, double tol, int maxit, int *cycles, double *resid)
{
    hipStream_t s = nullptr;
    try {
        if (!cls || !b || !phi || nx < 1 || ny < 1 || maxit < 0 || !(tol >= 0))
            throw Error(CSRST_EINVAL, "bad Poisson arguments");
        const size_t m = (size_t)nx * ny;
        for (size_t k = 0; k < m; ++k)
            if (cls[k] > pmg::kKnown) throw Error(CSRST_EINVAL, hu::fmt("node class %d at %zu", cls[k], k));
        HU_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        DBuf<uint8_t> dcls;
        DBuf<double> db, dphi;
        dcls.alloc(m); db.alloc(m); dphi.alloc(m);
        HU_CHECK(hipMemcpyAsync(dcls.p, cls, m, hipMemcpyHostToDevice, s));
        HU_CHECK(hipMemcpyAsync(db.p, b, m * sizeof(double), hipMemcpyHostToDevice, s));
        HU_CHECK(hipMemcpyAsync(dphi.p, phi, m * sizeof(double), hipMemcpyHostToDevice, s));
        pmg::Solver sol;
        sol.setup(nx, ny, hx, hy, dcls.p, s);
        const pmg::Stats st = sol.solve(db.p, dphi.p, tol, maxit);
        HU_CHECK(hipMemcpyAsync(phi, dphi.p, m * sizeof(double), hipMemcpyDeviceToHost, s));
        HU_CHECK(hipStreamSynchronize(s));
        HU_CHECK(hipStreamDestroy(s));
        s = nullptr;
        if (cycles) *cycles = st.cycles;
        if (resid) *resid = st.bnorm > 0 ? st.resid / st.bnorm : st.resid;
        return CSRST_OK;
    } catch (const std::exception &e) {
        if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
        const Error *he = dynamic_cast<const Error *>(&e);
        hu::global_error() = e.what();
        return he ? he->code : CSRST_EINVAL;
    }
}
