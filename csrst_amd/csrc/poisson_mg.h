// FP64 geometric multigrid for the cut-domain 5-point Poisson operator of
// PoissonSolver::cutoffStencil (PoissonSolver.cpp:678-766).
//
// Operator on an nx x ny node grid (row-major, x fastest), spacings hx
// (already multiplied by gamma for the boosted frame) and hy:
//   unknown rows:   2(cx+cy) phi_ij - cx (phi_i-1,j + phi_i+1,j) - cy (phi_i,j-1 + phi_i,j+1) = b_ij
//                   with cx = 1/hx^2, cy = 1/hy^2; neighbours that are not
//                   unknowns contribute zero (Dirichlet) unless they are
//                   "known" nodes, whose value b moves to the right-hand side;
//   other rows:     identity, phi = b.
// For gamma >> 1 the x coupling is weaker than the y coupling by
// (hy / (gamma hx))^2, so the smoother is an exact y-line solve (parallel
// cyclic reduction in LDS, one workgroup per column, red/black column order)
// and coarsening halves x only.  That pairing converges at a rate that does not
// depend on the anisotropy.  Lines longer than kMaxNy are solved by one
// thread per column (Thomas elimination) instead.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "hip_util.h"

namespace pmg {

enum NodeClass : uint8_t { kFixed = 0, kUnknown = 1, kKnown = 2 };

constexpr int kMaxNy = 2048;     // longest y-line solved in LDS; longer lines use per-thread Thomas

struct Level {
    int nx = 0, ny = 0;
    double cx = 0, cy = 0;
    uint8_t *act = nullptr;      // 1 = unknown on this level
    uint8_t *colact = nullptr;   // per column: any unknown
    double *phi = nullptr, *rhs = nullptr, *res = nullptr;
};

struct Stats {
    int cycles = 0;
    double resid0 = 0, resid = 0, bnorm = 0;   // ||r|| at start and end, ||b|| (unknown rows)
};

class Solver {
public:
    // cls: device array of NodeClass, nx*ny.  Allocates the level hierarchy.
    void setup(int nx, int ny, double hx, double hy, const uint8_t *cls, hipStream_t s);
    // b: device, nx*ny right-hand side of every row; phi: device, in = initial
    // guess, out = solution (identity rows set to b).  Iterates V(2,2) cycles
    // until ||b - A phi|| <= tol ||b|| over unknown rows, maxit cycles, or
    // three cycles in a row that reduce the residual by less than 5% (the
    // round-off floor of the grid).
    Stats solve(const double *b, double *phi, double tol, int maxit);
    // residual norm of phi (device) against b without iterating
    double residual(const double *b, const double *phi);
    // Step-wise use (a caller that acts on the iterate between V-cycles):
    // begin() loads b and the guess phi0 (device) and returns ||b||; resnorm()
    // computes the level-0 residual (level0().res) and returns its norm over
    // the unknowns; cycle() runs one V-cycle on level0().phi; store() writes
    // the solution (identity rows phi = b) to phi (device).
    double begin(const double *b, const double *phi0);
    double resnorm() { return norm_res(0); }
    void cycle() { vcycle(0); }
    void store(const double *b, double *phi);
    const Level &level0() const { return lv_[0]; }
    hipStream_t stream() const { return s_; }
    int nx() const { return nx_; }
    int ny() const { return ny_; }
    int levels() const { return (int)lv_.size(); }

    int nu1 = 2, nu2 = 2, coarse_sweeps = 8;

private:
    void vcycle(int l);
    void smooth(int l, int sweeps);
    double norm_res(int l);
    void prepare(const double *b);
    int nx_ = 0, ny_ = 0;
    hipStream_t s_ = nullptr;
    const uint8_t *cls_ = nullptr;
    std::vector<Level> lv_;
    hu::DBuf<uint8_t> bytes_;
    hu::DBuf<double> reals_;
    hu::DBuf<double> partial_;
    double bnorm_ = 0;
};

}  // namespace pmg
