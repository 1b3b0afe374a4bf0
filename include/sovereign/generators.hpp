// Seeded synthetic LP families for tests and benchmarks.  A portable RNG
// (splitmix64-seeded xoshiro256**, Box-Muller normals, Lemire bounded
// integers) so every platform draws the same instances.
//
//   random_lp       m rows, n columns, k nonzeros per column at distinct
//                   uniform rows, values N(0,1); x0 ~ U[0,1]^n, b = A x0,
//                   rows lo = hi = b; bounds [0, 10]; y0 ~ N(0,1)^m,
//                   s0 ~ U[0,1]^n, c = A'y0 + s0; minimize.  Feasible and
//                   bounded by construction.
//   random_lp_zipf  the same with k_j = clip(Zipf(a), kmin, kmax) nonzeros in
//                   column j (Zipf on {1,2,...}; the tail above kmax folds
//                   onto kmax, values below kmin onto kmin).
//   transportation  S sources x D sinks, x_ij >= 0 at column i*D + j; supply
//                   rows sum_j x_ij = s_i (s ~ U[50,150]), demand rows
//                   sum_i x_ij = d_j (d ~ U[50,150] rescaled so sum d = sum s);
//                   integer costs U{1..100}; minimize.
//   batch_sizes     m_i = round(exp(U[ln lo, ln hi])) for a batch.
//   TransportFamily the parametric transportation family of the learned warm
//                   start (fixed A, demand and cost parameters, see below).
#pragma once

#include <cstdint>
#include <vector>

#include "sovereign/problem.hpp"

namespace sov::gen {

class Rng {
public:
    explicit Rng(std::uint64_t seed);
    std::uint64_t next();
    double uniform();              ///< [0, 1)
    double uniform_open();         ///< (0, 1]
    double normal();               ///< N(0, 1), Box-Muller (one draw of two uniforms)
    std::uint64_t below(std::uint64_t bound);   ///< uniform integer in [0, bound)

private:
    std::uint64_t s_[4];
};

/// An LP as plain arrays (CSR with ascending columns).
struct LpArrays {
    Idx m = 0, n = 0;
    std::vector<Idx> row_ptr{0}, col_idx;
    std::vector<Real> val, c, col_lo, col_hi, row_lo, row_hi;
    ObjSense sense = ObjSense::Minimize;
    Real obj_const = 0.0;
    Problem to_problem() const;
};

LpArrays random_lp(Idx m, Idx n, int k, std::uint64_t seed);
LpArrays random_lp_zipf(Idx m, Idx n, double a, int kmin, int kmax, std::uint64_t seed);
LpArrays transportation(int sources, int sinks, std::uint64_t seed);
std::vector<Idx> batch_sizes(int count, Idx lo, Idx hi, std::uint64_t seed);

/// Parametric transportation family: fixed S x D structure; base supplies and
/// demands ~ U[50,150], base costs c0 ~ U[1,100], U in R^{(S*D) x P} with
/// entries N(0, 1/P).  An instance takes theta in R^P and demand log-factors
/// z in R^D:  d_j = base_d_j * exp(sigma_d * z_j), s_i = base_s_i * (sum d /
/// sum base_s), c = c0 * (1 + amp * tanh(U theta)).
struct TransportFamily {
    int S = 100, D = 200, P = 16;
    double sigma_d = 0.2, amp = 0.2;
    std::vector<Real> base_s, base_d, c0, U;   // U row-major (S*D) x P
    static TransportFamily make(int S, int D, int P, std::uint64_t seed);
    /// Draw theta ~ N(0, I_P) and z ~ N(0, I_D) for instance `index`.
    void draw(std::uint64_t seed, std::uint64_t index, std::vector<Real>& theta, std::vector<Real>& z) const;
    LpArrays instance(const std::vector<Real>& theta, const std::vector<Real>& z) const;
};

}  // namespace sov::gen
