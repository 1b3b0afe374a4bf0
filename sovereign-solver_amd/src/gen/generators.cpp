// Seeded synthetic LP families (API: include/sovereign/generators.hpp).
#include "sovereign/generators.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace sov::gen {

namespace {
std::uint64_t splitmix(std::uint64_t& x) {
    std::uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
std::uint64_t rotl(std::uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
constexpr double kTwoPi = 6.283185307179586476925286766559;
}  // namespace

Rng::Rng(std::uint64_t seed) {
    std::uint64_t x = seed;
    for (auto& w : s_) w = splitmix(x);
}

std::uint64_t Rng::next() {
    const std::uint64_t r = rotl(s_[1] * 5, 7) * 9;
    const std::uint64_t t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = rotl(s_[3], 45);
    return r;
}

double Rng::uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
double Rng::uniform_open() { return (static_cast<double>(next() >> 11) + 1.0) * 0x1.0p-53; }

double Rng::normal() {
    const double u1 = uniform_open(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(kTwoPi * u2);
}

std::uint64_t Rng::below(std::uint64_t bound) {
    return static_cast<std::uint64_t>((static_cast<unsigned __int128>(next()) * bound) >> 64);
}

Problem LpArrays::to_problem() const {
    return Problem::from_csr(CsrMatrix(m, n, row_ptr, col_idx, val), c, col_lo, col_hi, row_lo, row_hi, sense,
                             obj_const);
}

namespace {

/// k distinct values in [0, m), ascending (Floyd's sampling).
void distinct_rows(Rng& r, Idx m, int k, std::vector<Idx>& out) {
    out.clear();
    for (Idx j = m - k; j < m; ++j) {
        const Idx t = static_cast<Idx>(r.below(static_cast<std::uint64_t>(j) + 1));
        if (std::find(out.begin(), out.end(), t) == out.end()) out.push_back(t);
        else out.push_back(j);
    }
    std::sort(out.begin(), out.end());
}

/// Column lists -> CSR with ascending columns per row; plus the x0/y0/s0
/// construction shared by the random families.
LpArrays assemble_random(Idx m, Idx n, const std::vector<std::vector<Idx>>& rows,
                         const std::vector<std::vector<Real>>& vals, Rng& r) {
    LpArrays lp;
    lp.m = m;
    lp.n = n;
    lp.row_ptr.assign(static_cast<std::size_t>(m) + 1, 0);
    for (Idx j = 0; j < n; ++j)
        for (Idx i : rows[static_cast<std::size_t>(j)]) ++lp.row_ptr[static_cast<std::size_t>(i) + 1];
    std::partial_sum(lp.row_ptr.begin(), lp.row_ptr.end(), lp.row_ptr.begin());
    const auto nnz = static_cast<std::size_t>(lp.row_ptr.back());
    lp.col_idx.resize(nnz);
    lp.val.resize(nnz);
    std::vector<Idx> cur(lp.row_ptr.begin(), lp.row_ptr.end() - 1);
    for (Idx j = 0; j < n; ++j) {
        const auto& rr = rows[static_cast<std::size_t>(j)];
        const auto& vv = vals[static_cast<std::size_t>(j)];
        for (std::size_t t = 0; t < rr.size(); ++t) {
            const auto p = static_cast<std::size_t>(cur[static_cast<std::size_t>(rr[t])]++);
            lp.col_idx[p] = j;
            lp.val[p] = vv[t];
        }
    }
    std::vector<Real> x0(static_cast<std::size_t>(n)), y0(static_cast<std::size_t>(m)), s0(static_cast<std::size_t>(n));
    for (auto& v : x0) v = r.uniform();
    for (auto& v : y0) v = r.normal();
    for (auto& v : s0) v = r.uniform();
    lp.row_lo.resize(static_cast<std::size_t>(m));
    for (Idx i = 0; i < m; ++i) {
        Real acc = 0.0;
        for (Idx k = lp.row_ptr[static_cast<std::size_t>(i)]; k < lp.row_ptr[static_cast<std::size_t>(i) + 1]; ++k)
            acc += lp.val[static_cast<std::size_t>(k)] * x0[static_cast<std::size_t>(lp.col_idx[static_cast<std::size_t>(k)])];
        lp.row_lo[static_cast<std::size_t>(i)] = acc;
    }
    lp.row_hi = lp.row_lo;
    lp.c.resize(static_cast<std::size_t>(n));
    for (Idx j = 0; j < n; ++j) {
        Real acc = 0.0;
        const auto& rr = rows[static_cast<std::size_t>(j)];
        const auto& vv = vals[static_cast<std::size_t>(j)];
        for (std::size_t t = 0; t < rr.size(); ++t) acc += vv[t] * y0[static_cast<std::size_t>(rr[t])];
        lp.c[static_cast<std::size_t>(j)] = acc + s0[static_cast<std::size_t>(j)];
    }
    lp.col_lo.assign(static_cast<std::size_t>(n), 0.0);
    lp.col_hi.assign(static_cast<std::size_t>(n), 10.0);
    return lp;
}

}  // namespace

LpArrays random_lp(Idx m, Idx n, int k, std::uint64_t seed) {
    if (m <= 0 || n <= 0 || k < 0 || k > m) throw Error("random_lp: bad sizes");
    Rng r(seed);
    std::vector<std::vector<Idx>> rows(static_cast<std::size_t>(n));
    std::vector<std::vector<Real>> vals(static_cast<std::size_t>(n));
    for (Idx j = 0; j < n; ++j) {
        distinct_rows(r, m, k, rows[static_cast<std::size_t>(j)]);
        auto& v = vals[static_cast<std::size_t>(j)];
        v.resize(static_cast<std::size_t>(k));
        for (auto& e : v) e = r.normal();
    }
    return assemble_random(m, n, rows, vals, r);
}

LpArrays random_lp_zipf(Idx m, Idx n, double a, int kmin, int kmax, std::uint64_t seed) {
    if (m <= 0 || n <= 0 || kmin < 1 || kmax < kmin || kmax > m || a <= 1.0) throw Error("random_lp_zipf: bad sizes");
    // CDF of Zipf(a) on {1..kmax-1}; u beyond it maps to kmax (the folded tail)
    double zeta = 0.0;
    for (int t = 1; t < 2000000; ++t) zeta += std::pow(static_cast<double>(t), -a);
    zeta += std::pow(2000000.0, 1.0 - a) / (a - 1.0);   // integral tail estimate
    std::vector<double> cdf(static_cast<std::size_t>(kmax));
    double acc = 0.0;
    for (int t = 1; t < kmax; ++t) {
        acc += std::pow(static_cast<double>(t), -a) / zeta;
        cdf[static_cast<std::size_t>(t)] = acc;
    }
    Rng r(seed);
    std::vector<std::vector<Idx>> rows(static_cast<std::size_t>(n));
    std::vector<std::vector<Real>> vals(static_cast<std::size_t>(n));
    for (Idx j = 0; j < n; ++j) {
        const double u = r.uniform();
        int k = kmax;
        for (int t = 1; t < kmax; ++t)
            if (u < cdf[static_cast<std::size_t>(t)]) {
                k = t;
                break;
            }
        k = std::max(k, kmin);
        distinct_rows(r, m, k, rows[static_cast<std::size_t>(j)]);
        auto& v = vals[static_cast<std::size_t>(j)];
        v.resize(static_cast<std::size_t>(k));
        for (auto& e : v) e = r.normal();
    }
    return assemble_random(m, n, rows, vals, r);
}

LpArrays transportation(int S, int D, std::uint64_t seed) {
    if (S <= 0 || D <= 0) throw Error("transportation: bad sizes");
    Rng r(seed);
    std::vector<Real> sup(static_cast<std::size_t>(S)), dem(static_cast<std::size_t>(D));
    for (auto& v : sup) v = 50.0 + 100.0 * r.uniform();
    for (auto& v : dem) v = 50.0 + 100.0 * r.uniform();
    const Real ts = std::accumulate(sup.begin(), sup.end(), 0.0), td = std::accumulate(dem.begin(), dem.end(), 0.0);
    for (auto& v : dem) v *= ts / td;
    LpArrays lp;
    lp.m = S + D;
    lp.n = S * D;
    const auto n = static_cast<std::size_t>(lp.n);
    lp.c.resize(n);
    for (auto& v : lp.c) v = static_cast<Real>(1 + r.below(100));
    lp.col_lo.assign(n, 0.0);
    lp.col_hi.assign(n, kInfinity);
    lp.row_ptr.assign(static_cast<std::size_t>(lp.m) + 1, 0);
    lp.col_idx.reserve(2 * n);
    lp.val.assign(2 * n, 1.0);
    for (int i = 0; i < S; ++i) {
        for (int j = 0; j < D; ++j) lp.col_idx.push_back(i * D + j);
        lp.row_ptr[static_cast<std::size_t>(i) + 1] = static_cast<Idx>(lp.col_idx.size());
    }
    for (int j = 0; j < D; ++j) {
        for (int i = 0; i < S; ++i) lp.col_idx.push_back(i * D + j);
        lp.row_ptr[static_cast<std::size_t>(S + j) + 1] = static_cast<Idx>(lp.col_idx.size());
    }
    lp.row_lo = sup;
    lp.row_lo.insert(lp.row_lo.end(), dem.begin(), dem.end());
    lp.row_hi = lp.row_lo;
    return lp;
}

std::vector<Idx> batch_sizes(int count, Idx lo, Idx hi, std::uint64_t seed) {
    Rng r(seed);
    std::vector<Idx> out(static_cast<std::size_t>(count));
    const double a = std::log(static_cast<double>(lo)), b = std::log(static_cast<double>(hi));
    for (auto& v : out) v = static_cast<Idx>(std::lround(std::exp(a + (b - a) * r.uniform())));
    return out;
}

TransportFamily TransportFamily::make(int S, int D, int P, std::uint64_t seed) {
    TransportFamily f;
    f.S = S;
    f.D = D;
    f.P = P;
    Rng r(seed);
    f.base_s.resize(static_cast<std::size_t>(S));
    f.base_d.resize(static_cast<std::size_t>(D));
    for (auto& v : f.base_s) v = 50.0 + 100.0 * r.uniform();
    for (auto& v : f.base_d) v = 50.0 + 100.0 * r.uniform();
    f.c0.resize(static_cast<std::size_t>(S) * static_cast<std::size_t>(D));
    for (auto& v : f.c0) v = 1.0 + 99.0 * r.uniform();
    f.U.resize(f.c0.size() * static_cast<std::size_t>(P));
    const double sd = 1.0 / std::sqrt(static_cast<double>(P));
    for (auto& v : f.U) v = sd * r.normal();
    return f;
}

void TransportFamily::draw(std::uint64_t seed, std::uint64_t index, std::vector<Real>& theta,
                           std::vector<Real>& z) const {
    Rng r(seed ^ (0xA24BAED4963EE407ULL * (index + 1)));
    theta.resize(static_cast<std::size_t>(P));
    z.resize(static_cast<std::size_t>(D));
    for (auto& v : theta) v = r.normal();
    for (auto& v : z) v = r.normal();
}

LpArrays TransportFamily::instance(const std::vector<Real>& theta, const std::vector<Real>& z) const {
    LpArrays lp = transportation(S, D, 1);   // structure (values/costs replaced below)
    std::vector<Real> d(static_cast<std::size_t>(D)), s(static_cast<std::size_t>(S));
    for (int j = 0; j < D; ++j) d[static_cast<std::size_t>(j)] = base_d[static_cast<std::size_t>(j)] * std::exp(sigma_d * z[static_cast<std::size_t>(j)]);
    const Real td = std::accumulate(d.begin(), d.end(), 0.0), tb = std::accumulate(base_s.begin(), base_s.end(), 0.0);
    for (int i = 0; i < S; ++i) s[static_cast<std::size_t>(i)] = base_s[static_cast<std::size_t>(i)] * (td / tb);
    lp.row_lo = s;
    lp.row_lo.insert(lp.row_lo.end(), d.begin(), d.end());
    lp.row_hi = lp.row_lo;
    const std::size_t nv = c0.size();
    for (std::size_t v = 0; v < nv; ++v) {
        Real u = 0.0;
        for (int p = 0; p < P; ++p) u += U[v * static_cast<std::size_t>(P) + static_cast<std::size_t>(p)] * theta[static_cast<std::size_t>(p)];
        lp.c[v] = c0[v] * (1.0 + amp * std::tanh(u));
    }
    return lp;
}

}  // namespace sov::gen
