// Host Ruiz + Pock-Chambolle scaling (API: include/sovereign/scaling.hpp).
// Same arithmetic, in the same order, as reference src/l1/scaling.cpp: Ruiz
// sweeps over |a| r_i c_j inf-norms with early stop (:36-74), then
// Pock-Chambolle sums of m^(2-alpha) / m^alpha (:80-101); apply (:110-133);
// unscale (:139-149); the problem form rescales c, column and row bounds
// (:155-205).  The device version is sov_hip_scale_* (include/sov_hip.h).
#include "sovereign/scaling.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>

namespace sov {
namespace {
Real root_or_one(Real v) { return v > 0.0 ? std::sqrt(v) : 1.0; }
}  // namespace

Scaling Scaling::equilibrate(const CsrMatrix& A, const ScalingOptions& opt) {
    const auto m = static_cast<std::size_t>(A.rows()), n = static_cast<std::size_t>(A.cols());
    Scaling s;
    s.row_scale_.assign(m, 1.0);
    s.col_scale_.assign(n, 1.0);
    const auto rp = A.row_ptr();
    const auto ci = A.col_idx();
    const auto v = A.values();
    std::vector<Real> racc(m), cacc(n);
    if (opt.apply_ruiz) {
        for (int sweep = 0; sweep < opt.ruiz_iterations; ++sweep) {
            std::fill(racc.begin(), racc.end(), 0.0);
            std::fill(cacc.begin(), cacc.end(), 0.0);
            for (std::size_t r = 0; r < m; ++r)
                for (auto k = static_cast<std::size_t>(rp[r]); k < static_cast<std::size_t>(rp[r + 1]); ++k) {
                    const auto c = static_cast<std::size_t>(ci[k]);
                    const Real mag = std::abs(v[k]) * s.row_scale_[r] * s.col_scale_[c];
                    racc[r] = std::max(racc[r], mag);
                    cacc[c] = std::max(cacc[c], mag);
                }
            Real worst = 0.0;
            for (std::size_t r = 0; r < m; ++r)
                if (racc[r] > 0.0) {
                    s.row_scale_[r] /= root_or_one(racc[r]);
                    worst = std::max(worst, std::abs(racc[r] - 1.0));
                }
            for (std::size_t c = 0; c < n; ++c)
                if (cacc[c] > 0.0) {
                    s.col_scale_[c] /= root_or_one(cacc[c]);
                    worst = std::max(worst, std::abs(cacc[c] - 1.0));
                }
            if (worst < opt.ruiz_tolerance) break;
        }
    }
    if (opt.apply_pock_chambolle) {
        const Real a = opt.pock_chambolle_alpha;
        std::fill(racc.begin(), racc.end(), 0.0);
        std::fill(cacc.begin(), cacc.end(), 0.0);
        for (std::size_t r = 0; r < m; ++r)
            for (auto k = static_cast<std::size_t>(rp[r]); k < static_cast<std::size_t>(rp[r + 1]); ++k) {
                const auto c = static_cast<std::size_t>(ci[k]);
                const Real mag = std::abs(v[k]) * s.row_scale_[r] * s.col_scale_[c];
                if (mag <= 0.0) continue;
                racc[r] += std::pow(mag, 2.0 - a);
                cacc[c] += std::pow(mag, a);
            }
        for (std::size_t r = 0; r < m; ++r)
            if (racc[r] > 0.0) s.row_scale_[r] /= root_or_one(racc[r]);
        for (std::size_t c = 0; c < n; ++c)
            if (cacc[c] > 0.0) s.col_scale_[c] /= root_or_one(cacc[c]);
    }
    return s;
}

Scaling Scaling::from_factors(std::vector<Real> row_scale, std::vector<Real> col_scale) {
    Scaling s;
    s.row_scale_ = std::move(row_scale);
    s.col_scale_ = std::move(col_scale);
    return s;
}

CsrMatrix Scaling::apply(const CsrMatrix& A) const {
    if (A.rows() != num_rows() || A.cols() != num_cols()) throw Error(Status::DimensionMismatch, "Scaling::apply");
    const auto rp = A.row_ptr();
    const auto ci = A.col_idx();
    const auto v = A.values();
    std::vector<Real> nv(v.size());
    for (std::size_t r = 0; r + 1 < rp.size(); ++r)
        for (auto k = static_cast<std::size_t>(rp[r]); k < static_cast<std::size_t>(rp[r + 1]); ++k)
            nv[k] = v[k] * row_scale_[r] * col_scale_[static_cast<std::size_t>(ci[k])];
    return CsrMatrix(A.rows(), A.cols(), std::vector<Idx>(rp.begin(), rp.end()), std::vector<Idx>(ci.begin(), ci.end()),
                     std::move(nv));
}

void Scaling::unscale_primal(std::span<Real> x) const {
    if (static_cast<Idx>(x.size()) != num_cols()) throw Error(Status::DimensionMismatch, "Scaling::unscale_primal");
    for (std::size_t j = 0; j < x.size(); ++j) x[j] *= col_scale_[j];
}

void Scaling::unscale_dual(std::span<Real> y) const {
    if (static_cast<Idx>(y.size()) != num_rows()) throw Error(Status::DimensionMismatch, "Scaling::unscale_dual");
    for (std::size_t i = 0; i < y.size(); ++i) y[i] *= row_scale_[i];
}

std::pair<Problem, Scaling> Scaling::equilibrate(const Problem& p, const ScalingOptions& opt) {
    Scaling s = equilibrate(p.matrix(), opt);
    const auto n = static_cast<std::size_t>(p.num_cols()), m = static_cast<std::size_t>(p.num_rows());
    std::vector<Real> c(n), clo(n), chi(n), rlo(m), rhi(m);
    for (std::size_t j = 0; j < n; ++j) {
        const Real cs = s.col_scale_[j];
        c[j] = p.objective()[j] * cs;
        clo[j] = is_finite_bound(p.col_lower()[j]) ? p.col_lower()[j] / cs : p.col_lower()[j];
        chi[j] = is_finite_bound(p.col_upper()[j]) ? p.col_upper()[j] / cs : p.col_upper()[j];
    }
    for (std::size_t i = 0; i < m; ++i) {
        const Real rs = s.row_scale_[i];
        rlo[i] = is_finite_bound(p.row_lower()[i]) ? p.row_lower()[i] * rs : p.row_lower()[i];
        rhi[i] = is_finite_bound(p.row_upper()[i]) ? p.row_upper()[i] * rs : p.row_upper()[i];
    }
    // Scaled values, dropping magnitudes <= 1e-14 as the reference's rebuild does.
    const CsrMatrix scaled = s.apply(p.matrix());
    std::vector<Idx> rp(scaled.row_ptr().size(), 0), ci;
    std::vector<Real> vv;
    ci.reserve(static_cast<std::size_t>(scaled.nnz()));
    vv.reserve(static_cast<std::size_t>(scaled.nnz()));
    for (std::size_t r = 0; r < m; ++r) {
        for (auto k = static_cast<std::size_t>(scaled.row_ptr()[r]); k < static_cast<std::size_t>(scaled.row_ptr()[r + 1]); ++k) {
            if (std::abs(scaled.values()[k]) <= tol::numeric_zero) continue;
            ci.push_back(scaled.col_idx()[k]);
            vv.push_back(scaled.values()[k]);
        }
        rp[r + 1] = static_cast<Idx>(vv.size());
    }
    Problem sp = Problem::from_csr(CsrMatrix(p.num_rows(), p.num_cols(), std::move(rp), std::move(ci), std::move(vv)),
                                   std::move(c), std::move(clo), std::move(chi), std::move(rlo), std::move(rhi),
                                   p.sense(), p.objective_constant(), p.name());
    return {std::move(sp), std::move(s)};
}

Scaling::Conditioning Scaling::measure(const CsrMatrix& A) {
    Conditioning c;
    const auto e = A.coefficient_extremes();
    c.max_abs = e.max_abs;
    c.min_abs_nonzero = e.min_abs_nonzero;
    c.coeff_spread = e.min_abs_nonzero > 0.0 ? e.max_abs / e.min_abs_nonzero : 0.0;
    std::vector<Real> rn(static_cast<std::size_t>(A.rows()), 0.0), cn(static_cast<std::size_t>(A.cols()), 0.0);
    for (std::size_t r = 0; r < rn.size(); ++r)
        for (auto k = static_cast<std::size_t>(A.row_ptr()[r]); k < static_cast<std::size_t>(A.row_ptr()[r + 1]); ++k) {
            const Real a = std::abs(A.values()[k]);
            rn[r] = std::max(rn[r], a);
            auto& cc = cn[static_cast<std::size_t>(A.col_idx()[k])];
            cc = std::max(cc, a);
        }
    auto spread = [](const std::vector<Real>& v) {
        Real lo = kInfinity, hi = 0.0;
        for (Real x : v)
            if (x > 0.0) {
                lo = std::min(lo, x);
                hi = std::max(hi, x);
            }
        return (hi > 0.0 && std::isfinite(lo)) ? hi / lo : 0.0;
    };
    c.row_norm_spread = spread(rn);
    c.col_norm_spread = spread(cn);
    return c;
}

std::string Scaling::summary() const {
    auto mm = [](const std::vector<Real>& v) {
        Real lo = kInfinity, hi = 0.0;
        for (Real x : v) {
            lo = std::min(lo, x);
            hi = std::max(hi, x);
        }
        return std::pair<Real, Real>{lo, hi};
    };
    const auto [rlo, rhi] = mm(row_scale_);
    const auto [clo, chi] = mm(col_scale_);
    std::ostringstream os;
    os << "scaling: row factors in [" << rlo << ", " << rhi << "], column factors in [" << clo << ", " << chi << "]";
    return os.str();
}

}  // namespace sov
