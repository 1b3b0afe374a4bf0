// Ruiz equilibration then Pock-Chambolle diagonal preconditioning.
// API mirrors include/sovereign/scaling.hpp:39-103 of the reference.
// The host implementation here is used by the primitive loop; the fused
// device loop computes the same factors with HIP kernels (sov_hip_scale_*)
// and wraps them with Scaling::from_factors (additive).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "sovereign/numeric.hpp"
#include "sovereign/problem.hpp"
#include "sovereign/sparse.hpp"

namespace sov {

struct ScalingOptions {
    int ruiz_iterations = 8;
    Real ruiz_tolerance = 1e-2;
    Real pock_chambolle_alpha = 1.0;
    bool apply_ruiz = true;
    bool apply_pock_chambolle = true;
};

/// A_scaled = diag(row_scale) * A * diag(col_scale).
class Scaling {
public:
    Scaling() = default;
    static Scaling equilibrate(const CsrMatrix& matrix, const ScalingOptions& options = {});
    static std::pair<Problem, Scaling> equilibrate(const Problem& problem,
                                                   const ScalingOptions& options = {});
    /// Additive: wrap factors computed elsewhere (the device scaling kernels).
    static Scaling from_factors(std::vector<Real> row_scale, std::vector<Real> col_scale);

    Idx num_rows() const noexcept { return static_cast<Idx>(row_scale_.size()); }
    Idx num_cols() const noexcept { return static_cast<Idx>(col_scale_.size()); }
    std::span<const Real> row_scale() const noexcept { return row_scale_; }
    std::span<const Real> col_scale() const noexcept { return col_scale_; }

    CsrMatrix apply(const CsrMatrix& matrix) const;
    void unscale_primal(std::span<Real> x_in_scaled_space) const;
    void unscale_dual(std::span<Real> y_in_scaled_space) const;

    struct Conditioning {
        Real max_abs = 0.0;
        Real min_abs_nonzero = 0.0;
        Real coeff_spread = 0.0;
        Real row_norm_spread = 0.0;
        Real col_norm_spread = 0.0;
    };
    static Conditioning measure(const CsrMatrix& matrix);
    std::string summary() const;

private:
    std::vector<Real> row_scale_;
    std::vector<Real> col_scale_;
};

}  // namespace sov
