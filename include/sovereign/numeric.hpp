// Scalar / index policy shared by the whole engine.
// Mirrors include/sovereign/numeric.hpp:24-109 of the reference (same names,
// same values) so code written against the reference compiles unchanged.
// Additive: Precision::F32 is no longer "reserved" -- the MI355X engine runs
// its iterates in FP32 storage when PdhgOptions::precision asks for it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>

namespace sov {

using f64 = double;
using f32 = float;

/// Host-side working precision of every public API.
using Real = f64;

enum class Precision : std::uint8_t {
    F64,   ///< FP64 storage and arithmetic
    F32,   ///< FP32 storage/arithmetic for SpMV and iterate updates, FP64 reductions
    F16,   ///< declared by the reference; not used by the PDHG engine
};

constexpr const char* to_string(Precision p) noexcept {
    return p == Precision::F64 ? "f64" : p == Precision::F32 ? "f32" : "f16";
}

constexpr std::size_t byte_width(Precision p) noexcept {
    return p == Precision::F64 ? 8 : p == Precision::F32 ? 4 : 2;
}

#if defined(SOVEREIGN_INDEX64)
using Idx = std::int64_t;
inline constexpr bool kIndex64 = true;
#else
using Idx = std::int32_t;
inline constexpr bool kIndex64 = false;
#endif

inline constexpr Idx kIdxMax = std::numeric_limits<Idx>::max();

inline constexpr std::size_t max_addressable_nnz() noexcept {
    return static_cast<std::size_t>(kIdxMax);
}

namespace tol {
inline constexpr Real primal_feasibility = 1e-8;
inline constexpr Real dual_feasibility = 1e-8;
inline constexpr Real relative_gap = 1e-9;
inline constexpr Real numeric_zero = 1e-14;
inline constexpr Real oracle_relative = 1e-9;
}  // namespace tol

/// |a-b| <= abs_tol, or <= rel * max(|a|,|b|).
inline bool close(Real a, Real b, Real rel = tol::oracle_relative,
                  Real abs_tol = tol::numeric_zero) noexcept {
    const Real d = a > b ? a - b : b - a;
    if (d <= abs_tol) return true;
    const Real aa = a < 0 ? -a : a, bb = b < 0 ? -b : b;
    return d <= rel * (aa > bb ? aa : bb);
}

}  // namespace sov
