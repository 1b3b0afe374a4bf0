// Host backend (BackendKind::Host): "device" memory is host memory, every
// primitive runs on the calling thread.  It is the parity backend: the
// primitive PDHG loop on it reproduces the reference's host PDHG bit for bit
// (reference src/l0/backend_host.cpp), so the arithmetic below keeps the
// reference's operation order:
//   * dot: pairwise summation, leaves of at most 128 products added in order;
//   * spmv: per-row gather in storage order, y = alpha*s (beta == 0, so a
//     NaN already in y is overwritten) or alpha*s + beta*y;
//   * spmv_transpose: y scaled by beta (zeroed when beta == 0), then the
//     entries scattered row by row with alpha*x[r] folded in first;
//   * project_box: lower bound first, then upper; a null bound array means
//     that side is unbounded;
//   * sort_by_key: stable.
// Host-to-"device" and back copies are counted (TransferStats) like a
// discrete device's would be.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <tuple>

#include "sovereign/backend.hpp"

namespace sov {
namespace {

/// Allocations are rounded up to and aligned on 256 bytes, like device
/// allocations, so kernels-shaped host loops see the same alignment.
constexpr std::size_t kHostAlign = 256;

Real pairwise_dot(const Real* x, const Real* y, std::size_t n) {
    if (n > 128) {
        const std::size_t half = n / 2;
        return pairwise_dot(x, y, half) + pairwise_dot(x + half, y + half, n - half);
    }
    Real leaf = 0.0;
    for (std::size_t i = 0; i < n; ++i) leaf += x[i] * y[i];
    return leaf;
}

/// Ok when n >= 0 and, for n > 0, none of the pointers is null.
Status vector_args(Idx n, std::initializer_list<const void*> ptrs) {
    if (n < 0) return Status::InvalidArgument;
    if (n > 0)
        for (const void* p : ptrs)
            if (p == nullptr) return Status::InvalidArgument;
    return Status::Ok;
}

Status csr_args(const CsrView& A, const void* x, const void* y, Idx ylen) {
    if (A.rows < 0 || A.cols < 0) return Status::InvalidArgument;
    if (ylen > 0 && y == nullptr) return Status::InvalidArgument;
    if (A.rows > 0 && A.row_ptr == nullptr) return Status::InvalidArgument;
    if (A.nnz > 0 && (x == nullptr || A.col_idx == nullptr || A.values == nullptr)) return Status::InvalidArgument;
    return Status::Ok;
}

class HostBackend final : public Backend {
public:
    std::string_view name() const noexcept override { return "host"; }
    BackendKind kind() const noexcept override { return BackendKind::Host; }
    bool has_discrete_memory() const noexcept override { return false; }
    std::string_view device_description() const noexcept override {
        return "host CPU (reference semantics, single-threaded)";
    }

    // ---- memory -----------------------------------------------------------
    Status allocate(std::size_t bytes, void** out) override {
        if (out == nullptr) return Status::InvalidArgument;
        *out = nullptr;
        if (bytes == 0) return Status::Ok;
        const std::size_t rounded = (bytes + kHostAlign - 1) / kHostAlign * kHostAlign;
        *out = std::aligned_alloc(kHostAlign, rounded);
        return *out != nullptr ? Status::Ok : Status::OutOfMemory;
    }
    void deallocate(void* ptr) noexcept override { std::free(ptr); }

    Status copy_to_device(const void* src, void* dst, std::size_t bytes) override {
        return copy(src, dst, bytes, &transfers_.h2d_calls, &transfers_.h2d_bytes);
    }
    Status copy_to_host(const void* src, void* dst, std::size_t bytes) override {
        return copy(src, dst, bytes, &transfers_.d2h_calls, &transfers_.d2h_bytes);
    }
    Status copy_device_to_device(const void* src, void* dst, std::size_t bytes) override {
        return copy(src, dst, bytes, nullptr, nullptr);
    }
    Status fill_zero(void* dst, std::size_t bytes) override {
        if (bytes != 0) {
            if (dst == nullptr) return Status::InvalidArgument;
            std::memset(dst, 0, bytes);
        }
        return Status::Ok;
    }
    Status synchronize() override { return Status::Ok; }

    // ---- level 1 ----------------------------------------------------------
    Status axpy(Idx n, Real alpha, const Real* x, Real* y) override {
        if (const Status s = vector_args(n, {x, y}); s != Status::Ok) return s;
        for (Idx i = 0; i < n; ++i) y[i] += alpha * x[i];
        return Status::Ok;
    }
    Status scale(Idx n, Real alpha, Real* x) override {
        if (const Status s = vector_args(n, {x}); s != Status::Ok) return s;
        for (Idx i = 0; i < n; ++i) x[i] *= alpha;
        return Status::Ok;
    }
    Status dot(Idx n, const Real* x, const Real* y, Real* result) override {
        if (result == nullptr) return Status::InvalidArgument;
        if (const Status s = vector_args(n, {x, y}); s != Status::Ok) return s;
        *result = pairwise_dot(x, y, static_cast<std::size_t>(n));
        return Status::Ok;
    }
    Status norm2(Idx n, const Real* x, Real* result) override {
        Real sq = 0.0;
        const Status s = dot(n, x, x, &sq);
        if (s == Status::Ok) *result = std::sqrt(sq);
        return s;
    }
    Status max_abs(Idx n, const Real* x, Real* result) override {
        if (result == nullptr) return Status::InvalidArgument;
        if (const Status s = vector_args(n, {x}); s != Status::Ok) return s;
        Real biggest = 0.0;
        for (Idx i = 0; i < n; ++i) biggest = std::max(biggest, std::abs(x[i]));
        *result = biggest;
        return Status::Ok;
    }

    // ---- sparse level 2 ---------------------------------------------------
    Status spmv(const CsrView& A, Real alpha, const Real* x, Real beta, Real* y) override {
        if (const Status s = csr_args(A, x, y, A.rows); s != Status::Ok) return s;
        const bool overwrite = beta == 0.0;
        for (Idx r = 0; r < A.rows; ++r) {
            Real row = 0.0;
            for (Idx k = A.row_ptr[r], end = A.row_ptr[r + 1]; k < end; ++k) row += A.values[k] * x[A.col_idx[k]];
            y[r] = overwrite ? alpha * row : alpha * row + beta * y[r];
        }
        return Status::Ok;
    }
    Status spmv_transpose(const CsrView& A, Real alpha, const Real* x, Real beta, Real* y) override {
        if (const Status s = csr_args(A, x, y, A.cols); s != Status::Ok) return s;
        if (beta == 0.0) std::fill(y, y + A.cols, Real{0});
        else
            for (Idx c = 0; c < A.cols; ++c) y[c] *= beta;
        for (Idx r = 0; r < A.rows; ++r) {
            const Real weight = alpha * x[r];
            for (Idx k = A.row_ptr[r], end = A.row_ptr[r + 1]; k < end; ++k) y[A.col_idx[k]] += A.values[k] * weight;
        }
        return Status::Ok;
    }
    Status project_box(Idx n, const Real* lo, const Real* hi, Real* x) override {
        if (const Status s = vector_args(n, {x}); s != Status::Ok) return s;
        if (lo != nullptr)
            for (Idx i = 0; i < n; ++i) x[i] = x[i] < lo[i] ? lo[i] : x[i];
        if (hi != nullptr)
            for (Idx i = 0; i < n; ++i) x[i] = x[i] > hi[i] ? hi[i] : x[i];
        return Status::Ok;
    }
    Status sort_by_key(Idx n, Idx* keys, Real* values) override {
        if (const Status s = vector_args(n, {keys, values}); s != Status::Ok) return s;
        std::vector<std::pair<Idx, Real>> rows(static_cast<std::size_t>(n));
        for (Idx i = 0; i < n; ++i) rows[static_cast<std::size_t>(i)] = {keys[i], values[i]};
        std::stable_sort(rows.begin(), rows.end(),
                         [](const auto& a, const auto& b) { return a.first < b.first; });
        for (Idx i = 0; i < n; ++i) std::tie(keys[i], values[i]) = rows[static_cast<std::size_t>(i)];
        return Status::Ok;
    }

private:
    /// memmove (the regions may overlap for device-to-device copies), counted
    /// when calls/bytes are given.
    static Status copy(const void* src, void* dst, std::size_t bytes, std::size_t* calls, std::size_t* total) {
        if (bytes == 0) return Status::Ok;
        if (src == nullptr || dst == nullptr) return Status::InvalidArgument;
        std::memmove(dst, src, bytes);
        if (calls != nullptr) {
            ++*calls;
            *total += bytes;
        }
        return Status::Ok;
    }
};

}  // namespace

std::unique_ptr<Backend> make_host_backend() { return std::make_unique<HostBackend>(); }

}  // namespace sov
