// Host backend (BackendKind::Host).  Same semantics as reference
// src/l0/backend_host.cpp: memcpy "device" with counted transfers (:54-70),
// pairwise dot with 128-element leaves (:204-215), gather SpMV where beta == 0
// overwrites (:133-149), zero-fill-then-scatter SpMV^T (:151-172), clamp with
// null = unbounded (:174-182), stable sort_by_key (:184-200).  Single thread.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "sovereign/backend.hpp"

namespace sov {
namespace {

Real pairwise(const Real* x, const Real* y, std::size_t n) {
    if (n == 0) return 0.0;
    if (n <= 128) {
        Real s = 0.0;
        for (std::size_t i = 0; i < n; ++i) s += x[i] * y[i];
        return s;
    }
    const std::size_t h = n / 2;
    return pairwise(x, y, h) + pairwise(x + h, y + h, n - h);
}

class HostBackend final : public Backend {
public:
    std::string_view name() const noexcept override { return "host"; }
    BackendKind kind() const noexcept override { return BackendKind::Host; }
    bool has_discrete_memory() const noexcept override { return false; }
    std::string_view device_description() const noexcept override {
        return "host CPU (reference semantics, single-threaded)";
    }

    Status allocate(std::size_t bytes, void** out) override {
        if (out == nullptr) return Status::InvalidArgument;
        *out = nullptr;
        if (bytes == 0) return Status::Ok;
        void* p = std::aligned_alloc(64, (bytes + 63) / 64 * 64);
        if (p == nullptr) return Status::OutOfMemory;
        *out = p;
        return Status::Ok;
    }
    void deallocate(void* p) noexcept override { std::free(p); }

    Status copy_to_device(const void* s, void* d, std::size_t b) override {
        if (b == 0) return Status::Ok;
        if (s == nullptr || d == nullptr) return Status::InvalidArgument;
        std::memcpy(d, s, b);
        transfers_.h2d_calls += 1;
        transfers_.h2d_bytes += b;
        return Status::Ok;
    }
    Status copy_to_host(const void* s, void* d, std::size_t b) override {
        if (b == 0) return Status::Ok;
        if (s == nullptr || d == nullptr) return Status::InvalidArgument;
        std::memcpy(d, s, b);
        transfers_.d2h_calls += 1;
        transfers_.d2h_bytes += b;
        return Status::Ok;
    }
    Status copy_device_to_device(const void* s, void* d, std::size_t b) override {
        if (b == 0) return Status::Ok;
        if (s == nullptr || d == nullptr) return Status::InvalidArgument;
        std::memmove(d, s, b);
        return Status::Ok;
    }
    Status fill_zero(void* d, std::size_t b) override {
        if (b == 0) return Status::Ok;
        if (d == nullptr) return Status::InvalidArgument;
        std::memset(d, 0, b);
        return Status::Ok;
    }
    Status synchronize() override { return Status::Ok; }

    Status axpy(Idx n, Real a, const Real* x, Real* y) override {
        if (n < 0 || (n > 0 && (x == nullptr || y == nullptr))) return Status::InvalidArgument;
        for (Idx i = 0; i < n; ++i) y[i] += a * x[i];
        return Status::Ok;
    }
    Status scale(Idx n, Real a, Real* x) override {
        if (n < 0 || (n > 0 && x == nullptr)) return Status::InvalidArgument;
        for (Idx i = 0; i < n; ++i) x[i] *= a;
        return Status::Ok;
    }
    Status dot(Idx n, const Real* x, const Real* y, Real* r) override {
        if (n < 0 || r == nullptr || (n > 0 && (x == nullptr || y == nullptr))) return Status::InvalidArgument;
        *r = pairwise(x, y, static_cast<std::size_t>(n));
        return Status::Ok;
    }
    Status norm2(Idx n, const Real* x, Real* r) override {
        Real d = 0.0;
        const Status s = dot(n, x, x, &d);
        if (s != Status::Ok) return s;
        *r = std::sqrt(d);
        return Status::Ok;
    }
    Status max_abs(Idx n, const Real* x, Real* r) override {
        if (n < 0 || r == nullptr || (n > 0 && x == nullptr)) return Status::InvalidArgument;
        Real m = 0.0;
        for (Idx i = 0; i < n; ++i) m = std::max(m, std::abs(x[i]));
        *r = m;
        return Status::Ok;
    }

    Status spmv(const CsrView& A, Real alpha, const Real* x, Real beta, Real* y) override {
        if (A.rows < 0 || A.cols < 0) return Status::InvalidArgument;
        if (A.rows > 0 && (y == nullptr || A.row_ptr == nullptr)) return Status::InvalidArgument;
        if (A.nnz > 0 && (x == nullptr || A.col_idx == nullptr || A.values == nullptr)) return Status::InvalidArgument;
        for (Idx r = 0; r < A.rows; ++r) {
            Real s = 0.0;
            for (Idx k = A.row_ptr[r]; k < A.row_ptr[r + 1]; ++k) s += A.values[k] * x[A.col_idx[k]];
            y[r] = beta == 0.0 ? alpha * s : alpha * s + beta * y[r];
        }
        return Status::Ok;
    }
    Status spmv_transpose(const CsrView& A, Real alpha, const Real* x, Real beta, Real* y) override {
        if (A.rows < 0 || A.cols < 0) return Status::InvalidArgument;
        if (A.cols > 0 && y == nullptr) return Status::InvalidArgument;
        if (A.nnz > 0 && (x == nullptr || A.col_idx == nullptr || A.values == nullptr)) return Status::InvalidArgument;
        for (Idx c = 0; c < A.cols; ++c) y[c] = beta == 0.0 ? 0.0 : y[c] * beta;
        for (Idx r = 0; r < A.rows; ++r) {
            const Real ax = alpha * x[r];
            for (Idx k = A.row_ptr[r]; k < A.row_ptr[r + 1]; ++k) y[A.col_idx[k]] += A.values[k] * ax;
        }
        return Status::Ok;
    }
    Status project_box(Idx n, const Real* lo, const Real* hi, Real* x) override {
        if (n < 0 || (n > 0 && x == nullptr)) return Status::InvalidArgument;
        for (Idx i = 0; i < n; ++i) {
            if (lo != nullptr && x[i] < lo[i]) x[i] = lo[i];
            if (hi != nullptr && x[i] > hi[i]) x[i] = hi[i];
        }
        return Status::Ok;
    }
    Status sort_by_key(Idx n, Idx* keys, Real* values) override {
        if (n < 0 || (n > 0 && (keys == nullptr || values == nullptr))) return Status::InvalidArgument;
        std::vector<std::size_t> ord(static_cast<std::size_t>(n));
        std::iota(ord.begin(), ord.end(), std::size_t{0});
        std::stable_sort(ord.begin(), ord.end(), [&](std::size_t a, std::size_t b) { return keys[a] < keys[b]; });
        std::vector<Idx> k2(ord.size());
        std::vector<Real> v2(ord.size());
        for (std::size_t i = 0; i < ord.size(); ++i) {
            k2[i] = keys[ord[i]];
            v2[i] = values[ord[i]];
        }
        std::copy(k2.begin(), k2.end(), keys);
        std::copy(v2.begin(), v2.end(), values);
        return Status::Ok;
    }
};

}  // namespace

std::unique_ptr<Backend> make_host_backend() { return std::make_unique<HostBackend>(); }

}  // namespace sov
