// The device seam.  Mirrors include/sovereign/backend.hpp:45-263 of the
// reference: same virtual interface, factory and move-only owning handles,
// so src/l1 callers (Pdhg, race, crossover) compile unchanged.
//
// BackendKind::Hip is the MI355X backend of this project: a thin C++ wrapper
// (sovereign-solver_amd/src/l0/backend_hip.cpp) over the C-ABI of
// include/sov_hip.h.  It implements every primitive (no Unsupported), keeps
// A^T as its own CSR on the device, and synchronises its own stream only.
//
// Additive:
//  * Backend::native_context() -- the sov_hip_ctx* behind a Hip backend, so
//    Pdhg::solve can drive the fused device-resident loop; nullptr elsewhere.
//  * available_backends() lists Hip only when the HIP runtime reports a
//    device at run time (the reference decides at compile time), so a build
//    without a GPU still defaults to Host instead of a backend that fails.
#pragma once

#include <cstddef>
#include <memory>
#include <span>
#include <string>
#include <string_view>
#include <vector>

#include "sovereign/numeric.hpp"
#include "sovereign/sparse.hpp"
#include "sovereign/status.hpp"

namespace sov {

enum class BackendKind { Host, Cuda, Hip };

constexpr const char* to_string(BackendKind k) noexcept {
    switch (k) {
        case BackendKind::Host: return "host";
        case BackendKind::Cuda: return "cuda";
        case BackendKind::Hip: return "hip";
    }
    return "?";
}

struct CsrView {
    Idx rows = 0;
    Idx cols = 0;
    Idx nnz = 0;
    const Idx* row_ptr = nullptr;
    const Idx* col_idx = nullptr;
    const Real* values = nullptr;
};

struct TransferStats {
    std::size_t h2d_calls = 0;
    std::size_t d2h_calls = 0;
    std::size_t h2d_bytes = 0;
    std::size_t d2h_bytes = 0;
};

class Backend {
public:
    virtual ~Backend() = default;
    Backend(const Backend&) = delete;
    Backend& operator=(const Backend&) = delete;

    virtual std::string_view name() const noexcept = 0;
    virtual BackendKind kind() const noexcept = 0;
    virtual bool has_discrete_memory() const noexcept = 0;
    virtual std::string_view device_description() const noexcept = 0;

    virtual Status allocate(std::size_t bytes, void** out) = 0;
    virtual void deallocate(void* ptr) noexcept = 0;
    virtual Status copy_to_device(const void* host_src, void* dev_dst, std::size_t bytes) = 0;
    virtual Status copy_to_host(const void* dev_src, void* host_dst, std::size_t bytes) = 0;
    virtual Status copy_device_to_device(const void* src, void* dst, std::size_t bytes) = 0;
    virtual Status fill_zero(void* dev_dst, std::size_t bytes) = 0;
    virtual Status synchronize() = 0;

    virtual Status axpy(Idx n, Real alpha, const Real* x, Real* y) = 0;
    virtual Status scale(Idx n, Real alpha, Real* x) = 0;
    virtual Status dot(Idx n, const Real* x, const Real* y, Real* result) = 0;
    virtual Status norm2(Idx n, const Real* x, Real* result) = 0;
    virtual Status max_abs(Idx n, const Real* x, Real* result) = 0;

    /// y := alpha*A*x + beta*y   (beta == 0 overwrites y, NaN-safe)
    virtual Status spmv(const CsrView& A, Real alpha, const Real* x, Real beta, Real* y) = 0;
    /// y := alpha*A^T*x + beta*y
    virtual Status spmv_transpose(const CsrView& A, Real alpha, const Real* x, Real beta,
                                  Real* y) = 0;
    /// x := clamp(x, lo, hi); a null bound is unbounded on that side.
    virtual Status project_box(Idx n, const Real* lo, const Real* hi, Real* x) = 0;

    virtual Status sort_by_key(Idx n, Idx* keys, Real* values) {
        (void)n; (void)keys; (void)values;
        return Status::NotImplemented;
    }
    virtual Status factorize_ldlt(const CsrView& A, void** factor_out) {
        (void)A; (void)factor_out;
        return Status::NotImplemented;
    }

    /// Additive: the native device context (a sov_hip_ctx* for Hip), else null.
    virtual void* native_context() noexcept { return nullptr; }

    const TransferStats& transfers() const noexcept { return transfers_; }
    void reset_transfer_stats() noexcept { transfers_ = {}; }

protected:
    Backend() = default;
    TransferStats transfers_{};
};

std::vector<BackendKind> available_backends();
bool is_available(BackendKind kind);
/// Cuda > Hip > Host among the available ones.
BackendKind default_backend_kind();
/// Throws Error when the kind is not available.
std::unique_ptr<Backend> make_backend(BackendKind kind);
std::string backend_report();

namespace detail {
/// One device allocation and the backend that owns it, released once.
struct Block {
    Backend* backend = nullptr;
    void* ptr = nullptr;
    std::size_t bytes = 0;
    Block() = default;
    Block(Backend& backend, std::size_t bytes, const char* what);
    ~Block() { reset(); }
    Block(const Block&) = delete;
    Block& operator=(const Block&) = delete;
    Block(Block&& other) noexcept;
    Block& operator=(Block&& other) noexcept;
    void reset() noexcept;
};
}  // namespace detail

class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(Backend& backend, std::size_t count);
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& other) noexcept;
    DeviceBuffer& operator=(DeviceBuffer&& other) noexcept;

    static DeviceBuffer upload(Backend& backend, std::span<const Real> host_data);
    Status download(std::span<Real> host_dst) const;
    std::vector<Real> to_host() const;

    Real* data() noexcept { return static_cast<Real*>(mem_.ptr); }
    const Real* data() const noexcept { return static_cast<const Real*>(mem_.ptr); }
    std::size_t size() const noexcept { return count_; }
    bool empty() const noexcept { return count_ == 0; }
    Backend* backend() const noexcept { return mem_.backend; }

private:
    detail::Block mem_;
    std::size_t count_ = 0;
};

class DeviceCsr {
public:
    DeviceCsr() = default;
    ~DeviceCsr();
    DeviceCsr(const DeviceCsr&) = delete;
    DeviceCsr& operator=(const DeviceCsr&) = delete;
    DeviceCsr(DeviceCsr&& other) noexcept;
    DeviceCsr& operator=(DeviceCsr&& other) noexcept;

    /// Validates, then exactly three host->device copies.
    static DeviceCsr upload(Backend& backend, const CsrMatrix& host_matrix);

    CsrView view() const noexcept {
        return CsrView{rows_, cols_, nnz_, static_cast<const Idx*>(ptr_.ptr), static_cast<const Idx*>(idx_.ptr),
                       static_cast<const Real*>(val_.ptr)};
    }
    Idx rows() const noexcept { return rows_; }
    Idx cols() const noexcept { return cols_; }
    Idx nnz() const noexcept { return nnz_; }
    std::size_t device_bytes() const noexcept { return ptr_.bytes + idx_.bytes + val_.bytes; }
    bool empty() const noexcept { return val_.ptr == nullptr; }
    Backend* backend() const noexcept { return ptr_.backend; }

private:
    Idx rows_ = 0, cols_ = 0, nnz_ = 0;
    detail::Block ptr_, idx_, val_;   // row pointers, column indices, values
};

}  // namespace sov
