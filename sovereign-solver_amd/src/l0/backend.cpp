// Backend factory and owning handles (API: include/sovereign/backend.hpp).
// Same contract as reference src/l0/backend.cpp: DeviceBuffer zero-fills on
// allocation (:89-96), DeviceCsr::upload validates then makes exactly three
// host->device copies (:154-188), make_backend throws for a kind that is not
// available (:50-70).  Difference: Hip availability is decided at run time
// (a device must be present), see backend.hpp.
#include "sovereign/backend.hpp"

#include <utility>

#include "sov_hip.h"

namespace sov {

std::unique_ptr<Backend> make_host_backend();
std::unique_ptr<Backend> make_hip_backend(int device);

static_assert(sizeof(Idx) == 4, "the HIP C-ABI takes int32 CSR indices");

std::vector<BackendKind> available_backends() {
    std::vector<BackendKind> kinds{BackendKind::Host};
    int count = 0;
    if (sov_hip_device_count(&count) == SOV_OK && count > 0) kinds.push_back(BackendKind::Hip);
    return kinds;
}

bool is_available(BackendKind kind) {
    for (BackendKind k : available_backends())
        if (k == kind) return true;
    return false;
}

BackendKind default_backend_kind() {
    if (is_available(BackendKind::Cuda)) return BackendKind::Cuda;
    if (is_available(BackendKind::Hip)) return BackendKind::Hip;
    return BackendKind::Host;
}

std::unique_ptr<Backend> make_backend(BackendKind kind) {
    switch (kind) {
        case BackendKind::Host: return make_host_backend();
        case BackendKind::Cuda:
            throw Error("backend 'cuda' is not part of this build (MI355X/HIP only)");
        case BackendKind::Hip:
            if (!is_available(BackendKind::Hip))
                throw Error("backend 'hip' needs a visible HIP device; none was found");
            return make_hip_backend(0);
    }
    throw Error("unknown backend kind");
}

std::string backend_report() {
    std::string s = "backends:";
    const BackendKind d = default_backend_kind();
    for (BackendKind k : available_backends()) {
        s += ' ';
        s += to_string(k);
        if (k == d) s += "(default)";
    }
    s += kIndex64 ? "  index=int64" : "  index=int32";
    s += "  real=f64 (device iterates f64|f32)";
    return s;
}

DeviceBuffer::DeviceBuffer(Backend& backend, std::size_t count) : backend_(&backend), count_(count) {
    if (count == 0) return;
    void* raw = nullptr;
    check(backend.allocate(count * sizeof(Real), &raw), "DeviceBuffer::allocate");
    ptr_ = static_cast<Real*>(raw);
    check(backend.fill_zero(ptr_, count * sizeof(Real)), "DeviceBuffer::fill_zero");
}

DeviceBuffer DeviceBuffer::upload(Backend& backend, std::span<const Real> host) {
    DeviceBuffer b(backend, host.size());
    if (!host.empty())
        check(backend.copy_to_device(host.data(), b.ptr_, host.size() * sizeof(Real)), "DeviceBuffer::upload");
    return b;
}

DeviceBuffer::~DeviceBuffer() { release(); }

void DeviceBuffer::release() noexcept {
    if (backend_ != nullptr && ptr_ != nullptr) backend_->deallocate(ptr_);
    ptr_ = nullptr;
    count_ = 0;
}

DeviceBuffer::DeviceBuffer(DeviceBuffer&& o) noexcept : backend_(o.backend_), ptr_(o.ptr_), count_(o.count_) {
    o.ptr_ = nullptr;
    o.count_ = 0;
}

DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
        release();
        backend_ = o.backend_;
        ptr_ = std::exchange(o.ptr_, nullptr);
        count_ = std::exchange(o.count_, 0);
    }
    return *this;
}

Status DeviceBuffer::download(std::span<Real> dst) const {
    if (dst.size() != count_) return Status::DimensionMismatch;
    if (count_ == 0) return Status::Ok;
    return backend_->copy_to_host(ptr_, dst.data(), count_ * sizeof(Real));
}

std::vector<Real> DeviceBuffer::to_host() const {
    std::vector<Real> out(count_);
    check(download(out), "DeviceBuffer::to_host");
    return out;
}

DeviceCsr DeviceCsr::upload(Backend& backend, const CsrMatrix& m) {
    std::string why;
    const Status v = m.validate(&why);
    if (v != Status::Ok) throw Error(v, "DeviceCsr::upload: source matrix is malformed (" + why + ")");
    DeviceCsr d;
    d.backend_ = &backend;
    d.rows_ = m.rows();
    d.cols_ = m.cols();
    d.nnz_ = m.nnz();
    const std::size_t pb = m.row_ptr().size() * sizeof(Idx), ib = m.col_idx().size() * sizeof(Idx),
                      vb = m.values().size() * sizeof(Real);
    d.bytes_ = pb + ib + vb;
    void* raw = nullptr;
    check(backend.allocate(pb, &raw), "DeviceCsr: row_ptr allocate");
    d.row_ptr_ = static_cast<Idx*>(raw);
    check(backend.allocate(ib, &raw), "DeviceCsr: col_idx allocate");
    d.col_idx_ = static_cast<Idx*>(raw);
    check(backend.allocate(vb, &raw), "DeviceCsr: values allocate");
    d.values_ = static_cast<Real*>(raw);
    check(backend.copy_to_device(m.row_ptr().data(), d.row_ptr_, pb), "DeviceCsr: row_ptr upload");
    if (ib > 0) {
        check(backend.copy_to_device(m.col_idx().data(), d.col_idx_, ib), "DeviceCsr: col_idx upload");
        check(backend.copy_to_device(m.values().data(), d.values_, vb), "DeviceCsr: values upload");
    }
    return d;
}

DeviceCsr::~DeviceCsr() { release(); }

void DeviceCsr::release() noexcept {
    if (backend_ != nullptr) {
        backend_->deallocate(row_ptr_);
        backend_->deallocate(col_idx_);
        backend_->deallocate(values_);
    }
    row_ptr_ = nullptr;
    col_idx_ = nullptr;
    values_ = nullptr;
    rows_ = cols_ = nnz_ = 0;
    bytes_ = 0;
}

DeviceCsr::DeviceCsr(DeviceCsr&& o) noexcept
    : backend_(o.backend_), rows_(o.rows_), cols_(o.cols_), nnz_(o.nnz_), row_ptr_(o.row_ptr_),
      col_idx_(o.col_idx_), values_(o.values_), bytes_(o.bytes_) {
    o.row_ptr_ = nullptr;
    o.col_idx_ = nullptr;
    o.values_ = nullptr;
    o.rows_ = o.cols_ = o.nnz_ = 0;
    o.bytes_ = 0;
}

DeviceCsr& DeviceCsr::operator=(DeviceCsr&& o) noexcept {
    if (this != &o) {
        release();
        backend_ = o.backend_;
        rows_ = std::exchange(o.rows_, 0);
        cols_ = std::exchange(o.cols_, 0);
        nnz_ = std::exchange(o.nnz_, 0);
        row_ptr_ = std::exchange(o.row_ptr_, nullptr);
        col_idx_ = std::exchange(o.col_idx_, nullptr);
        values_ = std::exchange(o.values_, nullptr);
        bytes_ = std::exchange(o.bytes_, 0);
    }
    return *this;
}

}  // namespace sov
