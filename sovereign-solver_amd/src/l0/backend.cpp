// Backend factory and the owning device handles (API:
// include/sovereign/backend.hpp; the reference's include/sovereign/backend.hpp
// :172-263 and src/l0/backend.cpp).
//
// Contract kept from the reference: a DeviceBuffer is zero-filled when it is
// allocated; DeviceCsr::upload checks the host matrix first and then makes
// exactly three host->device copies (row pointers, column indices, values)
// so the residency test can count them; make_backend throws for a kind that
// is not available.  This build decides Hip availability at run time (a HIP
// device must be visible) and has no CUDA backend.
//
// Both handles keep their arrays as Block members: one device allocation
// with its owning backend, released exactly once, movable.  A partly built
// DeviceCsr (an allocation or copy that throws) frees what it already holds
// through the same members.
#include "sovereign/backend.hpp"

#include <algorithm>
#include <utility>

#include "sov_hip.h"

namespace sov {

std::unique_ptr<Backend> make_host_backend();
std::unique_ptr<Backend> make_hip_backend(int device);

static_assert(sizeof(Idx) == 4, "the HIP C-ABI takes int32 CSR indices");

// ------------------------------------------------------------------ factory

std::vector<BackendKind> available_backends() {
    int devices = 0;
    const bool hip = sov_hip_device_count(&devices) == SOV_OK && devices > 0;
    return hip ? std::vector<BackendKind>{BackendKind::Host, BackendKind::Hip}
               : std::vector<BackendKind>{BackendKind::Host};
}

bool is_available(BackendKind kind) {
    const std::vector<BackendKind> kinds = available_backends();
    return std::find(kinds.begin(), kinds.end(), kind) != kinds.end();
}

BackendKind default_backend_kind() {
    // accelerators first, in the reference's order of preference
    for (const BackendKind k : {BackendKind::Cuda, BackendKind::Hip})
        if (is_available(k)) return k;
    return BackendKind::Host;
}

std::unique_ptr<Backend> make_backend(BackendKind kind) {
    if (kind == BackendKind::Host) return make_host_backend();
    if (kind == BackendKind::Hip) {
        if (!is_available(kind)) throw Error("make_backend(hip): the HIP runtime reports no device");
        return make_hip_backend(0);
    }
    if (kind == BackendKind::Cuda) throw Error("make_backend(cuda): this build targets MI355X through HIP only");
    throw Error("make_backend: unknown backend kind " + std::to_string(static_cast<int>(kind)));
}

std::string backend_report() {
    const BackendKind preferred = default_backend_kind();
    std::string line = "backends:";
    for (const BackendKind k : available_backends())
        line += std::string(" ") + to_string(k) + (k == preferred ? "(default)" : "");
    line += kIndex64 ? "  index=int64" : "  index=int32";
    line += "  real=f64 (device iterates f64|f32)";
    return line;
}

// -------------------------------------------------------------------- Block

namespace detail {

Block::Block(Backend& backend, std::size_t bytes, const char* what) : backend(&backend), bytes(bytes) {
    if (bytes == 0) return;
    const std::string context = std::string(what) + ": device allocation of " + std::to_string(bytes) + " bytes";
    check(backend.allocate(bytes, &ptr), context.c_str());
}

Block::Block(Block&& o) noexcept
    : backend(o.backend), ptr(std::exchange(o.ptr, nullptr)), bytes(std::exchange(o.bytes, 0)) {}

Block& Block::operator=(Block&& o) noexcept {
    if (this != &o) {
        reset();
        backend = o.backend;
        ptr = std::exchange(o.ptr, nullptr);
        bytes = std::exchange(o.bytes, 0);
    }
    return *this;
}

void Block::reset() noexcept {
    if (ptr != nullptr && backend != nullptr) backend->deallocate(ptr);
    ptr = nullptr;
    bytes = 0;
}

}  // namespace detail

// ------------------------------------------------------------- DeviceBuffer

DeviceBuffer::DeviceBuffer(Backend& backend, std::size_t count)
    : mem_(backend, count * sizeof(Real), "DeviceBuffer"), count_(count) {
    if (count > 0) check(backend.fill_zero(mem_.ptr, mem_.bytes), "DeviceBuffer: zero fill");
}

DeviceBuffer DeviceBuffer::upload(Backend& backend, std::span<const Real> host) {
    DeviceBuffer buf(backend, host.size());
    if (!host.empty())
        check(backend.copy_to_device(host.data(), buf.mem_.ptr, host.size_bytes()), "DeviceBuffer::upload");
    return buf;
}

DeviceBuffer::~DeviceBuffer() = default;

DeviceBuffer::DeviceBuffer(DeviceBuffer&& o) noexcept : mem_(std::move(o.mem_)), count_(std::exchange(o.count_, 0)) {}

DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
        mem_ = std::move(o.mem_);
        count_ = std::exchange(o.count_, 0);
    }
    return *this;
}

Status DeviceBuffer::download(std::span<Real> dst) const {
    if (dst.size() != count_) return Status::DimensionMismatch;
    return count_ == 0 ? Status::Ok : mem_.backend->copy_to_host(mem_.ptr, dst.data(), dst.size_bytes());
}

std::vector<Real> DeviceBuffer::to_host() const {
    std::vector<Real> host(count_);
    check(download(host), "DeviceBuffer::to_host");
    return host;
}

// ---------------------------------------------------------------- DeviceCsr

DeviceCsr DeviceCsr::upload(Backend& backend, const CsrMatrix& m) {
    std::string why;
    if (const Status s = m.validate(&why); s != Status::Ok)
        throw Error(s, "DeviceCsr::upload refuses the host matrix: " + why);
    DeviceCsr d;
    d.rows_ = m.rows();
    d.cols_ = m.cols();
    d.nnz_ = m.nnz();
    d.ptr_ = detail::Block(backend, m.row_ptr().size_bytes(), "DeviceCsr row pointers");
    d.idx_ = detail::Block(backend, m.col_idx().size_bytes(), "DeviceCsr column indices");
    d.val_ = detail::Block(backend, m.values().size_bytes(), "DeviceCsr values");
    // the three host->device copies (an empty matrix has only row pointers)
    check(backend.copy_to_device(m.row_ptr().data(), d.ptr_.ptr, d.ptr_.bytes), "DeviceCsr::upload (row pointers)");
    if (d.nnz_ > 0) {
        check(backend.copy_to_device(m.col_idx().data(), d.idx_.ptr, d.idx_.bytes), "DeviceCsr::upload (indices)");
        check(backend.copy_to_device(m.values().data(), d.val_.ptr, d.val_.bytes), "DeviceCsr::upload (values)");
    }
    return d;
}

DeviceCsr::~DeviceCsr() = default;

DeviceCsr::DeviceCsr(DeviceCsr&& o) noexcept
    : rows_(std::exchange(o.rows_, 0)), cols_(std::exchange(o.cols_, 0)), nnz_(std::exchange(o.nnz_, 0)),
      ptr_(std::move(o.ptr_)), idx_(std::move(o.idx_)), val_(std::move(o.val_)) {}

DeviceCsr& DeviceCsr::operator=(DeviceCsr&& o) noexcept {
    if (this != &o) {
        rows_ = std::exchange(o.rows_, 0);
        cols_ = std::exchange(o.cols_, 0);
        nnz_ = std::exchange(o.nnz_, 0);
        ptr_ = std::move(o.ptr_);
        idx_ = std::move(o.idx_);
        val_ = std::move(o.val_);
    }
    return *this;
}

}  // namespace sov
