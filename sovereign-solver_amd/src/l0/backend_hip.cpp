// BackendKind::Hip -- the MI355X backend: a thin C++ wrapper over the C-ABI
// of include/sov_hip.h.  Replaces the reference's HIP skeleton
// (src/l0/backend_hip.cpp:56-188), whose spmv / spmv_transpose / max_abs
// returned Unsupported (:161-175): every primitive is implemented here.
//
// A CsrView carries raw device pointers, so the backend keeps a small cache
// of sov_hip_csr handles keyed by the view's arrays.  The first SpMV on a
// matrix builds its device-side A^T and row schedules (no host->device
// traffic); deallocate() of any of the matrix's arrays drops the handle.
#include <memory>
#include <string>
#include <vector>

#include "sov_hip.h"
#include "sovereign/backend.hpp"

namespace sov {
namespace {

Status to_status(int rc) { return static_cast<Status>(rc); }

class HipBackend final : public Backend {
public:
    explicit HipBackend(int device) {
        const int rc = sov_hip_ctx_create(device, &ctx_);
        if (rc != SOV_OK) throw Error(to_status(rc), std::string("HipBackend: ") + sov_hip_last_error());
        desc_ = sov_hip_ctx_device_name(ctx_);
    }
    ~HipBackend() override {
        for (auto& e : cache_) sov_hip_csr_destroy(e.h);
        sov_hip_ctx_destroy(ctx_);
    }

    std::string_view name() const noexcept override { return "hip"; }
    BackendKind kind() const noexcept override { return BackendKind::Hip; }
    bool has_discrete_memory() const noexcept override { return true; }
    std::string_view device_description() const noexcept override { return desc_; }
    void* native_context() noexcept override { return ctx_; }

    Status allocate(std::size_t bytes, void** out) override {
        if (out == nullptr) return Status::InvalidArgument;
        return to_status(sov_hip_malloc(ctx_, bytes, out));
    }
    void deallocate(void* p) noexcept override {
        if (p == nullptr) return;
        drop_handles_using(p);
        (void)sov_hip_free(ctx_, p);
    }
    Status copy_to_device(const void* s, void* d, std::size_t b) override {
        if (b == 0) return Status::Ok;
        const Status st = to_status(sov_hip_memcpy_h2d(ctx_, d, s, b));
        if (st == Status::Ok) {
            transfers_.h2d_calls += 1;
            transfers_.h2d_bytes += b;
        }
        return st;
    }
    Status copy_to_host(const void* s, void* d, std::size_t b) override {
        if (b == 0) return Status::Ok;
        const Status st = to_status(sov_hip_memcpy_d2h(ctx_, d, s, b));
        if (st == Status::Ok) {
            transfers_.d2h_calls += 1;
            transfers_.d2h_bytes += b;
        }
        return st;
    }
    Status copy_device_to_device(const void* s, void* d, std::size_t b) override {
        return to_status(sov_hip_memcpy_d2d(ctx_, d, s, b));
    }
    Status fill_zero(void* d, std::size_t b) override { return to_status(sov_hip_fill_zero(ctx_, d, b)); }
    Status synchronize() override { return to_status(sov_hip_sync(ctx_)); }

    Status axpy(Idx n, Real a, const Real* x, Real* y) override { return to_status(sov_hip_axpy_f64(ctx_, n, a, x, y)); }
    Status scale(Idx n, Real a, Real* x) override { return to_status(sov_hip_scal_f64(ctx_, n, a, x)); }
    Status dot(Idx n, const Real* x, const Real* y, Real* r) override {
        return to_status(sov_hip_dot_f64(ctx_, n, x, y, r));
    }
    Status norm2(Idx n, const Real* x, Real* r) override { return to_status(sov_hip_nrm2_f64(ctx_, n, x, r)); }
    Status max_abs(Idx n, const Real* x, Real* r) override { return to_status(sov_hip_amax_f64(ctx_, n, x, r)); }

    Status spmv(const CsrView& A, Real alpha, const Real* x, Real beta, Real* y) override {
        return product(A, 0, alpha, x, beta, y);
    }
    Status spmv_transpose(const CsrView& A, Real alpha, const Real* x, Real beta, Real* y) override {
        return product(A, 1, alpha, x, beta, y);
    }
    Status project_box(Idx n, const Real* lo, const Real* hi, Real* x) override {
        return to_status(sov_hip_project_box_f64(ctx_, n, lo, hi, x));
    }

private:
    struct Entry {
        const Idx* row_ptr;
        const Idx* col_idx;
        const Real* values;
        Idx rows, cols, nnz;
        sov_hip_csr* h;
    };

    Status product(const CsrView& A, int trans, Real alpha, const Real* x, Real beta, Real* y) {
        if (A.rows < 0 || A.cols < 0 || A.nnz < 0) return Status::InvalidArgument;
        if (A.row_ptr == nullptr) return Status::InvalidArgument;
        if ((trans ? A.cols : A.rows) > 0 && y == nullptr) return Status::InvalidArgument;
        if (A.nnz > 0 && (x == nullptr || A.col_idx == nullptr || A.values == nullptr)) return Status::InvalidArgument;
        sov_hip_csr* h = nullptr;
        const Status s = handle_for(A, &h);
        if (s != Status::Ok) return s;
        return to_status(sov_hip_spmv(ctx_, h, trans, SOV_PREC_F64, alpha, x, beta, y));
    }

    Status handle_for(const CsrView& A, sov_hip_csr** out) {
        for (const Entry& e : cache_)
            if (e.row_ptr == A.row_ptr && e.col_idx == A.col_idx && e.values == A.values && e.rows == A.rows &&
                e.cols == A.cols && e.nnz == A.nnz) {
                *out = e.h;
                return Status::Ok;
            }
        sov_hip_csr* h = nullptr;
        const int rc = sov_hip_csr_create(ctx_, A.rows, A.cols, A.nnz, A.row_ptr, A.col_idx, A.values, 0, &h);
        if (rc != SOV_OK) return to_status(rc);
        cache_.push_back(Entry{A.row_ptr, A.col_idx, A.values, A.rows, A.cols, A.nnz, h});
        *out = h;
        return Status::Ok;
    }

    void drop_handles_using(const void* p) noexcept {
        for (std::size_t i = 0; i < cache_.size();) {
            const Entry& e = cache_[i];
            if (e.row_ptr == p || e.col_idx == p || e.values == p) {
                sov_hip_csr_destroy(e.h);
                cache_[i] = cache_.back();
                cache_.pop_back();
            } else {
                ++i;
            }
        }
    }

    sov_hip_ctx* ctx_ = nullptr;
    std::string desc_;
    std::vector<Entry> cache_;
};

}  // namespace

std::unique_ptr<Backend> make_hip_backend(int device) { return std::make_unique<HipBackend>(device); }

}  // namespace sov
