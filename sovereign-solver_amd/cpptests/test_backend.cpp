// Backend conformance suite: the hand-computed cases of the reference's
// tests/test_l0_backend.cpp (:64-384), demanded of every available backend
// (Host always; Hip when a device is visible).  Set SOV_TEST_REQUIRE_HIP=1 to
// fail instead of silently testing Host only.
//
//        [ 2 0 -1 ]
//   A =  [ 0 3  4 ]     x = (1,2,3)   -> A x   = (-1, 18, 1, 15)
//        [ 1 0  0 ]     y = (1,2,3,4) -> A^T y = (5, 6, 27)
//        [ 0 0  5 ]
#include <cstdlib>
#include <limits>
#include <type_traits>

#include "sovereign/backend.hpp"
#include "test_support.hpp"

using namespace sov;

namespace {

constexpr Real kTol = 1e-12;

CsrMatrix reference_matrix() {
    const std::vector<Triplet> t = {{0, 0, 2.0}, {0, 2, -1.0}, {1, 1, 3.0}, {1, 2, 4.0}, {2, 0, 1.0}, {3, 2, 5.0}};
    return CsrMatrix::from_triplets(4, 3, t);
}

template <class F>
void for_each_backend(F&& body) {
    const char* req = std::getenv("SOV_TEST_REQUIRE_HIP");
    if (req != nullptr && req[0] == '1' && !is_available(BackendKind::Hip))
        throw tst::Failure{"SOV_TEST_REQUIRE_HIP=1 but no HIP device is visible"};
    for (const BackendKind kind : available_backends()) {
        std::unique_ptr<Backend> be = make_backend(kind);
        try {
            body(*be);
        } catch (const tst::Failure& f) {
            throw tst::Failure{std::string("[backend=") + to_string(kind) + "] " + f.message};
        }
    }
}

}  // namespace

TEST(backend, host_is_always_available_and_default_prefers_a_device) {
    CHECK(is_available(BackendKind::Host));
    const BackendKind d = default_backend_kind();
    CHECK(is_available(BackendKind::Hip) ? d == BackendKind::Hip : d == BackendKind::Host);
    CHECK_THROWS(make_backend(BackendKind::Cuda));
    if (!is_available(BackendKind::Hip)) CHECK_THROWS(make_backend(BackendKind::Hip));
}

TEST(backend, every_backend_identifies_itself) {
    for_each_backend([](Backend& be) {
        CHECK(!be.name().empty());
        CHECK(!be.device_description().empty());
        CHECK(be.kind() == BackendKind::Host ? !be.has_discrete_memory() : be.has_discrete_memory());
    });
}

TEST(backend, buffer_roundtrip_and_zero_fill) {
    for_each_backend([](Backend& be) {
        const std::vector<Real> host = {1.0, -2.5, 3.25, 0.0, 1e-7};
        DeviceBuffer buf = DeviceBuffer::upload(be, host);
        const std::vector<Real> back = buf.to_host();
        for (std::size_t i = 0; i < host.size(); ++i) CHECK_NEAR(back[i], host[i], kTol);
        DeviceBuffer z(be, 16);
        for (Real v : z.to_host()) CHECK_NEAR(v, 0.0, kTol);
    });
}

TEST(backend, matrix_crosses_the_bus_exactly_once) {
    for_each_backend([](Backend& be) {
        const CsrMatrix host_A = reference_matrix();
        be.reset_transfer_stats();
        DeviceCsr A = DeviceCsr::upload(be, host_A);
        CHECK_EQ(be.transfers().h2d_calls, std::size_t{3});
        CHECK_EQ(be.transfers().h2d_bytes, host_A.byte_size());
        DeviceBuffer x = DeviceBuffer::upload(be, std::vector<Real>{1.0, 2.0, 3.0});
        DeviceBuffer y(be, 4);
        DeviceBuffer t(be, 3);
        const std::size_t baseline = be.transfers().h2d_calls;
        for (int i = 0; i < 32; ++i) {
            CHECK(be.spmv(A.view(), 1.0, x.data(), 0.0, y.data()) == Status::Ok);
            CHECK(be.spmv_transpose(A.view(), 1.0, y.data(), 0.0, t.data()) == Status::Ok);
        }
        CHECK_MSG(be.transfers().h2d_calls == baseline, "an SpMV re-uploaded data");
    });
}

TEST(backend, device_handles_are_move_only) {
    static_assert(!std::is_copy_constructible_v<DeviceCsr>);
    static_assert(std::is_move_constructible_v<DeviceCsr>);
    static_assert(!std::is_copy_constructible_v<DeviceBuffer>);
    for_each_backend([](Backend& be) {
        DeviceCsr a = DeviceCsr::upload(be, reference_matrix());
        const Idx nnz = a.nnz();
        DeviceCsr b = std::move(a);
        CHECK_EQ(b.nnz(), nnz);
        CHECK(a.empty());
    });
}

TEST(backend, upload_rejects_a_malformed_matrix) {
    for_each_backend([](Backend& be) {
        const CsrMatrix bad(1, 3, {0, 2}, {2, 0}, {1.0, 2.0});
        CHECK_THROWS(DeviceCsr::upload(be, bad));
    });
}

TEST(backend, level1_hand_answers) {
    for_each_backend([](Backend& be) {
        DeviceBuffer x = DeviceBuffer::upload(be, std::vector<Real>{1.0, 2.0, 3.0, 4.0});
        DeviceBuffer y = DeviceBuffer::upload(be, std::vector<Real>{10.0, 20.0, 30.0, 40.0});
        CHECK(be.axpy(4, 2.0, x.data(), y.data()) == Status::Ok);
        const std::vector<Real> after = y.to_host();
        CHECK_NEAR(after[0], 12.0, kTol);
        CHECK_NEAR(after[3], 48.0, kTol);
        CHECK(be.scale(4, 0.5, y.data()) == Status::Ok);
        CHECK_NEAR(y.to_host()[0], 6.0, kTol);
        Real d = 0;
        CHECK(be.dot(4, x.data(), x.data(), &d) == Status::Ok);
        CHECK_NEAR(d, 30.0, kTol);
        CHECK(be.norm2(4, x.data(), &d) == Status::Ok);
        CHECK_NEAR(d, std::sqrt(30.0), 1e-12);
        DeviceBuffer z = DeviceBuffer::upload(be, std::vector<Real>{1.0, -7.5, 3.0, 2.0});
        CHECK(be.max_abs(4, z.data(), &d) == Status::Ok);   // no backend may decline
        CHECK_NEAR(d, 7.5, kTol);
        Real e = 123.0;
        CHECK(be.dot(0, nullptr, nullptr, &e) == Status::Ok);
        CHECK_NEAR(e, 0.0, kTol);
        CHECK(be.axpy(0, 1.0, nullptr, nullptr) == Status::Ok);
    });
}

TEST(backend, spmv_hand_answers_alpha_beta_and_nan_overwrite) {
    for_each_backend([](Backend& be) {
        DeviceCsr A = DeviceCsr::upload(be, reference_matrix());
        DeviceBuffer x = DeviceBuffer::upload(be, std::vector<Real>{1.0, 2.0, 3.0});
        DeviceBuffer y(be, 4);
        CHECK(be.spmv(A.view(), 1.0, x.data(), 0.0, y.data()) == Status::Ok);
        CHECK(be.synchronize() == Status::Ok);
        std::vector<Real> o = y.to_host();
        CHECK_NEAR(o[0], -1.0, kTol);
        CHECK_NEAR(o[1], 18.0, kTol);
        CHECK_NEAR(o[2], 1.0, kTol);
        CHECK_NEAR(o[3], 15.0, kTol);
        DeviceBuffer y2 = DeviceBuffer::upload(be, std::vector<Real>{1.0, 1.0, 1.0, 1.0});
        CHECK(be.spmv(A.view(), 2.0, x.data(), 3.0, y2.data()) == Status::Ok);
        o = y2.to_host();
        CHECK_NEAR(o[0], 1.0, kTol);
        CHECK_NEAR(o[1], 39.0, kTol);
        CHECK_NEAR(o[2], 5.0, kTol);
        CHECK_NEAR(o[3], 33.0, kTol);
        const Real nan = std::numeric_limits<Real>::quiet_NaN();
        DeviceBuffer yn = DeviceBuffer::upload(be, std::vector<Real>{nan, nan, nan, nan});
        CHECK(be.spmv(A.view(), 1.0, x.data(), 0.0, yn.data()) == Status::Ok);
        for (Real v : yn.to_host()) CHECK_MSG(!std::isnan(v), "beta=0 failed to overwrite y");
        DeviceBuffer yt = DeviceBuffer::upload(be, std::vector<Real>{1.0, 2.0, 3.0, 4.0});
        DeviceBuffer t(be, 3);
        CHECK(be.spmv_transpose(A.view(), 1.0, yt.data(), 0.0, t.data()) == Status::Ok);
        o = t.to_host();
        CHECK_NEAR(o[0], 5.0, kTol);
        CHECK_NEAR(o[1], 6.0, kTol);
        CHECK_NEAR(o[2], 27.0, kTol);
    });
}

TEST(backend, spmv_agrees_with_the_host_container) {
    const CsrMatrix host_A = reference_matrix();
    const std::vector<Real> x = {0.5, -1.25, 2.75};
    std::vector<Real> expected(4, 0.0);
    host_A.multiply(x, expected);
    for_each_backend([&](Backend& be) {
        DeviceCsr A = DeviceCsr::upload(be, host_A);
        DeviceBuffer dx = DeviceBuffer::upload(be, x);
        DeviceBuffer dy(be, 4);
        CHECK(be.spmv(A.view(), 1.0, dx.data(), 0.0, dy.data()) == Status::Ok);
        const std::vector<Real> got = dy.to_host();
        for (std::size_t i = 0; i < expected.size(); ++i) CHECK_NEAR(got[i], expected[i], 1e-14);
    });
}

TEST(backend, project_box_clamps_and_null_is_unbounded) {
    for_each_backend([](Backend& be) {
        DeviceBuffer lo = DeviceBuffer::upload(be, std::vector<Real>{0.0, 0.0, -1.0, -5.0});
        DeviceBuffer hi = DeviceBuffer::upload(be, std::vector<Real>{1.0, 10.0, 1.0, 5.0});
        DeviceBuffer x = DeviceBuffer::upload(be, std::vector<Real>{-3.0, 4.0, 7.0, 0.5});
        CHECK(be.project_box(4, lo.data(), hi.data(), x.data()) == Status::Ok);
        const std::vector<Real> o = x.to_host();
        CHECK_NEAR(o[0], 0.0, kTol);
        CHECK_NEAR(o[1], 4.0, kTol);
        CHECK_NEAR(o[2], 1.0, kTol);
        CHECK_NEAR(o[3], 0.5, kTol);
        DeviceBuffer x2 = DeviceBuffer::upload(be, std::vector<Real>{-3.0, 1e9});
        CHECK(be.project_box(2, lo.data(), nullptr, x2.data()) == Status::Ok);
        const std::vector<Real> o2 = x2.to_host();
        CHECK_NEAR(o2[0], 0.0, kTol);
        CHECK_NEAR(o2[1], 1e9, 1e-6);
    });
}

TEST(backend, unimplemented_primitives_say_so) {
    for_each_backend([](Backend& be) {
        void* f = nullptr;
        CHECK(be.factorize_ldlt(CsrView{}, &f) == Status::NotImplemented);
    });
}

TST_MAIN()
