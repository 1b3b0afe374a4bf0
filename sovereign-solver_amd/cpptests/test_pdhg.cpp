// PDHG engine suite: the hand-derived LPs of the reference's
// tests/test_pdhg.cpp (:29-147), run on every available backend and, on Hip,
// with both loops (primitive and fused).  Tolerance 5e-4 as in the reference.
#include <cmath>
#include <cstdlib>

#include "sovereign/backend.hpp"
#include "sovereign/io.hpp"
#include "sovereign/pdhg.hpp"
#include "test_support.hpp"

using namespace sov;

namespace {

constexpr Real kTol = 5e-4;

template <class F>
void for_each_engine(F&& body) {
    const char* req = std::getenv("SOV_TEST_REQUIRE_HIP");
    if (req != nullptr && req[0] == '1' && !is_available(BackendKind::Hip))
        throw tst::Failure{"SOV_TEST_REQUIRE_HIP=1 but no HIP device is visible"};
    for (const BackendKind kind : available_backends()) {
        for (const PdhgEngine eng : {PdhgEngine::Primitive, PdhgEngine::Fused}) {
            if (eng == PdhgEngine::Fused && kind != BackendKind::Hip) continue;
            auto be = make_backend(kind);
            try {
                body(*be, eng);
            } catch (const tst::Failure& f) {
                throw tst::Failure{std::string("[") + to_string(kind) + (eng == PdhgEngine::Fused ? "/fused] " : "/primitive] ") + f.message};
            }
        }
    }
}

PdhgResult solve_lp(Backend& be, PdhgEngine eng, const char* text, PdhgOptions opt = {}) {
    opt.engine = eng;
    return Pdhg(opt).solve(read_lp_string(text).problem, be);
}

const char* kTwoVar = "Maximize\n obj: x + y\nSubject To\n c1: x + 2 y <= 4\n c2: 4 x + 2 y <= 12\nEnd\n";

}  // namespace

TEST(pdhg, two_variable_optimum) {
    for_each_engine([](Backend& be, PdhgEngine e) {
        const PdhgResult r = solve_lp(be, e, kTwoVar);
        CHECK(r.status == PdhgStatus::Optimal);
        CHECK_NEAR(r.objective, 10.0 / 3.0, kTol);
        CHECK_NEAR(r.primal[0], 8.0 / 3.0, 1e-3);
        CHECK_NEAR(r.primal[1], 2.0 / 3.0, 1e-3);
        CHECK(r.primal_infeasibility < 1e-4);
        CHECK(r.dual_infeasibility < 1e-3);
    });
}

TEST(pdhg, binding_lower_bound) {
    for_each_engine([](Backend& be, PdhgEngine e) {
        const PdhgResult r = solve_lp(be, e, "Minimize\n obj: 3 x + 2 y\nSubject To\n c1: x + y >= 4\nEnd\n");
        CHECK(r.status == PdhgStatus::Optimal);
        CHECK_NEAR(r.objective, 8.0, kTol);
        CHECK_NEAR(r.primal[0], 0.0, 1e-3);
        CHECK_NEAR(r.primal[1], 4.0, 1e-3);
    });
}

TEST(pdhg, equality_row) {
    for_each_engine([](Backend& be, PdhgEngine e) {
        const PdhgResult r = solve_lp(be, e, "Minimize\n obj: x + y\nSubject To\n c1: x + y = 10\nBounds\n x <= 6\nEnd\n");
        CHECK(r.status == PdhgStatus::Optimal);
        CHECK_NEAR(r.objective, 10.0, kTol);
    });
}

TEST(pdhg, column_bound_and_reduced_cost_sign) {
    for_each_engine([](Backend& be, PdhgEngine e) {
        const PdhgResult r = solve_lp(be, e, "Minimize\n obj: -x\nSubject To\n c1: x + y <= 100\nBounds\n x <= 3\nEnd\n");
        CHECK(r.status == PdhgStatus::Optimal);
        CHECK_NEAR(r.objective, -3.0, kTol);
        CHECK_NEAR(r.primal[0], 3.0, 1e-3);
        CHECK(r.reduced_costs[0] < 1e-3);
    });
}

TEST(pdhg, degenerate_optimum) {
    for_each_engine([](Backend& be, PdhgEngine e) {
        const PdhgResult r = solve_lp(be, e, "Minimize\n obj: -x1 - x2\nSubject To\n c1: x1 <= 4\n c2: x1 + x2 <= 4\n c3: x2 <= 2\nEnd\n");
        CHECK(r.status == PdhgStatus::Optimal);
        CHECK_NEAR(r.objective, -4.0, kTol);
    });
}

TEST(pdhg, restart_counter_with_short_period) {
    for_each_engine([](Backend& be, PdhgEngine e) {
        PdhgOptions opt;
        opt.restart_check_period = 5;
        const PdhgResult r = solve_lp(be, e, kTwoVar, opt);
        CHECK(r.status == PdhgStatus::Optimal);
        CHECK(r.restarts >= 0);
    });
}

TEST(pdhg, time_limit_before_the_first_check_still_reports_a_point) {
    for_each_engine([](Backend& be, PdhgEngine e) {
        PdhgOptions opt;
        opt.time_limit_seconds = 1e-9;
        const PdhgResult r = solve_lp(be, e, kTwoVar, opt);
        CHECK(r.status == PdhgStatus::TimeLimit);
        CHECK(r.iterations == 0);
        CHECK(r.primal.size() == 2);
    });
}

TEST(pdhg, cancellation_stops_the_solve) {
    for_each_engine([](Backend& be, PdhgEngine e) {
        CancellationToken tok;
        tok.cancel();
        PdhgOptions opt;
        opt.engine = e;
        const PdhgResult r = Pdhg(opt).solve(read_lp_string(kTwoVar).problem, be, &tok);
        CHECK(r.status == PdhgStatus::Cancelled);
        CHECK(r.primal.size() == 2);
    });
}

TEST(pdhg, adaptive_restarts_and_relative_kkt_reach_tight_tolerance) {
    for_each_engine([](Backend& be, PdhgEngine e) {
        PdhgOptions opt;
        opt.tolerance = 1e-8;
        opt.termination = PdhgTermination::RelativeKkt;
        opt.restart_scheme = PdhgRestartScheme::Adaptive;
        opt.use_primal_weight = true;
        const PdhgResult r = solve_lp(be, e, kTwoVar, opt);
        CHECK(r.status == PdhgStatus::Optimal);
        CHECK_NEAR(r.objective, 10.0 / 3.0, 1e-7);
        CHECK(r.relative_primal_residual <= 1e-8 && r.relative_dual_residual <= 1e-8 && r.relative_gap <= 1e-8);
    });
}

TEST(pdhg, crossed_bounds_and_invalid_problems_fail_cleanly) {
    for_each_engine([](Backend& be, PdhgEngine e) {
        PdhgOptions opt;
        opt.engine = e;
        const Problem p = Problem::from_csr(CsrMatrix(1, 1, {0, 1}, {0}, {1.0}), {1.0}, {2.0}, {1.0}, {0.0}, {5.0});
        const PdhgResult r = Pdhg(opt).solve(p, be);
        CHECK(r.status == PdhgStatus::NumericalFailure);
    });
}

TST_MAIN()
