// Race integration on one device: the PDHG entrant of the reference's
// race_solve (src/l1/race.cpp:110-152, entrant :123-130) is a thread that
// makes its own backend and calls Pdhg(opt).solve(problem, *backend, &token);
// the first valid finisher cancels the other entrants' tokens
// (include/sovereign/cancellation.hpp:23-30).  On a Hip backend each entrant
// has its own context (its own stream), so two entrants share the GPU.
// These cases run two such threads on one device:
//   * a loser cancelled mid-solve returns Cancelled within one check
//     interval (at most one interval's iterations after the cancel became
//     visible) and still reports a point;
//   * the other entrant, solving concurrently, reaches the same iterations
//     and the same objective, bit for bit, as when it runs alone, and that
//     objective is the LP's exact optimum (HiGHS, tests/golden/exact_optima.json)
//     to 1e-6;
//   * the race pattern itself: first valid result wins, the loser is
//     cancelled and every thread is joined.
// Without a HIP device (SOV_TEST_REQUIRE_HIP unset or 0) the cases pass
// without running.
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <optional>
#include <thread>

#include "sovereign/backend.hpp"
#include "sovereign/cancellation.hpp"
#include "sovereign/generators.hpp"
#include "sovereign/pdhg.hpp"
#include "test_support.hpp"

using namespace sov;

namespace {

using Clock = std::chrono::steady_clock;

bool hip_here() {
    const char* req = std::getenv("SOV_TEST_REQUIRE_HIP");
    const bool need = req != nullptr && req[0] == '1';
    const bool have = is_available(BackendKind::Hip);
    if (need && !have) throw tst::Failure{"SOV_TEST_REQUIRE_HIP=1 but no HIP device is visible"};
    if (!have) std::printf("  (no HIP device: skipped)\n");
    return have;
}

/// The 2,000 x 5,000 LP (5 nonzeros per column, seed 0) and its exact
/// optimum from tests/golden/exact_optima.json ("lp_2000x5000_s0").
Problem lp_2000x5000() { return gen::random_lp(2000, 5000, 5, 0).to_problem(); }
constexpr Real kExact2000x5000 = 491.4052025167769;

PdhgOptions winner_options() {
    PdhgOptions o;
    o.engine = PdhgEngine::Fused;
    o.tolerance = 1e-6;
    o.termination = PdhgTermination::RelativeKkt;
    o.restart_scheme = PdhgRestartScheme::Adaptive;
    o.autotune = false;   // a fixed schedule: the solve is bit-reproducible
    o.max_iterations = 2000000;
    return o;
}

}  // namespace

TEST(race, cancelled_entrant_stops_within_one_check_interval_while_the_other_solves) {
    if (!hip_here()) return;
    const Problem big = gen::random_lp_zipf(100000, 200000, 1.5, 2, 200, 3).to_problem();
    const Problem small = lp_2000x5000();

    // the winner alone first: the reference run for bitwise comparison
    PdhgResult alone;
    {
        auto be = make_backend(BackendKind::Hip);
        alone = Pdhg(winner_options()).solve(small, *be);
    }
    CHECK(alone.status == PdhgStatus::Optimal);

    CancellationToken stop;
    std::atomic<long long> last_check{-1};
    PdhgOptions lo;
    lo.engine = PdhgEngine::Fused;
    lo.tolerance = 0.0;   // never converges: only the token ends it
    lo.termination = PdhgTermination::RelativeKkt;
    lo.max_iterations = 100000000;
    lo.checkpoint = [&](long long it, const std::vector<Real>&, const std::vector<Real>&, Real) {
        last_check.store(it, std::memory_order_relaxed);
    };
    // the callback at its own check (an asynchronous hand-off reports each
    // check one interval late), so `last_check` is where the loop stands
    lo.checkpoint_async = false;
    PdhgResult lost, won;
    Clock::time_point t_cancel{}, t_lost{};
    std::atomic<bool> loser_done{false};
    std::thread loser([&] {
        auto be = make_backend(BackendKind::Hip);
        lost = Pdhg(lo).solve(big, *be, &stop);
        t_lost = Clock::now();
        loser_done.store(true);
    });
    std::thread winner([&] {
        auto be = make_backend(BackendKind::Hip);
        won = Pdhg(winner_options()).solve(small, *be);
    });
    winner.join();
    // let the loser get well into its loop, then cancel it
    while (last_check.load(std::memory_order_relaxed) < 400 && !loser_done.load())
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
    stop.cancel();
    t_cancel = Clock::now();
    const long long seen = last_check.load(std::memory_order_relaxed);
    loser.join();
    const double latency = std::chrono::duration<double>(t_lost - t_cancel).count();
    std::printf("  loser: %s after %lld iterations (last check seen at cancel: iteration %lld), returned %.3f s "
                "after the cancel\n", to_string(lost.status), lost.iterations, seen, latency);
    std::printf("  winner: %s, %lld iterations (alone: %lld), objective %.15g (alone %.15g, exact %.15g)\n",
                to_string(won.status), won.iterations, alone.iterations, won.objective, alone.objective,
                kExact2000x5000);
    CHECK(lost.status == PdhgStatus::Cancelled);
    CHECK(lost.primal.size() == static_cast<std::size_t>(big.num_cols()));
    // checkpoints report iteration (count - 1); the token is read before
    // every interval, so at most the interval in flight runs after it
    CHECK_MSG(lost.iterations <= seen + 1 + lo.restart_check_period,
              std::to_string(lost.iterations) + " > " + std::to_string(seen + 1 + lo.restart_check_period));
    CHECK(won.status == PdhgStatus::Optimal);
    CHECK(won.iterations == alone.iterations);
    CHECK(won.objective == alone.objective);
    CHECK(std::abs(won.objective - kExact2000x5000) <= 1e-6 * (1.0 + std::abs(kExact2000x5000)));
}

TEST(race, first_valid_entrant_wins_and_the_other_is_cancelled) {
    if (!hip_here()) return;
    // two PDHG entrants on one device, as race_solve runs them: FP64 and
    // FP32 + FP64 polish iterates on the same LP; whichever reports a valid
    // (optimal) result first cancels the other
    const Problem p = lp_2000x5000();
    struct Shared {
        std::mutex mu;
        std::condition_variable cv;
        std::optional<int> winner;
        int remaining = 2;
    } shared;
    CancellationToken tokens[2];
    PdhgResult results[2];
    auto entrant = [&](int me) {
        PdhgOptions o = winner_options();
        o.precision = me == 0 ? Precision::F64 : Precision::F32;
        auto be = make_backend(BackendKind::Hip);
        results[me] = Pdhg(o).solve(p, *be, &tokens[me]);
        std::lock_guard<std::mutex> lock(shared.mu);
        --shared.remaining;
        if (results[me].status == PdhgStatus::Optimal && !shared.winner) {
            shared.winner = me;
            tokens[1 - me].cancel();
        }
        shared.cv.notify_all();
    };
    std::thread a(entrant, 0), b(entrant, 1);
    {
        std::unique_lock<std::mutex> lock(shared.mu);
        shared.cv.wait(lock, [&] { return shared.winner.has_value() || shared.remaining == 0; });
    }
    a.join();
    b.join();
    CHECK(shared.winner.has_value());
    const int w = *shared.winner;
    std::printf("  winner: %s iterates, %lld iterations, objective %.12g; other: %s after %lld iterations\n",
                w == 0 ? "FP64" : "FP32", results[w].iterations, results[w].objective,
                to_string(results[1 - w].status), results[1 - w].iterations);
    CHECK(results[w].status == PdhgStatus::Optimal);
    CHECK(std::abs(results[w].objective - kExact2000x5000) <= 1e-6 * (1.0 + std::abs(kExact2000x5000)));
    // the other either was cancelled or had already finished on its own
    CHECK(results[1 - w].status == PdhgStatus::Cancelled || results[1 - w].status == PdhgStatus::Optimal);
}

TST_MAIN()
