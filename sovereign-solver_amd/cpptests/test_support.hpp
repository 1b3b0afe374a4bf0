// Minimal self-registering test harness for the C++ test ports (same shape
// as the reference's tests/test_support.hpp: TEST / CHECK / CHECK_NEAR /
// CHECK_THROWS / TST_MAIN).  Exit code 0 = all passed.
#pragma once

#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

namespace tst {

struct Failure {
    std::string message;
};

struct Case {
    const char* suite;
    const char* name;
    std::function<void()> body;
};

inline std::vector<Case>& registry() {
    static std::vector<Case> r;
    return r;
}

struct Registrar {
    Registrar(const char* s, const char* n, std::function<void()> b) { registry().push_back({s, n, std::move(b)}); }
};

inline int run_all(const char* filter) {
    int failed = 0, ran = 0;
    for (const Case& c : registry()) {
        const std::string full = std::string(c.suite) + "." + c.name;
        if (filter != nullptr && full.find(filter) == std::string::npos) continue;
        ++ran;
        try {
            c.body();
            std::printf("[ ok ] %s\n", full.c_str());
        } catch (const Failure& f) {
            ++failed;
            std::printf("[FAIL] %s: %s\n", full.c_str(), f.message.c_str());
        } catch (const std::exception& e) {
            ++failed;
            std::printf("[FAIL] %s: exception %s\n", full.c_str(), e.what());
        }
    }
    std::printf("%d of %d passed\n", ran - failed, ran);
    return failed == 0 ? 0 : 1;
}

}  // namespace tst

#define TST_CAT2(a, b) a##b
#define TST_CAT(a, b) TST_CAT2(a, b)
#define TEST(suite, name)                                                                          \
    static void TST_CAT(tst_body_, __LINE__)();                                                    \
    static ::tst::Registrar TST_CAT(tst_reg_, __LINE__)(#suite, #name, TST_CAT(tst_body_, __LINE__)); \
    static void TST_CAT(tst_body_, __LINE__)()
#define CHECK_MSG(cond, msg)                                                                         \
    do {                                                                                             \
        if (!(cond)) throw ::tst::Failure{std::string(__FILE__ ":") + std::to_string(__LINE__) + ": " + \
                                          #cond + " -- " + std::string(msg)};                        \
    } while (0)
#define CHECK(cond) CHECK_MSG(cond, "")
#define CHECK_EQ(a, b) CHECK_MSG((a) == (b), "values differ")
#define CHECK_NEAR(a, b, tol) \
    CHECK_MSG(std::abs(static_cast<double>(a) - static_cast<double>(b)) <= (tol), std::to_string(a) + " vs " + std::to_string(b))
#define CHECK_THROWS(expr)                     \
    do {                                       \
        bool threw_ = false;                   \
        try {                                  \
            (void)(expr);                      \
        } catch (...) {                        \
            threw_ = true;                     \
        }                                      \
        CHECK_MSG(threw_, "expected a throw"); \
    } while (0)
#define TST_MAIN() \
    int main(int argc, char** argv) { return ::tst::run_all(argc > 1 ? argv[1] : nullptr); }
