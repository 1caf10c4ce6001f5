// Minimal test registry for the C++ host-layer tests: TEST(name) registers a
// case; expect_* record failures; main() runs all and returns non-zero on any
// failure.  Run on the GPU by tests/test_cpp_api.py.
#pragma once
#include <cmath>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace tcase {
struct Case { const char* name; std::function<void()> fn; };
inline std::vector<Case>& all() { static std::vector<Case> v; return v; }
inline int& failures() { static int f = 0; return f; }
struct Reg { Reg(const char* n, std::function<void()> f) { all().push_back({n, std::move(f)}); } };
inline void fail(const std::string& m) { throw std::runtime_error(m); }
inline void expect_true(bool c, const std::string& m) { if (!c) fail("expect_true: " + m); }
inline void expect_near(double a, double b, double tol, const std::string& m) {
    if (!(std::abs(a - b) <= tol)) fail("expect_near: " + m + " (" + std::to_string(a) + " vs " + std::to_string(b) + ")");
}
inline void expect_less(double a, double b, const std::string& m) {
    if (!(a < b)) fail("expect_less: " + m + " (" + std::to_string(a) + " !< " + std::to_string(b) + ")");
}
template <class E, class F> void expect_throw(F f, const std::string& m) {
    try { f(); } catch (const E&) { return; } catch (...) { fail("expect_throw (wrong type): " + m); }
    fail("expect_throw: " + m);
}
inline int run_all() {
    int fails = 0;
    for (auto& c : all()) {
        try { c.fn(); std::printf("[PASS] %s\n", c.name); }
        catch (const std::exception& e) { ++fails; std::printf("[FAIL] %s: %s\n", c.name, e.what()); }
        std::fflush(stdout);
    }
    std::printf("%zu tests, %d failed\n", all().size(), fails);
    return fails ? 1 : 0;
}
}  // namespace tcase
#define TEST(name) static void name(); static tcase::Reg reg_##name(#name, name); static void name()
using namespace tcase;
