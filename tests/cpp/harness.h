// Minimal test harness for the C++ host-API tests: TEST_CASE registers a
// function, REQUIRE / REQUIRE_THROWS_AS / REQUIRE_NOTHROW check conditions
// (reporting file:line and the expression), main() runs every test (or the
// ones whose name contains argv[1]; "--list" only lists them) and returns
// the number of failed tests.
#ifndef CUMAT_AMD_TEST_HARNESS_H
#define CUMAT_AMD_TEST_HARNESS_H

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace harness {

struct Failure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Test {
    const char* name;
    std::function<void()> fn;
};

inline std::vector<Test>& registry()
{
    static std::vector<Test> r;
    return r;
}

struct Registrar {
    Registrar(const char* name, std::function<void()> fn) { registry().push_back({name, std::move(fn)}); }
};

inline void fail(const char* file, int line, const std::string& what)
{
    std::ostringstream s;
    s << file << ":" << line << ": " << what;
    throw Failure(s.str());
}

// |a - b| <= tol * max(1, |b|)
inline bool close(double a, double b, double tol) { return std::fabs(a - b) <= tol * std::fmax(1.0, std::fabs(b)); }

}  // namespace harness

#define HARNESS_CAT2(a, b) a##b
#define HARNESS_CAT(a, b) HARNESS_CAT2(a, b)
#define TEST_CASE(name) \
    static void HARNESS_CAT(test_fn_, __LINE__)(); \
    static harness::Registrar HARNESS_CAT(test_reg_, __LINE__)(name, HARNESS_CAT(test_fn_, __LINE__)); \
    static void HARNESS_CAT(test_fn_, __LINE__)()

#define REQUIRE(cond) \
    do { \
        if (!(cond)) harness::fail(__FILE__, __LINE__, "REQUIRE(" #cond ") failed"); \
    } while (0)
#define REQUIRE_MSG(cond, msg) \
    do { \
        if (!(cond)) { \
            std::ostringstream os_; \
            os_ << "REQUIRE(" #cond ") failed: " << msg; \
            harness::fail(__FILE__, __LINE__, os_.str()); \
        } \
    } while (0)
#define REQUIRE_THROWS_AS(expr, type) \
    do { \
        bool caught_ = false; \
        try { \
            (void)(expr); \
        } catch (const type&) { \
            caught_ = true; \
        } catch (...) { \
        } \
        if (!caught_) harness::fail(__FILE__, __LINE__, "REQUIRE_THROWS_AS(" #expr ", " #type ") failed"); \
    } while (0)
#define REQUIRE_NOTHROW(expr) \
    do { \
        try { \
            (void)(expr); \
        } catch (const std::exception& e_) { \
            harness::fail(__FILE__, __LINE__, std::string("REQUIRE_NOTHROW(" #expr ") threw: ") + e_.what()); \
        } \
    } while (0)

#define HARNESS_MAIN() \
    int main(int argc, char** argv) \
    { \
        const char* filter = argc > 1 ? argv[1] : nullptr; \
        bool list = filter && std::strcmp(filter, "--list") == 0; \
        int failed = 0, ran = 0; \
        for (auto& t : harness::registry()) { \
            if (list) { \
                std::printf("%s\n", t.name); \
                continue; \
            } \
            if (filter && !std::strstr(t.name, filter)) continue; \
            ++ran; \
            std::printf("[ RUN  ] %s\n", t.name); \
            std::fflush(stdout); \
            try { \
                t.fn(); \
                std::printf("[  OK  ] %s\n", t.name); \
            } catch (const std::exception& e) { \
                ++failed; \
                std::printf("[ FAIL ] %s\n         %s\n", t.name, e.what()); \
            } \
            std::fflush(stdout); \
        } \
        if (!list) std::printf("%d test(s), %d failed\n", ran, failed); \
        return failed; \
    }

#endif
