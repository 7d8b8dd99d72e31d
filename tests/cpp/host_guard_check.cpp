// The exception barrier of the C entry points (icempm_amd/csrc/host_guard.h):
// a body's result passes through, a failed allocation and any other
// exception become ICEMPM_E_HOST with their text.  Host code only; prints
// one line per case and exits nonzero on a mismatch.
#include <cstdio>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_guard.h"

using icempm::host_guard;

static int fails = 0;
static void expect(bool ok, const char *what) {
    std::printf("%s %s\n", ok ? "ok" : "FAIL", what);
    if (!ok) ++fails;
}

int main() {
    std::string err;
    expect(host_guard(&err, [] { return 7; }) == 7 && err.empty(), "result passes through");
    int rc = host_guard(&err, []() -> int { throw std::bad_alloc(); });
    expect(rc == ICEMPM_E_HOST && err == "host memory allocation failed", "bad_alloc");
    rc = host_guard(&err, []() -> int {
        std::vector<double> v;
        v.reserve(v.max_size());   // throws std::length_error or std::bad_alloc
        return 0;
    });
    expect(rc == ICEMPM_E_HOST && !err.empty(), "oversized container");
    rc = host_guard(&err, []() -> int { throw std::runtime_error("boom"); });
    expect(rc == ICEMPM_E_HOST && err == "boom", "std::exception text");
    rc = host_guard(&err, []() -> int { throw 3; });
    expect(rc == ICEMPM_E_HOST && err == "unexpected host failure", "other exception");
    expect(host_guard(nullptr, []() -> int { throw std::bad_alloc(); }) == ICEMPM_E_HOST, "no error text");
    return fails ? 1 : 0;
}
