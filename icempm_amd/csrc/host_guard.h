// host_guard.h -- the exception barrier of the C entry points of
// libicempm_hip.so (host code only).  Not part of the C ABI.
#pragma once

#include <exception>
#include <new>
#include <string>

#include "../../include/icempm.h"

namespace icempm {

// The body of a C entry point, run so that no C++ exception crosses the C
// interface: a failed host allocation (std::bad_alloc from a container, a
// string or new) or any other exception becomes ICEMPM_E_HOST, with its
// text in *err when there is one.
inline void note_error(std::string *err, const char *msg) noexcept {
    if (!err) return;
    try {
        *err = msg;
    } catch (...) {
    }
}
template <class F>
int host_guard(std::string *err, F &&body) noexcept {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        note_error(err, "host memory allocation failed");
    } catch (const std::exception &e) {
        note_error(err, e.what());
    } catch (...) {
        note_error(err, "unexpected host failure");
    }
    return ICEMPM_E_HOST;
}

}  // namespace icempm
