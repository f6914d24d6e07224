// Exported entry points run their host-side work through zk::guard: a C++
// exception (a failed host allocation, a thread that could not be started)
// becomes an error code instead of unwinding into the C, Rust or Python
// caller, where it would end the process.
#pragma once
#include <hip/hip_runtime.h>

#include <new>

namespace zk {

template <class F>
inline hipError_t guard(F &&f) noexcept {
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return hipErrorOutOfMemory;
    } catch (...) {
        return hipErrorUnknown;
    }
}

// the same for entry points that report failure by a sentinel value
template <class T, class F>
inline T guard_or(T failed, F &&f) noexcept {
    try {
        return f();
    } catch (...) {
        return failed;
    }
}

}  // namespace zk
