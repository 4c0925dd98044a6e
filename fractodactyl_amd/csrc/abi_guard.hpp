// Exception barrier of the C ABI: every exported entry point runs its body through
// guarded(), so no C++ exception unwinds into the caller (a Rust or C frame, where that
// is undefined behaviour).  std::bad_alloc becomes FDG_ENOMEM and any other exception
// FDG_EDEVICE, with the message kept for fdg_last_error.
#pragma once
#include <exception>
#include <new>
#include <string>

#include "fdg.h"

namespace fdg {

// fdg_last_error(NULL): the message of the calling thread's last failed call that had no
// context (fdg_create, host-only helpers).
extern thread_local std::string g_create_error;

inline void note_error(std::string* err, const char* m) noexcept {
    try {
        if (err)
            *err = m;
        else
            g_create_error = m;
    } catch (...) {
        // no memory left even for the message: the status code still tells
    }
}

template <class R, class F>
R guarded(std::string* err, F&& body) noexcept {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        note_error(err, "out of host memory");
        return (R)FDG_ENOMEM;
    } catch (const std::exception& e) {
        note_error(err, e.what());
        return (R)FDG_EDEVICE;
    } catch (...) {
        note_error(err, "unexpected C++ exception");
        return (R)FDG_EDEVICE;
    }
}

// Teardown and setters without a status: an exception there is dropped.
template <class F>
void guarded_void(F&& body) noexcept {
    try {
        body();
    } catch (...) {
    }
}

}  // namespace fdg
