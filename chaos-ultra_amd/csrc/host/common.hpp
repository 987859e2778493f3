// common.hpp — error plumbing shared by the C-ABI translation units.
//
// Internally a failing call throws CuError{status, message}; every extern "C"
// entry point runs its body through guarded(), which stores the message in a
// thread-local slot (cu_last_error) and returns the status code.
#pragma once
#include <hip/hip_runtime_api.h>

#include <exception>
#include <string>

#include "../../../include/chaosultra.h"

namespace cu {

struct CuError : std::exception {
    int status;
    std::string message;
    CuError(int s, std::string m) : status(s), message(std::move(m)) {}
    const char *what() const noexcept override { return message.c_str(); }
};

[[noreturn]] inline void fail(int status, const std::string &message) {
    throw CuError(status, message);
}

void set_last_error(const std::string &message);

// Throws CU_ERR_DEVICE (or `status`) with the HIP error string when rc != hipSuccess.
inline void check(hipError_t rc, const char *what, int status = CU_ERR_DEVICE) {
    if (rc != hipSuccess) {
        (void)hipGetLastError();  // clear the sticky per-thread error
        fail(status, std::string(what) + ": " + hipGetErrorString(rc));
    }
}

template <class F>
int guarded(F &&body) {
    try {
        body();
        set_last_error("");
        return CU_OK;
    } catch (const CuError &e) {
        set_last_error(e.message);
        return e.status;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return CU_ERR_DEVICE;
    } catch (...) {
        set_last_error("unknown error");
        return CU_ERR_DEVICE;
    }
}

// Device selected by cu_init; every device-touching call makes it current on
// the calling thread (HIP's current device is per thread).
void ensure_device();
int current_device();
int compute_units();
// Register a per-device set-up step: cu_init runs each registered step once,
// with the device it selects current (deep_static.hip registers the address
// lookup of its static scratch).  A step that throws is skipped; the code that
// registered it repeats the step on first use and reports the error there.
void on_device_init(void (*setup)(int device));

}  // namespace cu
