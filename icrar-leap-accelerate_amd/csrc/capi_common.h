// Helpers shared by the C ABI's source files: status codes from exceptions,
// the last error message, and the device scope of a call.
#pragma once

#include <new>
#include <string>

#include "leap_internal.h"

namespace leap {

// Message of the last failure on this thread that had no context.
inline thread_local std::string g_error;

inline void set_error(const leap_ctx* ctx, const std::string& msg) {
    if (ctx)
        const_cast<leap_ctx*>(ctx)->error = msg;
    g_error = msg;
}

// Runs f and turns what it throws into a status code and a message.
template <typename F>
int guarded(const leap_ctx* ctx, F&& f) {
    try {
        f();
        return LEAP_OK;
    } catch (const Error& e) {
        set_error(ctx, e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_error(ctx, "host allocation failed");
        return LEAP_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_error(ctx, e.what());
        return LEAP_ERR_INTERNAL;
    }
}

// Makes the context's device current for the scope of a call.
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev) {
        LEAP_HIP(hipGetDevice(&prev));
        if (prev != dev) LEAP_HIP(hipSetDevice(dev));
    }
    ~DeviceScope() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

// The no-allocation entry point's scratch array (noalloc.hip): resolves its
// address on the current device, once per device.  Called by leap_hip_create.
void noalloc_init(int device);

}  // namespace leap
