// capi.hpp - internals shared by the C ABI's source files (c_api.cpp, noalloc.hip):
// the handle, the exception -> return code translation and argument helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>

#include "../../include/oz2/oz2_c.h"
#include "engine.hpp"
#include "selector.hpp"

struct oz2_engine {
    oz2::Engine eng;
    oz2::Selector sel;
    explicit oz2_engine(hipStream_t s) : eng(s) {}
};

namespace oz2 {
namespace capi {

// message of the calling thread's last failure (oz2_last_error)
std::string& last_error();

// runs f, translating the reference's exception types into the C ABI's return codes
template <class F>
int guard(F&& f) {
    try {
        f();
        return OZ2_OK;
    } catch (const std::invalid_argument& e) {
        last_error() = e.what();
        return OZ2_INVALID_ARGUMENT;
    } catch (const std::overflow_error& e) {
        last_error() = e.what();
        return OZ2_OVERFLOW_ERROR;
    } catch (const std::logic_error& e) {
        last_error() = e.what();
        return OZ2_LOGIC_ERROR;
    } catch (const std::bad_alloc&) {
        last_error() = "oz2: host allocation failed";
        return OZ2_RUNTIME_ERROR;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return OZ2_RUNTIME_ERROR;
    } catch (...) {
        last_error() = "oz2: unknown error";
        return OZ2_RUNTIME_ERROR;
    }
}

inline Op to_op(int t) {
    if (t == OZ2_TRANS_N) return Op::N;
    if (t == OZ2_TRANS_T) return Op::T;
    if (t == OZ2_TRANS_C) return Op::C;
    throw std::invalid_argument("oz2: bad transpose code");
}

inline void need(oz2_handle h) {
    if (!h) throw std::invalid_argument("oz2: null handle");
}

// the stream waits for the caller's events first (the reference's deps vector)
inline void wait_deps(hipStream_t st, void* const* deps, int ndeps) {
    for (int i = 0; i < ndeps; ++i)
        if (deps && deps[i])
            if (hipStreamWaitEvent(st, static_cast<hipEvent_t>(deps[i]), 0) != hipSuccess)
                throw std::runtime_error("oz2: hipStreamWaitEvent failed");
}

// records the caller's completion event (the returned sycl::event)
inline void signal(hipStream_t st, void* done) {
    if (done && hipEventRecord(static_cast<hipEvent_t>(done), st) != hipSuccess)
        throw std::runtime_error("oz2: hipEventRecord failed");
}

// oz2_create calls this with the handle's device current (noalloc.hip)
void noalloc_attach(int dev);

}  // namespace capi
}  // namespace oz2
