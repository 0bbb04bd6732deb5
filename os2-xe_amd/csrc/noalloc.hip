// noalloc.hip - oz2_dgemm_noalloc: the DGEMM entry point that allocates nothing.
//
// For callers that cannot call hipMalloc (work under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).  It runs oz2_dgemm's path
// with every intermediate (exponents, row maxima, residue planes, residues) in
// `scratch`, a fixed device array of kScratchBytes that each GPU holds from the
// library's load until the process ends.  The panel planner cuts C into panels
// whose intermediates fit it; a product whose smallest panels do not fit is
// rejected with OZ2_INVALID_ARGUMENT before anything is enqueued.
//
// Every call on one GPU shares that GPU's array and leaves its contents behind.
// Nothing in it needs to start at zero: each launch writes all it reads later
// (the split writes its padding rows and k as zeros, the GEMM every residue of
// its padded tiles).  Calls that overlap on one GPU - from two handles on two
// streams, or two threads - race on the array; one in-order stream per GPU, or
// the caller's own ordering, keeps them apart.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <stdexcept>
#include <string>

#include "capi.hpp"

namespace {

constexpr size_t kScratchBytes = OZ2_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch size");
constexpr int kMaxDevices = 64;

// the workspace of oz2_dgemm_noalloc (one copy per GPU).  Host code never uses this
// name as an address: it is a host-side stand-in; scratch_ptr[] holds each device's
// own address of the array, from hipGetSymbolAddress.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* scratch_ptr[kMaxDevices] = {};
hipError_t scratch_err[kMaxDevices] = {};

void* scratch_of(int dev) {
    if (dev < 0 || dev >= kMaxDevices)
        throw std::runtime_error("oz2: oz2_dgemm_noalloc supports device ordinals below " +
                                 std::to_string(kMaxDevices));
    std::lock_guard<std::mutex> lock(g_mu);
    if (!scratch_ptr[dev])
        throw std::runtime_error(std::string("oz2: no scratch address for device ") + std::to_string(dev) + ": " +
                                 (scratch_err[dev] != hipSuccess ? hipGetErrorString(scratch_err[dev])
                                                                 : "the handle was created on another device"));
    return scratch_ptr[dev];
}

}  // namespace

namespace oz2 {
namespace capi {

void noalloc_attach(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return;  // oz2_dgemm_noalloc reports it
    std::lock_guard<std::mutex> lock(g_mu);
    if (scratch_ptr[dev]) return;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != dev) {
        (void)hipGetLastError();
        return;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        scratch_err[dev] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        return;
    }
    scratch_ptr[dev] = p;
}

}  // namespace capi
}  // namespace oz2

extern "C" int oz2_dgemm_noalloc(oz2_handle h, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
                                 const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                                 int64_t ldc, int num_moduli, void* const* deps, int ndeps, void* done) {
    using namespace oz2::capi;
    return guard([&] {
        need(h);
        const oz2::Op a = to_op(ta), b = to_op(tb);
        std::byte* ws = static_cast<std::byte*>(scratch_of(h->eng.device()));
        wait_deps(h->eng.stream(), deps, ndeps);
        // Trans::C equals T for real data (hpp:1001)
        h->eng.gemm_fixed(ws, kScratchBytes, a == oz2::Op::C ? oz2::Op::T : a, b == oz2::Op::C ? oz2::Op::T : b, m, n,
                          k, alpha, A, lda, B, ldb, beta, C, ldc, num_moduli);
        signal(h->eng.stream(), done);
    });
}
