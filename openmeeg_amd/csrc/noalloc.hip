// Allocation-free mixed-precision solve: omh_mixed_solve with its
// refinement buffers (FP64 residual, its FP32 copy, norm partial sums) in a
// fixed device array instead of the library's growable scratch.  For
// callers that cannot allocate device memory: code recorded under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.
//
// The array holds NOALLOC_MAX_ELEMS = 2^22 elements of the n x nrhs
// residual in FP64 and in FP32, the norm partials and the buffers of the
// split-FP16 solve products (the group's split rows of W and one scale per
// row of L: at most 17.3 MB for any n x nrhs <= 2^22): 66 MiB + 8 KiB per
// device, held until the process ends.  Each device has its own copy at its
// own address, looked up once per device with hipGetSymbolAddress.  Every
// call on a device shares that device's array: calls that overlap on one
// device (other streams, other threads) race.  Nothing in it needs to start
// at zero: each call writes the residual, its FP32 copy and the partial sums
// before reading them.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "common.hpp"
#include "linalg.hpp"
#include "noalloc.hpp"

namespace {
constexpr size_t R_BYTES = sizeof(double) * (size_t)omh::NOALLOC_MAX_ELEMS;
constexpr size_t R32_BYTES = sizeof(float) * (size_t)omh::NOALLOC_MAX_ELEMS;
constexpr size_t PART_BYTES = sizeof(double) * (size_t)omh::NORM_PARTS;
constexpr size_t SPLIT_BYTES = size_t(18) << 20;
constexpr size_t SCRATCH_BYTES = R_BYTES + R32_BYTES + SPLIT_BYTES + PART_BYTES;
static_assert(R_BYTES % 256 == 0 && R32_BYTES % 256 == 0 && SPLIT_BYTES % 256 == 0,
              "sub-buffers keep the array's 256-byte alignment");
}  // namespace

// The fixed buffer (device memory of every GPU the library runs on).  In
// host code its name is only a handle for hipGetSymbolAddress.  It has
// external linkage (in a namespace of its own): with internal linkage
// (`static`) the device code object keeps it as a local symbol, which
// hipGetSymbolAddress cannot find ("invalid device symbol").  No kernel
// names it (they take its address as an argument), so `used` keeps it in the
// device code object.
namespace omh_noalloc {
__device__ __align__(256) __attribute__((used)) unsigned char scratch[SCRATCH_BYTES];
}  // namespace omh_noalloc

namespace omh {

// A kernel of this file.  The array's symbol is found only once the runtime
// has loaded this file's code object on the device: with the HIP runtime
// that PyTorch bundles, a lookup made after other files' kernels had run
// (and none of this file's) failed with "invalid device symbol".  Querying
// this kernel's attributes, or else launching it (it does nothing), loads
// that code object first.
__global__ void k_noalloc_anchor() {}

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};
}  // namespace

// Device address of this device's array, looked up on the first call made
// with the device current (stream st: where the anchor kernel runs if the
// attribute query did not load the code object).
static unsigned char* noalloc_base(hipStream_t st) {
    int dev = 0;
    OMH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) throw std::runtime_error("device index out of range for the fixed buffer");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        hipFuncAttributes fa;
        OMH_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_noalloc_anchor)));
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(omh_noalloc::scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();   // clear the failed lookup's error
            k_noalloc_anchor<<<1, 64, 0, st>>>();
            OMH_LAUNCH_CHECK();
            p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(omh_noalloc::scratch));
        }
        if (e != hipSuccess || !p)
            throw HipError(std::string("hipGetSymbolAddress of the fixed buffer: ") + hipGetErrorString(e));
        g_ptr[dev] = p;
    }
    return (unsigned char*)g_ptr[dev];
}

int mixed_refine_noalloc(const double* H, int64_t n, int64_t ldh, const float* L32, int64_t ld32, const void* work32,
                         const double* B, int64_t ldb, double* X, int64_t ldx, int64_t r, int max_iter, double tol,
                         double* final_rel_update, bool* converged, hipStream_t st) {
    if (n * r > NOALLOC_MAX_ELEMS)
        throw std::invalid_argument("n x nrhs = " + std::to_string(n * r) + " exceeds the fixed buffer's " +
                                    std::to_string(NOALLOC_MAX_ELEMS) + " elements (use omh_mixed_solve)");
    if (split_solve_bytes(n, r) > SPLIT_BYTES)   // not reached for n x r <= 2^22
        throw std::invalid_argument("the split solve buffers exceed the fixed buffer's share");
    unsigned char* base = noalloc_base(st);
    RefineBuffers buf;
    buf.R = (double*)base;
    buf.R32 = (float*)(base + R_BYTES);
    buf.split = base + R_BYTES + R32_BYTES;
    buf.part = (double*)(base + R_BYTES + R32_BYTES + SPLIT_BYTES);
    return mixed_refine_in(H, n, ldh, L32, ld32, work32, B, ldb, X, ldx, r, max_iter, tol, final_rel_update, converged,
                           buf, st, nullptr);
}

}  // namespace omh
