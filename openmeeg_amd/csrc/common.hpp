// Shared host helpers: error checking and the per-device scratch arena.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace omh {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Iterative refinement stopped without reaching its tolerance (the results
// of the last correction are still written).
struct NotConverged : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define OMH_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::omh::HipError(std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

#define OMH_LAUNCH_CHECK() OMH_HIP(hipGetLastError())

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Grow-only device scratch buffer per slot and per device, handed out for
// work on stream st.  Growth frees and allocates on st (stream ordered, the
// caller's other streams do not wait); a slot last used on another stream is
// handed out after the device has finished its work.  Calls that overlap on
// one device and use the same slot from unrelated host threads race.
enum ScratchSlot {
    SCR_ASSEMBLY = 0,
    SCR_DTMP = 1,
    SCR_SOURCES = 2,
    SCR_LINALG = 3,
    SCR_LINALG2 = 4,
    SCR_SMALL = 5,
    SCR_PAIRQ = 6,
    SCR_SRCCOEF = 7,   // source-matrix coefficients (the source matrix may run beside the factorisation)
    SCR_GEMM = 8,      // split-K partial products of the C-ABI GEMM entry points
    SCR_RESID_XT = 9,  // the refinement's transposed solution (structured residual products)
    SCR_ROWMAP = 10,   // row map of a reordered factor (segment order)
    SCR_SOLVE_SCALES = 11,   // the refinement's cache of the split solve products' scales
    SCR_NSLOTS = 12
};
void* scratch(ScratchSlot slot, size_t bytes, hipStream_t st);
void release_scratch();

// A private set of scratch slots (e.g. owned by a batch plan whose launches
// are captured into a graph: the buffers must not move once captured).  While
// an ArenaScope is alive, scratch() on this thread serves from it; a frozen
// arena refuses to grow.
struct ScratchArena {
    void* p[SCR_NSLOTS] = {};
    size_t sz[SCR_NSLOTS] = {};
    hipStream_t last[SCR_NSLOTS] = {};
    bool frozen = false;
    void release();
    ~ScratchArena() { release(); }
};
class ArenaScope {
public:
    explicit ArenaScope(ScratchArena* a);
    ~ArenaScope();
    ArenaScope(const ArenaScope&) = delete;
    ArenaScope& operator=(const ArenaScope&) = delete;
private:
    ScratchArena* prev_;
};

// Fill kernel launcher (a kernel, not hipMemsetAsync, so that it is safe in
// captured graphs).
void fill_f64(double* p, int64_t rows, int64_t cols, int64_t ld, double v, hipStream_t s);

// Optional kernel timing: a ProfScope records a HIP event pair around the
// launches it encloses on stream st, with `work` units of algorithmic work
// (flops or evaluations), aggregated per name.
bool prof_enabled();
void prof_enable(bool on);
bool prof_read(const std::string& name, int64_t& calls, double& seconds, double& work);
class ProfScope {
public:
    ProfScope(const char* name, double work, hipStream_t st);
    ~ProfScope();
    void cancel();   // drop the record (timing not wanted after all)
    ProfScope(const ProfScope&) = delete;
    ProfScope& operator=(const ProfScope&) = delete;
private:
    hipStream_t st_;
    bool active_ = false;
    std::string name_;
    double work_ = 0.0;
    hipEvent_t a_{}, b_{};
};

}  // namespace omh
