// Unknown-order conversions between the reference's storage (packed upper
// SymMatrix, symmatrix.h:52-62; vertices-first numbering, geometry.cpp:256-295)
// and the device matrices, plus the MFMA lane-layout probe.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "device_geometry.hpp"
#include "layout.hpp"

namespace omh {

namespace {

inline unsigned grid_for(int64_t total) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(total, 256), 16384)); }

__device__ __forceinline__ int64_t pidx(int64_t i, int64_t j) {
    if (i > j) { const int64_t t = i; i = j; j = t; }
    return i + j * (j + 1) / 2;
}

__global__ void k_unpack(const double* __restrict__ P, int64_t n, double* __restrict__ H, Layout L) {
    const int64_t total = n * n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / n, i = t - j * n;   // reference indices
        H[dev_index(L, i) + dev_index(L, j) * L.ld] = P[pidx(i, j)];
    }
}

__global__ void k_pack(const double* __restrict__ H, int64_t n, Layout L, double* __restrict__ P) {
    const int64_t total = n * (n + 1) / 2;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (int64_t)gridDim.x * blockDim.x) {
        int64_t j = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
        while (j * (j + 1) / 2 > p) --j;
        while ((j + 1) * (j + 2) / 2 <= p) ++j;
        const int64_t i = p - j * (j + 1) / 2;
        // read the lower triangle of the device matrix (valid after a factor
        // that left only the upper one intact is not assumed here)
        int64_t di = dev_index(L, i), dj = dev_index(L, j);
        if (di < dj) { const int64_t t2 = di; di = dj; dj = t2; }
        P[p] = H[di + dj * L.ld];
    }
}

__global__ void k_permute_rows(const double* __restrict__ src, int64_t lds, double* __restrict__ dst, int64_t ldd,
                               int64_t n, int64_t ncol, Layout L, int to_ref) {
    const int64_t total = n * ncol;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = t / n, i = t - c * n;   // i: reference row
        const int64_t d = dev_index(L, i);
        if (to_ref) dst[i + c * ldd] = src[d + c * lds];
        else dst[d + c * ldd] = src[i + c * lds];
    }
}

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// One wave: D = A B with A(i,k) = 4i + k + 1, B(k,j) = (k+1)(j%5+1); the raw
// accumulator element r of lane l goes to out[4l + r] (FP64) and
// out[256 + 4l + r] (FP32), so that the host can read off the layout.
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));

__global__ void k_probe(double* out) {
    const int l = threadIdx.x;
    const int i = l % 16, k = l / 16, j = l % 16;
    const double a = 4.0 * i + k + 1.0;
    const double b = (k + 1.0) * (j % 5 + 1.0);
    double4_t c64 = {0, 0, 0, 0};
    c64 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c64, 0, 0, 0);
    float4_t c32 = {0, 0, 0, 0};
    c32 = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a, (float)b, c32, 0, 0, 0);
    // fp16 16x16x32: lane l holds A(l % 16, 8 (l / 16) + e) and
    // B(8 (l / 16) + e, l % 16), e = 0..7 (the operand layout under test)
    half8_t a16, b16;
    for (int e = 0; e < 8; ++e) {
        const int kk = 8 * (l / 16) + e;
        a16[e] = (_Float16)(float)((i + 2 * kk) % 7 - 3);
        b16[e] = (_Float16)(float)((3 * kk + j) % 5 - 2);
    }
    float4_t c16 = {0, 0, 0, 0};
    c16 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a16, b16, c16, 0, 0, 0);
    for (int r = 0; r < 4; ++r) {
        out[4 * l + r] = c64[r];
        out[256 + 4 * l + r] = (double)c32[r];
        out[512 + 4 * l + r] = (double)c16[r];
    }
}

}  // namespace

void sym_unpack(const double* P, int64_t n, double* H, const Layout& L, hipStream_t st) {
    k_unpack<<<grid_for(n * n), 256, 0, st>>>(P, n, H, L);
    OMH_LAUNCH_CHECK();
}

void sym_pack(const double* H, int64_t n, const Layout& L, double* P, hipStream_t st) {
    k_pack<<<grid_for(n * (n + 1) / 2), 256, 0, st>>>(H, n, L, P);
    OMH_LAUNCH_CHECK();
}

void permute_rows(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t n, int64_t ncol, const Layout& L,
                  bool to_ref, hipStream_t st) {
    k_permute_rows<<<grid_for(n * ncol), 256, 0, st>>>(src, lds, dst, ldd, n, ncol, L, to_ref ? 1 : 0);
    OMH_LAUNCH_CHECK();
}

void probe_mfma_layout(double* out, hipStream_t st) {
    k_probe<<<1, 64, 0, st>>>(out);
    OMH_LAUNCH_CHECK();
}

}  // namespace omh
