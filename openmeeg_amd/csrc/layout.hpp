// Unknown-order conversions and the MFMA layout probe (device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_geometry.hpp"

namespace omh {

void sym_unpack(const double* P, int64_t n, double* H, const Layout& L, hipStream_t st);
void sym_pack(const double* H, int64_t n, const Layout& L, double* P, hipStream_t st);
void permute_rows(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t n, int64_t ncol, const Layout& L,
                  bool to_ref, hipStream_t st);
void probe_mfma_layout(double* out, hipStream_t st);

}  // namespace omh
