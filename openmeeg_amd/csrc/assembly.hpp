// HeadMat assembly entry point (device).
#pragma once

#include <hip/hip_runtime.h>

#include "device_geometry.hpp"
#include "geometry.hpp"

namespace omh {

// Assembles the symmetric HeadMat (Details::HeadMatrix + deflate,
// assembleHeadMat.cpp:28-115) in FP64 into the device matrix H (n x n,
// column-major, both halves written) in the unknown order of L, with the
// non-adaptive Gauss rule of the given order (Integrator(order,0,.),
// integrator.h:33-48).
// far_field: pairs at least FAR_NEAR diameters apart use Gauss quadrature
// on both triangles instead of the analytic inner integral (the mixed
// path; the plain FP64 path keeps the reference's integrals everywhere).
void assemble_headmat(Geometry& g, int order, double* H, const Layout& L, hipStream_t st, bool far_field = false);

}  // namespace omh
