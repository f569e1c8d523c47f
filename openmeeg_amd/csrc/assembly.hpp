// HeadMat assembly entry point (device).
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>

#include "device_geometry.hpp"
#include "geometry.hpp"

namespace omh {

// Assembles the symmetric HeadMat (Details::HeadMatrix + deflate,
// assembleHeadMat.cpp:28-115) in FP64 into the device matrix H (n x n,
// column-major, both halves written) in the unknown order of L, with the
// Gauss rule of the given order (Integrator(order,levels,tol),
// integrator.h:33-93; levels > 0: adaptive subdivision of the S and D pair
// integrals, 0..11).
// far_field: pairs at least FAR_NEAR diameters apart use Gauss quadrature
// on both triangles instead of the analytic inner integral (the mixed
// path; the plain FP64 path keeps the reference's integrals everywhere;
// ignored with levels > 0).
void assemble_headmat(Geometry& g, int order, double* H, const Layout& L, hipStream_t st, bool far_field = false,
                      int levels = 0, double tol = 0.0);

// OverlappingSourceMesh (OMExceptions.H): a source mesh meets the geometry.
struct OverlappingSourceMesh : std::runtime_error {
    OverlappingSourceMesh() : std::runtime_error("the source mesh intersects a mesh of the geometry") {}
};

// Mesh::intersection (mesh.cpp:246-252) of every mesh of g with the source
// mesh src (a one-mesh Geometry from make_source_mesh).
bool meshes_intersect(Geometry& g, Geometry& src, hipStream_t st);

// SurfSourceMat (assembleSourceMat.cpp:22-56) into the device matrix out
// (n x nsrc, column-major, leading dimension L.ld): rows are the unknowns
// of g in the order of L, columns the vertices of the source mesh src.
// The source mesh lies in the domain of its first vertex; for each boundary
// mesh of that domain, the N block (coefficient +-K x orientation, from an
// S block of coefficient 1) and, unless the mesh is a current barrier, the
// D block with that coefficient x -1/sigma.  Throws OverlappingSourceMesh
// when a source triangle meets a geometry triangle.
void surf_source_mat(Geometry& g, Geometry& src, int order, double* out, const Layout& L, hipStream_t st,
                     bool far_field = false);

}  // namespace omh
