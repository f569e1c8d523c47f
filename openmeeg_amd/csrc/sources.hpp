// Dipole right-hand side, point location and EEG sensor projection (device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device_geometry.hpp"
#include "geometry.hpp"

namespace omh {

// Geometry::domain(point) for npts device points (stride doubles apart):
// domain index per point (geometry.cpp:146-154).
std::vector<int> locate_domains(Geometry& g, const double* d_pts, int64_t stride, int64_t npts, hipStream_t st);

// SourceMatrix<Dipole> (assembleSourceMat.cpp:66-104): d_dip is n_dip x 6
// row-major on the device (x y z qx qy qz); domain[s] the domain of dipole s.
// Writes the n x n_dip matrix d_out (column-major, leading dimension ld_out)
// with rows in the unknown order of L.
void dip_source_mat(Geometry& g, const double* d_dip, int64_t n_dip, const std::vector<int>& domain, int order,
                    int levels, double tol, double* d_out, int64_t ld_out, const Layout& L, hipStream_t st);

// The two halves of dip_source_mat: per (dipole, mesh) coefficients on the
// host (cD, cP: n_dip x n_meshes row-major; used[m]: mesh m has a non-zero
// coefficient), then the device launches with those coefficients already
// in device memory (no host synchronisation: capture-safe).
void dsm_coefficients(const Geometry& g, int64_t n_dip, const std::vector<int>& domain, std::vector<double>& cD,
                      std::vector<double>& cP, std::vector<char>& used);
void dsm_launch(Geometry& g, const double* d_dip, int64_t n_dip, const double* d_cD, const double* d_cP,
                const std::vector<char>& used, int order, int levels, double tol, double* d_out, int64_t ld_out,
                const Layout& L, hipStream_t st);

// Adjoint of SourceMatrix<Dipole> with the non-adaptive rule of `order`:
// d_grad (n_dip x 6, row-major) = for each dipole s the gradient of
// W(:,s) . rhs_s with respect to (x y z) and (qx qy qz), W being n x n_dip
// (column-major, leading dimension ldw, rows in the unknown order of L).
void dip_source_adjoint(Geometry& g, const double* d_dip, int64_t n_dip, const std::vector<int>& domain, int order,
                        const double* d_W, int64_t ldw, const Layout& L, double* d_grad, hipStream_t st);

// Head2EEGMat (assembleSensors.cpp:25-39): for each electrode (d_el, n_el x 3
// row-major on the device) the three vertex unknowns (reference order) of the
// closest triangle and their barycentric weights (host arrays n_el x 3).
void head2eeg(Geometry& g, const double* d_el, int64_t n_el, int64_t* cols, double* vals, hipStream_t st);

// Head2MEGMat before the sensor weights (assembleSensors.cpp:69-99): for
// npts integration points (d_pts, d_dirs: npts x 3 row-major on the device)
// the npts x n matrix d_out (column-major, leading dimension ld_out) mapping
// the unknowns (order of L; triangle columns zero) to the field along each
// point's direction, MagFactor x each mesh's conductivity jump x the
// Ferguson operator.
void head2meg(Geometry& g, const double* d_pts, const double* d_dirs, int64_t npts, double* d_out, int64_t ld_out,
              const Layout& L, hipStream_t st);

// SurfSource2MEGMat before the sensor weights (assembleSensors.cpp:102-124):
// d_out (npts x nsrc, column-major) = the Ferguson operator of the source
// mesh src (a one-mesh Geometry) with coefficient 1 at each point along its
// direction; columns are the source mesh's vertices.
void surf2meg(Geometry& src, const double* d_pts, const double* d_dirs, int64_t npts, double* d_out, int64_t ld_out,
              hipStream_t st);

// Surf2VolMat (assembleHeadMat.cpp:325-363): for the npts points d_pts
// (npts x 3 row-major) the rows of the points that lie in a domain of
// non-zero conductivity (the others are dropped; their input indices are
// returned, in order), n columns in the unknown order of L, written to
// d_out (leading dimension L.ld >= number of kept points).
std::vector<int64_t> surf2vol(Geometry& g, const double* d_pts, int64_t npts, double* d_out, const Layout& L,
                              hipStream_t st);

// DipSource2InternalPotMat (assembleSourceMat.cpp:154-186): kept points x
// n_dip, K / sigma x the dipole potential at the points of the dipole's
// domain (domain[j]).  Returns the kept points as surf2vol.
std::vector<int64_t> dip2internal_pot(Geometry& g, const double* d_dip, int64_t n_dip, const std::vector<int>& domain,
                                      const double* d_pts, int64_t npts, double* d_out, int64_t ld_out,
                                      hipStream_t st);

// EITSourceMat (assembleSourceMat.cpp:109-152): d_out (n x ne, column-major,
// leading dimension L.ld, rows in the unknown order of L) for EIT
// electrodes at d_pos (ne x 3 row-major, device) with radii (host, NULL =
// all zero: currents given as intensities): for each injection triangle t
// of an electrode (Sensors::findInjectionTriangles, sensors.cpp:167-220),
// weight x row t of the transfer matrix [D*_23 S_23 -I_33/2 + D*_33]
// built from the mesh pairs whose first mesh is t's current-barrier mesh.
void eit_source_mat(Geometry& g, const double* d_pos, const double* radii, int64_t ne, int order, double* d_out,
                    const Layout& L, hipStream_t st);

// DipSource2MEGMat before the sensor weights (assembleSensors.cpp:133-156):
// d_out (npts x n_dip, column-major) for dipoles d_dip (n_dip x 6 row-major).
void dip2meg(const double* d_dip, int64_t n_dip, const double* d_pts, const double* d_dirs, int64_t npts,
             double* d_out, int64_t ld_out, hipStream_t st);

}  // namespace omh
