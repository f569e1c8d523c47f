/*
 * openmeeg_amd: MI355X-native symmetric-BEM EEG forward path, C ABI.
 *
 * Every entry point returns 0 on success and a non-zero status otherwise;
 * omh_last_error() then gives the message (per thread).  No C++ exception
 * crosses this boundary.  Device pointers are HIP device memory; `stream` is
 * a hipStream_t (NULL = the null stream).  Dense matrices are column-major.
 *
 * Unknown orderings of device matrices (`ordering`):
 *   OMH_ORDER_REFERENCE       vertices first, then triangles (the reference's
 *                             numbering, OpenMEEG/src/geometry.cpp:256-295);
 *   OMH_ORDER_TRIANGLES_FIRST the same numbering rotated so that triangle
 *                             unknowns come first; the HeadMat factors without
 *                             pivoting in this order.
 *
 * Each declaration cites the reference interface it replaces (file:line in
 * OpenMEEG 2.5.16).
 */
#ifndef OPENMEEG_AMD_H
#define OPENMEEG_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMH_ORDER_REFERENCE 0
#define OMH_ORDER_TRIANGLES_FIRST 1

/* Status returned when a source mesh intersects the geometry
 * (OverlappingSourceMesh, OMExceptions.H). */
#define OMH_STATUS_OVERLAPPING_SOURCE_MESH 5
/* Status of a mixed-precision solve whose iterative refinement stopped
 * without reaching its tolerance (see omh_refine_state); the outputs of
 * the last correction are written all the same. */
#define OMH_STATUS_NOT_CONVERGED 6

typedef struct omh_geometry omh_geometry;
typedef struct omh_mesh omh_mesh;

/* Message of the last failed call on this thread (never NULL). */
const char* omh_last_error(void);

/* ---------------------------------------------------------------- geometry */

/* Geometry(geomFileName, condFileName): OpenMEEG/include/geometry.h:67-69,
 * GeometryIOs/GeomFile.h:158-300 (format 1.1, .tri meshes). */
int omh_geometry_load(const char* geom_file, const char* cond_file, omh_geometry** out);

/* make_nested_geometry(meshes, conductivity):
 * wrapping/python/openmeeg/_make_geometry.py:131-211.  Meshes inner to
 * outer; vertices[l] is n_vertices[l] x 3 row-major, triangles[l] is
 * n_triangles[l] x 3 (0-based, mesh-local).  1 or 3 layers name the domains
 * as the reference (Brain/Skull/Scalp/Air); other counts use Brain,
 * Layer1..Layer{L-1}, Air. */
int omh_geometry_nested(int n_layers, const int64_t* n_vertices, const double* const* vertices,
                        const int64_t* n_triangles, const int64_t* const* triangles,
                        const double* conductivities, omh_geometry** out);

/* make_geometry(meshes, interfaces, domains):
 * wrapping/python/openmeeg/_make_geometry.py:15-128.  Interface i is the
 * list interface_meshes[interface_offsets[i] .. interface_offsets[i+1]) with
 * orientations +1/-1; domain d is bounded by domain_interfaces[...] with
 * domain_inside[...] = 1 for SimpleDomain::Inside, 0 for Outside. */
int omh_geometry_create(int n_meshes, const char* const* mesh_names, const int64_t* n_vertices,
                        const double* const* vertices, const int64_t* n_triangles,
                        const int64_t* const* triangles, int n_interfaces, const char* const* interface_names,
                        const int64_t* interface_offsets, const int32_t* interface_meshes,
                        const int32_t* interface_orientations, int n_domains, const char* const* domain_names,
                        const double* domain_conductivities, const int64_t* domain_offsets,
                        const int32_t* domain_interfaces, const int32_t* domain_inside, omh_geometry** out);

void omh_geometry_free(omh_geometry* g);

/* out[0] size = nb_parameters - nb_current_barrier_triangles (HeadMat order,
 * assembleHeadMat.cpp:77), out[1] nb_parameters (geometry.h:160),
 * out[2] nb_current_barrier_triangles (geometry.h:253), out[3] vertex
 * unknowns, out[4] non-barrier triangle unknowns, out[5] meshes,
 * out[6] domains, out[7] is_nested (geometry.h:91). */
int omh_geometry_sizes(const omh_geometry* g, int64_t out[8]);

/* Geometry::domain(name) index (geometry.cpp:156-165); domain name by index. */
int omh_geometry_domain_index(const omh_geometry* g, const char* name, int32_t* index);
int omh_geometry_domain_name(const omh_geometry* g, int32_t index, char* buf, int64_t buflen);

/* ---------------------------------------------------------------- assembly */

/* HeadMat(geo, Integrator(order,0,.)): OpenMEEG/include/assemble.h:32,
 * src/assembleHeadMat.cpp:28-135.  Writes the full symmetric n x n FP64
 * matrix (both halves) into d_H (leading dimension ld >= n) in `ordering`.
 * Only non-adaptive integration (levels = 0, the reference default) runs
 * here; order 0..3 selects the 3/6/7/16-point rule. */
int omh_head_mat(omh_geometry* g, int order, double* d_H, int64_t ld, int ordering, void* stream);

/* omh_head_mat with flags.  OMH_HEADMAT_FAR_FIELD (no reference
 * counterpart; for the mixed-precision path): triangle pairs whose centroid
 * distance is at least 4 times the longer of their longest edges use Gauss
 * quadrature on both triangles (16 inner points, 7 from 10 diameters on)
 * instead of the analytic inner integral.  Relative change of such entries:
 * below 1e-10 (S) and at the analytic formula's own cancellation level
 * (D).  Order 3 only; other orders ignore the flag. */
#define OMH_HEADMAT_FAR_FIELD 1
int omh_head_mat_ex(omh_geometry* g, int order, int32_t flags, double* d_H, int64_t ld, int ordering, void* stream);

/* HeadMat(geo, Integrator(order, levels, tol)) with adaptive integration
 * (assemble.h:32, integrator.h:44-93): every S and D pair integral is
 * refined by 4-way subdivision of its Gauss triangle while the refined sum
 * differs from the coarse one by more than tol (relative), to at most
 * `levels` (0..11) subdivisions; levels = 0 is omh_head_mat. */
int omh_head_mat_integrator(omh_geometry* g, int order, int levels, double tol, double* d_H, int64_t ld, int ordering,
                            void* stream);

/* Geometry::domain(point) for n device points (x y z every `stride`
 * doubles): geometry.cpp:146-154.  Writes domain indices to h_domains. */
int omh_locate_domains(omh_geometry* g, const double* d_points, int64_t stride, int64_t n, int32_t* h_domains,
                       void* stream);

/* SourceMatrix<Dipole>(geo, dipoles, Integrator(order,levels,tol), domain):
 * assembleSourceMat.h:24-29, assembleSourceMat.cpp:66-104; the Python
 * DipSourceMat of wrapping/python/openmeeg/_openmeeg.i:399-413.  d_dipoles
 * is n_dip x 6 row-major (x y z qx qy qz).  h_domains gives the domain of
 * every dipole, or NULL to locate each dipole (domain name ""). */
int omh_dip_source_mat(omh_geometry* g, const double* d_dipoles, int64_t n_dip, const int32_t* h_domains,
                       int order, int levels, double tol, double* d_out, int64_t ld, int ordering, void* stream);

/* Adjoint of omh_dip_source_mat (no reference counterpart; the gradient
 * path of the forward model): with the non-adaptive rule of `order`,
 * d_grad[6s + 0..2] = d/d(x y z) and d_grad[6s + 3..5] = d/d(qx qy qz) of
 * W(:,s) . rhs_s, where rhs_s is the DipSourceMat column of dipole s and W
 * an n x n_dip device matrix (leading dimension ldw) in `ordering`. */
int omh_dip_source_adjoint(omh_geometry* g, const double* d_dipoles, int64_t n_dip, const int32_t* h_domains,
                           int order, const double* d_W, int64_t ldw, int ordering, double* d_grad, void* stream);

/* Conductivity of a domain (Domain::conductivity, domain.h:77-79) and its
 * change with the pair coefficients recomputed (a non-conductive domain
 * cannot change).  Pair i of the communicating mesh pairs
 * (geometry.cpp:405-420): meshes[2], relative orientation, coeffs[3] =
 * (S, N, D) coefficients of assembleHeadMat.cpp:92-97, domain_mask = bit d
 * set for each common domain d (at most 31 domains).  Mesh m (info[8]):
 * info[0] first triangle unknown, info[1] triangles, info[2..3] vertex
 * unknown range (-1 if not contiguous), info[4] vertices, info[5] flags
 * (1 current barrier, 2 isolated, 4 outermost), info[6] its part
 * (connected component), info[7] the unknown whose diagonal entry sets the
 * part's deflation constant (assembleHeadMat.cpp:28-60) if this mesh holds
 * it, else -1. */
int omh_geometry_domain_conductivity(const omh_geometry* g, int32_t domain, double* value);
int omh_geometry_set_conductivity(omh_geometry* g, int32_t domain, double value);
int omh_geometry_pair_count(const omh_geometry* g, int32_t* count);
int omh_geometry_pair_info(const omh_geometry* g, int32_t i, int32_t* meshes, int32_t* orientation, double* coeffs,
                           int32_t* domain_mask);
int omh_geometry_mesh_info(const omh_geometry* g, int32_t m, int64_t* info);

/* Head2EEGMat(geo, electrodes): assemble.h:35, assembleSensors.cpp:25-39,
 * danielsson.cpp:26-129.  d_electrodes n_el x 3 row-major.  Row e of the
 * sparse matrix has the three entries (h_cols[3e+k], h_vals[3e+k]); columns
 * are reference-order vertex unknowns. */
int omh_head2eeg_mat(omh_geometry* g, const double* d_electrodes, int64_t n_el, int64_t* h_cols, double* h_vals,
                     void* stream);

/* Head2MEGMat (assembleSensors.cpp:69-99) before the sensor weights: for
 * npts integration points (d_pts, d_dirs: npts x 3 row-major) the npts x n
 * matrix d_out (column-major, leading dimension ld) of MagFactor x the
 * meshes' conductivity jumps x the Ferguson operator (assembleFerguson.cpp:
 * 18-30, operators.h:40-59) along each point's direction; columns in the
 * unknown order `ordering`, triangle columns zero.  The caller sums points
 * into sensors with Sensors::getWeightsMatrix (sensors.h:191-196). */
int omh_head2meg_mat(omh_geometry* g, const double* d_pts, const double* d_dirs, int64_t npts, double* d_out,
                     int64_t ld, int32_t ordering, void* stream);

/* ------------------------------------------------------- surface sources */

/* Mesh(filename) of a distributed source surface (mesh.h, Mesh::load
 * mesh.cpp:255-265; vertices x y z, 0-based triangles, both row-major).
 * Vertex i has index i (Mesh::generate_indices, mesh.cpp:266-272); vertices
 * are not de-duplicated.  The device tables are built on the first device
 * use and kept with the handle. */
int omh_mesh_create(const double* vertices, int64_t n_vertices, const int64_t* triangles, int64_t n_triangles,
                    omh_mesh** out);
/* Mesh(filename) for a .tri file (MeshIOs/tri.h:26-56). */
int omh_mesh_load(const char* path, omh_mesh** out);
void omh_mesh_free(omh_mesh* m);
/* out[0] = vertices, out[1] = triangles. */
int omh_mesh_sizes(const omh_mesh* m, int64_t out[2]);

/* Geometry::check(mesh) (geometry.cpp:329-343): *intersects = 1 when a
 * triangle of the mesh meets a triangle of the geometry (Mesh::intersection,
 * mesh.cpp:246-252), else 0.  Mesh::has_self_intersection is not part of
 * it: its test never reports a pair there (mesh.cpp:227-237). */
int omh_mesh_intersects(omh_geometry* g, omh_mesh* m, int32_t* intersects, void* stream);

/* SurfSourceMat(geo, sources, Integrator(order,0,.)) (assembleSourceMat.cpp:
 * 22-56): d_out (n x n_vertices of the mesh, column-major, leading dimension
 * ld >= n) with rows in the unknown order `ordering`.  Returns
 * OMH_STATUS_OVERLAPPING_SOURCE_MESH when the mesh meets the geometry.
 * flags: OMH_HEADMAT_FAR_FIELD as for omh_head_mat_ex.  The D block of a
 * current-barrier boundary mesh is left out (its triangles have no rows). */
int omh_surf_source_mat(omh_geometry* g, omh_mesh* m, int order, int flags, double* d_out, int64_t ld,
                        int32_t ordering, void* stream);

/* SurfSource2MEGMat (assembleSensors.cpp:102-124) before the sensor weights:
 * d_out (npts x n_vertices, column-major, leading dimension ld >= npts) =
 * the mesh's Ferguson operator (coefficient 1) at each point along its
 * direction (d_pts, d_dirs: npts x 3 row-major). */
int omh_surf_source2meg_mat(omh_mesh* m, const double* d_pts, const double* d_dirs, int64_t npts, double* d_out,
                            int64_t ld, void* stream);

/* ------------------------------------------------------ internal potentials */

/* Surf2VolMat(geo, points) (assembleHeadMat.cpp:325-363): the potential at
 * points inside the head from the unknowns.  Points (d_points, npts x 3
 * row-major) in a zero-conductivity domain are dropped; *n_kept rows are
 * written to d_out (column-major, leading dimension ld >= npts, n columns in
 * the unknown order `ordering`), and h_kept (npts entries, may be NULL)
 * receives the input index of each row. */
int omh_surf2vol_mat(omh_geometry* g, const double* d_points, int64_t npts, double* d_out, int64_t ld,
                     int32_t ordering, int64_t* h_kept, int64_t* n_kept, void* stream);

/* DipSource2InternalPotMat(geo, dipoles, points, domain) (assembleSourceMat.
 * cpp:154-186): kept points x n_dip (as omh_surf2vol_mat), K / sigma x the
 * dipole potential at the points of the dipole's domain (h_domains, or NULL
 * to locate each dipole), zero elsewhere. */
int omh_dip_source2internal_pot_mat(omh_geometry* g, const double* d_dipoles, int64_t n_dip, const int32_t* h_domains,
                                    const double* d_points, int64_t npts, double* d_out, int64_t ld, int64_t* h_kept,
                                    int64_t* n_kept, void* stream);

/* EITSourceMat(geo, electrodes, Integrator(order,0,.)) (assembleSourceMat.
 * cpp:109-152): d_out (n x n_el, column-major, leading dimension ld >= n,
 * rows in the unknown order `ordering`) for EIT electrodes at d_positions
 * (n_el x 3 row-major) with radii h_radii (host; NULL = all zero, the
 * injected currents are intensities).  Injection triangles as
 * Sensors::findInjectionTriangles (sensors.cpp:167-220). */
int omh_eit_source_mat(omh_geometry* g, const double* d_positions, const double* h_radii, int64_t n_el, int order,
                       double* d_out, int64_t ld, int32_t ordering, void* stream);

/* DipSource2MEGMat (assembleSensors.cpp:133-156) before the sensor weights:
 * d_out (npts x n_dip, column-major) = MagFactor (q x (x - r)) . d /
 * (|x - r|^3 |d|) for dipoles d_dipoles (n_dip x 6 row-major). */
int omh_dip_source2meg_mat(const double* d_dipoles, int64_t n_dip, const double* d_pts, const double* d_dirs,
                           int64_t npts, double* d_out, int64_t ld, void* stream);

/* ------------------------------------------------------- dense linear algebra */

/* Workspace bytes for omh_sym_factor_* / omh_sym_solve_* (element size 8 or 4). */
int64_t omh_factor_workspace_bytes(int64_t n, int32_t elem_size);

/* Factorisation behind SymMatrix::invert/inverse/solveLin
 * (OpenMEEGMaths/include/symmatrix.h:259-297, src/symmatrix.cpp:133-152;
 * DSPTRF there): pivot-free signed Cholesky H = L S L^T of the lower triangle
 * of d_A in place (matrix in OMH_ORDER_TRIANGLES_FIRST).  *info = 0, or
 * 1 + index of the first vanishing pivot (then the factor is unusable). */
int omh_sym_factor_f64(double* d_A, int64_t n, int64_t ld, void* d_work, int32_t* info, void* stream);
int omh_sym_factor_f32(float* d_A, int64_t n, int64_t ld, void* d_work, int32_t* info, void* stream);
/* The FP32 factorisation with its trailing updates as split-FP16 MFMA
 * products: each panel row is scaled by a power of two and split into two
 * FP16 parts, three FP16 products accumulate in FP32 (about 22 significant
 * bits per operand).  Same factor layout and workspace as
 * omh_sym_factor_f32; for the refined (mixed-precision) solve.  No
 * reference counterpart. */
int omh_sym_factor_f32_split(float* d_A, int64_t n, int64_t ld, void* d_work, int32_t* info, void* stream);

/* The same factorisation told the block structure of the matrix, so that
 * the trailing updates skip tiles that only multiply structural zeros (no
 * counterpart in the reference, which factors the packed matrix densely,
 * symmatrix.h:259-297; the result is the same).  The unknowns are cut into
 * nseg consecutive segments starting at seg_start[0] = 0 < seg_start[1] <
 * ... < n; hfill[s * nseg + t] (s > t; the other entries are not read) is 0
 * when the block of d_A with the rows of segment s and the columns of
 * segment t is zero.  L's pattern is its symbolic factorisation.  The
 * blocks declared zero are checked on the device (invalid argument when
 * one holds a non-zero).  HeadMat's pattern comes from its mesh pairs
 * (api.py Geometry.factor_segments). */
int omh_sym_factor_f32_split_blocks(float* d_A, int64_t n, int64_t ld, void* d_work, int32_t nseg,
                                    const int64_t* seg_start, const uint8_t* hfill, int32_t* info, void* stream);
/* The same for omh_sym_factor_f32 / omh_sym_factor_f64 (the trailing update
 * of each block runs as products over the pairs of row intervals its panel
 * can reach). */
int omh_sym_factor_f32_blocks(float* d_A, int64_t n, int64_t ld, void* d_work, int32_t nseg, const int64_t* seg_start,
                              const uint8_t* hfill, int32_t* info, void* stream);
int omh_sym_factor_f64_blocks(double* d_A, int64_t n, int64_t ld, void* d_work, int32_t nseg,
                              const int64_t* seg_start, const uint8_t* hfill, int32_t* info, void* stream);

/* Pivot signs of a finished omh_sym_factor_* factorisation (workspace
 * d_work, element size 4 or 8) against the pattern +1 for the first
 * n_positive unknowns and -1 for the rest (the triangle, then the vertex
 * unknowns of a HeadMat in OMH_ORDER_TRIANGLES_FIRST): *first = the first
 * unknown whose sign differs, or -1.  In that pattern the matrix is
 * quasi-definite and the pivot-free factor is stable; a mismatch means the
 * order needs pivoting (omh_lu_factor_f64).  No reference counterpart. */
int omh_factor_sign_mismatch(const void* d_work, int64_t n, int32_t elem_size, int64_t n_positive, int64_t* first,
                             void* stream);
/* The same against one expected sign per segment (seg_sign[s] = +1 or -1
 * for the unknowns of segment s; segments as for the *_blocks entry points,
 * in the factor's own order).  No reference counterpart. */
int omh_factor_sign_mismatch_segments(const void* d_work, int64_t n, int32_t elem_size, int32_t nseg,
                                      const int64_t* seg_start, const int8_t* seg_sign, int64_t* first, void* stream);

/* Factor order.  The factorisations may run on a symmetric reordering of the
 * matrix that moves whole segments (seg_start as for the *_blocks entry
 * points; the reordered matrix takes the segments in the order
 * seg_order[0..nseg-1], a permutation of 0..nseg-1).  A HeadMat taken mesh
 * by mesh (each mesh's triangle unknowns, then its vertex unknowns, the
 * innermost mesh first) is block tridiagonal over the meshes, and its
 * pivot-free factor fills no block outside that band: a quarter fewer
 * trailing-update flops than the triangles-first order at 4 layers.  None of
 * these has a reference counterpart (the reference factors the packed
 * matrix in its own order, symmatrix.cpp:133-152).
 *
 * omh_sym_reorder: the lower triangle of the reordered matrix, d_dst(i, j) =
 * d_src(p(i), p(j)) for i >= j, from the full symmetric FP64 d_src (both
 * triangles stored), as FP64 (dst_elem_size 8) or as the FP32 copy of the
 * mixed factorisations (4).  d_dst must not overlap d_src. */
int omh_sym_reorder(const double* d_src, int64_t lds, void* d_dst, int64_t ldd, int32_t dst_elem_size, int64_t n,
                    int32_t nseg, const int64_t* seg_start, const int32_t* seg_order, void* stream);
/* Rows of an n x ncol FP64 matrix into the factor order (inverse 0:
 * d_dst(i, :) = d_src(p(i), :)) or back (inverse 1: d_dst(p(i), :) =
 * d_src(i, :)), for the solves of a factor of the reordered matrix. */
int omh_reorder_rows_f64(const double* d_src, int64_t lds, double* d_dst, int64_t ldd, int64_t n, int64_t ncol,
                         int32_t nseg, const int64_t* seg_start, const int32_t* seg_order, int32_t inverse,
                         void* stream);

/* Pivoted fallback of SymMatrix::invert/inverse/solveLin (DSPTRF/DSPTRS,
 * symmatrix.cpp:133-152, symmatrix.h:259-297): FP64 LU with partial row
 * pivoting, P A = L U in place (any unknown order; for matrices on which the
 * pivot-free factorisation breaks down or shows a sign mismatch).  d_ipiv:
 * n device int32 (row j exchanged with row d_ipiv[j] at step j, 0-based);
 * *info = 0, or 1 + the first column with an exactly zero pivot.
 * omh_lu_solve_f64: d_B <- A^-1 d_B for nrhs columns.  Workspace bytes from
 * omh_lu_workspace_bytes (-1 for n <= 0). */
int64_t omh_lu_workspace_bytes(int64_t n);
int omh_lu_factor_f64(double* d_A, int64_t n, int64_t ld, int32_t* d_ipiv, void* d_work, int32_t* info,
                      void* stream);
int omh_lu_solve_f64(const double* d_LU, int64_t n, int64_t ld, const int32_t* d_ipiv, const void* d_work,
                     double* d_B, int64_t ldb, int64_t nrhs, void* stream);

/* d_B <- H^-1 d_B for nrhs columns with a stored factor (DSPTRS,
 * symmatrix.cpp:133-152; linsolve gain.h:52-57). */
int omh_sym_solve_f64(const double* d_L, int64_t n, int64_t ld, const void* d_work, double* d_B, int64_t ldb,
                      int64_t nrhs, void* stream);
int omh_sym_solve_f32(const float* d_L, int64_t n, int64_t ld, const void* d_work, float* d_B, int64_t ldb,
                      int64_t nrhs, void* stream);
/* The same with the block structure the factor was computed with (the
 * arguments of the *_blocks factorisations): the updates below and above
 * each group of diagonal blocks run only over the rows that L's non-zero
 * blocks reach.  Not checked against d_L: pass the structure the factor
 * was made with.  No reference counterpart. */
int omh_sym_solve_f64_blocks(const double* d_L, int64_t n, int64_t ld, const void* d_work, double* d_B, int64_t ldb,
                             int64_t nrhs, int32_t nseg, const int64_t* seg_start, const uint8_t* hfill, void* stream);
int omh_sym_solve_f32_blocks(const float* d_L, int64_t n, int64_t ld, const void* d_work, float* d_B, int64_t ldb,
                             int64_t nrhs, int32_t nseg, const int64_t* seg_start, const uint8_t* hfill, void* stream);

/* Stopping rule of the refinement in omh_mixed_solve* after correction
 * `it` (0-based) of relative size rel = ||dX||_F / ||X||_F, the one before
 * being prev (no reference counterpart).  Returns 1 (converged) when
 * rel <= tol, or when the corrections contract (rel < prev) and the
 * estimated error left, rel x rel / prev, is <= tol, or when they no longer
 * contract (rel > prev / 2, from it = 2 on) at the FP64 rounding floor
 * rel <= min(1e3 tol, 1e-10); 2 (stalled: no longer contracting above that
 * floor, or NaN); 0 (continue). */
int32_t omh_refine_state(int32_t it, double rel, double prev, double tol);

/* Mixed-precision solve X = H^-1 B: FP32 factor (d_L32, d_work32) with FP64
 * residuals on the FP64 matrix d_H and iterative refinement, stopped by
 * omh_refine_state or after max_iter corrections.  Returns
 * OMH_STATUS_NOT_CONVERGED (X, *iterations and *rel_update still written)
 * unless omh_refine_state reached "converged".  No reference counterpart:
 * it stands beside DSPTRS for the mixed-precision path. */
int omh_mixed_solve(const double* d_H, int64_t n, int64_t ldh, const float* d_L32, int64_t ld32,
                    const void* d_work32, const double* d_B, int64_t ldb, double* d_X, int64_t ldx, int64_t nrhs,
                    int32_t max_iter, double tol, int32_t* iterations, double* rel_update, void* stream);
/* omh_mixed_solve without any device allocation (no reference counterpart;
 * for callers that cannot allocate: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The refinement buffers
 * live in a fixed device array of the library (66 MiB + 8 KiB per GPU, held
 * until the process ends) instead of its growable scratch: n x nrhs must
 * not exceed 4194304 (2^22), else invalid argument.  Same arguments,
 * results and status as omh_mixed_solve.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU race. */
int omh_mixed_solve_noalloc(const double* d_H, int64_t n, int64_t ldh, const float* d_L32, int64_t ld32,
                            const void* d_work32, const double* d_B, int64_t ldb, double* d_X, int64_t ldx,
                            int64_t nrhs, int32_t max_iter, double tol, int32_t* iterations, double* rel_update,
                            void* stream);

/* The same with the block structure of d_H (as for the *_blocks
 * factorisations, and the one the factor d_L32 was made with): the FP64
 * residual skips d_H's zero blocks, the FP32 solves L's.  Not checked
 * against d_H. */
int omh_mixed_solve_blocks(const double* d_H, int64_t n, int64_t ldh, const float* d_L32, int64_t ld32,
                           const void* d_work32, const double* d_B, int64_t ldb, double* d_X, int64_t ldx,
                           int64_t nrhs, int32_t max_iter, double tol, int32_t nseg, const int64_t* seg_start,
                           const uint8_t* hfill, int32_t* iterations, double* rel_update, void* stream);
/* omh_mixed_solve_blocks with a factor d_L32 of the reordered matrix (the
 * FP32 copy from omh_sym_reorder, factored with the reordered structure):
 * seg_start / hfill describe d_H in its own order, seg_order the factor's
 * segment order.  Residuals in d_H's order, each correction solved in the
 * factor's (the residual's rows gathered into it, the correction's rows
 * scattered back); d_B and d_X in d_H's order. */
int omh_mixed_solve_reordered(const double* d_H, int64_t n, int64_t ldh, const float* d_L32, int64_t ld32,
                              const void* d_work32, const double* d_B, int64_t ldb, double* d_X, int64_t ldx,
                              int64_t nrhs, int32_t max_iter, double tol, int32_t nseg, const int64_t* seg_start,
                              const uint8_t* hfill, const int32_t* seg_order, int32_t* iterations,
                              double* rel_update, void* stream);

/* C = alpha op(A) op(B) + beta C: the DGEMM products of
 * OpenMEEGMaths/include/matrix.h:333-350 and gain.h:69-89. */
int omh_gemm_f64(int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k, double alpha, const double* d_A,
                 int64_t lda, const double* d_B, int64_t ldb, double beta, double* d_C, int64_t ldc, void* stream);
int omh_gemm_f32(int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k, float alpha, const float* d_A,
                 int64_t lda, const float* d_B, int64_t ldb, float beta, float* d_C, int64_t ldc, void* stream);

/* FP64 -> FP32 copy (lower_only: only entries with row >= col). */
int omh_convert_f64_f32(const double* d_src, int64_t lds, float* d_dst, int64_t ldd, int64_t m, int64_t n,
                        int32_t lower_only, void* stream);

/* SparseMatrix * SymMatrix of sparse_matrix.cpp:36-51 for a Head2EEGMat
 * (3 non-zeros per row): out(e,:) = sum_k vals(e,k) Hm(cols(e,k),:).
 * d_cols/d_vals are device copies of omh_head2eeg_mat's output with the
 * columns already mapped to the row order of d_Hm. */
int omh_sparse_rows(const int64_t* d_cols, const double* d_vals, int64_t ne, const double* d_Hm, int64_t ldh,
                    int64_t n, double* d_out, int64_t ldo, void* stream);

/* The right-hand sides of the adjoint solve, linsolve(HeadMat,
 * Head2EEGMat) (gain.h:28-49, used at gain.h:84): d_out (n x ne, leading
 * dimension ldo) = the dense transpose of the sparse ne x n matrix whose
 * row e holds d_vals(e,k) at column d_rows(e,k), k < 3 (device arrays of
 * 3 ne entries, rows already in the row order of d_out; repeated rows add,
 * rows outside [0, n) are ignored). */
int omh_sparse_transpose_dense(const int64_t* d_rows, const double* d_vals, int64_t ne, int64_t n, double* d_out,
                               int64_t ldo, void* stream);

/* Unknown-order conversions between the reference's storage and device
 * matrices.  omh_sym_unpack: packed upper SymMatrix data (symmatrix.h:52-62,
 * reference order, device pointer) -> full symmetric matrix in `ordering`.
 * omh_sym_pack: the inverse.  omh_permute_rows: rows of an n x ncol matrix
 * from reference order to `ordering` (to_reference = 0) or back (1). */
int omh_sym_unpack(const double* d_packed, int64_t n, int64_t nv, double* d_H, int64_t ld, int ordering, void* stream);
int omh_sym_pack(const double* d_H, int64_t n, int64_t nv, int64_t ld, int ordering, double* d_packed, void* stream);
int omh_permute_rows(const double* d_src, int64_t lds, double* d_dst, int64_t ldd, int64_t n, int64_t nv,
                     int64_t ncol, int ordering, int32_t to_reference, void* stream);

/* Kernel timing for benchmarks (no reference counterpart): when enabled,
 * HIP event pairs are recorded on the launch stream around the named
 * kernels: "factor_update_f32_t128", "factor_update_f32_t64",
 * "factor_update_f64_t64" (interior launches of the trailing-update GEMM
 * with 128- or 64-wide tiles), "diag_factor_f64", "diag_factor_f32",
 * "s_block", "d_pairs".  omh_profile_enable clears the
 * records; omh_profile_read returns the launches, the summed seconds and the
 * summed algorithmic work (flop for the factor kernels, Gauss-point
 * evaluations for the assembly kernels) of one name. */
int omh_profile_enable(int32_t on);
int omh_profile_read(const char* name, int64_t* calls, double* seconds, double* work);

/* Batched EEG gains (no reference counterpart; the reference computes one
 * head per GainEEGadjoint call, gain.h:78-89).  A batch holds n_heads heads,
 * head h being geometry geoms[h] (kept alive by the caller) with n_dipoles[h]
 * dipoles (h_dipoles[h]: n x 6 host array, x y z qx qy qz, each located in
 * its domain) and n_electrodes[h] electrodes (h_electrodes[h]: n x 3).
 * precision: 0 FP64 factor and solve, 1 FP32 factor with FP64 refinement.
 * A pass assembles HeadMat and DipSourceMat (Integrator(3,10,0.001)) of
 * every head and computes G = (H^-1 E^T)^T DSM with the dense algebra of
 * all heads grouped in lock step.
 *
 * omh_batch_run: one pass after the work on `stream`.  The first pass (and
 * every pass with use_graph = 0) runs eagerly and fixes the number of
 * refinement steps; with use_graph = 1 later passes replay one captured HIP
 * graph of the whole pass (*used_graph tells which ran).  Returns
 * OMH_STATUS_NOT_CONVERGED when the last eager pass's refinement did not
 * converge (omh_refine_state) for every head.
 * omh_batch_tune: times the candidates of the grouped trailing update on the
 * batch's own shapes and keeps the fastest: FP32 products on 64- or 128-wide
 * tiles with grid caps, and (precision 1) the split-FP16 update (the panel
 * split into FP16 parts, three products per 128 x 128 tile, FP32 sums).
 * n_results x {tile, grid_cap, split16 (0/1), seconds} are written to
 * h_results if not NULL (at most max_results); the next pass runs eagerly
 * again.
 * omh_batch_gain: device pointer to head h's gain (n_el x n_dip,
 * column-major), valid until omh_batch_free.
 * omh_batch_stats: factor status (0 or 1 + first vanishing pivot) and the
 * last relative refinement update of head h.
 * omh_batch_config: refinement steps, trailing tile and grid cap, grouped
 * launches per pass, nodes of the captured graph (0 before capture) and
 * whether the trailing update is the split-FP16 one. */
typedef struct omh_batch omh_batch;
int omh_batch_create(int32_t n_heads, omh_geometry* const* geoms, const double* const* h_dipoles,
                     const int64_t* n_dipoles, const double* const* h_electrodes, const int64_t* n_electrodes,
                     int32_t precision, omh_batch** out);
int omh_batch_run(omh_batch* b, int32_t use_graph, int32_t* used_graph, void* stream);
int omh_batch_tune(omh_batch* b, double* h_results, int32_t max_results, int32_t* n_results, void* stream);
int omh_batch_gain(const omh_batch* b, int32_t head, const double** d_gain, int64_t* n_el, int64_t* n_dip);
int omh_batch_stats(omh_batch* b, int32_t head, int32_t* factor_info, double* rel_update, void* stream);
int omh_batch_config(const omh_batch* b, int32_t* refine_steps, int32_t* tile, int64_t* grid_cap,
                     int64_t* launches, int64_t* graph_nodes, int32_t* split16);
int omh_batch_free(omh_batch* b);

/* Frees the library's cached device scratch buffers. */
int omh_release_scratch(void);

/* Writes the raw outputs of one FP64 16x16x4, one FP32 16x16x4 and one
 * FP16 16x16x32 MFMA block on known operands into d_out (3 x 256 doubles:
 * lane l, element r at 4 l + r) for the lane-layout self-test. */
int omh_probe_mfma_layout(double* d_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OPENMEEG_AMD_H */
