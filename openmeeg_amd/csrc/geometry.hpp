// Host-side geometry model of the symmetric BEM: meshes, interfaces, domains,
// conductivities, current barriers, unknown numbering and the communicating
// mesh pairs with their conductivity coefficients.
//
// Reference behaviour (file:line in OpenMEEG 2.5.16):
//   Geometry::import                geometry.cpp:197-229, add_vertex geometry.h:100-109
//   .geom / .cond / .tri readers    GeometryIOs/GeomFile.h:158-300, MeshIOs/tri.h:26-56
//   interface orientation check     interface.cpp:40-95
//   finalize                        geometry.h:228-249
//   mark_current_barriers           geometry.cpp:424-536
//   generate_indices                geometry.cpp:256-295
//   make_mesh_pairs                 geometry.cpp:405-420
//   sigma / sigma_inv / indicator   geometry.h:182-184, 289-310
//   relative_orientation            geometry.h:201-206
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace omh {

using i64 = int64_t;

struct Mesh {
    std::string name;
    std::vector<i64> verts;                  // global vertex ids, mesh order
    std::vector<std::array<i64, 3>> tris;    // global vertex ids
    bool current_barrier = false;
    bool isolated = false;
    bool outermost = false;
    i64 tri_first_index = -1;                // unknown index of the first triangle
};

struct Interface {
    std::string name;
    std::vector<std::pair<int, int>> omeshes;   // (mesh id, +1 / -1)
};

struct Domain {
    std::string name;
    std::vector<std::pair<Interface, bool>> boundaries;   // (interface, inside)
    double cond = -1.0;
    int mesh_orientation(int m) const;
};

struct MeshPair {
    int m1, m2, orientation;
    double cS, cN, cD;
};

struct DeviceGeometry;   // device tables, defined in device_geometry.hpp

struct Geometry {
    std::vector<std::array<double, 3>> vertices;
    std::map<std::array<double, 3>, i64> vmap;
    std::vector<Mesh> meshes;
    std::vector<Domain> domains;
    std::vector<std::vector<int>> parts;
    std::vector<MeshPair> pairs;
    std::vector<i64> vert_index;             // unknown index per vertex (-1 invalid)
    std::vector<char> invalid;
    int outer_domain = -1;
    bool nested = false;
    i64 nb_parameters = 0;
    i64 nb_current_barrier_triangles = 0;
    i64 n_vert_unknowns = 0;                 // vertex unknowns (first in reference order)
    i64 n_tri_unknowns = 0;                  // non-barrier triangle unknowns

    std::shared_ptr<DeviceGeometry> dev;     // built on first device use

    i64 size() const { return nb_parameters - nb_current_barrier_triangles; }

    // construction
    i64 add_vertex(const std::array<double, 3>& v);
    void import_meshes(const std::vector<std::string>& names,
                       const std::vector<std::vector<std::array<double, 3>>>& pts,
                       const std::vector<std::vector<std::array<i64, 3>>>& tris);
    int mesh_index(const std::string& name) const;
    int domain_index(const std::string& name) const;
    void finalize();

    // Changes the conductivity of a domain and recomputes the pair
    // coefficients.  A zero conductivity changes the structure (current
    // barriers, unknowns) and is refused here.
    void set_conductivity(int domain, double value);

    // queries
    std::vector<int> domains_of(int m) const;
    std::vector<int> common_domains(int m1, int m2) const;
    double sigma(int m1, int m2) const;
    double sigma_inv(int m1, int m2) const;
    double indicator(int m1, int m2) const;
    int relative_orientation(int m1, int m2) const;
    double mesh_solid_angle(int m, const double p[3]) const;
    double interface_solid_angle(const Interface& itf, const double p[3]) const;
    bool interface_contains_sa(double solid_angle) const;
    void fix_interface_orientation(Interface& itf) const;
    // Geometry::domain(point) (geometry.cpp:146-154) on the host: the first
    // domain whose boundaries all hold the point as they should; -1 if none.
    int domain_of_point(const double p[3]) const;

private:
    void mark_current_barriers();
    void check_nested();
    void generate_indices();

public:
    void make_pairs();
};

// Geometry(geom_file, cond_file), GeomFile.h:158-300 (format 1.1).
std::unique_ptr<Geometry> load_geometry(const std::string& geom_file, const std::string& cond_file);

// A standalone mesh (a distributed source surface) held in a one-mesh
// Geometry so that the device tables can be built for it: vertices in the
// given order (no de-duplication), vertex i has index i and the triangles
// follow (Mesh::generate_indices, mesh.cpp:266-272); the mesh is marked
// outermost and a current barrier, as SurfSourceMat does
// (assembleSourceMat.cpp:33-36).  No domains, no pairs.
std::unique_ptr<Geometry> make_source_mesh(const std::vector<std::array<double, 3>>& pts,
                                           const std::vector<std::array<i64, 3>>& tris);

// Mesh(filename) of a .tri file (MeshIOs/tri.h:26-56) as a source mesh.
std::unique_ptr<Geometry> load_source_mesh(const std::string& path);

}  // namespace omh
