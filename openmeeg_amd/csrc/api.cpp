// extern "C" boundary of the library (declared in include/openmeeg_amd.h).
// Every entry point catches C++ exceptions and turns them into a status code
// plus a per-thread message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "../../include/openmeeg_amd.h"
#include "assembly.hpp"
#include "batch.hpp"
#include "common.hpp"
#include "geometry.hpp"
#include "layout.hpp"
#include "linalg.hpp"
#include "lu.hpp"
#include "noalloc.hpp"
#include "sources.hpp"

struct omh_geometry {
    std::unique_ptr<omh::Geometry> g;
};

struct omh_mesh {
    std::unique_ptr<omh::Geometry> g;   // one-mesh geometry (make_source_mesh)
};

struct omh_batch {
    std::unique_ptr<omh::BatchPlan> plan;
};

namespace {

thread_local std::string g_err = "no error";

template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const omh::OverlappingSourceMesh& e) {
        g_err = e.what();
        return OMH_STATUS_OVERLAPPING_SOURCE_MESH;
    } catch (const omh::NotConverged& e) {
        g_err = e.what();
        return OMH_STATUS_NOT_CONVERGED;
    } catch (const std::invalid_argument& e) {
        g_err = std::string("invalid argument: ") + e.what();
        return 2;
    } catch (const omh::HipError& e) {
        g_err = std::string("HIP error: ") + e.what();
        return 3;
    } catch (const std::bad_alloc&) {
        g_err = "out of host memory";
        return 4;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    } catch (...) {
        g_err = "unknown error";
        return 1;
    }
}

hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

omh::Layout layout_of(const omh::Geometry& g, int64_t ld, int ordering) {
    if (ordering != OMH_ORDER_REFERENCE && ordering != OMH_ORDER_TRIANGLES_FIRST)
        throw std::invalid_argument("unknown ordering");
    if (ld < g.size()) throw std::invalid_argument("leading dimension smaller than the number of unknowns");
    omh::Layout L;
    L.ld = ld;
    L.nv = g.n_vert_unknowns;
    L.nt = g.n_tri_unknowns;
    L.tri_first = ordering == OMH_ORDER_TRIANGLES_FIRST ? 1 : 0;
    return L;
}

[[noreturn]] void not_converged(int it, double rel, double tol) {
    char buf[200];
    std::snprintf(buf, sizeof buf,
                  "iterative refinement did not converge: relative update %.3e after %d corrections (tol %.3e); "
                  "the FP32 factor is too coarse for this matrix",
                  rel, it, tol);
    throw omh::NotConverged(buf);
}

void need(const void* p, const char* what) {
    if (!p) throw std::invalid_argument(std::string(what) + " is NULL");
}

std::vector<std::vector<std::array<double, 3>>> gather_pts(int n, const int64_t* nv, const double* const* v) {
    std::vector<std::vector<std::array<double, 3>>> pts(n);
    for (int m = 0; m < n; ++m) {
        if (nv[m] <= 0) throw std::invalid_argument("mesh without vertices");
        need(v[m], "vertices");
        pts[m].resize(nv[m]);
        for (int64_t i = 0; i < nv[m]; ++i) pts[m][i] = {v[m][3 * i], v[m][3 * i + 1], v[m][3 * i + 2]};
    }
    return pts;
}

std::vector<std::vector<std::array<int64_t, 3>>> gather_tris(int n, const int64_t* nt, const int64_t* const* t) {
    std::vector<std::vector<std::array<int64_t, 3>>> tris(n);
    for (int m = 0; m < n; ++m) {
        if (nt[m] <= 0) throw std::invalid_argument("mesh without triangles");
        need(t[m], "triangles");
        tris[m].resize(nt[m]);
        for (int64_t i = 0; i < nt[m]; ++i) tris[m][i] = {t[m][3 * i], t[m][3 * i + 1], t[m][3 * i + 2]};
    }
    return tris;
}

}  // namespace

extern "C" {

const char* omh_last_error(void) { return g_err.c_str(); }

int omh_geometry_load(const char* geom_file, const char* cond_file, omh_geometry** out) {
    return guard([&] {
        need(geom_file, "geometry file");
        need(cond_file, "conductivity file");
        need(out, "out");
        auto h = std::make_unique<omh_geometry>();
        h->g = omh::load_geometry(geom_file, cond_file);
        *out = h.release();
    });
}

int omh_geometry_nested(int n_layers, const int64_t* n_vertices, const double* const* vertices,
                        const int64_t* n_triangles, const int64_t* const* triangles, const double* conductivities,
                        omh_geometry** out) {
    return guard([&] {
        if (n_layers < 1) throw std::invalid_argument("at least one layer is needed");
        need(n_vertices, "n_vertices");
        need(vertices, "vertices");
        need(n_triangles, "n_triangles");
        need(triangles, "triangles");
        need(conductivities, "conductivities");
        need(out, "out");
        const int L = n_layers;
        std::vector<std::string> names;
        if (L == 3) names = {"Cortex", "Skull", "Head"};
        else if (L == 1) names = {"Cortex"};
        else
            for (int k = 0; k < L; ++k) names.push_back("Layer" + std::to_string(k));
        auto g = std::make_unique<omh::Geometry>();
        g->import_meshes(names, gather_pts(L, n_vertices, vertices), gather_tris(L, n_triangles, triangles));
        std::vector<omh::Interface> itf;
        for (int k = 0; k < L; ++k) itf.push_back(omh::Interface{names[k], {{k, 1}}});
        auto dom = [&](const std::string& name, double c, std::vector<std::pair<int, bool>> b) {
            omh::Domain d;
            d.name = name;
            d.cond = c;
            for (auto& x : b) d.boundaries.push_back({itf[x.first], x.second});
            g->domains.push_back(d);
        };
        // _make_geometry.py:167-204
        if (L == 3) {
            dom("Scalp", conductivities[2], {{1, false}, {2, true}});
            dom("Brain", conductivities[0], {{0, true}});
            dom("Air", 0.0, {{2, false}});
            dom("Skull", conductivities[1], {{0, false}, {1, true}});
        } else if (L == 1) {
            dom("Brain", conductivities[0], {{0, true}});
            dom("Air", 0.0, {{0, false}});
        } else {
            dom("Brain", conductivities[0], {{0, true}});
            for (int k = 1; k < L; ++k) dom("Layer" + std::to_string(k), conductivities[k], {{k - 1, false}, {k, true}});
            dom("Air", 0.0, {{L - 1, false}});
        }
        g->finalize();
        auto h = std::make_unique<omh_geometry>();
        h->g = std::move(g);
        *out = h.release();
    });
}

int omh_geometry_create(int n_meshes, const char* const* mesh_names, const int64_t* n_vertices,
                        const double* const* vertices, const int64_t* n_triangles, const int64_t* const* triangles,
                        int n_interfaces, const char* const* interface_names, const int64_t* interface_offsets,
                        const int32_t* interface_meshes, const int32_t* interface_orientations, int n_domains,
                        const char* const* domain_names, const double* domain_conductivities,
                        const int64_t* domain_offsets, const int32_t* domain_interfaces, const int32_t* domain_inside,
                        omh_geometry** out) {
    return guard([&] {
        if (n_meshes < 1 || n_interfaces < 1 || n_domains < 1)
            throw std::invalid_argument("meshes, interfaces and domains must be non-empty");
        need(out, "out");
        std::vector<std::string> names;
        for (int m = 0; m < n_meshes; ++m) names.push_back(mesh_names[m] ? mesh_names[m] : std::to_string(m + 1));
        auto g = std::make_unique<omh::Geometry>();
        g->import_meshes(names, gather_pts(n_meshes, n_vertices, vertices),
                         gather_tris(n_meshes, n_triangles, triangles));
        std::vector<omh::Interface> itf(n_interfaces);
        for (int i = 0; i < n_interfaces; ++i) {
            itf[i].name = interface_names && interface_names[i] ? interface_names[i] : std::to_string(i + 1);
            for (int64_t p = interface_offsets[i]; p < interface_offsets[i + 1]; ++p) {
                const int m = interface_meshes[p], o = interface_orientations[p];
                if (m < 0 || m >= n_meshes) throw std::invalid_argument("interface mesh index out of range");
                if (o != 1 && o != -1) throw std::invalid_argument("interface orientation must be +1 or -1");
                itf[i].omeshes.push_back({m, o});
            }
            if (itf[i].omeshes.empty()) throw std::invalid_argument("empty interface " + itf[i].name);
        }
        for (int d = 0; d < n_domains; ++d) {
            omh::Domain dm;
            dm.name = domain_names && domain_names[d] ? domain_names[d] : std::to_string(d + 1);
            dm.cond = domain_conductivities[d];
            for (int64_t p = domain_offsets[d]; p < domain_offsets[d + 1]; ++p) {
                const int i = domain_interfaces[p];
                if (i < 0 || i >= n_interfaces) throw std::invalid_argument("domain interface index out of range");
                dm.boundaries.push_back({itf[i], domain_inside[p] != 0});
            }
            if (dm.boundaries.empty()) throw std::invalid_argument("domain " + dm.name + " has no boundary");
            g->domains.push_back(dm);
        }
        g->finalize();
        auto h = std::make_unique<omh_geometry>();
        h->g = std::move(g);
        *out = h.release();
    });
}

void omh_geometry_free(omh_geometry* g) { delete g; }

int omh_geometry_sizes(const omh_geometry* g, int64_t out[8]) {
    return guard([&] {
        need(g, "geometry");
        need(out, "out");
        const omh::Geometry& G = *g->g;
        out[0] = G.size();
        out[1] = G.nb_parameters;
        out[2] = G.nb_current_barrier_triangles;
        out[3] = G.n_vert_unknowns;
        out[4] = G.n_tri_unknowns;
        out[5] = (int64_t)G.meshes.size();
        out[6] = (int64_t)G.domains.size();
        out[7] = G.nested ? 1 : 0;
    });
}

int omh_geometry_domain_index(const omh_geometry* g, const char* name, int32_t* index) {
    return guard([&] {
        need(g, "geometry");
        need(name, "name");
        need(index, "index");
        *index = g->g->domain_index(name);
    });
}

int omh_geometry_domain_name(const omh_geometry* g, int32_t index, char* buf, int64_t buflen) {
    return guard([&] {
        need(g, "geometry");
        need(buf, "buf");
        const auto& d = g->g->domains.at(index);
        if ((int64_t)d.name.size() + 1 > buflen) throw std::invalid_argument("buffer too small");
        std::memcpy(buf, d.name.c_str(), d.name.size() + 1);
    });
}

int omh_geometry_set_conductivity(omh_geometry* g, int32_t domain, double value) {
    return guard([&] {
        need(g, "geometry");
        g->g->set_conductivity(domain, value);
    });
}

int omh_geometry_pair_count(const omh_geometry* g, int32_t* count) {
    return guard([&] {
        need(g, "geometry");
        need(count, "count");
        *count = (int32_t)g->g->pairs.size();
    });
}

int omh_geometry_pair_info(const omh_geometry* g, int32_t i, int32_t* meshes, int32_t* orientation, double* coeffs,
                           int32_t* domain_mask) {
    return guard([&] {
        need(g, "geometry");
        const auto& G = *g->g;
        if (i < 0 || i >= (int)G.pairs.size()) throw std::invalid_argument("pair index out of range");
        const auto& p = G.pairs[i];
        if (meshes) { meshes[0] = p.m1; meshes[1] = p.m2; }
        if (orientation) *orientation = p.orientation;
        if (coeffs) { coeffs[0] = p.cS; coeffs[1] = p.cN; coeffs[2] = p.cD; }
        if (domain_mask) {
            int32_t mask = 0;
            for (int d : G.common_domains(p.m1, p.m2)) {
                if (d >= 31) throw std::invalid_argument("domain masks hold at most 31 domains");
                mask |= (1 << d);
            }
            *domain_mask = mask;
        }
    });
}

int omh_geometry_mesh_info(const omh_geometry* g, int32_t m, int64_t* info) {
    return guard([&] {
        need(g, "geometry");
        need(info, "info");
        const auto& G = *g->g;
        if (m < 0 || m >= (int)G.meshes.size()) throw std::invalid_argument("mesh index out of range");
        const auto& M = G.meshes[m];
        int64_t v0 = -1, v1 = -1;
        bool contiguous = true;
        for (size_t k = 0; k < M.verts.size(); ++k) {
            const int64_t u = G.vert_index[M.verts[k]];
            if (k == 0) v0 = u;
            else if (u != v0 + (int64_t)k) contiguous = false;
            v1 = u + 1;
        }
        info[0] = M.tri_first_index;
        info[1] = (int64_t)M.tris.size();
        info[2] = contiguous ? v0 : -1;
        info[3] = contiguous ? v1 : -1;
        info[4] = (int64_t)M.verts.size();
        info[5] = (M.current_barrier ? 1 : 0) | (M.isolated ? 2 : 0) | (M.outermost ? 4 : 0);
        // part and deflation pivot (assembleHeadMat.cpp:28-60: the first
        // vertex of the first outermost mesh of the part)
        info[6] = -1;
        info[7] = -1;
        for (size_t p = 0; p < G.parts.size(); ++p) {
            int64_t i_first = 0;
            int pivot = -1;
            for (int mid : G.parts[p]) {
                if (mid == m) info[6] = (int64_t)p;
                const auto& Q = G.meshes[mid];
                if (Q.outermost && i_first == 0 && !Q.verts.empty()) {
                    i_first = G.vert_index[Q.verts.front()];
                    pivot = mid;
                }
            }
            if (pivot == m) info[7] = i_first;
        }
    });
}

int omh_geometry_domain_conductivity(const omh_geometry* g, int32_t domain, double* value) {
    return guard([&] {
        need(g, "geometry");
        need(value, "value");
        if (domain < 0 || domain >= (int)g->g->domains.size()) throw std::invalid_argument("domain index out of range");
        *value = g->g->domains[domain].cond;
    });
}

int omh_head_mat(omh_geometry* g, int order, double* d_H, int64_t ld, int ordering, void* stream) {
    return guard([&] {
        need(g, "geometry");
        need(d_H, "d_H");
        omh::assemble_headmat(*g->g, order, d_H, layout_of(*g->g, ld, ordering), S(stream));
    });
}

int omh_head_mat_ex(omh_geometry* g, int order, int32_t flags, double* d_H, int64_t ld, int ordering, void* stream) {
    return guard([&] {
        need(g, "geometry");
        need(d_H, "d_H");
        if (flags & ~OMH_HEADMAT_FAR_FIELD) throw std::invalid_argument("unknown HeadMat flags");
        omh::assemble_headmat(*g->g, order, d_H, layout_of(*g->g, ld, ordering), S(stream),
                              (flags & OMH_HEADMAT_FAR_FIELD) != 0);
    });
}

int omh_head_mat_integrator(omh_geometry* g, int order, int levels, double tol, double* d_H, int64_t ld,
                            int ordering, void* stream) {
    return guard([&] {
        need(g, "geometry");
        need(d_H, "d_H");
        omh::assemble_headmat(*g->g, order, d_H, layout_of(*g->g, ld, ordering), S(stream), false, levels, tol);
    });
}

int omh_locate_domains(omh_geometry* g, const double* d_points, int64_t stride, int64_t n, int32_t* h_domains,
                       void* stream) {
    return guard([&] {
        need(g, "geometry");
        need(h_domains, "h_domains");
        if (n > 0) need(d_points, "d_points");
        const auto d = omh::locate_domains(*g->g, d_points, stride, n, S(stream));
        for (int64_t i = 0; i < n; ++i) h_domains[i] = d[i];
    });
}

int omh_dip_source_mat(omh_geometry* g, const double* d_dipoles, int64_t n_dip, const int32_t* h_domains, int order,
                       int levels, double tol, double* d_out, int64_t ld, int ordering, void* stream) {
    return guard([&] {
        need(g, "geometry");
        need(d_out, "d_out");
        if (n_dip < 0) throw std::invalid_argument("negative number of dipoles");
        if (n_dip > 0) need(d_dipoles, "d_dipoles");
        std::vector<int> dom;
        if (h_domains) {
            dom.assign(h_domains, h_domains + n_dip);
            for (int d : dom)
                if (d < 0 || d >= (int)g->g->domains.size()) throw std::invalid_argument("domain index out of range");
        } else {
            dom = omh::locate_domains(*g->g, d_dipoles, 6, n_dip, S(stream));
        }
        omh::dip_source_mat(*g->g, d_dipoles, n_dip, dom, order, levels, tol, d_out, ld,
                            layout_of(*g->g, ld, ordering), S(stream));
    });
}

int omh_dip_source_adjoint(omh_geometry* g, const double* d_dipoles, int64_t n_dip, const int32_t* h_domains, int order,
                           const double* d_W, int64_t ldw, int ordering, double* d_grad, void* stream) {
    return guard([&] {
        need(g, "geometry");
        need(d_W, "d_W");
        need(d_grad, "d_grad");
        if (n_dip < 0) throw std::invalid_argument("negative number of dipoles");
        if (n_dip > 0) need(d_dipoles, "d_dipoles");
        std::vector<int> dom;
        if (h_domains) {
            dom.assign(h_domains, h_domains + n_dip);
            for (int d : dom)
                if (d < 0 || d >= (int)g->g->domains.size()) throw std::invalid_argument("domain index out of range");
        } else {
            dom = omh::locate_domains(*g->g, d_dipoles, 6, n_dip, S(stream));
        }
        omh::dip_source_adjoint(*g->g, d_dipoles, n_dip, dom, order, d_W, ldw, layout_of(*g->g, ldw, ordering), d_grad,
                                S(stream));
    });
}

int omh_head2eeg_mat(omh_geometry* g, const double* d_electrodes, int64_t n_el, int64_t* h_cols, double* h_vals,
                     void* stream) {
    return guard([&] {
        need(g, "geometry");
        if (n_el > 0) {
            need(d_electrodes, "d_electrodes");
            need(h_cols, "h_cols");
            need(h_vals, "h_vals");
        }
        omh::head2eeg(*g->g, d_electrodes, n_el, h_cols, h_vals, S(stream));
    });
}

int omh_head2meg_mat(omh_geometry* g, const double* d_pts, const double* d_dirs, int64_t npts, double* d_out,
                     int64_t ld, int32_t ordering, void* stream) {
    return guard([&] {
        need(g, "geometry");
        need(d_out, "d_out");
        if (npts < 0 || ld < std::max<int64_t>(1, npts)) throw std::invalid_argument("bad sizes");
        if (npts > 0) {
            need(d_pts, "d_pts");
            need(d_dirs, "d_dirs");
        }
        // the unknown order maps columns; rows are the points (ld >= npts)
        omh::Layout L = layout_of(*g->g, g->g->size(), ordering);
        L.ld = ld;
        omh::head2meg(*g->g, d_pts, d_dirs, npts, d_out, ld, L, S(stream));
    });
}

int omh_mesh_create(const double* vertices, int64_t n_vertices, const int64_t* triangles, int64_t n_triangles,
                    omh_mesh** out) {
    return guard([&] {
        need(out, "out");
        if (n_vertices <= 0 || n_triangles <= 0) throw std::invalid_argument("empty mesh");
        need(vertices, "vertices");
        need(triangles, "triangles");
        std::vector<std::array<double, 3>> pts(n_vertices);
        for (int64_t i = 0; i < n_vertices; ++i) pts[i] = {vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]};
        std::vector<std::array<int64_t, 3>> tris(n_triangles);
        for (int64_t i = 0; i < n_triangles; ++i) tris[i] = {triangles[3 * i], triangles[3 * i + 1], triangles[3 * i + 2]};
        auto h = std::make_unique<omh_mesh>();
        h->g = omh::make_source_mesh(pts, tris);
        *out = h.release();
    });
}

int omh_mesh_load(const char* path, omh_mesh** out) {
    return guard([&] {
        need(path, "path");
        need(out, "out");
        auto h = std::make_unique<omh_mesh>();
        h->g = omh::load_source_mesh(path);
        *out = h.release();
    });
}

void omh_mesh_free(omh_mesh* m) { delete m; }

int omh_mesh_sizes(const omh_mesh* m, int64_t out[2]) {
    return guard([&] {
        need(m, "mesh");
        need(out, "out");
        out[0] = (int64_t)m->g->vertices.size();
        out[1] = (int64_t)m->g->meshes[0].tris.size();
    });
}

int omh_mesh_intersects(omh_geometry* g, omh_mesh* m, int32_t* intersects, void* stream) {
    return guard([&] {
        need(g, "geometry");
        need(m, "mesh");
        need(intersects, "intersects");
        *intersects = omh::meshes_intersect(*g->g, *m->g, S(stream)) ? 1 : 0;
    });
}

int omh_surf_source_mat(omh_geometry* g, omh_mesh* m, int order, int flags, double* d_out, int64_t ld,
                        int32_t ordering, void* stream) {
    return guard([&] {
        need(g, "geometry");
        need(m, "mesh");
        need(d_out, "d_out");
        if (flags & ~OMH_HEADMAT_FAR_FIELD) throw std::invalid_argument("unknown flags");
        omh::surf_source_mat(*g->g, *m->g, order, d_out, layout_of(*g->g, ld, ordering), S(stream),
                             (flags & OMH_HEADMAT_FAR_FIELD) != 0);
    });
}

int omh_surf_source2meg_mat(omh_mesh* m, const double* d_pts, const double* d_dirs, int64_t npts, double* d_out,
                            int64_t ld, void* stream) {
    return guard([&] {
        need(m, "mesh");
        need(d_out, "d_out");
        if (npts < 0 || ld < std::max<int64_t>(1, npts)) throw std::invalid_argument("bad sizes");
        if (npts > 0) {
            need(d_pts, "d_pts");
            need(d_dirs, "d_dirs");
        }
        omh::surf2meg(*m->g, d_pts, d_dirs, npts, d_out, ld, S(stream));
    });
}

int omh_surf2vol_mat(omh_geometry* g, const double* d_points, int64_t npts, double* d_out, int64_t ld,
                     int32_t ordering, int64_t* h_kept, int64_t* n_kept, void* stream) {
    return guard([&] {
        need(g, "geometry");
        need(n_kept, "n_kept");
        if (npts < 0 || ld < std::max<int64_t>(1, npts)) throw std::invalid_argument("bad sizes");
        if (npts > 0) {
            need(d_points, "d_points");
            need(d_out, "d_out");
        }
        omh::Layout L = layout_of(*g->g, g->g->size(), ordering);
        L.ld = ld;
        const std::vector<int64_t> kept = omh::surf2vol(*g->g, d_points, npts, d_out, L, S(stream));
        *n_kept = (int64_t)kept.size();
        if (h_kept)
            for (size_t i = 0; i < kept.size(); ++i) h_kept[i] = kept[i];
    });
}

int omh_dip_source2internal_pot_mat(omh_geometry* g, const double* d_dipoles, int64_t n_dip, const int32_t* h_domains,
                                    const double* d_points, int64_t npts, double* d_out, int64_t ld, int64_t* h_kept,
                                    int64_t* n_kept, void* stream) {
    return guard([&] {
        need(g, "geometry");
        need(n_kept, "n_kept");
        if (npts < 0 || n_dip < 0 || ld < std::max<int64_t>(1, npts)) throw std::invalid_argument("bad sizes");
        if (npts > 0) need(d_points, "d_points");
        if (n_dip > 0) need(d_dipoles, "d_dipoles");
        if (npts > 0 && n_dip > 0) need(d_out, "d_out");
        std::vector<int> dom;
        if (h_domains) {
            dom.assign(h_domains, h_domains + n_dip);
            for (int d : dom)
                if (d < 0 || d >= (int)g->g->domains.size()) throw std::invalid_argument("domain index out of range");
        } else if (n_dip > 0) {
            dom = omh::locate_domains(*g->g, d_dipoles, 6, n_dip, S(stream));
        }
        const std::vector<int64_t> kept =
            omh::dip2internal_pot(*g->g, d_dipoles, n_dip, dom, d_points, npts, d_out, ld, S(stream));
        *n_kept = (int64_t)kept.size();
        if (h_kept)
            for (size_t i = 0; i < kept.size(); ++i) h_kept[i] = kept[i];
    });
}

int omh_eit_source_mat(omh_geometry* g, const double* d_positions, const double* h_radii, int64_t n_el, int order,
                       double* d_out, int64_t ld, int32_t ordering, void* stream) {
    return guard([&] {
        need(g, "geometry");
        if (n_el < 0) throw std::invalid_argument("negative number of electrodes");
        if (n_el > 0) {
            need(d_positions, "d_positions");
            need(d_out, "d_out");
        }
        omh::eit_source_mat(*g->g, d_positions, h_radii, n_el, order, d_out, layout_of(*g->g, ld, ordering),
                            S(stream));
    });
}

int omh_dip_source2meg_mat(const double* d_dipoles, int64_t n_dip, const double* d_pts, const double* d_dirs,
                           int64_t npts, double* d_out, int64_t ld, void* stream) {
    return guard([&] {
        need(d_out, "d_out");
        if (npts < 0 || n_dip < 0 || ld < std::max<int64_t>(1, npts)) throw std::invalid_argument("bad sizes");
        if (npts > 0 && n_dip > 0) {
            need(d_dipoles, "d_dipoles");
            need(d_pts, "d_pts");
            need(d_dirs, "d_dirs");
        }
        omh::dip2meg(d_dipoles, n_dip, d_pts, d_dirs, npts, d_out, ld, S(stream));
    });
}

int64_t omh_factor_workspace_bytes(int64_t n, int32_t elem_size) {
    if (n <= 0 || (elem_size != 4 && elem_size != 8)) return -1;
    return (int64_t)omh::factor_workspace_bytes(n, (size_t)elem_size);
}

int omh_sym_factor_f64(double* d_A, int64_t n, int64_t ld, void* d_work, int32_t* info, void* stream) {
    return guard([&] {
        need(d_A, "d_A");
        need(d_work, "d_work");
        if (n <= 0 || ld < n) throw std::invalid_argument("bad matrix size");
        const auto fi = omh::sym_factor<double>(d_A, n, ld, d_work, S(stream));
        if (info) *info = fi.info;
    });
}

int omh_sym_factor_f32(float* d_A, int64_t n, int64_t ld, void* d_work, int32_t* info, void* stream) {
    return guard([&] {
        need(d_A, "d_A");
        need(d_work, "d_work");
        if (n <= 0 || ld < n) throw std::invalid_argument("bad matrix size");
        const auto fi = omh::sym_factor<float>(d_A, n, ld, d_work, S(stream));
        if (info) *info = fi.info;
    });
}

int omh_sym_factor_f32_split(float* d_A, int64_t n, int64_t ld, void* d_work, int32_t* info, void* stream) {
    return guard([&] {
        need(d_A, "d_A");
        need(d_work, "d_work");
        if (n <= 0 || ld < n) throw std::invalid_argument("bad matrix size");
        const auto fi = omh::sym_factor<float>(d_A, n, ld, d_work, S(stream), true);
        if (info) *info = fi.info;
    });
}

extern "C++" {
// Block structure of the *_blocks entry points: segments from seg_start, the
// matrix's block pattern from hfill (lower triangle, s > t), L's from its
// symbolic factorisation.
static omh::BlockFill block_pattern(int64_t n, int32_t nseg, const int64_t* seg_start, const uint8_t* hfill) {
    need(seg_start, "seg_start");
    need(hfill, "hfill");
    if (nseg < 1 || nseg > 4096) throw std::invalid_argument("bad segment count");
    omh::BlockFill bf;
    bf.start.assign(seg_start, seg_start + nseg);
    bf.start.push_back(n);
    if (bf.start[0] != 0) throw std::invalid_argument("the first segment must start at 0");
    for (int s = 0; s < nseg; ++s)
        if (bf.start[s + 1] <= bf.start[s]) throw std::invalid_argument("segment starts must increase below n");
    bf.hfill.assign(hfill, hfill + (size_t)nseg * nseg);
    bf.symbolic();
    return bf;
}

// Segment starts with n appended (checked as in block_pattern) and a checked
// segment permutation.
static std::vector<int64_t> segment_starts(int64_t n, int32_t nseg, const int64_t* seg_start) {
    need(seg_start, "seg_start");
    if (nseg < 1 || nseg > 4096) throw std::invalid_argument("bad segment count");
    std::vector<int64_t> start(seg_start, seg_start + nseg);
    start.push_back(n);
    if (start[0] != 0) throw std::invalid_argument("the first segment must start at 0");
    for (int s = 0; s < nseg; ++s)
        if (start[s + 1] <= start[s]) throw std::invalid_argument("segment starts must increase below n");
    return start;
}

static std::vector<int32_t> segment_order(int32_t nseg, const int32_t* seg_order) {
    need(seg_order, "seg_order");
    std::vector<int32_t> order(seg_order, seg_order + nseg);
    std::vector<char> seen((size_t)nseg, 0);
    for (const int32_t s : order) {
        if (s < 0 || s >= nseg || seen[(size_t)s]) throw std::invalid_argument("seg_order is not a permutation of the segments");
        seen[(size_t)s] = 1;
    }
    return order;
}

// The row map on the device (its own scratch slot; the upload has finished
// when this returns).
static const int32_t* device_rowmap(const std::vector<int32_t>& map, hipStream_t st) {
    int32_t* d = (int32_t*)omh::scratch(omh::SCR_ROWMAP, sizeof(int32_t) * map.size(), st);
    OMH_HIP(hipMemcpyAsync(d, map.data(), sizeof(int32_t) * map.size(), hipMemcpyHostToDevice, st));
    OMH_HIP(hipStreamSynchronize(st));
    return d;
}

// The matrix's block structure (segments start, lower pattern hfill) in the
// segment order `order`: segment q of the reordered matrix is segment
// order[q] of the matrix.
static omh::BlockFill reordered_pattern(const std::vector<int64_t>& start, const uint8_t* hfill,
                                        const std::vector<int32_t>& order) {
    need(hfill, "hfill");
    const int ns = (int)order.size();
    std::vector<int64_t> fstart(1, 0);
    for (int q = 0; q + 1 < ns; ++q) fstart.push_back(fstart.back() + start[order[q] + 1] - start[order[q]]);
    std::vector<uint8_t> fh((size_t)ns * ns, 0);
    for (int q = 0; q < ns; ++q)
        for (int p = 0; p < q; ++p) {
            const int a = std::max(order[q], order[p]), b = std::min(order[q], order[p]);
            fh[(size_t)q * ns + p] = hfill[(size_t)a * ns + b] ? 1 : 0;
        }
    return block_pattern(start.back(), ns, fstart.data(), fh.data());
}

template <typename T>
static int factor_blocks(T* d_A, int64_t n, int64_t ld, void* d_work, int32_t nseg, const int64_t* seg_start,
                         const uint8_t* hfill, int32_t* info, void* stream, bool split16) {
    return guard([&] {
        need(d_A, "d_A");
        need(d_work, "d_work");
        if (n <= 0 || ld < n) throw std::invalid_argument("bad matrix size");
        hipStream_t st = S(stream);
        const omh::BlockFill bf = block_pattern(n, nseg, seg_start, hfill);
        if (omh::block_fill_violated<T>(d_A, n, ld, bf, st))
            throw std::invalid_argument("a block declared zero holds non-zero entries");
        const auto fi = omh::sym_factor<T>(d_A, n, ld, d_work, st, split16, &bf);
        if (info) *info = fi.info;
    });
}
}  // extern "C++"

int omh_sym_factor_f32_split_blocks(float* d_A, int64_t n, int64_t ld, void* d_work, int32_t nseg,
                                    const int64_t* seg_start, const uint8_t* hfill, int32_t* info, void* stream) {
    return factor_blocks<float>(d_A, n, ld, d_work, nseg, seg_start, hfill, info, stream, true);
}

int omh_sym_factor_f32_blocks(float* d_A, int64_t n, int64_t ld, void* d_work, int32_t nseg, const int64_t* seg_start,
                              const uint8_t* hfill, int32_t* info, void* stream) {
    return factor_blocks<float>(d_A, n, ld, d_work, nseg, seg_start, hfill, info, stream, false);
}

int omh_sym_factor_f64_blocks(double* d_A, int64_t n, int64_t ld, void* d_work, int32_t nseg,
                              const int64_t* seg_start, const uint8_t* hfill, int32_t* info, void* stream) {
    return factor_blocks<double>(d_A, n, ld, d_work, nseg, seg_start, hfill, info, stream, false);
}

int omh_sym_solve_f64(const double* d_L, int64_t n, int64_t ld, const void* d_work, double* d_B, int64_t ldb,
                      int64_t nrhs, void* stream) {
    return guard([&] {
        need(d_L, "d_L");
        need(d_work, "d_work");
        need(d_B, "d_B");
        if (n <= 0 || ld < n || ldb < n || nrhs < 0) throw std::invalid_argument("bad sizes");
        omh::sym_solve<double>(d_L, n, ld, d_work, d_B, ldb, nrhs, S(stream));
    });
}

int omh_sym_solve_f32(const float* d_L, int64_t n, int64_t ld, const void* d_work, float* d_B, int64_t ldb,
                      int64_t nrhs, void* stream) {
    return guard([&] {
        need(d_L, "d_L");
        need(d_work, "d_work");
        need(d_B, "d_B");
        if (n <= 0 || ld < n || ldb < n || nrhs < 0) throw std::invalid_argument("bad sizes");
        omh::sym_solve<float>(d_L, n, ld, d_work, d_B, ldb, nrhs, S(stream));
    });
}

int omh_sym_solve_f64_blocks(const double* d_L, int64_t n, int64_t ld, const void* d_work, double* d_B, int64_t ldb,
                             int64_t nrhs, int32_t nseg, const int64_t* seg_start, const uint8_t* hfill, void* stream) {
    return guard([&] {
        need(d_L, "d_L");
        need(d_work, "d_work");
        need(d_B, "d_B");
        if (n <= 0 || ld < n || ldb < n || nrhs < 0) throw std::invalid_argument("bad sizes");
        const omh::BlockFill bf = block_pattern(n, nseg, seg_start, hfill);
        omh::sym_solve<double>(d_L, n, ld, d_work, d_B, ldb, nrhs, S(stream), &bf);
    });
}

int omh_sym_solve_f32_blocks(const float* d_L, int64_t n, int64_t ld, const void* d_work, float* d_B, int64_t ldb,
                             int64_t nrhs, int32_t nseg, const int64_t* seg_start, const uint8_t* hfill, void* stream) {
    return guard([&] {
        need(d_L, "d_L");
        need(d_work, "d_work");
        need(d_B, "d_B");
        if (n <= 0 || ld < n || ldb < n || nrhs < 0) throw std::invalid_argument("bad sizes");
        const omh::BlockFill bf = block_pattern(n, nseg, seg_start, hfill);
        omh::sym_solve<float>(d_L, n, ld, d_work, d_B, ldb, nrhs, S(stream), &bf);
    });
}

int omh_factor_sign_mismatch(const void* d_work, int64_t n, int32_t elem_size, int64_t n_positive, int64_t* first,
                             void* stream) {
    return guard([&] {
        need(d_work, "d_work");
        need(first, "first");
        if (n <= 0 || (elem_size != 4 && elem_size != 8) || n_positive < 0 || n_positive > n)
            throw std::invalid_argument("bad sizes");
        *first = omh::factor_sign_mismatch(d_work, n, (size_t)elem_size, n_positive, S(stream));
    });
}

int omh_factor_sign_mismatch_segments(const void* d_work, int64_t n, int32_t elem_size, int32_t nseg,
                                      const int64_t* seg_start, const int8_t* seg_sign, int64_t* first, void* stream) {
    return guard([&] {
        need(d_work, "d_work");
        need(first, "first");
        need(seg_sign, "seg_sign");
        if (n <= 0 || (elem_size != 4 && elem_size != 8)) throw std::invalid_argument("bad sizes");
        const std::vector<int64_t> start = segment_starts(n, nseg, seg_start);
        std::vector<int8_t> expect((size_t)n);
        for (int s = 0; s < nseg; ++s) {
            if (seg_sign[s] != 1 && seg_sign[s] != -1) throw std::invalid_argument("segment signs must be +1 or -1");
            for (int64_t i = start[s]; i < start[s + 1]; ++i) expect[(size_t)i] = seg_sign[s];
        }
        *first = omh::factor_sign_mismatch(d_work, n, (size_t)elem_size, expect.data(), S(stream));
    });
}

int omh_sym_reorder(const double* d_src, int64_t lds, void* d_dst, int64_t ldd, int32_t dst_elem_size, int64_t n,
                    int32_t nseg, const int64_t* seg_start, const int32_t* seg_order, void* stream) {
    return guard([&] {
        need(d_src, "d_src");
        need(d_dst, "d_dst");
        if (n <= 0 || lds < n || ldd < n || (dst_elem_size != 4 && dst_elem_size != 8))
            throw std::invalid_argument("bad sizes");
        if ((const void*)d_src == d_dst) throw std::invalid_argument("d_dst must not be d_src");
        if (n > INT32_MAX) throw std::invalid_argument("n too large for a row map");
        const std::vector<int64_t> start = segment_starts(n, nseg, seg_start);
        const std::vector<int32_t> order = segment_order(nseg, seg_order);
        hipStream_t st = S(stream);
        if (omh::sym_reorder_blocks(d_src, lds, d_dst, ldd, dst_elem_size, start, order, st)) return;
        const int32_t* rowmap = device_rowmap(omh::segment_rowmap(start, order), st);
        omh::sym_reorder(d_src, lds, d_dst, ldd, dst_elem_size, n, rowmap, st);
    });
}

int omh_reorder_rows_f64(const double* d_src, int64_t lds, double* d_dst, int64_t ldd, int64_t n, int64_t ncol,
                         int32_t nseg, const int64_t* seg_start, const int32_t* seg_order, int32_t inverse,
                         void* stream) {
    return guard([&] {
        need(d_src, "d_src");
        need(d_dst, "d_dst");
        if (n <= 0 || ncol < 0 || lds < n || ldd < n) throw std::invalid_argument("bad sizes");
        if (d_src == d_dst) throw std::invalid_argument("d_dst must not be d_src");
        if (n > INT32_MAX) throw std::invalid_argument("n too large for a row map");
        const std::vector<int64_t> start = segment_starts(n, nseg, seg_start);
        const std::vector<int32_t> order = segment_order(nseg, seg_order);
        hipStream_t st = S(stream);
        const int32_t* rowmap = device_rowmap(omh::segment_rowmap(start, order), st);
        omh::rows_reorder_f64(d_src, lds, d_dst, ldd, n, ncol, rowmap, inverse != 0, st);
    });
}

int64_t omh_lu_workspace_bytes(int64_t n) { return n > 0 ? (int64_t)omh::lu_workspace_bytes(n) : -1; }

int omh_lu_factor_f64(double* d_A, int64_t n, int64_t ld, int32_t* d_ipiv, void* d_work, int32_t* info,
                      void* stream) {
    return guard([&] {
        need(d_A, "d_A");
        need(d_ipiv, "d_ipiv");
        need(d_work, "d_work");
        if (n <= 0 || ld < n) throw std::invalid_argument("bad sizes");
        if (n > INT32_MAX) throw std::invalid_argument("n exceeds the 32-bit pivot indices");
        const int i = omh::lu_factor(d_A, n, ld, d_ipiv, d_work, S(stream));
        if (info) *info = i;
    });
}

int omh_lu_solve_f64(const double* d_LU, int64_t n, int64_t ld, const int32_t* d_ipiv, const void* d_work,
                     double* d_B, int64_t ldb, int64_t nrhs, void* stream) {
    return guard([&] {
        need(d_LU, "d_LU");
        need(d_ipiv, "d_ipiv");
        need(d_work, "d_work");
        need(d_B, "d_B");
        if (n <= 0 || ld < n || ldb < n || nrhs < 0) throw std::invalid_argument("bad sizes");
        omh::lu_solve(d_LU, n, ld, d_ipiv, d_work, d_B, ldb, nrhs, S(stream));
    });
}

int32_t omh_refine_state(int32_t it, double rel, double prev, double tol) {
    return (int32_t)omh::refine_state(it, rel, prev, tol);
}

int omh_mixed_solve(const double* d_H, int64_t n, int64_t ldh, const float* d_L32, int64_t ld32, const void* d_work32,
                    const double* d_B, int64_t ldb, double* d_X, int64_t ldx, int64_t nrhs, int32_t max_iter,
                    double tol, int32_t* iterations, double* rel_update, void* stream) {
    return guard([&] {
        need(d_H, "d_H");
        need(d_L32, "d_L32");
        need(d_work32, "d_work32");
        need(d_B, "d_B");
        need(d_X, "d_X");
        if (n <= 0 || ldh < n || ld32 < n || ldb < n || ldx < n || nrhs < 0 || max_iter < 1)
            throw std::invalid_argument("bad sizes");
        double rel = 0.0;
        bool ok = false;
        const int it = omh::mixed_refine(d_H, n, ldh, d_L32, ld32, d_work32, d_B, ldb, d_X, ldx, nrhs, max_iter, tol,
                                         &rel, &ok, S(stream));
        if (iterations) *iterations = it;
        if (rel_update) *rel_update = rel;
        if (!ok) not_converged(it, rel, tol);
    });
}

int omh_mixed_solve_noalloc(const double* d_H, int64_t n, int64_t ldh, const float* d_L32, int64_t ld32,
                            const void* d_work32, const double* d_B, int64_t ldb, double* d_X, int64_t ldx,
                            int64_t nrhs, int32_t max_iter, double tol, int32_t* iterations, double* rel_update,
                            void* stream) {
    return guard([&] {
        need(d_H, "d_H");
        need(d_L32, "d_L32");
        need(d_work32, "d_work32");
        need(d_B, "d_B");
        need(d_X, "d_X");
        if (n <= 0 || ldh < n || ld32 < n || ldb < n || ldx < n || nrhs < 0 || max_iter < 1)
            throw std::invalid_argument("bad sizes");
        double rel = 0.0;
        bool ok = false;
        const int it = omh::mixed_refine_noalloc(d_H, n, ldh, d_L32, ld32, d_work32, d_B, ldb, d_X, ldx, nrhs,
                                                 max_iter, tol, &rel, &ok, S(stream));
        if (iterations) *iterations = it;
        if (rel_update) *rel_update = rel;
        if (!ok) not_converged(it, rel, tol);
    });
}

int omh_mixed_solve_blocks(const double* d_H, int64_t n, int64_t ldh, const float* d_L32, int64_t ld32,
                           const void* d_work32, const double* d_B, int64_t ldb, double* d_X, int64_t ldx,
                           int64_t nrhs, int32_t max_iter, double tol, int32_t nseg, const int64_t* seg_start,
                           const uint8_t* hfill, int32_t* iterations, double* rel_update, void* stream) {
    return guard([&] {
        need(d_H, "d_H");
        need(d_L32, "d_L32");
        need(d_work32, "d_work32");
        need(d_B, "d_B");
        need(d_X, "d_X");
        if (n <= 0 || ldh < n || ld32 < n || ldb < n || ldx < n || nrhs < 0 || max_iter < 1)
            throw std::invalid_argument("bad sizes");
        const omh::BlockFill bf = block_pattern(n, nseg, seg_start, hfill);
        double rel = 0.0;
        bool ok = false;
        const int it = omh::mixed_refine(d_H, n, ldh, d_L32, ld32, d_work32, d_B, ldb, d_X, ldx, nrhs, max_iter, tol,
                                         &rel, &ok, S(stream), &bf);
        if (iterations) *iterations = it;
        if (rel_update) *rel_update = rel;
        if (!ok) not_converged(it, rel, tol);
    });
}

int omh_mixed_solve_reordered(const double* d_H, int64_t n, int64_t ldh, const float* d_L32, int64_t ld32,
                              const void* d_work32, const double* d_B, int64_t ldb, double* d_X, int64_t ldx,
                              int64_t nrhs, int32_t max_iter, double tol, int32_t nseg, const int64_t* seg_start,
                              const uint8_t* hfill, const int32_t* seg_order, int32_t* iterations,
                              double* rel_update, void* stream) {
    return guard([&] {
        need(d_H, "d_H");
        need(d_L32, "d_L32");
        need(d_work32, "d_work32");
        need(d_B, "d_B");
        need(d_X, "d_X");
        if (n <= 0 || ldh < n || ld32 < n || ldb < n || ldx < n || nrhs < 0 || max_iter < 1)
            throw std::invalid_argument("bad sizes");
        if (n > INT32_MAX) throw std::invalid_argument("n too large for a row map");
        const omh::BlockFill bf = block_pattern(n, nseg, seg_start, hfill);
        const std::vector<int32_t> order = segment_order(nseg, seg_order);
        hipStream_t st = S(stream);
        omh::FactorOrder fo;
        fo.fill = reordered_pattern(bf.start, hfill, order);
        fo.rowmap = device_rowmap(omh::segment_rowmap(bf.start, order), st);
        double rel = 0.0;
        bool ok = false;
        const int it = omh::mixed_refine(d_H, n, ldh, d_L32, ld32, d_work32, d_B, ldb, d_X, ldx, nrhs, max_iter, tol,
                                         &rel, &ok, st, &bf, &fo);
        if (iterations) *iterations = it;
        if (rel_update) *rel_update = rel;
        if (!ok) not_converged(it, rel, tol);
    });
}

int omh_gemm_f64(int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k, double alpha, const double* d_A,
                 int64_t lda, const double* d_B, int64_t ldb, double beta, double* d_C, int64_t ldc, void* stream) {
    return guard([&] {
        if (m < 0 || n < 0 || k < 0) throw std::invalid_argument("negative size");
        omh::gemm_entry<double>(trans_a != 0, trans_b != 0, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc,
                                S(stream));
    });
}

int omh_gemm_f32(int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k, float alpha, const float* d_A,
                 int64_t lda, const float* d_B, int64_t ldb, float beta, float* d_C, int64_t ldc, void* stream) {
    return guard([&] {
        if (m < 0 || n < 0 || k < 0) throw std::invalid_argument("negative size");
        omh::gemm_entry<float>(trans_a != 0, trans_b != 0, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc,
                               S(stream));
    });
}

int omh_convert_f64_f32(const double* d_src, int64_t lds, float* d_dst, int64_t ldd, int64_t m, int64_t n,
                        int32_t lower_only, void* stream) {
    return guard([&] {
        need(d_src, "d_src");
        need(d_dst, "d_dst");
        omh::convert_f64_to_f32(d_src, lds, d_dst, ldd, m, n, lower_only != 0, S(stream));
    });
}

int omh_sparse_rows(const int64_t* d_cols, const double* d_vals, int64_t ne, const double* d_Hm, int64_t ldh, int64_t n,
                    double* d_out, int64_t ldo, void* stream) {
    return guard([&] {
        need(d_cols, "d_cols");
        need(d_vals, "d_vals");
        need(d_Hm, "d_Hm");
        need(d_out, "d_out");
        omh::sparse_rows(d_cols, d_vals, ne, d_Hm, ldh, n, d_out, ldo, S(stream));
    });
}

int omh_sparse_transpose_dense(const int64_t* d_rows, const double* d_vals, int64_t ne, int64_t n, double* d_out,
                               int64_t ldo, void* stream) {
    return guard([&] {
        need(d_rows, "d_rows");
        need(d_vals, "d_vals");
        need(d_out, "d_out");
        if (ne < 0 || n < 0 || ldo < std::max<int64_t>(1, n)) throw std::invalid_argument("bad sizes");
        omh::sparse_transpose_dense(d_rows, d_vals, ne, n, d_out, ldo, S(stream));
    });
}

static omh::Layout raw_layout(int64_t n, int64_t nv, int64_t ld, int ordering) {
    if (nv < 0 || nv > n) throw std::invalid_argument("bad vertex-unknown count");
    if (ld < n) throw std::invalid_argument("leading dimension too small");
    omh::Layout L;
    L.ld = ld;
    L.nv = nv;
    L.nt = n - nv;
    L.tri_first = ordering == OMH_ORDER_TRIANGLES_FIRST ? 1 : 0;
    return L;
}

int omh_sym_unpack(const double* d_packed, int64_t n, int64_t nv, double* d_H, int64_t ld, int ordering, void* stream) {
    return guard([&] {
        need(d_packed, "d_packed");
        need(d_H, "d_H");
        omh::sym_unpack(d_packed, n, d_H, raw_layout(n, nv, ld, ordering), S(stream));
    });
}

int omh_sym_pack(const double* d_H, int64_t n, int64_t nv, int64_t ld, int ordering, double* d_packed, void* stream) {
    return guard([&] {
        need(d_packed, "d_packed");
        need(d_H, "d_H");
        omh::sym_pack(d_H, n, raw_layout(n, nv, ld, ordering), d_packed, S(stream));
    });
}

int omh_permute_rows(const double* d_src, int64_t lds, double* d_dst, int64_t ldd, int64_t n, int64_t nv,
                     int64_t ncol, int ordering, int32_t to_reference, void* stream) {
    return guard([&] {
        need(d_src, "d_src");
        need(d_dst, "d_dst");
        if (lds < n || ldd < n) throw std::invalid_argument("leading dimension too small");
        omh::permute_rows(d_src, lds, d_dst, ldd, n, ncol, raw_layout(n, nv, n, ordering), to_reference != 0,
                          S(stream));
    });
}

int omh_batch_create(int32_t n_heads, omh_geometry* const* geoms, const double* const* h_dipoles,
                     const int64_t* n_dipoles, const double* const* h_electrodes, const int64_t* n_electrodes,
                     int32_t precision, omh_batch** out) {
    return guard([&] {
        need(out, "out");
        *out = nullptr;
        if (n_heads <= 0) throw std::invalid_argument("a batch needs at least one head");
        need(geoms, "geoms");
        need(h_dipoles, "h_dipoles");
        need(n_dipoles, "n_dipoles");
        need(h_electrodes, "h_electrodes");
        need(n_electrodes, "n_electrodes");
        std::vector<omh::BatchHeadInput> in(n_heads);
        for (int h = 0; h < n_heads; ++h) {
            need(geoms[h], "geometry");
            if (n_dipoles[h] <= 0 || n_electrodes[h] <= 0)
                throw std::invalid_argument("each head needs dipoles and electrodes");
            need(h_dipoles[h], "dipoles");
            need(h_electrodes[h], "electrodes");
            in[h].g = geoms[h]->g.get();
            in[h].dipoles.assign(h_dipoles[h], h_dipoles[h] + 6 * n_dipoles[h]);
            in[h].electrodes.assign(h_electrodes[h], h_electrodes[h] + 3 * n_electrodes[h]);
        }
        auto b = std::make_unique<omh_batch>();
        b->plan = std::make_unique<omh::BatchPlan>(in, precision);
        *out = b.release();
    });
}

int omh_batch_run(omh_batch* b, int32_t use_graph, int32_t* used_graph, void* stream) {
    return guard([&] {
        need(b, "batch");
        const bool g = b->plan->run(S(stream), use_graph != 0);
        if (used_graph) *used_graph = g ? 1 : 0;
    });
}

int omh_batch_tune(omh_batch* b, double* h_results, int32_t max_results, int32_t* n_results, void* stream) {
    return guard([&] {
        need(b, "batch");
        const auto res = b->plan->tune(S(stream));
        if (n_results) *n_results = (int32_t)res.size();
        if (h_results)
            for (int i = 0; i < (int)res.size() && i < max_results; ++i) {
                h_results[4 * i] = res[i].tile;
                h_results[4 * i + 1] = (double)res[i].grid_cap;
                h_results[4 * i + 2] = res[i].split16 ? 1.0 : 0.0;
                h_results[4 * i + 3] = res[i].seconds;
            }
    });
}

int omh_batch_gain(const omh_batch* b, int32_t head, const double** d_gain, int64_t* n_el, int64_t* n_dip) {
    return guard([&] {
        need(b, "batch");
        need(d_gain, "d_gain");
        if (head < 0 || head >= b->plan->n_heads()) throw std::invalid_argument("head index out of range");
        int64_t ne = 0, nd = 0;
        *d_gain = b->plan->gain(head, ne, nd);
        if (n_el) *n_el = ne;
        if (n_dip) *n_dip = nd;
    });
}

int omh_batch_stats(omh_batch* b, int32_t head, int32_t* factor_info, double* rel_update, void* stream) {
    return guard([&] {
        need(b, "batch");
        if (head < 0 || head >= b->plan->n_heads()) throw std::invalid_argument("head index out of range");
        int fi = 0;
        double ru = 0.0;
        b->plan->stats(head, fi, ru, S(stream));
        if (factor_info) *factor_info = fi;
        if (rel_update) *rel_update = ru;
    });
}

int omh_batch_config(const omh_batch* b, int32_t* refine_steps, int32_t* tile, int64_t* grid_cap, int64_t* launches,
                     int64_t* graph_nodes, int32_t* split16) {
    return guard([&] {
        need(b, "batch");
        if (split16) *split16 = b->plan->trailing_split16() ? 1 : 0;
        if (refine_steps) *refine_steps = b->plan->refine_steps();
        if (tile) *tile = b->plan->trailing_tile();
        if (grid_cap) *grid_cap = b->plan->trailing_grid_cap();
        if (launches) *launches = b->plan->launches();
        if (graph_nodes) *graph_nodes = b->plan->graph_nodes();
    });
}

int omh_batch_free(omh_batch* b) {
    return guard([&] { delete b; });
}

int omh_release_scratch(void) {
    return guard([&] { omh::release_scratch(); });
}

int omh_profile_enable(int32_t on) {
    return guard([&] { omh::prof_enable(on != 0); });
}

int omh_profile_read(const char* name, int64_t* calls, double* seconds, double* work) {
    return guard([&] {
        need(name, "name");
        need(calls, "calls");
        need(seconds, "seconds");
        need(work, "work");
        omh::prof_read(name, *calls, *seconds, *work);
    });
}

int omh_probe_mfma_layout(double* d_out, void* stream) {
    return guard([&] {
        need(d_out, "d_out");
        omh::probe_mfma_layout(d_out, S(stream));
    });
}

}  // extern "C"
