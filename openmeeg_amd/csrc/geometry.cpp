// Host-side geometry model; see geometry.hpp for the reference map.
#include "geometry.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <deque>
#include <fstream>
#include <limits>
#include <set>
#include <sstream>
#include <stdexcept>

namespace omh {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kK = 1.0 / (4.0 * kPi);   // constants.h:23

// om_common.h:24-26
bool almost_equal(double x, double y, double eps = 1e3) {
    return std::abs(x - y) < std::numeric_limits<double>::epsilon() * std::abs(x + y) * eps ||
           std::abs(x - y) < std::numeric_limits<double>::min();
}

// Triangle::solid_angle, triangle.h:57-62,123-136.
double tri_solid_angle(const std::array<double, 3>& a, const std::array<double, 3>& b,
                       const std::array<double, 3>& c, const double* v) {
    double e1[3], e2[3], on[3];
    for (int k = 0; k < 3; ++k) { e1[k] = b[k] - a[k]; e2[k] = c[k] - a[k]; }
    on[0] = e1[1] * e2[2] - e1[2] * e2[1];
    on[1] = e1[2] * e2[0] - e1[0] * e2[2];
    on[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double nr = std::sqrt(on[0] * on[0] + on[1] * on[1] + on[2] * on[2]);
    const double area = 0.5 * nr;
    const double n[3] = {on[0] / nr, on[1] / nr, on[2] / nr};
    double r[3][3];
    for (int k = 0; k < 3; ++k) { r[0][k] = a[k] - v[k]; r[1][k] = b[k] - v[k]; r[2][k] = c[k] - v[k]; }
    auto dot = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    const double d0 = std::sqrt(dot(r[0], r[0])), d1 = std::sqrt(dot(r[1], r[1])), d2 = std::sqrt(dot(r[2], r[2]));
    const double p0 = dot(r[1], r[2]), p1 = dot(r[2], r[0]), p2 = dot(r[0], r[1]);
    const double prod = d0 * d1 * d2;
    const double det = 2 * area * dot(r[0], n);
    return (std::fabs(det) < 1e-10) ? 0.0 : 2 * std::atan2(det, prod + (d0 * p0 + d1 * p1 + d2 * p2));
}

std::string trim(const std::string& s) {
    const auto b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos) return "";
    const auto e = s.find_last_not_of(" \t\r\n");
    return s.substr(b, e - b + 1);
}

std::string dirname_of(const std::string& path) {
    const auto p = path.find_last_of('/');
    return p == std::string::npos ? std::string(".") : path.substr(0, p);
}

// MeshIOs/tri.h:26-56
void read_tri(const std::string& path, std::vector<std::array<double, 3>>& pts,
              std::vector<std::array<i64, 3>>& tris) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open mesh file " + path);
    char ch;
    i64 np;
    f >> ch >> np;
    pts.resize(np);
    for (i64 i = 0; i < np; ++i) {
        double n[3];
        f >> pts[i][0] >> pts[i][1] >> pts[i][2] >> n[0] >> n[1] >> n[2];
    }
    i64 nt;
    f >> ch >> nt >> nt >> nt;
    tris.resize(nt);
    for (i64 i = 0; i < nt; ++i) f >> tris[i][0] >> tris[i][1] >> tris[i][2];
    if (!f) throw std::runtime_error("malformed mesh file " + path);
    for (const auto& t : tris)
        for (int k = 0; k < 3; ++k)
            if (t[k] < 0 || t[k] >= np) throw std::runtime_error("triangle index out of range in " + path);
}

// section_name + filename of GeomFile.h:86-130: "Keyword: "path"" (unnamed,
// name = position+1) or "Keyword name: "path"".
void named_entry(const std::string& line, const std::string& kw, int k, std::string& name, std::string& path) {
    std::string rest = trim(line.substr(kw.size()));
    std::string val;
    if (!rest.empty() && rest[0] == ':') {
        name = std::to_string(k + 1);
        val = trim(rest.substr(1));
    } else {
        const auto c = rest.find(':');
        if (c == std::string::npos) throw std::runtime_error("malformed line: " + line);
        name = trim(rest.substr(0, c));
        val = trim(rest.substr(c + 1));
    }
    if (val.size() >= 2 && val.front() == '"' && val.back() == '"') val = val.substr(1, val.size() - 2);
    path = val;
}

}  // namespace

int Domain::mesh_orientation(int m) const {   // domain.h:95-101, 43-45
    for (const auto& b : boundaries)
        for (const auto& om : b.first.omeshes)
            if (om.first == m) return b.second ? om.second : -om.second;
    return 0;
}

i64 Geometry::add_vertex(const std::array<double, 3>& v) {
    auto it = vmap.find(v);
    if (it != vmap.end()) return it->second;
    const i64 idx = (i64)vertices.size();
    vertices.push_back(v);
    vmap.emplace(v, idx);
    return idx;
}

void Geometry::import_meshes(const std::vector<std::string>& names,
                             const std::vector<std::vector<std::array<double, 3>>>& pts,
                             const std::vector<std::vector<std::array<i64, 3>>>& tris) {
    std::vector<std::vector<i64>> indmaps(names.size());
    for (size_t m = 0; m < names.size(); ++m) {
        indmaps[m].reserve(pts[m].size());
        for (const auto& p : pts[m]) indmaps[m].push_back(add_vertex(p));
    }
    for (size_t m = 0; m < names.size(); ++m) {
        Mesh mesh;
        mesh.name = names[m];
        mesh.verts = indmaps[m];
        mesh.tris.reserve(tris[m].size());
        for (const auto& t : tris[m]) {
            for (int k = 0; k < 3; ++k)
                if (t[k] < 0 || t[k] >= (i64)indmaps[m].size())
                    throw std::invalid_argument("triangle vertex index out of range in mesh " + names[m]);
            mesh.tris.push_back({indmaps[m][t[0]], indmaps[m][t[1]], indmaps[m][t[2]]});
        }
        if (mesh.tris.empty()) throw std::invalid_argument("mesh " + names[m] + " has no triangles");
        meshes.push_back(std::move(mesh));
    }
}

int Geometry::mesh_index(const std::string& name) const {
    for (size_t i = 0; i < meshes.size(); ++i)
        if (meshes[i].name == name) return (int)i;
    throw std::invalid_argument("unknown mesh " + name);
}

int Geometry::domain_index(const std::string& name) const {
    for (size_t i = 0; i < domains.size(); ++i)
        if (domains[i].name == name) return (int)i;
    throw std::invalid_argument("unknown domain " + name);
}

double Geometry::mesh_solid_angle(int m, const double p[3]) const {   // mesh.cpp:239-244
    double s = 0.0;
    for (const auto& t : meshes[m].tris) s += tri_solid_angle(vertices[t[0]], vertices[t[1]], vertices[t[2]], p);
    return s;
}

double Geometry::interface_solid_angle(const Interface& itf, const double p[3]) const {   // interface.cpp:40-45
    double s = 0.0;
    for (const auto& om : itf.omeshes) s += om.second * mesh_solid_angle(om.first, p);
    return s;
}

bool Geometry::interface_contains_sa(double sa) const {   // interface.cpp:18-35
    if (almost_equal(sa, -4 * kPi)) return true;
    if (almost_equal(sa, 0.0)) return false;
    if (almost_equal(sa, 4 * kPi)) return true;
    return std::abs(sa) > 2 * kPi;
}

int Geometry::domain_of_point(const double p[3]) const {
    std::vector<double> sa(meshes.size());
    for (size_t m = 0; m < meshes.size(); ++m) sa[m] = mesh_solid_angle((int)m, p);
    for (size_t k = 0; k < domains.size(); ++k) {
        bool inside = true;
        for (const auto& b : domains[k].boundaries) {
            double a = 0.0;
            for (const auto& om : b.first.omeshes) a += om.second * sa[om.first];
            inside = inside && (interface_contains_sa(a) == b.second);
        }
        if (inside) return (int)k;
    }
    return -1;
}

// Interface::is_mesh_orientations_coherent, interface.cpp:55-95, at the
// bounding-box centre.
void Geometry::fix_interface_orientation(Interface& itf) const {
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (const auto& om : itf.omeshes)
        for (i64 v : meshes[om.first].verts)
            for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], vertices[v][k]); hi[k] = std::max(hi[k], vertices[v][k]); }
    const double c[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
    const double sa = interface_solid_angle(itf, c);
    if (almost_equal(sa, 0.0)) throw std::invalid_argument("interface " + itf.name + ": bounding-box centre is outside");
    if (almost_equal(sa, 4 * kPi))
        for (auto& om : itf.omeshes) om.second = -om.second;
}

std::vector<int> Geometry::domains_of(int m) const {
    std::vector<int> r;
    for (size_t k = 0; k < domains.size(); ++k)
        if (domains[k].mesh_orientation(m) != 0) r.push_back((int)k);
    return r;
}

std::vector<int> Geometry::common_domains(int m1, int m2) const {
    const auto a = domains_of(m1), b = domains_of(m2);
    std::vector<int> r;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(r));
    return r;
}

double Geometry::sigma(int m1, int m2) const {
    double s = 0.0;
    for (int k : common_domains(m1, m2)) s += domains[k].cond;
    return s;
}

double Geometry::sigma_inv(int m1, int m2) const {
    double s = 0.0;
    for (int k : common_domains(m1, m2)) s += 1.0 / domains[k].cond;
    return s;
}

double Geometry::indicator(int m1, int m2) const { return (double)common_domains(m1, m2).size(); }

int Geometry::relative_orientation(int m1, int m2) const {
    if (m1 == m2) return 1;
    const auto d = common_domains(m1, m2);
    if (d.empty()) return 0;
    return domains[d[0]].mesh_orientation(m1) == domains[d[0]].mesh_orientation(m2) ? 1 : -1;
}

void Geometry::mark_current_barriers() {
    std::set<i64> inval;
    for (auto& d : domains)
        if (almost_equal(d.cond, 0.0))
            for (auto& b : d.boundaries)
                for (auto& om : b.first.omeshes) {
                    Mesh& m = meshes[om.first];
                    const bool fully = m.current_barrier;
                    m.current_barrier = true;
                    if (fully) {
                        m.isolated = true;
                        m.outermost = false;
                        inval.insert(m.verts.begin(), m.verts.end());
                    }
                }
    for (auto& d : domains)
        if (almost_equal(d.cond, 0.0))
            for (auto& b : d.boundaries)
                if (!b.second)
                    for (auto& om : b.first.omeshes) {
                        Mesh& m = meshes[om.first];
                        if (m.current_barrier && !m.isolated) m.outermost = true;
                    }
    std::set<i64> shared;
    for (i64 v : inval)
        for (const auto& m : meshes)
            if (!m.isolated && std::find(m.verts.begin(), m.verts.end(), v) != m.verts.end()) shared.insert(v);
    invalid.assign(vertices.size(), 0);
    for (i64 v : inval)
        if (!shared.count(v)) invalid[v] = 1;
    // connected parts separated by non-conductive domains
    std::deque<int> cand;
    for (size_t i = 0; i < meshes.size(); ++i) cand.push_back((int)i);
    parts.clear();
    while (!cand.empty()) {
        const int c = cand.front();
        cand.pop_front();
        std::vector<int> conn{c};
        std::deque<int> pending{c};
        while (!pending.empty()) {
            const int cur = pending.front();
            pending.pop_front();
            std::deque<int> rest;
            for (int o : cand) {
                if (!common_domains(cur, o).empty()) { conn.push_back(o); pending.push_back(o); }
                else rest.push_back(o);
            }
            cand.swap(rest);
        }
        if (conn.size() > 1 || !meshes[conn[0]].isolated) parts.push_back(conn);
    }
}

void Geometry::check_nested() {
    nested = true;
    for (size_t k = 0; k < domains.size(); ++k) {
        if ((int)k == outer_domain) continue;
        int nin = 0;
        for (const auto& b : domains[k].boundaries) nin += b.second ? 1 : 0;
        if (nin >= 2) { nested = false; return; }
    }
}

void Geometry::generate_indices() {
    if (invalid.size() != vertices.size()) invalid.assign(vertices.size(), 0);
    vert_index.assign(vertices.size(), -1);
    i64 idx = 0;
    for (size_t v = 0; v < vertices.size(); ++v)
        if (!invalid[v]) vert_index[v] = idx++;
    n_vert_unknowns = idx;
    for (auto& m : meshes)
        if (!m.isolated && !m.current_barrier) {
            m.tri_first_index = idx;
            idx += (i64)m.tris.size();
        }
    n_tri_unknowns = idx - n_vert_unknowns;
    nb_current_barrier_triangles = 0;
    for (auto& m : meshes)
        if (m.current_barrier) {
            if (!m.isolated) {
                nb_current_barrier_triangles += (i64)m.tris.size();
                m.tri_first_index = idx;
                idx += (i64)m.tris.size();
            } else {
                m.tri_first_index = -1;
            }
        }
    nb_parameters = idx;
}

void Geometry::make_pairs() {
    pairs.clear();
    for (size_t i = 0; i < meshes.size(); ++i) {
        if (meshes[i].isolated) continue;
        for (size_t j = 0; j <= i; ++j) {
            const int o = relative_orientation((int)i, (int)j);
            if (!meshes[j].isolated && sigma((int)i, (int)j) != 0.0 && o != 0) {
                const double f = o * kK;
                pairs.push_back({(int)i, (int)j, o, f * sigma_inv((int)i, (int)j), f * sigma((int)i, (int)j),
                                 -f * indicator((int)i, (int)j)});
            }
        }
    }
}

void Geometry::set_conductivity(int d, double value) {
    if (d < 0 || d >= (int)domains.size()) throw std::invalid_argument("domain index out of range");
    if (!(value > 0.0)) throw std::invalid_argument("conductivity must stay positive");
    if (almost_equal(domains[d].cond, 0.0)) throw std::invalid_argument("cannot change a non-conductive domain");
    domains[d].cond = value;
    make_pairs();
}

void Geometry::finalize() {   // geometry.h:228-249
    bool has_cond = true;
    for (const auto& d : domains) has_cond = has_cond && d.cond != -1.0;
    if (has_cond) mark_current_barriers();
    if (!domains.empty()) {
        outer_domain = -1;
        for (size_t k = 0; k < domains.size(); ++k) {
            bool outer = true;
            for (const auto& b : domains[k].boundaries)
                if (b.second) { outer = false; break; }
            if (outer) { outer_domain = (int)k; break; }
        }
        if (outer_domain < 0) throw std::invalid_argument("the geometry has no outermost domain");
        for (auto& b : domains[outer_domain].boundaries)
            for (auto& om : b.first.omeshes) meshes[om.first].outermost = true;
        check_nested();
    }
    generate_indices();
    make_pairs();
}

std::unique_ptr<Geometry> load_geometry(const std::string& geom_file, const std::string& cond_file) {
    std::ifstream f(geom_file);
    if (!f) throw std::runtime_error("cannot open geometry file " + geom_file);
    std::vector<std::string> lines;
    std::string header;
    std::getline(f, header);
    if (trim(header) != "# Domain Description 1.1")
        throw std::invalid_argument("unsupported geometry format (expected '# Domain Description 1.1'): " + geom_file);
    for (std::string ln; std::getline(f, ln);) {
        ln = trim(ln);
        if (!ln.empty() && ln[0] != '#') lines.push_back(ln);
    }
    const std::string dir = dirname_of(geom_file);
    auto full = [&](const std::string& p) { return (!p.empty() && p[0] == '/') ? p : dir + "/" + p; };
    size_t i = 0;
    auto need = [&](const char* kw) {
        if (i >= lines.size() || lines[i].rfind(kw, 0) != 0)
            throw std::invalid_argument(std::string("geometry file: expected '") + kw + "' in " + geom_file);
    };
    std::vector<std::string> names;
    std::vector<std::vector<std::array<double, 3>>> pts;
    std::vector<std::vector<std::array<i64, 3>>> tris;
    auto add_mesh_file = [&](const std::string& name, const std::string& path) {
        names.push_back(name);
        pts.emplace_back();
        tris.emplace_back();
        read_tri(full(path), pts.back(), tris.back());
    };
    if (i < lines.size() && lines[i].rfind("Meshes", 0) == 0) {
        const int nm = std::stoi(trim(lines[i].substr(6)));
        ++i;
        for (int k = 0; k < nm; ++k, ++i) {
            std::string name, path;
            named_entry(lines.at(i), "Mesh", k, name, path);
            add_mesh_file(name, path);
        }
    }
    need("Interfaces");
    std::istringstream is(lines[i].substr(10));
    int nif = 0;
    is >> nif;
    ++i;
    std::vector<std::string> iface_lines;
    const bool mesh_as_interfaces = names.empty();
    for (int k = 0; k < nif; ++k, ++i) {
        if (mesh_as_interfaces) {
            std::string name, path;
            named_entry(lines.at(i), "Interface", k, name, path);
            add_mesh_file(name, path);
        } else {
            iface_lines.push_back(lines.at(i));
        }
    }
    auto g = std::make_unique<Geometry>();
    g->import_meshes(names, pts, tris);
    std::map<std::string, Interface> interfaces;
    if (mesh_as_interfaces) {
        for (size_t m = 0; m < g->meshes.size(); ++m)
            interfaces[g->meshes[m].name] = Interface{g->meshes[m].name, {{(int)m, 1}}};
    } else {
        for (int k = 0; k < nif; ++k) {
            const std::string& ln = iface_lines[k];
            const auto c = ln.find(':');
            std::istringstream hs(ln.substr(0, c));
            std::string kw, name;
            hs >> kw >> name;
            if (name.empty()) name = std::to_string(k + 1);
            Interface itf{name, {}};
            std::istringstream ts(c == std::string::npos ? "" : ln.substr(c + 1));
            for (std::string t; ts >> t;) {
                int sgn = 1;
                if (t[0] == '-' || t[0] == '+') { sgn = t[0] == '-' ? -1 : 1; t = t.substr(1); }
                itf.omeshes.push_back({g->mesh_index(t), sgn});
            }
            interfaces[name] = itf;
        }
    }
    for (auto& kv : interfaces) g->fix_interface_orientation(kv.second);
    need("Domains");
    const int nd = std::stoi(trim(lines[i].substr(7)));
    ++i;
    for (int k = 0; k < nd; ++k, ++i) {
        const std::string& ln = lines.at(i);
        if (ln.rfind("Domain", 0) != 0) throw std::invalid_argument("geometry file: expected 'Domain'");
        const auto c = ln.find(':');
        Domain d;
        d.name = trim(ln.substr(6, c == std::string::npos ? std::string::npos : c - 6));
        std::istringstream ts(c == std::string::npos ? "" : ln.substr(c + 1));
        for (std::string t; ts >> t;) {
            if (t == "shared") break;
            bool inside = false;
            if (t[0] == '-' || t[0] == '+') { inside = t[0] == '-'; t = t.substr(1); }
            auto it = interfaces.find(t);
            if (it == interfaces.end()) throw std::invalid_argument("domain " + d.name + " uses unknown interface " + t);
            d.boundaries.push_back({it->second, inside});
        }
        g->domains.push_back(d);
    }
    std::ifstream cf(cond_file);
    if (!cf) throw std::runtime_error("cannot open conductivity file " + cond_file);
    std::map<std::string, double> conds;
    for (std::string ln; std::getline(cf, ln);) {
        ln = trim(ln);
        if (ln.empty() || ln[0] == '#') continue;
        std::istringstream ls(ln);
        std::string name;
        double v;
        if (ls >> name >> v) conds[name] = v;
    }
    for (auto& d : g->domains) {
        auto it = conds.find(d.name);
        if (it == conds.end()) throw std::invalid_argument("no conductivity for domain " + d.name);
        d.cond = it->second;
    }
    g->finalize();
    return g;
}

std::unique_ptr<Geometry> make_source_mesh(const std::vector<std::array<double, 3>>& pts,
                                           const std::vector<std::array<i64, 3>>& tris) {
    if (pts.empty()) throw std::invalid_argument("source mesh without vertices");
    if (tris.empty()) throw std::invalid_argument("source mesh without triangles");
    const i64 nv = (i64)pts.size();
    for (const auto& t : tris)
        for (int k = 0; k < 3; ++k)
            if (t[k] < 0 || t[k] >= nv) throw std::invalid_argument("source mesh triangle index out of range");
    auto g = std::make_unique<Geometry>();
    g->vertices = pts;
    Mesh m;
    m.name = "source";
    m.verts.resize(nv);
    for (i64 i = 0; i < nv; ++i) m.verts[i] = i;
    m.tris = tris;
    m.current_barrier = true;
    m.outermost = true;
    m.tri_first_index = nv;
    g->meshes.push_back(std::move(m));
    g->vert_index.resize(nv);
    for (i64 i = 0; i < nv; ++i) g->vert_index[i] = i;
    g->invalid.assign(nv, 0);
    g->nb_parameters = nv + (i64)tris.size();
    g->nb_current_barrier_triangles = (i64)tris.size();
    g->n_vert_unknowns = nv;
    return g;
}

std::unique_ptr<Geometry> load_source_mesh(const std::string& path) {
    if (path.size() < 4 || path.compare(path.size() - 4, 4, ".tri") != 0)
        throw std::invalid_argument("only .tri source meshes are read: " + path);
    std::vector<std::array<double, 3>> pts;
    std::vector<std::array<i64, 3>> tris;
    read_tri(path, pts, tris);
    return make_source_mesh(pts, tris);
}

}  // namespace omh
