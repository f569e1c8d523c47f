"""Geometry model of the CPU oracle (TEST INFRASTRUCTURE ONLY).

A pure-Python restatement of OpenMEEG's geometry layer: the .geom/.cond/.tri
readers, vertex de-duplication, current-barrier marking, outermost domain,
unknown numbering, communicating mesh pairs and conductivity coefficients.
Only tests/, the smoke() check and bench.py's cpu_baseline leg may import
this package; the product path never does.

Reference behaviour followed (file:line in the reference tree):
  * .geom parsing          OpenMEEG/include/GeometryIOs/GeomFile.h:158-300
  * .tri loading           OpenMEEG/include/MeshIOs/tri.h:26-56
  * vertex de-duplication  OpenMEEG/include/geometry.h:100-109 (exact equality)
  * mesh adjacency         OpenMEEG/include/mesh.h:252-257
  * interface orientation  OpenMEEG/src/interface.cpp:40-95
  * current barriers/parts OpenMEEG/src/geometry.cpp:424-536
  * outermost domain       OpenMEEG/src/geometry.cpp:22-41, geometry.h:167-171
  * unknown numbering      OpenMEEG/src/geometry.cpp:256-295
  * mesh pairs             OpenMEEG/src/geometry.cpp:405-420
  * sigma/sigma_inv/indicator, relative_orientation  geometry.h:182-206,289-310
  * HeadMat coefficients   OpenMEEG/src/assembleHeadMat.cpp:92-97
"""
import math
import os
import re
import sys

import numpy as np

K = 1.0 / (4.0 * math.pi)   # OpenMEEG/include/constants.h:23
DBL_EPS = sys.float_info.epsilon
DBL_MIN = sys.float_info.min


def almost_equal(x, y, eps=1e3):
    """om_common.h:24-26."""
    return abs(x - y) < DBL_EPS * abs(x + y) * eps or abs(x - y) < DBL_MIN


class Mesh:
    def __init__(self, name):
        self.name = name
        self.verts = []          # global vertex ids in mesh order
        self.tris = None         # (T,3) global vertex ids
        self.current_barrier = False
        self.isolated = False
        self.outermost = False


class Interface:
    def __init__(self, name, omeshes):
        self.name = name
        self.omeshes = list(omeshes)   # [(mesh id, +1/-1)]


class Domain:
    def __init__(self, name):
        self.name = name
        self.boundaries = []     # [(Interface, inside: bool)]
        self.cond = -1.0

    def mesh_orientation(self, m):
        """domain.h:95-101 with SimpleDomain::mesh_orientation domain.h:43-45."""
        for iface, inside in self.boundaries:
            for mid, o in iface.omeshes:
                if mid == m:
                    return o if inside else -o
        return 0


def read_tri(path):
    """tri.h:26-56: '- N' then N lines 'x y z nx ny nz', '- T T T' then T lines."""
    with open(path) as f:
        tok = f.read().split()
    assert tok[0] == "-"
    nv = int(tok[1])
    vals = np.array(tok[2:2 + 6 * nv], dtype=np.float64).reshape(nv, 6)
    pos = 2 + 6 * nv
    assert tok[pos] == "-"
    nt = int(tok[pos + 1])
    tris = np.array(tok[pos + 4:pos + 4 + 3 * nt], dtype=np.int64).reshape(nt, 3)
    return vals[:, :3].copy(), tris


class Geometry:
    """Oracle geometry.  Build with from_files() or from_meshes()."""

    def __init__(self):
        self.vertices = []       # list of tuples
        self._vmap = {}
        self.meshes = []
        self.domains = []
        self.parts = []
        self.outer_domain = None
        self.nested = False

    # -- construction ---------------------------------------------------
    def _add_vertex(self, v):
        """geometry.h:100-109: first exactly-equal vertex, else append."""
        key = (float(v[0]), float(v[1]), float(v[2]))
        idx = self._vmap.get(key)
        if idx is None:
            idx = len(self.vertices)
            self.vertices.append(key)
            self._vmap[key] = idx
        return idx

    def _import(self, named_meshes):
        """Geometry::import, geometry.cpp:197-229 (all points first)."""
        indmaps = []
        for name, (pts, tris) in named_meshes:
            indmaps.append([self._add_vertex(p) for p in pts])
        for (name, (pts, tris)), imap in zip(named_meshes, indmaps):
            m = Mesh(name)
            imap = np.asarray(imap, dtype=np.int64)
            m.verts = list(imap)                     # mesh.cpp:76-79
            m.tris = imap[np.asarray(tris, dtype=np.int64)]
            self.meshes.append(m)

    def mesh_index(self, name):
        for i, m in enumerate(self.meshes):
            if m.name == name:
                return i
        raise KeyError(name)

    def _interface_solid_angle(self, iface, p, lib):
        return sum(o * lib.mesh_solid_angle(self, mid, p) for mid, o in iface.omeshes)

    def _fix_orientation(self, iface, lib):
        """Interface::is_mesh_orientations_coherent, interface.cpp:55-95 (the
        bounding-box centre; the random-point retry is not needed for closed
        star-shaped surfaces and is not restated)."""
        pts = np.array([self.vertices[v] for mid, _ in iface.omeshes for v in self.meshes[mid].verts])
        c = 0.5 * (pts.min(axis=0) + pts.max(axis=0))
        sa = self._interface_solid_angle(iface, c, lib)
        if almost_equal(sa, 0.0):
            raise ValueError("bounding-box centre on or outside interface %s" % iface.name)
        if almost_equal(sa, 4 * math.pi):
            iface.omeshes = [(mid, -o) for mid, o in iface.omeshes]

    @classmethod
    def from_files(cls, geom_file, cond_file, lib):
        g = cls()
        d = os.path.dirname(os.path.abspath(geom_file))
        with open(geom_file) as f:
            lines = [ln.strip() for ln in f]
        hdr = lines[0]
        m = re.match(r"# Domain Description (\d+)\.(\d+)", hdr)
        if not m or (m.group(1), m.group(2)) != ("1", "1"):
            raise ValueError("only Domain Description 1.1 is restated")
        body = [ln for ln in lines[1:] if ln and not ln.startswith("#")]
        i = 0
        named = []
        if body[i].startswith("Meshes"):
            nmesh = int(body[i].split()[1]); i += 1
            for k in range(nmesh):
                name, path = _named_entry(body[i], "Mesh", k); i += 1
                named.append((name, read_tri(os.path.join(d, path))))
        assert body[i].startswith("Interfaces")
        nif = int(body[i].split()[1]); i += 1
        iface_lines = []
        if not named:
            for k in range(nif):
                name, path = _named_entry(body[i], "Interface", k); i += 1
                named.append((name, read_tri(os.path.join(d, path))))
        else:
            for k in range(nif):
                iface_lines.append(body[i]); i += 1
        g._import(named)
        interfaces = {}
        if not iface_lines:
            for mi, mesh in enumerate(g.meshes):
                interfaces[mesh.name] = Interface(mesh.name, [(mi, 1)])
        else:
            for k, ln in enumerate(iface_lines):
                head, _, rest = ln.partition(":")
                parts = head.split()
                name = parts[1] if len(parts) > 1 else str(k + 1)
                toks = rest.split()
                oms = []
                for t in toks:
                    sgn = -1 if t.startswith("-") else 1
                    oms.append((g.mesh_index(t.lstrip("+-")), sgn))
                interfaces[name] = Interface(name, oms)
        for iface in interfaces.values():
            g._fix_orientation(iface, lib)
        assert body[i].startswith("Domains")
        ndom = int(body[i].split()[1]); i += 1
        for k in range(ndom):
            ln = body[i]; i += 1
            head, _, rest = ln.partition(":")
            dom = Domain(head.split()[1])
            for t in rest.split():
                inside = t.startswith("-")          # GeomFile.h:106-113,281
                dom.boundaries.append((interfaces[t.lstrip("+-")], inside))
            g.domains.append(dom)
        conds = {}
        with open(cond_file) as f:
            for ln in f:
                ln = ln.strip()
                if not ln or ln.startswith("#"):
                    continue
                a, b = ln.split()[:2]
                conds[a] = float(b)
        for dom in g.domains:
            dom.cond = conds[dom.name]
        g.finalize(lib)
        return g

    @classmethod
    def nested(cls, meshes, conductivity, lib):
        """make_nested_geometry, wrapping/python/openmeeg/_make_geometry.py:131-211
        (meshes inner to outer, 1 or 3 layers), generalised to L layers with
        domains named Brain, Skull, ..., Scalp/Air as in the 3-layer case."""
        L = len(meshes)
        names = ["Cortex", "Skull", "Head"] if L == 3 else ["Cortex"] if L == 1 else \
            ["Layer%d" % k for k in range(L)]
        g = cls()
        g._import(list(zip(names, meshes)))
        ifaces = [Interface(n, [(k, 1)]) for k, n in enumerate(names)]
        # make_geometry does not reorient interfaces (_make_geometry.py:97-121)
        doms = []
        if L == 3:
            spec = [("Scalp", [(1, False), (2, True)], conductivity[2]),
                    ("Brain", [(0, True)], conductivity[0]),
                    ("Air", [(2, False)], 0.0),
                    ("Skull", [(0, False), (1, True)], conductivity[1])]
        elif L == 1:
            spec = [("Brain", [(0, True)], conductivity[0]), ("Air", [(0, False)], 0.0)]
        else:
            spec = [("Brain", [(0, True)], conductivity[0])]
            for k in range(1, L):
                spec.append(("Layer%d" % k, [(k - 1, False), (k, True)], conductivity[k]))
            spec.append(("Air", [(L - 1, False)], 0.0))
        for name, bnds, c in spec:
            dom = Domain(name)
            dom.cond = float(c)
            dom.boundaries = [(ifaces[k], inside) for k, inside in bnds]
            doms.append(dom)
        g.domains = doms
        g.finalize(lib)
        return g

    # -- finalize: geometry.h:228-249 -------------------------------------
    def domains_of(self, m):
        return [k for k, d in enumerate(self.domains) if d.mesh_orientation(m) != 0]

    def common_domains(self, m1, m2):
        s2 = set(self.domains_of(m2))
        return [k for k in self.domains_of(m1) if k in s2]

    def sigma(self, m1, m2):
        return sum(self.domains[k].cond for k in self.common_domains(m1, m2))

    def sigma_inv(self, m1, m2):
        s = 0.0
        for k in self.common_domains(m1, m2):
            c = self.domains[k].cond
            s += (1.0 / c) if c != 0.0 else math.inf
        return s

    def indicator(self, m1, m2):
        return float(len(self.common_domains(m1, m2)))

    def relative_orientation(self, m1, m2):
        if m1 == m2:
            return 1
        doms = self.common_domains(m1, m2)
        if not doms:
            return 0
        d = self.domains[doms[0]]
        return 1 if d.mesh_orientation(m1) == d.mesh_orientation(m2) else -1

    def finalize(self, lib):
        if all(d.cond != -1.0 for d in self.domains):
            self._mark_current_barriers()
        if self.domains:
            for k, d in enumerate(self.domains):           # geometry.cpp:22-41
                if all(not inside for _, inside in d.boundaries):
                    self.outer_domain = k
                    break
            else:
                raise ValueError("no outermost domain")
            for iface, _ in self.domains[self.outer_domain].boundaries:
                for mid, _ in iface.omeshes:
                    self.meshes[mid].outermost = True
            self._check_nested()
        self._generate_indices()
        self._make_pairs()

    def _mark_current_barriers(self):
        invalid = set()
        for d in self.domains:
            if almost_equal(d.cond, 0.0):
                for iface, _ in d.boundaries:
                    for mid, _ in iface.omeshes:
                        m = self.meshes[mid]
                        fully = m.current_barrier
                        m.current_barrier = True
                        if fully:
                            m.isolated = True
                            m.outermost = False
                            invalid.update(m.verts)
        for d in self.domains:
            if almost_equal(d.cond, 0.0):
                for iface, inside in d.boundaries:
                    if not inside:
                        for mid, _ in iface.omeshes:
                            m = self.meshes[mid]
                            if m.current_barrier and not m.isolated:
                                m.outermost = True
        shared = set()
        for v in invalid:
            for m in self.meshes:
                if not m.isolated and v in m.verts:
                    shared.add(v)
        self.invalid = invalid - shared
        cand = list(range(len(self.meshes)))
        parts = []
        while cand:
            c = cand.pop(0)
            conn = [c]
            pending = [c]
            while pending:
                cur = pending.pop(0)
                rest = []
                for o in cand:
                    if self.common_domains(cur, o):
                        conn.append(o)
                        pending.append(o)
                    else:
                        rest.append(o)
                cand = rest
            if len(conn) > 1 or not self.meshes[conn[0]].isolated:
                parts.append(conn)
        self.parts = parts

    def _check_nested(self):
        self.nested = True
        for k, d in enumerate(self.domains):
            if k != self.outer_domain:
                if sum(1 for _, inside in d.boundaries if inside) >= 2:
                    self.nested = False
                    return

    def _generate_indices(self):
        invalid = getattr(self, "invalid", set())
        nv = len(self.vertices)
        self.vert_index = np.full(nv, -1, dtype=np.int64)
        idx = 0
        for v in range(nv):
            if v not in invalid:
                self.vert_index[v] = idx
                idx += 1
        self.nv_unknowns = idx
        tri_index = []
        for m in self.meshes:
            if not m.isolated and not m.current_barrier:
                m.tri_index = np.arange(idx, idx + len(m.tris), dtype=np.int64)
                idx += len(m.tris)
            else:
                m.tri_index = None
        self.nb_current_barrier_triangles = 0
        for m in self.meshes:
            if m.current_barrier:
                if not m.isolated:
                    self.nb_current_barrier_triangles += len(m.tris)
                    m.tri_index = np.arange(idx, idx + len(m.tris), dtype=np.int64)
                    idx += len(m.tris)
                else:
                    m.tri_index = np.full(len(m.tris), -1, dtype=np.int64)
        self.nb_parameters = idx
        self.size = idx - self.nb_current_barrier_triangles

    def _make_pairs(self):
        self.pairs = []
        for i, m1 in enumerate(self.meshes):
            if m1.isolated:
                continue
            for j in range(i + 1):
                o = self.relative_orientation(i, j)
                if not self.meshes[j].isolated and self.sigma(i, j) != 0.0 and o != 0:
                    f = o * K
                    self.pairs.append((i, j, o, f * self.sigma_inv(i, j),
                                       f * self.sigma(i, j), -f * self.indicator(i, j)))

    # -- point location: Domain::contains domain.cpp:5-10, Interface::contains
    #    interface.cpp:18-35 ------------------------------------------------
    def interface_contains(self, iface, p, lib):
        sa = self._interface_solid_angle(iface, p, lib)
        if almost_equal(sa, -4 * math.pi):
            return True
        if almost_equal(sa, 0.0):
            return False
        if almost_equal(sa, 4 * math.pi):
            return True
        return abs(sa) > 2 * math.pi

    def domain_of_point(self, p, lib):
        """Geometry::domain(const Vect3&), geometry.cpp:146-154."""
        for k, d in enumerate(self.domains):
            if all(self.interface_contains(iface, p, lib) == inside for iface, inside in d.boundaries):
                return k
        raise ValueError("point in no domain")

    def domain_index(self, name):
        for k, d in enumerate(self.domains):
            if d.name == name:
                return k
        raise KeyError(name)


def _named_entry(line, keyword, k):
    """section_name + filename of GeomFile.h:86-130."""
    rest = line[len(keyword):]
    if rest.startswith(":"):
        name = str(k + 1)
        path = rest[1:].strip()
    else:
        name, _, path = rest.strip().partition(":")
        name = name.strip()
        path = path.strip()
    return name, path.strip('"')
