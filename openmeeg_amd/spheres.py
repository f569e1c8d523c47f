"""Synthetic nested-sphere heads (seeded) for the parity tests and the bench.

Surfaces are class-I geodesic spheres: the 20 faces of an icosahedron are
split into f*f triangles and projected on the sphere, giving 10 f^2 + 2
vertices and 20 f^2 triangles (f = 8: 642, f = 16: 2562, f = 32: 10242).
Triangles are wound like the reference's data/Head* meshes: the mesh solid
angle seen from an inside point is -4 pi (OpenMEEG/src/interface.cpp:55-95
treats that as the coherent orientation).

The configurations:

* ``c0``: 3 layers, f=8 (642 V), radii 0.87/0.92/1.0, sigma 1/0.0125/1,
  100 dipoles uniform in r<0.8 with unit random moments, 64 electrodes
  uniform on the outer sphere.
* ``c1``: as c0 with f=16 (2562 V): n = 17926 unknowns, 2000 dipoles on a
  jittered grid in r<0.8, 128 electrodes.
* ``c2``: 4 layers, f=32 (10242 V), radii 0.80/0.84/0.90/1.0, sigma
  0.33/1.79/0.0042/0.33, n = 102408 unknowns, 5000 dipoles in r<=0.72,
  256 electrodes.
* ``c3``: c1 mesh, 4096 dipoles, 128 electrodes.
* ``batched``: heads with radially perturbed surfaces (real spherical
  harmonics of degree 2..4, 3% amplitude) and uneven sizes.
"""
import math

import numpy as np

_PHI = (1.0 + math.sqrt(5.0)) / 2.0


def _icosahedron():
    v = np.array([[-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0],
                  [0, -1, _PHI], [0, 1, _PHI], [0, -1, -_PHI], [0, 1, -_PHI],
                  [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def geodesic_sphere(freq):
    """Unit geodesic sphere of frequency `freq` (10 f^2 + 2 vertices)."""
    V0, F0 = _icosahedron()
    verts = {}
    pts = []

    def key(a, b, c, i, j):
        # barycentric lattice point (i, j, f-i-j) on face (a, b, c), made
        # canonical over shared edges/corners by the sorted (vertex, weight)
        w = {a: freq - i - j, b: i, c: j}
        return tuple(sorted((v, x) for v, x in w.items() if x))

    def vid(a, b, c, i, j):
        k = key(a, b, c, i, j)
        idx = verts.get(k)
        if idx is None:
            p = ((freq - i - j) * V0[a] + i * V0[b] + j * V0[c]) / freq
            idx = len(pts)
            pts.append(p / np.linalg.norm(p))
            verts[k] = idx
        return idx

    tris = []
    for a, b, c in F0:
        for i in range(freq):
            for j in range(freq - i):
                p0 = vid(a, b, c, i, j)
                p1 = vid(a, b, c, i + 1, j)
                p2 = vid(a, b, c, i, j + 1)
                tris.append((p0, p1, p2))
                if i + j < freq - 1:
                    p3 = vid(a, b, c, i + 1, j + 1)
                    tris.append((p1, p3, p2))
    P = np.array(pts)
    T = np.array(tris, dtype=np.int64)
    # wind so that the geometric normal (p1-p0)x(p2-p0) points inward
    c = P[T].mean(axis=1)
    nrm = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    flip = np.einsum("ij,ij->i", nrm, c) > 0
    T[flip] = T[flip][:, [0, 2, 1]]
    return P, T


def _real_sph_harm(l, m, theta, phi):
    from scipy.special import lpmv
    mm = abs(m)
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - mm) / math.factorial(l + mm))
    P = lpmv(mm, l, np.cos(theta))
    if m > 0:
        return math.sqrt(2) * norm * P * np.cos(mm * phi)
    if m < 0:
        return math.sqrt(2) * norm * P * np.sin(mm * phi)
    return norm * P


def perturbation(unit_pts, rng, amplitude=0.03, lmin=2, lmax=4):
    """Radial factor 1 + amplitude * s(x) with s a random real spherical
    harmonic field of degrees lmin..lmax scaled to max |s| = 1."""
    x, y, z = unit_pts.T
    theta = np.arccos(np.clip(z, -1, 1))
    phi = np.arctan2(y, x)
    coeffs = {(l, m): rng.standard_normal() for l in range(lmin, lmax + 1) for m in range(-l, l + 1)}

    def field(th, ph):
        s = np.zeros_like(th)
        for (l, m), c in coeffs.items():
            s += c * _real_sph_harm(l, m, th, ph)
        return s

    # scale on a fine reference grid so every surface of a head shares it
    g_th = np.arccos(np.linspace(-1, 1, 181))
    g_ph = np.linspace(-math.pi, math.pi, 361)
    TH, PH = np.meshgrid(g_th, g_ph)
    smax = np.abs(field(TH.ravel(), PH.ravel())).max()
    return 1.0 + amplitude * field(theta, phi) / smax


def nested_sphere_meshes(freq, radii, rng=None, amplitude=0.0):
    """List of (vertices, triangles), inner to outer."""
    P, T = geodesic_sphere(freq)
    fac = perturbation(P, rng, amplitude) if amplitude > 0 else np.ones(len(P))
    return [(P * (r * fac)[:, None], T.copy()) for r in radii]


def random_dipoles(rng, n, rmax):
    """n dipoles uniform in the ball r < rmax with unit random moments."""
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rmax * rng.random(n) ** (1.0 / 3.0)
    q = rng.standard_normal((n, 3))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return np.hstack([d * r[:, None], q])


def grid_dipoles(rng, n, rmax):
    """n dipoles on a jittered cubic grid inside r < rmax (unit moments)."""
    side = 2
    while True:
        g = (np.arange(side) + 0.5) / side * 2 - 1
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) * rmax
        inside = np.linalg.norm(pts, axis=1) < rmax * (1 - 1.0 / side)
        if inside.sum() >= n:
            break
        side += 1
    pts = pts[inside]
    pts = pts[rng.permutation(len(pts))[:n]]
    h = 2 * rmax / side
    jit = rng.uniform(-0.25 * h, 0.25 * h, size=pts.shape)
    pts = pts + jit
    r = np.linalg.norm(pts, axis=1)
    pts[r >= rmax] *= (0.999 * rmax / r[r >= rmax])[:, None]
    q = rng.standard_normal((n, 3))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return np.hstack([pts, q])


def sphere_electrodes(rng, n, radius=1.0):
    e = rng.standard_normal((n, 3))
    return radius * e / np.linalg.norm(e, axis=1, keepdims=True)


def sphere_squids(rng, n, radius=1.15, tilt=0.5):
    """n MEG magnetometer points on a sphere outside the head: positions,
    unit orientations (radial plus `tilt` times a random tangential
    direction: a purely radial field would not see the volume currents of a
    spherical head), weights 1."""
    p = sphere_electrodes(rng, n, radius)
    r = p / np.linalg.norm(p, axis=1, keepdims=True)
    t = rng.standard_normal((n, 3))
    t -= (t * r).sum(axis=1, keepdims=True) * r
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    o = r + tilt * t
    return p, o / np.linalg.norm(o, axis=1, keepdims=True), np.ones(n)


CONFIGS = {
    "c0": dict(freq=8, radii=(0.87, 0.92, 1.0), cond=(1.0, 0.0125, 1.0), n_dip=100,
               dip="uniform", rmax=0.8, n_elec=64, seed=0),
    "c1": dict(freq=16, radii=(0.87, 0.92, 1.0), cond=(1.0, 0.0125, 1.0), n_dip=2000,
               dip="grid", rmax=0.8, n_elec=128, seed=1),
    "c2": dict(freq=32, radii=(0.80, 0.84, 0.90, 1.0), cond=(0.33, 1.79, 0.0042, 0.33),
               n_dip=5000, dip="uniform", rmax=0.72, n_elec=256, seed=2),
    "c3": dict(freq=16, radii=(0.87, 0.92, 1.0), cond=(1.0, 0.0125, 1.0), n_dip=4096,
               dip="uniform", rmax=0.8, n_elec=128, seed=3),
}


def make_config(name, **over):
    """Return (meshes, conductivities, dipoles (n,6), electrodes (n,3))."""
    c = dict(CONFIGS[name])
    c.update(over)
    rng = np.random.default_rng(c["seed"])
    meshes = nested_sphere_meshes(c["freq"], c["radii"])
    if c["dip"] == "grid":
        dip = grid_dipoles(rng, c["n_dip"], c["rmax"])
    else:
        dip = random_dipoles(rng, c["n_dip"], c["rmax"])
    el = sphere_electrodes(rng, c["n_elec"], c["radii"][-1])
    return meshes, tuple(c["cond"]), dip, el


def noisy_targets(G, rng, snr=10.0):
    """Measurement targets G + noise with sigma_noise = rms(G) / snr."""
    G = np.asarray(G, dtype=np.float64)
    return G + rng.standard_normal(G.shape) * (np.sqrt(np.mean(G * G)) / snr)


def make_batched_heads(n_heads=32, seed=4, freqs=(8, 10, 12, 14, 16)):
    """Uneven perturbed 3-layer heads: list of (meshes, cond, dipoles, electrodes)."""
    rng = np.random.default_rng(seed)
    heads = []
    for h in range(n_heads):
        f = int(freqs[rng.integers(len(freqs))])
        meshes = nested_sphere_meshes(f, (0.87, 0.92, 1.0), rng=rng, amplitude=0.03)
        n_dip = int(rng.integers(200, 2001))
        n_el = int(rng.integers(64, 257))
        # dipoles well inside the perturbed cortex (min radius >= 0.87*0.97)
        dip = random_dipoles(rng, n_dip, 0.78)
        # electrodes on the unit sphere; Head2EEGMat projects each onto the
        # closest point of the perturbed scalp
        e = sphere_electrodes(rng, n_el, 1.0)
        heads.append((meshes, (1.0, 0.0125, 1.0), dip, e))
    return heads
