"""Closed-form multilayer-sphere EEG potential (TEST INFRASTRUCTURE ONLY).

Restates the series of Zhang, Phys. Med. Biol. 40 (1995) 335-349, formulas
(1I)/(2I), in the form the reference uses to make its analytic EEG oracles
(apps/tools/matlab/om_spher_pot_iso.m:22-64, om_vangle.m).  The reference's
tests/analytic/eeg_head{1,2,3}_analytic.txt were produced by that function;
matching them pins this restatement, which then serves as an independent
oracle for the synthetic nested-sphere heads.
"""
import math

import numpy as np
from scipy.special import lpmv


def _vangle(r, s):
    mag = np.sqrt(np.einsum("ij,ij->i", r, r) * np.einsum("ij,ij->i", s, s))
    mag = np.maximum(mag, 1e-30)
    return np.arccos(np.clip(np.einsum("ij,ij->i", r, s) / mag, -1.0, 1.0))


def sphere_potential(x, q, p, sigmas, radii, relerr=1e-20, nmax=200):
    """Potential at points x (n,3) on the outer sphere of a layered sphere
    (radii ascending, conductivities inner to outer) due to a current dipole
    of moment q at position p.  om_spher_pot_iso.m:22-64."""
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    q = np.asarray(q, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    n_pts = len(x)
    P = np.tile(p, (n_pts, 1))
    Q = np.tile(q, (n_pts, 1))
    N = len(sigmas)
    rk = np.asarray(radii, dtype=np.float64)
    alpha = _vangle(P, Q)
    cosg = np.cos(_vangle(P, x))
    beta = _vangle(np.cross(P, Q), np.cross(P, x))
    counter = np.zeros(n_pts)
    n = 1
    while True:
        term = (2 * n + 1) / n * (np.linalg.norm(p) / rk[N - 1]) ** (n - 1)
        M = np.eye(2) / (2 * n + 1) ** (N - 1)
        for k in range(N - 1):
            rs = sigmas[k] / sigmas[k + 1]
            re = (rk[N - 1] / rk[k]) ** (2 * n + 1)
            M = M @ np.array([[n + (n + 1) * rs, (n + 1) * (rs - 1) * re],
                              [n * (rs - 1) / re, (n + 1) + n * rs]])
        f = n / (n * M[1, 1] + (1 + n) * M[1, 0])
        P0 = lpmv(0, n, cosg)
        P1 = lpmv(1, n, cosg)
        t = term * f * (n * np.cos(alpha) * P0 - np.cos(beta) * np.sin(alpha) * P1)
        counter += t
        if (np.all(np.abs(t) < np.abs(counter) * relerr) and n > 20) or n > nmax:
            break
        n += 1
    return counter * np.linalg.norm(q) / (4 * math.pi * sigmas[N - 1] * rk[N - 1] ** 2)


def sphere_gain(electrodes, dipoles, sigmas, radii):
    """Analytic gain (n_electrodes x n_dipoles)."""
    G = np.zeros((len(electrodes), len(dipoles)))
    for j, d in enumerate(dipoles):
        G[:, j] = sphere_potential(electrodes, d[3:6], d[0:3], sigmas, radii)
    return G


def rdm(a, b):
    """Relative difference measure of tests/test_compare_matrix.cpp:305-345."""
    a = a - a.mean()
    b = b - b.mean()
    return float(np.linalg.norm(a / np.linalg.norm(a) - b / np.linalg.norm(b)))


def mag(a, b):
    """Magnitude ratio error of tests/test_compare_matrix.cpp:348-386."""
    a = a - a.mean()
    b = b - b.mean()
    return float(abs(1.0 - np.linalg.norm(a) / np.linalg.norm(b)))
