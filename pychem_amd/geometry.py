"""Synthetic molecular geometries for the workloads (coordinates in Angstrom).

* ``water()``           H2O, O at the origin, O-H 0.9572 A, H-O-H 104.52 deg
* ``alkane(n, ...)``    linear C_nH_{2n+2}: ideal zig-zag carbon chain,
  C-C 1.54 A, C-H 1.09 A, tetrahedral angles, optional Gaussian jitter drawn
  from ``numpy.random.default_rng(seed)``
* ``alcohol(n, ...)``   C_nH_{2n+1}OH built on the same chain
* ``small_molecule(name)`` H2O, NH3, CH4, C2H6, CH3OH, C3H8
"""
import math

import numpy as np

TET = math.radians(109.4712206)
CC, CH, CO, OH, NH = 1.54, 1.09, 1.43, 0.96, 1.012


def water():
    th = math.radians(104.52) / 2.0
    r = 0.9572
    return [["O", 8.0, 0.0, 0.0, 0.0],
            ["H", 1.0, r * math.sin(th), 0.0, r * math.cos(th)],
            ["H", 1.0, -r * math.sin(th), 0.0, r * math.cos(th)]]


def _unit(v):
    return v / np.linalg.norm(v)


def _tet_dirs_two(u1, u2):
    """Two remaining tetrahedral directions at an atom bonded along u1, u2."""
    d = _unit(-(u1 + u2))
    n = _unit(np.cross(u1, u2))
    half = TET / 2.0
    return [d * math.cos(half) + n * math.sin(half), d * math.cos(half) - n * math.sin(half)]


def _tet_dirs_three(u1, ref):
    """Three tetrahedral directions at an atom bonded along u1; ``ref`` fixes
    the azimuth (one direction lies in the plane of u1 and ref)."""
    w = ref - np.dot(ref, u1) * u1
    if np.linalg.norm(w) < 1e-8:
        w = np.cross(u1, [0.0, 0.0, 1.0])
        if np.linalg.norm(w) < 1e-8:
            w = np.cross(u1, [0.0, 1.0, 0.0])
    w = _unit(w)
    v = np.cross(u1, w)
    out = []
    for k in range(3):
        phi = 2.0 * math.pi * k / 3.0
        perp = math.cos(phi) * w + math.sin(phi) * v
        out.append(-u1 / 3.0 + math.sqrt(8.0) / 3.0 * perp)
    return out


def _chain(n):
    """Zig-zag heavy-atom chain positions: C_k = (k*d*sin(t/2), (k%2)*d*cos(t/2), 0)."""
    s, c = math.sin(TET / 2.0), math.cos(TET / 2.0)
    return [np.array([k * CC * s, (k % 2) * CC * c, 0.0]) for k in range(n)]


def _jitter(rows, sigma, seed):
    if sigma and sigma > 0.0:
        rng = np.random.default_rng(seed)
        for r in rows:
            d = rng.normal(0.0, sigma, 3)
            r[2] += float(d[0]); r[3] += float(d[1]); r[4] += float(d[2])
    return rows


def alkane(n, sigma=0.0, seed=None):
    """Linear alkane C_nH_{2n+2} (carbons first, then hydrogens)."""
    if n == 1:
        rows = [["C", 6.0, 0.0, 0.0, 0.0]]
        u1 = np.array([0.0, 0.0, 1.0])
        dirs = [u1] + _tet_dirs_three(-u1, np.array([1.0, 0.0, 0.0]))
        for d in dirs:
            p = CH * d
            rows.append(["H", 1.0, float(p[0]), float(p[1]), float(p[2])])
        return _jitter(rows, sigma, seed)
    C = _chain(n)
    rows = [["C", 6.0, float(p[0]), float(p[1]), float(p[2])] for p in C]
    H = []
    for k in range(n):
        nb = [C[j] for j in (k - 1, k + 1) if 0 <= j < n]
        us = [_unit(x - C[k]) for x in nb]
        if len(us) == 2:
            dirs = _tet_dirs_two(us[0], us[1])
        else:
            other = C[k + 1] if k == 0 else C[k - 1]
            nxt = C[k + 2] if (k == 0 and n > 2) else (C[k - 2] if n > 2 else C[k] + np.array([0.0, 1.0, 0.0]))
            ref = -(nxt - other)
            dirs = _tet_dirs_three(us[0], ref if n > 2 else np.array([0.0, 1.0, 0.0]))
        for d in dirs:
            H.append(C[k] + CH * d)
    rows += [["H", 1.0, float(p[0]), float(p[1]), float(p[2])] for p in H]
    return _jitter(rows, sigma, seed)


def alcohol(n, sigma=0.0, seed=None):
    """Primary alcohol C_nH_{2n+1}OH: the chain's last heavy atom is O."""
    C = _chain(n + 1)
    rows = [["C", 6.0, float(p[0]), float(p[1]), float(p[2])] for p in C[:n]]
    O = C[n - 1] + CO * _unit(C[n] - C[n - 1])
    rows.append(["O", 8.0, float(O[0]), float(O[1]), float(O[2])])
    heavy = C[:n] + [O]
    H = []
    for k in range(n):
        nb = [heavy[j] for j in (k - 1, k + 1) if 0 <= j < n + 1]
        us = [_unit(x - heavy[k]) for x in nb]
        if len(us) == 2:
            dirs = _tet_dirs_two(us[0], us[1])
        else:
            dirs = _tet_dirs_three(us[0], np.array([0.0, 1.0, 0.0]))
        for d in dirs:
            H.append(heavy[k] + CH * d)
    u = _unit(heavy[n - 1] - O)
    dirs = _tet_dirs_three(u, np.array([0.0, 1.0, 0.0]) if n % 2 == 0 else np.array([0.0, -1.0, 0.0]))
    H.append(O + OH * dirs[0])
    rows += [["H", 1.0, float(p[0]), float(p[1]), float(p[2])] for p in H]
    return _jitter(rows, sigma, seed)


def ammonia():
    th = math.radians(106.7)
    # pyramidal NH3 with H-N-H = 106.7 deg
    c = math.cos(th)
    s_alpha = math.sqrt((1.0 - c) * 2.0 / 3.0)      # sin of angle from the C3 axis
    ca = math.sqrt(1.0 - s_alpha ** 2)
    rows = [["N", 7.0, 0.0, 0.0, 0.0]]
    for k in range(3):
        phi = 2.0 * math.pi * k / 3.0
        rows.append(["H", 1.0, NH * s_alpha * math.cos(phi), NH * s_alpha * math.sin(phi), -NH * ca])
    return rows


def small_molecule(name, sigma=0.0, seed=None):
    name = name.upper()
    if name == "H2O":
        rows = water()
    elif name == "NH3":
        rows = ammonia()
    elif name == "CH4":
        rows = alkane(1)
    elif name == "C2H6":
        rows = alkane(2)
    elif name == "C3H8":
        rows = alkane(3)
    elif name == "CH3OH":
        rows = alcohol(1)
    else:
        raise KeyError(name)
    return _jitter([list(r) for r in rows], sigma, seed)
