"""Independent McMurchie-Davidson integrals (TEST INFRASTRUCTURE ONLY).

A second algorithm for the same integrals, used to pin the HGP oracle on the
d-shell classes no reference fixture covers.  Hermite expansion coefficients
E_t^{ij} and Hermite Coulomb integrals R_{tuv}; Boys F_n from
scipy.special.hyp1f1 (F_n(T) = 1F1(n+1/2; n+3/2; -T)/(2n+1)).  Plain numpy,
meant for tiny inputs only.
"""
import math
from functools import lru_cache

import numpy as np
from scipy.special import hyp1f1


def boys(n, T):
    return hyp1f1(n + 0.5, n + 1.5, -T) / (2.0 * n + 1.0)


def E(i, j, t, Qx, a, b):
    """Hermite expansion coefficient (recursive definition)."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == j == t == 0:
        return math.exp(-q * Qx * Qx)
    if j == 0:
        return (1.0 / (2 * p)) * E(i - 1, j, t - 1, Qx, a, b) - (q * Qx / a) * E(i - 1, j, t, Qx, a, b) \
            + (t + 1) * E(i - 1, j, t + 1, Qx, a, b)
    return (1.0 / (2 * p)) * E(i, j - 1, t - 1, Qx, a, b) + (q * Qx / b) * E(i, j - 1, t, Qx, a, b) \
        + (t + 1) * E(i, j - 1, t + 1, Qx, a, b)


def R(t, u, v, n, p, PCx, PCy, PCz, RPC):
    T = p * RPC * RPC
    if t == u == v == 0:
        return (-2 * p) ** n * boys(n, T)
    if t == u == 0:
        val = 0.0
        if v > 1:
            val += (v - 1) * R(t, u, v - 2, n + 1, p, PCx, PCy, PCz, RPC)
        val += PCz * R(t, u, v - 1, n + 1, p, PCx, PCy, PCz, RPC)
        return val
    if t == 0:
        val = 0.0
        if u > 1:
            val += (u - 1) * R(t, u - 2, v, n + 1, p, PCx, PCy, PCz, RPC)
        val += PCy * R(t, u - 1, v, n + 1, p, PCx, PCy, PCz, RPC)
        return val
    val = 0.0
    if t > 1:
        val += (t - 1) * R(t - 2, u, v, n + 1, p, PCx, PCy, PCz, RPC)
    val += PCx * R(t - 1, u, v, n + 1, p, PCx, PCy, PCz, RPC)
    return val


def prim_eri(a, lmn1, A, b, lmn2, B, c, lmn3, C, d, lmn4, D):
    """Unnormalised primitive (ab|cd)."""
    l1, m1, n1 = lmn1; l2, m2, n2 = lmn2; l3, m3, n3 = lmn3; l4, m4, n4 = lmn4
    p = a + b
    q = c + d
    alpha = p * q / (p + q)
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    RPQ = np.linalg.norm(P - Q)
    val = 0.0
    for t in range(l1 + l2 + 1):
        et = E(l1, l2, t, A[0] - B[0], a, b)
        for u in range(m1 + m2 + 1):
            eu = E(m1, m2, u, A[1] - B[1], a, b)
            for v in range(n1 + n2 + 1):
                ev = E(n1, n2, v, A[2] - B[2], a, b)
                if et * eu * ev == 0.0:
                    continue
                for tau in range(l3 + l4 + 1):
                    ft = E(l3, l4, tau, C[0] - D[0], c, d)
                    for nu in range(m3 + m4 + 1):
                        fu = E(m3, m4, nu, C[1] - D[1], c, d)
                        for phi in range(n3 + n4 + 1):
                            fv = E(n3, n4, phi, C[2] - D[2], c, d)
                            if ft * fu * fv == 0.0:
                                continue
                            val += et * eu * ev * ft * fu * fv * (-1) ** (tau + nu + phi) * \
                                R(t + tau, u + nu, v + phi, 0, alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], RPQ)
    return val * 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q))


def prim_norm(a, lmn):
    """Normalisation of a primitive Cartesian Gaussian x^l y^m z^n e^{-a r^2}."""
    l, m, n = lmn
    df = lambda k: 1 if k <= 0 else k * df(k - 2)
    return (2 * a / math.pi) ** 0.75 * (4 * a) ** ((l + m + n) / 2.0) / math.sqrt(df(2 * l - 1) * df(2 * m - 1) * df(2 * n - 1))


def cart_comps(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def contracted_eri(shells):
    """shells: 4 tuples (l, center[3], exps, coefs); contraction coefficients
    multiply normalised primitives (no contracted renormalisation).  Returns
    the Cartesian block [nA][nB][nC][nD]."""
    comps = [cart_comps(s[0]) for s in shells]
    out = np.zeros(tuple(len(c) for c in comps))
    for ia, ca in enumerate(comps[0]):
        for ib, cb in enumerate(comps[1]):
            for ic, cc in enumerate(comps[2]):
                for idd, cd in enumerate(comps[3]):
                    tot = 0.0
                    for ea, wa in zip(shells[0][2], shells[0][3]):
                        na = wa * prim_norm(ea, ca)
                        for eb, wb in zip(shells[1][2], shells[1][3]):
                            nb = wb * prim_norm(eb, cb)
                            for ec, wc in zip(shells[2][2], shells[2][3]):
                                nc = wc * prim_norm(ec, cc)
                                for ed, wd in zip(shells[3][2], shells[3][3]):
                                    nd = wd * prim_norm(ed, cd)
                                    tot += na * nb * nc * nd * prim_eri(
                                        ea, ca, np.asarray(shells[0][1]), eb, cb, np.asarray(shells[1][1]),
                                        ec, cc, np.asarray(shells[2][1]), ed, cd, np.asarray(shells[3][1]))
                    out[ia, ib, ic, idd] = tot
    return out


def overlap_1d_E(l1, l2, Qx, a, b):
    return E(l1, l2, 0, Qx, a, b)


def contracted_overlap(sa, sb):
    """Cartesian overlap block via Hermite E_0 (S = E0x E0y E0z (pi/p)^1.5)."""
    ca, cb = cart_comps(sa[0]), cart_comps(sb[0])
    A, B = np.asarray(sa[1]), np.asarray(sb[1])
    out = np.zeros((len(ca), len(cb)))
    for i, x in enumerate(ca):
        for j, y in enumerate(cb):
            tot = 0.0
            for ea, wa in zip(sa[2], sa[3]):
                for eb, wb in zip(sb[2], sb[3]):
                    p = ea + eb
                    s = 1.0
                    for k in range(3):
                        s *= E(x[k], y[k], 0, A[k] - B[k], ea, eb)
                    tot += wa * wb * prim_norm(ea, x) * prim_norm(eb, y) * s * (math.pi / p) ** 1.5
            out[i, j] = tot
    return out


def contracted_nuclear(sa, sb, C, Z):
    """Cartesian nuclear-attraction block -Z <a|1/|r-C||b> via Hermite R."""
    ca, cb = cart_comps(sa[0]), cart_comps(sb[0])
    A, B, C = np.asarray(sa[1]), np.asarray(sb[1]), np.asarray(C, dtype=float)
    out = np.zeros((len(ca), len(cb)))
    for i, x in enumerate(ca):
        for j, y in enumerate(cb):
            tot = 0.0
            for ea, wa in zip(sa[2], sa[3]):
                for eb, wb in zip(sb[2], sb[3]):
                    p = ea + eb
                    P = (ea * A + eb * B) / p
                    RPC = np.linalg.norm(P - C)
                    v = 0.0
                    for t in range(x[0] + y[0] + 1):
                        for u in range(x[1] + y[1] + 1):
                            for w in range(x[2] + y[2] + 1):
                                v += E(x[0], y[0], t, A[0] - B[0], ea, eb) * E(x[1], y[1], u, A[1] - B[1], ea, eb) * \
                                    E(x[2], y[2], w, A[2] - B[2], ea, eb) * \
                                    R(t, u, w, 0, p, P[0] - C[0], P[1] - C[1], P[2] - C[2], RPC)
                    tot += wa * wb * prim_norm(ea, x) * prim_norm(eb, y) * v * 2 * math.pi / p
            out[i, j] = -Z * tot
    return out
