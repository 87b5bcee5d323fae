"""Pins the CPU oracle (oracle/oracle.c + oracle/scf.py) before it is trusted
as the GPU checker: reference golden energies, literature integrals, closed
forms, an independent McMurchie-Davidson implementation, symmetry and
invariance properties."""
import math

import mpmath
import numpy as np
import pytest
from scipy.special import gammainc, gamma

from pychem_amd import geometry, structures
from oracle import cint, md, scf

# ---------------------------------------------------------------- Boys ----


def test_boys_against_mpmath_and_gammainc():
    mpmath.mp.dps = 30
    T = np.array([0.0, 1e-12, 1e-3, 0.5, 1.0, 3.7, 10.0, 17.3, 29.9, 34.99, 35.0, 35.01, 50.0, 120.0])
    F = cint.boys(T, 8)
    for i, t in enumerate(T):
        for m in range(9):
            ref = float(mpmath.hyp1f1(m + 0.5, m + 1.5, -t) / (2 * m + 1))
            assert abs(F[i, m] - ref) <= 1e-14 * abs(ref), (t, m)
            if t > 0:
                g = gamma(m + 0.5) * gammainc(m + 0.5, t) / (2 * t ** (m + 0.5))
                assert abs(F[i, m] - g) <= 1e-12 * abs(g)
    # closed forms: F_m(0) = 1/(2m+1); F_0(T) = sqrt(pi/T)/2 erf(sqrt T)
    assert np.allclose(F[0], [1.0 / (2 * m + 1) for m in range(9)], rtol=0, atol=1e-16)
    assert abs(F[4, 0] - 0.5 * math.sqrt(math.pi) * math.erf(1.0)) < 1e-15


def test_boys_downward_recursion_identity():
    T = np.linspace(0.01, 60, 97)
    F = cint.boys(T, 8)
    for m in range(8):
        assert np.allclose((2 * T * F[:, m + 1] + np.exp(-T)) / (2 * m + 1), F[:, m], rtol=1e-13, atol=0)


# ------------------------------------------------------ golden energies ----

GOLDENS = [  # Tests/example5.out:18, example4.out:34, example6.out:39, example7.out:44
    ([["H", 1.0, 0, 0, 0], ["H", 1.0, 1.398397232, 0, 0]], "STO-3G", "BOHR", -1.11675932391, 3e-8),
    ([["H", 1.0, 0, 0, 0], ["H", 1.0, 1.0, 0, 0]], "6-311G", "ANGSTROM", -1.09686463292, 3e-7),
    ([["Li", 3.0, 0, 0, 0], ["H", 1.0, 3.013923978561, 0, 0]], "STO-3G", "BOHR", -7.86202531689, 3e-6),
    ([["B", 5.0, 0, 0, 0], ["H", 1.0, 3.013923978561, 0, 0]], "STO-3G", "BOHR", -24.7030877843, 1e-7),
]


@pytest.mark.parametrize("rows,basis,units,ref,tol", GOLDENS)
def test_reference_golden_energies(rows, basis, units, ref, tol):
    m = structures.Molecule(rows, basis=basis, units=units)
    r = scf.run_rhf(m, conv=1e-7, maxiter=50)
    assert abs(r["energy"] - ref) < tol


def test_h2_sto3g_orbitals_and_density_golden():
    """Tests/example5.out:22,25-26,29-30."""
    m = structures.Molecule([["H", 1.0, 0, 0, 0], ["H", 1.0, 1.398397232, 0, 0]], basis="STO-3G", units="BOHR")
    r = scf.run_rhf(m, conv=1e-10)
    assert np.allclose(r["ea"], [-0.57855389, 0.67114356], atol=2e-8)
    assert np.allclose(np.abs(r["Ca"]), [[0.54884227, 1.21245199], [0.54884227, 1.21245199]], atol=2e-8)
    assert np.allclose(r["Da"], 0.30122784, atol=2e-8)


def test_h2_szabo_ostlund_integrals():
    """Szabo & Ostlund sec. 3.5.2: H2/STO-3G at R = 1.4 bohr."""
    m = structures.Molecule([["H", 1.0, 0, 0, 0], ["H", 1.0, 1.4, 0, 0]], basis="STO-3G", units="BOHR")
    eri, _ = cint.eri_dense(m)
    S, T, V = cint.one_electron(m)
    assert abs(eri[0, 0, 0, 0] - 0.7746) < 5e-5
    assert abs(eri[0, 0, 1, 1] - 0.5697) < 5e-5
    assert abs(eri[0, 1, 0, 1] - 0.2970) < 5e-5
    assert abs(eri[0, 0, 0, 1] - 0.4441) < 5e-5
    assert abs(S[0, 1] - 0.6593) < 5e-5
    assert abs((T + V)[0, 0] - (-1.1204)) < 5e-5
    assert abs((T + V)[0, 1] - (-0.9584)) < 5e-5


# -------------------------------------------------- closed forms, MD -------


def test_ss_closed_form():
    """Primitive (ss|ss) = 2 pi^2.5/(pq sqrt(p+q)) K_AB K_CD F0(T)."""
    rng = np.random.default_rng(5)
    rows = [["H", 1.0] + list(rng.normal(size=3)) for _ in range(2)]
    m = structures.Molecule(rows, basis="6-311G")
    sh = [s for s in m.Shells if s.Cgtf.NPrimitives == 1]
    a, b = sh[0], sh[1]
    blk = cint.eri_quartet(m, a.Index, b.Index, a.Index, b.Index)[0, 0, 0, 0]
    al, be = a.Cgtf.Exponents[0], b.Cgtf.Exponents[0]
    A, B = a.Coords, b.Coords
    p = al + be
    K = math.exp(-al * be / p * np.dot(A - B, A - B))
    n = (2 * al / math.pi) ** 0.75 * (2 * be / math.pi) ** 0.75
    ref = 2 * math.pi ** 2.5 / (p * p * math.sqrt(2 * p)) * K * K * n * n
    assert abs(blk - ref) < 1e-14 * abs(ref)


def _tup(s):
    return (s.Cgtf.AngularMomentum, s.Coords, s.Cgtf.Exponents, s.Cgtf.Coefficients)


@pytest.mark.parametrize("seed", [0, 1])
def test_eri_vs_mcmurchie_davidson_all_classes(seed):
    """Independent algorithm on s..d quartets, including (dd|dd)."""
    rng = np.random.default_rng(seed)
    rows = [["C", 6.0] + list(rng.normal(scale=0.8, size=3)), ["O", 8.0] + list(rng.normal(scale=0.8, size=3))]
    m = structures.Molecule(rows, basis="6-31G*", cartesian_l=[2])
    shells = m.Shells
    by_l = {l: [s.Index for s in shells if s.Cgtf.AngularMomentum == l and s.Cgtf.NPrimitives <= 3] for l in (0, 1, 2)}
    combos = [(2, 2, 2, 2), (2, 1, 2, 1), (1, 2, 0, 2), (0, 0, 2, 1), (1, 1, 1, 1), (2, 0, 0, 0), (0, 1, 0, 1)]
    for ls in combos:
        q = tuple(int(rng.choice(by_l[l])) for l in ls)
        ref = md.contracted_eri([_tup(shells[i]) for i in q])
        mine = cint.eri_quartet(m, *q)
        assert np.abs(ref - mine).max() < 1e-12, (ls, q)


def test_one_electron_vs_mcmurchie_davidson():
    m = structures.Molecule([["C", 6.0, 0, 0, 0], ["H", 1.0, 0.3, 0.9, -0.4], ["O", 8.0, -0.7, 0.2, 0.5]],
                            basis="6-31G*", cartesian_l=[2], multiplicity=2)
    S, T, V = cint.one_electron(m)
    sh = m.Shells
    for a, b in [(5, 5), (5, 13), (2, 5), (13, 12), (4, 5), (0, 13)]:
        ia, ib = sh[a].Ivec, sh[b].Ivec
        s_ref = md.contracted_overlap(_tup(sh[a]), _tup(sh[b]))
        v_ref = sum(md.contracted_nuclear(_tup(sh[a]), _tup(sh[b]), at.Coordinates, at.NuclearCharge)
                    for at in m.Atoms)
        assert np.abs(S[np.ix_(ia, ib)] - s_ref).max() < 1e-13
        assert np.abs(V[np.ix_(ia, ib)] - v_ref).max() < 1e-12


def test_kinetic_by_laplacian_of_overlaps():
    """T_ab = -1/2 <a|lap|b> assembled from MD overlaps of shifted Cartesians."""
    m = structures.Molecule([["C", 6.0, 0, 0, 0], ["O", 8.0, 0.5, -0.9, 0.7]], basis="6-31G*", cartesian_l=[2])
    S, T, V = cint.one_electron(m)
    for a, b in [(5, 11), (4, 2), (11, 11), (0, 10), (2, 5)]:
        sa, sb = m.Shells[a], m.Shells[b]
        la, lb = sa.Cgtf.AngularMomentum, sb.Cgtf.AngularMomentum
        ca, cb = md.cart_comps(la), md.cart_comps(lb)
        ref = np.zeros((len(ca), len(cb)))
        for i, x in enumerate(ca):
            for j, y in enumerate(cb):
                tot = 0.0
                for ea, wa in zip(sa.Cgtf.Exponents, sa.Cgtf.Coefficients):
                    for eb, wb in zip(sb.Cgtf.Exponents, sb.Cgtf.Coefficients):
                        p = ea + eb
                        def s1d(k, i1, j1):
                            if i1 < 0 or j1 < 0:
                                return 0.0
                            return md.E(i1, j1, 0, sa.Coords[k] - sb.Coords[k], ea, eb) * math.sqrt(math.pi / p)
                        lap = 0.0
                        for k in range(3):
                            other = 1.0
                            for kk in range(3):
                                if kk != k:
                                    other *= s1d(kk, x[kk], y[kk])
                            l = y[k]
                            d2 = l * (l - 1) * s1d(k, x[k], l - 2) - 2 * eb * (2 * l + 1) * s1d(k, x[k], l) \
                                + 4 * eb * eb * s1d(k, x[k], l + 2)
                            lap += other * d2
                        tot += wa * wb * md.prim_norm(ea, x) * md.prim_norm(eb, y) * (-0.5) * lap
                ref[i, j] = tot
        assert np.abs(T[np.ix_(sa.Ivec, sb.Ivec)] - ref).max() < 1e-12


def test_normalisation_conventions():
    """Self-overlap of one normalised primitive is 1; Cartesian d components
    are normalised separately: <xx|xx> = 1, <xx|yy> = 1/3."""
    m = structures.Molecule([["C", 6.0, 0, 0, 0]], basis="6-31G*", cartesian_l=[2])
    S, _, _ = cint.one_electron(m)
    d = [s for s in m.Shells if s.Cgtf.AngularMomentum == 2][0]
    i = d.Ivec
    assert abs(S[i[0], i[0]] - 1.0) < 1e-14 and abs(S[i[3], i[3]] - 1.0) < 1e-14
    assert abs(S[i[0], i[3]] - 1.0 / 3.0) < 1e-14
    m2 = structures.Molecule([["C", 6.0, 0, 0, 0]], basis="6-31G*", cartesian_l=[])
    S2, _, _ = cint.one_electron(m2)
    d2 = [s for s in m2.Shells if s.Cgtf.AngularMomentum == 2][0]
    blk = S2[np.ix_(d2.Ivec, d2.Ivec)]
    assert np.allclose(np.diag(blk), 1.0, atol=1e-14)


def test_eri_eightfold_symmetry_and_schwarz():
    m = structures.Molecule(geometry.small_molecule("CH3OH", 0.05, 11), basis="6-31G*", cartesian_l=[2])
    eri, _ = cint.eri_dense(m, 0.0)
    for perm in ["bacd", "abdc", "cdab", "dcba"]:
        assert np.abs(eri - np.einsum("abcd->" + perm, eri)).max() < 1e-14
    diag = np.einsum("abab->ab", eri)
    bound = np.sqrt(np.abs(diag))
    assert np.all(np.abs(eri) <= np.einsum("ab,cd->abcd", bound, bound) + 1e-13)


def test_rotation_invariance_of_energy():
    rows = geometry.small_molecule("CH3OH", 0.0)
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    rot = [[r[0], r[1]] + list(R @ np.array(r[2:])) for r in rows]
    e1 = scf.run_rhf(structures.Molecule(rows, basis="6-31G*", cartesian_l=[2]), diis=scf.PULAY)["energy"]
    e2 = scf.run_rhf(structures.Molecule(rot, basis="6-31G*", cartesian_l=[2]), diis=scf.PULAY)["energy"]
    assert abs(e1 - e2) < 1e-8


def test_jk_against_numpy_einsum():
    m = structures.Molecule(geometry.water(), basis="6-31G*", cartesian_l=[2])
    eri, _ = cint.eri_dense(m)
    rng = np.random.default_rng(3)
    A = rng.normal(size=(m.NOrbitals,) * 2)
    D = A @ A.T
    J, K = cint.jk(eri, 2 * D, D)
    assert np.allclose(J, np.einsum("cd,abcd->ab", 2 * D, eri), atol=1e-11)
    assert np.allclose(K, np.einsum("cb,abcd->ad", -D, eri), atol=1e-11)


# ------------------------------------------------------------------ MP2 ----


def test_mp2_closed_shell_equals_reference_spin_orbital_form():
    m = structures.Molecule(geometry.water(), basis="STO-3G")
    r = scf.run_rhf(m)
    e_cs, _ = scf.mp2_closed_shell(r["eri"], r["Ca"], r["ea"], 5)
    e_so = scf.mp2_spin_orbital(r["eri"], r["Ca"], r["Cb"], r["ea"], r["eb"], 5, 5)
    assert abs(e_cs - e_so) < 1e-12
    assert -0.04 < e_cs < -0.03


def test_mp2_invariant_to_occupied_rotations():
    m = structures.Molecule(geometry.water(), basis="STO-3G")
    r = scf.run_rhf(m)
    e0, _ = scf.mp2_closed_shell(r["eri"], r["Ca"], r["ea"], 5)
    # canonical orbitals are required by the energy denominator; a rotation
    # among *degenerate-free* orbitals changes it -- but the transform itself
    # is invariant: (ia|jb) rotates as a tensor
    _, ov = scf.mp2_closed_shell(r["eri"], r["Ca"], r["ea"], 5)
    Q, _ = np.linalg.qr(np.random.default_rng(0).normal(size=(5, 5)))
    C2 = r["Ca"].copy()
    C2[:, :5] = C2[:, :5] @ Q
    _, ov2 = scf.mp2_closed_shell(r["eri"], C2, r["ea"], 5)
    ov_rot = np.einsum("iajb,ik,jl->kalb", ov, Q, Q)
    assert np.abs(ov2 - ov_rot).max() < 1e-12
    assert e0 < 0
