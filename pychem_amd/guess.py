"""SCF initial guesses: CORE (the reference default, Methods/hartree_fock.py:55-57)
and SAD, superposition of atomic densities.

The reference accepts "SAD" as a guess keyword (Util/structures.py:211) but
never implements it; this is a from-scratch implementation.  For every
distinct (element, basis) an atomic SCF is run in the molecule's own basis on
the HIP engine: spin-restricted, spherically averaged, with the electrons
placed by Aufbau and the last partially filled (near-)degenerate set sharing
its electrons evenly (fractional occupation).  The molecular guess is the
block-diagonal sum of the atomic densities; D_alpha = D_beta = D_total / 2.
"""
import numpy as np

from . import structures

_ATOM_CACHE = {}


def _atomic_density(label, Z, basis, cartesian_l, maxiter=60, conv=1e-8):
    key = (label, basis, tuple(cartesian_l))
    if key in _ATOM_CACHE:
        return _ATOM_CACHE[key]
    from .engine import Engine
    import torch
    mol = structures.Molecule([[label, Z, 0.0, 0.0, 0.0]], basis=basis, cartesian_l=cartesian_l,
                              units="BOHR", multiplicity=(structures.N_ELECTRONS[label] % 2) + 1)
    eng = Engine(mol)
    S, T, V = (x.cpu().numpy() for x in eng.one_electron())
    H = T + V
    eng.build_eri(1e-12)
    n = mol.NOrbitals
    eri = eng.eri_dense().cpu().numpy().reshape(n, n, n, n)
    eng.close()
    s, U = np.linalg.eigh(S)
    keep = s > structures.LINEAR_DEPENDENCE
    X = U[:, keep] * s[keep] ** -0.5
    nel = float(structures.N_ELECTRONS[label])

    def occupations(e):
        occ = np.zeros(len(e))
        left = nel
        i = 0
        while left > 1e-12 and i < len(e):
            j = i
            while j + 1 < len(e) and abs(e[j + 1] - e[i]) < 1e-3:
                j += 1
            g = j - i + 1
            put = min(left, 2.0 * g)
            occ[i:j + 1] = put / g
            left -= put
            i = j + 1
        return occ

    # J_ab = sum_cd (ab|cd) D_cd and K_ad = sum_bc (ab|cd) D_cb as two
    # matrix-vector products over the reshaped tensor
    eri_j = eri.reshape(n * n, n * n)
    eri_k = np.ascontiguousarray(eri.transpose(0, 3, 2, 1)).reshape(n * n, n * n)
    F = H.copy()
    D = np.zeros((n, n))
    E_old = 0.0
    for _ in range(maxiter):
        e, C = np.linalg.eigh(X.T @ F @ X)
        C = X @ C
        occ = occupations(e)
        D_new = (C * occ) @ C.T
        D = D_new if _ == 0 else 0.5 * (D + D_new)
        J = (eri_j @ D.reshape(-1)).reshape(n, n)
        K = (eri_k @ D.reshape(-1)).reshape(n, n)
        F = H + J - 0.5 * K
        E = 0.5 * np.sum(D * (H + F))
        if abs(E - E_old) < conv:
            break
        E_old = E
    _ATOM_CACHE[key] = D
    return D


def sad_density(molecule):
    """Block-diagonal superposition of atomic densities (total density)."""
    n = molecule.NOrbitals
    D = np.zeros((n, n))
    off = 0
    basis = molecule.Basis
    for atom in molecule.Atoms:
        Da = _atomic_density(atom.Label, atom.NuclearCharge, basis, molecule.CartesianL)
        m = Da.shape[0]
        D[off:off + m, off:off + m] = Da
        off += m
    return D


def core_guess(molecule):
    """Density from diagonalising the core Hamiltonian (closed-shell alpha)."""
    e, C = np.linalg.eigh(molecule.Xt @ molecule.Core @ molecule.X)
    C = molecule.X @ C
    na = molecule.NAlphaElectrons
    return C[:, :na] @ C[:, :na].T
