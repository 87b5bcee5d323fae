"""Hartree-Fock SCF on the HIP engine, with the reference's call surface
(Methods/hartree_fock.py).

Integrals, Schwarz screening, ERI storage and every J/K build run on the GPU
(``Engine``); the FP64 Fock diagonalisation, DIIS and energy sums run on the
host in numpy, as in the reference.  RHF is the UHF code path with equal
alpha/beta occupations, exactly as the reference codes it
(Methods/hartree_fock.py:23-154), so the same loop serves UHF.

Functions keep the reference names: ``make_core_matrices``,
``evaluate_2e_ints``, ``make_coulomb_exchange_matrices``,
``make_fock_matrices``, ``make_MOs``, ``make_density_matrices``,
``calculate_energy``, ``do``.
"""
import time

import numpy as np
import torch

from . import structures
from .diis import DIIS
from .engine import Engine


class Matrices:
    """Util/structures.py:882-901."""

    def __init__(self, n, nocc):
        self.NOcc = nocc
        self.Density = np.zeros((n, n))
        self.Fock = np.zeros((n, n))
        self.Exchange = np.zeros((n, n))
        self.MOs = None
        self.Energies = None

    def sort_occupied(self):
        """Sort the occupied orbitals by energy (Util/structures.py:895-901)."""
        n = self.NOcc
        idx = np.append(np.argsort(self.Energies[:n]), np.arange(n, len(self.Energies)))
        self.Energies = self.Energies[idx]
        self.MOs = self.MOs[:, idx]


class ElectronicState:
    """Ground state with nalpha/nbeta electrons (Util/structures.py:862-876)."""

    def __init__(self, n, nalpha, nbeta):
        self.NAlpha, self.NBeta = nalpha, nbeta
        self.Alpha = Matrices(n, nalpha)
        self.Beta = Matrices(n, nbeta)
        self.Coulomb = np.zeros((n, n))
        self.Energy = None
        self.TotalEnergy = None
        self.AlphaDIIS = DIIS()
        self.BetaDIIS = DIIS()


def nuclear_repulsion(molecule):
    """sum_{i>j} Z_i Z_j / r_ij (Methods/integrals.py:196-207)."""
    e = 0.0
    for i, a in enumerate(molecule.Atoms):
        for j in range(i):
            b = molecule.Atoms[j]
            r = np.linalg.norm(np.array(a.Coordinates) - np.array(b.Coordinates))
            e += a.NuclearCharge * b.NuclearCharge / r
    return e


def get_engine(molecule):
    eng = getattr(molecule, "Engine", None)
    if eng is None:
        eng = Engine(molecule)
        molecule.Engine = eng
    return eng


def make_core_matrices(molecule):
    """Core = T + V, Overlap, canonical X = U s^-1/2 (s > 1e-6, descending),
    S^1/2 (Methods/hartree_fock.py:207-237)."""
    eng = get_engine(molecule)
    S, T, V = eng.one_electron()
    molecule.Overlap = S.cpu().numpy()
    molecule.Core = (T + V).cpu().numpy()
    s, U = np.linalg.eigh(molecule.Overlap)
    order = s.argsort()[::-1]
    U, s = U[:, order], s[order]
    keep = s > structures.LINEAR_DEPENDENCE
    molecule.X = U[:, keep] * (s[keep] ** -0.5)
    molecule.Xt = molecule.X.T
    molecule.S = (U[:, keep] * s[keep] ** 0.5) @ U[:, keep].T


def evaluate_2e_ints(molecule, threshold=structures.INTEGRAL_THRESHOLD, fp32_threshold=0.0):
    """Schwarz bounds, then every non-negligible unique ERI into HBM
    (Methods/hartree_fock.py:241-325).  Quartets with bound product below
    ``fp32_threshold`` are computed and stored in FP32."""
    eng = get_engine(molecule)
    return eng.build_eri(threshold, fp32_threshold)


def make_coulomb_exchange_matrices(molecule, state):
    """J from the total density, K^sigma = -sum D^sigma_cb (ab|cd)
    (Methods/hartree_fock.py:329-347)."""
    eng = get_engine(molecule)
    dev = eng.device
    Da = torch.from_numpy(state.Alpha.Density).to(dev)
    Dt = torch.from_numpy(state.Alpha.Density + state.Beta.Density).to(dev)
    J, Ka = eng.jk(Dt, Da)
    state.Coulomb = J.cpu().numpy()
    state.Alpha.Exchange = Ka.cpu().numpy()
    if state.NAlpha == state.NBeta and np.array_equal(state.Alpha.Density, state.Beta.Density):
        state.Beta.Exchange = state.Alpha.Exchange
    else:
        _, Kb = eng.jk(Dt, torch.from_numpy(state.Beta.Density).to(dev))
        state.Beta.Exchange = Kb.cpu().numpy()


def make_fock_matrices(molecule, state):
    """Methods/hartree_fock.py:352-355."""
    state.Alpha.Fock = molecule.Core + state.Coulomb + state.Alpha.Exchange
    state.Beta.Fock = molecule.Core + state.Coulomb + state.Beta.Exchange


def make_MOs(molecule, state):
    """eigh of Xt F X (Methods/hartree_fock.py:168-175)."""
    X = molecule.X
    for m in (state.Alpha, state.Beta):
        e, C = np.linalg.eigh(molecule.Xt @ m.Fock @ X)
        m.Energies, m.MOs = e, X @ C


def make_density_matrices(molecule, state):
    """D^sigma = C_occ C_occ^T (Methods/hartree_fock.py:179-184)."""
    for m in (state.Alpha, state.Beta):
        C = m.MOs[:, :m.NOcc]
        m.Density = C @ C.T


def calculate_energy(molecule, state):
    """E = 1/2 [sum Dt H + sum Da Fa + sum Db Fb] (Methods/hartree_fock.py:188-201)."""
    core = np.sum((state.Alpha.Density + state.Beta.Density) * molecule.Core)
    state.Energy = 0.5 * (core + np.sum(state.Alpha.Density * state.Alpha.Fock)
                          + np.sum(state.Beta.Density * state.Beta.Fock))


def do(molecule, conv=structures.ENERGY_CONVERGENCE, maxiter=25, use_diis=True, diis_start=1,
       guess="CORE", threshold=structures.INTEGRAL_THRESHOLD, fp32_threshold=0.0, initial_run=True,
       verbose=False):
    """SCF for the ground state (Methods/hartree_fock.py:23-154).

    ``guess`` is "CORE" (the reference default), or an (N x N) density matrix
    per spin given as a tuple (Da, Db), or a single density used for both.
    Returns a dict with the energy, iteration count, convergence flag, MOs,
    orbital energies, densities and per-stage timings.
    """
    t0 = time.perf_counter()
    timings = {}
    molecule.NuclearRepulsion = nuclear_repulsion(molecule)
    make_core_matrices(molecule)
    timings["one_electron"] = time.perf_counter() - t0
    n = molecule.NOrbitals
    state = ElectronicState(n, molecule.NAlphaElectrons, molecule.NBetaElectrons)
    state.Alpha.Fock = molecule.Core.copy()
    state.Beta.Fock = molecule.Core.copy()
    t1 = time.perf_counter()
    if initial_run:
        evaluate_2e_ints(molecule, threshold, fp32_threshold)
        torch.cuda.synchronize()
    timings["two_electron"] = time.perf_counter() - t1

    if isinstance(guess, str):
        if guess != "CORE":
            raise ValueError("unknown guess %r" % guess)
        make_MOs(molecule, state)
        make_density_matrices(molecule, state)
    else:
        Da, Db = guess if isinstance(guess, tuple) else (guess, guess)
        state.Alpha.Density = np.array(Da, dtype=float)
        state.Beta.Density = np.array(Db, dtype=float)
    calculate_energy(molecule, state)
    dE = 1.0
    it = 0
    history = []
    diis_error = None
    t2 = time.perf_counter()
    while conv < abs(dE):
        it += 1
        if it > maxiter:
            break
        make_coulomb_exchange_matrices(molecule, state)
        make_fock_matrices(molecule, state)
        if use_diis and it > diis_start:
            for m, d in ((state.Alpha, state.AlphaDIIS), (state.Beta, state.BetaDIIS)):
                r = DIIS.residual(molecule.Overlap, m.Density, m.Fock, molecule.X)
                m.Fock = d.extrapolate(r, m.Fock)
            diis_error = max(state.AlphaDIIS.error, state.BetaDIIS.error)
        make_MOs(molecule, state)
        state.Alpha.sort_occupied()
        state.Beta.sort_occupied()
        make_density_matrices(molecule, state)
        old = state.Energy
        calculate_energy(molecule, state)
        dE = state.Energy - old
        state.TotalEnergy = state.Energy + molecule.NuclearRepulsion
        history.append(state.TotalEnergy)
        if verbose:
            print(" Cycle: %d  Total energy: %.12f  Change: %.3e  DIIS error: %s" % (it, state.TotalEnergy, dE, diis_error))
    timings["scf_loop"] = time.perf_counter() - t2
    timings["total"] = time.perf_counter() - t0
    if state.TotalEnergy is None:
        state.TotalEnergy = state.Energy + molecule.NuclearRepulsion
    molecule.State = state
    return dict(energy=state.TotalEnergy, electronic=state.Energy, iterations=it,
                converged=abs(dE) <= conv, Ca=state.Alpha.MOs, Cb=state.Beta.MOs,
                ea=state.Alpha.Energies, eb=state.Beta.Energies, Da=state.Alpha.Density,
                Db=state.Beta.Density, history=history, timings=timings, nuclear=molecule.NuclearRepulsion)
