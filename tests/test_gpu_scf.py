"""GPU RHF vs the oracle SCF and the reference's golden energies."""
import numpy as np
import pytest

from pychem_amd import geometry, structures
from oracle import scf as oscf

pytestmark = pytest.mark.gpu

# Tests/example{5,4,6,7}.out; converged there only to |dE| < 1e-7
GOLDENS = [
    ("H2_sto3g", [["H", 1.0, 0, 0, 0], ["H", 1.0, 1.398397232, 0, 0]], "STO-3G", "BOHR", -1.11675932391, 3e-8),
    ("H2_6311g", [["H", 1.0, 0, 0, 0], ["H", 1.0, 1.0, 0, 0]], "6-311G", "ANGSTROM", -1.09686463292, 3e-7),
    ("LiH_sto3g", [["Li", 3.0, 0, 0, 0], ["H", 1.0, 3.013923978561, 0, 0]], "STO-3G", "BOHR", -7.86202531689, 3e-6),
    ("BH_sto3g", [["B", 5.0, 0, 0, 0], ["H", 1.0, 3.013923978561, 0, 0]], "STO-3G", "BOHR", -24.7030877843, 1e-7),
]


@pytest.mark.parametrize("name,rows,basis,units,ref,tol", GOLDENS, ids=[g[0] for g in GOLDENS])
def test_reference_goldens(gpu, name, rows, basis, units, ref, tol):
    from pychem_amd import hartree_fock
    m = structures.Molecule(rows, basis=basis, units=units)
    r = hartree_fock.do(m, conv=1e-7, maxiter=50)
    assert abs(r["energy"] - ref) < tol


def test_water_sto3g_tight(gpu):
    """Config: water / STO-3G, converged to 1e-10; energy parity 1e-8 vs oracle."""
    from pychem_amd import hartree_fock
    m = structures.Molecule(geometry.water(), basis="STO-3G")
    r = hartree_fock.do(m, conv=1e-10, maxiter=100)
    o = oscf.run_rhf(m, conv=1e-10)
    assert r["converged"]
    assert abs(r["energy"] - o["energy"]) < 1e-8
    assert abs(r["energy"] - (-74.96292752)) < 1e-6


@pytest.mark.parametrize("n", [4])
def test_alkane_631gs_vs_oracle(gpu, n):
    from pychem_amd import hartree_fock
    m = structures.Molecule(geometry.alkane(n, 0.02, 1000 * n), basis="6-31G*", cartesian_l=[2])
    r = hartree_fock.do(m, conv=1e-10, maxiter=100)
    o = oscf.run_rhf(m, conv=1e-10)
    assert abs(r["energy"] - o["energy"]) < 1e-8
    assert np.abs(r["ea"] - o["ea"]).max() < 1e-7
