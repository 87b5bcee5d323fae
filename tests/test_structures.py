"""Data model: basis data, component order, normalisation factors, input parser."""
import math
import os

import numpy as np
import pytest

from pychem_amd import geometry, structures

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_basis_sizes_match_reference_counts():
    """6-31G* with Cartesian d: C has 15 functions, H 2; C10H22 has 194."""
    m = structures.Molecule(geometry.alkane(10), basis="6-31G*", cartesian_l=[2])
    assert m.NOrbitals == 194 and len(m.Shells) == 104
    w = structures.Molecule(geometry.water(), basis="STO-3G")
    assert w.NOrbitals == 7
    assert structures.Molecule(geometry.alkane(20), basis="6-31G*", cartesian_l=[2]).NOrbitals == 384


def test_component_order_and_scaling():
    assert structures.cart_components(2) == [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)]
    cg = structures.ContractedGaussian([2, [0.8, 1.0]], [2])
    assert abs(cg.ScaledCCs[0] - (1.6) ** 1.75) < 1e-15
    g = math.gamma
    assert abs(cg.ContractionScaling[0] - (g(2.5) * g(0.5) * g(0.5)) ** -0.5) < 1e-15
    assert abs(cg.ContractionScaling[1] - (g(1.5) * g(1.5) * g(0.5)) ** -0.5) < 1e-15


def test_remove_punctuation():
    assert structures.remove_punctuation("6-31G*") == "631GS"
    assert structures.remove_punctuation("STO-3G") == "STO3G"


def test_alkane_geometry_bonds():
    rows = geometry.alkane(6)
    xyz = np.array([r[2:] for r in rows])
    C = xyz[:6]
    assert np.allclose(np.linalg.norm(np.diff(C, axis=0), axis=1), 1.54, atol=1e-9)
    H = xyz[6:]
    d = np.linalg.norm(H[:, None, :] - C[None, :, :], axis=2).min(axis=1)
    assert np.allclose(d, 1.09, atol=1e-9)
    assert len(rows) == 6 + 14


def test_read_input_reference_format(tmp_path):
    p = tmp_path / "ex.inp"
    p.write_text(open(os.path.join(ROOT, "tests", "golden", "example5.inp")).read())
    jobs = structures.read_input(str(p))
    (name, mol, st), = jobs
    assert name == "example5" and mol.NOrbitals == 2
    assert st["Reference"] == "RHF" and st["MaxIter"] == 50
    assert abs(mol.Atoms[1].Coordinates[0] - 1.398397232) < 1e-12   # BOHR units


def test_read_input_mo_file_keys(tmp_path):
    """SCF_Guess = READ names one (alpha, beta) file pair per listed state
    from MO_Read_Basis without punctuation (Util/structures.py:217-227);
    Dump_MOs is a boolean (Util/structures.py:170-174)."""
    text = open(os.path.join(ROOT, "tests", "golden", "example5.inp")).read()
    text += "SCF_Guess = 'READ'\nMO_Read_Basis = '6-31G*'\nMO_Read_State = [0, 1]\nDump_MOs = True\n"
    p = tmp_path / "ex.inp"
    p.write_text(text)
    (_, _, st), = structures.read_input(str(p))
    assert st["Guess"] == "READ" and st["DumpMOs"] is True
    assert st["MOFiles"] == [("631GS_0.alpha_MOs", "631GS_0.beta_MOs"), ("631GS_1.alpha_MOs", "631GS_1.beta_MOs")]
    p.write_text(text.replace("MO_Read_State = [0, 1]\n", ""))
    with pytest.raises(ValueError):
        structures.read_input(str(p))


def test_shell_table_offsets():
    m = structures.Molecule(geometry.small_molecule("CH3OH"), basis="6-31G*")
    t = m.shell_table()
    assert t["nao"] == m.NOrbitals
    assert t["ncart_total"] >= t["nao"]
    assert np.all(np.diff(t["ao_off"]) == t["nfun"][:-1])
    T = m.cart_to_final()
    assert T.shape == (t["nao"], t["ncart_total"])
