"""Molecule / basis / shell-pair data model.

Mirrors the reference's data model (Util/structures.py) for the parts on the
integral + Fock-build path:

* ``ContractedGaussian``  Util/structures.py:834-856
* ``Atom``                Util/structures.py:807-828
* ``ShellPair`` / ``Shell`` Util/structures.py:918-968 (shell-pair quantities of
  Methods/c_ints/shellpair_quantities.c:5-38 computed with numpy here)
* ``Molecule``            Util/structures.py:319-523 (ground state only)
* ``read_input``          the INI input format, Util/structures.py:23-47

Besides the reference attributes, ``Molecule`` exposes flat shell tables
(``shell_table()``) that the HIP library consumes.
"""
import ast
import configparser
import json
import math
import os

import numpy as np

# Data/constants.py:2,6-9,31-36
TO_BOHR = 1.8897161646320724
N_ELECTRONS = {'H': 1, 'HE': 2, 'LI': 3, 'BE': 4, 'B': 5, 'C': 6, 'N': 7, 'O': 8, 'F': 9, 'NE': 10}
INTEGRAL_THRESHOLD = 1.0e-8
ENERGY_CONVERGENCE = 1.0e-7
LINEAR_DEPENDENCE = 1.0e-6
DIIS_MAX_CONDITION = 1.0e7

MAX_L = 2   # s, p, d shells


def ncart(l):
    return (l + 1) * (l + 2) // 2


def cart_components(l):
    """Cartesian components (lx, ly, lz), lx descending then ly descending
    (Methods/c_ints/angmom_index.c:3-15)."""
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


# Data/transform_basis.py:3-12 (s, p identity; d rows on xx,xy,xz,yy,yz,zz)
_S3 = math.sqrt(3.0) / 2.0
CART_TO_SPHER = {
    0: np.array([[1.0]]),
    1: np.identity(3),
    2: np.array([[_S3, 0, 0, -_S3, 0, 0],
                 [0, 1, 0, 0, 0, 0],
                 [0, 0, 1, 0, 0, 0],
                 [0, 0, 0, 0, 1, 0],
                 [-0.5, 0, 0, -0.5, 0, 1]], dtype=float),
}

_BASIS_CACHE = None


def basis_library():
    global _BASIS_CACHE
    if _BASIS_CACHE is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "basis_sets.json")
        with open(path) as fh:
            _BASIS_CACHE = json.load(fh)
    return _BASIS_CACHE


def remove_punctuation(basis_set):
    """Util/structures.py:975-977: '6-31G*' -> '631GS'."""
    return (basis_set.replace('*', 's').replace('-', '').replace('(', '').replace(')', '')
            .replace(',', '').upper())


class ContractedGaussian:
    """One contracted shell (Util/structures.py:834-856)."""

    def __init__(self, function, cartesian_l):
        self.AngularMomentum = int(function[0])
        if self.AngularMomentum > MAX_L:
            raise ValueError("shells above d are not supported (l=%d)" % self.AngularMomentum)
        self.NAngMomCart = ncart(self.AngularMomentum)
        self.NAngMomSpher = 2 * self.AngularMomentum + 1
        self.Primitives = [list(p) for p in function[1:]]
        self.NPrimitives = len(self.Primitives)
        self.Exponents = np.array([e for e, _ in self.Primitives], dtype=float)
        self.Coefficients = np.array([c for _, c in self.Primitives], dtype=float)
        self.DoubleExponents = 2.0 * self.Exponents
        l = self.AngularMomentum
        self.ScaledCCs = np.array([c * (2.0 * e) ** ((l + 1.5) / 2.0) for e, c in self.Primitives])
        if l in cartesian_l:
            self.NAngMom = self.NAngMomCart
            self.CartToSpher = np.identity(self.NAngMomCart)
            self.Spherical = False
        else:
            self.NAngMom = self.NAngMomSpher
            self.CartToSpher = CART_TO_SPHER[l]
            self.Spherical = l >= 2
        self.ContractionScaling = np.array(
            [(math.gamma(lx + 0.5) * math.gamma(ly + 0.5) * math.gamma(lz + 0.5)) ** -0.5
             for lx, ly, lz in cart_components(l)])


# Contracted shells depend only on (basis set, element, Cartesian l's), not
# on the atom's position: atoms of one element share the same (read-only)
# ContractedGaussian objects.
_SHELLS = {}


def _element_shells(basis_set, label, cartesian_l, lib):
    key = (basis_set, label, tuple(sorted(set(int(l) for l in cartesian_l))))
    shells = _SHELLS.get(key)
    if shells is None:
        shells = [ContractedGaussian(f, cartesian_l) for f in lib[basis_set][label]]
        _SHELLS[key] = shells
    return shells


class Atom:
    """Util/structures.py:807-828."""

    def __init__(self, index, row, basis_set, cartesian_l, to_bohr):
        label, Z, x, y, z = row
        self.Index = index
        self.Label = str(label).upper()
        self.NuclearCharge = float(Z)
        self.Coordinates = [x * to_bohr, y * to_bohr, z * to_bohr]
        lib = basis_library()
        if basis_set not in lib:
            raise KeyError("basis set %s not available" % basis_set)
        if self.Label not in lib[basis_set]:
            raise KeyError("element %s not in basis %s" % (self.Label, basis_set))
        self.Basis = list(_element_shells(basis_set, self.Label, cartesian_l, lib))
        self.NFunctions = sum(c.NAngMom for c in self.Basis)


class Shell:
    """Util/structures.py:963-968."""

    def __init__(self, coords, cgtf, index, index_vec):
        self.Coords = np.array(coords, dtype=float)
        self.Cgtf = cgtf
        self.Index = index
        self.Ivec = index_vec


class ShellPair:
    """Shell-pair quantities (Util/structures.py:918-956 with
    Methods/c_ints/shellpair_quantities.c:5-38)."""

    def __init__(self, coords_a, cgtf_a, ia, ia_vec, coords_b, cgtf_b, ib, ib_vec):
        self.Centre1 = Shell(coords_a, cgtf_a, ia, ia_vec)
        self.Centre2 = Shell(coords_b, cgtf_b, ib, ib_vec)
        A = np.array(coords_a, dtype=float)
        B = np.array(coords_b, dtype=float)
        a = cgtf_a.Exponents[:, None]
        b = cgtf_b.Exponents[None, :]
        sig = 1.0 / (a + b)
        r2 = float(np.dot(A - B, A - B))
        self.Ltot = cgtf_a.AngularMomentum + cgtf_b.AngularMomentum
        self.PrimitivePairSigmas = sig
        self.PrimitivePairOverlaps = (np.pi * sig) ** 1.5 * np.exp(-a * b * sig * r2)
        self.PrimitivePairCentres = (a[..., None] * A + b[..., None] * B) * sig[..., None]
        self.PrimitivePairHalfSigmas = 0.5 * sig
        self.CentreDisplacement = A - B
        self.Normalization = np.outer(cgtf_a.ContractionScaling, cgtf_b.ContractionScaling)
        self.ContractionCoeffs = np.outer(cgtf_a.ScaledCCs, cgtf_b.ScaledCCs)
        self.BasisTransform = np.kron(cgtf_a.CartToSpher, cgtf_b.CartToSpher)


class Molecule:
    """Ground-state molecule in a single basis (Util/structures.py:319-523).

    ``coords`` rows are ``[label, Z, x, y, z]``; units 'ANGSTROM' (default,
    scaled by TO_BOHR) or 'BOHR'/'ATOMIC'.
    """

    def __init__(self, coords, basis="STO3G", charge=0, multiplicity=1, cartesian_l=(),
                 units="ANGSTROM", build_pairs=False):
        self.Coords = [list(r) for r in coords]
        self.NAtom = len(coords)
        self.Charge = int(charge)
        self.Multiplicity = int(multiplicity)
        self.CartesianL = list(cartesian_l)
        self.CoordsScaleFactor = TO_BOHR if str(units).upper() == "ANGSTROM" else 1.0
        self.Basis = remove_punctuation(basis)
        self.Atoms = [Atom(i, r, self.Basis, self.CartesianL, self.CoordsScaleFactor)
                      for i, r in enumerate(self.Coords)]
        self.NElectrons = sum(N_ELECTRONS[a.Label] for a in self.Atoms) - self.Charge
        self.NOrbitals = sum(a.NFunctions for a in self.Atoms)
        self.NCgtf = sum(len(a.Basis) for a in self.Atoms)
        # Util/structures.py:484-492
        self.NAlphaElectrons = (self.NElectrons + (self.Multiplicity - 1)) // 2
        self.NBetaElectrons = (self.NElectrons - (self.Multiplicity - 1)) // 2
        if self.NAlphaElectrons - self.NBetaElectrons + 1 != self.Multiplicity or \
                self.NAlphaElectrons + self.NBetaElectrons != self.NElectrons:
            raise ValueError("charge and multiplicity inconsistent with specified molecule")
        # flat shell list in global order: atoms, then shells in basis order
        self.Shells = []
        off = 0
        for atom in self.Atoms:
            for cg in atom.Basis:
                self.Shells.append(Shell(atom.Coordinates, cg, len(self.Shells),
                                         list(range(off, off + cg.NAngMom))))
                off += cg.NAngMom
        self._shell_pairs = None
        if build_pairs:
            self.build_shell_pairs()
        self.Core = None
        self.Overlap = None
        self.NuclearRepulsion = None

    def build_shell_pairs(self):
        """All pairs a <= b (Util/structures.py:509-522)."""
        sp = {}
        for a, sa in enumerate(self.Shells):
            for b in range(a, len(self.Shells)):
                sb = self.Shells[b]
                sp[(a, b)] = ShellPair(sa.Coords, sa.Cgtf, a, sa.Ivec, sb.Coords, sb.Cgtf, b, sb.Ivec)
                sp[(a, b)].Molecule = self
        self._shell_pairs = sp
        return sp

    @property
    def ShellPairs(self):
        """Python ShellPair objects (built on first use; the HIP engine builds
        its own device-side pair data)."""
        if self._shell_pairs is None:
            self.build_shell_pairs()
        return self._shell_pairs

    # ------------------------------------------------------------------ #
    def shell_table(self):
        """Flat per-shell arrays consumed by the HIP library and the oracle."""
        ns = len(self.Shells)
        l = np.array([s.Cgtf.AngularMomentum for s in self.Shells], dtype=np.int32)
        nprim = np.array([s.Cgtf.NPrimitives for s in self.Shells], dtype=np.int32)
        prim_off = np.zeros(ns, dtype=np.int32)
        prim_off[1:] = np.cumsum(nprim)[:-1]
        spher = np.array([1 if s.Cgtf.Spherical else 0 for s in self.Shells], dtype=np.int32)
        center = np.ascontiguousarray(np.array([s.Coords for s in self.Shells], dtype=float).reshape(-1))
        exps = np.concatenate([s.Cgtf.Exponents for s in self.Shells]).astype(float)
        cc = np.concatenate([s.Cgtf.ScaledCCs for s in self.Shells]).astype(float)
        ao_off = np.array([s.Ivec[0] for s in self.Shells], dtype=np.int32)
        nfun = np.array([len(s.Ivec) for s in self.Shells], dtype=np.int32)
        cart_off = np.zeros(ns, dtype=np.int32)
        nc = np.array([ncart(int(x)) for x in l], dtype=np.int32)
        cart_off[1:] = np.cumsum(nc)[:-1]
        atom_xyz = np.ascontiguousarray(np.array([a.Coordinates for a in self.Atoms], dtype=float).reshape(-1))
        atom_z = np.array([a.NuclearCharge for a in self.Atoms], dtype=float)
        return dict(nshell=ns, l=l, nprim=nprim, prim_off=prim_off, spher=spher, center=center,
                    exps=exps, cc=cc, ao_off=ao_off, nfun=nfun, cart_off=cart_off,
                    ncart_total=int(nc.sum()), nao=int(self.NOrbitals),
                    natom=self.NAtom, atom_xyz=atom_xyz, atom_z=atom_z)

    def cart_to_final(self):
        """Matrix T (nao x ncart) with final-basis functions = T @ cartesian ones."""
        tab = self.shell_table()
        T = np.zeros((tab["nao"], tab["ncart_total"]))
        for s, sh in enumerate(self.Shells):
            o, c = tab["ao_off"][s], tab["cart_off"][s]
            m = sh.Cgtf.CartToSpher
            T[o:o + m.shape[0], c:c + m.shape[1]] = m
        return T


def read_input(path):
    """Parse a reference-format INI input (Util/structures.py:23-47) into a
    list of (section, Molecule, settings-dict).  Keys are case-insensitive;
    values go through ``ast.literal_eval``."""
    parser = configparser.ConfigParser()
    parser.optionxform = str
    parser.read(path)
    jobs = []
    for section in parser.sections():
        raw = {k.upper(): v for k, v in parser.items(section)}

        def get(key, default=None):
            if key.upper() not in raw:
                return default
            v = ast.literal_eval(raw[key.upper()])
            return v.upper() if isinstance(v, str) else v

        if "COORDS" not in raw:
            raise ValueError("coordinates not found in input file")
        units = get("Coords_Units", "ANGSTROM")
        basis_sets = get("Basis_Sets")
        if not basis_sets:
            raise ValueError("must include Basis_Sets")
        settings = {
            "Method": get("Method", "HF"),
            "Reference": get("Reference", "UHF"),
            "BasisSets": [remove_punctuation(b) for b in basis_sets],
            "MaxIter": get("Max_SCF_Iterations", 25),
            "UseDIIS": get("Use_DIIS", True),
            "DIISSize": get("DIIS_Size", 15),
            "DIISStart": get("DIIS_Start", 1),
            "Guess": get("SCF_Guess", "CORE"),
            "DumpMOs": get("Dump_MOs", False) is True,
        }
        if settings["Guess"] == "READ":
            # Util/structures.py:217-227: one file pair per listed state
            if "MO_READ_STATE" not in raw or "MO_READ_BASIS" not in raw:
                raise ValueError("must give the files to read alpha and beta MOs from, using the keywords "
                                 "MO_Read_Basis and MO_Read_State")
            states = get("MO_Read_State")
            states = [str(i) for i in states] if isinstance(states, (list, tuple)) else str(states)
            mo_basis = remove_punctuation(get("MO_Read_Basis"))
            settings["MOFiles"] = [(mo_basis + "_" + st + ".alpha_MOs", mo_basis + "_" + st + ".beta_MOs")
                                   for st in states]
        mol = Molecule(ast.literal_eval(raw["COORDS"]), basis=settings["BasisSets"][0],
                       charge=get("Charge", 0), multiplicity=get("Multiplicity", 1),
                       cartesian_l=get("Cartesian_L", []) or [], units=units)
        jobs.append((section, mol, settings))
    return jobs
