"""Molecule / basis model with the reference's conventions.

Mirrors ``Molecule``, ``Atom`` and ``ContractedGaussianFunction`` of the
reference (src/hf/molecule.py:14-226) without its ``ase`` dependency:

* xyz coordinates in angstrom are centred at their mean and divided by
  0.52917720859 to give bohr (molecule.py:68-73);
* ``e_nuc`` is the classical nuclear repulsion (molecule.py:75-80);
* a Gaussian94 basis block is parsed per element (molecule.py:115-165) and
  one contracted function is emitted per Cartesian component in the order
  s; px,py,pz; dxx,dxy,dxz,dyy,dyz,dzz; fxxx,... (molecule.py:138-165);
* each component gets per-primitive factors
  N_i = sqrt(2^(2l+1.5) a_i^(l+1.5) / ((2lx-1)!!(2ly-1)!!(2lz-1)!! pi^1.5))
  and contraction coefficients rescaled so that its self-overlap is 1
  (molecule.py:210-226).

Deviation (deliberate): the reference treats an element header line such as
``S     0`` as an S shell with 0 primitives (molecule.py:128-133), which breaks
sulfur, phosphorus and fluorine.  Here only lines after the header are shell
lines.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

import numpy as np

BOHR_ANGSTROM = 0.52917720859  # molecule.py:73

_ELEMENTS = [
    "X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si",
    "P", "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu",
    "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr",
]
ATOMIC_NUMBER = {s: i for i, s in enumerate(_ELEMENTS) if i > 0}

# Cartesian component order per angular momentum (molecule.py:138-165).
CART = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)],
    3: [(3, 0, 0), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 1, 1), (1, 0, 2), (0, 3, 0),
        (0, 2, 1), (0, 1, 2), (0, 0, 3)],
}
_LETTER_L = {"S": 0, "P": 1, "D": 2, "F": 3}


class BasisSetNotAvailableError(Exception):
    """No basis block for an element (molecule.py:14-20)."""

    def __init__(self, msg: str = "no basis set found for this element", *args):
        super().__init__(msg, *args)
        self.msg = msg

    def __str__(self):
        return self.msg


def factorial2(n: int) -> float:
    """(n)!! with (-1)!! = (0)!! = 1, as utils.safe_factorial2 (utils.py:87-89)."""
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


class ContractedGaussianFunction:
    """One Cartesian component of a contracted shell (molecule.py:173-226)."""

    def __init__(self, npgf, expns, coefs, orbital_type, shell, center):
        self.npgf = int(npgf)
        self.expns = [float(e) for e in expns]
        self.coefs = [float(c) for c in coefs]
        self.orbital_type = orbital_type
        self.shell = [int(s) for s in shell]
        self.center = np.asarray(center, dtype=np.float64)
        l = sum(self.shell)
        f2 = (factorial2(2 * self.shell[0] - 1) * factorial2(2 * self.shell[1] - 1)
              * factorial2(2 * self.shell[2] - 1))
        ex = np.asarray(self.expns)
        self.normalize_factors = np.sqrt(2.0 ** (2 * l + 1.5) * ex ** (l + 1.5) / f2 / np.pi ** 1.5)
        prefactor = np.pi ** 1.5 * f2 / 2.0 ** l
        n = 0.0
        for ia in range(len(ex)):
            for ib in range(len(ex)):
                n += (self.normalize_factors[ia] * self.normalize_factors[ib] * self.coefs[ia]
                      * self.coefs[ib] / (ex[ia] + ex[ib]) ** (l + 1.5))
        n = (n * prefactor) ** -0.5
        self.coefs = [c * n for c in self.coefs]


@dataclass
class Shell:
    """A contracted shell: the unit the integral engine works on.

    ``coefs`` holds c_i * N_contraction * sqrt(2^(2l+1.5) a_i^(l+1.5) / pi^1.5),
    i.e. everything except the per-component factor
    1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!) which ``component_norms`` carries.
    """
    L: int
    center: np.ndarray
    expns: np.ndarray
    coefs: np.ndarray
    atom: int
    offset: int = 0  # index of the first Cartesian function

    @property
    def nfunc(self) -> int:
        return (self.L + 1) * (self.L + 2) // 2

    @property
    def nprim(self) -> int:
        return len(self.expns)


def component_norm(lx: int, ly: int, lz: int) -> float:
    return 1.0 / np.sqrt(factorial2(2 * lx - 1) * factorial2(2 * ly - 1) * factorial2(2 * lz - 1))


def parse_gaussian94(text: str, symbol: str):
    """Shells of one element from Gaussian94 text: list of (L, expns, coefs)."""
    target = None
    for block in text.split("****"):
        block = block.strip()
        if re.match(rf"^{symbol}\s+0", block) is not None:
            target = block
            break
    if target is None:
        raise BasisSetNotAvailableError()
    lines = [ln for ln in target.split("\n")]
    shells = []
    i = 1  # line 0 is the element header
    while i < len(lines):
        line = lines[i].strip()
        if line and line[0] in _LETTER_L:
            tok = line.split()
            n = int(tok[1])
            ex, cf = [], []
            for t in lines[i + 1:i + 1 + n]:
                a, c = (float(v.replace("D", "E")) for v in t.split()[:2])
                ex.append(a)
                cf.append(c)
            shells.append((_LETTER_L[line[0]], ex, cf))
            i += 1 + n
        else:
            i += 1
    return shells


class Atom:
    """One atom and its contracted functions (molecule.py:84-168)."""

    def __init__(self, bas: str, symbol: str, number: int, xyz, bas_text: str | None = None):
        self.bas = bas
        self.symbol = symbol
        self.number = int(number)
        self.xyz = np.asarray(xyz, dtype=np.float64)
        if bas_text is None:
            with open(bas) as fh:
                bas_text = fh.read()
        self.shell_defs = parse_gaussian94(bas_text, symbol)
        self.cgfs: list[ContractedGaussianFunction] = []
        for L, ex, cf in self.shell_defs:
            letter = "SPDF"[L]
            for comp in CART[L]:
                self.cgfs.append(ContractedGaussianFunction(len(ex), ex, list(cf), letter, comp,
                                                            self.xyz))
        # per-shell radial data (fixed by element and basis, not by position):
        # L, exponents and coefficients with the contraction normalisation and
        # the x^L component's angular factor folded in
        self.shell_L, self.shell_expns, self.shell_coefs = [], [], []
        k = 0
        for L, ex, _cf in self.shell_defs:
            cg = self.cgfs[k]  # first component (x^L) of this shell
            self.shell_L.append(L)
            self.shell_expns.append(np.asarray(ex, dtype=np.float64))
            self.shell_coefs.append(np.asarray(cg.coefs) * np.asarray(cg.normalize_factors)
                                    * np.sqrt(factorial2(2 * L - 1)))
            k += (L + 1) * (L + 2) // 2
        self.shell_nprim = [len(e) for e in self.shell_expns]

    def __len__(self):
        return len(self.cgfs)


def read_xyz(path: str):
    with open(path) as fh:
        lines = fh.read().strip().split("\n")
    n = int(lines[0].split()[0])
    syms, xyz = [], []
    for ln in lines[2:2 + n]:
        tok = ln.split()
        syms.append(tok[0])
        xyz.append([float(t) for t in tok[1:4]])
    return syms, np.asarray(xyz, dtype=np.float64)


class Molecule:
    """Molecule in the reference's conventions (molecule.py:23-80).

    ``Molecule(xyz_filename, bas_filename)`` reads files like the reference;
    ``Molecule.from_geometry(symbols, xyz_angstrom, basis_text)`` builds from
    arrays (used by the synthetic workloads).
    """

    def __init__(self, xyz_filename: str | None = None, bas_filename: str | None = None, *,
                 symbols=None, xyz_angstrom=None, basis_text: str | None = None):
        if xyz_filename is not None:
            symbols, xyz_angstrom = read_xyz(xyz_filename)
        if basis_text is None:
            with open(bas_filename) as fh:
                basis_text = fh.read()
        self.bas = bas_filename
        self.basis_text = basis_text
        self.symbols = list(symbols)
        self.numbers = np.array([ATOMIC_NUMBER[s] for s in self.symbols], dtype=np.int64)
        self.xyz = np.asarray(xyz_angstrom, dtype=np.float64)
        self.xyz_ = self.xyz - self.xyz.mean(axis=0)
        self.xyz__ = self.xyz_ / BOHR_ANGSTROM
        self.natom = len(self.symbols)
        self.e_nuc = self.nuclear_repulsion()
        self.atoms = [Atom(bas_filename, s, z, r, bas_text=basis_text)
                      for s, z, r in zip(self.symbols, self.numbers, self.xyz__)]

    @classmethod
    def from_geometry(cls, symbols, xyz_angstrom, basis_text: str) -> "Molecule":
        return cls(symbols=symbols, xyz_angstrom=xyz_angstrom, basis_text=basis_text)

    def __len__(self):
        return self.natom

    def nuclear_repulsion(self) -> float:
        e = 0.0
        for i in range(len(self.symbols) - 1):
            for j in range(i + 1, len(self.symbols)):
                e += self.numbers[i] * self.numbers[j] / np.linalg.norm(self.xyz__[i] - self.xyz__[j])
        return float(e)

    # ---- flattened views -------------------------------------------------
    @property
    def cgfs(self) -> list[ContractedGaussianFunction]:
        return [c for a in self.atoms for c in a.cgfs]

    @property
    def nbf(self) -> int:
        return sum(len(a) for a in self.atoms)

    def shells(self) -> list[Shell]:
        """Contracted shells in AO order, with radial normalisation folded in."""
        out = []
        off = 0
        for ia, atom in enumerate(self.atoms):
            for L, ex, c in zip(atom.shell_L, atom.shell_expns, atom.shell_coefs):
                out.append(Shell(L=L, center=atom.xyz.copy(), expns=ex.copy(), coefs=c.copy(),
                                 atom=ia, offset=off))
                off += (L + 1) * (L + 2) // 2
        return out

    def function_norms(self) -> np.ndarray:
        """Per-function angular factors 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!)."""
        return np.array([component_norm(*c.shell) for c in self.cgfs])

    def cgf_arrays(self):
        """Flat per-function arrays in the layout of the reference's CGFs."""
        cg = self.cgfs
        npgf = np.array([c.npgf for c in cg], dtype=np.int32)
        poff = np.zeros(len(cg), dtype=np.int32)
        poff[1:] = np.cumsum(npgf)[:-1]
        coefs = np.concatenate([np.asarray(c.coefs) for c in cg])
        expns = np.concatenate([np.asarray(c.expns) for c in cg])
        norms = np.concatenate([np.asarray(c.normalize_factors) for c in cg])
        shell = np.array([c.shell for c in cg], dtype=np.int32).reshape(-1)
        center = np.array([c.center for c in cg], dtype=np.float64).reshape(-1)
        return npgf, poff, coefs, expns, norms, shell, center
