"""Synthetic molecules and even-tempered basis sets for the ERI / J-K workloads.

Frozen decisions (the workload text leaves them open):

* Even-tempered exponents a_k = 0.15 * 3^k (bohr^-2).  Shell j of angular
  type L on an element takes the DISJOINT triple a_{3j}, a_{3j+1}, a_{3j+2}
  with raw coefficients 1/3 each (then the reference normalisation of
  molecule.py:210-226).  The alternative reading -- overlapping windows
  a_j, a_{j+1}, a_{j+2} -- is rejected: it makes shells of one type nearly
  linearly dependent.
* A d shell is a single primitive with a = 0.8 (water O in the s/p/d set,
  and C in the alkane set); an f shell (the f-shell test set only) a single
  primitive with a = 1.1.
* Water: O at the origin, H at 0.96 A along x and at 0.96 A rotated by
  104.5 deg in the xy plane.
* Clusters: per water a rotation from a normalised Gaussian quaternion and an
  O position uniform in [0, 12]^3 A, redrawn while any O-O distance is below
  2.7 A; all draws from numpy.random.default_rng(seed).
* Alkanes: zigzag carbon chain in the xy plane (C-C 1.54 A, C-C-C 109.47 deg),
  C-H 1.09 A in tetrahedral directions, then a seeded N(0, 0.02 A) jitter on
  every atom.
"""
from __future__ import annotations

import numpy as np

from .molecule import Molecule

ALPHA0 = 0.15
RATIO = 3.0
D_EXPONENT = 0.8
F_EXPONENT = 1.1

# number of (s, p, d[, f]) shells per element for each basis family
BASIS_SP = {"O": (4, 2, 0), "H": (3, 0, 0), "C": (4, 3, 0)}
BASIS_SPD = {"O": (4, 2, 1), "H": (3, 0, 0), "C": (4, 3, 1)}
BASIS_ALKANE = {"C": (4, 3, 1), "H": (3, 1, 0)}
# f shells on O and H (every one of the 55 quartet classes, f on two centres)
BASIS_SPDF = {"O": (4, 2, 1, 1), "H": (3, 1, 0, 1)}


def even_tempered_text(shell_counts: dict) -> str:
    """Gaussian94 text for the even-tempered family described above."""
    blocks = []
    for elem, counts in shell_counts.items():
        ns, npp, nd = counts[:3]
        nf = counts[3] if len(counts) > 3 else 0
        lines = [f"{elem}     0"]
        for letter, count in (("S", ns), ("P", npp)):
            for j in range(count):
                lines.append(f"{letter}   3   1.00")
                for k in range(3 * j, 3 * j + 3):
                    lines.append(f"  {ALPHA0 * RATIO ** k:.10f}   {1.0 / 3.0:.17f}")
        for _ in range(nd):
            lines.append("D   1   1.00")
            lines.append(f"  {D_EXPONENT:.10f}   1.0")
        for _ in range(nf):
            lines.append("F   1   1.00")
            lines.append(f"  {F_EXPONENT:.10f}   1.0")
        blocks.append("\n".join(lines))
    return "\n****\n".join(blocks) + "\n****\n"


def water_geometry() -> np.ndarray:
    """O, H, H positions (angstrom) of the reference water geometry."""
    th = np.deg2rad(104.5)
    return np.array([[0.0, 0.0, 0.0], [0.96, 0.0, 0.0],
                     [0.96 * np.cos(th), 0.96 * np.sin(th), 0.0]])


def _quat_rotation(q: np.ndarray) -> np.ndarray:
    q = q / np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(n: int, seed: int, box: float = 12.0, min_oo: float = 2.7):
    rng = np.random.default_rng(seed)
    base = water_geometry()
    centers = []
    syms, xyz = [], []
    while len(centers) < n:
        rot = _quat_rotation(rng.standard_normal(4))
        o = rng.uniform(0.0, box, size=3)
        if any(np.linalg.norm(o - c) < min_oo for c in centers):
            continue
        centers.append(o)
        for s, r in zip("OHH", base @ rot.T + o):
            syms.append(s)
            xyz.append(r)
    return syms, np.asarray(xyz)


def alkane(n_carbon: int, seed: int, jitter: float = 0.02):
    """Zigzag CnH(2n+2) (angstrom).

    Carbon i sits at (i dx, (i mod 2) dy, 0).  Every carbon gets two H in the
    plane perpendicular to the chain plane, along -b cos(54.735) +- z sin(54.735)
    with b the bisector of its two C-C bonds; a chain end gets one more H
    where the zigzag would put the next carbon.  Virtual carbons at i = -1 and
    i = n stand in for missing neighbours, so every centre is tetrahedral.
    """
    rng = np.random.default_rng(seed)
    cc, ch = 1.54, 1.09
    half = np.deg2rad(109.47) / 2.0
    dx, dy = cc * np.sin(half), cc * np.cos(half)

    def cpos(i):
        return np.array([i * dx, (i % 2) * dy, 0.0])

    syms = ["C"] * n_carbon
    xyz = [cpos(i) for i in range(n_carbon)]
    zhat = np.array([0.0, 0.0, 1.0])
    for i in range(n_carbon):
        c = cpos(i)
        n1, n2 = cpos(i - 1) - c, cpos(i + 1) - c
        b = n1 / np.linalg.norm(n1) + n2 / np.linalg.norm(n2)
        out = -b / np.linalg.norm(b)
        dirs = [np.cos(half) * out + np.sin(half) * zhat, np.cos(half) * out - np.sin(half) * zhat]
        if i == 0:
            dirs.append(n1 / np.linalg.norm(n1))
        if i == n_carbon - 1:
            dirs.append(n2 / np.linalg.norm(n2))
        for d in dirs:
            syms.append("H")
            xyz.append(c + ch * d)
    xyz = np.asarray(xyz) + rng.normal(0.0, jitter, size=(len(syms), 3))
    return syms, xyz


def config_molecule(index: int) -> Molecule:
    """Molecules of the five workloads, by index 0..3 (index 4: see batch())."""
    if index == 0:
        return Molecule.from_geometry(list("OHH"), water_geometry(), even_tempered_text(BASIS_SP))
    if index == 1:
        s, x = water_cluster(20, seed=2)
        return Molecule.from_geometry(s, x, even_tempered_text(BASIS_SPD))
    if index == 2:
        s, x = alkane(16, seed=3)
        return Molecule.from_geometry(s, x, even_tempered_text(BASIS_ALKANE))
    if index == 3:
        s, x = water_cluster(10, seed=4)
        return Molecule.from_geometry(s, x, even_tempered_text(BASIS_SPD))
    raise ValueError(f"no single molecule for workload index {index}")


def water_spdf() -> Molecule:
    """The reference water with f shells on O and both H (BASIS_SPDF, N = 58):
    every quartet class of the engine, f shells on different centres."""
    return Molecule.from_geometry(list("OHH"), water_geometry(), even_tempered_text(BASIS_SPDF))


def batch(n: int = 256, seed: int = 5) -> list[Molecule]:
    """Uneven molecules: water clusters of 1-8 waters or CkH(2k+2), k = 1..7."""
    rng = np.random.default_rng(seed)
    text = even_tempered_text(BASIS_SP)
    mols = []
    for i in range(n):
        if rng.uniform() < 0.5:
            k = int(rng.integers(1, 9))
            s, x = water_cluster(k, seed=int(rng.integers(1 << 30)), box=4.0 + 2.0 * k ** (1 / 3))
        else:
            k = int(rng.integers(1, 8))
            s, x = alkane(k, seed=int(rng.integers(1 << 30)))
        mols.append(Molecule.from_geometry(s, x, text))
    return mols
