"""Function lists shaped like the reference's own objects, for the
drop-in boundary tests.

``ref_cgfs(mol)`` builds, for a Molecule of this package, the list
``scf_diis`` hands to ``build_eri_parallel`` (src/hf/hf.py:254-259): one
object per Cartesian component with the attributes of
``ContractedGaussianFunction`` (src/hf/molecule.py:173-226).  As in the
reference's Atom (molecule.py:138-165), all components of a shell are
constructed from ONE coefficient list, which every constructor rescales in
place by its contraction normalisation -- so the objects carry the
reference's rounding, not this package's.  No reference code is imported.
"""
import numpy as np


def _f2(n):
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


class RefShapedCGF:
    def __init__(self, npgf, expns, coefs, orbital_type, shell, center):
        self.npgf = npgf
        self.expns = expns
        self.coefs = coefs            # shared with the shell's other components
        self.orbital_type = orbital_type
        self.shell = shell
        self.center = center
        l = sum(shell)
        f2 = _f2(2 * shell[0] - 1) * _f2(2 * shell[1] - 1) * _f2(2 * shell[2] - 1)
        self.normalize_factors = np.sqrt(np.power(2, 2 * l + 1.5) * np.power(expns, l + 1.5)
                                         / f2 / np.power(np.pi, 1.5))
        pre = np.power(np.pi, 1.5) * f2 / np.power(2.0, l)
        n = 0.0
        for a in range(len(expns)):
            for b in range(len(expns)):
                n += (self.normalize_factors[a] * self.normalize_factors[b] * coefs[a] * coefs[b]
                      / np.power(expns[a] + expns[b], l + 1.5))
        n = np.power(n * pre, -0.5)
        for a in range(len(expns)):
            coefs[a] *= n


CART = {0: [[0, 0, 0]], 1: [[1, 0, 0], [0, 1, 0], [0, 0, 1]],
        2: [[2, 0, 0], [1, 1, 0], [1, 0, 1], [0, 2, 0], [0, 1, 1], [0, 0, 2]]}


def ref_cgfs(mol):
    out = []
    for atom in mol.atoms:
        center = atom.xyz            # one array object per atom, as the reference
        for L, ex, cf in atom.shell_defs:
            expns, coefs = list(ex), list(cf)
            for comp in CART[L]:
                out.append(RefShapedCGF(len(expns), expns, coefs, "SPD"[L], list(comp), center))
    return out
