"""CPU checks of the host side of the nuclear gradient (no GPU work)."""
import numpy as np

from segfault_py_amd import grad, workloads


def test_nuclear_repulsion_gradient_matches_finite_differences():
    """d E_nuc / dR from grad.nuclear_repulsion_gradient equals central
    differences of Molecule.nuclear_repulsion (its pair loop), and the
    forces sum to zero (translation invariance)."""
    mol = workloads.config_molecule(3)  # ten waters, 30 atoms
    g = grad.nuclear_repulsion_gradient(mol)
    R = mol.xyz__
    h = 1e-4
    fd = np.zeros_like(g)
    for a in range(mol.natom):
        for k in range(3):
            x0 = R[a, k]
            R[a, k] = x0 + h
            ep = mol.nuclear_repulsion()
            R[a, k] = x0 - h
            em = mol.nuclear_repulsion()
            R[a, k] = x0
            fd[a, k] = (ep - em) / (2 * h)
    assert np.abs(g - fd).max() < 1e-6 * max(1.0, np.abs(g).max())
    assert np.abs(g.sum(axis=0)).max() < 1e-10
