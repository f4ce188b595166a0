"""The SCF drivers' host logic (segfault_py_amd/hf.py) on the CPU.

hf._run takes its integrals from a basis object; here a stand-in basis
serves the oracle's S, T, V and dense-tensor J/K as CPU tensors, so the
drivers' own control flow -- DIIS bookkeeping, the singular-system
fall-back, the iteration limit, the overlap check -- runs without a GPU and
is compared with the oracle's restatement of scf_diis (hf.py:236-353).
Also: error paths at the library boundary (bad nuclei, bad primitives,
elements missing from the basis)."""
import ctypes
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN


class OracleBasis:
    """Integrals from the oracle, as CPU tensors, behind the Basis API the
    SCF drivers use (one_electron, jk, jk_graph)."""

    def __init__(self, oracle, mol, jk_scale=1.0, duplicate_first=False):
        S, T, V = oracle.one_electron(mol, 8)
        self.eri = oracle.eri_tensor(mol, 8)
        if duplicate_first:       # a repeated function: S singular
            S[1, :] = S[0, :]
            S[:, 1] = S[:, 0]
        self.S, self.T, self.V = (torch.tensor(m) for m in (S, T, V))
        self.jk_scale = jk_scale
        self.nbf = mol.nbf
        self.calls = 0

    def one_electron(self):
        return self.S.clone(), self.T.clone(), self.V.clone()

    def jk(self, P, precision=0, out=None):
        self.calls += 1
        Pn = P.numpy()
        J = np.einsum("pqrs,rs->pq", self.eri, Pn) * self.jk_scale
        K = np.einsum("prqs,rs->pq", self.eri, Pn) * self.jk_scale
        if out is None:
            return torch.tensor(J), torch.tensor(K)
        out[0].copy_(torch.tensor(J))
        out[1].copy_(torch.tensor(K))
        return out

    def jk_graph(self, P, out, precision=0):
        return self.jk(P, precision, out)


def _oracle_energy(oracle, mol, tol, maxiter=100):
    S, T, V = oracle.one_electron(mol, 8)
    eri = oracle.eri_tensor(mol, 8)
    return oracle.scf_diis(S, T + V, int(mol.numbers.sum()) // 2,
                           lambda P: oracle.jk_dense(eri, P), mol.e_nuc, tol=tol, maxiter=maxiter)


def test_diis_driver_matches_oracle_restatement(oracle, water_sp):
    from segfault_py_amd import hf
    E, _, conv, it = _oracle_energy(oracle, water_sp, 1e-10)
    r = hf.scf_diis(water_sp, tol=1e-10, basis=OracleBasis(oracle, water_sp))
    assert r.converged and conv
    assert r.iterations == it
    assert abs(r.energy - E) < 1e-10


def test_plain_and_damped_drivers_converge_to_the_same_energy(oracle, water_sp):
    from segfault_py_amd import hf
    e = [f(water_sp, tol=1e-10, basis=OracleBasis(oracle, water_sp)).energy
         for f in (hf.scf, hf.scf_damping, hf.scf_diis)]
    assert max(e) - min(e) < 1e-8


def test_iteration_limit_reports_not_converged(oracle, water_sp):
    """maxiter = 0: one Fock build, no convergence (etot_old = inf), stop."""
    from segfault_py_amd import hf
    b = OracleBasis(oracle, water_sp)
    r = hf.scf_diis(water_sp, maxiter=0, basis=b)
    assert not r.converged and r.iterations == 1 and b.calls == 1
    r = hf.scf_diis(water_sp, maxiter=3, tol=1e-14, basis=OracleBasis(oracle, water_sp))
    assert not r.converged and r.iterations == 4


def test_singular_pulay_system_falls_back(oracle, water_sp, monkeypatch):
    """Where the Pulay system is singular the reference keeps the
    iteration's undamped Fock matrix (hf.py:341-346): with every
    extrapolation refused, the DIIS driver follows the plain iterations."""
    from segfault_py_amd import hf
    Fs = [torch.eye(3, dtype=torch.float64)] * 2
    _, ok = hf._diis_extrapolate(Fs, [torch.ones(3, 3, dtype=torch.float64)] * 2)
    assert not bool(ok)                               # identical residuals: singular
    r1, r2 = torch.ones(3, 3, dtype=torch.float64), torch.eye(3, dtype=torch.float64)
    F, ok = hf._diis_extrapolate([2 * Fs[0], 3 * Fs[0]], [r1, r2])
    assert bool(ok)
    B = np.array([[9.0, 3.0, -1.0], [3.0, 3.0, -1.0], [-1.0, -1.0, 0.0]])
    c = np.linalg.solve(B, [0.0, 0.0, -1.0])
    assert torch.allclose(F, (2 * c[0] + 3 * c[1]) * Fs[0], rtol=1e-14, atol=0)
    refused = lambda self: (self.Fs[-1], torch.tensor(False))  # noqa: E731
    monkeypatch.setattr(hf._Pulay, "extrapolate", refused)
    rd = hf.scf_diis(water_sp, tol=1e-10, maxiter=60, basis=OracleBasis(oracle, water_sp))
    rp = hf.scf(water_sp, tol=1e-10, maxiter=60, basis=OracleBasis(oracle, water_sp))
    assert rd.history == rp.history


def test_singular_overlap_is_refused(oracle, water_sp):
    from segfault_py_amd import hf
    with pytest.raises(ValueError, match="not positive definite"):
        hf.scf_diis(water_sp, basis=OracleBasis(oracle, water_sp, duplicate_first=True))


def test_element_missing_from_the_basis():
    from segfault_py_amd.molecule import BasisSetNotAvailableError, Molecule
    with pytest.raises(BasisSetNotAvailableError):
        Molecule.from_geometry(["S", "H", "H"], np.eye(3),
                               open(os.path.join(GOLDEN, "cc-pvdz-HCNO.dat")).read())


def test_library_rejects_bad_primitives_and_nuclei():
    from segfault_py_amd import engine
    L = engine.lib()
    h = ctypes.c_void_p()
    one = np.ones(1)
    for ex, cf, cen in ((0.0, 1.0, 0.0), (-1.0, 1.0, 0.0), (np.nan, 1.0, 0.0),
                        (1.0, np.inf, 0.0), (1.0, 1.0, np.nan)):
        rc = L.sf_basis_create(1, np.zeros(1, dtype=np.int32), np.ones(1, dtype=np.int32),
                               np.full(3, cen), np.array([ex]), np.array([cf]), 1,
                               np.zeros(1, dtype=np.int32), ctypes.byref(h))
        assert rc == 1 and not h.value, (ex, cf, cen)
    assert L.sf_basis_set_nuclei(None, 1, one, np.zeros(3)) == 1


def test_batched_pulay_equals_per_system():
    """The batched DIIS store (one (b, n, n) system per molecule) gives each
    molecule the extrapolation of its own store, and the Gram matrix kept
    row by row equals the full one, also after the oldest vectors drop."""
    from segfault_py_amd import hf
    g = torch.Generator().manual_seed(3)
    b, n = 3, 5
    batched = hf._Pulay(4)
    single = [hf._Pulay(4) for _ in range(b)]
    for _ in range(6):
        F = torch.randn(b, n, n, generator=g, dtype=torch.float64)
        r = torch.randn(b, n, n, generator=g, dtype=torch.float64)
        batched.push(F, r)
        for m in range(b):
            single[m].push(F[m], r[m])
    R = torch.stack(batched.res, dim=1).reshape(b, 4, -1)
    assert torch.allclose(batched.B, R @ R.transpose(1, 2), rtol=1e-13, atol=1e-12)
    Fb, okb = batched.extrapolate()
    for m in range(b):
        Fm, okm = single[m].extrapolate()
        assert bool(okb[m]) and bool(okm)
        assert torch.allclose(Fb[m], Fm, rtol=1e-12, atol=1e-12)
