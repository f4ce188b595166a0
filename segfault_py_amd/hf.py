"""Closed-shell Hartree-Fock drivers on the GPU, drop-in for the reference's
``scf``, ``scf_damping`` and ``scf_diis`` (src/hf/hf.py:65-353).

Same algorithm and constants as the reference: core-Hamiltonian guess,
symmetric orthogonalisation S^-1/2 (hf.py:278-279), P = 2 C_occ C_occ^T,
F = H + J - K/2, E = 1/2 sum P (H + F), stop when |dE| <= tol; damping mixes
0.1 of the previous density (hf.py:205-210); DIIS starts after two plain
iterations (hf.py:317-319) with the commutator error S^-1/2 (FPS - SPF) S^-1/2
(hf.py:327), at most 7 stored vectors (hf.py:324-328) and a fall-back to the
undamped Fock matrix when the Pulay system is singular (hf.py:341-346).

Differences: S, T, V and J/K come from the HIP engine (no N^4 tensor, Schwarz
screening), everything stays on the GPU in FP64 (eigh and the DIIS solve
included), and the functions return an ``SCFResult`` instead of printing
only.  ``precision=1`` runs quartets below ``tau_mixed`` in FP32 with FP64
accumulation.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

from . import engine


@dataclass
class SCFResult:
    energy: float            # total energy (electronic + nuclear), Eh
    e_elec: float
    e_nuc: float
    converged: bool
    iterations: int
    P: object                # density (CUDA tensor)
    C: object                # MO coefficients
    eps: object              # orbital energies
    F: object                # final Fock matrix
    S: object
    H: object
    nocc: int
    wall_s: float = 0.0
    jk_s: float = 0.0
    history: list = field(default_factory=list)

    def energy_weighted_density(self):
        """W = 2 sum_occ eps_i C_i C_i^T (for the overlap-derivative term)."""
        Co = self.C[:, :self.nocc]
        return 2.0 * (Co * self.eps[:self.nocc]) @ Co.T


def _run(mol, mode: str, tol: float, maxiter: int, tau: float, precision: int, tau_mixed: float,
         basis=None, verbose: bool = False) -> SCFResult:
    import torch

    t0 = time.perf_counter()
    b = basis if basis is not None else engine.Basis(mol).screen(tau, tau_mixed)
    S, T, V = b.one_electron()
    H = T + V
    vals, vecs = torch.linalg.eigh(S)
    X = vecs @ torch.diag(1.0 / torch.sqrt(vals)) @ vecs.T
    nocc = int(sum(int(z) for z in mol.numbers)) // 2
    F = H.clone()
    Fs, res = [], []
    P_old = torch.zeros_like(S)
    etot_old = float("inf")
    counter = 0
    converged = False
    jk_s = 0.0
    history = []
    while True:
        eps, Cp = torch.linalg.eigh(X @ F @ X)
        C = X @ Cp
        Co = C[:, :nocc]
        P_new = 2.0 * Co @ Co.T
        if mode == "damping":
            P = 0.1 * P_old + 0.9 * P_new
            P_old = P_new
        else:
            P = P_new
        tj = time.perf_counter()
        J, K = b.jk(P, precision)
        torch.cuda.synchronize()
        jk_s += time.perf_counter() - tj
        F = H + J - 0.5 * K
        etot = float(0.5 * torch.sum(P * (H + F)))
        history.append(etot + mol.e_nuc)
        if verbose:
            print(f"Iter {counter:>4}\tenergy difference {etot - etot_old:.10f}\t\t"
                  f"energy {etot + mol.e_nuc:.10f}")
        if abs(etot - etot_old) <= tol:
            converged = True
            break
        counter += 1
        etot_old = etot
        if counter > maxiter:
            if verbose:
                print(f"SCF iteration did not coverge in {maxiter} steps!")
            break
        if mode != "diis" or counter <= 2:
            continue
        if len(Fs) > 6:
            Fs.pop(0)
            res.pop(0)
        res.append(X @ (F @ P @ S - S @ P @ F) @ X)
        Fs.append(F)
        n = len(Fs)
        Bm = torch.zeros((n + 1, n + 1), dtype=torch.float64, device=S.device)
        Bm[-1, :] = -1.0
        Bm[:, -1] = -1.0
        Bm[-1, -1] = 0.0
        for i in range(n):
            for j in range(n):
                Bm[i, j] += torch.sum(res[i] * res[j])
        rhs = torch.zeros(n + 1, dtype=torch.float64, device=S.device)
        rhs[-1] = -1.0
        try:
            coef = torch.linalg.solve(Bm, rhs)
        except torch.linalg.LinAlgError:
            if verbose:
                print("DIIS solve Pulay equation failed, last step Fock matrix is resoted")
            continue
        if not bool(torch.isfinite(coef).all()):
            continue
        F = sum(Fs[i] * coef[i] for i in range(n))
    if verbose:
        print(f"SCF energy {etot + mol.e_nuc}")
    return SCFResult(energy=etot + mol.e_nuc, e_elec=etot, e_nuc=mol.e_nuc, converged=converged,
                     iterations=counter, P=P, C=C, eps=eps, F=F, S=S, H=H, nocc=nocc,
                     wall_s=time.perf_counter() - t0, jk_s=jk_s, history=history)


def scf(mol, tol=1e-8, maxiter=30, *, tau=1e-12, precision=0, tau_mixed=1e-5, basis=None,
        verbose=False) -> SCFResult:
    """Plain SCF iterations (hf.py:65-160)."""
    return _run(mol, "plain", tol, maxiter, tau, precision, tau_mixed, basis, verbose)


def scf_damping(mol, tol=1e-8, maxiter=100, *, tau=1e-12, precision=0, tau_mixed=1e-5, basis=None,
                verbose=False) -> SCFResult:
    """SCF with density damping 0.1 (hf.py:164-232)."""
    return _run(mol, "damping", tol, maxiter, tau, precision, tau_mixed, basis, verbose)


def scf_diis(mol, tol=1e-8, maxiter=100, *, tau=1e-12, precision=0, tau_mixed=1e-5, basis=None,
             verbose=False) -> SCFResult:
    """SCF with Pulay DIIS (hf.py:236-353)."""
    return _run(mol, "diis", tol, maxiter, tau, precision, tau_mixed, basis, verbose)
