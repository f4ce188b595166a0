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


class _Pulay:
    """The DIIS store of hf.py:324-346 -- at most 7 Fock matrices and their
    commutator residuals, oldest first -- with the residuals' Gram matrix
    B_ij = <r_i, r_j> kept up to date: a new residual adds one row, the
    dots of every stored residual with it (one broadcast multiply and one
    row sum; rocBLAS takes ~30 ms for the equivalent (7 x n^2)(n^2 x 7)
    GEMM at n = 440).  Works on one (n, n) system or a batch (b, n, n)."""

    def __init__(self, nmax: int = 7):
        self.nmax = nmax
        self.Fs, self.res = [], []
        self.B = None

    def push(self, F, r):
        import torch

        if len(self.Fs) >= self.nmax:
            self.Fs.pop(0)
            self.res.pop(0)
            self.B = self.B[..., 1:, 1:]
        self.Fs.append(F)
        self.res.append(r)
        k = len(self.res)
        lead = r.shape[:-2]
        R = torch.stack(self.res, dim=len(lead)).reshape(*lead, k, -1)
        d = (R * r.reshape(*lead, 1, -1)).sum(dim=-1)               # (..., k)
        B = torch.empty((*lead, k, k), dtype=r.dtype, device=r.device)
        if k > 1:
            B[..., :k - 1, :k - 1] = self.B
        B[..., k - 1, :] = d
        B[..., :, k - 1] = d
        self.B = B

    def extrapolate(self):
        """(F_diis, ok): the Pulay combination from the bordered system,
        solved with solve_ex (no exception, so no host synchronisation); ok
        is false where the reference's solve raises (exactly singular) or
        the coefficients are not finite."""
        import torch

        k = len(self.Fs)
        lead = self.B.shape[:-2]
        Bm = torch.full((*lead, k + 1, k + 1), -1.0, dtype=self.B.dtype, device=self.B.device)
        Bm[..., :k, :k] = self.B
        Bm[..., k, k] = 0.0
        rhs = torch.zeros((*lead, k + 1), dtype=self.B.dtype, device=self.B.device)
        rhs[..., k] = -1.0
        coef, info = torch.linalg.solve_ex(Bm, rhs)
        ok = (info == 0) & torch.isfinite(coef).all(dim=-1)
        Fst = torch.stack(self.Fs, dim=len(lead))                     # (..., k, n, n)
        F = (coef[..., :k, None, None] * Fst).sum(dim=len(lead))
        return F, ok


def _diis_extrapolate(Fs, res):
    """Pulay extrapolation of the stored matrices (oldest first); see _Pulay."""
    d = _Pulay(len(Fs))
    for F, r in zip(Fs, res):
        d.push(F, r)
    return d.extrapolate()


def _run(mol, mode: str, tol: float, maxiter: int, tau: float, precision: int, tau_mixed: float,
         basis=None, verbose: bool = False, graph: bool = True) -> SCFResult:
    import torch

    t0 = time.perf_counter()
    b = basis if basis is not None else engine.Basis(mol).screen(tau, tau_mixed)
    S, T, V = b.one_electron()
    H = T + V
    vals, vecs = torch.linalg.eigh(S)
    lam_min, lam_max = float(vals[0]), float(vals[-1])
    if not lam_min > 1e-12 * lam_max:
        # S^-1/2 of the reference (hf.py:278-279) would amplify rounding by
        # 1/sqrt(lam_min): refuse instead of iterating on noise
        raise ValueError(f"overlap matrix is not positive definite to working precision "
                         f"(eigenvalues {lam_min:.3e} .. {lam_max:.3e}): linearly dependent "
                         "basis functions")
    X = vecs @ torch.diag(1.0 / torch.sqrt(vals)) @ vecs.T
    nocc = int(sum(int(z) for z in mol.numbers)) // 2
    on_gpu = S.is_cuda
    F = H.clone()
    pulay = _Pulay(7)
    P_old = torch.zeros_like(S)
    # the J/K build reads P from, and writes J and K to, fixed buffers, so
    # every iteration replays one captured graph (sf_jk_graph)
    P_buf, J_buf, K_buf = torch.zeros_like(S), torch.zeros_like(S), torch.zeros_like(S)
    etot_old = float("inf")
    counter = 0
    converged = False
    events = []
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
        P_buf.copy_(P)
        if on_gpu:
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ev[0].record()
        if graph:
            b.jk_graph(P_buf, (J_buf, K_buf), precision)
        else:
            b.jk(P_buf, precision, out=(J_buf, K_buf))
        if on_gpu:
            ev[1].record()
            events.append(ev)
        F = H + J_buf - 0.5 * K_buf
        e_dev = 0.5 * torch.sum(P * (H + F))
        # the DIIS step of this iteration (it runs unless this iteration
        # converges) is queued before the one host read of the iteration
        diis = mode == "diis" and counter + 1 > 2 and counter + 1 <= maxiter
        if diis:
            pulay.push(F, X @ (F @ P @ S - S @ P @ F) @ X)
            F_diis, ok = pulay.extrapolate()
            host = torch.stack([e_dev, ok.to(e_dev.dtype)]).cpu()
        else:
            host = e_dev.reshape(1).cpu()
        etot = float(host[0])
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
                print(f"SCF stopped after {maxiter} iterations without converging")
            break
        if not diis:
            continue
        if bool(host[1]):
            F = F_diis
        elif verbose:
            print("singular Pulay system: keeping this iteration's Fock matrix")
    if on_gpu:
        torch.cuda.synchronize()
    jk_s = 1e-3 * sum(a.elapsed_time(z) for a, z in events)
    if verbose:
        print(f"SCF energy {etot + mol.e_nuc}")
    return SCFResult(energy=etot + mol.e_nuc, e_elec=etot, e_nuc=mol.e_nuc, converged=converged,
                     iterations=counter, P=P, C=C, eps=eps, F=F, S=S, H=H, nocc=nocc,
                     wall_s=time.perf_counter() - t0, jk_s=jk_s, history=history)


def scf(mol, tol=1e-8, maxiter=30, *, tau=1e-12, precision=0, tau_mixed=1e-5, basis=None,
        verbose=False, graph=True) -> SCFResult:
    """Plain SCF iterations (hf.py:65-160)."""
    return _run(mol, "plain", tol, maxiter, tau, precision, tau_mixed, basis, verbose, graph)


def scf_damping(mol, tol=1e-8, maxiter=100, *, tau=1e-12, precision=0, tau_mixed=1e-5, basis=None,
                verbose=False, graph=True) -> SCFResult:
    """SCF with density damping 0.1 (hf.py:164-232)."""
    return _run(mol, "damping", tol, maxiter, tau, precision, tau_mixed, basis, verbose, graph)


def scf_diis(mol, tol=1e-8, maxiter=100, *, tau=1e-12, precision=0, tau_mixed=1e-5, basis=None,
             verbose=False, graph=True) -> SCFResult:
    """SCF with Pulay DIIS (hf.py:236-353)."""
    return _run(mol, "diis", tol, maxiter, tau, precision, tau_mixed, basis, verbose, graph)


@dataclass
class BatchSCFResult:
    energies: list           # total energy per molecule, Eh
    converged: list          # per molecule
    iterations: list         # iteration at which each molecule converged (or stopped)
    wall_s: float            # whole batch, setup included
    scf_s: float             # SCF iterations only
    jk_s: float              # J/K builds (graph replays), summed
    n_iter: int              # iterations run for the batch
    tuning: object = None    # autotuned launch shapes (or None)


def scf_batch(mols, tol=1e-8, maxiter=100, *, tau=1e-12, precision=0, tau_mixed=1e-5,
              autotune=True, graph=True, basis=None) -> BatchSCFResult:
    """Pulay-DIIS SCF (hf.py:236-353) of many independent molecules at once.

    Each molecule follows the reference algorithm on its own -- core guess,
    S^-1/2, DIIS after two plain iterations with at most 7 vectors and the
    undamped fall-back when its Pulay system is singular, stop when its
    |dE| <= tol -- but all molecules advance together: their matrices are
    padded to (nmol, nmax, nmax), eigh/bmm/solve run batched, J and K for
    every molecule come from one persistent launch per quartet class of a
    single BatchBasis, replayed from a captured hipGraph every iteration.
    A converged molecule is frozen (its density, Fock matrix and energy stop
    changing) while the others continue.
    """
    import torch

    t0 = time.perf_counter()
    b = basis if basis is not None else engine.BatchBasis(mols).screen(tau, tau_mixed)
    B, n = b.nmol, b.nmax
    dev = torch.device("cuda")
    nbf = torch.tensor(b.nbf_mol, device=dev)
    idx = torch.arange(n, device=dev)
    real = idx[None, :] < nbf[:, None]                          # (B, n) real functions
    mask2 = (real[:, :, None] & real[:, None, :]).double()      # (B, n, n)
    pad = (~real).double()
    nocc = torch.tensor([int(sum(int(z) for z in m.numbers)) // 2 for m in mols], device=dev)
    occ = (idx[None, :] < nocc[:, None]).double()               # (B, n)
    e_nuc = torch.tensor([m.e_nuc for m in mols], dtype=torch.float64, device=dev)

    S, T, V = b.one_electron()
    H = T + V
    # S^-1/2 per molecule: identity on the padding keeps S invertible, the
    # mask removes it again
    lam, U = torch.linalg.eigh(S + torch.diag_embed(pad))
    bad = lam[:, 0] <= 1e-12 * lam[:, -1]
    if bool(bad.any()):
        raise ValueError(f"molecules {torch.nonzero(bad).flatten().tolist()}: overlap matrix not "
                         "positive definite to working precision (linearly dependent functions)")
    X = (U * (1.0 / torch.sqrt(lam))[:, None, :]) @ U.transpose(1, 2)
    X = (X * mask2).contiguous()
    shift = torch.diag_embed(pad * 1e6)                          # padded orbitals last
    P_buf = torch.zeros((B, n, n), dtype=torch.float64, device=dev)
    J_buf = torch.zeros_like(P_buf)
    K_buf = torch.zeros_like(P_buf)
    tuning = None
    if autotune:
        tuning = b.autotune(P_buf, precision)
    F = H.clone()
    P = torch.zeros_like(P_buf)
    E = torch.zeros(B, dtype=torch.float64, device=dev)
    E_old = torch.full((B,), float("inf"), dtype=torch.float64, device=dev)
    done = torch.zeros(B, dtype=torch.bool, device=dev)
    done_iter = torch.full((B,), -1, dtype=torch.long, device=dev)
    pulay = _Pulay(7)
    jk_s = 0.0
    counter = 0
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    while True:
        eps, Cp = torch.linalg.eigh(X @ F @ X + shift)
        C = X @ Cp
        P_new = 2.0 * (C * occ[:, None, :]) @ C.transpose(1, 2)
        keep = done[:, None, None]
        P = torch.where(keep, P, P_new)
        P_buf.copy_(P)
        tj = time.perf_counter()
        if graph:
            b.jk_graph(P_buf, (J_buf, K_buf), precision)
        else:
            b.jk(P_buf, precision, out=(J_buf, K_buf))
        torch.cuda.synchronize()
        jk_s += time.perf_counter() - tj
        F_new = H + J_buf - 0.5 * K_buf
        E_new = 0.5 * torch.sum(P * (H + F_new), dim=(1, 2))
        F = torch.where(keep, F, F_new)
        E = torch.where(done, E, E_new)
        newly = (~done) & (torch.abs(E - E_old) <= tol)
        done_iter = torch.where(newly, torch.full_like(done_iter, counter), done_iter)
        done = done | newly
        if bool(done.all()):
            break
        counter += 1
        E_old = E.clone()
        if counter > maxiter:
            break
        if counter <= 2:
            continue
        pulay.push(F, X @ (F @ P @ S - S @ P @ F) @ X)
        F_diis, ok = pulay.extrapolate()
        ok = ok & (~done)
        F = torch.where(ok[:, None, None], F_diis, F)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    energies = (E + e_nuc).cpu().tolist()
    its = torch.where(done, done_iter, torch.full_like(done_iter, counter)).cpu().tolist()
    return BatchSCFResult(energies=energies, converged=done.cpu().tolist(), iterations=its,
                          wall_s=t2 - t0, scf_s=t2 - t1, jk_s=jk_s, n_iter=counter,
                          tuning=tuning)
