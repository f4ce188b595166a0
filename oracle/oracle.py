"""ctypes front-end of the CPU FP64 oracle (liboracle.so) plus the dense J/K
digestion and the DIIS SCF loop of the reference, in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg.  The product package never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return path


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_boys.argtypes = [ctypes.c_int, ctypes.c_double, _f64p]
        L.orc_rys_roots.argtypes = [ctypes.c_int, ctypes.c_double, _f64p, _f64p]
        L.orc_rys_roots.restype = ctypes.c_int
        for name in ("orc_eri_rys", "orc_eri_mmd"):
            f = getattr(L, name)
            f.argtypes = [_i32p, _f64p, _f64p, _f64p, _i32p, _f64p]
            f.restype = ctypes.c_double
        L.orc_hermite_e.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3
        L.orc_hermite_e.restype = ctypes.c_double
        L.orc_coulomb_r.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 5
        L.orc_coulomb_r.restype = ctypes.c_double
        L.orc_eri_3d_rys.argtypes = [_i32p, _f64p, _f64p]
        L.orc_eri_3d_rys.restype = ctypes.c_double
        basis = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _i32p, _f64p]
        L.orc_eri_tensor.argtypes = basis + [_f64p, ctypes.c_int, ctypes.c_int]
        L.orc_eri_list.argtypes = basis + [ctypes.c_long, _i32p, _f64p, ctypes.c_int, ctypes.c_int]
        L.orc_rys5_fast_init.argtypes = []
        L.orc_rys5_fast.argtypes = [ctypes.c_double, _f64p, _f64p]
        L.orc_time_rys5_fast.argtypes = [ctypes.c_long, ctypes.c_double]
        L.orc_time_rys5_fast.restype = ctypes.c_double
        L.orc_one_electron.argtypes = basis + [ctypes.c_int, _f64p, _f64p, _f64p, _f64p, _f64p,
                                               ctypes.c_int]
        _LIB = L
    return _LIB


def boys(mmax: int, T: float) -> np.ndarray:
    F = np.zeros(mmax + 1)
    lib().orc_boys(mmax, float(T), F)
    return F


def rys_roots(n: int, X: float):
    u = np.zeros(n)
    w = np.zeros(n)
    if lib().orc_rys_roots(n, float(X), u, w) != 0:
        raise ValueError("bad root count")
    return u, w


def _quartet_args(cgfs):
    npgf = np.array([c.npgf for c in cgfs], dtype=np.int32)
    coefs = np.concatenate([np.asarray(c.coefs, dtype=np.float64) for c in cgfs])
    expns = np.concatenate([np.asarray(c.expns, dtype=np.float64) for c in cgfs])
    norms = np.concatenate([np.asarray(c.normalize_factors, dtype=np.float64) for c in cgfs])
    shells = np.array([c.shell for c in cgfs], dtype=np.int32).reshape(-1)
    centers = np.array([c.center for c in cgfs], dtype=np.float64).reshape(-1)
    return npgf, coefs, expns, norms, shells, centers


def eri_rys(c1, c2, c3, c4) -> float:
    """rys.eri_rys on four contracted functions (core.f90:967-1002)."""
    return lib().orc_eri_rys(*_quartet_args([c1, c2, c3, c4]))


def eri_mmd(c1, c2, c3, c4) -> float:
    """mmd.eri on four contracted functions (core.f90:167-198)."""
    return lib().orc_eri_mmd(*_quartet_args([c1, c2, c3, c4]))


def _basis_args(mol):
    npgf, poff, coefs, expns, norms, shell, center = mol.cgf_arrays()
    return (len(npgf), npgf, poff, np.ascontiguousarray(coefs), np.ascontiguousarray(expns),
            np.ascontiguousarray(norms), shell, center)


def rys5_fast(X: float):
    """Five nodes and weights from the fitted tables (orc_rys5_fast)."""
    lib().orc_rys5_fast_init()
    u = np.zeros(5)
    w = np.zeros(5)
    lib().orc_rys5_fast(float(X), u, w)
    return u, w


def eri_tensor(mol, nthreads: int = 0, fast_roots: bool = False) -> np.ndarray:
    """Dense N^4 ERI tensor, like build_eri_parallel (hf.py:33-61).
    fast_roots: nodes from the fitted tables (orc_rys5_fast) instead of
    the exact quadrature."""
    args = _basis_args(mol)
    n = args[0]
    out = np.zeros(n ** 4)
    lib().orc_eri_tensor(*args, out, nthreads, int(fast_roots))
    return out.reshape(n, n, n, n)


def eri_list(mol, quartets: np.ndarray, nthreads: int = 0, fast_roots: bool = False) -> np.ndarray:
    q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 4)
    out = np.zeros(len(q))
    lib().orc_eri_list(*_basis_args(mol), len(q), q.reshape(-1), out, nthreads, int(fast_roots))
    return out


def one_electron(mol, nthreads: int = 0):
    """S, T, V of mmd2.py:14-84 as used by hf.py:266-274."""
    args = _basis_args(mol)
    n = args[0]
    S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    Z = np.asarray(mol.numbers, dtype=np.float64)
    xyz = np.ascontiguousarray(mol.xyz__, dtype=np.float64).reshape(-1)
    lib().orc_one_electron(*args, len(Z), Z, xyz, S, T, V, nthreads)
    return S, T, V


def jk_dense(eri: np.ndarray, P: np.ndarray):
    """J and K of hf.py:300: J = einsum('pqrs,rs'), K = einsum('prqs,rs')."""
    J = np.einsum("pqrs,rs->pq", eri, P, optimize=True)
    K = np.einsum("prqs,rs->pq", eri, P, optimize=True)
    return J, K


def scf_diis(S, Hcore, nocc: int, jk, e_nuc: float = 0.0, tol: float = 1e-8, maxiter: int = 100):
    """Restatement of scf_diis (hf.py:236-353) with the J/K call abstracted.

    ``jk(P) -> (J, K)``.  Returns (energy, P, converged, iterations).
    """
    vals, vecs = np.linalg.eigh(S)
    X = vecs @ np.diag(1.0 / np.sqrt(vals)) @ vecs.T
    F = Hcore.copy()
    Fs, res = [], []
    etot_old = float("inf")
    counter = 0
    converged = False
    while True:
        _, Cp = np.linalg.eigh(X @ F @ X)
        C = X @ Cp
        Cocc = C[:, :nocc]
        P = 2.0 * Cocc @ Cocc.T
        J, K = jk(P)
        F = Hcore + J - 0.5 * K
        etot = 0.5 * np.sum(P * (Hcore + F))
        if abs(etot - etot_old) <= tol:
            converged = True
            break
        counter += 1
        etot_old = etot
        if counter > maxiter:
            break
        if counter <= 2:
            continue
        if len(Fs) > 6:
            Fs.pop(0)
            res.pop(0)
        res.append(X @ (F @ P @ S - S @ P @ F) @ X)
        Fs.append(F)
        n = len(Fs)
        B = np.zeros((n + 1, n + 1))
        B[-1, :] = -1.0
        B[:, -1] = -1.0
        B[-1, -1] = 0.0
        for i in range(n):
            for j in range(n):
                B[i, j] += np.trace(res[i].T @ res[j])
        rhs = np.zeros(n + 1)
        rhs[-1] = -1.0
        try:
            c = np.linalg.solve(B, rhs)
        except np.linalg.LinAlgError:
            continue
        F = sum(Fs[i] * c[i] for i in range(n))
    return etot + e_nuc, P, converged, counter
