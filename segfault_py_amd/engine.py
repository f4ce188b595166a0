"""ctypes front-end of libsfeng.so (the HIP ERI / J-K engine).

This is the Python side of the drop-in boundary: the reference imports its
integral engine as ``from core.core import rys`` and calls
``rys.eri_rys(...)`` per quartet (src/hf/hf.py:14-15, :26-30), builds the
N^4 tensor with ``build_eri_parallel`` (hf.py:33-61) and digests it with two
einsums (hf.py:300).  Here:

* ``rys.eri_rys(...)`` -- same argument order and meaning, computed on the GPU;
* ``Basis(mol).screen(tau).jk(P)`` -- direct-SCF J and K (device tensors);
* ``Basis(mol).eri_dense()`` -- the N^4 tensor for small N.

The library must be built (``python -m segfault_py_amd.build``); there is no
CPU fallback: a missing library or GPU raises ``EngineError``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsfeng.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

NUM_QUARTET_CLASSES = 21
PAIR_CLASSES = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]


def quartet_class_ls(c: int):
    """(la, lb, lc, ld) of quartet class c."""
    pcb = int((np.sqrt(8 * c + 1) - 1) // 2)
    pck = c - pcb * (pcb + 1) // 2
    return PAIR_CLASSES[pcb] + PAIR_CLASSES[pck]


class EngineError(RuntimeError):
    """A non-zero status from libsfeng, with its message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libsfeng status {code}: {msg}")
        self.code = code


_LIB = None


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise EngineError(-1, f"{LIB_PATH} is not built (python -m segfault_py_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    L.sf_last_error.restype = ctypes.c_char_p
    L.sf_version.restype = ctypes.c_int
    vp = ctypes.c_void_p
    L.sf_boys.argtypes = [ctypes.c_longlong, ctypes.c_int, vp, vp, vp]
    L.sf_rys_roots.argtypes = [ctypes.c_longlong, ctypes.c_int, vp, vp, vp, vp]
    L.sf_eri_rys.argtypes = [_i32p, _f64p, _f64p, _f64p, _i32p, _f64p, _f64p]
    L.sf_eri_components.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _i32p, _f64p,
                                    ctypes.c_longlong, _i32p, _f64p]
    L.sf_basis_create.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, ctypes.c_int,
                                  _i32p, ctypes.POINTER(vp)]
    L.sf_basis_destroy.argtypes = [vp]
    L.sf_basis_nbf.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.sf_schwarz.argtypes = [vp, _f64p]
    L.sf_screen.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.sf_quartet_counts.argtypes = [vp, _i64p, ctypes.POINTER(ctypes.c_longlong),
                                    ctypes.POINTER(ctypes.c_longlong)]
    L.sf_pair_list.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), vp, vp, vp, vp]
    L.sf_quartet_list.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), vp, vp]
    L.sf_eri_dense.argtypes = [vp, _f64p]
    L.sf_jk.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp]
    L.sf_basis_set_nuclei.argtypes = [vp, ctypes.c_int, _f64p, _f64p]
    L.sf_one_electron.argtypes = [vp, vp, vp, vp, vp]
    L.sf_grad.argtypes = [vp, vp, vp, _f64p, vp]
    L.sf_jk_profile.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp, _f64p]
    _LIB = L
    return L


def check(code: int):
    if code != 0:
        raise EngineError(code, lib().sf_last_error().decode(errors="replace"))


def exported_symbols() -> list[str]:
    """Names declared in include/sfeng.h (for the load/export test)."""
    return ["sf_last_error", "sf_version", "sf_boys", "sf_rys_roots", "sf_eri_rys",
            "sf_eri_components", "sf_basis_create", "sf_basis_destroy", "sf_basis_nbf",
            "sf_schwarz", "sf_screen", "sf_quartet_counts", "sf_pair_list", "sf_quartet_list",
            "sf_eri_dense", "sf_jk", "sf_jk_profile", "sf_basis_set_nuclei", "sf_one_electron",
            "sf_grad"]


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise EngineError(-1, "no GPU visible to torch (HIP device required)")
    return torch


def _stream_ptr(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# --------------------------------------------------------- batch kernels --
def boys(T, mmax: int):
    """F_m(T) for m = 0..mmax on the GPU; T is a float64 CUDA tensor."""
    torch = _torch()
    T = T.contiguous()
    F = torch.empty((T.numel(), mmax + 1), dtype=torch.float64, device=T.device)
    check(lib().sf_boys(T.numel(), mmax, ctypes.c_void_p(T.data_ptr()),
                        ctypes.c_void_p(F.data_ptr()), _stream_ptr(torch)))
    return F


def rys_roots(X, nroots: int):
    """Rys nodes (as t^2) and weights on the GPU; X is a float64 CUDA tensor."""
    torch = _torch()
    X = X.contiguous()
    u = torch.empty((X.numel(), nroots), dtype=torch.float64, device=X.device)
    w = torch.empty_like(u)
    check(lib().sf_rys_roots(X.numel(), nroots, ctypes.c_void_p(X.data_ptr()),
                             ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(w.data_ptr()),
                             _stream_ptr(torch)))
    return u, w


# ------------------------------------------ reference-compatible quartets --
class _Rys:
    """Drop-in for ``from core.core import rys`` (core.f90:967-1002)."""

    @staticmethod
    def eri_rys(coefs1, coefs2, coefs3, coefs4, expns1, expns2, expns3, expns4,
                shell1, shell2, shell3, shell4, normalize_factors1, normalize_factors2,
                normalize_factors3, normalize_factors4, center1, center2, center3, center4) -> float:
        coefs = [coefs1, coefs2, coefs3, coefs4]
        npgf = np.array([len(c) for c in coefs], dtype=np.int32)
        out = np.zeros(1)
        check(lib().sf_eri_rys(
            npgf,
            np.ascontiguousarray(np.concatenate(coefs), dtype=np.float64),
            np.ascontiguousarray(np.concatenate([expns1, expns2, expns3, expns4]), dtype=np.float64),
            np.ascontiguousarray(np.concatenate([normalize_factors1, normalize_factors2,
                                                 normalize_factors3, normalize_factors4]),
                                 dtype=np.float64),
            np.ascontiguousarray(np.concatenate([shell1, shell2, shell3, shell4]), dtype=np.int32),
            np.ascontiguousarray(np.concatenate([center1, center2, center3, center4]),
                                 dtype=np.float64),
            out))
        return float(out[0])


rys = _Rys()


def eri_components(mol, quartets) -> np.ndarray:
    """Contracted component quartets (function indices) of a Molecule."""
    npgf, poff, coefs, expns, norms, shell, center = mol.cgf_arrays()
    q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 4)
    out = np.zeros(len(q))
    check(lib().sf_eri_components(len(npgf), npgf, poff, np.ascontiguousarray(coefs),
                                  np.ascontiguousarray(expns), np.ascontiguousarray(norms), shell,
                                  center, len(q), q.reshape(-1), out))
    return out


# ------------------------------------------------------------ the basis --
class Basis:
    """A shell basis on the GPU: Schwarz screening, quartet lists, J/K."""

    def __init__(self, mol):
        shells = mol.shells()
        self.mol = mol
        self.nshell = len(shells)
        L = np.array([s.L for s in shells], dtype=np.int32)
        nprim = np.array([s.nprim for s in shells], dtype=np.int32)
        center = np.ascontiguousarray(np.concatenate([s.center for s in shells]), dtype=np.float64)
        expns = np.ascontiguousarray(np.concatenate([s.expns for s in shells]), dtype=np.float64)
        coefs = np.ascontiguousarray(np.concatenate([s.coefs for s in shells]), dtype=np.float64)
        atom = np.array([s.atom for s in shells], dtype=np.int32)
        h = ctypes.c_void_p()
        check(lib().sf_basis_create(self.nshell, L, nprim, center, expns, coefs, mol.natom, atom,
                                    ctypes.byref(h)))
        self._h = h
        n = ctypes.c_int()
        check(lib().sf_basis_nbf(self._h, ctypes.byref(n)))
        self.nbf = n.value
        check(lib().sf_basis_set_nuclei(self._h, mol.natom,
                                        np.ascontiguousarray(mol.numbers, dtype=np.float64),
                                        np.ascontiguousarray(mol.xyz__, dtype=np.float64).reshape(-1)))
        self.tau = None
        self.tau_mixed = None

    def close(self):
        if getattr(self, "_h", None):
            lib().sf_basis_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def one_electron(self):
        """S, T, V (float64 CUDA tensors, N x N)."""
        torch = _torch()
        S = torch.empty((self.nbf, self.nbf), dtype=torch.float64, device="cuda")
        T = torch.empty_like(S)
        V = torch.empty_like(S)
        check(lib().sf_one_electron(self._h, ctypes.c_void_p(S.data_ptr()),
                                    ctypes.c_void_p(T.data_ptr()), ctypes.c_void_p(V.data_ptr()),
                                    _stream_ptr(torch)))
        return S, T, V

    def gradient(self, P, W) -> np.ndarray:
        """Electronic nuclear gradient (natom x 3, Eh/bohr) for density P and
        energy-weighted density W (float64 CUDA tensors)."""
        torch = _torch()
        P = P.contiguous()
        W = W.contiguous()
        g = np.zeros(3 * self.mol.natom)
        check(lib().sf_grad(self._h, ctypes.c_void_p(P.data_ptr()), ctypes.c_void_p(W.data_ptr()), g,
                            _stream_ptr(torch)))
        return g.reshape(-1, 3)

    def schwarz(self) -> np.ndarray:
        """Q[a*(a+1)/2 + b] for shell pairs a >= b."""
        Q = np.zeros(self.nshell * (self.nshell + 1) // 2)
        check(lib().sf_schwarz(self._h, Q))
        return Q

    def screen(self, tau: float = 1e-12, tau_mixed: float = 0.0) -> "Basis":
        check(lib().sf_screen(self._h, float(tau), float(tau_mixed)))
        self.tau, self.tau_mixed = tau, tau_mixed
        return self

    def counts(self):
        """Per quartet class [surviving, FP64, FP32] shell quartets, plus the
        surviving unique function quartets and primitive quartets."""
        c = np.zeros(3 * NUM_QUARTET_CLASSES, dtype=np.int64)
        fq, pq = ctypes.c_longlong(), ctypes.c_longlong()
        check(lib().sf_quartet_counts(self._h, c, ctypes.byref(fq), ctypes.byref(pq)))
        return c.reshape(NUM_QUARTET_CLASSES, 3), fq.value, pq.value

    def pair_list(self, pc: int, with_nprim: bool = False):
        """Sorted significant pairs of pair class pc: (sa, sb, Q[, nprim])."""
        n = ctypes.c_int()
        check(lib().sf_pair_list(self._h, pc, ctypes.byref(n), None, None, None, None))
        sa = np.zeros(max(n.value, 1), dtype=np.int32)
        sb = np.zeros_like(sa)
        npp = np.zeros_like(sa)
        Q = np.zeros(max(n.value, 1))
        vp = ctypes.c_void_p
        check(lib().sf_pair_list(self._h, pc, ctypes.byref(n), sa.ctypes.data_as(vp),
                                 sb.ctypes.data_as(vp), Q.ctypes.data_as(vp),
                                 npp.ctypes.data_as(vp)))
        k = n.value
        if with_nprim:
            return sa[:k], sb[:k], Q[:k], npp[:k]
        return sa[:k], sb[:k], Q[:k]

    def quartet_list(self, c: int):
        nb = ctypes.c_int()
        check(lib().sf_quartet_list(self._h, c, ctypes.byref(nb), None, None))
        na = np.zeros(max(nb.value, 1), dtype=np.int32)
        n64 = np.zeros_like(na)
        check(lib().sf_quartet_list(self._h, c, ctypes.byref(nb), na.ctypes.data_as(ctypes.c_void_p),
                                    n64.ctypes.data_as(ctypes.c_void_p)))
        return na[:nb.value], n64[:nb.value]

    def eri_dense(self) -> np.ndarray:
        n = self.nbf
        out = np.zeros(n ** 4)
        check(lib().sf_eri_dense(self._h, out))
        return out.reshape(n, n, n, n)

    def jk(self, P, precision: int = 0, out=None):
        """J, K for the density P (float64 CUDA tensor, N x N)."""
        torch = _torch()
        P = P.contiguous()
        if P.dtype != torch.float64 or P.shape != (self.nbf, self.nbf) or not P.is_cuda:
            raise EngineError(1, "P must be a float64 CUDA tensor of shape (nbf, nbf)")
        if out is None:
            J = torch.empty_like(P)
            K = torch.empty_like(P)
        else:
            J, K = out
        check(lib().sf_jk(self._h, ctypes.c_void_p(P.data_ptr()), ctypes.c_void_p(J.data_ptr()),
                          ctypes.c_void_p(K.data_ptr()), int(precision), _stream_ptr(torch)))
        return J, K

    def jk_profile(self, P, precision: int = 0, out=None):
        """jk() plus per-launch-group milliseconds (HIP events on the stream):
        returns J, K, ms[44] (see sf_jk_profile in include/sfeng.h)."""
        torch = _torch()
        P = P.contiguous()
        J = torch.empty_like(P) if out is None else out[0]
        K = torch.empty_like(P) if out is None else out[1]
        ms = np.zeros(44)
        check(lib().sf_jk_profile(self._h, ctypes.c_void_p(P.data_ptr()),
                                  ctypes.c_void_p(J.data_ptr()), ctypes.c_void_p(K.data_ptr()),
                                  int(precision), _stream_ptr(torch), ms))
        return J, K, ms
