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

PAIR_CLASSES = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2), (3, 0), (3, 1), (3, 2), (3, 3)]
NUM_QUARTET_CLASSES = 55      # classes 0..20: s, p, d shells; 21..54: with an f shell
NUM_CLASSES_SPD = 21          # classes of s, p, d shells only (the rest have an f shell)
# sf_jk_profile slots: [2c], [2c + 1] the FP64 / FP32 segment of class c,
# then density scaling and reduction + finalisation
SLOT_SCALE = 2 * NUM_QUARTET_CLASSES
SLOT_REDUCE = SLOT_SCALE + 1
NUM_SLOTS = SLOT_SCALE + 2


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
    L.sf_eri_components_noalloc.argtypes = L.sf_eri_components.argtypes
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
    L.sf_grad.argtypes = [vp, vp, vp, _f64p, ctypes.c_int, vp]
    L.sf_jk_profile.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp, _f64p]
    L.sf_basis_create_batch.argtypes = [ctypes.c_int, _i32p, ctypes.c_int, _i32p, _i32p, _f64p,
                                        _f64p, _f64p, ctypes.c_int, _i32p, _i32p,
                                        ctypes.POINTER(vp)]
    L.sf_basis_set_layout.argtypes = [vp, _i64p, _i32p]
    L.sf_basis_layout.argtypes = [vp, ctypes.POINTER(ctypes.c_int), vp, vp, vp,
                                  ctypes.POINTER(ctypes.c_longlong)]
    L.sf_jk_graph.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp]
    L.sf_autotune.argtypes = [vp, vp, ctypes.c_int, vp, _i32p, _f64p]
    L.sf_set_tuning.argtypes = [vp, _i32p]
    L.sf_set_prim_tau.argtypes = [vp, ctypes.c_double]
    L.sf_set_prim_tau_fp64.argtypes = [vp, ctypes.c_double]
    L.sf_prim_quartet_counts.argtypes = [vp, ctypes.c_int, _i64p]
    L.sf_far_prim_quartet_counts.argtypes = [vp, ctypes.c_int, _i64p]
    L.sf_set_engine.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.sf_set_streams.argtypes = [vp, ctypes.c_int]
    _LIB = L
    return L


def check(code: int):
    if code != 0:
        raise EngineError(code, lib().sf_last_error().decode(errors="replace"))


def exported_symbols() -> list[str]:
    """Names declared in include/sfeng.h (for the load/export test)."""
    return ["sf_last_error", "sf_version", "sf_boys", "sf_rys_roots", "sf_eri_rys",
            "sf_eri_components", "sf_eri_components_noalloc", "sf_basis_create", "sf_basis_destroy", "sf_basis_nbf",
            "sf_schwarz", "sf_screen", "sf_quartet_counts", "sf_pair_list", "sf_quartet_list",
            "sf_eri_dense", "sf_jk", "sf_jk_profile", "sf_basis_set_nuclei", "sf_one_electron",
            "sf_grad", "sf_basis_create_batch", "sf_basis_set_layout", "sf_basis_layout",
            "sf_jk_graph", "sf_autotune", "sf_set_tuning", "sf_set_prim_tau", "sf_set_prim_tau_fp64",
            "sf_set_engine", "sf_set_streams", "sf_prim_quartet_counts",
            "sf_far_prim_quartet_counts"]


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise EngineError(-1, "no GPU visible to torch (HIP device required)")
    return torch


def _stream_ptr(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# --------------------------------------------------------- batch kernels --
def _check_vector(x, name):
    torch = _torch()
    if not isinstance(x, torch.Tensor) or x.dtype != torch.float64 or not x.is_cuda:
        raise EngineError(1, f"{name} must be a float64 CUDA tensor")
    return x.contiguous()


def boys(T, mmax: int):
    """F_m(T) for m = 0..mmax on the GPU; T is a float64 CUDA tensor."""
    torch = _torch()
    T = _check_vector(T, "T")
    F = torch.empty((T.numel(), mmax + 1), dtype=torch.float64, device=T.device)
    check(lib().sf_boys(T.numel(), mmax, ctypes.c_void_p(T.data_ptr()),
                        ctypes.c_void_p(F.data_ptr()), _stream_ptr(torch)))
    return F


def rys_roots(X, nroots: int):
    """Rys nodes (as t^2) and weights on the GPU; X is a float64 CUDA tensor."""
    torch = _torch()
    X = _check_vector(X, "X")
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


def eri_components(mol, quartets, noalloc: bool = False) -> np.ndarray:
    """Contracted component quartets (function indices) of a Molecule;
    noalloc stages through the library's static device array
    (sf_eri_components_noalloc) instead of allocating per call."""
    npgf, poff, coefs, expns, norms, shell, center = mol.cgf_arrays()
    q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 4)
    out = np.zeros(len(q))
    fn = lib().sf_eri_components_noalloc if noalloc else lib().sf_eri_components
    check(fn(len(npgf), npgf, poff, np.ascontiguousarray(coefs), np.ascontiguousarray(expns),
             np.ascontiguousarray(norms), shell, center, len(q), q.reshape(-1), out))
    return out


# ------------------------------------------- the reference's CGF list --
# Cartesian component order inside a shell (src/hf/molecule.py:138-165).
_CART = {L: [(lx, ly, L - lx - ly) for lx in range(L, -1, -1) for ly in range(L - lx, -1, -1)]
         for L in range(4)}


def _dfact2(n: int) -> float:
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def shells_from_cgfs(cgfs, rtol: float = 1e-12):
    """Group the reference's per-component function list into shells.

    ``cgfs`` is the list ``scf_diis`` builds (src/hf/hf.py:254-259): one
    ``ContractedGaussianFunction`` per Cartesian component (attributes
    ``npgf, expns, coefs, shell, center, normalize_factors``,
    src/hf/molecule.py:173-226), in the AO order of molecule.py:138-165.
    A shell is a run of consecutive components in that order with the same
    centre and exponents; every component's ``coefs * normalize_factors *
    sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!)`` must agree (to ``rtol``: the
    reference rescales one coefficient list shared by a shell's components
    once per component) and is the shell's coefficient.  Atoms are the
    distinct centres in order of first appearance.

    Returns (L, nprim, center[3*nshell], expns, coefs, shell_atom,
    atom_xyz[natom, 3]); raises ValueError when the list is not made of
    whole shells.
    """
    L_, nprim, center, expns, coefs, sh_atom = [], [], [], [], [], []
    atoms = []
    k = 0
    n = len(cgfs)
    while k < n:
        c0 = cgfs[k]
        comp = tuple(int(v) for v in c0.shell)
        L = sum(comp)
        if L not in _CART or comp != _CART[L][0]:
            raise ValueError(f"function {k}: component {comp} does not start a shell "
                             f"(expected {(L, 0, 0)})")
        nc = len(_CART[L])
        if k + nc > n:
            raise ValueError(f"function {k}: shell of L={L} is cut off by the end of the list")
        ex = np.asarray(c0.expns, dtype=np.float64)
        cen = np.asarray(c0.center, dtype=np.float64).reshape(3)
        if int(c0.npgf) != len(ex):
            raise ValueError(f"function {k}: npgf {c0.npgf} != {len(ex)} exponents")
        ref = None
        for m, want in enumerate(_CART[L]):
            c = cgfs[k + m]
            got = tuple(int(v) for v in c.shell)
            if got != want:
                raise ValueError(f"function {k + m}: component {got}, expected {want}")
            if not np.array_equal(np.asarray(c.expns, dtype=np.float64), ex):
                raise ValueError(f"function {k + m}: exponents differ from its shell's")
            if not np.array_equal(np.asarray(c.center, dtype=np.float64).reshape(3), cen):
                raise ValueError(f"function {k + m}: centre differs from its shell's")
            f2 = _dfact2(2 * got[0] - 1) * _dfact2(2 * got[1] - 1) * _dfact2(2 * got[2] - 1)
            eff = (np.asarray(c.coefs, dtype=np.float64)
                   * np.asarray(c.normalize_factors, dtype=np.float64) * np.sqrt(f2))
            if ref is None:
                ref = eff
            elif not np.allclose(eff, ref, rtol=rtol, atol=0.0):
                raise ValueError(f"function {k + m}: normalised coefficients differ from "
                                 "its shell's first component")
        for ia, a in enumerate(atoms):
            if np.array_equal(a, cen):
                break
        else:
            ia = len(atoms)
            atoms.append(cen)
        L_.append(L)
        nprim.append(len(ex))
        center.append(cen)
        expns.append(ex)
        coefs.append(ref)
        sh_atom.append(ia)
        k += nc
    return (np.array(L_, dtype=np.int32), np.array(nprim, dtype=np.int32),
            np.ascontiguousarray(np.concatenate(center)), np.ascontiguousarray(np.concatenate(expns)),
            np.ascontiguousarray(np.concatenate(coefs)), np.array(sh_atom, dtype=np.int32),
            np.array(atoms, dtype=np.float64).reshape(-1, 3))


def build_eri_parallel(cgfs, ncgf: int) -> np.ndarray:
    """Drop-in for the reference's ``build_eri_parallel(cgfs, ncgf)``
    (src/hf/hf.py:33-61, called at :264): the dense (ncgf,)*4 tensor in
    chemists' notation for the reference's own function list, every unique
    quartet computed (no screening) on the GPU."""
    if ncgf != len(cgfs):
        raise ValueError(f"ncgf = {ncgf} but the list holds {len(cgfs)} functions")
    return Basis.from_cgfs(cgfs).screen(0.0, 0.0).eri_dense()


# ------------------------------------------------------------ the basis --
def _shell_arrays(mols):
    """Concatenated shell arrays of a list of molecules (AO order kept)."""
    # per atom: its shells' radial data (Atom.shell_*), centred on the atom
    L, nprim, center, expns, coefs, atom = [], [], [], [], [], []
    mol_shell, mol_atom = [0], [0]
    for mol in mols:
        nsh = 0
        for ia, a in enumerate(mol.atoms):
            k = len(a.shell_L)
            L += a.shell_L
            nprim += a.shell_nprim
            center.append(np.broadcast_to(np.asarray(a.xyz, dtype=np.float64), (k, 3)))
            expns += a.shell_expns
            coefs += a.shell_coefs
            atom += [mol_atom[-1] + ia] * k
            nsh += k
        mol_shell.append(mol_shell[-1] + nsh)
        mol_atom.append(mol_atom[-1] + mol.natom)
    cat = lambda xs: np.ascontiguousarray(np.concatenate(xs), dtype=np.float64)  # noqa: E731
    return (np.array(L, dtype=np.int32), np.array(nprim, dtype=np.int32),
            cat(center).reshape(-1), cat(expns), cat(coefs), np.array(atom, dtype=np.int32),
            np.array(mol_shell, dtype=np.int32), np.array(mol_atom, dtype=np.int32))


class _CgfMolecule:
    """What a Basis built from a function list knows of its molecule: the
    atom centres (bohr) and, if given, their nuclear charges."""

    def __init__(self, xyz, numbers):
        self.xyz__ = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.natom = len(self.xyz__)
        if numbers is not None:
            numbers = np.asarray(numbers, dtype=np.float64).reshape(-1)
            if len(numbers) != self.natom:
                raise ValueError(f"{len(numbers)} nuclear charges for {self.natom} centres")
        self.numbers = numbers


class Basis:
    """A shell basis on the GPU: Schwarz screening, quartet lists, J/K.

    Matrices (P, J, K, S, T, V) are float64 CUDA tensors of shape
    ``self.shape``: (nbf, nbf) here, (nmol, nmax, nmax) for a BatchBasis.
    """

    def __init__(self, mol):
        self.mol = mol
        self.mols = [mol]
        L, nprim, center, expns, coefs, atom, _, _ = _shell_arrays([mol])
        self.nshell = len(L)
        h = ctypes.c_void_p()
        check(lib().sf_basis_create(self.nshell, L, nprim, center, expns, coefs, mol.natom, atom,
                                    ctypes.byref(h)))
        self._h = h
        self._finish([mol])

    @classmethod
    def from_cgfs(cls, cgfs, numbers=None) -> "Basis":
        """A basis from the reference's per-component function list
        (``cgfs`` of src/hf/hf.py:254-259; see shells_from_cgfs).  Atoms are
        the distinct function centres in order of first appearance;
        ``numbers`` (their nuclear charges, e.g. ``[a.number for a in
        mol.atoms]`` of the reference's Molecule) places point charges there
        for the one-electron integrals and gradients; without it only the
        two-electron entry points have nuclei to use (none)."""
        L, nprim, center, expns, coefs, atom, axyz = shells_from_cgfs(cgfs)
        self = cls.__new__(cls)
        self.mol = _CgfMolecule(axyz, numbers)
        self.mols = [self.mol]
        self.nshell = len(L)
        h = ctypes.c_void_p()
        check(lib().sf_basis_create(self.nshell, L, nprim, center, expns, coefs, len(axyz), atom,
                                    ctypes.byref(h)))
        self._h = h
        self._finish([self.mol])
        return self

    def _finish(self, mols):
        n = ctypes.c_int()
        check(lib().sf_basis_nbf(self._h, ctypes.byref(n)))
        self.nbf = n.value
        if all(m.numbers is not None for m in mols):
            Z = np.ascontiguousarray(np.concatenate([m.numbers for m in mols]), dtype=np.float64)
            xyz = np.ascontiguousarray(np.concatenate([np.asarray(m.xyz__, dtype=np.float64).reshape(-1, 3)
                                                       for m in mols]).reshape(-1))
            check(lib().sf_basis_set_nuclei(self._h, len(Z), Z, xyz))
        self.shape = (self.nbf, self.nbf)
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

    def _zeros(self):
        torch = _torch()
        return torch.zeros(self.shape, dtype=torch.float64, device="cuda")

    def _check_matrix(self, P, name="P"):
        torch = _torch()
        if (not isinstance(P, torch.Tensor) or P.dtype != torch.float64
                or tuple(P.shape) != tuple(self.shape) or not P.is_cuda):
            raise EngineError(1, f"{name} must be a float64 CUDA tensor of shape {self.shape}")
        return P.contiguous()

    def _check_out(self, out):
        """Output pair (J, K): float64 CUDA tensors of ``self.shape``, each
        contiguous (the library writes them in place, row-major)."""
        if not isinstance(out, (tuple, list)) or len(out) != 2:
            raise EngineError(1, "out must be a pair (J, K)")
        for name, M in zip(("J", "K"), out):
            self._check_matrix(M, name)
            if not M.is_contiguous():
                raise EngineError(1, f"{name} must be contiguous (written in place)")
        return out

    def one_electron(self):
        """S, T, V (float64 CUDA tensors of shape ``self.shape``)."""
        torch = _torch()
        S = self._zeros()
        T = self._zeros()
        V = self._zeros()
        check(lib().sf_one_electron(self._h, ctypes.c_void_p(S.data_ptr()),
                                    ctypes.c_void_p(T.data_ptr()), ctypes.c_void_p(V.data_ptr()),
                                    _stream_ptr(torch)))
        return S, T, V

    def gradient(self, P, W, precision: int = 0) -> np.ndarray:
        """Electronic nuclear gradient (natom x 3, Eh/bohr) for density P and
        energy-weighted density W (float64 CUDA tensors); precision 1 runs
        the derivative quartets below tau_mixed in FP32."""
        torch = _torch()
        P = self._check_matrix(P, "P")
        W = self._check_matrix(W, "W")
        g = np.zeros(3 * self.mol.natom)
        check(lib().sf_grad(self._h, ctypes.c_void_p(P.data_ptr()), ctypes.c_void_p(W.data_ptr()), g,
                            int(precision), _stream_ptr(torch)))
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

    def prim_quartet_counts(self, precision: int = 0) -> np.ndarray:
        """Primitive quartets the Rys kernels evaluate per class and segment
        under the current thresholds (sf_prim_quartet_counts): [55, 2]."""
        c = np.zeros(2 * NUM_QUARTET_CLASSES, dtype=np.int64)
        check(lib().sf_prim_quartet_counts(self._h, int(precision), c))
        return c.reshape(NUM_QUARTET_CLASSES, 2)

    def far_prim_quartet_counts(self, precision: int = 0) -> np.ndarray:
        """The part of prim_quartet_counts in far (ss|ss) / (ps|ss) quartets,
        evaluated in the point-charge form (sf_far_prim_quartet_counts)."""
        c = np.zeros(2 * NUM_QUARTET_CLASSES, dtype=np.int64)
        check(lib().sf_far_prim_quartet_counts(self._h, int(precision), c))
        return c.reshape(NUM_QUARTET_CLASSES, 2)

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
        """J, K for the density P (float64 CUDA tensor of shape ``self.shape``)."""
        torch = _torch()
        P = self._check_matrix(P)
        J, K = (self._zeros(), self._zeros()) if out is None else self._check_out(out)
        check(lib().sf_jk(self._h, ctypes.c_void_p(P.data_ptr()), ctypes.c_void_p(J.data_ptr()),
                          ctypes.c_void_p(K.data_ptr()), int(precision), _stream_ptr(torch)))
        return J, K

    def jk_graph(self, P, out, precision: int = 0):
        """jk() replayed from a captured hipGraph (sf_jk_graph).  P and the
        output pair ``out = (J, K)`` must stay the same tensors between calls
        for the capture to be reused."""
        torch = _torch()
        if not P.is_contiguous():
            raise EngineError(1, "P must be contiguous (the graph keeps its address)")
        self._check_matrix(P)
        J, K = self._check_out(out)
        check(lib().sf_jk_graph(self._h, ctypes.c_void_p(P.data_ptr()),
                                ctypes.c_void_p(J.data_ptr()), ctypes.c_void_p(K.data_ptr()),
                                int(precision), _stream_ptr(torch)))
        return J, K

    def autotune(self, P, precision: int = 0):
        """Pick launch shapes per quartet class (sf_autotune).  Returns
        (shapes[55, 3, 3] as (block, grid, arith) per tuning slot -- 0: FP64
        builds, 1/2: FP64/FP32 segment of mixed builds -- and ms[55, 3],
        zero for the slots not tuned by this call)."""
        torch = _torch()
        P = self._check_matrix(P)
        shapes = np.zeros(9 * NUM_QUARTET_CLASSES, dtype=np.int32)
        ms = np.zeros(3 * NUM_QUARTET_CLASSES)
        check(lib().sf_autotune(self._h, ctypes.c_void_p(P.data_ptr()), int(precision),
                                _stream_ptr(torch), shapes, ms))
        torch.cuda.synchronize()
        return shapes.reshape(NUM_QUARTET_CLASSES, 3, 3), ms.reshape(NUM_QUARTET_CLASSES, 3)

    def set_prim_tau(self, tau_prim: float) -> "Basis":
        """Primitive-quartet threshold of the FP32 quartets (< 0: default
        1e-8 * tau_mixed, 0: off); see sf_set_prim_tau."""
        check(lib().sf_set_prim_tau(self._h, float(tau_prim)))
        return self

    def set_prim_tau_fp64(self, tau_prim: float) -> "Basis":
        """Primitive-quartet threshold of the FP64 quartets (< 0: default
        1e-3 * tau, 0: off); see sf_set_prim_tau_fp64."""
        check(lib().sf_set_prim_tau_fp64(self._h, float(tau_prim)))
        return self

    def set_engine(self, engine: int, qcls: int = -1) -> "Basis":
        """Integral engine of one quartet class (or all, qcls = -1): 0 = Rys,
        1 = McMurchie-Davidson with MFMA E.R.E transforms (sf_set_engine;
        s, p, d classes only -- qcls = -1 leaves the f-shell classes on Rys)."""
        check(lib().sf_set_engine(self._h, int(qcls), int(engine)))
        return self

    def set_streams(self, nside: int) -> "Basis":
        """Side streams of jk()/jk_graph() (0..3): class launches and
        reductions spread over the current stream and nside more; 0 runs
        them all in order on the current stream (sf_set_streams)."""
        check(lib().sf_set_streams(self._h, int(nside)))
        return self

    def set_tuning(self, shapes):
        """Launch shapes of every class and tuning slot, as autotune()
        returns them: (NUM_QUARTET_CLASSES, 3, 3) ints (sf_set_tuning)."""
        flat = np.ascontiguousarray(shapes, dtype=np.int32).reshape(-1)
        if flat.size != 9 * NUM_QUARTET_CLASSES:
            raise EngineError(1, f"set_tuning: {flat.size} entries, expected "
                                 f"{9 * NUM_QUARTET_CLASSES} (classes x slots x (block, grid, arith))")
        check(lib().sf_set_tuning(self._h, flat))

    def jk_profile(self, P, precision: int = 0, out=None):
        """jk() plus per-launch-group milliseconds (HIP events on the stream):
        returns J, K, ms[NUM_SLOTS] (see sf_jk_profile in include/sfeng.h)."""
        torch = _torch()
        P = self._check_matrix(P)
        J, K = (self._zeros(), self._zeros()) if out is None else self._check_out(out)
        ms = np.zeros(NUM_SLOTS)
        check(lib().sf_jk_profile(self._h, ctypes.c_void_p(P.data_ptr()),
                                  ctypes.c_void_p(J.data_ptr()), ctypes.c_void_p(K.data_ptr()),
                                  int(precision), _stream_ptr(torch), ms))
        return J, K, ms


class BatchBasis(Basis):
    """Many independent molecules in one handle (sf_basis_create_batch).

    Matrices use a padded layout: tensors of shape (nmol, nmax, nmax),
    molecule m's n_m x n_m block in the top-left corner of slice m; the
    padding is never read or written by the library.
    """

    def __init__(self, mols):
        self.mols = list(mols)
        self.mol = None
        L, nprim, center, expns, coefs, atom, mol_shell, mol_atom = _shell_arrays(self.mols)
        self.nshell = len(L)
        h = ctypes.c_void_p()
        check(lib().sf_basis_create_batch(len(self.mols), mol_shell, self.nshell, L, nprim, center,
                                          expns, coefs, int(mol_atom[-1]), mol_atom, atom,
                                          ctypes.byref(h)))
        self._h = h
        self._finish(self.mols)
        nmol = ctypes.c_int()
        nbf_mol = np.zeros(len(self.mols), dtype=np.int32)
        check(lib().sf_basis_layout(self._h, ctypes.byref(nmol), nbf_mol.ctypes.data_as(ctypes.c_void_p),
                                    None, None, ctypes.byref(ctypes.c_longlong())))
        self.nmol = nmol.value
        self.nbf_mol = nbf_mol
        self.nmax = int(nbf_mol.max())
        off = np.arange(self.nmol, dtype=np.int64) * self.nmax * self.nmax
        ld = np.full(self.nmol, self.nmax, dtype=np.int32)
        check(lib().sf_basis_set_layout(self._h, off, ld))
        self.shape = (self.nmol, self.nmax, self.nmax)

    def gradient(self, P, W, precision: int = 0):
        raise EngineError(5, "gradients are computed per molecule (Basis)")

    def eri_dense(self):
        raise EngineError(5, "the dense ERI tensor is computed per molecule (Basis)")

    def schwarz(self):
        raise EngineError(5, "Schwarz factors of a batch: use pair_list()")
