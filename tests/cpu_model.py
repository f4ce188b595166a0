"""CPU model of the GPU J/K pipeline, for tests.

Drives the kernel math headers compiled for the host (tests/harness) through
the same steps as libsfeng's sf_screen / sf_jk: canonical shell pairs,
Schwarz factors, Q-sorted significant pairs, per-bra ket prefixes, digestion
tiles, and J = D (X + X^T) D, K = D (Y + Y^T) D.  It checks the algorithm
(degeneracy factors, tile bookkeeping, list construction) against the
oracle on the CPU, and gives the CPU list builder that the GPU lists must
match bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np

from conftest import build_harness

_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
PAIR_CLASSES = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
_H = None


def harness():
    global _H
    if _H is None:
        H = ctypes.CDLL(build_harness())
        H.h_eri.argtypes = [ctypes.c_int] * 4 + [_f64p, ctypes.c_int, _f64p, ctypes.c_int, _f64p,
                                                 _f64p, _f64p, _f64p, _f64p, ctypes.c_int]
        H.h_rys_kernel.argtypes = [ctypes.c_int, ctypes.c_double, _f64p, _f64p]
        H.h_digest.argtypes = [ctypes.c_int] * 4 + [_f64p, _f64p, ctypes.c_int, _i32p, _f64p]
        H.h_eri_split.argtypes = H.h_eri.argtypes
        H.h_eri_split_pairs.argtypes = H.h_eri.argtypes
        H.h_digest_tiles.argtypes = [ctypes.c_int] * 4 + [_f64p, ctypes.c_double, _f64p,
                                                          ctypes.c_int, _i32p, _f64p]
        H.h_digest_tile_rows.argtypes = H.h_digest_tiles.argtypes
        H.h_digest_tile_groups.argtypes = H.h_digest_tiles.argtypes
        H.h_rys.argtypes = [ctypes.c_int, ctypes.c_double, _f64p, _f64p]
        H.h_boys.argtypes = [ctypes.c_int, ctypes.c_double, _f64p]
        H.h_boys01.argtypes = [ctypes.c_double, _f64p]
        H.h_boys01_kernel.argtypes = [ctypes.c_double, _f64p]
        H.h_grad.argtypes = [ctypes.c_int] * 4 + [_f64p, ctypes.c_int, _f64p, ctypes.c_int, _f64p,
                                                  _f64p, _f64p, _f64p, _f64p, _f64p]
        H.h_grad_slices.argtypes = H.h_grad.argtypes
        H.h_grad_onee.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int, _f64p, _f64p,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, _f64p, _f64p,
                                  _f64p, ctypes.c_double, _f64p]
        H.h_eri_gen.argtypes = [ctypes.c_int] * 4 + [_f64p, ctypes.c_int, _f64p, ctypes.c_int,
                                                 _f64p, _f64p, _f64p, _f64p, _f64p]
        H.h_gen_tiles.argtypes = H.h_digest_tiles.argtypes
        H.h_eri_bra_runtime.argtypes = H.h_eri_gen.argtypes
        H.h_eri_far.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, _f64p, ctypes.c_int, _f64p,
                                _f64p]
        H.h_one_electron.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int, _f64p, _f64p,
                                     ctypes.c_int, _f64p, _f64p, _f64p, _f64p, _f64p]
        _H = H
    return _H


def ncart(L):
    return (L + 1) * (L + 2) // 2


def prim_pairs(sa, sb):
    A, B = sa.center, sb.center
    AB2 = float(((A - B) ** 2).sum())
    out = []
    for a, ca in zip(sa.expns, sa.coefs):
        for b, cb in zip(sb.expns, sb.coefs):
            p = a + b
            P = (a * A + b * B) / p
            K = ca * cb * np.exp(-a * b / p * AB2)
            out.append([p, *P, K, 1.0 / p, K / p, b])
    return np.ascontiguousarray(np.array(out).reshape(-1))


def shell_quartet(shells, a, b, c, d, fp32=False, split=False, pairs=False):
    """Integrals of one shell quartet through the kernels' code paths: the
    full-class path, the split path (per component of A where the class
    allows, as the GPU), or with pairs the per-bra-component-pair path."""
    S = [shells[a], shells[b], shells[c], shells[d]]
    out = np.zeros(int(np.prod([s.nfunc for s in S])))
    H = harness()
    fn = H.h_eri_split_pairs if pairs else H.h_eri_split if split else H.h_eri
    rc = fn(S[0].L, S[1].L, S[2].L, S[3].L, prim_pairs(S[0], S[1]),
                         S[0].nprim * S[1].nprim, prim_pairs(S[2], S[3]), S[2].nprim * S[3].nprim,
                         S[0].center, S[1].center, S[2].center, S[3].center, out, int(fp32))
    assert rc == 0
    return out


def uses_split(la, lb, lc, ld):
    """Classes the GPU computes one bra component pair per wavefront (more
    than 54 integrals per quartet); the rule of QuartetShape::SPLIT."""
    return ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) > 54


def canonical_pairs(shells):
    """All pairs a >= b, oriented so L(sa) >= L(sb); grouped by pair class."""
    groups = {pc: [] for pc in range(6)}
    for a in range(len(shells)):
        for b in range(a + 1):
            sa, sb = (a, b) if shells[a].L >= shells[b].L else (b, a)
            la, lb = shells[sa].L, shells[sb].L
            groups[la * (la + 1) // 2 + lb].append((sa, sb))
    return groups


def schwarz(shells, groups):
    Q = {}
    for pc, pairs in groups.items():
        for sa, sb in pairs:
            v = shell_quartet(shells, sa, sb, sa, sb)
            na, nb = shells[sa].nfunc, shells[sb].nfunc
            v = v.reshape(na, nb, na, nb)
            Q[(sa, sb)] = float(np.sqrt(max(abs(v[i, j, i, j]) for i in range(na) for j in range(nb))))
    return Q


def sorted_significant(groups, Q, tau):
    qmax = max(Q.values())
    out = {}
    for pc, pairs in groups.items():
        keep = [p for p in pairs if Q[p] * qmax >= tau]
        keep = sorted(keep, key=lambda p: -Q[p])  # stable: ties keep input order
        out[pc] = keep
    return out


def ket_counts(Qb, Qk, same, tau, tau_mixed):
    """The CPU quartet-list builder: per bra, surviving ket prefix length and
    its FP64 part (the GPU builder must agree exactly)."""
    Qb = np.asarray(Qb, dtype=np.float64)
    Qk = np.asarray(Qk, dtype=np.float64)
    n_all = np.zeros(len(Qb), dtype=np.int32)
    n64 = np.zeros(len(Qb), dtype=np.int32)
    for i, qi in enumerate(Qb):
        prod = qi * Qk
        n = int(np.sum(prod >= tau))
        # prefix property: products are non-increasing along the sorted kets
        assert np.all(prod[:n] >= tau)
        if same:
            n = min(n, i + 1)
        m = n
        if tau_mixed > tau:
            m = min(int(np.sum(prod >= tau_mixed)), n)
        n_all[i], n64[i] = n, m
    return n_all, n64


def jk_model(mol, P, tau=1e-12, tau_mixed=0.0, fp32_below=False):
    shells = mol.shells()
    fn = mol.function_norms()
    groups = canonical_pairs(shells)
    Q = schwarz(shells, groups)
    lists = sorted_significant(groups, Q, tau)
    n = mol.nbf
    Ps = np.ascontiguousarray(P * fn[:, None] * fn[None, :])
    X = np.zeros((n, n))
    Y = np.zeros((n, n))
    for pcb in range(6):
        for pck in range(pcb + 1):
            bra, ket = lists[pcb], lists[pck]
            if not bra or not ket:
                continue
            same = pcb == pck
            n_all, n64 = ket_counts([Q[p] for p in bra], [Q[p] for p in ket], same, tau, tau_mixed)
            for i, (a, b) in enumerate(bra):
                for j in range(n_all[i]):
                    c, d = ket[j]
                    Ls = [shells[s].L for s in (a, b, c, d)]
                    split = uses_split(*Ls)
                    v = shell_quartet(shells, a, b, c, d, fp32=fp32_below and j >= n64[i], split=split)
                    fac = (0.5 if a == b else 1.0) * (0.5 if c == d else 1.0) * (0.5 if same and i == j else 1.0)
                    nA, nB, nC, nD = [ncart(L) for L in Ls]
                    t = np.zeros(nA * nB + nC * nD + nA * nC + nA * nD + nB * nC + nB * nD)
                    f0 = np.array([shells[s].offset for s in (a, b, c, d)], dtype=np.int32)
                    if split:
                        assert harness().h_digest_tiles(*Ls, np.ascontiguousarray(v), fac, Ps, n, f0, t) == 0
                    else:
                        v = np.ascontiguousarray(v * fac)
                        assert harness().h_digest(*Ls, v, Ps, n, f0, t) == 0
                    o = 0
                    for (s1, s2, n1, n2, M) in ((a, b, nA, nB, X), (c, d, nC, nD, X), (a, c, nA, nC, Y),
                                                (a, d, nA, nD, Y), (b, c, nB, nC, Y), (b, d, nB, nD, Y)):
                        r, q0 = shells[s1].offset, shells[s2].offset
                        M[r:r + n1, q0:q0 + n2] += t[o:o + n1 * n2].reshape(n1, n2)
                        o += n1 * n2
    D = fn[:, None] * fn[None, :]
    return D * (X + X.T), D * (Y + Y.T)


def one_electron_model(mol):
    """S, T, V assembled from the kernels' per-shell-pair routine on the CPU."""
    shells = mol.shells()
    fn = mol.function_norms()
    n = mol.nbf
    S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    Z = np.ascontiguousarray(mol.numbers, dtype=np.float64)
    C = np.ascontiguousarray(mol.xyz__, dtype=np.float64).reshape(-1)
    for a in range(len(shells)):
        for b in range(a + 1):
            sa, sb = (a, b) if shells[a].L >= shells[b].L else (b, a)
            A, B = shells[sa], shells[sb]
            nA, nB = A.nfunc, B.nfunc
            s, t, v = np.zeros(nA * nB), np.zeros(nA * nB), np.zeros(nA * nB)
            assert harness().h_one_electron(A.L, B.L, prim_pairs(A, B), A.nprim * B.nprim, A.center,
                                            B.center, len(Z), Z, C, s, t, v) == 0
            r, c = A.offset, B.offset
            for M, x in ((S, s), (T, t), (V, v)):
                blk = x.reshape(nA, nB) * fn[r:r + nA, None] * fn[None, c:c + nB]
                M[r:r + nA, c:c + nB] = blk
                M[c:c + nB, r:r + nA] = blk.T
    return S, T, V


def quartet_grad(shells, a, b, c, d, Dfac, slices=False):
    """sum_{ijkl} Dfac * d(ij|kl)/d(A,B,C) from the kernels' derivative code
    (their decomposition, or with slices one bra component pair at a time)."""
    S = [shells[x] for x in (a, b, c, d)]
    g = np.zeros(9)
    fn = harness().h_grad_slices if slices else harness().h_grad
    assert fn(S[0].L, S[1].L, S[2].L, S[3].L, prim_pairs(S[0], S[1]),
                            S[0].nprim * S[1].nprim, prim_pairs(S[2], S[3]), S[2].nprim * S[3].nprim,
                            S[0].center, S[1].center, S[2].center, S[3].center,
                            np.ascontiguousarray(Dfac, dtype=np.float64), g) == 0
    return g


def one_electron_grad_model(mol, P, W):
    """d/dR [sum P (T + V) - sum W S] per atom (natom x 3), from the kernels'
    per-pair routine on the CPU."""
    shells = mol.shells()
    fn = mol.function_norms()
    Ps = P * fn[:, None] * fn[None, :]
    Ws = W * fn[:, None] * fn[None, :]
    Z = np.ascontiguousarray(mol.numbers, dtype=np.float64)
    C = np.ascontiguousarray(mol.xyz__, dtype=np.float64).reshape(-1)
    g = np.zeros(3 * mol.natom)
    for a in range(len(shells)):
        for b in range(a + 1):
            sa, sb = (a, b) if shells[a].L >= shells[b].L else (b, a)
            A, B = shells[sa], shells[sb]
            r, c = A.offset, B.offset
            Pb = np.ascontiguousarray(Ps[r:r + A.nfunc, c:c + B.nfunc])
            Wb = np.ascontiguousarray(Ws[r:r + A.nfunc, c:c + B.nfunc])
            row = np.zeros(3 * mol.natom)
            assert harness().h_grad_onee(A.L, B.L, prim_pairs(A, B), A.nprim * B.nprim, A.center,
                                         B.center, A.atom, B.atom, mol.natom, Z, C, Pb, Wb,
                                         1.0 if sa == sb else 2.0, row) == 0
            g += row
    return g.reshape(-1, 3)


def two_electron_grad_model(mol, P):
    """d/dR of E2 = 1/2 sum P_pq P_rs [(pq|rs) - 1/2 (pr|qs)] per atom, over
    the unique shell quartets (unscreened) with the kernels' derivative code."""
    shells = mol.shells()
    fn = mol.function_norms()
    Ps = P * fn[:, None] * fn[None, :]
    groups = canonical_pairs(shells)
    pairs = [p for pc in range(6) for p in groups[pc]]
    g = np.zeros((mol.natom, 3))
    for ib, (a, b) in enumerate(pairs):
        for (c, d) in pairs[:ib + 1]:
            la, lb, lc, ld = (shells[x].L for x in (a, b, c, d))
            # canonical class order: bra pair class >= ket pair class
            if (la * (la + 1) // 2 + lb) < (lc * (lc + 1) // 2 + ld):
                A_, B_, C_, D_ = c, d, a, b
            else:
                A_, B_, C_, D_ = a, b, c, d
            sh = [shells[x] for x in (A_, B_, C_, D_)]
            fac = (0.5 if A_ == B_ else 1.0) * (0.5 if C_ == D_ else 1.0) * (0.5 if {A_, B_} == {C_, D_} and (A_, B_) == (C_, D_) else 1.0)
            o = [x.offset for x in sh]
            n = [x.nfunc for x in sh]
            Pab = Ps[o[0]:o[0] + n[0], o[1]:o[1] + n[1]]
            Pcd = Ps[o[2]:o[2] + n[2], o[3]:o[3] + n[3]]
            Pac = Ps[o[0]:o[0] + n[0], o[2]:o[2] + n[2]]
            Pbd = Ps[o[1]:o[1] + n[1], o[3]:o[3] + n[3]]
            Pad = Ps[o[0]:o[0] + n[0], o[3]:o[3] + n[3]]
            Pbc = Ps[o[1]:o[1] + n[1], o[2]:o[2] + n[2]]
            Dfac = fac * (4 * np.einsum('ab,cd->abcd', Pab, Pcd) - np.einsum('ac,bd->abcd', Pac, Pbd)
                          - np.einsum('ad,bc->abcd', Pad, Pbc))
            gq = quartet_grad(shells, A_, B_, C_, D_, Dfac.reshape(-1))
            gD = -(gq[0:3] + gq[3:6] + gq[6:9])
            for k, x in enumerate((A_, B_, C_)):
                g[shells[x].atom] += gq[3 * k:3 * k + 3]
            g[shells[D_].atom] += gD
    return g
