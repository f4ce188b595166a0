"""The HIP kernels' math headers (sf_math.hpp, eri_core.hpp) compiled for
the CPU and checked against mpmath golden vectors and the oracle, and the
J/K pipeline modelled on the CPU against the oracle's dense einsum."""
import json
import os

import numpy as np
import pytest

import cpu_model as CM
from conftest import GOLDEN


def _golden(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


def test_rys_tables_match_mpmath():
    H = CM.harness()
    worst_u = worst_w = 0.0
    for pt in _golden("rys_mpmath.json")["points"]:
        n = pt["n"]
        u, w = np.zeros(n), np.zeros(n)
        assert H.h_rys(n, float(pt["X"]), u, w) == 0
        ur = np.array([float(v) for v in pt["u"]])
        wr = np.array([float(v) for v in pt["w"]])
        worst_u = max(worst_u, float(np.max(np.abs(u - ur) / ur)))
        worst_w = max(worst_w, float(np.max(np.abs(w - wr) / wr)))
    assert worst_u < 1e-14 and worst_w < 1e-14, (worst_u, worst_w)


def test_kernel_node_rules_match_mpmath():
    """The node rules inside the ERI kernels (one node from F0/F1 of the
    Boys grid, two to five from the Chebyshev tables) against 50-digit
    mpmath nodes and weights; the one-node rule also against the Chebyshev
    rule over a dense range of X."""
    H = CM.harness()
    worst_u = worst_w = 0.0
    for pt in _golden("rys_mpmath.json")["points"]:
        n = pt["n"]
        u, w = np.zeros(n), np.zeros(n)
        assert H.h_rys_kernel(n, float(pt["X"]), u, w) == 0
        ur = np.array([float(v) for v in pt["u"]])
        wr = np.array([float(v) for v in pt["w"]])
        worst_u = max(worst_u, float(np.max(np.abs(u - ur) / ur)))
        worst_w = max(worst_w, float(np.max(np.abs(w - wr) / wr)))
    assert worst_u < 1e-14 and worst_w < 1e-14, (worst_u, worst_w)
    rng = np.random.default_rng(0)
    X = np.concatenate([rng.uniform(0, 50, 4000), np.geomspace(1e-12, 1e4, 200)])
    for x in X:
        u, w, uc, wc = np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(1)
        H.h_rys_kernel(1, float(x), u, w)
        H.h_rys(1, float(x), uc, wc)
        assert np.max(np.abs(u - uc) / uc) < 5e-14 and np.max(np.abs(w - wc) / wc) < 5e-14, x


def test_boys_grid_matches_mpmath():
    H = CM.harness()
    worst = 0.0
    for pt in _golden("boys_mpmath.json")["points"]:
        F = np.zeros(25)
        H.h_boys(24, float(pt["T"]), F)
        ref = np.array([float(v) for v in pt["F"]])
        worst = max(worst, float(np.max(np.abs(F - ref) / ref)))
    assert worst < 1e-14, worst  # 24 recursion steps of rounding


def test_boys01_closed_form_matches_mpmath():
    H = CM.harness()
    worst = 0.0
    for pt in _golden("boys_mpmath.json")["points"]:
        F = np.zeros(2)
        H.h_boys01(float(pt["T"]), F)
        ref = np.array([float(v) for v in pt["F"][:2]])
        worst = max(worst, float(np.max(np.abs(F - ref) / ref)))
    assert worst < 5e-15, worst


def test_boys01_kernel_taylor_rows_match_mpmath():
    """The one-node kernels' F_0, F_1: LDS rows at X = 0.2 k of Taylor
    coefficients F_m / m! (m <= 11), F_1 as the derivative of the F_0
    polynomial; against the 50-digit mpmath values, and a dense sweep over
    [0, 45] against the 0.1-grid evaluation."""
    H = CM.harness()
    worst = 0.0
    for pt in _golden("boys_mpmath.json")["points"]:
        F = np.zeros(2)
        H.h_boys01_kernel(float(pt["T"]), F)
        ref = np.array([float(v) for v in pt["F"][:2]])
        worst = max(worst, float(np.max(np.abs(F - ref) / ref)))
    assert worst < 5e-15, worst
    F, G = np.zeros(2), np.zeros(2)
    for x in np.linspace(0.0, 45.0, 9001):
        H.h_boys01_kernel(float(x), F)
        H.h_boys01(float(x), G)
        assert np.max(np.abs(F - G) / G) < 5e-15, x


def _check_quartets(oracle, mol, nsample, seed, fp32=False):
    import itertools
    shells = mol.shells()
    fn = mol.function_norms()
    rng = np.random.default_rng(seed)
    worst = 0.0
    seen = set()
    # one quartet per angular class present (last shell of each L, so
    # centres differ where the basis allows), then random quartets
    last = {}
    for k, sh in enumerate(shells):
        last[sh.L] = k
    picks = []
    for la in last:
        for lb in last:
            for lc in last:
                for ld in last:
                    picks.append((last[la], last[lb], last[lc], last[ld]))
    picks += [tuple(int(x) for x in rng.integers(0, len(shells), 4)) for _ in range(nsample)]
    for a, b, c, d in picks:
        if shells[a].L < shells[b].L:
            a, b = b, a
        if shells[c].L < shells[d].L:
            c, d = d, c
        if (shells[a].L, shells[b].L) < (shells[c].L, shells[d].L):
            a, b, c, d = c, d, a, b
        seen.add(tuple(shells[s].L for s in (a, b, c, d)))
        v = CM.shell_quartet(shells, a, b, c, d, fp32=fp32)
        idx = np.array([(shells[a].offset + i, shells[b].offset + j, shells[c].offset + k,
                         shells[d].offset + l) for i, j, k, l in itertools.product(
                            *[range(shells[s].nfunc) for s in (a, b, c, d)])])
        ref = oracle.eri_list(mol, idx, 8)
        got = v * fn[idx[:, 0]] * fn[idx[:, 1]] * fn[idx[:, 2]] * fn[idx[:, 3]]
        worst = max(worst, float(np.max(np.abs(got - ref))))
    return worst, seen


def test_shell_quartets_match_oracle_sp(oracle, water_sp):
    worst, _ = _check_quartets(oracle, water_sp, 200, 0)
    assert worst < 1e-12, worst


def test_shell_quartets_match_oracle_spd(oracle, water_ccpvdz):
    worst, seen = _check_quartets(oracle, water_ccpvdz, 400, 1)
    assert worst < 1e-12, worst
    assert (2, 2, 2, 2) in seen and (2, 1, 1, 0) in seen


def test_fp32_quartets_error_bound(oracle, water_ccpvdz):
    worst, _ = _check_quartets(oracle, water_ccpvdz, 150, 2, fp32=True)
    # FP32 recurrences: relative ~1e-7 of integrals of size <= ~5
    assert worst < 5e-6, worst


def _density(n, seed):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, n // 3 + 1))
    return np.ascontiguousarray(C @ C.T / n)


@pytest.mark.parametrize("fixture", ["water_sp", "water_ccpvdz"])
def test_jk_pipeline_model_matches_einsum(oracle, request, fixture):
    mol = request.getfixturevalue(fixture)
    eri = oracle.eri_tensor(mol, 8)
    P = _density(mol.nbf, 7)
    Jr, Kr = oracle.jk_dense(eri, P)
    J, K = CM.jk_model(mol, P, tau=0.0)
    assert np.max(np.abs(J - Jr)) < 1e-11
    assert np.max(np.abs(K - Kr)) < 1e-11
    # screened at 1e-12: deviation bounded by what was dropped
    J2, K2 = CM.jk_model(mol, P, tau=1e-12)
    assert np.max(np.abs(J2 - Jr)) < 1e-9 and np.max(np.abs(K2 - Kr)) < 1e-9


def test_ket_count_builder_prefix_and_caps():
    Qb = np.array([3.0, 2.0, 1e-3, 1e-7])
    Qk = np.array([3.0, 2.0, 1e-3, 1e-7])
    n_all, n64 = CM.ket_counts(Qb, Qk, True, 1e-9, 1e-5)
    assert list(n_all) == [1, 2, 3, 2]
    assert list(n64) == [1, 2, 2, 0]
    n_all, n64 = CM.ket_counts(Qb, Qk[:2], False, 1e-9, 0.0)
    assert list(n_all) == [2, 2, 2, 2] and list(n64) == list(n_all)


def test_split_bra_component_path_equals_full(water_ccpvdz):
    """The split path -- eri_bra_row (one wavefront per component of A) or
    eri_bra_component (one per bra component pair) -- reproduces
    eri_shell_quartet for every class, FP64 and FP32."""
    shells = water_ccpvdz.shells()
    last = {}
    for k, sh in enumerate(shells):
        last[sh.L] = k
    n = 0
    for la in (0, 1, 2):
        for lb in range(la + 1):
            for lc in (0, 1, 2):
                for ld in range(lc + 1):
                    if (la, lb) < (lc, ld):
                        continue
                    q = (last[la], last[lb], last[lc], last[ld])
                    full = CM.shell_quartet(shells, *q)
                    split = CM.shell_quartet(shells, *q, split=True)
                    assert np.max(np.abs(full - split)) <= 1e-14 * max(1.0, np.max(np.abs(full)))
                    f32 = CM.shell_quartet(shells, *q, fp32=True, split=True)
                    assert np.max(np.abs(f32 - full)) <= 1e-5 * max(1.0, np.max(np.abs(full)))
                    # per component of A (the kernels' split path where the class
                    # allows) vs per bra component pair: the same operations in the
                    # same order, bit for bit on the host (no FMA contraction here)
                    pairs = CM.shell_quartet(shells, *q, split=True, pairs=True)
                    assert np.array_equal(split, pairs)
                    n += 1
    assert n == 21


def test_digest_tiles_equal_digest_quartet():
    rng = np.random.default_rng(9)
    nbf = 40
    P = rng.standard_normal((nbf, nbf))
    P = np.ascontiguousarray(P + P.T)
    H = CM.harness()
    for ls in [(0, 0, 0, 0), (1, 0, 1, 0), (2, 1, 1, 0), (2, 2, 2, 1), (2, 2, 2, 2), (1, 1, 1, 1),
               (2, 1, 2, 0), (2, 1, 1, 1), (2, 2, 1, 0), (2, 2, 2, 0), (2, 1, 2, 1), (2, 2, 1, 1)]:
        nA, nB, nC, nD = [CM.ncart(l) for l in ls]
        v = rng.standard_normal(nA * nB * nC * nD)
        f0 = np.array([0, 7, 15, 30], dtype=np.int32)
        nt = nA * nB + nC * nD + nA * nC + nA * nD + nB * nC + nB * nD
        t1, t2 = np.zeros(nt), np.zeros(nt)
        assert H.h_digest(*ls, np.ascontiguousarray(v * 0.25), P, nbf, f0, t1) == 0
        assert H.h_digest_tiles(*ls, v, 0.25, P, nbf, f0, t2) == 0
        assert np.max(np.abs(t1 - t2)) < 1e-12
        # the split-class digestion's row mode: same sums, same order
        t3 = np.zeros(nt)
        assert H.h_digest_tile_rows(*ls, v, 0.25, P, nbf, f0, t3) == 0
        assert np.array_equal(t2, t3)
        # and its tile mode: groups of tiles in one pass over the integrals
        t4 = np.zeros(nt)
        assert H.h_digest_tile_groups(*ls, v, 0.25, P, nbf, f0, t4) == 0
        assert np.array_equal(t2, t4), ls


@pytest.mark.parametrize("fixture", ["water_sp", "water_ccpvdz"])
def test_one_electron_model_matches_oracle(oracle, request, fixture):
    mol = request.getfixturevalue(fixture)
    S, T, V = CM.one_electron_model(mol)
    Sr, Tr, Vr = oracle.one_electron(mol, 8)
    assert np.max(np.abs(S - Sr)) < 1e-13
    assert np.max(np.abs(T - Tr)) < 1e-12
    assert np.max(np.abs(V - Vr)) < 1e-12


def _shifted(shell, axis, h):
    import copy
    s2 = copy.copy(shell)
    s2.center = shell.center.copy()
    s2.center[axis] += h
    return s2


@pytest.mark.parametrize("ls", [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 1, 1),
                                (2, 0, 1, 0), (2, 1, 1, 0), (2, 2, 1, 1), (2, 1, 2, 1)])
def test_derivative_integrals_match_finite_differences(water_ccpvdz, ls):
    """grad_bra_component (the gradient kernel's math) vs central differences
    of the contracted integrals w.r.t. each centre (h = 1e-5 bohr)."""
    shells = water_ccpvdz.shells()
    byL = {}
    for k, sh in enumerate(shells):
        byL.setdefault(sh.L, []).append(k)
    # distinct shells on different atoms where possible
    pick = [byL[ls[0]][-1], byL[ls[1]][0], byL[ls[2]][len(byL[ls[2]]) // 2], byL[ls[3]][-1]]
    rng = np.random.default_rng(sum(ls))
    nint = int(np.prod([CM.ncart(l) for l in ls]))
    Dfac = rng.standard_normal(nint)
    g = CM.quartet_grad(shells, *pick, Dfac)
    h = 1e-5
    base = [shells[x] for x in pick]
    fd = np.zeros(12)
    for ctr in range(4):
        for ax in range(3):
            vals = []
            for sgn in (1, -1):
                sh = list(base)
                sh[ctr] = _shifted(base[ctr], ax, sgn * h)
                v = CM.shell_quartet(sh, 0, 1, 2, 3)
                vals.append(float(np.dot(Dfac, v)))
            fd[3 * ctr + ax] = (vals[0] - vals[1]) / (2 * h)
    got = np.concatenate([g, -(g[0:3] + g[3:6] + g[6:9])])
    scale = max(1.0, np.max(np.abs(fd)))
    assert np.max(np.abs(got - fd)) < 2e-7 * scale, (got, fd)


def _displaced(mol, atom, axis, h_bohr):
    from segfault_py_amd.molecule import Molecule, BOHR_ANGSTROM
    xyz = mol.xyz.copy()
    xyz[atom, axis] += h_bohr * BOHR_ANGSTROM
    return Molecule.from_geometry(mol.symbols, xyz, _basis_text(mol))


def _basis_text(mol):
    from segfault_py_amd import workloads as W
    return W.even_tempered_text(W.BASIS_SP)


def test_gradient_all_components_of_a_equals_slices(water_ccpvdz):
    """grad_all_a (every component of A in one thread, B an s shell) against
    the per-bra-component-pair path, for every class that uses it."""
    shells = water_ccpvdz.shells()
    last = {}
    for k, sh in enumerate(shells):
        last[sh.L] = k
    rng = np.random.default_rng(21)
    n = 0
    # the canonical classes with B an s shell and at most 36 density factors
    for la, lc, ld in [(0, 0, 0), (1, 0, 0), (1, 1, 0), (2, 0, 0), (2, 1, 0), (2, 2, 0)]:
        q = (last[la], last[0], last[lc], last[ld])
        nA, nC, nD = CM.ncart(la), CM.ncart(lc), CM.ncart(ld)
        Dfac = rng.standard_normal(nA * nC * nD)
        g1 = CM.quartet_grad(shells, *q, Dfac)
        g2 = CM.quartet_grad(shells, *q, Dfac, slices=True)
        assert np.max(np.abs(g1 - g2)) <= 1e-13 * max(1.0, np.max(np.abs(g2))), (la, lc, ld)
        n += 1
    assert n == 6


def test_one_and_two_electron_gradient_models_match_fd(oracle, water_sp):
    """Per-pair one-electron derivative routine and per-quartet derivative
    ERIs, assembled on the CPU, vs central differences (h = 1e-4 bohr) of
    E1 = sum P(T+V) - sum W S and E2 = 1/2 sum P P [(pq|rs) - (pr|qs)/2]
    from the oracle at displaced geometries (P, W held fixed)."""
    mol = water_sp
    rng = np.random.default_rng(21)
    C = rng.standard_normal((mol.nbf, 5))
    P = C @ C.T / 5
    Wm = rng.standard_normal((mol.nbf, mol.nbf))
    Wm = (Wm + Wm.T) / 10
    g1 = CM.one_electron_grad_model(mol, P, Wm)
    g2 = CM.two_electron_grad_model(mol, P)
    h = 1e-4
    fd1 = np.zeros_like(g1)
    fd2 = np.zeros_like(g2)
    for at in range(mol.natom):
        for ax in range(3):
            e1, e2 = [], []
            for sgn in (1, -1):
                m = _displaced(mol, at, ax, sgn * h)
                S, T, V = oracle.one_electron(m, 8)
                e1.append(np.sum(P * (T + V)) - np.sum(Wm * S))
                eri = oracle.eri_tensor(m, 8)
                J, K = oracle.jk_dense(eri, P)
                e2.append(0.5 * np.sum(P * (J - 0.5 * K)))
            fd1[at, ax] = (e1[0] - e1[1]) / (2 * h)
            fd2[at, ax] = (e2[0] - e2[1]) / (2 * h)
    assert np.max(np.abs(g1 - fd1)) < 1e-6, (g1, fd1)
    assert np.max(np.abs(g2 - fd2)) < 1e-6, (g2, fd2)


def test_far_point_charge_form_equals_full_path():
    """(ss|ss) and (ps|ss) quartets whose every primitive quartet has
    X = p q / (p + q) |P - Q|^2 >= 40 (the GPU's far lists): the point-charge
    form eri_one_node_far (S_b S_k / R, and the dipole term for p) equals
    the Taylor/asymptotic Boys path of eri_shell_quartet to rounding, on a
    20-water cluster; and such quartets exist in number."""
    from segfault_py_amd import workloads as W
    mol = W.config_molecule(1)
    shells = mol.shells()
    H = CM.harness()
    rng = np.random.default_rng(3)
    s_sh = [k for k, sh in enumerate(shells) if sh.L == 0]
    p_sh = [k for k, sh in enumerate(shells) if sh.L == 1]
    nfar = nnear = 0
    worst = 0.0
    for t in range(400):
        la = t % 2
        a = int(rng.choice(p_sh if la else s_sh))
        b, c, d = (int(rng.choice(s_sh)) for _ in range(3))
        bra, ket = CM.prim_pairs(shells[a], shells[b]), CM.prim_pairs(shells[c], shells[d])
        pb, pk = bra.reshape(-1, 8), ket.reshape(-1, 8)
        p, q = pb[:, 0][:, None], pk[:, 0][None, :]
        d2 = ((pb[:, None, 1:4] - pk[None, :, 1:4]) ** 2).sum(-1)
        if (p * q / (p + q) * d2).min() < 40.0:
            nnear += 1
            continue
        nfar += 1
        ref = CM.shell_quartet(shells, a, b, c, d)
        out = np.zeros(3 if la else 1)
        assert H.h_eri_far(la, bra, len(pb), ket, len(pk), shells[a].center, out) == 0
        # relative to the quartet's largest integral (values below 1e-250,
        # far beneath any screening threshold, against that floor)
        worst = max(worst, float(np.max(np.abs(out - ref)) / max(np.max(np.abs(ref)), 1e-250)))
    assert nfar > 100 and nnear > 0, (nfar, nnear)
    assert worst < 1e-14, worst
