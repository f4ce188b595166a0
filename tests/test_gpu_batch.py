"""Many molecules in one handle (sf_basis_create_batch), graph replay of the
J/K build (sf_jk_graph), launch-shape autotuning and persistent grids, and
the batched SCF driver.

Bars: a molecule's J, K, S, T, V inside a batch are bit-identical to those of
its own single-molecule handle (same screening, same kernels, same fixed
reduction order); graph replays and every launch shape are bit-identical to
the eager default build; batched SCF energies within 1e-8 Eh of the
single-molecule DIIS SCF.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mols():
    from segfault_py_amd import workloads as W
    from segfault_py_amd.molecule import Molecule
    sp = W.even_tempered_text(W.BASIS_SP)
    spd = W.even_tempered_text(W.BASIS_SPD)
    out = [W.config_molecule(0)]
    s, x = W.water_cluster(2, seed=7, box=5.0)
    out.append(Molecule.from_geometry(s, x, spd))       # s/p/d: all 21 classes
    s, x = W.alkane(2, seed=8)
    out.append(Molecule.from_geometry(s, x, sp))
    s, x = W.water_cluster(3, seed=9, box=6.0)
    out.append(Molecule.from_geometry(s, x, sp))
    return out


def _density(n, seed):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, n // 3 + 1))
    return C @ C.T / n


@pytest.fixture(scope="module")
def batch(gpu):
    from segfault_py_amd import engine
    mols = _mols()
    bb = engine.BatchBasis(mols).screen(1e-12, 1e-5)
    singles = [engine.Basis(m).screen(1e-12, 1e-5) for m in mols]
    P = bb._zeros()
    Ps = []
    for k, m in enumerate(mols):
        n = singles[k].nbf
        Pk = gpu.tensor(_density(n, 100 + k), dtype=gpu.float64, device="cuda")
        P[k, :n, :n] = Pk
        Ps.append(Pk)
    return mols, bb, singles, P, Ps


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_jk_bitwise_equal_single(batch, precision):
    mols, bb, singles, P, Ps = batch
    J, K = bb.jk(P, precision)
    for k, b in enumerate(singles):
        n = b.nbf
        Js, Ks = b.jk(Ps[k], precision)
        assert bool((J[k, :n, :n] == Js).all()), k
        assert bool((K[k, :n, :n] == Ks).all()), k
        # padding is never written
        assert float(J[k, n:, :].abs().sum()) == 0.0 and float(K[k, :, n:].abs().sum()) == 0.0


def test_batch_mixed_one_launch_bitwise_equal_single(batch):
    """Mixed builds with every FP32 segment in FP64 arithmetic (one launch
    over both segments of every class, thresholds chosen per quartet from
    each bra's n64 within its molecule): the batch equals each molecule's
    own handle bit for bit."""
    mols, bb, singles, P, Ps = batch
    from segfault_py_amd.engine import NUM_QUARTET_CLASSES as NQ
    sh = np.tile(np.array([256, 0, 0], dtype=np.int32), 3 * NQ).reshape(NQ, 3, 3)
    sh[:, 2, 2] = 1
    default = np.tile(np.array([256, 0, 0], dtype=np.int32), 3 * NQ)
    bb.set_tuning(sh)
    J, K = bb.jk(P, 1)
    bb.set_tuning(default)
    Jd, Kd = bb.jk(P, 1)
    assert float((J - Jd).abs().max()) < 1e-9 and float((K - Kd).abs().max()) < 1e-9
    for k, b in enumerate(singles):
        n = b.nbf
        b.set_tuning(sh)
        Js, Ks = b.jk(Ps[k], 1)
        b.set_tuning(default)
        assert bool((J[k, :n, :n] == Js).all()), k
        assert bool((K[k, :n, :n] == Ks).all()), k


def test_batch_one_electron_bitwise_equal_single(batch):
    mols, bb, singles, P, Ps = batch
    S, T, V = bb.one_electron()
    for k, b in enumerate(singles):
        n = b.nbf
        for X, Xs in zip((S, T, V), b.one_electron()):
            assert bool((X[k, :n, :n] == Xs).all()), k


def test_batch_counts_are_sums_of_singles(batch):
    mols, bb, singles, P, Ps = batch
    c, fq, pq = bb.counts()
    cs = [b.counts() for b in singles]
    assert np.array_equal(c, sum(x[0] for x in cs))
    assert fq == sum(x[1] for x in cs) and pq == sum(x[2] for x in cs)


def test_graph_replays_match_eager(batch, gpu):
    """Replays of the captured build, with eager work and synchronisation
    between them and the density changed in place, each equal to the eager
    build of the same density."""
    mols, bb, singles, P, Ps = batch
    Pg = P.clone()
    Jg, Kg = bb._zeros(), bb._zeros()
    for rep in range(4):
        Pg.mul_(1.0 + 0.25 * rep)
        J0, K0 = bb.jk(Pg, rep % 2)
        gpu.cuda.synchronize()
        bb.jk_graph(Pg, (Jg, Kg), rep % 2)
        gpu.cuda.synchronize()
        assert bool((Jg == J0).all()) and bool((Kg == K0).all()), rep
        _ = (Jg @ Kg).sum()  # unrelated eager work between replays


def test_single_molecule_graph_and_tuning(gpu, water_ccpvdz):
    from segfault_py_amd import engine
    b = engine.Basis(water_ccpvdz).screen(1e-12, 1e-5)
    P = gpu.tensor(_density(b.nbf, 3), dtype=gpu.float64, device="cuda")
    Jd, Kd = b.jk(P, 1)
    shapes, ms = b.autotune(P, 1)
    assert shapes.shape == (engine.NUM_QUARTET_CLASSES, 3, 3) and np.all(shapes[:, :, 0] % 64 == 0)
    assert set(np.unique(shapes[:, :, 2])) <= {0, 1} and not shapes[:, :2, 2].any()
    J0, K0 = b.jk(P, 1)
    Jg, Kg = b._zeros(), b._zeros()
    for _ in range(3):
        b.jk_graph(P, (Jg, Kg), 1)
        gpu.cuda.synchronize()
        assert bool((Jg == J0).all()) and bool((Kg == K0).all())
    # FP64 arithmetic chosen for some FP32 segments only moves J/K by FP32 rounding
    assert float((J0 - Jd).abs().max()) < 1e-9 and float((K0 - Kd).abs().max()) < 1e-9
    # FP64 arithmetic in every FP32 segment: closer to the FP64 build
    s64 = shapes.copy()
    s64[:, 2, 2] = 1
    b.set_tuning(s64)
    Jx, Kx = b.jk(P, 1)
    J64, K64 = b.jk(P, 0)
    assert float((Jx - J64).abs().max()) <= float((Jd - J64).abs().max()) + 1e-12


@pytest.mark.parametrize("block,grid", [(64, 1), (128, 3), (256, 7), (64, 0)])
def test_persistent_launch_shapes_bitwise(batch, block, grid):
    """Capped (persistent) grids stride over the quartets; every shape gives
    the default build's J and K bit for bit, FP64 and mixed, split classes
    included."""
    from segfault_py_amd.engine import NUM_QUARTET_CLASSES as NQ
    mols, bb, singles, P, Ps = batch
    default = np.tile(np.array([256, 0, 0], dtype=np.int32), 3 * NQ)
    bb.set_tuning(default)
    ref = [bb.jk(P, p) for p in (0, 1)]
    bb.set_tuning(np.tile(np.array([block, grid, 0], dtype=np.int32), 3 * NQ))
    for p in (0, 1):
        J, K = bb.jk(P, p)
        assert bool((J == ref[p][0]).all()) and bool((K == ref[p][1]).all()), p
    bb.set_tuning(default)


def test_layout_errors(batch):
    from segfault_py_amd import engine
    mols, bb, singles, P, Ps = batch
    off = np.zeros(bb.nmol, dtype=np.int64)          # every block at 0: overlap
    ld = np.full(bb.nmol, bb.nmax, dtype=np.int32)
    with pytest.raises(engine.EngineError) as e:
        engine.check(engine.lib().sf_basis_set_layout(bb._h, off, ld))
    assert e.value.code == 1
    with pytest.raises(engine.EngineError):
        bb.gradient(P, P)


def test_scf_batch_matches_single(gpu):
    from segfault_py_amd import hf
    mols = _mols()
    res = hf.scf_batch(mols, tol=1e-10, autotune=True, graph=True)
    assert all(res.converged)
    for k, m in enumerate(mols):
        e = hf.scf_diis(m, tol=1e-10).energy
        assert abs(res.energies[k] - e) < 1e-8, (k, res.energies[k], e)
    resm = hf.scf_batch(mols, tol=1e-10, precision=1, autotune=False, graph=False)
    assert max(abs(a - b) for a, b in zip(resm.energies, res.energies)) < 1e-6
