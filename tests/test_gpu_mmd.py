"""The McMurchie-Davidson engine (E.R.E transforms on the FP64 matrix cores)
against the Rys engine and the oracle, through the C-ABI.

Two independent methods on the GPU (Rys quadrature; Hermite expansion with
MFMA GEMMs) agreeing pins both: a wrong MFMA operand/accumulator lane
layout, a dropped (-1)^(tau+nu+phi) sign or a wrong Hermite index would
break the agreement.  Bars: ERIs within 1e-12 of Rys and 1e-11 of the
oracle; J/K within 1e-10 of Rys.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _density(n, seed):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, n // 3 + 1))
    return C @ C.T / n


@pytest.mark.parametrize("fixture", ["water_sp", "water_ccpvdz"])
def test_mmd_dense_matches_rys_and_oracle(gpu, oracle, request, fixture):
    from segfault_py_amd import engine
    mol = request.getfixturevalue(fixture)
    b = engine.Basis(mol).screen(0.0, 0.0)
    rys = b.eri_dense()
    b.set_engine(1)
    mmd = b.eri_dense()
    assert np.max(np.abs(mmd - rys)) < 1e-12
    assert np.max(np.abs(mmd - oracle.eri_tensor(mol, 8))) < 1e-11
    # permutational symmetry of the MMD tensor itself
    assert np.max(np.abs(mmd - mmd.transpose(1, 0, 2, 3))) < 1e-13
    assert np.max(np.abs(mmd - mmd.transpose(2, 3, 0, 1))) < 1e-13


def test_mmd_jk_every_class_matches_rys(gpu):
    """Three waters with the s/p/d basis (all 21 quartet classes): J/K with
    every class on MMD, and with only the d-shell classes on MMD, against
    Rys; FP64 and the mixed-precision segments."""
    from segfault_py_amd import engine, workloads as W
    from segfault_py_amd.molecule import Molecule
    s, x = W.water_cluster(3, seed=7, box=5.0)
    mol = Molecule.from_geometry(s, x, W.even_tempered_text(W.BASIS_SPD))
    b = engine.Basis(mol).screen(1e-12, 1e-5)
    assert (b.counts()[0][:engine.NUM_CLASSES_SPD, 0] > 0).all()
    P = gpu.tensor(_density(mol.nbf, 9), dtype=gpu.float64, device="cuda")
    J0, K0 = b.jk(P, 0)
    b.set_engine(1)
    J1, K1 = b.jk(P, 0)
    assert float((J1 - J0).abs().max()) < 1e-10 and float((K1 - K0).abs().max()) < 1e-10
    b.set_engine(0)
    for c in range(6, 21):            # classes with a d shell
        b.set_engine(1, c)
    J2, K2 = b.jk(P, 1)
    b.set_engine(0)
    Jm, Km = b.jk(P, 1)
    assert float((J2 - J0).abs().max()) < 1e-9 and float((K2 - K0).abs().max()) < 1e-9
    assert float((Jm - J0).abs().max()) < 1e-9


def test_mmd_graph_and_batch(gpu):
    """The MMD engine inside a captured graph and in a molecule batch gives
    the eager single-molecule result bit for bit."""
    from segfault_py_amd import engine, workloads as W
    mols = [W.config_molecule(0)]
    s, x = W.water_cluster(2, seed=3, box=5.0)
    from segfault_py_amd.molecule import Molecule
    mols.append(Molecule.from_geometry(s, x, W.even_tempered_text(W.BASIS_SPD)))
    bb = engine.BatchBasis(mols).screen(1e-12, 0.0).set_engine(1)
    P = bb._zeros()
    singles = []
    for k, m in enumerate(mols):
        n = int(bb.nbf_mol[k])
        P[k, :n, :n] = gpu.tensor(_density(n, 20 + k), dtype=gpu.float64, device="cuda")
        singles.append(engine.Basis(m).screen(1e-12, 0.0).set_engine(1))
    J, K = bb.jk(P, 0)
    Jg, Kg = bb._zeros(), bb._zeros()
    for _ in range(2):
        bb.jk_graph(P, (Jg, Kg), 0)
        gpu.cuda.synchronize()
        assert bool((Jg == J).all()) and bool((Kg == K).all())
    for k, b in enumerate(singles):
        n = b.nbf
        Js, Ks = b.jk(P[k, :n, :n].contiguous(), 0)
        assert bool((J[k, :n, :n] == Js).all()) and bool((K[k, :n, :n] == Ks).all())
