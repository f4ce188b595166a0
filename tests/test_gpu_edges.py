"""Edge cases of the J/K path on the GPU: nothing survives screening, a
single atom, a batch of one molecule, the same molecule repeated in a batch,
and the unscreened dense export of the smallest basis."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _h_atom():
    from segfault_py_amd import workloads as W
    from segfault_py_amd.molecule import Molecule
    return Molecule.from_geometry(["H"], np.zeros((1, 3)), W.even_tempered_text(W.BASIS_SP))


def test_nothing_survives_screening(gpu, water_sp):
    from segfault_py_amd import engine
    b = engine.Basis(water_sp).screen(1e6, 1e7)
    c, fq, pq = b.counts()
    assert c.sum() == 0 and fq == 0 and pq == 0
    P = gpu.eye(b.nbf, dtype=gpu.float64, device="cuda")
    for prec in (0, 1):
        J, K = b.jk(P, prec)
        assert float(J.abs().max()) == 0.0 and float(K.abs().max()) == 0.0
    J, K = b._zeros(), b._zeros()
    b.jk_graph(P, (J, K))
    gpu.cuda.synchronize()
    assert float(J.abs().max()) == 0.0
    shapes, ms = b.autotune(P, 1)       # nothing to time: shapes stay the defaults
    assert (shapes[:, :, 0] == 256).all()


def test_single_atom_matches_oracle(gpu, oracle):
    from segfault_py_amd import engine
    mol = _h_atom()
    b = engine.Basis(mol).screen(0.0, 0.0)
    assert b.nbf == 3
    eri = oracle.eri_tensor(mol, 1)
    assert np.max(np.abs(b.eri_dense() - eri)) < 1e-12
    rng = np.random.default_rng(2)
    C = rng.standard_normal((3, 2))
    Pn = C @ C.T
    J, K = b.jk(gpu.tensor(Pn, dtype=gpu.float64, device="cuda"))
    Jr, Kr = oracle.jk_dense(eri, Pn)
    assert np.max(np.abs(J.cpu().numpy() - Jr)) < 1e-12
    assert np.max(np.abs(K.cpu().numpy() - Kr)) < 1e-12


def test_batch_of_one_and_duplicates(gpu, water_ccpvdz):
    """A batch of one molecule is the single-molecule handle; three copies
    of one molecule give three identical blocks (bit for bit)."""
    from segfault_py_amd import engine
    mol = water_ccpvdz
    single = engine.Basis(mol).screen(1e-12, 1e-5)
    rng = np.random.default_rng(3)
    C = rng.standard_normal((single.nbf, 5))
    P1 = gpu.tensor(C @ C.T, dtype=gpu.float64, device="cuda")
    Js, Ks = single.jk(P1, 1)
    one = engine.BatchBasis([mol]).screen(1e-12, 1e-5)
    J1, K1 = one.jk(P1[None].contiguous(), 1)
    assert bool((J1[0] == Js).all()) and bool((K1[0] == Ks).all())
    three = engine.BatchBasis([mol, mol, mol]).screen(1e-12, 1e-5)
    P3 = P1[None].repeat(3, 1, 1).contiguous()
    J3, K3 = three.jk(P3, 1)
    for k in range(3):
        assert bool((J3[k] == Js).all()) and bool((K3[k] == Ks).all())
    c1, c3 = one.counts(), three.counts()
    assert np.array_equal(3 * c1[0], c3[0]) and 3 * c1[1] == c3[1] and 3 * c1[2] == c3[2]
