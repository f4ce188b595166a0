"""The drop-in boundary on the reference's own objects, on the GPU:
build_eri_parallel(cgfs, ncgf) and Basis.from_cgfs take the function list
scf_diis builds (src/hf/hf.py:254-264) -- reference-shaped
ContractedGaussianFunction objects -- and reproduce the oracle's tensor,
one-electron integrals and J/K."""
import numpy as np
import pytest

from refshape import ref_cgfs

pytestmark = pytest.mark.gpu


def test_build_eri_parallel_on_reference_cgfs(gpu, oracle, water_ccpvdz):
    from segfault_py_amd import engine
    cg = ref_cgfs(water_ccpvdz)
    eri = engine.build_eri_parallel(cg, len(cg))
    ref = oracle.eri_tensor(water_ccpvdz, 8)
    assert eri.shape == ref.shape and np.max(np.abs(eri - ref)) < 1e-11


def test_from_cgfs_jk_and_one_electron(gpu, oracle, water_ccpvdz):
    from segfault_py_amd import engine
    mol = water_ccpvdz
    b = engine.Basis.from_cgfs(ref_cgfs(mol), numbers=mol.numbers).screen(1e-12)
    rng = np.random.default_rng(41)
    C = rng.standard_normal((mol.nbf, 6))
    Pn = C @ C.T / mol.nbf
    J, K = b.jk(gpu.tensor(Pn, dtype=gpu.float64, device="cuda"))
    Jr, Kr = oracle.jk_dense(oracle.eri_tensor(mol, 8), Pn)
    assert np.max(np.abs(J.cpu().numpy() - Jr)) < 1e-10
    assert np.max(np.abs(K.cpu().numpy() - Kr)) < 1e-10
    S, T, V = b.one_electron()
    Sr, Tr, Vr = oracle.one_electron(mol, 8)
    assert np.max(np.abs(S.cpu().numpy() - Sr)) < 1e-13
    assert np.max(np.abs(T.cpu().numpy() - Tr)) < 1e-12
    assert np.max(np.abs(V.cpu().numpy() - Vr)) < 1e-12
