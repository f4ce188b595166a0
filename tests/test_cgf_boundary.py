"""The drop-in boundary on the reference's own function list (CPU part):
shells_from_cgfs groups reference-shaped ContractedGaussianFunction
objects (src/hf/molecule.py:173-226) into the shells the library takes,
and rejects lists that are not made of whole shells."""
import copy

import numpy as np
import pytest

from refshape import ref_cgfs


def test_grouping_matches_the_package_shells(water_ccpvdz):
    from segfault_py_amd import engine
    mol = water_ccpvdz
    L, nprim, center, expns, coefs, atom, axyz = engine.shells_from_cgfs(ref_cgfs(mol))
    shells = mol.shells()
    assert L.tolist() == [s.L for s in shells]
    assert nprim.tolist() == [s.nprim for s in shells]
    assert atom.tolist() == [s.atom for s in shells]
    assert np.array_equal(axyz, np.asarray(mol.xyz__))
    assert np.array_equal(center.reshape(-1, 3), np.array([s.center for s in shells]))
    assert np.array_equal(expns, np.concatenate([s.expns for s in shells]))
    want = np.concatenate([s.coefs for s in shells])
    assert np.max(np.abs(coefs - want) / np.abs(want)) < 1e-13


def test_grouping_rejects_broken_lists(water_ccpvdz):
    from segfault_py_amd import engine
    cg = ref_cgfs(water_ccpvdz)
    p0 = next(k for k, c in enumerate(cg) if sum(c.shell) == 1)
    with pytest.raises(ValueError, match="does not start a shell"):
        engine.shells_from_cgfs(cg[p0 + 1:])                  # starts at p_y
    with pytest.raises(ValueError, match="expected"):
        engine.shells_from_cgfs(cg[:p0 + 1] + cg[p0 + 2:])    # p_y missing
    with pytest.raises(ValueError, match="cut off"):
        engine.shells_from_cgfs(cg[:p0 + 2])
    bad = copy.deepcopy(cg)
    bad[p0 + 1].expns = [e * 1.001 for e in bad[p0 + 1].expns]
    with pytest.raises(ValueError, match="exponents"):
        engine.shells_from_cgfs(bad)
    bad = copy.deepcopy(cg)
    bad[p0 + 2].coefs = [c * 1.01 for c in bad[p0 + 2].coefs]
    with pytest.raises(ValueError, match="coefficients"):
        engine.shells_from_cgfs(bad)
    with pytest.raises(ValueError, match="ncgf"):
        engine.build_eri_parallel(cg, len(cg) + 1)
