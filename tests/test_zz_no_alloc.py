"""sf_eri_components_noalloc (the allocation-free entry point) against
sf_eri_components through the C-ABI: two calls in a row on one device, the
second with fewer quartets, each bit-identical to the allocating entry
point; a call beyond the staging array is refused with status 5 before any
device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(fn, mol, q):
    npgf, poff, coefs, expns, norms, shell, center = mol.cgf_arrays()
    q = np.ascontiguousarray(q, dtype=np.int32).reshape(-1, 4)
    out = np.full(len(q), np.nan)
    rc = fn(len(npgf), npgf, poff, np.ascontiguousarray(coefs), np.ascontiguousarray(expns),
            np.ascontiguousarray(norms), shell, center, len(q), q.reshape(-1), out)
    return rc, out


def test_noalloc_equals_allocating_entry_point(gpu, water_ccpvdz):
    from segfault_py_amd import engine
    L = engine.lib()
    rng = np.random.default_rng(21)
    first = rng.integers(0, water_ccpvdz.nbf, (20000, 4))
    second = rng.integers(0, water_ccpvdz.nbf, (3000, 4))
    for q in (first, second):
        rc, got = _call(L.sf_eri_components_noalloc, water_ccpvdz, q)
        assert rc == 0, L.sf_last_error()
        rc, ref = _call(L.sf_eri_components, water_ccpvdz, q)
        assert rc == 0, L.sf_last_error()
        assert np.array_equal(got, ref)
    # beyond the 32 MiB staging array: refused, output untouched
    big = np.zeros((1_500_000, 4), dtype=np.int32)
    rc, out = _call(L.sf_eri_components_noalloc, water_ccpvdz, big)
    assert rc == 5 and b"staging array" in L.sf_last_error()
    assert np.isnan(out).all()
