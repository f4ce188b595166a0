"""Analytic RHF nuclear gradients (derivative-ERI backward kernels) against
central finite differences (h = 1e-4 bohr) of GPU and oracle SCF energies.
Bar: max |error| < 1e-5 Eh/bohr."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 1e-4


def _energy(mol, tol=1e-11, precision=0):
    from segfault_py_amd import hf
    return hf.scf_diis(mol, tol=tol, precision=precision).energy


def _fd(mol, comps, energy=_energy):
    from segfault_py_amd import grad as G
    R = np.asarray(mol.xyz__)
    out = {}
    for (a, k) in comps:
        e = []
        for sgn in (1, -1):
            Rd = R.copy()
            Rd[a, k] += sgn * H
            e.append(energy(G.molecule_at(mol, Rd)))
        out[(a, k)] = (e[0] - e[1]) / (2 * H)
    return out


def _analytic(mol, precision=0):
    """SCF and gradient in the same precision (mixed: quartets and
    derivative quartets below tau_mixed = 1e-5 in FP32)."""
    from segfault_py_amd import engine, grad as G, hf
    b = engine.Basis(mol).screen(1e-12, 1e-5)
    r = hf.scf_diis(mol, tol=1e-11, basis=b, precision=precision)
    return G.rhf_gradient(mol, r, b, precision)


@pytest.mark.parametrize("fixture", ["water_sp", "water_ccpvdz"])
def test_gradient_vs_gpu_finite_differences(gpu, request, fixture):
    mol = request.getfixturevalue(fixture)
    g = _analytic(mol)
    assert np.max(np.abs(g.sum(axis=0))) < 1e-8  # translational invariance
    comps = [(a, k) for a in range(mol.natom) for k in range(3)]
    fd = _fd(mol, comps)
    err = max(abs(g[a, k] - v) for (a, k), v in fd.items())
    assert err < 1e-5, err


def test_gradient_vs_oracle_finite_differences(gpu, oracle, water_sp):
    """FD of the oracle's own SCF energy (CPU S, T, V, ERIs, DIIS)."""
    def e_oracle(m):
        S, T, V = oracle.one_electron(m, 8)
        eri = oracle.eri_tensor(m, 8)
        E, _, conv, _ = oracle.scf_diis(S, T + V, int(m.numbers.sum()) // 2,
                                        lambda P: oracle.jk_dense(eri, P), m.e_nuc, tol=1e-11)
        assert conv
        return E
    g = _analytic(water_sp)
    fd = _fd(water_sp, [(a, k) for a in range(3) for k in range(3)], energy=e_oracle)
    err = max(abs(g[a, k] - v) for (a, k), v in fd.items())
    assert err < 1e-5, err


def test_autograd_function(gpu, water_sp):
    from segfault_py_amd import grad as G
    coords = gpu.tensor(np.asarray(water_sp.xyz__), dtype=gpu.float64, requires_grad=True)
    E = G.RHFEnergy.apply(coords, water_sp)
    E.backward()
    g = _analytic(water_sp)
    assert abs(float(E) - _energy(water_sp)) < 1e-9
    assert np.max(np.abs(coords.grad.numpy() - g)) < 1e-9


def test_cluster_gradient_fp64_and_mixed(gpu):
    """Ten waters with the s/p/d basis (30 atoms): analytic vs FD on a few
    components; mixed-precision SCF gradient within 1e-5 of FP64."""
    from segfault_py_amd import workloads as W
    mol = W.config_molecule(3)
    g64 = _analytic(mol, 0)
    gmx = _analytic(mol, 1)
    assert np.max(np.abs(g64.sum(axis=0))) < 1e-7
    assert np.max(np.abs(gmx - g64)) < 1e-5
    fd = _fd(mol, [(0, 0), (1, 1), (2, 2), (17, 0)], energy=lambda m: _energy(m, tol=1e-10))
    err = max(abs(g64[a, k] - v) for (a, k), v in fd.items())
    assert err < 1e-5, err
    # the mixed gradient against finite differences of the mixed energy
    fdm = _fd(mol, [(0, 0), (17, 0)], energy=lambda m: _energy(m, tol=1e-10, precision=1))
    errm = max(abs(gmx[a, k] - v) for (a, k), v in fdm.items())
    assert errm < 1e-5, errm


def test_gradient_primitive_screening(gpu):
    """The derivative kernels skip primitive quartets below their segment's
    threshold (FP64 quartets: 1e-3 tau; FP32 quartets of a mixed gradient:
    1e-8 tau_mixed), as the J/K kernels do: against the same gradients with
    primitive screening off, ten waters with the s/p/d basis."""
    from segfault_py_amd import engine, hf, workloads as W
    mol = W.config_molecule(3)
    b = engine.Basis(mol).screen(1e-12, 1e-5)
    r = hf.scf_diis(mol, basis=b, tol=1e-10)
    Wd = r.energy_weighted_density()
    g = b.gradient(r.P, Wd, 0)
    gm = b.gradient(r.P, Wd, 1)
    b.set_prim_tau_fp64(0.0)
    b.set_prim_tau(0.0)
    g0 = b.gradient(r.P, Wd, 0)
    gm0 = b.gradient(r.P, Wd, 1)
    b.set_prim_tau_fp64(-1.0)
    b.set_prim_tau(-1.0)
    assert np.max(np.abs(g - g0)) < 1e-10
    assert np.max(np.abs(gm - gm0)) < 1e-8
    assert np.max(np.abs(gm - g)) < 1e-8
