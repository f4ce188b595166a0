"""Analytic RHF nuclear gradients and their torch.autograd wrapper.

The reference computes single-point energies only (src/hf/hf.py:236-353);
this module adds the derivative the engine's backward kernels provide:

    dE/dR = sum P dH/dR - sum W dS/dR + 1/2 sum Gamma d(pq|rs)/dR + dE_nuc/dR

with W = 2 sum_occ e_i C_i C_i^T and Gamma_pqrs = P_pq P_rs - P_pr P_qs / 2.
The electronic part runs on the GPU (sf_grad: derivative-ERI kernels per
angular-momentum class, one-electron derivative kernels, fixed-order
reductions); the nuclear-repulsion part is O(natom^2) on the host.

``RHFEnergy.apply(coords_bohr)`` is a ``torch.autograd.Function``: forward
runs the SCF at the given geometry (and raises if it does not converge: the
analytic gradient assumes a stationary SCF), backward runs the derivative
kernels and returns the analytic gradient, so a forward without a backward
pays for no gradient.
"""
from __future__ import annotations

import numpy as np

from . import engine, hf
from .molecule import BOHR_ANGSTROM, Molecule


def nuclear_repulsion_gradient(mol) -> np.ndarray:
    """d E_nuc / dR_A = -sum_B Z_A Z_B (R_A - R_B) / |R_A - R_B|^3 (natom x 3)."""
    R = np.asarray(mol.xyz__, dtype=np.float64).reshape(-1, 3)
    Z = np.asarray(mol.numbers, dtype=np.float64)
    d = R[:, None, :] - R[None, :, :]
    r = np.sqrt(np.einsum("abk,abk->ab", d, d))
    np.fill_diagonal(r, np.inf)
    w = np.outer(Z, Z) / r ** 3
    return -np.einsum("ab,abk->ak", w, d)


def rhf_gradient(mol, result: hf.SCFResult, basis: engine.Basis, precision: int = 0) -> np.ndarray:
    """Total RHF gradient (natom x 3, Eh/bohr) at a converged SCF result;
    precision 1 runs the derivative quartets below tau_mixed in FP32."""
    W = result.energy_weighted_density()
    g_el = basis.gradient(result.P, W, precision)
    return g_el + nuclear_repulsion_gradient(mol)


def molecule_at(template: Molecule, coords_bohr) -> Molecule:
    """The template's atoms and basis at new coordinates (bohr)."""
    xyz = np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3) * BOHR_ANGSTROM
    text = getattr(template, "basis_text", None)
    if text is None:
        raise ValueError("template molecule must carry basis_text")
    return Molecule.from_geometry(template.symbols, xyz, text)


class RHFEnergy:
    """torch.autograd.Function: E_RHF(coords) with an analytic backward.

    Usage: ``E = RHFEnergy.apply(coords, template)`` with ``coords`` a float64
    tensor (natom, 3) in bohr; options via ``RHFEnergy.options``.
    """

    options = {"tau": 1e-12, "tol": 1e-10, "precision": 0, "tau_mixed": 1e-5}

    @staticmethod
    def _fn():
        import torch

        class _F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, coords, template):
                mol = molecule_at(template, coords.detach().cpu().numpy())
                o = dict(RHFEnergy.options)
                basis = engine.Basis(mol).screen(o["tau"], o["tau_mixed"])
                res = hf.scf_diis(mol, tol=o["tol"], basis=basis, precision=o["precision"])
                if not res.converged:
                    raise RuntimeError(f"RHFEnergy: SCF did not converge in {res.iterations} "
                                       "iterations; the analytic gradient needs a converged SCF")
                # the gradient is computed in backward, from what forward leaves here
                ctx.sf_state = (mol, res, basis, o["precision"])
                return torch.tensor(res.energy, dtype=coords.dtype, device=coords.device)

            @staticmethod
            def backward(ctx, grad_out):
                mol, res, basis, precision = ctx.sf_state
                g = rhf_gradient(mol, res, basis, precision)
                del ctx.sf_state
                gt = torch.tensor(g, dtype=grad_out.dtype, device=grad_out.device)
                return grad_out * gt, None

        return _F

    @classmethod
    def apply(cls, coords, template):
        return cls._fn().apply(coords, template)
