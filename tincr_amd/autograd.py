"""Excitation energies as a differentiable function of gamma and gamma_lr.

``excitation_energies(frozen, gamma, gamma_lr, n_states)`` returns the lowest
LC-TDA excitation energies omega_1..n (a torch tensor) for the frozen orbitals
in ``frozen`` (a ``FrozenOrbitals``) and the Coulomb / long-range matrices
given as torch tensors.  Its backward pass is the device adjoint
(``TdaSystem.adjoint``, tda_adjoint): by Hellmann-Feynman on the converged CI
vectors, d omega_k / d gamma_AB = 2 u_A u_B and d omega_k / d gamma_lr,AB =
-sum_ij q_oo^A_ij P^B_ij -- the gamma terms of tda_gradient_lc
(src/gradients/excited_state.rs:259-487) at frozen orbitals.  Chained with
d gamma / dR (gamma_gradients_atomwise, src/scc/gamma_approximation.rs:256-282)
this gives the frozen-charge nuclear gradient.

The energies require a symmetric operator, so gamma and gamma_lr must be
symmetric; perturb them symmetrically (e.g. build them as (G + G^T) / 2 in
torch) when differentiating.
"""
from dataclasses import dataclass

import numpy as np
import torch

from .tda import TdaSystem


@dataclass
class FrozenOrbitals:
    """The inputs of a TDA solve that are held fixed: geometry, AO blocks per
    atom, occupied / virtual MO indices, MO coefficients, AO overlap and
    orbital energies (the arguments of tda_set_system besides gamma)."""
    n_atoms: int
    orbs_per_atom: np.ndarray
    xyz: np.ndarray
    occ_indices: list
    virt_indices: list
    orbs: np.ndarray
    s: np.ndarray
    orbe: np.ndarray


class _ExcitationEnergies(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gamma, gamma_lr, frozen, n_states, tol, max_iter):
        g = gamma.detach().cpu().numpy().astype(np.float64)
        glr = None if gamma_lr is None else gamma_lr.detach().cpu().numpy().astype(np.float64)
        s = TdaSystem(frozen.n_atoms, frozen.orbs_per_atom, frozen.xyz, frozen.occ_indices, frozen.virt_indices,
                      frozen.orbs, frozen.s, frozen.orbe, g, glr)
        try:
            w, _ = s.run_tda(n_states, max_iter, tol)
            adj = s.adjoint(n_states)
        finally:
            s.close()
        ctx.has_lr = gamma_lr is not None
        saved = [torch.from_numpy(adj["d_gamma"])]
        if ctx.has_lr:
            saved.append(torch.from_numpy(adj["d_gamma_lr"]))
        ctx.save_for_backward(*saved)
        ctx.dtypes = (gamma.dtype, gamma.device)
        return torch.from_numpy(np.ascontiguousarray(w)).to(dtype=gamma.dtype, device=gamma.device)

    @staticmethod
    def backward(ctx, grad_w):
        dtype, device = ctx.dtypes
        gw = grad_w.detach().to(torch.float64).cpu()
        saved = ctx.saved_tensors
        d_gamma = torch.einsum("k,kab->ab", gw, saved[0]).to(dtype=dtype, device=device)
        d_gamma_lr = None
        if ctx.has_lr:
            d_gamma_lr = torch.einsum("k,kab->ab", gw, saved[1]).to(dtype=dtype, device=device)
        return d_gamma, d_gamma_lr, None, None, None, None


def excitation_energies(frozen, gamma, gamma_lr, n_states, tol=1e-8, max_iter=300):
    """omega_1..n_states (Ha) of the LC-TDA problem, differentiable in gamma
    and gamma_lr (torch tensors [N, N]; gamma_lr None: no long-range exchange)."""
    return _ExcitationEnergies.apply(gamma, gamma_lr, frozen, n_states, tol, max_iter)
