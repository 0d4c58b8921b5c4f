"""Adjoint (Hellmann-Feynman) derivatives of the excitation energies on the
GPU: against the oracle's closed form on the fixture molecules, and the
frozen-charge nuclear gradient against central finite differences (step
1e-4 bohr) of GPU-solved energies on 32 seeded coordinates."""
import numpy as np
import pytest

from oracle import tda_oracle as orc
from tests._fixtures import molecule

pytestmark = pytest.mark.gpu


def _tda():
    from tincr_amd import tda
    return tda


def _oracle_engine(m):
    q_ov, q_oo, q_vv = orc.trans_charges(m["n_atoms"], m["orbs_per_atom"], m["orbs"], m["s"],
                                         m["occ_indices"], m["virt_indices"])
    orbe = np.ravel(m["orbe"])
    homo, lumo = m["occ_indices"][-1], m["virt_indices"][0]
    omega = orc.orbe_differences(orbe[:homo + 1], orbe[lumo:])
    return orc.SystemEngine(q_ov, q_oo, q_vv, m["gamma"], m["gamma_lr"], omega, len(m["occ_indices"]),
                            len(m["virt_indices"]), faithful=False)


def _system(m, precision="fp64", operator="ao", gamma=None, gamma_lr=None):
    return _tda().TdaSystem(m["n_atoms"], m["orbs_per_atom"], m["xyz"], m["occ_indices"], m["virt_indices"],
                            m["orbs"], m["s"], m["orbe"], m["gamma"] if gamma is None else gamma,
                            m["gamma_lr"] if gamma_lr is None else gamma_lr, precision=precision,
                            operator=operator)


@pytest.mark.parametrize("mol", ["h2o", "benzene", "uracil"])
@pytest.mark.parametrize("op", ["ao", "charges"])
def test_adjoint_vs_oracle(mol, op):
    print(f"adjoint {mol} op={op}", flush=True)
    m = molecule(mol)
    s = _system(m, operator=op)
    n = 4
    w, x = s.run_tda(n, 200, 1e-8)
    eng = _oracle_engine(m)
    got = s.adjoint(n, charges=True)
    got_x = s.adjoint(n, X=x)
    for k in range(n):
        ref = orc.hf_derivatives(eng, x[:, k])
        for key in ("d_gamma", "d_gamma_lr", "d_qoo", "d_qvv"):
            scale = max(np.abs(ref[key]).max(), 1e-3)
            assert np.abs(got[key][k] - ref[key]).max() < 1e-12 * scale, (key, k)
        dq = got["c_qov"][k][:, None] * x[:, k][None, :]
        assert np.abs(dq - ref["d_qov"]).max() < 1e-12 * max(np.abs(ref["d_qov"]).max(), 1e-3)
        assert np.array_equal(got_x["d_gamma"][k], got["d_gamma"][k])
        # without the MO-charge derivatives the AO operator takes its own
        # route (transition charges and gamma_lr derivatives from the AO
        # transition densities, no q_oo / q_vv): the same closed form
        for key in ("d_gamma", "d_gamma_lr"):
            scale = max(np.abs(ref[key]).max(), 1e-3)
            assert np.abs(got_x[key][k] - ref[key]).max() < 1e-12 * scale, (key, k)
        dq = got_x["c_qov"][k][:, None] * x[:, k][None, :]
        assert np.abs(dq - ref["d_qov"]).max() < 1e-12 * max(np.abs(ref["d_qov"]).max(), 1e-3)
    # sum rule: the Coulomb energy 2 u^T gamma u = sum_AB gamma_AB dw/dgamma_AB
    u = eng.q_ov @ x[:, 0]
    assert abs((got["d_gamma"][0] * m["gamma"]).sum() - 2.0 * u @ m["gamma"] @ u) < 1e-12
    s.close()


def _config4(name):
    from tincr_amd import systems as sy
    mol = sy.polyene(60, seed=0) if name == "polyene60" else sy.water_cluster(64, seed=0)
    return mol, sy.tda_inputs(mol, seed=0)


@pytest.mark.parametrize("name", ["polyene60", "water64"])
def test_nuclear_gradient_vs_finite_differences(name):
    """d omega_1..4 / dR at frozen orbitals and charges (gamma and gamma_lr
    follow R) vs central differences, step 1e-4 bohr, on 32 seeded
    coordinates; bar 1e-6 Ha/bohr."""
    from tincr_amd import systems as sy
    mol, inp = _config4(name)
    print(f"gradient {name}: {mol.n_atoms} atoms", flush=True)
    n, tol = 4, 1e-7
    s = _system(inp)
    w0, x0 = s.run_tda(n, 300, tol)
    dg = sy.gamma_gradients_atomwise(mol.z, mol.xyz, 0.0)
    dglr = sy.gamma_gradients_atomwise(mol.z, mol.xyz, sy.LONG_RANGE_RADIUS)
    grad = s.nuclear_gradient(n, dg, dglr)
    s.close()
    if name == "polyene60":
        q = orc.trans_charges(inp["n_atoms"], inp["orbs_per_atom"], inp["orbs"], inp["s"], inp["occ_indices"],
                              inp["virt_indices"])
        o, v = len(inp["occ_indices"]), len(inp["virt_indices"])
        eng = orc.SystemEngine(*q, inp["gamma"], inp["gamma_lr"],
                               orc.orbe_differences(inp["orbe"][:o], inp["orbe"][o:]), o, v, faithful=False)
        ref = orc.frozen_charge_gradient(eng, x0[:, 0], dg, dglr)
        assert np.abs(grad[0] - ref).max() < 1e-12 * max(np.abs(ref).max(), 1.0)
    rng = np.random.default_rng(4)
    coords = rng.choice(3 * mol.n_atoms, 32, replace=False)
    h = 1e-4
    worst = 0.0
    for c in coords:
        vals = []
        for sgn in (1.0, -1.0):
            xyz = mol.xyz.copy()
            xyz[c // 3, c % 3] += sgn * h
            g = sy.gamma_atomwise(mol.z, xyz, 0.0)
            glr = sy.gamma_atomwise(mol.z, xyz, sy.LONG_RANGE_RADIUS)
            sd = _system(dict(inp, xyz=xyz), gamma=g, gamma_lr=glr)
            wd, _ = sd.run_tda(n, 300, tol)
            sd.close()
            vals.append(wd)
        fd = (vals[0] - vals[1]) / (2 * h)
        worst = max(worst, float(np.abs(fd - grad[:, c]).max()))
    print(f"max |FD - adjoint| = {worst:.2e} Ha/bohr", flush=True)
    assert worst < 1e-6


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_nuclear_gradient_mixed_precision_solve(prec):
    """The gradient from a mixed-precision solve (FP64-verified vectors) equals
    the FP64 solve's within 1e-6 Ha/bohr."""
    from tincr_amd import systems as sy
    mol, inp = _config4("polyene60")
    dg = sy.gamma_gradients_atomwise(mol.z, mol.xyz, 0.0)
    dglr = sy.gamma_gradients_atomwise(mol.z, mol.xyz, sy.LONG_RANGE_RADIUS)
    s64 = _system(inp)
    s64.run_tda(4, 300, 1e-7)
    g64 = s64.nuclear_gradient(4, dg, dglr)
    slp = _system(inp, precision=prec)
    slp.run_tda(4, 300, 1e-7)
    glp = slp.nuclear_gradient(4, dg, dglr)
    assert np.abs(glp - g64).max() < 1e-6
    s64.close()
    slp.close()


@pytest.mark.parametrize("mol,lc", [("uracil", True), ("benzene", True), ("h2o", False)])
def test_autograd_function_vs_finite_differences(mol, lc):
    """tincr_amd.autograd.excitation_energies: torch backward (the device
    adjoint) against central differences of the solved energies along a
    random symmetric direction of gamma and gamma_lr."""
    import torch
    from tincr_amd.autograd import FrozenOrbitals, excitation_energies
    m = molecule(mol)
    print(f"autograd {mol} lc={lc}", flush=True)
    fr = FrozenOrbitals(m["n_atoms"], m["orbs_per_atom"], m["xyz"], m["occ_indices"], m["virt_indices"],
                        m["orbs"], m["s"], m["orbe"])
    n = 3
    rng = np.random.default_rng(5)
    c = torch.tensor(rng.standard_normal(n))
    g0 = torch.tensor(m["gamma"], requires_grad=True)
    glr0 = torch.tensor(m["gamma_lr"], requires_grad=True) if lc else None
    g = 0.5 * (g0 + g0.T)
    glr = 0.5 * (glr0 + glr0.T) if lc else None
    loss = excitation_energies(fr, g, glr, n) @ c
    loss.backward()
    d = rng.standard_normal(g0.shape)
    d = 0.5 * (d + d.T)
    dl = rng.standard_normal(g0.shape) if lc else None
    if lc:
        dl = 0.5 * (dl + dl.T)

    def f(h):
        gg = torch.tensor(m["gamma"] + h * d)
        gl = torch.tensor(m["gamma_lr"] + h * dl) if lc else None
        return float(excitation_energies(fr, gg, gl, n) @ c)

    h = 1e-5
    fd = (f(h) - f(-h)) / (2 * h)
    an = float((g0.grad * torch.tensor(d)).sum())
    if lc:
        an += float((glr0.grad * torch.tensor(dl)).sum())
    assert abs(an - fd) < 1e-7 * max(1.0, abs(fd)), (an, fd)
