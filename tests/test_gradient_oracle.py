"""Pins of the excited-state gradient oracle and its host inputs (CPU).

* dS / dH0 (tincr_amd.scc.h0_and_s_gradients) and the gamma gradient rows
  against central differences of scc.h0_and_s / systems.gamma_atomwise;
* f_v(v) : W and f_lr(v) : W (helpers.rs:47-79, :130-186) are the nuclear
  derivatives, at fixed v and W, of the Coulomb and long-range exchange Fock
  contractions F_v[v] : W and F_lr[v] : W (the matrices helpers.rs:56-76 and
  :144-182 differentiate), checked by central differences;
* the whole gradient against central differences of the lowest excitation
  energy of SCC + dense TDA (the reference's own check,
  src/gradients/numerical.rs:67-73: step 1e-4 bohr, tolerance 1e-6): the
  no-LC gradient and the exact LC form pass; the reference's LC form misses
  by about 1e-4 Ha/bohr on butadiene (its A-like Z-vector operator and the
  doubled exchange pattern of W_ij's H+[Z]), which the last test records.
"""
import numpy as np
import pytest

from oracle import gradient_oracle as go
from oracle import tda_oracle as to
from tincr_amd import scc
from tincr_amd import systems as sy


def _formaldehyde(seed=0):
    rng = np.random.default_rng(seed)
    xyz = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 2.28], [1.77, 0.0, -1.1], [-1.77, 0.0, -1.1]])
    return np.array([6, 8, 1, 1]), xyz + rng.uniform(-0.1, 0.1, xyz.shape)


def _molecules():
    w = sy.water_cluster(2, seed=1)
    p = sy.polyene(4, seed=1)
    return [("water2", w.z, w.xyz), ("butadiene", p.z, p.xyz), ("formaldehyde",) + _formaldehyde()]


@pytest.mark.parametrize("name,z,xyz", _molecules(), ids=[m[0] for m in _molecules()])
def test_sk_gradients_match_central_differences(name, z, xyz):
    ds, dh = scc.h0_and_s_gradients(z, xyz)
    atom, _, _, _ = scc.ao_layout(z)
    full_s = go.expand_rows(ds, atom, len(z))
    full_h = go.expand_rows(dh, atom, len(z))
    h = 1e-5
    for c in range(3 * len(z)):
        xp = xyz.copy()
        xp.reshape(-1)[c] += h
        xm = xyz.copy()
        xm.reshape(-1)[c] -= h
        sp, hp = scc.h0_and_s(z, xp)
        sm, hm = scc.h0_and_s(z, xm)
        assert np.abs((sp - sm) / (2 * h) - full_s[c]).max() < 1e-9
        assert np.abs((hp - hm) / (2 * h) - full_h[c]).max() < 1e-9


@pytest.mark.parametrize("r_lr", [0.0, sy.LONG_RANGE_RADIUS])
def test_gamma_gradient_rows(r_lr):
    mol = sy.water_cluster(3, seed=2)
    hub = scc.hubbard_of(mol.z)
    full = sy.gamma_gradients_atomwise(mol.z, mol.xyz, r_lr, hubbard=hub)
    rows = scc.gamma_gradients_rows(mol.z, mol.xyz, r_lr, hubbard=hub)
    n = len(mol.z)
    for a in range(n):
        for k in range(3):
            assert np.abs(rows[k, a] - full[3 * a + k, a]).max() < 1e-15
            full[3 * a + k, a] = 0.0
    assert not full.any()                     # every other entry of the reference layout is zero
    # and the rows are d gamma_AB / d R_A
    h = 1e-5
    for c in (0, 4, 8, 13):
        xp = mol.xyz.copy()
        xp.reshape(-1)[c] += h
        xm = mol.xyz.copy()
        xm.reshape(-1)[c] -= h
        fd = (sy.gamma_atomwise(mol.z, xp, r_lr, hubbard=hub) - sy.gamma_atomwise(mol.z, xm, r_lr, hubbard=hub)) / (2 * h)
        a, k = divmod(c, 3)
        assert np.abs(fd[a] - rows[k, a]).max() < 1e-9


def test_f_v_and_f_lr_are_fock_derivatives():
    mol = sy.polyene(4, seed=1)
    z, xyz = mol.z, mol.xyz
    atom, _, _, _ = scc.ao_layout(z)
    n, n_at = len(atom), len(z)
    hub = scc.hubbard_of(z)
    rng = np.random.default_rng(0)
    v = rng.standard_normal((n, n))
    w = rng.standard_normal((n, n))

    def mats(x):
        s, _ = scc.h0_and_s(z, x)
        g = sy.gamma_atomwise(z, x, 0.0, hubbard=hub)[np.ix_(atom, atom)]
        gl = sy.gamma_atomwise(z, x, sy.LONG_RANGE_RADIUS, hubbard=hub)[np.ix_(atom, atom)]
        return s, g, gl

    def fock_v(x):             # helpers.rs:56-76 at zero derivative
        s, g, _ = mats(x)
        gsv = g @ np.sum(s * (v + v.T), axis=0)
        return 0.25 * s * (gsv[:, None] + gsv[None, :])

    def fock_lr(x):            # helpers.rs:144-182 at zero derivative
        s, _, gl = mats(x)
        return 0.25 * (gl * (s @ v.T @ s) + ((s @ v.T) * gl) @ s + s @ ((v.T @ s) * gl) + s @ (gl * v.T) @ s)

    s, g, gl = mats(xyz)
    ds, _ = scc.h0_and_s_gradients(z, xyz)
    grad_s = go.expand_rows(ds, atom, n_at)
    g1 = go.gamma_ao_gradients(sy.gamma_gradients_atomwise(z, xyz, 0.0, hubbard=hub), atom)
    g1l = go.gamma_ao_gradients(sy.gamma_gradients_atomwise(z, xyz, sy.LONG_RANGE_RADIUS, hubbard=hub), atom)
    an_v = go.contract(go.f_v(v, s, grad_s, g, g1), w)
    an_lr = go.contract(go.f_lr(v, s, grad_s, gl, g1l), w)
    h = 1e-5
    for c in range(3 * n_at):
        xp = xyz.copy()
        xp.reshape(-1)[c] += h
        xm = xyz.copy()
        xm.reshape(-1)[c] -= h
        assert abs(np.sum((fock_v(xp) - fock_v(xm)) * w) / (2 * h) - an_v[c]) < 1e-8
        assert abs(np.sum((fock_lr(xp) - fock_lr(xm)) * w) / (2 * h) - an_lr[c]) < 1e-8


def _lowest_state(z, xyz, lc):
    r = scc.run_scc(z, xyz, lc=lc, charge_conv=1e-13, energy_conv=1e-13, max_iter=500)
    assert r["converged"]
    occ = [i for i, f in enumerate(r["occupation"]) if f > 1.0]
    virt = [i for i, f in enumerate(r["occupation"]) if f <= 1.0]
    n_at = len(z)
    q_ov, q_oo, q_vv = to.trans_charges(n_at, np.bincount(r["atom"], minlength=n_at), r["orbs"], r["s"], occ, virt)
    om = to.orbe_differences(r["orbe"][occ], r["orbe"][virt])
    eng = to.SystemEngine(q_ov, q_oo, q_vv, r["gamma"], r["gamma_lr"], om, len(occ), len(virt))
    w, x = np.linalg.eigh(eng.dense())
    return r, occ, virt, w, x


@pytest.fixture(scope="module", params=["lc", "nolc"])
def butadiene_fd(request):
    lc = request.param == "lc"
    mol = sy.polyene(4, seed=0)
    z, xyz = mol.z, mol.xyz.copy()
    r, occ, virt, w, x = _lowest_state(z, xyz, lc)
    assert w[1] - w[0] > 1e-3                 # no crossing within the finite-difference steps
    g = go.gradient_inputs(z, xyz, r, occ, virt)
    h = 1e-4
    fd = np.zeros(3 * len(z))
    for c in range(3 * len(z)):
        xp = xyz.copy()
        xp.reshape(-1)[c] += h
        xm = xyz.copy()
        xm.reshape(-1)[c] -= h
        fd[c] = (_lowest_state(z, xp, lc)[3][0] - _lowest_state(z, xm, lc)[3][0]) / (2 * h)
    return lc, g, x[:, 0].reshape(len(occ), len(virt)), w[0], fd


def test_gradient_matches_central_differences(butadiene_fd):
    lc, g, x, omega, fd = butadiene_fd
    grad = go.tda_gradient(g, x, omega, exact=True)
    assert np.abs(grad - fd).max() < 1e-6
    assert np.abs(fd).max() > 1e-2
    if not lc:                                 # without LC the reference's gradient is the exact one
        assert np.abs(go.tda_gradient(g, x, omega, exact=False) - grad).max() < 1e-12


def test_reference_lc_form_misses_central_differences(butadiene_fd):
    lc, g, x, omega, fd = butadiene_fd
    if not lc:
        pytest.skip("the two forms coincide without LC")
    err = np.abs(go.tda_gradient(g, x, omega, exact=False) - fd).max()
    assert 1e-5 < err < 1e-3
