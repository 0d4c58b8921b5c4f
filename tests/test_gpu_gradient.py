"""Excited-state nuclear gradient on the GPU (tda_gradient) on converged SCC
states: against the oracle's restatement of tda_gradient_lc / _nolc
(src/gradients/excited_state.rs:56-487) in both forms and with both exchange
operators, and the exact form against central differences (step 1e-4 bohr,
the reference's src/gradients/numerical.rs:67-73 check, tolerance 1e-6
Ha/bohr) of SCC + GPU TDA excitation energies on 32 seeded coordinates of
each differentiation system, C60H62 and (H2O)64, with the state from an FP64
and from an FP32-ladder solve.  The displaced SCCs start from the converged
density of the undisplaced geometry (the reference's restart from kept P and
dq) and run their BLAS on 4 threads (small matrices: more threads only add
synchronisation)."""
import numpy as np
import pytest

from oracle import gradient_oracle as go
from tincr_amd import scc
from tincr_amd import systems as sy

pytestmark = pytest.mark.gpu


def _scc(mol, lc, conv=1e-12, p_init=None):
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=4, user_api="blas"):
        r = scc.run_scc(mol.z, mol.xyz, lc=lc, charge_conv=conv, energy_conv=conv, max_iter=500, p_init=p_init)
    assert r["converged"]
    occ = [i for i, f in enumerate(r["occupation"]) if f > 1.0]
    virt = [i for i, f in enumerate(r["occupation"]) if f <= 1.0]
    return r, occ, virt


def _system(mol, r, occ, virt, operator="ao", precision="fp64"):
    from tincr_amd import tda
    return tda.TdaSystem(mol.n_atoms, mol.orbs_per_atom, mol.xyz, occ, virt, r["orbs"], r["s"], r["orbe"],
                         r["gamma"], r["gamma_lr"], operator=operator, precision=precision)


CASES = {
    "butadiene": (lambda: sy.polyene(4, seed=0), True),
    "water3": (lambda: sy.water_cluster(3, seed=2), True),
    "C10H12": (lambda: sy.polyene(10, seed=0), True),
    "hexatriene-nolc": (lambda: sy.polyene(6, seed=0), False),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("op", ["ao", "charges"])
def test_gradient_matches_oracle(case, op):
    print(f"gradient vs oracle {case} op={op}", flush=True)
    make, lc = CASES[case]
    mol = make()
    r, occ, virt = _scc(mol, lc)
    s = _system(mol, r, occ, virt, op)
    w, x = s.run_tda(4, 300, 1e-6)     # the same x and omega go to both sides
    inputs = scc.excited_gradient_inputs(mol.z, mol.xyz, r)
    g = go.gradient_inputs(mol.z, mol.xyz, r, occ, virt)
    o, v = len(occ), len(virt)
    for state in (0, 2):
        for form in ("reference", "exact"):
            got, z, its = s.tda_gradient(state, inputs, form=form, conv=1e-22, max_iter=3000, return_z=True)
            ref, parts = go.tda_gradient(g, x[:, state].reshape(o, v), w[state], exact=form == "exact",
                                         zvector_conv=1e-22, return_parts=True)
            assert its < 3000
            scale = max(1.0, np.abs(ref).max())
            assert np.abs(got - ref).max() < 1e-10 * scale, (state, form, np.abs(got - ref).max())
            assert np.abs(z - parts["z"].ravel()).max() < 1e-10 * max(1.0, np.abs(z).max()), (state, form)
    s.close()


def test_gradient_reference_tolerances_and_errors():
    """The reference's CG limits (1e4 iterations, r.r <= 1e-19) converge; a
    context without the needed inputs reports an error instead of running."""
    mol = sy.polyene(4, seed=0)
    r, occ, virt = _scc(mol, True)
    s = _system(mol, r, occ, virt)
    s.run_tda(2, 300, 1e-7)
    inputs = scc.excited_gradient_inputs(mol.z, mol.xyz, r)
    g1, _, its = s.tda_gradient(0, inputs, return_z=True)
    assert 0 < its < 10000
    g2 = s.tda_gradient(0, inputs, conv=1e-24, max_iter=5000)
    assert np.abs(g1 - g2).max() < 1e-7
    from tincr_amd import _native as nat
    bad = dict(inputs, dgamma_lr=None)
    with pytest.raises(nat.TdaError):
        s.tda_gradient(0, bad)
    s.close()


def _pick_state(w, n):
    """Lowest of the first n states whose neighbours lie at least 1e-4 Ha
    away (no crossing within the finite-difference steps)."""
    for k in range(n):
        lo = w[k] - w[k - 1] if k > 0 else np.inf
        if min(lo, w[k + 1] - w[k]) > 1e-4:
            return k
    pytest.skip("no isolated state among the lowest ones")


@pytest.mark.parametrize("name", ["C60H62", "(H2O)64"])
def test_exact_gradient_matches_central_differences(name):
    print(f"gradient vs central differences {name}", flush=True)
    mol = sy.polyene(60, seed=0) if name == "C60H62" else sy.water_cluster(64, seed=0)
    conv = 1e-11
    r, occ, virt = _scc(mol, True, conv)
    s = _system(mol, r, occ, virt)
    w, _ = s.run_tda(6, 300, 1e-7)
    state = _pick_state(w, 5)
    inputs = scc.excited_gradient_inputs(mol.z, mol.xyz, r)
    exact = s.tda_gradient(state, inputs, form="exact")
    reference = s.tda_gradient(state, inputs, form="reference")
    s.close()
    # FP32 mode: the state solved by the FP32 ladder (FP32 products, FP64
    # verification); the gradient itself is FP64
    s32 = _system(mol, r, occ, virt, precision="fp32")
    w32, _ = s32.run_tda(6, 300, 1e-7)
    assert abs(w32[state] - w[state]) < 1e-9
    exact32 = s32.tda_gradient(state, inputs, form="exact")
    s32.close()
    assert np.abs(exact32 - exact).max() < 1e-7
    rng = np.random.default_rng(7)
    coords = rng.choice(3 * mol.n_atoms, 32, replace=False)
    h = 1e-4
    err_exact, err_ref = [], []
    for c in coords:
        e = []
        for sign in (1.0, -1.0):
            m = sy.Molecule(mol.z, mol.xyz.copy())
            m.xyz.reshape(-1)[c] += sign * h
            rr, oc, vi = _scc(m, True, conv, p_init=r["p"])
            assert (oc, vi) == (occ, virt)
            sd = _system(m, rr, oc, vi)
            wd, _ = sd.run_tda(6, 300, 1e-7)
            sd.close()
            e.append(wd[state])
        fd = (e[0] - e[1]) / (2 * h)
        err_exact.append(abs(exact[c] - fd))
        err_ref.append(abs(reference[c] - fd))
        print(f"  coord {c}: fd {fd:+.9f} exact {exact[c]:+.9f} reference {reference[c]:+.9f}", flush=True)
    print(f"  state {state}: max |exact - fd| {max(err_exact):.2e}, max |reference - fd| {max(err_ref):.2e}",
          flush=True)
    assert max(err_exact) < 1e-6
