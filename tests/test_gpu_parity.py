"""End-to-end GPU parity of the LC-TDA path against the CPU oracle.

Tolerances (FP64 path): transition charges 1e-14 absolute; A.x 1e-12
relative to |A| |x|; excitation energies 1e-10 Ha against the dense closed form
(the reference's tda_with_lc intent) and 1e-6 eV against the oracle Davidson;
identical state ordering.  Mixed precision: max |omega - omega_FP64| <= 1 meV.
"""
import numpy as np
import pytest

from oracle import tda_oracle as orc
from tests._fixtures import MOLECULES, legacy, molecule

pytestmark = pytest.mark.gpu

HARTREE_TO_EV = orc.HARTREE_TO_EV


def _tda():
    from tincr_amd import tda
    return tda


def _oracle_engine(m, lc=True):
    q_ov, q_oo, q_vv = orc.trans_charges(m["n_atoms"], m["orbs_per_atom"], m["orbs"], m["s"],
                                         m["occ_indices"], m["virt_indices"])
    orbe = np.ravel(m["orbe"])
    homo, lumo = m["occ_indices"][-1], m["virt_indices"][0]
    omega = orc.orbe_differences(orbe[:homo + 1], orbe[lumo:])
    return orc.SystemEngine(q_ov, q_oo, q_vv, m["gamma"], m["gamma_lr"] if lc else None, omega,
                            len(m["occ_indices"]), len(m["virt_indices"]), faithful=False)


def _system(m, precision="fp64", lc=True, **kw):
    tda = _tda()
    return tda.TdaSystem(m["n_atoms"], m["orbs_per_atom"], m["xyz"], m["occ_indices"], m["virt_indices"],
                         m["orbs"], m["s"], m["orbe"], m["gamma"], m["gamma_lr"] if lc else None,
                         precision=precision, **kw)


def _synthetic(builder, n, seed=0):
    from tincr_amd import systems as sy
    mol = builder(n, seed=seed)
    d = sy.tda_inputs(mol, seed=seed)
    d["name"] = mol.name
    d["orbs_per_atom"] = mol.orbs_per_atom
    return d


def _cases():
    return MOLECULES


@pytest.mark.parametrize("mol", MOLECULES)
def test_trans_charges(mol):
    print(f"trans_charges {mol}", flush=True)
    m = molecule(mol)
    tda = _tda()
    got = tda.trans_charges(m["n_atoms"], m["orbs_per_atom"], m["orbs"], m["s"], m["occ_indices"],
                            m["virt_indices"])
    ref = orc.trans_charges(m["n_atoms"], m["orbs_per_atom"], m["orbs"], m["s"], m["occ_indices"],
                            m["virt_indices"])
    for g, r in zip(got, ref):
        assert g.shape == r.shape
        assert np.abs(g - r).max() < 1e-14


def test_trans_charges_legacy_water_golden_path():
    """Charges from the legacy water orbitals reproduce the 8 golden omegas
    (not_refactored_tests/solver.rs:33-199) when fed to the oracle's dense A."""
    g = legacy("legacy_water_tda.json")
    w = molecule("h2o")
    q = _tda().trans_charges(3, [4, 1, 1], g["orbs"], g["S"], [0, 1, 2, 3], [4, 5])
    omega = orc.orbe_differences(g["orbe"][:4], g["orbe"][4:])
    eng = orc.SystemEngine(*q, w["gamma"], w["gamma_lr"], omega, 4, 2)
    assert np.abs(np.linalg.eigh(eng.dense())[0] - g["omega_ref_out"]).max() < 1e-10


@pytest.mark.parametrize("mol", MOLECULES)
@pytest.mark.parametrize("lc", [True, False])
@pytest.mark.parametrize("op", ["ao", "charges"])
def test_products_fixtures(mol, lc, op):
    print(f"products {mol} lc={lc} op={op}", flush=True)
    m = molecule(mol)
    sysg = _system(m, lc=lc, operator=op)
    sysg.prepare_tda()
    eng = _oracle_engine(m, lc)
    ov = eng.get_size()
    x = np.random.default_rng(3).standard_normal((ov, 7))
    ax = sysg.compute_products(x)
    ref = eng.apply(x)
    scale = np.abs(ref).max()
    assert np.abs(ax - ref).max() < 1e-12 * max(scale, 1.0)
    assert np.abs(sysg.properties["q_ov"] - eng.q_ov).max() < 1e-14
    if lc:
        assert np.abs(sysg.properties["q_oo"] - eng.q_oo).max() < 1e-14
        assert np.abs(sysg.properties["q_vv"] - eng.q_vv).max() < 1e-14
    assert np.array_equal(sysg.properties["omega"], eng.omega)
    r = np.random.default_rng(4).standard_normal(ov)
    for w in (eng.omega[3], 0.5):
        assert np.abs(sysg.precondition(r, w) - eng.precondition(r, w)).max() == 0.0
    sysg.close()


def test_products_benzene_apbv_golden():
    """A.x through the GPU charges of benzene's fixture orbitals equals the
    oracle; the oracle itself is pinned by the (A+B).v golden."""
    m = molecule("benzene")
    sysg = _system(m)
    eng = _oracle_engine(m)
    x = np.random.default_rng(5).standard_normal((eng.get_size(), 3))
    assert np.abs(sysg.compute_products(x) - eng.apply(x)).max() < 1e-13
    sysg.close()


@pytest.mark.parametrize("mol", MOLECULES)
@pytest.mark.parametrize("op", ["ao", "charges"])
def test_run_tda_fixtures(mol, op):
    print(f"run_tda {mol} op={op}", flush=True)
    m = molecule(mol)
    sysg = _system(m, operator=op)
    sysg.prepare_tda()
    n = 4
    w, x = sysg.run_tda(n, 200, 1e-8)
    eng = _oracle_engine(m)
    dense = orc.tda_full_matrix(eng, m["xyz"])
    assert np.abs(w - dense["omega"][:n]).max() < 1e-10
    for s in range(n):
        assert abs(abs(x[:, s] @ dense["X"][:, s]) - 1.0) < 1e-6
    ref = orc.run_tda(_oracle_engine(m), m["xyz"], n, 200, 1e-8)
    assert np.abs(w - ref["omega"]).max() * HARTREE_TO_EV < 1e-6
    f = sysg.properties["oscillator_strengths"]
    assert np.abs(f - ref["f"]).max() < 1e-8
    assert np.abs(sysg.properties["q_trans"] - eng.q_ov @ x).max() < 1e-13
    sysg.close()


def test_run_tda_not_converged_raises():
    m = molecule("benzene")
    sysg = _system(m)
    tda = _tda()
    with pytest.raises(tda.DavidsonError):
        sysg.run_tda(4, 1, 1e-8)
    sysg.close()


def test_run_tda_bad_args():
    from tincr_amd import _native as nat
    m = molecule("h2o")
    sysg = _system(m)
    with pytest.raises(nat.TdaError) as e:
        sysg.run_tda(100, 10, 1e-5)   # n_roots > ov
    assert e.value.code == nat.TDA_ERR_BAD_ARG
    sysg.close()


@pytest.mark.parametrize("case", ["water16", "polyene20"])
@pytest.mark.parametrize("op", ["ao", "charges"])
def test_run_tda_synthetic_vs_oracle(case, op):
    """Config-0-sized (H2O)_16 and a C20H22 polyene with synthetic orbitals:
    GPU Davidson vs the oracle Davidson (same guess, same rules)."""
    from tincr_amd import systems as sy
    print(f"synthetic {case} op={op}", flush=True)
    m = _synthetic(sy.water_cluster, 16) if case == "water16" else _synthetic(sy.polyene, 20)
    n = 10
    sysg = _system(m, operator=op)
    w, x = sysg.run_tda(n, 100, 1e-5)
    ref = orc.run_tda(_oracle_engine(m), m["xyz"], n, 100, 1e-5)
    assert np.abs(w - ref["omega"]).max() * HARTREE_TO_EV < 1e-6
    for s in range(n):
        assert abs(abs(x[:, s] @ ref["X"][:, s]) - 1.0) < 1e-4
    st = sysg.last_stats
    assert st["converged"] == 1 and st["n_iter"] == ref["n_iter"]
    sysg.close()


@pytest.mark.parametrize("case,n", [("polyene20", 70), ("polyene20", 100), ("water16", 200)])
def test_run_tda_many_roots(case, n):
    """More roots than the one-workgroup subspace eigensolver holds (64
    roots, 1024 basis vectors; the reference's default is nstates = 1000,
    src/defaults.rs:85): the Davidson basis of up to 15 n vectors is
    diagonalised by rocSOLVER and the new vectors enter in blocks of 64.
    The lowest n excitation energies equal the dense closed form, the CI
    vectors are orthonormal, and the FP16 ladder agrees to 1 meV."""
    from tincr_amd import systems as sy
    print(f"many roots {case} n={n}", flush=True)
    m = _synthetic(sy.water_cluster, 16) if case == "water16" else _synthetic(sy.polyene, 20)
    eng = _oracle_engine(m)
    exact = np.linalg.eigvalsh(eng.dense())[:n]
    sysg = _system(m)
    w, x = sysg.run_tda(n, 300, 1e-6)
    assert sysg.last_stats["converged"] == 1
    assert np.abs(w - exact).max() < 1e-9
    assert np.abs(x.T @ x - np.eye(n)).max() < 1e-10
    sysg.close()
    slp = _system(m, precision="fp16")
    w16, _ = slp.run_tda(n, 300, 1e-6)
    assert np.abs(w16 - exact).max() * HARTREE_TO_EV < 1e-3
    slp.close()


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_ao_pair_factor_table_matches_gather(monkeypatch, prec):
    """The 16-bit density GEMM reads its gamma_lr factors from the AO-pair
    table (FP32, built once per system) instead of gathering them per tile
    through the AO -> atom map (TINCR_TDA_GFULL=0): the same FP32 factors, so
    the same products and the same solve, bit for bit."""
    from tincr_amd import systems as sy
    m = _synthetic(sy.water_cluster, 16)
    ov = len(m["occ_indices"]) * len(m["virt_indices"])
    x = np.random.default_rng(9).standard_normal((ov, 5))
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("TINCR_TDA_GFULL", flag)
        s = _system(m, precision=prec)
        ax = s.compute_products(x)
        w, _ = s.run_tda(6, 200, 1e-6)
        out.append((ax, w, dict(s.last_stats)))
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2]["n_iter"] == out[1][2]["n_iter"]


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
@pytest.mark.parametrize("op", ["ao", "charges"])
def test_products_mixed_precision(prec, op):
    from tincr_amd import systems as sy
    m = _synthetic(sy.water_cluster, 16)
    sysg = _system(m, precision=prec, operator=op)
    eng = _oracle_engine(m)
    x = np.random.default_rng(9).standard_normal((eng.get_size(), 5))
    x /= np.linalg.norm(x, axis=0)
    err = np.abs(sysg.compute_products(x) - eng.apply(x)).max()
    tol = {"fp32": 1e-6, "bf16": 2e-2, "fp16": 3e-3}[prec]
    assert err < tol, err
    sysg.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
@pytest.mark.parametrize("mol", ["uracil", "benzene"])
@pytest.mark.parametrize("op", ["ao", "charges"])
def test_run_tda_mixed_precision(prec, mol, op):
    """Mixed-precision Davidson + FP64 continuation: omega within 1 meV of
    FP64 (in practice the FP64 continuation gives ~1e-10 Ha)."""
    m = molecule(mol)
    s64 = _system(m)
    w64, x64 = s64.run_tda(4, 200, 1e-5)
    slp = _system(m, precision=prec, operator=op)
    w, x = slp.run_tda(4, 200, 1e-5)
    assert np.abs(w - w64).max() * HARTREE_TO_EV < 1e-3
    assert slp.last_stats["converged"] == 1
    assert slp.last_stats["max_residual"] < 1e-5
    s64.close()
    slp.close()


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
@pytest.mark.parametrize("mol", ["uracil", "benzene"])
def test_fp32_basis_of_the_16bit_phase(prec, mol, monkeypatch):
    """The FP16 / BF16 phase keeps its basis V and products AV as FP32
    (TINCR_TDA_BASIS32, default on; every product with them in FP64): the
    solve lands within 1 meV of FP64 with the FP32 basis and with the FP64
    one, and the 16-bit phase did run."""
    m = molecule(mol)
    s64 = _system(m)
    w64, _ = s64.run_tda(6, 300, 1e-6)
    s64.close()
    for b32 in ("1", "0"):
        monkeypatch.setenv("TINCR_TDA_BASIS32", b32)
        print(f"fp32 basis={b32} {mol} {prec}", flush=True)
        slp = _system(m, precision=prec)
        w, _ = slp.run_tda(6, 300, 1e-6)
        st = slp.last_stats
        slp.close()
        assert st["converged"] == 1 and st["max_residual"] < 1e-6
        assert st["n_iter"] - st["n_iter_fp32"] - st["n_iter_fp64"] > 0
        assert np.abs(w - w64).max() * HARTREE_TO_EV < 1e-3


def test_small_system_mixed_context_runs_fp64(monkeypatch):
    """A mixed-precision context on a system whose products cost less than
    TINCR_TDA_LADDER_MIN_FLOPS runs the reference FP64 path: same omega as an
    FP64 context to the last bits, no low-precision phase."""
    m = molecule("uracil")
    s64 = _system(m)
    w64, x64 = s64.run_tda(4, 200, 1e-5)
    monkeypatch.setenv("TINCR_TDA_LADDER_MIN_FLOPS", "1e12")
    slp = _system(m, precision="fp16")
    w, x = slp.run_tda(4, 200, 1e-5)
    st = slp.last_stats
    assert st["n_iter_fp64"] == st["n_iter"] == s64.last_stats["n_iter"]
    assert st["n_products_fp64"] == st["n_products"] and st["split_fp32"] == 0
    assert np.abs(w - w64).max() < 1e-14
    s64.close()
    slp.close()


def test_run_tda_without_host_x():
    """copy_x=False (X = NULL through the C ABI): the CI vectors stay on the
    device; omega, f and the adjoint of the device vectors equal those of a
    run that copies X out."""
    m = molecule("uracil")
    s = _system(m, precision="fp16")
    w1, x1 = s.run_tda(4, 200, 1e-6)
    f1 = s.properties["oscillator_strengths"].copy()
    a1 = s.adjoint(4, x1)
    w2, x2 = s.run_tda(4, 200, 1e-6, copy_x=False)
    assert x2 is None
    assert np.array_equal(w1, w2)
    assert np.array_equal(f1, s.properties["oscillator_strengths"])
    a2 = s.adjoint(4)
    assert np.abs(a1["d_gamma"] - a2["d_gamma"]).max() < 1e-12
    s.close()


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_second_gram_schmidt_pass(monkeypatch, prec):
    """TINCR_TDA_REORTH=1: the mixed phases give every block of new vectors
    the second Gram-Schmidt pass; the solve still lands within 1 meV of FP64."""
    print(f"second gram-schmidt pass {prec}", flush=True)
    m = molecule("benzene")
    s64 = _system(m)
    w64, _ = s64.run_tda(4, 200, 1e-5)
    s64.close()
    monkeypatch.setenv("TINCR_TDA_REORTH", "1")
    slp = _system(m, precision=prec)
    w, _ = slp.run_tda(4, 200, 1e-5)
    assert slp.last_stats["converged"] == 1
    assert np.abs(w - w64).max() * HARTREE_TO_EV < 1e-3
    slp.close()


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
@pytest.mark.parametrize("op", ["ao", "charges"])
def test_fp64_handover_without_second_pass(monkeypatch, prec, op):
    """TINCR_TDA_REORTH=0 (no second Gram-Schmidt pass in the 16-bit phases,
    the setting under which a BF16 phase on benzene once built a nearly
    dependent basis): the FP64 phase orthonormalises its start block and
    accepts only unit Ritz vectors, so the solve still lands within 1 meV of
    FP64 with orthonormal CI vectors."""
    print(f"fp64 hand-over without reorth {prec} {op}", flush=True)
    m = molecule("benzene")
    s64 = _system(m)
    w64, _ = s64.run_tda(4, 200, 1e-5)
    s64.close()
    monkeypatch.setenv("TINCR_TDA_REORTH", "0")
    slp = _system(m, precision=prec, operator=op)
    w, x = slp.run_tda(4, 200, 1e-5)
    assert slp.last_stats["converged"] == 1
    assert np.abs(w - w64).max() * HARTREE_TO_EV < 1e-3
    assert w.min() > 0.1
    assert np.abs(x.T @ x - np.eye(4)).max() < 1e-10
    slp.close()


@pytest.mark.parametrize("prec,n", [("fp64", 10), ("fp16", 10), ("fp64", 40)])
def test_block_gram_schmidt_forms_agree(monkeypatch, prec, n):
    """The in-block Gram-Schmidt of the new Davidson vectors in coefficient
    form (Gram matrices, the default) and vector by vector
    (TINCR_TDA_GS_GRAM=0, the reference's loop) give the same solve; 40
    roots take blocks of more than 32 new vectors."""
    from tincr_amd import systems as sy
    print(f"gram-schmidt forms {prec} n={n}", flush=True)
    m = _synthetic(sy.polyene, 20)
    out = {}
    for form in ("0", "1"):
        monkeypatch.setenv("TINCR_TDA_GS_GRAM", form)
        sysg = _system(m, precision=prec)
        w, x = sysg.run_tda(n, 300, 1e-6)
        out[form] = (w, x, dict(sysg.last_stats))
        sysg.close()
    (w0, x0, st0), (w1, x1, st1) = out["0"], out["1"]
    assert st0["converged"] == st1["converged"] == 1
    assert abs(st0["n_iter"] - st1["n_iter"]) <= (0 if prec == "fp64" else 2)
    assert np.abs(w0 - w1).max() < (1e-12 if prec == "fp64" else 1e-9)
    for s in range(n):
        assert abs(abs(x0[:, s] @ x1[:, s]) - 1.0) < 1e-8


@pytest.mark.parametrize("mol", ["h2o", "benzene", "uracil"])
@pytest.mark.parametrize("lc,op", [(True, "ao"), (True, "charges"), (False, "charges")])
def test_zvector_vs_oracle(mol, lc, op):
    """tda_zvector (CG on the forward product, helpers.rs:1130-1223 /
    :1053-1128) against the oracle's CG and a dense solve (without gamma_lr
    the operator is the charges form)."""
    m = molecule(mol)
    print(f"zvector {mol} lc={lc} op={op}", flush=True)
    eng = _oracle_engine(m, lc=lc)
    n = eng.get_size()
    rhs = np.random.default_rng(n + 1).standard_normal((n, 3))
    z_ref, its_ref = orc.zvector(eng, rhs)
    s = _system(m, lc=lc, operator=op)
    z, its = s.zvector(rhs)
    assert np.abs(z - z_ref).max() < 1e-9 * np.abs(z_ref).max()
    a = eng.dense()
    if not lc:
        a = a + 2.0 * eng.q_ov.T @ eng.gamma @ eng.q_ov
    assert np.linalg.norm(a @ z - rhs, axis=0).max() < 1e-9
    assert np.all(np.abs(np.asarray(its) - np.asarray(its_ref)) <= 2)
    # a single column and a capped iteration count
    z1, its1 = s.zvector(rhs[:, 0], max_iter=3)
    assert its1[0] == 3 and np.isfinite(z1).all()
    s.close()


def test_exchange_atom_blocking():
    """A tiny HBM budget forces several atom blocks in the exchange."""
    from tincr_amd import systems as sy
    m = _synthetic(sy.water_cluster, 16)
    eng = _oracle_engine(m)
    x = np.random.default_rng(11).standard_normal((eng.get_size(), 4))
    full = _system(m, operator="charges")
    blocked = _system(m, hbm_budget_bytes=4 * 64 * 64 * 8 * 5, operator="charges")
    a1 = full.compute_products(x)
    a2 = blocked.compute_products(x)
    assert np.abs(a1 - eng.apply(x)).max() < 1e-12
    assert np.abs(a2 - a1).max() < 1e-13
    full.close()
    blocked.close()


def test_repeated_solves_switch_precision():
    """One context, several solves: FP16 -> FP16 (charges cleared) -> FP64 ->
    FP16.  The exchange intermediate is reused across precisions; its pad rows
    (uracil: n_occ = 21, o_p = 64) must never leak stale bytes into a sum."""
    m = molecule("uracil")
    ref = _system(m)
    w64, _ = ref.run_tda(4, 200, 1e-5)
    s = _system(m, precision="fp16", operator="charges")
    for prec, op in [("fp16", "charges"), ("fp16", "charges"), ("fp64", "charges"), ("fp16", "charges"),
                     ("bf16", "ao"), ("fp32", "ao"), ("fp16", "ao"), ("fp64", "ao"), ("fp16", "charges")]:
        s.ctx.set_precision(prec)
        s.ctx.set_operator(op)
        s.clear_charges()
        w, _ = s.run_tda(4, 200, 1e-5)
        assert np.abs(w - w64).max() * HARTREE_TO_EV < 1e-3, (prec, op)
    ref.close()
    s.close()


def test_ao_small_workspace_chunks():
    """A tiny HBM budget forces the AO operator to split the vectors into
    several workspace chunks."""
    from tincr_amd import systems as sy
    m = _synthetic(sy.water_cluster, 16)
    eng = _oracle_engine(m)
    x = np.random.default_rng(12).standard_normal((eng.get_size(), 7))
    s = _system(m, hbm_budget_bytes=3 * 1024 * 1024)
    assert np.abs(s.compute_products(x) - eng.apply(x)).max() < 1e-12
    s.close()


# ---------------------------------------------------------------------------
# charge-transfer pairs (tda/ct_pair.rs, solvers/mod.rs:165-274)
# ---------------------------------------------------------------------------
_FRAG_KEYS = ("n_atoms", "orbs_per_atom", "xyz", "occ_indices", "virt_indices", "orbs", "s", "orbe", "gamma")


def _ct_case(n_waters, n_h, seed=0):
    from tincr_amd import systems as sy
    return sy.ct_pair_inputs(sy.water_cluster(n_waters, seed=seed), 3 * n_h, seed=seed)


def _ct_pair(ct, precision="fp64", **kw):
    tda = _tda()
    h = tda.Fragment(**{k: ct["h"][k] for k in _FRAG_KEYS})
    l = tda.Fragment(**{k: ct["l"][k] for k in _FRAG_KEYS})
    return tda.TdaCTPair(h, l, ct["s_hl"], ct["gamma_hl"], ct["gamma_lr_hl"], precision=precision, **kw)


def _ct_oracle(ct):
    return orc.ct_pair_engine(ct["h"], ct["l"], ct["s_hl"], ct["gamma_hl"], ct["gamma_lr_hl"])


def _ct_xyz(ct):
    return np.concatenate([ct["h"]["xyz"], ct["l"]["xyz"]])


@pytest.mark.parametrize("case", [(3, 1), (4, 3), (12, 5)])
def test_ct_pair_products(case):
    """Pair charges (q_ov over H u L, q_oo of H, q_vv of L), omega and A.x
    against the oracle's CT engine; N_H != N_L in every case."""
    print(f"ct products {case}", flush=True)
    ct = _ct_case(*case)
    p = _ct_pair(ct)
    eng = _ct_oracle(ct)
    p.prepare_ct_tda()
    assert np.abs(p.properties["q_ov"] - eng.q_ov).max() < 1e-14
    assert np.abs(p.properties["q_oo"] - eng.q_oo).max() < 1e-14
    assert np.abs(p.properties["q_vv"] - eng.q_vv).max() < 1e-14
    assert np.array_equal(p.properties["omega"], eng.omega)
    x = np.random.default_rng(21).standard_normal((eng.get_size(), 6))
    ref = eng.apply(x)
    assert np.abs(p.compute_products(x) - ref).max() < 1e-12 * max(np.abs(ref).max(), 1.0)
    r = np.random.default_rng(22).standard_normal(eng.get_size())
    assert np.abs(p.precondition(r, eng.omega[2]) - eng.precondition(r, eng.omega[2])).max() == 0.0
    p.close()


def test_ct_pair_run_vs_oracle():
    """run_ct_tda (ct_pair.rs:125-180) against the oracle Davidson on the same
    engine: same iteration count, omega, X, q_trans, f."""
    ct = _ct_case(12, 5)
    p = _ct_pair(ct)
    n = 6
    w, x = p.run_ct_tda(n, 100, 1e-6)
    ref = orc.run_tda(_ct_oracle(ct), _ct_xyz(ct), n, 100, 1e-6)
    assert np.abs(w - ref["omega"]).max() * HARTREE_TO_EV < 1e-6
    for s in range(n):
        assert abs(abs(x[:, s] @ ref["X"][:, s]) - 1.0) < 1e-5
    assert p.last_stats["n_iter"] == ref["n_iter"]
    assert np.abs(p.properties["oscillator_strengths"] - ref["f"]).max() < 1e-8
    dense = orc.tda_full_matrix(_ct_oracle(ct), _ct_xyz(ct))
    assert np.abs(w - dense["omega"][:n]).max() < 1e-9
    p.close()


def test_ct_pair_equals_supermolecule_on_gpu():
    """The GPU CT pair equals the GPU monomer path on the supermolecule with
    block-diagonal orbitals diag(C_H, C_L) (both operator forms)."""
    ct = _ct_case(4, 3)
    h, l = ct["h"], ct["l"]
    nao_h, nao_l = h["orbs"].shape[0], l["orbs"].shape[0]
    c = np.zeros((nao_h + nao_l, nao_h + nao_l))
    c[:nao_h, :nao_h] = h["orbs"]
    c[nao_h:, nao_h:] = l["orbs"]
    sup = dict(n_atoms=h["n_atoms"] + l["n_atoms"],
               orbs_per_atom=np.concatenate([h["orbs_per_atom"], l["orbs_per_atom"]]), xyz=_ct_xyz(ct),
               occ_indices=list(h["occ_indices"]), virt_indices=[nao_h + a for a in l["virt_indices"]],
               orbs=c, s=ct["s"], orbe=np.concatenate([h["orbe"], l["orbe"]]), gamma=ct["gamma"],
               gamma_lr=ct["gamma_lr"])
    p = _ct_pair(ct)
    wp, _ = p.run_ct_tda(4, 100, 1e-7)
    for op in ("ao", "charges"):
        m = _system(sup, operator=op)
        wm, _ = m.run_tda(4, 100, 1e-7)
        assert np.abs(wp - wm).max() < 1e-10, op
        m.close()
    p.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_ct_pair_mixed_precision(prec):
    ct = _ct_case(12, 5)
    p64 = _ct_pair(ct)
    w64, _ = p64.run_ct_tda(4, 200, 1e-5)
    plp = _ct_pair(ct, precision=prec)
    w, _ = plp.run_ct_tda(4, 200, 1e-5)
    assert np.abs(w - w64).max() * HARTREE_TO_EV < 1e-3
    assert plp.last_stats["converged"] == 1 and plp.last_stats["max_residual"] < 1e-5
    p64.close()
    plp.close()


def test_ct_pair_atom_blocking_and_context_reuse():
    """A small HBM budget splits the pair exchange over L-atom blocks; the same
    context then loads a monomer and runs it with the AO operator again."""
    from tincr_amd import _native as nat
    ct = _ct_case(12, 5)
    eng = _ct_oracle(ct)
    x = np.random.default_rng(23).standard_normal((eng.get_size(), 4))
    p = _ct_pair(ct, hbm_budget_bytes=4 * 64 * 64 * 8 * 3)
    assert np.abs(p.compute_products(x) - eng.apply(x)).max() < 1e-12
    m = molecule("h2o")
    ctx = p.ctx
    opa = np.ascontiguousarray(m["orbs_per_atom"], dtype=np.int32)
    orbs, s_, orbe = nat.f64(m["orbs"]), nat.f64(m["s"]), nat.f64(np.ravel(m["orbe"]))
    g, glr, xyz = nat.f64(m["gamma"]), nat.f64(m["gamma_lr"]), nat.f64(np.asarray(m["xyz"]).reshape(-1, 3))
    o, v = len(m["occ_indices"]), len(m["virt_indices"])
    ctx.check(ctx.lib.tda_set_system(ctx.handle, m["n_atoms"], orbs.shape[0], nat.iptr(opa), nat.dptr(orbs),
                                     nat.dptr(s_), nat.dptr(orbe), m["occ_indices"][0], o, m["virt_indices"][0], v,
                                     nat.dptr(g), nat.dptr(glr), nat.dptr(xyz)))
    xm = np.random.default_rng(24).standard_normal((o * v, 3))
    axm = np.empty_like(xm)
    ctx.check(ctx.lib.tda_compute_products(ctx.handle, nat.dptr(xm), 3, nat.dptr(axm)))
    assert np.abs(axm - _oracle_engine(m).apply(xm)).max() < 1e-12
    p.close()


# ---------------------------------------------------------------------------
# dense-matrix engine (solvers/mod.rs:36-51, as fmo/lcmo/basis.rs:497-503 uses it)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("dim", [100, 333])
def test_dense_engine_vs_oracle(seed, dim):
    """Diagonally dominant matrices of solvers/utils.rs:29-42: products,
    preconditioner (no floor) and the Davidson against the oracle Davidson on
    the oracle's dense engine (same iteration count) and numpy eigh."""
    print(f"dense seed={seed} dim={dim}", flush=True)
    tda = _tda()
    a = orc.generate_diagonal_dominant(np.random.default_rng(seed), dim, 0.005)
    eng = tda.DenseEngine(a)
    ref = orc.DenseEngine(a)
    x = np.random.default_rng(seed + 10).standard_normal((dim, 5))
    assert np.abs(eng.compute_products(x) - ref.compute_products(x)).max() < 1e-13
    r = np.random.default_rng(seed + 20).standard_normal(dim)
    w0 = np.diag(a)[3] + 1e-6
    assert np.abs(eng.precondition(r, w0) - ref.precondition(r, w0)).max() == 0.0
    n = 4
    w, v = eng.davidson(n, 1e-7, 200)
    guess = orc.initial_subspace(np.diag(a), n)
    w_o, v_o, it_o = orc.davidson(ref, guess, n, 1e-7, 200)
    assert np.abs(w - w_o).max() < 1e-12
    assert eng.last_stats["n_iter"] == it_o
    w_ref, x_ref = np.linalg.eigh(a)
    assert np.allclose(w, w_ref[:n], rtol=1e-6, atol=0)
    for s in range(n):
        assert abs(abs(v[:, s] @ x_ref[:, s]) - 1.0) < 1e-6
    eng.close()


def test_dense_engine_rejects_tda_outputs():
    from tincr_amd import _native as nat
    a = orc.generate_diagonal_dominant(np.random.default_rng(5), 50, 0.005)
    eng = _tda().DenseEngine(a)
    f = np.empty(2)
    rc = eng.ctx.lib.tda_run(eng.ctx.handle, 2, 50, 1e-6, None, None, None, None, nat.dptr(f), None)
    assert rc == nat.TDA_ERR_BAD_ARG
    rc = eng.ctx.lib.tda_get_charges(eng.ctx.handle, None, None, None, None)
    assert rc == nat.TDA_ERR_STATE
    eng.close()


def test_fp16_ao_product_scaled_into_range(monkeypatch):
    """FP16 AO products on SCC orbitals ((H2O)64): the power-of-two scale of
    the first GEMM's output (TINCR_TDA_AO_SCALE=1; off by default) keeps the
    middle matrix in FP16's normal range.  Against the FP64 product of the
    same unit vectors (random, localised on the lowest orbital pairs,
    smooth in the orbital-energy gaps), the scaled product's relative error
    is no larger than the unscaled one and below 5e-5."""
    from tincr_amd import scc, systems as sy
    m = scc.tda_inputs_scc(sy.water_cluster(64, seed=0), blas_threads=4)
    o, v = len(m["occ_indices"]), len(m["virt_indices"])
    de = m["orbe"][m["virt_indices"]][None, :] - m["orbe"][m["occ_indices"]][:, None]
    rng = np.random.default_rng(3)
    x = np.zeros((o * v, 3))
    x[:, 0] = rng.standard_normal(o * v)
    low = np.argsort(de.ravel())[:32]
    x[low, 1] = rng.standard_normal(32)
    x[:, 2] = (np.exp(-(de - de.min()) / 0.02) * rng.standard_normal((o, v))).ravel()
    x /= np.linalg.norm(x, axis=0)
    s64 = _system(m)
    ref = s64.compute_products(x)
    s64.close()
    err = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("TINCR_TDA_AO_SCALE", flag)
        s = _system(m, precision="fp16", operator="ao")
        y = s.compute_products(x)
        s.close()
        err[flag] = np.linalg.norm(y - ref, axis=0) / np.linalg.norm(ref, axis=0)
    print("relative error scaled", err["1"], "unscaled", err["0"], flush=True)
    assert np.all(err["1"] <= err["0"] * 1.05)
    assert err["1"].max() < 5e-5
