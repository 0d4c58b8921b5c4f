"""Batched FMO fragments (fmo/lcmo/basis.rs:172-238, 462-465) on the GPU:
every item of a batch equals the same fragment solved alone and the oracle."""
import numpy as np
import pytest

from oracle import tda_oracle as orc

pytestmark = pytest.mark.gpu

HARTREE_TO_EV = orc.HARTREE_TO_EV
_FRAG_KEYS = ("n_atoms", "orbs_per_atom", "xyz", "occ_indices", "virt_indices", "orbs", "s", "orbe", "gamma")


def _workload(n_frag, n_pairs, seed=0):
    from tincr_amd import systems as sy
    frags = sy.fmo_fragments(n_frag, seed)
    return sy.fmo_inputs(frags, sy.fmo_pairs(frags, n_pairs), seed)


def _frag(d):
    from tincr_amd import tda
    return tda.Fragment(**{k: d[k] for k in _FRAG_KEYS})


def _n_roots(o, v, n=4):
    return min(n, o * v)


def _fill(batch, mons, cts, n, max_iter, tol):
    for m in mons:
        batch.add_system(m["n_atoms"], m["orbs_per_atom"], m["xyz"], m["occ_indices"], m["virt_indices"],
                         m["orbs"], m["s"], m["orbe"], m["gamma"], m["gamma_lr"],
                         _n_roots(len(m["occ_indices"]), len(m["virt_indices"]), n), max_iter, tol)
    for c in cts:
        batch.add_ct_pair(_frag(c["h"]), _frag(c["l"]), c["s_hl"], c["gamma_hl"], c["gamma_lr_hl"],
                          _n_roots(len(c["h"]["occ_indices"]), len(c["l"]["virt_indices"]), n), max_iter, tol)


@pytest.mark.parametrize("mode", ["grouped", "pool"])
def test_batch_equals_single_solves_and_oracle(mode):
    """pool: every item equals the fragment solved alone (same operator,
    same Davidson) to 1e-12.  grouped (one Davidson for all fragments,
    charges-form operator, two-pass Gram-Schmidt): the same roots to 1e-9 Ha
    and f to 1e-8 (a different rounding history converged to the same
    tolerance), the oracle's iteration counts on the sampled items."""
    from tincr_amd import tda
    mons, cts = _workload(24, 16)
    print(f"batch {mode} {len(mons)} monomers {len(cts)} pairs", flush=True)
    b = tda.TdaBatch("fp64", n_workers=4, mode=mode)
    _fill(b, mons, cts, 4, 150, 1e-6)
    res = b.run()
    assert len(res) == len(mons) + len(cts)
    tw, tf = (1e-12, 1e-12) if mode == "pool" else (1e-9, 1e-8)
    for k, m in enumerate(mons):
        r = res[k]
        n = len(r["ci_eigenvalues"])
        s = tda.TdaSystem(m["n_atoms"], m["orbs_per_atom"], m["xyz"], m["occ_indices"], m["virt_indices"],
                          m["orbs"], m["s"], m["orbe"], m["gamma"], m["gamma_lr"])
        w, x = s.run_tda(n, 150, 1e-6)
        assert np.abs(r["ci_eigenvalues"] - w).max() < tw
        assert np.abs(r["oscillator_strengths"] - s.properties["oscillator_strengths"]).max() < tf
        assert r["status"] == 0
        if mode == "pool":
            assert r["stats"]["n_iter"] == s.last_stats["n_iter"]
        # same CI vectors up to a sign per root
        ov = np.abs(np.sum(r["ci_coefficients"] * x, axis=0))
        assert np.abs(ov - 1.0).max() < 1e-6
        s.close()
        if k < 6:
            q_ov, q_oo, q_vv = orc.trans_charges(m["n_atoms"], m["orbs_per_atom"], m["orbs"], m["s"],
                                                 m["occ_indices"], m["virt_indices"])
            o, v = len(m["occ_indices"]), len(m["virt_indices"])
            omega = orc.orbe_differences(m["orbe"][:o], m["orbe"][o:])
            eng = orc.SystemEngine(q_ov, q_oo, q_vv, m["gamma"], m["gamma_lr"], omega, o, v, faithful=False)
            ref = orc.run_tda(eng, m["xyz"], n, 150, 1e-6)
            assert np.abs(r["ci_eigenvalues"] - ref["omega"]).max() * HARTREE_TO_EV < 1e-6
            assert r["stats"]["n_iter"] == ref["n_iter"]
    for k, c in enumerate(cts):
        r = res[len(mons) + k]
        n = len(r["ci_eigenvalues"])
        p = tda.TdaCTPair(_frag(c["h"]), _frag(c["l"]), c["s_hl"], c["gamma_hl"], c["gamma_lr_hl"])
        w, _ = p.run_ct_tda(n, 150, 1e-6)
        assert np.abs(r["ci_eigenvalues"] - w).max() < tw
        assert r["status"] == 0
        p.close()
        if k < 4:
            eng = orc.ct_pair_engine(c["h"], c["l"], c["s_hl"], c["gamma_hl"], c["gamma_lr_hl"])
            ref = orc.run_tda(eng, np.concatenate([c["h"]["xyz"], c["l"]["xyz"]]), n, 150, 1e-6)
            assert np.abs(r["ci_eigenvalues"] - ref["omega"]).max() * HARTREE_TO_EV < 1e-6
            assert r["stats"]["n_iter"] == ref["n_iter"]
    b.close()


@pytest.mark.parametrize("mode", ["grouped", "pool"])
@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_batch_mixed_precision(prec, mode):
    from tincr_amd import tda
    mons, cts = _workload(24, 16, seed=1)
    print(f"batch {mode} {prec}", flush=True)
    b64 = tda.TdaBatch("fp64", n_workers=4, mode=mode)
    _fill(b64, mons, cts, 8, 150, 1e-5)
    r64 = b64.run()
    blp = tda.TdaBatch(prec, n_workers=4, mode=mode)
    _fill(blp, mons, cts, 8, 150, 1e-5)
    rlp = blp.run()
    dw = max(np.abs(a["ci_eigenvalues"] - c["ci_eigenvalues"]).max() for a, c in zip(rlp, r64))
    assert dw * HARTREE_TO_EV < 1e-3
    assert all(r["stats"]["converged"] == 1 for r in rlp)
    b64.close()
    blp.close()


@pytest.mark.parametrize("mode", ["grouped", "pool"])
def test_batch_reports_unconverged_items(mode):
    from tincr_amd import tda
    mons, cts = _workload(6, 2, seed=2)
    b = tda.TdaBatch("fp64", n_workers=2, mode=mode)
    _fill(b, mons, [], 4, 150, 1e-6)
    big = max(range(len(mons)), key=lambda i: len(mons[i]["occ_indices"]) * len(mons[i]["virt_indices"]))
    m = mons[big]
    b.add_system(m["n_atoms"], m["orbs_per_atom"], m["xyz"], m["occ_indices"], m["virt_indices"], m["orbs"],
                 m["s"], m["orbe"], m["gamma"], m["gamma_lr"], 4, 1, 1e-6)
    with pytest.raises(tda.DavidsonError):
        b.run()
    b.close()
