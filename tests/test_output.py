"""Output formats (tincr_amd/output.py) on the CPU: the TdaStates summary,
the .npy results, the spectrum files, natural transition orbitals and their
Molden export, checked on oracle solutions of the fixture molecules.

No reference output file exists to diff against (the reference cannot run
here), so the tests pin the content: every number printed reads back to the
solution at its printed precision, the amplitude lines are exactly those
above the 0.1 threshold, the NTO weights sum to |x|^2 = 1 and reconstruct the
transition density, NTOs are S-orthonormal, and the Molden MO block reads
back to the NTO coefficients in Cartesian p order.  Rust's float formats are
checked on known strings.
"""
import re

import numpy as np
import pytest

from oracle import tda_oracle as orc
from tests._fixtures import load_dat, molecule
from tincr_amd import output


def _solution(mol, n=4):
    m = molecule(mol)
    q = orc.trans_charges(m["n_atoms"], m["orbs_per_atom"], m["orbs"], m["s"], m["occ_indices"],
                          m["virt_indices"])
    o, v = len(m["occ_indices"]), len(m["virt_indices"])
    orbe = m["orbe"]
    eng = orc.SystemEngine(*q, m["gamma"], m["gamma_lr"], orc.orbe_differences(orbe[:o], orbe[o:]), o, v,
                           faithful=False)
    r = orc.run_tda(eng, m["xyz"], n, 100, 1e-8)
    return m, r, o, v


def test_rust_exponent_format():
    assert output.rust_exp(3.425250914, 14) == "3.42525091400000e0"
    assert output.rust_exp(-0.09996722919, 14) == "-9.99672291900000e-2"
    assert output.rust_exp(1.0, 2) == "1.00e0"
    assert output.rust_exp(12345.0, 3) == "1.234e4"


@pytest.mark.parametrize("mol", ["h2o", "uracil"])
def test_tda_states_text(mol):
    m, r, o, v = _solution(mol)
    tdm = r["X"].reshape(o, v, -1)
    e_tot = -12.3456789
    txt = output.tda_states_text(e_tot, r["omega"], tdm, r["f"], r["tr_dipoles"])
    lines = txt.split("\n")
    assert lines[0] == " " * 80 and lines[1].strip() == "TDA Excitation Energies" and lines[2] == "-" * 75
    exc = [float(x) for x in re.findall(r"Excitation energy = +([0-9.]+) eV", txt)]
    assert np.allclose(exc, r["omega"] * output.HARTREE_TO_EV, atol=5e-7, rtol=0)
    tot = [float(x) for x in re.findall(r"Total energy for state +\d+: +([-0-9.]+) Hartree", txt)]
    assert np.allclose(tot, e_tot + r["omega"], atol=5e-13, rtol=0)
    osc = [float(x) for x in re.findall(r"Oscillator Strength: +([-0-9.]+)", txt)]
    assert np.allclose(osc, r["f"], atol=5e-9, rtol=0)
    blocks = txt.split("Excited state")[1:]
    for k, blk in enumerate(blocks):
        amps = re.findall(r"(H(?:-\d+)?) +--> (L(?:\+\d+)?) +Amplitude: ([0-9.]+) => +([0-9.]+) %", blk)
        big = np.argwhere(np.abs(tdm[::-1, :, k]) > 0.1)       # HOMO first
        assert len(amps) == len(big)
        for (h, l_, a, pct), (i, j) in zip(amps, big):
            assert h == ("H" if i == 0 else f"H-{i}") and l_ == ("L" if j == 0 else f"L+{j}")
            assert abs(float(a) - abs(tdm[o - 1 - i, j, k])) <= 5e-5
            assert abs(float(pct) - 100 * tdm[o - 1 - i, j, k] ** 2) <= 0.05 + 1e-9
    assert txt.endswith("All transition with amplitudes > 0.10000000 were printed.\n" + "-" * 75 + "\n")


def test_npy_and_spectrum_files(tmp_path):
    _, r, _, _ = _solution("uracil")
    output.write_tda_npy(tmp_path, r["omega"], r["f"])
    assert np.array_equal(np.load(tmp_path / "full_energies.npy"), r["omega"])
    assert np.array_equal(np.load(tmp_path / "oscillator_strengths.npy"), r["f"])
    output.spectrum_to_npy(tmp_path / "spec.npy", r["omega"], r["f"])
    spec = np.load(tmp_path / "spec.npy")
    assert spec.shape == (4, 2) and np.array_equal(spec[:, 1], r["f"])
    assert np.allclose(spec[:, 0], r["omega"] * output.HARTREE_TO_EV, rtol=1e-15)
    output.spectrum_to_txt(tmp_path / "spec.txt", r["omega"], r["f"])
    rows = open(tmp_path / "spec.txt").read().split("\n")
    assert rows[:2] == ["# Absorption spectrum", "# exc. energy / eV    osc. strength"]
    back = np.loadtxt(tmp_path / "spec.txt")
    assert np.allclose(back, spec, atol=1e-13, rtol=0)


@pytest.mark.parametrize("mol", ["h2o", "uracil", "benzene"])
def test_natural_transition_orbitals(mol):
    m, r, o, v = _solution(mol)
    s = m["s"]
    lumo = m["virt_indices"][0]
    for k in range(2):
        t = r["X"][:, k].reshape(o, v)
        lam, nto = output.natural_transition_orbitals(t, m["orbs"], lumo)
        p = min(o, v)
        assert lam.shape == (2 * p,) and nto.shape == (m["orbs"].shape[0], 2 * p)
        assert abs(lam[p:].sum() - 1.0) < 1e-12                  # |x|^2 = 1
        assert np.array_equal(lam[:p], lam[p:][::-1])
        assert np.abs(nto.T @ s @ nto - np.eye(2 * p)).max() < 1e-10
        # the hole / particle pair of weight lam reconstructs T in the MO basis
        c_occ = m["orbs"][:, :lumo]
        c_vir = m["orbs"][:, lumo:]
        hole = np.linalg.solve(c_occ.T @ s @ c_occ, c_occ.T @ s @ nto[:, :p][:, ::-1])     # U columns, by weight
        part = np.linalg.solve(c_vir.T @ s @ c_vir, c_vir.T @ s @ nto[:, p:])
        t_back = (hole * np.sqrt(lam[p:])[None, :]) @ part.T
        assert np.abs(t_back - t).max() < 1e-10


def test_nto_crop_rule():
    rng = np.random.default_rng(1)
    for o, v in ((5, 3), (3, 5), (4, 4)):
        c, _ = np.linalg.qr(rng.standard_normal((o + v, o + v)))
        lam, nto = output.natural_transition_orbitals(rng.standard_normal((o, v)), c, o)
        assert lam.shape == (2 * min(o, v),) and nto.shape == (o + v, 2 * min(o, v))


def test_molden_export(tmp_path):
    m, r, o, v = _solution("uracil")
    lumo = m["virt_indices"][0]
    t = r["X"][:, 0].reshape(o, v)
    path = tmp_path / "nto.molden"
    lam, nto = output.ntos_to_molden(path, m["z"], m["xyz"], t, m["orbs"], lumo, title="uracil S1")
    txt = open(path).read()
    assert txt.startswith("[Molden Format]\n[Title]\nuracil S1\n[Atoms] AU\n")
    atoms = re.findall(r"^([A-Z]) (\d+) (\d+) ([-0-9.]+) ([-0-9.]+) ([-0-9.]+)$", txt, re.M)
    assert len(atoms) == m["n_atoms"]
    assert np.allclose(np.array([[float(a[3]), float(a[4]), float(a[5])] for a in atoms]), m["xyz"], atol=5e-8)
    n_gto = txt.split("[GTO]\n")[1].split("[5D]")[0].count(" 0\n")
    assert n_gto == m["n_atoms"]
    mos = txt.split("[MO]\n")[1].split(" Sym=")[1:]
    assert len(mos) == nto.shape[1]
    order = []
    for za in m["z"]:
        base = len(order)
        order += [base] if za == 1 else [base, base + 3, base + 1, base + 2]
    for i, blk in enumerate(mos):
        coef = np.array([float(x) for x in re.findall(r"^ +\d+ ([-0-9.]+)$", blk, re.M)])
        assert np.allclose(coef, nto[order, i], atol=5e-8)
        occ = float(re.search(r"Occup=([-0-9.]+)", blk).group(1))
        assert abs(occ - lam[i]) <= 5e-8
