"""Output formats of the excited-state path (host side).

What the reference writes after a TDA solve:

* the ``TdaStates`` summary printed by System::run_tda
  (src/excited_states/tda/new_mod.rs:52-142, printed at tda/mod.rs:128):
  per state the excitation energy (eV), total energy, transition dipole,
  oscillator strength and the CI amplitudes above 0.1 as H-i --> L+a lines;
* ``full_energies.npy`` and ``oscillator_strengths.npy`` (tda/mod.rs:118-119);
* the absorption spectrum as .npy ([n, 2]: eV, f) and as text
  (src/excited_states/mod.rs:46-81);
* natural transition orbitals of a state (src/excited_states/ntos.rs:38-85)
  written as a Molden file (src/excited_states/mod.rs:84-113,
  src/io/molden.rs:8-172, STO-3G valence shells of src/param/basis_sets).

The numeric formats follow Rust's formatting of the reference's format
strings: ``{:w.p}`` of an f64 is fixed notation right-aligned in w columns,
``{:w.pe}`` is Rust's exponent notation (mantissa with p decimals, then
``e`` and the exponent without sign padding or zero fill).
"""
import datetime
import json
import os

import numpy as np

HARTREE_TO_EV = 27.211386245988     # src/constants.rs:5
AMPLITUDE_THRESHOLD = 0.1           # tda/new_mod.rs:54
SYMBOL = {1: "H", 6: "C", 7: "N", 8: "O"}
_PARAMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "params")


def rust_exp(x, prec):
    """Rust's ``{:.prec e}`` of an f64, e.g. 3.42525091400000e0, 1.5e-1."""
    if x == 0.0:
        return f"{0.0:.{prec}f}e0"
    m, e = f"{x:.{prec}e}".split("e")
    return f"{m}e{int(e)}"


# ---------------------------------------------------------------------------
# TdaStates text (tda/new_mod.rs:52-142)
# ---------------------------------------------------------------------------
def tda_states_text(total_energy, energies, tdm, f, tr_dip, threshold=AMPLITUDE_THRESHOLD):
    """The TdaStates summary.  energies [n] (Ha), tdm [n_occ, n_virt, n]
    (CI vectors as transition densities), f [n], tr_dip [3, n] (a.u.);
    total_energy: ground-state total energy (Ha)."""
    energies = np.asarray(energies)
    tdm = np.asarray(tdm)
    txt = f"{'':^80}\n"
    txt += f"{'TDA Excitation Energies':^80}\n"
    txt += "-" * 75 + "\n"
    n = len(energies)
    for k in range(n):
        abs_energy = total_energy + energies[k]
        txt += f"Excited state {k + 1:>5}: Excitation energy = {energies[k] * HARTREE_TO_EV:>8.6f} eV\n"
        txt += f"Total energy for state {k + 1:>5}: {abs_energy:22.12f} Hartree\n"
        txt += "  Multiplicity: Singlet\n"
        d = tr_dip[:, k]
        txt += f"  Trans. Mom. (a.u.): {d[0]:10.6f} X  {d[1]:10.6f} Y  {d[2]:10.6f} Z\n"
        txt += f"  Oscillator Strength:  {f[k]:12.8f}\n"
        t = tdm[:, :, k]
        # occupied orbitals from the HOMO down, virtual ones from the LUMO up
        for h, row in enumerate(t[::-1]):
            occ = "H" if h == 0 else f"H-{h}"
            for a, val in enumerate(row):
                if abs(val) > threshold:
                    virt = "L" if a == 0 else f"L+{a}"
                    txt += f"  {occ:<4} --> {virt:<4}  Amplitude: {abs(val):6.4f} => {val * val * 1e2:>4.1f} %\n"
        if k + 1 < n:
            txt += f"{'':^80}\n"
    txt += f"All transition with amplitudes > {threshold:10.8f} were printed.\n"
    txt += "-" * 75 + "\n"
    return txt


def write_tda_npy(directory, energies, f):
    """full_energies.npy and oscillator_strengths.npy (tda/mod.rs:118-119)."""
    np.save(os.path.join(directory, "full_energies.npy"), np.asarray(energies, dtype=np.float64))
    np.save(os.path.join(directory, "oscillator_strengths.npy"), np.asarray(f, dtype=np.float64))


# ---------------------------------------------------------------------------
# spectrum (excited_states/mod.rs:46-81)
# ---------------------------------------------------------------------------
def spectrum_to_npy(filename, energies, f):
    """[n, 2]: excitation energy (eV), oscillator strength."""
    data = np.stack([np.asarray(energies) * HARTREE_TO_EV, np.asarray(f)], axis=1)
    np.save(filename, data)


def spectrum_text(energies, f):
    txt = "# Absorption spectrum\n# exc. energy / eV    osc. strength\n"
    for e, s in zip(energies, f):
        txt += f"{e * HARTREE_TO_EV:16.14f}      {s:16.14f}\n"
    return txt


def spectrum_to_txt(filename, energies, f):
    with open(filename, "w") as fh:
        fh.write(spectrum_text(energies, f))


# ---------------------------------------------------------------------------
# natural transition orbitals (ntos.rs:38-85)
# ---------------------------------------------------------------------------
def natural_transition_orbitals(tdm, orbs, lumo):
    """SVD T = U diag(sigma) V^T of the transition density [n_occ, n_virt]:
    hole NTOs C_occ U (columns reversed: the dominant pair next to the gap),
    particle NTOs C_virt V, weights lambda = sigma^2 as (reversed, forward).
    The longer orbital space is cropped to the number of singular values
    (CropCoeffs, ntos.rs:93-109).  Returns (lambda [2 min(o, v)], NTO
    coefficients [n_ao, 2 min(o, v)])."""
    tdm = np.asarray(tdm)
    n_occ, n_virt = tdm.shape
    u, sigma, vt = np.linalg.svd(tdm, full_matrices=True)
    u = u[:, ::-1]
    lam = sigma * sigma
    lam = np.concatenate([lam[::-1], lam])
    ntos = np.zeros_like(orbs)
    ntos[:, :n_occ] = orbs[:, :lumo] @ u
    ntos[:, n_occ:] = orbs[:, lumo:] @ vt.T
    if n_occ > n_virt:
        ntos = ntos[:, n_occ - n_virt:]
    elif n_occ < n_virt:
        ntos = ntos[:, :2 * n_occ]
    return lam, ntos


# ---------------------------------------------------------------------------
# Molden export (io/molden.rs, io/basis_set.rs)
# ---------------------------------------------------------------------------
def _sto3g():
    with open(os.path.join(_PARAMS, "sto3g.json")) as fh:
        return json.load(fh)["elements"]


def _ao_order(z):
    """Row permutation to Molden's Cartesian order: per atom s, then the p
    shell (py, pz, px) -> (px, py, pz) (molden.rs:149-172)."""
    idx, mu = [], 0
    for za in z:
        if int(za) == 1:
            idx.append(mu)
            mu += 1
        else:
            idx += [mu, mu + 3, mu + 1, mu + 2]
            mu += 4
    return np.array(idx)


def molden_text(z, xyz, orbs, orbe, f, title=None):
    """[Molden Format] of MO coefficients orbs [n_ao, n_mo] (columns: occupied
    (f > 0) first, then virtual), orbital energies and occupations."""
    z = np.asarray(z)
    if title is None:
        title = "file created at: " + datetime.datetime.now(datetime.timezone.utc).strftime(
            "%a, %d %b %Y %H:%M:%S +0000")
    txt = "[Molden Format]\n[Title]\n" + title + "\n"
    txt += "[Atoms] AU\n"
    for i, (za, r) in enumerate(zip(z, np.asarray(xyz))):
        txt += f"{SYMBOL[int(za)]} {i + 1} {int(za)} {r[0]:2.7f} {r[1]:2.7f} {r[2]:2.7f}\n"
    txt += "[GTO]\n"
    basis = _sto3g()
    for i, za in enumerate(z):
        txt += f"{i + 1} 0\n"
        shell = basis[SYMBOL[int(za)].lower()]
        for l_, coef in zip(shell["angular_momentum"], shell["coefficients"]):
            txt += f" {'spdfghi'[l_]}    {len(shell['exponents'])} {1.0:1.2f}\n"
            for e, c in zip(shell["exponents"], coef):
                txt += f"{rust_exp(e, 14):>18} {rust_exp(c, 14):>18}\n"
        txt += "\n"
    txt += "[5D]\n[MO]\n"
    ordered = np.asarray(orbs)[_ao_order(z)]
    f = list(f)
    occs = [i for i, x in enumerate(f) if x > 0.0]
    virts = [i for i, x in enumerate(f) if not x > 0.0]
    for i, _ in enumerate(occs + virts):
        txt += f" Sym=  {i + 1}a\n"
        txt += f" Ene=  {orbe[i]:2.7f}\n"
        txt += " Spin= Alpha\n"
        txt += f" Occup={f[i]:2.7f}\n"
        for j, c in enumerate(ordered[:, i]):
            txt += f"    {j + 1} {c:2.7f}\n"
    return txt


def ntos_to_molden(filename, z, xyz, tdm, orbs, lumo, title=None):
    """NTOs of one state as a Molden file: hole NTOs at dummy energies -h..-1,
    particle NTOs at +1..; the NTO weights as occupations
    (excited_states/mod.rs:84-113)."""
    lam, ntos = natural_transition_orbitals(tdm, orbs, lumo)
    half = len(lam) // 2
    dummy = -float(half) + np.arange(len(lam), dtype=np.float64)
    dummy[half:] += 1.0
    with open(filename, "w") as fh:
        fh.write(molden_text(z, xyz, ntos, dummy, lam, title))
    return lam, ntos
