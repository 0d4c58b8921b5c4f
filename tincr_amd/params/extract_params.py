"""Write the DFTB parameter data the SCC input generator needs.

Reads, as text, the reference's parameter files (RON):

* ``src/param/slaterkoster/confined_pseudo_atom/<el>.ron`` -- Hubbard U,
  valence shells (principal / angular quantum numbers, occupations);
* ``src/param/slaterkoster/free_pseudo_atom/<el>.ron`` -- the valence orbital
  energies (the on-site H0 elements, src/initialization/atom.rs:62-94);
* ``src/param/slaterkoster/slako_tables/<a>_<b>.ron`` -- the distance grid
  ``d`` and the tabulated S and H0 Slater-Koster integrals ``s`` / ``h``,
  keyed by (l1, l2, i) where i is the integral index the transformation
  rules use (src/param/slako_transformations.rs:57);
* ``src/param/basis_sets/sto-3g.json`` -- the STO-3G valence shell (the
  last electron shell) of each element, which the Molden export writes as
  the [GTO] section (src/io/basis_set.rs:38-67, :87-97).

Nothing is executed; the numbers are parsed from the text.  Output:
``atoms.json``, ``slako_<a>_<b>.npz`` and ``sto3g.json`` beside this script.

    python tincr_amd/params/extract_params.py [REF_ROOT]
"""
import json
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
ELEMENTS = ["h", "c", "n", "o"]
PAIRS = ["h_h", "h_c", "h_n", "h_o", "c_c", "c_n", "c_o", "n_n", "n_o", "o_o"]


def _field(text, key):
    """Raw text of a top-level `key: value` field (value a scalar or [...])."""
    m = re.search(r"\b" + key + r"\s*:\s*", text)
    if m is None:
        return None
    i = m.end()
    if text[i] == "[":
        depth = 0
        for j in range(i, len(text)):
            depth += text[j] == "["
            depth -= text[j] == "]"
            if depth == 0:
                return text[i:j + 1]
    return re.match(r"[^,\n)]+", text[i:]).group(0).strip()


def _numbers(s):
    return [float(x) for x in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", s)]


def pseudo_atom(kind, el):
    with open(os.path.join(REF, "src/param/slaterkoster", kind, el + ".ron")) as fh:
        t = fh.read()
    out = {}
    for key in ["z", "hubbard_u", "n_elec"]:
        out[key] = float(_field(t, key))
    for key in ["energies", "angular_momenta", "valence_orbitals", "nshell", "orbital_occupation"]:
        out[key] = _numbers(_field(t, key))
    return out


def _dict_block(text, key):
    """{ (l1, l2, i): [..], ... } of a top-level field -> {(l1,l2,i): array}."""
    m = re.search(r"\n\s*" + key + r"\s*:\s*\{", text)
    i = m.end()
    depth = 1
    j = i
    while depth:
        depth += text[j] == "{"
        depth -= text[j] == "}"
        j += 1
    body = text[i:j - 1]
    out = {}
    for mm in re.finditer(r"\((\d+),\s*(\d+),\s*(\d+)\)\s*:\s*\[([^\]]*)\]", body):
        out[(int(mm.group(1)), int(mm.group(2)), int(mm.group(3)))] = np.array(_numbers(mm.group(4)))
    return out


def slako(pair):
    with open(os.path.join(REF, "src/param/slaterkoster/slako_tables", pair + ".ron")) as fh:
        t = fh.read()
    d = np.array(_numbers(_field(t, "d")))
    tabs = {"d": d, "z1": float(_field(t, "z1")), "z2": float(_field(t, "z2"))}
    for kind in ("s", "h"):
        for (l1, l2, i), v in _dict_block(t, kind).items():
            assert len(v) == len(d), (pair, kind, l1, l2, i)
            tabs[f"{kind}_{l1}_{l2}_{i}"] = v
    return tabs


def main():
    atoms = {}
    for el in ELEMENTS:
        conf = pseudo_atom("confined_pseudo_atom", el)
        free = pseudo_atom("free_pseudo_atom", el)
        atoms[el] = {"confined": conf, "free": {"energies": free["energies"],
                                                "valence_orbitals": free["valence_orbitals"]}}
    with open(os.path.join(HERE, "atoms.json"), "w") as fh:
        json.dump({"source": "src/param/slaterkoster/{confined,free}_pseudo_atom/*.ron", "atoms": atoms}, fh,
                  indent=1)
    for pair in PAIRS:
        np.savez_compressed(os.path.join(HERE, f"slako_{pair}.npz"), **slako(pair))
    with open(os.path.join(REF, "src/param/basis_sets/sto-3g.json")) as fh:
        bse = json.load(fh)
    z_of = {"h": "1", "c": "6", "n": "7", "o": "8"}
    sto = {}
    for el in ELEMENTS:
        shell = bse["elements"][z_of[el]]["electron_shells"][-1]
        sto[el] = {"angular_momentum": shell["angular_momentum"],
                   "exponents": [float(x) for x in shell["exponents"]],
                   "coefficients": [[float(x) for x in c] for c in shell["coefficients"]]}
    with open(os.path.join(HERE, "sto3g.json"), "w") as fh:
        json.dump({"source": "src/param/basis_sets/sto-3g.json (last shell per element)", "elements": sto}, fh,
                  indent=1)
    print("wrote", HERE)


if __name__ == "__main__":
    main()
