"""Write the golden fixtures used by the parity tests.

Reads, as text only, the reference's own test data:

* ``tests/data/<mol>/*.dat`` -- SCC outputs and gamma matrices for h2, h2o,
  ammonia, benzene, uracil (loaded by ``src/utils/tests.rs:76-134``);
* ``not_refactored_tests/solver.rs`` -- embedded golden arrays of the legacy
  TDA tests: water full pipeline (``:33-199``), water 2x2 active space
  (``:888-1006``) and the benzene (A+B).v product (``:2265-12655``).

Nothing of the reference is imported or executed; ``array![...]`` literals are
parsed as numbers.  Output: ``tests/golden/data/<mol>/*.dat`` (verbatim data
copies) and ``tests/golden/legacy_*.json``.

Run from the repo root:  ``python tests/golden/extract_goldens.py [REF_ROOT]``.
"""
import json
import os
import re
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

MOLECULES = ["h2", "h2o", "ammonia", "benzene", "uracil"]
DAT_FILES = [
    "S", "orbs_after_scc", "orbe_after_scc", "occupation", "orbs_per_atom",
    "gamma_atomwise", "gamma_atomwise_lc", "gamma_atomwise_grad",
    "gamma_atomwise_grad_lc", "distance_matrix",
    # SCC pieces (src/scc/helpers.rs:302-518 tests): H0, reference and
    # converged density matrices, H1, long-range exchange, charges, energies
    "H0", "P0", "P_after_scc", "H1_after_scc", "g0_lr_ao", "lc_exact_exchange", "lc_exchange_energy",
    "q_after_scc", "dq_after_scc", "E_band_structure", "E_coulomb_after_scc",
]


def copy_dat():
    for mol in MOLECULES:
        src = os.path.join(REF, "tests", "data", mol)
        dst = os.path.join(HERE, "data", mol)
        os.makedirs(dst, exist_ok=True)
        for name in DAT_FILES:
            p = os.path.join(src, name + ".dat")
            if os.path.exists(p):
                shutil.copyfile(p, os.path.join(dst, name + ".dat"))
        xyz = os.path.join(src, mol + ".xyz")
        if os.path.exists(xyz):
            shutil.copyfile(xyz, os.path.join(dst, mol + ".xyz"))


def _parse_literal(text, start):
    """Parse the nested ``[...]`` literal beginning at text[start] == '['."""
    depth = 0
    i = start
    while True:
        c = text[i]
        if c == "[":
            depth += 1
        elif c == "]":
            depth -= 1
            if depth == 0:
                break
        i += 1
    body = text[start:i + 1]
    body = re.sub(r"//[^\n]*", "", body)
    body = re.sub(r"(\d)\.(?=[\s,\]])", r"\1.0", body)   # "2." -> "2.0"
    body = re.sub(r",\s*\]", "]", body)                   # trailing commas
    return json.loads(body)


def arrays_in(text, lo, hi):
    """All ``let NAME ... = array![...]`` literals whose line lies in [lo, hi]."""
    lines = text.split("\n")
    offsets = [0]
    for ln in lines:
        offsets.append(offsets[-1] + len(ln) + 1)
    out = {}
    pat = re.compile(r"let\s+(?:mut\s+)?(\w+)\s*:[^=]*=\s*array!\[")
    for m in pat.finditer(text):
        line_no = text.count("\n", 0, m.start()) + 1
        if lo <= line_no <= hi:
            out[m.group(1)] = _parse_literal(text, m.end() - 1)
    return out


def main():
    copy_dat()
    with open(os.path.join(REF, "not_refactored_tests", "solver.rs")) as fh:
        text = fh.read()
    water = arrays_in(text, 33, 199)
    water["source"] = "not_refactored_tests/solver.rs:33-199"
    water2 = arrays_in(text, 888, 1006)
    water2["source"] = "not_refactored_tests/solver.rs:888-1006"
    benz = arrays_in(text, 2265, 12655)
    benz["source"] = "not_refactored_tests/solver.rs:2265-12655"
    for name, obj in [("legacy_water_tda.json", water),
                      ("legacy_water_2x2.json", water2),
                      ("legacy_benzene_apbv.json", benz)]:
        with open(os.path.join(HERE, name), "w") as fh:
            json.dump(obj, fh)
        print(name, sorted(k for k in obj if k != "source"))


if __name__ == "__main__":
    main()
