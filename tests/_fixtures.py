"""Loaders for the committed golden fixtures (tests/golden)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOHR_TO_ANGS = 0.529177210903   # src/constants.rs:4
# The tests/data fixtures were generated with the older factor used by the
# legacy tests (not_refactored_tests/solver.rs:34) and these Hubbard values;
# with them gamma_atomwise{,_lc,_grad,_grad_lc}.dat reproduce to 1e-16.
FIXTURE_BOHR = 0.529177249
FIXTURE_HUBBARD = {1: 0.47201583989641355, 6: 0.36742624125200113, 7: 0.5299999999999999,
                   8: 0.4467609798860577}
MOLECULES = ["h2o", "ammonia", "benzene", "uracil"]
ELEMENT_Z = {"H": 1, "C": 6, "N": 7, "O": 8}


def load_dat(mol, name):
    return np.loadtxt(os.path.join(GOLDEN, "data", mol, name + ".dat"), comments="%")


def load_xyz(mol):
    """Coordinates in bohr (src/io/coordinates.rs:8-20) and atomic numbers."""
    with open(os.path.join(GOLDEN, "data", mol, mol + ".xyz")) as fh:
        lines = fh.read().strip().split("\n")
    n = int(lines[0])
    z, xyz = [], []
    for ln in lines[2:2 + n]:
        p = ln.split()
        z.append(ELEMENT_Z[p[0]])
        xyz.append([float(t) for t in p[1:4]])
    return np.array(z), np.array(xyz) / FIXTURE_BOHR


def molecule(mol):
    occ = load_dat(mol, "occupation")
    opa = load_dat(mol, "orbs_per_atom").astype(np.int64).ravel()
    z, xyz = load_xyz(mol)
    occ_idx = [i for i, f in enumerate(occ) if f > 1.0]
    virt_idx = [i for i, f in enumerate(occ) if f <= 1.0]
    return dict(
        name=mol, z=z, xyz=xyz, orbs_per_atom=opa, n_atoms=len(opa),
        orbs=load_dat(mol, "orbs_after_scc"), s=load_dat(mol, "S"),
        orbe=load_dat(mol, "orbe_after_scc").ravel(),
        gamma=np.atleast_2d(load_dat(mol, "gamma_atomwise")),
        gamma_lr=np.atleast_2d(load_dat(mol, "gamma_atomwise_lc")),
        occ_indices=occ_idx, virt_indices=virt_idx,
    )


def legacy(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        d = json.load(fh)
    return {k: (np.array(v, dtype=np.float64) if k != "source" else v) for k, v in d.items()}
