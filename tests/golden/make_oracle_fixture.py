"""Write oracle fixtures for the large single-system parity tests.

The oracle (oracle/tda_oracle.py, the FP64 restatement of the reference's
trans_charges + DavidsonEngine + Davidson::new + run_tda) solves a seeded
system once, here on the CPU, and the result is committed as a small JSON
fixture: the GPU test then needs only the fixture, not a CPU solve of minutes.

Stored per state: omega (Ha), the oscillator strength f, and the projections
x_k . g_j of the CI vector on four seeded Gaussian vectors g_j (X itself,
ov x n doubles, is too large to commit; the projections pin it up to the
sign of each column).  Also n_iter, the Davidson iteration count.

    python tests/golden/make_oracle_fixture.py polyene200 20 1e-5

writes tests/golden/oracle_polyene200_n20.json (C200H202, 402 atoms,
1002 AO, 501 x 501, synthetic orbitals of seed 0).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import tda_oracle as orc  # noqa: E402
from tincr_amd import systems as sy  # noqa: E402

N_PROJ = 4
PROJ_SEED = 123


def projection_vectors(ov):
    return np.random.default_rng(PROJ_SEED).standard_normal((N_PROJ, ov))


class ChunkedEngine(orc.SystemEngine):
    """The oracle engine applying A in chunks of columns (same arithmetic per
    column; bounds the batched exchange intermediate)."""
    chunk = 4

    def apply(self, x):
        x = np.asarray(x, dtype=np.float64)
        cols = [super(ChunkedEngine, self).apply(x[:, c:c + self.chunk]) for c in range(0, x.shape[1], self.chunk)]
        return np.concatenate(cols, axis=1)


def inputs(system, seed=0):
    if system.startswith("polyene"):
        mol = sy.polyene(int(system[len("polyene"):]), seed=seed)
    elif system.startswith("water"):
        mol = sy.water_cluster(int(system[len("water"):]), seed=seed)
    else:
        raise SystemExit(f"unknown system {system}")
    return mol, sy.tda_inputs(mol, seed=seed)


def main():
    system, n, tol = sys.argv[1], int(sys.argv[2]), float(sys.argv[3])
    max_iter = 450
    mol, inp = inputs(system)
    o, v = len(inp["occ_indices"]), len(inp["virt_indices"])
    t0 = time.time()
    q_ov, q_oo, q_vv = orc.trans_charges(inp["n_atoms"], inp["orbs_per_atom"], inp["orbs"], inp["s"],
                                         inp["occ_indices"], inp["virt_indices"])
    omega = orc.orbe_differences(inp["orbe"][:o], inp["orbe"][o:])
    eng = ChunkedEngine(q_ov, q_oo, q_vv, inp["gamma"], inp["gamma_lr"], omega, o, v, faithful=False)
    res = orc.run_tda(eng, inp["xyz"], n, max_iter, tol)
    g = projection_vectors(o * v)
    proj = (g @ res["X"]).T                        # [n][N_PROJ]
    out = {
        "source": ("oracle/tda_oracle.py run_tda (FP64 restatement of tda/mod.rs:74-131, davidson.rs:53-185) "
                   f"on {mol.name}: {mol.n_atoms} atoms, {mol.n_ao} AO, {o} occ x {v} virt, synthetic orbitals "
                   f"seed 0, lowest {n} states, tol {tol}, max_iter {max_iter}"),
        "system": system, "n_roots": n, "tol": tol, "max_iter": max_iter,
        "omega": res["omega"].tolist(), "f": res["f"].tolist(), "n_iter": int(res["n_iter"]),
        "n_products": int(eng.n_products), "proj_seed": PROJ_SEED, "proj": proj.tolist(),
        "cpu_seconds": round(time.time() - t0, 1),
    }
    path = os.path.join(HERE, f"oracle_{system}_n{n}.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {path}: n_iter {out['n_iter']}, {out['n_products']} products, {out['cpu_seconds']} s")


if __name__ == "__main__":
    main()
