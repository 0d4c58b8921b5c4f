"""Seeded molecular systems and TDA inputs for tests and benchmarks.

* Geometries: water clusters on a jittered cubic lattice and planar all-trans
  polyenes C_nH_{n+2} (the workloads of the excited-state benchmarks).
* Minimal valence basis bookkeeping: H 1 AO / 1 e, C 4 AO / 4 e, N 4 / 5,
  O 4 / 6 (tincr's confined pseudo-atom data), closed-shell occupation
  occ = n_elec / 2 (src/initialization/system.rs:131-140).
* Gaussian gamma functions gamma / gamma_lr (src/scc/gamma_approximation.rs:
  29-35, 52-124, 175-193) with tincr's Hubbard parameters and
  r_lr = 3.03 bohr (src/defaults.rs:39).
* Synthetic orbitals for throughput runs: Haar-random orthogonal C with S = I
  and sorted orbital energies with a HOMO-LUMO gap.
"""
import math

import numpy as np

BOHR_TO_ANGS = 0.529177210903          # src/constants.rs:4
LONG_RANGE_RADIUS = 3.03               # src/defaults.rs:39 (bohr)
# Hubbard U (Ha) of the confined pseudo atoms (src/param/slaterkoster/confined_pseudo_atom/*.ron:8)
HUBBARD_U = {1: 0.39292032, 6: 0.34937503, 7: 0.42231745, 8: 0.49118359}
VALENCE_ORBS = {1: 1, 6: 4, 7: 4, 8: 4}
VALENCE_ELEC = {1: 1, 6: 4, 7: 5, 8: 6}
SYMBOL = {1: "H", 6: "C", 7: "N", 8: "O"}


class Molecule:
    """Atomic numbers + coordinates (bohr)."""

    def __init__(self, z, xyz_bohr, name=""):
        self.z = np.asarray(z, dtype=np.int64)
        self.xyz = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        self.name = name

    @property
    def n_atoms(self):
        return len(self.z)

    @property
    def orbs_per_atom(self):
        return np.array([VALENCE_ORBS[int(a)] for a in self.z], dtype=np.int32)

    @property
    def n_ao(self):
        return int(self.orbs_per_atom.sum())

    @property
    def n_elec(self):
        return int(sum(VALENCE_ELEC[int(a)] for a in self.z))

    @property
    def n_occ(self):
        return self.n_elec // 2

    @property
    def n_virt(self):
        return self.n_ao - self.n_occ

    def to_xyz(self):
        lines = [str(self.n_atoms), self.name]
        for a, r in zip(self.z, self.xyz * BOHR_TO_ANGS):
            lines.append(f"{SYMBOL[int(a)]} {r[0]:.8f} {r[1]:.8f} {r[2]:.8f}")
        return "\n".join(lines) + "\n"


def _random_rotation(rng):
    """Uniform random rotation (Shoemake 1992, 'Uniform random rotations')."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
                  math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3),
                  math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(n, seed=0, spacing=3.1, jitter=0.3):
    """(H2O)_n: O-H 0.9572 A, HOH 104.52 deg, centres on a cubic lattice of
    side ceil(n^(1/3)) filled lexicographically (first n sites), +-jitter A
    uniform per centre coordinate, uniform random orientation."""
    rng = np.random.default_rng(seed)
    side = 1
    while side ** 3 < n:
        side += 1
    r_oh, ang = 0.9572, math.radians(104.52)
    local = np.array([[0.0, 0.0, 0.0],
                      [r_oh * math.sin(ang / 2), r_oh * math.cos(ang / 2), 0.0],
                      [-r_oh * math.sin(ang / 2), r_oh * math.cos(ang / 2), 0.0]])
    z, xyz = [], []
    count = 0
    for ix in range(side):
        for iy in range(side):
            for iz in range(side):
                if count == n:
                    break
                centre = spacing * np.array([ix, iy, iz], dtype=float)
                centre = centre + rng.uniform(-jitter, jitter, 3)
                rot = _random_rotation(rng)
                for a, p in zip((8, 1, 1), local @ rot.T + centre):
                    z.append(a)
                    xyz.append(p)
                count += 1
    return Molecule(z, np.array(xyz) / BOHR_TO_ANGS, f"(H2O)_{n} seed {seed}")


def polyene(n_c, seed=0, jitter=0.02):
    """Planar all-trans C_nH_{n+2}: alternating C-C 1.34/1.46 A starting with a
    double bond, C-H 1.09 A, 120 deg angles, CH2 termini, +-jitter A uniform
    Cartesian noise."""
    rng = np.random.default_rng(seed)
    carbons = [np.zeros(2)]
    # zig-zag chain along x: bond directions alternate +-30 deg
    for k in range(1, n_c):
        length = 1.34 if (k - 1) % 2 == 0 else 1.46
        ang = math.radians(30.0 if k % 2 == 1 else -30.0)
        carbons.append(carbons[-1] + length * np.array([math.cos(ang), math.sin(ang)]))
    carbons = np.array(carbons)
    hyd = []
    r_ch = 1.09
    for k in range(n_c):
        nbrs = [carbons[j] for j in (k - 1, k + 1) if 0 <= j < n_c]
        if len(nbrs) == 2:
            d = (carbons[k] - nbrs[0]) / np.linalg.norm(carbons[k] - nbrs[0])
            d = d + (carbons[k] - nbrs[1]) / np.linalg.norm(carbons[k] - nbrs[1])
            hyd.append(carbons[k] + r_ch * d / np.linalg.norm(d))
        else:
            # terminal CH2: two hydrogens at +-120 deg from the C-C bond
            b = (carbons[k] - nbrs[0]) / np.linalg.norm(carbons[k] - nbrs[0])
            for sgn in (1.0, -1.0):
                a = math.radians(60.0) * sgn
                rot = np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
                hyd.append(carbons[k] + r_ch * rot @ b)
    pos = np.vstack([carbons, np.array(hyd)])
    pos3 = np.zeros((len(pos), 3))
    pos3[:, :2] = pos
    pos3 += rng.uniform(-jitter, jitter, pos3.shape)
    z = [6] * n_c + [1] * len(hyd)
    return Molecule(z, pos3 / BOHR_TO_ANGS, f"C{n_c}H{n_c + 2} seed {seed}")


# ---------------------------------------------------------------------------
# gamma functions (src/scc/gamma_approximation.rs)
# ---------------------------------------------------------------------------
def _erf(x):
    from scipy.special import erf
    return erf(x)


def _sigmas(z, hubbard):
    hub = HUBBARD_U if hubbard is None else hubbard
    return np.array([1.0 / (hub[int(a)] * math.sqrt(math.pi)) for a in z])


def gamma_atomwise(z, xyz, r_lr=0.0, hubbard=None):
    """gamma_AB = erf(c_AB r)/r, c_AB = 1/sqrt(2(s_A^2 + s_B^2 + r_lr^2/2)),
    gamma_AA = 1/sqrt(pi (s_A^2 + r_lr^2/4)), s = 1/(U sqrt(pi)).
    ``hubbard`` maps atomic number -> U (default: HUBBARD_U)."""
    z = np.asarray(z)
    sig = _sigmas(z, hubbard)
    c = 1.0 / np.sqrt(2.0 * (sig[:, None] ** 2 + sig[None, :] ** 2 + 0.5 * r_lr ** 2))
    d = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=-1)
    n = len(z)
    g = np.empty((n, n))
    off = ~np.eye(n, dtype=bool)
    g[off] = _erf(c[off] * d[off]) / d[off]
    g[np.diag_indices(n)] = 1.0 / np.sqrt(math.pi * (sig ** 2 + 0.25 * r_lr ** 2))
    return g


def gamma_gradients_atomwise(z, xyz, r_lr=0.0, hubbard=None):
    """dgamma/dR as [3N, N, N] (gamma_approximation.rs:126-171, 256-282)."""
    z = np.asarray(z)
    n = len(z)
    sig = _sigmas(z, hubbard)
    c = 1.0 / np.sqrt(2.0 * (sig[:, None] ** 2 + sig[None, :] ** 2 + 0.5 * r_lr ** 2))
    diff = xyz[:, None, :] - xyz[None, :, :]
    d = np.linalg.norm(diff, axis=-1)
    g1 = np.zeros((3 * n, n, n))
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            r = d[i, j]
            cv = c[i, j]
            dv = 2.0 * cv / math.sqrt(math.pi) * math.exp(-(cv * r) ** 2) / r - math.erf(cv * r) / r ** 2
            g1[3 * i:3 * i + 3, i, j] = diff[i, j] / r * dv
    return g1


# ---------------------------------------------------------------------------
# synthetic orbitals
# ---------------------------------------------------------------------------
def synthetic_orbitals(mol, seed=0, gap=0.1):
    """Haar-random orthogonal C (S = I) and sorted orbital energies with a
    HOMO-LUMO gap >= `gap` Ha.  For throughput runs: A.x cost does not depend
    on the data."""
    rng = np.random.default_rng(seed + 1000)
    n = mol.n_ao
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    q = q * np.sign(np.diag(r))[None, :]
    o = mol.n_occ
    occ = np.sort(rng.uniform(-1.0, -0.3, o))
    virt = np.sort(rng.uniform(-0.3 + gap, 0.6, n - o))
    return q, np.eye(n), np.concatenate([occ, virt])


def tda_inputs(mol, seed=0, lc=True):
    """Dict of TDA inputs for a Molecule (synthetic orbitals, real gammas)."""
    c, s, orbe = synthetic_orbitals(mol, seed)
    o = mol.n_occ
    return dict(
        n_atoms=mol.n_atoms, orbs_per_atom=mol.orbs_per_atom, xyz=mol.xyz,
        occ_indices=list(range(o)), virt_indices=list(range(o, mol.n_ao)),
        orbs=c, s=s, orbe=orbe,
        gamma=gamma_atomwise(mol.z, mol.xyz, 0.0),
        gamma_lr=gamma_atomwise(mol.z, mol.xyz, LONG_RANGE_RADIUS) if lc else None,
    )


# ---------------------------------------------------------------------------
# charge-transfer pair inputs
# ---------------------------------------------------------------------------
def ao_centres(mol, seed=0):
    """Model AO centres: each atom's position repeated over its AOs, each
    displaced by 0.5 bohr in a seeded random direction."""
    rng = np.random.default_rng(seed + 2000)
    centres = np.repeat(mol.xyz, mol.orbs_per_atom, axis=0)
    d = rng.standard_normal(centres.shape)
    return centres + 0.5 * d / np.linalg.norm(d, axis=1)[:, None]


def overlap_block(centres_a, centres_b, beta=0.3, sigma=1.6):
    """beta * exp(-|p_mu - p_nu|^2 / (2 sigma^2)): the off-diagonal part of
    ``model_overlap`` between two sets of AO centres."""
    r2 = ((centres_a[:, None, :] - centres_b[None, :, :]) ** 2).sum(-1)
    return beta * np.exp(-r2 / (2.0 * sigma * sigma))


def model_overlap(mol, seed=0, beta=0.3, sigma=1.6, centres=None):
    """A symmetric positive definite AO overlap with unit diagonal whose
    off-diagonal elements decay with distance: S = (1 - beta) I + beta K with
    the Gaussian kernel K_{mu,nu} = exp(-|p_mu - p_nu|^2 / (2 sigma^2)) over
    the AO centres p_mu of ``ao_centres``.  Its eigenvalues are >= 1 - beta.
    A stand-in for the Slater-Koster overlap where only its structure (SPD,
    decaying inter-fragment block) matters."""
    if centres is None:
        centres = ao_centres(mol, seed)
    return (1.0 - beta) * np.eye(len(centres)) + overlap_block(centres, centres, beta, sigma)


def fragment_orbitals(s, n_occ, seed=0, gap=0.1):
    """S-orthonormal orbitals C = S^-1/2 Q (Q Haar-random) and sorted orbital
    energies with a HOMO-LUMO gap >= `gap` Ha."""
    rng = np.random.default_rng(seed + 3000)
    n = s.shape[0]
    w, u = np.linalg.eigh(s)
    s_mhalf = (u / np.sqrt(w)[None, :]) @ u.T
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    q = q * np.sign(np.diag(r))[None, :]
    occ = np.sort(rng.uniform(-1.0, -0.3, n_occ))
    virt = np.sort(rng.uniform(-0.3 + gap, 0.6, n - n_occ))
    return s_mhalf @ q, np.concatenate([occ, virt])


def ct_pair_inputs(mol, n_atoms_h, seed=0, lc=True):
    """Inputs of a charge-transfer pair H -> L cut from `mol`: H = its first
    `n_atoms_h` atoms, L = the rest (each must hold an even electron count).
    Overlap from ``model_overlap`` of the whole molecule (H-L block
    included), S-orthonormal orbitals per fragment, real gamma / gamma_lr of
    the geometry.  Returns dict(h=..., l=..., s_hl, gamma_hl, gamma_lr_hl,
    s, gamma, gamma_lr) with h / l dicts of the fragment's n_atoms,
    orbs_per_atom, xyz, occ_indices, virt_indices, orbs, s, orbe, gamma."""
    nh = n_atoms_h
    frag_h = Molecule(mol.z[:nh], mol.xyz[:nh], mol.name + "_H")
    frag_l = Molecule(mol.z[nh:], mol.xyz[nh:], mol.name + "_L")
    s = model_overlap(mol, seed)
    gamma = gamma_atomwise(mol.z, mol.xyz, 0.0)
    gamma_lr = gamma_atomwise(mol.z, mol.xyz, LONG_RANGE_RADIUS) if lc else None
    nao_h = frag_h.n_ao
    out = dict(s=s, gamma=gamma, gamma_lr=gamma_lr, s_hl=s[:nao_h, nao_h:].copy(),
               gamma_hl=gamma[:nh, nh:].copy(), gamma_lr_hl=None if gamma_lr is None else gamma_lr[:nh, nh:].copy())
    for key, frag, ao, at, sd in (("h", frag_h, slice(0, nao_h), slice(0, nh), seed),
                                  ("l", frag_l, slice(nao_h, None), slice(nh, None), seed + 1)):
        sf = s[ao, ao].copy()
        c, orbe = fragment_orbitals(sf, frag.n_occ, sd)
        out[key] = dict(n_atoms=frag.n_atoms, orbs_per_atom=frag.orbs_per_atom, xyz=frag.xyz,
                        occ_indices=list(range(frag.n_occ)), virt_indices=list(range(frag.n_occ, frag.n_ao)),
                        orbs=c, s=sf, orbe=orbe, gamma=gamma[at, at].copy())
    return out


# ---------------------------------------------------------------------------
# batched FMO workload: seeded fragments in a box, nearest pairs
# ---------------------------------------------------------------------------
FMO_KINDS = ["H2O", "(H2O)2"] + [f"C{n}H{n + 2}" for n in range(4, 41, 2)]


def _fragment(kind, seed):
    if kind == "H2O":
        return water_cluster(1, seed=seed)
    if kind == "(H2O)2":
        return water_cluster(2, seed=seed)
    return polyene(int(kind[1:kind.index("H")]), seed=seed)


def fmo_fragments(n_frag=1024, seed=0, min_dist=3.0, box=None):
    """`n_frag` fragments drawn uniformly from FMO_KINDS (H2O, (H2O)2,
    C_nH_{n+2} for n = 4, 6, ..., 40), each with a uniform random orientation,
    placed by rejection sampling in a cube so that every interatomic distance
    between fragments is >= `min_dist` A.  Returns a list of Molecules (bohr)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    kinds = [FMO_KINDS[k] for k in rng.integers(0, len(FMO_KINDS), n_frag)]
    if box is None:
        n_atoms = sum(_fragment(k, 0).n_atoms for k in set(kinds)) / len(set(kinds)) * n_frag
        box = (100.0 * n_atoms) ** (1.0 / 3.0)
    tree, recent, frags = None, [], []
    for i, kind in enumerate(kinds):
        mol = _fragment(kind, seed * 100003 + i)
        local = mol.xyz * BOHR_TO_ANGS
        local = local - local.mean(axis=0)
        r = np.vstack(recent) if recent else None
        for _ in range(100000):
            pos = local @ _random_rotation(rng).T + rng.uniform(0.0, box, 3)
            if r is not None:
                if (((pos[:, None, :] - r[None, :, :]) ** 2).sum(-1) < min_dist ** 2).any():
                    continue
            if tree is not None and np.isfinite(tree.query(pos, distance_upper_bound=min_dist)[0]).any():
                continue
            break
        else:
            raise RuntimeError("fmo_fragments: box too crowded")
        recent.append(pos)
        frags.append(Molecule(mol.z, pos / BOHR_TO_ANGS, f"{kind} #{i}"))
        if len(recent) == 32:
            tree = cKDTree(np.vstack([f.xyz for f in frags]) * BOHR_TO_ANGS)
            recent = []
    return frags


def fmo_pairs(frags, n_pairs=2048, cutoff=8.0):
    """The `n_pairs` fragment pairs (I < J) with the smallest minimum
    interatomic distance (A), nearest first (ties by index): list of
    (I, J, distance).  The search radius starts at `cutoff` A and grows until
    it holds enough pairs."""
    from scipy.spatial import cKDTree
    pts = np.vstack([m.xyz for m in frags]) * BOHR_TO_ANGS
    owner = np.concatenate([np.full(m.n_atoms, i) for i, m in enumerate(frags)])
    tree = cKDTree(pts)
    while True:
        ab = tree.query_pairs(cutoff, output_type="ndarray")
        fi, fj = owner[ab[:, 0]], owner[ab[:, 1]]
        if len(np.unique(np.minimum(fi, fj)[fi != fj] * len(frags) + np.maximum(fi, fj)[fi != fj])) >= n_pairs:
            break
        cutoff *= 1.5
    keep = fi != fj
    ab, fi, fj = ab[keep], fi[keep], fj[keep]
    d = np.linalg.norm(pts[ab[:, 0]] - pts[ab[:, 1]], axis=1)
    lo, hi = np.minimum(fi, fj), np.maximum(fi, fj)
    key = lo * len(frags) + hi
    order = np.lexsort((d, key))
    key, d = key[order], d[order]
    first = np.r_[True, key[1:] != key[:-1]]
    key, d = key[first], d[first]
    rank = np.lexsort((key, d))
    n = len(frags)
    return [(int(key[r] // n), int(key[r] % n), float(d[r])) for r in rank[:n_pairs]]


def gamma_block(z_a, xyz_a, z_b, xyz_b, r_lr=0.0, hubbard=None):
    """gamma between two disjoint atom sets [N_a, N_b] (erf(c r)/r, r > 0)."""
    sa, sb = _sigmas(np.asarray(z_a), hubbard), _sigmas(np.asarray(z_b), hubbard)
    c = 1.0 / np.sqrt(2.0 * (sa[:, None] ** 2 + sb[None, :] ** 2 + 0.5 * r_lr ** 2))
    d = np.linalg.norm(xyz_a[:, None, :] - xyz_b[None, :, :], axis=-1)
    return _erf(c * d) / d


def fmo_inputs(frags, pairs, seed=0, lc=True):
    """Monomer inputs (model overlap, S-orthonormal seeded orbitals, real
    gamma / gamma_lr) for every fragment and CT-pair inputs (hole on I,
    particle on J) for every pair.  Returns (monomers, ct_pairs): monomer
    dicts with tda_inputs' keys plus ``centres``; pair dicts with h, l (the
    monomer dicts), s_hl, gamma_hl, gamma_lr_hl."""
    mons = []
    for i, mol in enumerate(frags):
        centres = ao_centres(mol, seed * 100003 + i)
        s = model_overlap(mol, centres=centres)
        c, orbe = fragment_orbitals(s, mol.n_occ, seed * 100003 + i)
        mons.append(dict(n_atoms=mol.n_atoms, orbs_per_atom=mol.orbs_per_atom, xyz=mol.xyz, z=mol.z,
                         occ_indices=list(range(mol.n_occ)), virt_indices=list(range(mol.n_occ, mol.n_ao)),
                         orbs=c, s=s, orbe=orbe, centres=centres,
                         gamma=gamma_atomwise(mol.z, mol.xyz, 0.0),
                         gamma_lr=gamma_atomwise(mol.z, mol.xyz, LONG_RANGE_RADIUS) if lc else None))
    cts = []
    for i, j, _ in pairs:
        h, l = mons[i], mons[j]
        cts.append(dict(h=h, l=l, s_hl=overlap_block(h["centres"], l["centres"]),
                        gamma_hl=gamma_block(h["z"], h["xyz"], l["z"], l["xyz"], 0.0),
                        gamma_lr_hl=gamma_block(h["z"], h["xyz"], l["z"], l["xyz"], LONG_RANGE_RADIUS)
                        if lc else None))
    return mons, cts
