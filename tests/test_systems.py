"""Host-side input generation: gamma functions and geometries (CPU)."""
import numpy as np
import pytest

from tincr_amd import systems as sy
from tests._fixtures import FIXTURE_HUBBARD, load_dat, load_xyz

MOLS = ["h2", "h2o", "ammonia", "benzene", "uracil"]


@pytest.mark.parametrize("mol", MOLS)
def test_gamma_matches_reference_fixtures(mol):
    """gamma_approximation.rs:498-660 compares at 1e-15 against these files."""
    z, xyz = load_xyz(mol)
    for r_lr, name in [(0.0, "gamma_atomwise"), (3.03, "gamma_atomwise_lc")]:
        g = sy.gamma_atomwise(z, xyz, r_lr, hubbard=FIXTURE_HUBBARD)
        assert np.abs(g - np.atleast_2d(load_dat(mol, name))).max() < 1e-15
    for r_lr, name in [(0.0, "gamma_atomwise_grad"), (3.03, "gamma_atomwise_grad_lc")]:
        g1 = sy.gamma_gradients_atomwise(z, xyz, r_lr, hubbard=FIXTURE_HUBBARD)
        assert np.abs(g1 - load_dat(mol, name).reshape(g1.shape)).max() < 1e-15


def test_gamma_gradient_is_derivative():
    z, xyz = load_xyz("h2o")
    g1 = sy.gamma_gradients_atomwise(z, xyz, 3.03)
    h = 1e-5
    for c in range(3 * len(z)):
        xp, xm = xyz.copy(), xyz.copy()
        xp[c // 3, c % 3] += h
        xm[c // 3, c % 3] -= h
        fd = (sy.gamma_atomwise(z, xp, 3.03) - sy.gamma_atomwise(z, xm, 3.03)) / (2 * h)
        # the reference stores only the row-derivative block d gamma_ij / d R_i
        a = c // 3
        assert np.abs(fd[a, :] - g1[c, a, :]).max() < 1e-8


@pytest.mark.parametrize("builder,n,sizes", [
    (sy.water_cluster, 16, (48, 96, 64, 32)),
    (sy.water_cluster, 64, (192, 384, 256, 128)),
    (sy.water_cluster, 512, (1536, 3072, 2048, 1024)),
    (sy.polyene, 200, (402, 1002, 501, 501)),
    (sy.polyene, 60, (122, 302, 151, 151)),
])
def test_config_sizes(builder, n, sizes):
    m = builder(n, seed=0)
    assert (m.n_atoms, m.n_ao, m.n_occ, m.n_virt) == sizes
    d = np.linalg.norm(m.xyz[:, None] - m.xyz[None], axis=-1) + 1e3 * np.eye(m.n_atoms)
    assert d.min() * sy.BOHR_TO_ANGS > 0.9


def test_polyene_geometry():
    m = sy.polyene(10, jitter=0.0)
    c = m.xyz[:10] * sy.BOHR_TO_ANGS
    bonds = np.linalg.norm(np.diff(c, axis=0), axis=1)
    assert np.allclose(bonds, [1.34, 1.46] * 4 + [1.34])
    h = m.xyz[10:] * sy.BOHR_TO_ANGS
    dch = np.linalg.norm(h[:, None] - c[None], axis=-1).min(1)
    assert np.allclose(dch, 1.09)


def test_synthetic_orbitals_orthonormal():
    m = sy.water_cluster(8)
    c, s, e = sy.synthetic_orbitals(m)
    assert np.abs(c.T @ s @ c - np.eye(m.n_ao)).max() < 1e-12
    assert e[m.n_occ] - e[m.n_occ - 1] >= 0.1 - 1e-12
