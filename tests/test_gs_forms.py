"""The in-block Gram-Schmidt of the Davidson's new vectors in coefficient form
(the recurrence that gs_gram_kernel, tincr_amd/csrc/stream.hip, runs on two
Gram matrices) against the reference's loop (davidson.rs:162-168: each new
vector minus its projection on the growing basis, kept when the remainder's
norm exceeds 1e-8).  CPU only: this pins the algebra of the coefficient form
and its fall-back rule; the GPU test test_block_gram_schmidt_forms_agree runs
the kernels both ways.
"""
import numpy as np
import pytest


def reference_loop(basis, p):
    """davidson.rs:162-168 on a basis with orthonormal columns and new vectors p (rows)."""
    guess = basis.copy()
    kept, norms = [], []
    for vec in p:
        orth_v = vec - guess @ (guess.T @ vec)
        norm = np.linalg.norm(orth_v)
        norms.append(norm)
        if norm > 1.0e-8:
            guess = np.concatenate([guess, (orth_v / norm)[:, None]], axis=1)
            kept.append(orth_v / norm)
    return np.array(kept).reshape(-1, p.shape[1]), np.array(norms)


def coefficient_form(basis, p):
    """W = P - P V V^T, Gt = [P; W] W^T, then the recurrence of gs_gram_kernel:
    u_j = sum_r T[j][r] w_r, c_s = (T_s . Gwp[:, j]) / |u_s|^2 over accepted s < j,
    T_j = e_j - sum_s c_s T_s, |u_j|^2 = T_j Gww T_j^T; ill when some
    |u_j|^2 < 1e-4 |w_j|^2."""
    nc = len(p)
    w = p - (p @ basis) @ basis.T
    gt = np.vstack([p, w]) @ w.T
    gpw, gww = gt[:nc], gt[nc:]
    t = np.zeros((nc, nc))
    nrm2 = np.zeros(nc)
    acc = np.zeros(nc, bool)
    ill = False
    for j in range(nc):
        cs = np.zeros(nc)
        for s in range(j):
            if acc[s]:
                cs[s] = (t[s, :s + 1] @ gpw[j, :s + 1]) / nrm2[s]
        t[j] = -(cs[:j] @ t[:j]) if j else 0.0
        t[j, j] = 1.0
        nrm2[j] = t[j, :j + 1] @ gww[:j + 1, :j + 1] @ t[j, :j + 1]
        acc[j] = np.sqrt(max(nrm2[j], 0.0)) > 1.0e-8
        ill = ill or not (nrm2[j] >= 1e-4 * gww[j, j])
    tq = np.array([t[j] / np.sqrt(nrm2[j]) for j in range(nc) if acc[j]]).reshape(-1, nc)
    return tq @ w, np.sqrt(np.maximum(nrm2, 0.0)), ill


def _basis(rng, dim, m):
    q, _ = np.linalg.qr(rng.standard_normal((dim, m)))
    return q


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_coefficient_form_equals_reference_loop(seed):
    rng = np.random.default_rng(seed)
    dim, m, nc = 3000, 40, 16
    v = _basis(rng, dim, m)
    # preconditioned residuals: very different scales, some correlation
    p = rng.standard_normal((nc, dim)) * rng.uniform(1e-4, 3.0, (nc, 1))
    p[3] += 0.5 * p[1]
    q_ref, n_ref = reference_loop(v, p)
    q, n, ill = coefficient_form(v, p)
    assert not ill
    assert q.shape == q_ref.shape
    assert np.abs(q - q_ref).max() < 1e-12
    assert np.abs(n / n_ref - 1).max() < 1e-12
    full = np.concatenate([v, q.T], axis=1)
    assert np.abs(full.T @ full - np.eye(full.shape[1])).max() < 1e-13


def test_dependent_vectors_raise_the_fallback_flag():
    """A duplicate and a nearly dependent vector: the Gram estimate of |u_j|
    is no longer reliable, so the flag sends the block to the loop."""
    rng = np.random.default_rng(5)
    dim = 2000
    v = _basis(rng, dim, 20)
    p = rng.standard_normal((8, dim))
    p[4] = p[2]                                            # duplicate: the loop drops it
    p[6] = 0.3 * p[0] - 2.0 * p[1] + 1e-9 * rng.standard_normal(dim)
    q_ref, n_ref = reference_loop(v, p)
    assert len(q_ref) in (6, 7)
    _, _, ill = coefficient_form(v, p)
    assert ill


def test_vectors_in_the_basis_span_are_dropped_without_fallback():
    """New vectors that lie in the old basis project to zero: both forms drop
    them (|u| <= 1e-8) and the zero remnant does not trip the flag."""
    rng = np.random.default_rng(7)
    dim = 2000
    v = _basis(rng, dim, 20)
    p = rng.standard_normal((5, dim))
    p[2] = v @ rng.standard_normal(20)                     # inside span(V)
    q_ref, n_ref = reference_loop(v, p)
    q, n, ill = coefficient_form(v, p)
    assert not ill
    assert len(q_ref) == len(q) == 4
    assert np.abs(q - q_ref).max() < 1e-12
    assert n_ref[2] < 1e-8 and n[2] < 1e-8


def _rayleigh_ritz(a, basis, n):
    """One Davidson step on a basis assumed orthonormal (davidson.rs:89-107):
    eigh of V^T A V, Ritz vectors V y, residual norms |A z - theta z|."""
    h = basis.T @ a @ basis
    theta, y = np.linalg.eigh(h)
    z = basis @ y[:, :n]
    r = a @ z - z * theta[:n]
    return theta[:n], z, np.linalg.norm(r, axis=0)


def test_fp64_handover_model_rejects_nearly_dependent_basis():
    """Model of the FP64 hand-over of the precision ladder.  A basis that a
    low-precision phase left nearly dependent (two columns equal to 1e-9)
    gives a spurious Ritz pair: theta ~ 0 below every eigenvalue of A, a Ritz
    vector of norm << 1 and a residual far below any tolerance -- it would
    pass the convergence test.  The two guards of the FP64 phase: the start
    block is orthonormalised first (the block Gram-Schmidt modelled above,
    1e-8 drop), and converged roots must have |z| = 1 to 1e-10."""
    rng = np.random.default_rng(11)
    dim, m, n = 400, 12, 3
    q, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
    lam = np.sort(rng.uniform(0.2, 1.5, dim))
    a = (q * lam) @ q.T
    basis = _basis(rng, dim, m)
    basis[:, 5] = basis[:, 2] + 1e-9 * basis[:, 7]       # nearly dependent pair
    basis[:, 5] /= np.linalg.norm(basis[:, 5])
    theta, z, res = _rayleigh_ritz(a, basis, n)
    spurious = theta < lam[0] - 1e-3
    assert spurious.any()                                  # theta ~ 0: below the spectrum
    norms = np.linalg.norm(z, axis=0)
    assert (res[spurious] < 1e-5).all()                    # "converged" by the residual test alone
    assert (np.abs(norms[spurious] - 1) > 1e-10).all()     # the norm guard catches it
    # orthonormalised start block: columns one at a time against the kept ones
    kept, _ = reference_loop(np.zeros((dim, 0)), basis.T)
    ortho = kept.T
    assert ortho.shape[1] == m - 1 or np.abs(ortho.T @ ortho - np.eye(ortho.shape[1])).max() < 1e-12
    theta2, z2, res2 = _rayleigh_ritz(a, ortho, n)
    assert (theta2 >= lam[0] - 1e-12).all()                # Cauchy interlacing: no spurious root
    assert np.abs(np.linalg.norm(z2, axis=0) - 1).max() < 1e-10


def _cgs_block(basis, p):
    """One classical pass of a block against the basis, then in-block, 1e-8
    drop (the reference loop); returns accepted vectors (columns), |p|, |u|."""
    q, n_u = reference_loop(basis, p.T)
    return q.T, np.linalg.norm(p, axis=0), n_u


def test_orthogonality_loss_grows_by_p_over_u_and_the_tracked_second_pass_stops_it():
    """Model of the Davidson expansion on a spectrum where every new
    preconditioned residual lies largely in the basis (|u| ~ |p| / 30, as on
    the SCC orbitals of (H2O)512): with one classical pass the basis loses
    orthogonality by about that factor per block until it is dependent;
    tracking eta <- (eta + eps) max |p|/|u| and taking a second pass once eta
    would exceed 1e-10 (tda_api.hip, davidson) keeps it orthonormal."""
    rng = np.random.default_rng(2)
    dim, nb, blocks = 4000, 4, 12

    def run(tracked):
        v, _ = np.linalg.qr(rng.standard_normal((dim, nb)))
        eta, worst = 1e-15, 0.0
        for _ in range(blocks):
            p = v @ rng.standard_normal((v.shape[1], nb)) * 30.0 + rng.standard_normal((dim, nb)) / np.sqrt(dim)
            q, n_p, n_u = _cgs_block(v, p)
            grow = max(a / b for a, b in zip(n_p, n_u) if b > 1e-8)
            eta_new = (eta + 2.3e-16) * grow
            if tracked and eta_new > 1e-10:
                q, _, _ = _cgs_block(v, q)          # second pass
            else:
                eta = max(eta, eta_new)
            v = np.concatenate([v, q], axis=1)
            worst = np.abs(v.T @ v - np.eye(v.shape[1])).max()
        return worst

    assert run(tracked=False) > 1e-6          # one pass per block: orthogonality lost
    assert run(tracked=True) < 1e-8          # bounded near eta_tol (the estimate is an order of magnitude)
