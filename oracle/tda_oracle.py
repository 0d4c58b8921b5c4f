"""CPU oracle for the LC-TDA hot path -- TEST INFRASTRUCTURE ONLY.

This module is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it; the product path
(``tincr_amd``) never does, and never falls back to it.

It restates, in FP64 numpy, the reference's algorithm (tincr, Rust):

* ``trans_charges``        src/excited_states/transition_charges.rs:8-98
* ``orbe_differences``     src/excited_states/utils.rs:8-24
* ``SystemEngine``         src/excited_states/solvers/mod.rs:278-385
  (``compute_products`` with the ProductCache semantics of
  src/excited_states/utils.rs:29-66, per-vector exchange loop :335-358,
  ``precondition`` :373-380); ``CTPairEngine`` follows :166-274 (exchange
  always on, gamma_lr^T, rectangular).
* ``initial_subspace``     src/excited_states/solvers/utils.rs:17-26,76-80
* ``davidson``             src/excited_states/solvers/davidson.rs:53-199
* ``run_tda``              src/excited_states/tda/mod.rs:74-131 with
  ``mulliken_dipoles`` / ``oscillator_strength`` tda/moments.rs:7-24
* ``tda_full_matrix``      src/excited_states/tda/mod.rs:133-215 (dense A)
* ``ct_q_ov``              src/excited_states/tda/ct_pair.rs:78-123
* ``ct_pair_engine``       src/excited_states/tda/ct_pair.rs:12-76 +
  solvers/mod.rs:165-274
* ``frozen_charge_gradient`` the gamma terms of d omega / dR
* ``zvector``              src/gradients/helpers.rs:1130-1223 (LC) and
  :1053-1128 (no LC): conjugate gradients on A z = rhs with the operator of
  mult_av_lc :1384-1437 / mult_av_nolc :1351-1382 (``zvector_operator``)
* ``hf_derivatives``       d omega / d(gamma, gamma_lr, q) by Hellmann-Feynman
  on the dense A of tda/mod.rs:180-215 (no reference function; the gamma and
  charge terms of src/gradients/excited_state.rs:259-487 at frozen orbitals)

Pinning (each checked in tests/test_oracle.py):

* trans_charges + orbe_differences + products + dense eigh: legacy water
  full-pipeline golden (not_refactored_tests/solver.rs:33-199, 8 omegas, 1e-10),
  water 2x2 golden (:888-1006, omegas + CI coefficients 1e-14), benzene (A+B).v
  golden (:2265-12655, 1e-12), Mulliken invariants (sum_A q_oo^A_ii = 1, etc.).
* davidson: dense ``numpy.linalg.eigh`` of the same operator on the five
  fixture molecules (intent of src/excited_states/tda/system.rs:105-159) and on
  seeded diagonally dominant matrices (solvers/utils.rs:29-42, 52-74).
* run_tda post-processing: f = 2/3 w |mu|^2 against a hand value and the
  monopole zero for the water n->pi* states.
* ct_q_ov: reduces to trans_charges' q_ov when hole and particle fragment
  coincide (exactly the a = b case of transition_charges.rs:58).
* ct_pair_engine: equals the monomer engine of the supermolecule whose MO
  coefficients are block-diagonal diag(C_H, C_L) with the full overlap, the
  occupied MOs of H and the virtual MOs of L: then the Mulliken charges of
  transition_charges.rs vanish on the other fragment's atoms and reduce to the
  CT formulas, so the pair A must equal that monomer A (n_atoms_h != n_atoms_l,
  so a transposed gamma_lr block cannot pass).
* frozen_charge_gradient: central differences (step 1e-4 bohr) of the dense
  eigenvalues with gamma / gamma_lr recomputed at displaced geometries.
* hf_derivatives: central finite differences of the dense eigenvalues of A
  with one input perturbed along a random direction (tests/test_oracle.py);
  A is linear in gamma / gamma_lr and bilinear in the charges, so a dropped
  factor 2, a sign or a transposed P^B fails at 1e-6 relative.
"""
import numpy as np

HARTREE_TO_EV = 27.211386245988   # src/constants.rs:5
BOHR_TO_ANGS = 0.529177210903     # src/constants.rs:4


# ---------------------------------------------------------------------------
# transition charges                      transition_charges.rs:8-98
# ---------------------------------------------------------------------------
def ao_offsets(orbs_per_atom):
    off = np.zeros(len(orbs_per_atom) + 1, dtype=np.int64)
    off[1:] = np.cumsum(np.asarray(orbs_per_atom, dtype=np.int64))
    return off


def trans_charges(n_atoms, orbs_per_atom, orbs, s, occ_indices, virt_indices):
    """q_ov [N, o*v], q_oo [N, o*o], q_vv [N, v*v]; idx = i*v+a etc.

    q^A_pq = 1/2 sum_{mu in A} (C_mu,p (SC)_mu,q + C_mu,q (SC)_mu,p)
    (transition_charges.rs:35, :58, :70, :90).  Occupied and virtual index
    sets are used as contiguous ranges (first..last+1), as at :21-27.
    """
    orbs = np.asarray(orbs, dtype=np.float64)
    s_c = np.asarray(s, dtype=np.float64) @ orbs          # :35
    i_o, f_o = occ_indices[0], occ_indices[-1] + 1
    i_v, f_v = virt_indices[0], virt_indices[-1] + 1
    o, v = f_o - i_o, f_v - i_v
    off = ao_offsets(orbs_per_atom)
    q_ov = np.zeros((n_atoms, o * v))
    q_oo = np.zeros((n_atoms, o * o))
    q_vv = np.zeros((n_atoms, v * v))
    for n in range(n_atoms):
        mu = slice(off[n], off[n + 1])
        c_o, sc_o = orbs[mu, i_o:f_o], s_c[mu, i_o:f_o]
        c_v, sc_v = orbs[mu, i_v:f_v], s_c[mu, i_v:f_v]
        q_ov[n] = (0.5 * (c_o.T @ sc_v + sc_o.T @ c_v)).ravel()
        q_oo[n] = (0.5 * (c_o.T @ sc_o + sc_o.T @ c_o)).ravel()
        q_vv[n] = (0.5 * (c_v.T @ sc_v + sc_v.T @ c_v)).ravel()
    return q_ov, q_oo, q_vv


def trans_charges_loops(n_atoms, orbs_per_atom, orbs, s, occ_indices, virt_indices):
    """Scalar-loop form of trans_charges (same accumulation order as the
    reference's triple loops, transition_charges.rs:37-95).  Tiny inputs only."""
    orbs = np.asarray(orbs, dtype=np.float64)
    s_c = np.asarray(s, dtype=np.float64) @ orbs
    occ = list(range(occ_indices[0], occ_indices[-1] + 1))
    virt = list(range(virt_indices[0], virt_indices[-1] + 1))
    o, v = len(occ), len(virt)
    q_ov = np.zeros((n_atoms, o * v))
    q_oo = np.zeros((n_atoms, o * o))
    q_vv = np.zeros((n_atoms, v * v))
    mu = 0
    for n in range(n_atoms):
        for _ in range(orbs_per_atom[n]):
            for i, p in enumerate(occ):
                for a, q in enumerate(virt):
                    q_ov[n, i * v + a] += 0.5 * (orbs[mu, p] * s_c[mu, q] + orbs[mu, q] * s_c[mu, p])
                for j, q in enumerate(occ):
                    q_oo[n, i * o + j] += 0.5 * (orbs[mu, p] * s_c[mu, q] + orbs[mu, q] * s_c[mu, p])
            for a, p in enumerate(virt):
                for b, q in enumerate(virt):
                    q_vv[n, a * v + b] += 0.5 * (orbs[mu, p] * s_c[mu, q] + orbs[mu, q] * s_c[mu, p])
            mu += 1
    return q_ov, q_oo, q_vv


def orbe_differences(occ_quant, virt_quant):
    """d[i*v+a] = virt[a] - occ[i]   (utils.rs:8-24)."""
    occ_quant = np.asarray(occ_quant, dtype=np.float64)
    virt_quant = np.asarray(virt_quant, dtype=np.float64)
    return (virt_quant[None, :] - occ_quant[:, None]).ravel()


def ct_q_ov(orbs_h, orbs_l, s_hl, orbs_per_atom_h, orbs_per_atom_l, homo, lumo):
    """CT-pair q_ov over H u L atoms (ct_pair.rs:78-123).

    q[A in H, i, a] = 1/2 sum_{mu in A} C^H_mu,i (S_HL C^L_v)_mu,a
    q[A in L, i, a] = 1/2 sum_{mu in A} C^L_mu,a (S_HL^T C^H_o)_mu,i
    """
    occs = np.asarray(orbs_h)[:, :homo + 1]
    virts = np.asarray(orbs_l)[:, lumo:]
    s_c_l = s_hl @ virts
    s_c_h = s_hl.T @ occs
    nh, nl = len(orbs_per_atom_h), len(orbs_per_atom_l)
    o, v = occs.shape[1], virts.shape[1]
    q = np.zeros((nh + nl, o, v))
    off = ao_offsets(orbs_per_atom_h)
    for n in range(nh):
        mu = slice(off[n], off[n + 1])
        q[n] = occs[mu].T @ s_c_l[mu]
    off = ao_offsets(orbs_per_atom_l)
    for n in range(nl):
        mu = slice(off[n], off[n + 1])
        q[nh + n] = s_c_h[mu].T @ virts[mu]
    return (0.5 * q).reshape(nh + nl, o * v)


def ct_pair_engine(h, l, s_hl, gamma_hl, gamma_lr_hl, faithful=False):
    """PairChargeTransfer::prepare_ct_tda (ct_pair.rs:12-76) with its
    DavidsonEngine (solvers/mod.rs:165-274).  h / l: dicts of the monomers'
    n_atoms, orbs_per_atom, occ_indices, virt_indices, orbs, s, orbe, gamma.
    Pair gamma from the monomer blocks and gamma_hl (:20-27); omega =
    orbe_l[lumo:] - orbe_h[:homo+1] (:43-58); q_ov over H u L (:78-123); q_oo
    of H and q_vv of L (:70-72); exchange with gamma_lr_hl^T (mod.rs:219-222)."""
    nh, nl = h["n_atoms"], l["n_atoms"]
    gamma = np.zeros((nh + nl, nh + nl))
    gamma[:nh, :nh] = h["gamma"]
    gamma[nh:, nh:] = l["gamma"]
    gamma[:nh, nh:] = gamma_hl
    gamma[nh:, :nh] = np.asarray(gamma_hl).T
    homo, lumo = h["occ_indices"][-1], l["virt_indices"][0]
    orbe_h, orbe_l = np.ravel(h["orbe"]), np.ravel(l["orbe"])
    omega = orbe_differences(orbe_h[:homo + 1], orbe_l[lumo:])
    q_ov = ct_q_ov(h["orbs"], l["orbs"], s_hl, h["orbs_per_atom"], l["orbs_per_atom"], homo, lumo)
    q_oo = trans_charges(nh, h["orbs_per_atom"], h["orbs"], h["s"], h["occ_indices"], h["virt_indices"])[1]
    q_vv = trans_charges(nl, l["orbs_per_atom"], l["orbs"], l["s"], l["occ_indices"], l["virt_indices"])[2]
    o, v = homo + 1, len(orbe_l) - lumo
    return SystemEngine(q_ov, q_oo, q_vv, gamma, gamma_lr_hl, omega, o, v, faithful=faithful,
                        gamma_lr_transposed=True)


# ---------------------------------------------------------------------------
# Davidson engines                         solvers/mod.rs:20-385
# ---------------------------------------------------------------------------
class SystemEngine:
    """DavidsonEngine for System/Monomer (solvers/mod.rs:278-385).

    ``faithful=True`` reproduces the reference's costs: gamma_lr.q_oo is
    recomputed on every call (:328-331) and the exchange runs one vector at a
    time (:335-358).  ``faithful=False`` batches the exchange over vectors
    (same arithmetic, different loop order) -- the "tuned CPU" variant.
    """

    def __init__(self, q_ov, q_oo, q_vv, gamma, gamma_lr, omega, n_occ, n_virt,
                 faithful=True, gamma_lr_transposed=False):
        self.q_ov = np.ascontiguousarray(q_ov, dtype=np.float64)
        self.q_oo = None if q_oo is None else np.ascontiguousarray(q_oo, dtype=np.float64)
        self.q_vv = None if q_vv is None else np.ascontiguousarray(q_vv, dtype=np.float64)
        self.gamma = np.asarray(gamma, dtype=np.float64)
        self.gamma_lr = None if gamma_lr is None else np.asarray(gamma_lr, dtype=np.float64)
        self.omega = np.asarray(omega, dtype=np.float64)
        self.n_occ, self.n_virt = n_occ, n_virt
        self.faithful = faithful
        # CT pairs contract gamma_lr^T (solvers/mod.rs:219-222)
        self.gamma_lr_t = gamma_lr_transposed
        self.cache = None          # ProductCache "TDA" entry, [ov, n]
        self.n_products = 0        # single-vector products computed

    def reset_cache(self):
        self.cache = None

    def _gamma_oo(self):
        glr = self.gamma_lr.T if self.gamma_lr_t else self.gamma_lr
        n_b = glr.shape[0]
        return (glr @ self.q_oo).reshape(n_b * self.n_occ, self.n_occ)

    def apply(self, x):
        """A x for the columns of x [ov, k] (no cache semantics)."""
        x = np.asarray(x, dtype=np.float64)
        o, v = self.n_occ, self.n_virt
        fock = self.omega[:, None] * x                                    # :307-308
        two_el = 2.0 * self.q_ov.T @ (self.gamma @ (self.q_ov @ x))       # :311
        if self.gamma_lr is not None:                                     # :314
            gamma_oo = self._gamma_oo()                                   # :328-331
            n_b = gamma_oo.shape[0] // o
            q_vv_r = self.q_vv.reshape(n_b * v, v)
            k_x = np.zeros_like(two_el)
            if self.faithful:
                for c in range(x.shape[1]):                               # :335-358
                    # the column as a contiguous o x v block (the reference's
                    # to_owned().into_shape), so the product runs on BLAS
                    xi = np.ascontiguousarray(x[:, c]).reshape(o, v)
                    t = (xi @ q_vv_r.T).reshape(o, n_b, v).transpose(1, 0, 2)
                    t = np.ascontiguousarray(t).reshape(n_b * o, v)
                    k_x[:, c] = (gamma_oo.T @ t).reshape(o * v)
            else:
                k = x.shape[1]
                xs = x.T.reshape(k * o, v)
                t = (xs @ q_vv_r.T).reshape(k, o, n_b, v).transpose(2, 1, 0, 3)
                t = np.ascontiguousarray(t).reshape(n_b * o, k * v)
                kx = (gamma_oo.T @ t).reshape(o, k, v).transpose(1, 0, 2)
                k_x[:] = kx.reshape(k, o * v).T
            two_el = two_el - k_x                                         # :360
        self.n_products += x.shape[1]
        return fock + two_el

    def compute_products(self, x):
        """ProductCache semantics (solvers/mod.rs:290-302, :365-366)."""
        n_prod = x.shape[1]
        n_old = 0 if self.cache is None else self.cache.shape[1]
        if n_prod <= n_old:
            self.cache = None
            comp = x
        else:
            comp = x[:, n_old:]
        new = self.apply(comp)
        self.cache = new if self.cache is None else np.concatenate([self.cache, new], axis=1)
        return self.cache.copy()

    def precondition(self, r_k, w_k):
        denom = w_k - self.omega                                          # :375-376
        denom = np.where(np.abs(denom) < 0.0001, 1.0, denom)              # :378
        return r_k / denom

    def get_size(self):
        return self.omega.shape[0]

    def dense(self):
        """tda_full_matrix operator (tda/mod.rs:180-215)."""
        o, v = self.n_occ, self.n_virt
        a = np.diag(self.omega) + 2.0 * self.q_ov.T @ (self.gamma @ self.q_ov)
        if self.gamma_lr is not None:
            glr = self.gamma_lr.T if self.gamma_lr_t else self.gamma_lr
            # K[(i,a),(j,b)] = sum_AB q_oo^A[i,j] glr[A,B] q_vv^B[a,b]  (tda/mod.rs:192-201)
            n_b = glr.shape[0]
            k = (self.q_oo.T @ (glr.T @ self.q_vv.reshape(n_b, v * v)) if self.gamma_lr_t
                 else self.q_oo.T @ (glr @ self.q_vv))
            k = k.reshape(o, o, v, v).transpose(0, 2, 1, 3).reshape(o * v, o * v)
            a = a - k
        return a


class DenseEngine:
    """DavidsonEngine for a dense matrix (solvers/mod.rs:36-51)."""

    def __init__(self, a):
        self.a = np.asarray(a, dtype=np.float64)

    def compute_products(self, x):
        return self.a @ x

    def precondition(self, r_k, w_k):
        return r_k / (w_k - np.diag(self.a))

    def get_size(self):
        return self.a.shape[0]


# ---------------------------------------------------------------------------
# Davidson driver                          solvers/davidson.rs:53-199
# ---------------------------------------------------------------------------
class DavidsonError(RuntimeError):
    """davidson.rs:28-37: 'Davidson Algorithm did not converge!'"""


def argsort(v):
    # solvers/utils.rs:76-80 uses sort_unstable_by (tie order unspecified);
    # a stable sort is one admissible order.
    return np.argsort(np.asarray(v), kind="stable")


def initial_subspace(diag, dim):
    """Unit vectors at the dim smallest diagonal entries (utils.rs:17-26)."""
    order = argsort(diag)
    mtx = np.zeros((len(diag), dim))
    for idx, i in enumerate(order):
        if idx < dim:
            mtx[i, idx] = 1.0
    return mtx


def davidson(engine, guess, n_roots, tolerance, max_iter, trace=None):
    """Block Davidson exactly as davidson.rs:53-185.

    Returns (eigenvalues [n_roots], eigenvectors [dim, n_roots], n_iter);
    raises DavidsonError after max_iter iterations without convergence.
    ``trace`` (a list) receives one dict per iteration.
    """
    dim = engine.get_size()
    guess = np.array(guess, dtype=np.float64)
    dim_sub_origin = guess.shape[1]
    dim_sub = dim_sub_origin
    max_space = 15 * n_roots                                               # :74
    for i in range(max_iter):                                              # :83
        ax = engine.compute_products(guess)                                # :86
        a_proj = guess.T @ ax                                              # :89
        u, v = np.linalg.eigh(a_proj, UPLO="U")                            # :93
        u = u[:n_roots]                                                    # :96
        v = v[:, :n_roots]                                                 # :97
        ritz = guess @ v                                                   # :101
        rk = ax @ v - ritz * u[None, :]                                    # :104
        errors = np.linalg.norm(rk, axis=0)                                # :107
        roots_cvd = int(np.sum(errors < tolerance))                        # :115-117
        roots_lft = n_roots - roots_cvd
        if trace is not None:
            trace.append(dict(iter=i, roots_cvd=roots_cvd, dim_sub=dim_sub,
                              n_basis=guess.shape[1], errors=errors.copy(), u=u.copy()))
        if roots_lft == 0 and i > 0:                                       # :121
            return u.copy(), np.ascontiguousarray(ritz[:, :n_roots]), i + 1
        if dim_sub + roots_lft <= max_space:                               # :146
            add_space = np.zeros((dim, roots_lft))
            cols = [idx for idx, x in enumerate(errors) if x > tolerance]  # :150-153
            for c, idx in enumerate(cols):
                add_space[:, c] = engine.precondition(rk[:, idx], u[idx])  # :156
            dim_sub += roots_lft                                           # :159
            for c in range(add_space.shape[1]):                            # :162-168
                vec = add_space[:, c]
                orth_v = vec - guess @ (guess.T @ vec)
                norm = np.linalg.norm(orth_v)
                if norm > 1.0e-8:
                    guess = np.concatenate([guess, (orth_v / norm)[:, None]], axis=1)
        else:                                                              # :172-176
            dim_sub = dim_sub_origin
            guess = ritz[:, :dim_sub].copy()
    raise DavidsonError("Davidson Algorithm did not converge!")


# ---------------------------------------------------------------------------
# post-processing                          tda/mod.rs:74-131, moments.rs:7-24
# ---------------------------------------------------------------------------
def mulliken_dipoles(q_trans, xyz):
    return np.asarray(xyz).T @ q_trans                                     # moments.rs:13


def oscillator_strength(exc_energies, tr_dipoles):
    return 2.0 / 3.0 * exc_energies * np.sum(tr_dipoles * tr_dipoles, axis=0)  # :23


def run_tda(engine, xyz, n_roots, max_iter, tolerance, trace=None):
    """System::run_tda (tda/mod.rs:74-131) minus printing/npy output."""
    engine.reset_cache()                                                   # :76
    guess = initial_subspace(engine.omega, n_roots)                        # :82
    w, x, n_iter = davidson(engine, guess, n_roots, tolerance, max_iter, trace)
    q_trans = engine.q_ov @ x                                              # :91
    mu = mulliken_dipoles(q_trans, xyz)                                    # :94
    f = oscillator_strength(w, mu)                                         # :97
    return dict(omega=w, X=x, q_trans=q_trans, tr_dipoles=mu, f=f, n_iter=n_iter)


def tda_full_matrix(engine, xyz):
    """Dense closed form (tda/mod.rs:133-178)."""
    w, x = np.linalg.eigh(engine.dense(), UPLO="U")
    q_trans = engine.q_ov @ x
    mu = mulliken_dipoles(q_trans, xyz)
    return dict(omega=w, X=x, q_trans=q_trans, tr_dipoles=mu, f=oscillator_strength(w, mu))


def hf_derivatives(engine, x):
    """Hellmann-Feynman derivatives of w = x^T A x (x normalised, A of
    SystemEngine.dense, tda/mod.rs:180-215) w.r.t. the operator's inputs,
    entries treated as independent.  Returns dict d_gamma [N, N],
    d_gamma_lr [N, N], d_qov [N, ov], d_qoo [N, o*o], d_qvv [N, v*v]."""
    o, v = engine.n_occ, engine.n_virt
    q_ov, gam = engine.q_ov, engine.gamma
    u = q_ov @ x
    out = dict(d_gamma=2.0 * np.outer(u, u), d_qov=4.0 * np.outer(gam @ u, x))
    if engine.gamma_lr is not None:
        xm = x.reshape(o, v)
        n_b = engine.q_vv.shape[0]
        qvv = engine.q_vv.reshape(n_b, v, v)
        p = np.einsum("ia,bab,jb->bij".replace("bab", "Bab").replace("->bij", "->Bij"), xm, qvv, xm)
        p = p.reshape(n_b, o * o)                     # P^B_ij = (x q_vv^B x^T)_ij
        out["d_gamma_lr"] = -engine.q_oo @ p.T        # [A, B]
        out["d_qoo"] = -engine.gamma_lr @ p           # [A, ij]
        g = (engine.gamma_lr @ engine.q_oo).reshape(-1, o, o)   # G^B = sum_A gamma_lr[B, A] q_oo^A
        out["d_qvv"] = -np.einsum("ia,Bij,jb->Bab", xm, g, xm).reshape(-1, v * v)
    return out


def zvector_operator(engine):
    """x -> A x of the Z-vector equation: mult_av_lc (helpers.rs:1384-1437,
    de + 2 J - K, the forward product) with LC; mult_av_nolc (:1351-1382,
    de + 4 J, Coulomb factor 4 at :1377) without."""
    if engine.gamma_lr is not None:
        return engine.apply

    def op(x):
        x = np.asarray(x, dtype=np.float64)
        return engine.apply(x) + 2.0 * engine.q_ov.T @ (engine.gamma @ (engine.q_ov @ x))
    return op


def zvector(engine, rhs, max_iter=10000, conv=1.0e-19):
    """tda_zvector_lc / tda_zvector_no_lc (helpers.rs:1130-1223, :1053-1128)
    for each column of rhs [ov, k]: z0 = rhs / omega (:1154-1155),
    r0 = rhs - A z0, p0 = r0; CG until r.r <= conv (:1212-1216).
    Returns z [ov, k] and the iteration count per column."""
    op = zvector_operator(engine)
    rhs = np.asarray(rhs, dtype=np.float64)
    if rhs.ndim == 1:
        rhs = rhs[:, None]
    z = np.zeros_like(rhs)
    its = []
    for c in range(rhs.shape[1]):
        b = rhs[:, c]
        zc = b / engine.omega
        r = b - op(zc[:, None])[:, 0]
        p = r.copy()
        n_it = max_iter
        for it in range(max_iter):
            ap = op(p[:, None])[:, 0]
            t1 = r @ r
            t2 = p @ ap
            zc = zc + (t1 / t2) * p
            r = r - (t1 / t2) * ap
            t2 = r @ r
            if t2 <= conv:
                n_it = it + 1
                break
            p = (t2 / t1) * p + r
        z[:, c] = zc
        its.append(n_it)
    return z, its


def frozen_charge_gradient(engine, x, dgamma, dgamma_lr):
    """d omega / dR at frozen orbitals and charges: the gamma terms of the
    LC-TDA gradient.  dgamma[_lr]: [3N, N, N] in the layout of
    gamma_gradients_atomwise (gamma_approximation.rs:256-282, row i of
    d gamma / d R_i only), hence the symmetrised derivatives."""
    d = hf_derivatives(engine, x)
    n3 = dgamma.shape[0]
    g = dgamma.reshape(n3, -1) @ (d["d_gamma"] + d["d_gamma"].T).ravel()
    if engine.gamma_lr is not None and dgamma_lr is not None:
        g = g + dgamma_lr.reshape(n3, -1) @ (d["d_gamma_lr"] + d["d_gamma_lr"].T).ravel()
    return g


def generate_diagonal_dominant(rng, dim, sparsity):
    """solvers/utils.rs:29-36 with a seeded uniform [0,1) generator."""
    diag = 10.0 * rng.random(dim)
    off = rng.random((dim, dim))
    arr = (off + off.T) * sparsity
    arr[np.diag_indices(dim)] = diag
    return arr
