"""CPU oracle for the LC-TDA nuclear gradient -- TEST INFRASTRUCTURE ONLY.

This module is a checker: only ``tests/`` may import it; the product path
(``tincr_amd``) never does and never falls back to it.

FP64 numpy restatement of the reference's excited-state gradient
(tincr, Rust), with the reference's [3N, n_AO, n_AO] derivative arrays:

* ``expand_rows``            the grad_s / grad_h0 layout of
  h0_and_s_gradients (src/scc/h0_and_s.rs:196-300) from the row form that
  ``tincr_amd.scc.h0_and_s_gradients`` returns
* ``gamma_ao_gradients``     gamma_gradients_ao_wise(_from_atomwise)
  (src/scc/gamma_approximation.rs:392-458)
* ``f_v``                    src/gradients/helpers.rs:47-79
* ``f_lr``                   src/gradients/helpers.rs:130-186
* ``Hav``                    src/gradients/helpers.rs:621-843 (HplusType
  Tab, Tij, Qia_Xpy, Qia_Tab, Qia_Tij, Qai, Wij; every type carries the
  factor 2 of Hav::compute)
* ``ground_state_lr``        the LC part of ground_state_gradient that the
  excited gradient reuses (src/gradients/ground_state.rs:96-131: flr_dmd0)
* ``tda_gradient_lc``        src/gradients/excited_state.rs:259-487 with
  the Z-vector of tda_oracle.zvector (tda_zvector_lc,
  src/gradients/helpers.rs:1130-1223)

Pinning (tests/test_gradient_oracle.py):

* ``h0_and_s_gradients`` (host input): central differences of
  ``tincr_amd.scc.h0_and_s`` to 1e-9;
* ``gamma_ao_gradients`` and ``f_v`` / ``f_lr``: f_v(v) : W and f_lr(v) : W
  are the nuclear derivatives of the Coulomb and long-range-exchange Fock
  contractions at frozen v and W (helpers.rs:47 is d/dR of
  1/4 sum (S o vp)... written out), checked against central differences of
  those contractions computed from S and gamma at displaced geometries;
* ``tda_gradient_lc`` as a whole: central differences (step 1e-4 bohr) of the
  lowest excitation energy of SCC + dense TDA at displaced geometries, the
  reference's own check (src/gradients/numerical.rs:67-73, tolerance 1e-6).
  The Hav terms, the W matrix and the Z-vector enter only through this
  check: a dropped factor, a sign or a transposed charge block in any of
  them moves the gradient by far more than 1e-6 Ha/bohr on the test
  molecules.
"""
import numpy as np

from . import tda_oracle as to


# ---------------------------------------------------------------------------
# derivative arrays in the reference's layout
# ---------------------------------------------------------------------------
def expand_rows(rows, atom, n_atoms):
    """[3, n, n] row form -> the reference's [3N, n, n]: entry (3A+k, mu, nu)
    is rows[k, mu, nu] for A(mu) == A and rows[k, nu, mu] for A(nu) == A
    (h0_and_s.rs:279-290 assigns both (mu, nu) and (nu, mu) under the row
    atom's coordinates)."""
    atom = np.asarray(atom)
    n = len(atom)
    out = np.zeros((3 * n_atoms, n, n))
    for a in range(n_atoms):
        m = atom == a
        for k in range(3):
            out[3 * a + k][m, :] = rows[k][m, :]
            out[3 * a + k][:, m] = rows[k][m, :].T
    return out


def gamma_ao_gradients(g1, atom):
    """g1_ao[3A+k, mu, nu] = g1[3A+k, A, B(nu)] for A(mu) == A != B(nu), and
    its transpose (gamma_approximation.rs:425-458)."""
    atom = np.asarray(atom)
    n3 = g1.shape[0]
    n = len(atom)
    out = np.zeros((n3, n, n))
    for a in range(n3 // 3):
        m = atom == a
        for k in range(3):
            row = g1[3 * a + k, a, atom]            # [n]: g1[., A, B(nu)]
            row = np.where(atom == a, 0.0, row)
            out[3 * a + k][m, :] = row[None, :]
            out[3 * a + k][:, m] = row[:, None]
    return out


# ---------------------------------------------------------------------------
# f_v / f_lr                                          helpers.rs:47-79, 130-186
# ---------------------------------------------------------------------------
def f_v(v, s, grad_s, g0_ao, g1_ao):
    """[3N, n, n]: 1/4 [dS (gsv_a + gsv_b) + S (dgsv + gdsv)_a + (.)_b] with
    vp = v + v^T, sv = colsum(S o vp), gsv = g0 sv, gdsv = g0 colsum(dS o vp),
    dgsv = dg sv (helpers.rs:56-76)."""
    vp = v + v.T
    sv = np.sum(s * vp, axis=0)
    gsv = g0_ao @ sv
    dsv = np.sum(grad_s * vp[None], axis=1)           # [3N, n]: sum_a dS[a, b] vp[a, b]
    gdsv = dsv @ g0_ao.T
    dgsv = g1_ao @ sv                                  # [3N, n]
    t = dgsv + gdsv
    out = grad_s * (gsv[None, :, None] + gsv[None, None, :])
    out += s[None] * (t[:, :, None] + t[:, None, :])
    return 0.25 * out


def f_lr(v, s, grad_s, g0_lr_ao, g1_lr_ao):
    """[3N, n, n]: the twelve terms of helpers.rs:144-182, times 1/4."""
    sv = s @ v
    v_t = v.T
    sv_t = s @ v_t
    gv = g0_lr_ao * v
    t_sv = sv.T
    svg_t = (sv * g0_lr_ao).T
    sgv_t = (s @ gv).T
    d_s, d_g = grad_s, g1_lr_ao
    d_sv_t = d_s @ v_t
    d_sv = d_s @ v
    d_gv = d_g * v[None]
    f = g0_lr_ao[None] * (d_s @ t_sv)                               # 1
    f += (d_sv_t * g0_lr_ao[None]) @ s                               # 2
    f += d_s @ svg_t                                                 # 3
    f += d_s @ sgv_t                                                 # 4
    f += g0_lr_ao[None] * (s @ np.transpose(d_sv, (0, 2, 1)))        # 5
    f += (sv_t * g0_lr_ao)[None] @ np.transpose(d_s, (0, 2, 1))      # 6
    f += s @ np.transpose(d_sv * g0_lr_ao[None], (0, 2, 1))          # 7
    f += s @ np.transpose(d_s @ gv, (0, 2, 1))                       # 8
    f += d_g * (s @ t_sv)[None]                                      # 9
    f += ((sv_t)[None] * d_g) @ s                                    # 10
    f += s @ np.transpose(sv[None] * d_g, (0, 2, 1))                 # 11
    f += s @ np.transpose(s @ d_gv, (0, 2, 1))                       # 12
    return 0.25 * f


def contract(f3, m):
    """f3 [3N, n, n] : m [n, n] -> [3N] (the into_shape + dot of
    excited_state.rs:452-485)."""
    return f3.reshape(f3.shape[0], -1) @ np.asarray(m).reshape(-1)


# ---------------------------------------------------------------------------
# Hav                                                   helpers.rs:621-843
# ---------------------------------------------------------------------------
class Hav:
    """Charge-space contractions of the gradient.  q_ov [N, o, v],
    q_oo [N, o, o], q_vv [N, v, v]; g0 = gamma, g0_lr = gamma_lr (atomwise).
    Each method returns 2 x (term 1 - term 2) as Hav::compute does."""

    def __init__(self, q_ov, q_oo, q_vv, g0, g0_lr):
        self.ov, self.oo, self.vv = q_ov, q_oo, q_vv
        self.g, self.l = g0, g0_lr

    def _lr(self, t):
        if self.l is None:                  # no long-range correction: exchange terms vanish
            return np.zeros_like(t)
        return np.einsum("AB,B...->A...", self.l, t)

    def tab(self, v):                       # [v, v] -> [o, o]      :683-704
        t1 = 2.0 * np.einsum("A,Aji->ji", self.g @ np.einsum("Bab,ab->B", self.vv, v), self.oo)
        t2 = self._lr(np.einsum("Bib,ab->Bia", self.ov, v))
        return 2.0 * (t1 - np.einsum("Aja,Aia->ji", self.ov, t2))

    def tij(self, v):                       # [o, o] -> [o, o]      :706-726
        t1 = 2.0 * np.einsum("A,Aji->ji", self.g @ np.einsum("Bkl,kl->B", self.oo, v), self.oo)
        t2 = self._lr(np.einsum("Bil,kl->Bik", self.oo, v))
        return 2.0 * (t1 - np.einsum("Ajk,Aik->ji", self.oo, t2))

    def qia_x(self, v):                     # [o, v] -> [v, v]      :728-749
        t1 = 2.0 * np.einsum("A,Aab->ab", self.g @ np.einsum("Bia,ia->B", self.ov, v), self.vv)
        t2 = self._lr(np.einsum("Bac,ic->Bai", self.vv, v))
        return 2.0 * (t1 - np.einsum("Aib,Aai->ba", self.ov, t2))

    def qia_tab(self, v):                   # [v, v] -> [o, v]      :751-772
        t1 = 2.0 * np.einsum("A,Aia->ia", self.g @ np.einsum("Bab,ab->B", self.vv, v), self.ov)
        t2 = self._lr(np.einsum("Bac,bc->Bab", self.vv, v))
        return 2.0 * (t1 - np.einsum("Aib,Aab->ia", self.ov, t2))

    def qia_tij(self, v):                   # [o, o] -> [o, v]      :774-795
        t1 = 2.0 * np.einsum("A,Aia->ia", self.g @ np.einsum("Bkl,kl->B", self.oo, v), self.ov)
        t2 = self._lr(np.einsum("Bka,jk->Baj", self.ov, v))
        return 2.0 * (t1 - np.einsum("Aij,Aaj->ia", self.oo, t2))

    def qai(self, v):                       # [o, v] -> [o, o]      :797-818 (also Wij)
        t1 = 2.0 * np.einsum("A,Aki->ki", self.g @ np.einsum("Bia,ia->B", self.ov, v), self.oo)
        t2 = self._lr(np.einsum("Bia,ja->Bij", self.ov, v))
        return 2.0 * (t1 - np.einsum("Akj,Aij->ki", self.oo, t2))

    def wij(self, v, exact=False):
        """W_ij response term (helpers.rs:797-818 via HplusType::Wij).  The
        reference reuses the Qai contraction, whose exchange part
        e1[k, i] = sum_A,j q_oo^A_kj (gamma_lr t)^A_ij enters twice; the
        derivative of the ground-state Fock matrix has (kj|ib) z_jb +
        (kb|ij) z_jb, i.e. e1 + e1^T (``exact``)."""
        if not exact:
            return self.qai(v)
        t1 = 2.0 * np.einsum("A,Aki->ki", self.g @ np.einsum("Bia,ia->B", self.ov, v), self.oo)
        t2 = self._lr(np.einsum("Bia,ja->Bij", self.ov, v))
        e1 = np.einsum("Akj,Aij->ki", self.oo, t2)
        return 2.0 * t1 - (e1 + e1.T)


def apb_operator(eng):
    """x -> (A + B) x of the ground-state orbital response: de + 4 J - K - K'
    with (K' x)_ia = sum_jb sum_AB q_ov^A_ib gamma_lr,AB q_ov^B_ja x_jb.  The
    reference's LC Z-vector (tda_zvector_lc) uses the TDA A (de + 2 J - K)
    instead; its no-LC Z-vector (factor 4, helpers.rs:1377) is this operator
    without exchange."""
    o, v = eng.n_occ, eng.n_virt
    qov = eng.q_ov.reshape(-1, o, v)

    def op(x):
        x = np.asarray(x, dtype=np.float64)
        out = eng.apply(x) + 2.0 * eng.q_ov.T @ (eng.gamma @ (eng.q_ov @ x))
        if eng.gamma_lr is not None:
            for c in range(x.shape[1]):
                xm = x[:, c].reshape(o, v)
                t = eng.gamma_lr @ np.einsum("Bja,jb->Bab", qov, xm).reshape(qov.shape[0], -1)
                out[:, c] -= np.einsum("Aib,Aab->ia", qov, t.reshape(-1, v, v)).ravel()
        return out
    return op


# ---------------------------------------------------------------------------
# assembly
# ---------------------------------------------------------------------------
def ground_state_lr(p, p0, s, grad_s, g0_lr_ao, g1_lr_ao):
    """flr_dmd0 = f_lr(P - P0) (ground_state.rs:101-116)."""
    return f_lr(p - p0, s, grad_s, g0_lr_ao, g1_lr_ao)


def gradient_inputs(z, xyz, scc_result, occ, virt):
    """Everything tda_gradient_lc / tda_gradient_nolc read from
    System::properties for a converged SCC (ground_state_gradient +
    prepare_excited_grad, ground_state.rs:12-131, excited_state.rs:11-54), in
    the reference's [3N, n, n] layout; the LC entries are None for an SCC
    without long-range correction (scc_result["gamma_lr"] None).  Host inputs
    come from tincr_amd.scc / systems."""
    from tincr_amd import scc as hscc, systems as sy
    z = np.asarray(z)
    atom = np.asarray(scc_result["atom"])
    n_atoms = len(z)
    hub = hscc.hubbard_of(z)
    ds, dh = hscc.h0_and_s_gradients(z, xyz)
    s = scc_result["s"]
    gamma, gamma_lr = scc_result["gamma"], scc_result["gamma_lr"]
    lc = gamma_lr is not None
    g1 = sy.gamma_gradients_atomwise(z, xyz, 0.0, hubbard=hub)
    g1_lr = sy.gamma_gradients_atomwise(z, xyz, sy.LONG_RANGE_RADIUS, hubbard=hub) if lc else None
    _, _, _, occ_ao = hscc.ao_layout(z)
    orbs_per_atom = np.bincount(atom, minlength=n_atoms)
    q_ov, q_oo, q_vv = to.trans_charges(n_atoms, orbs_per_atom, scc_result["orbs"], s, occ, virt)
    o, v = len(occ), len(virt)
    return dict(n_atoms=n_atoms, atom=atom, s=s, p=scc_result["p"], p0=np.diag(occ_ao), orbs=scc_result["orbs"],
                orbe=scc_result["orbe"], occ=list(occ), virt=list(virt), gamma=gamma, gamma_lr=gamma_lr,
                g0_ao=gamma[np.ix_(atom, atom)], g0_lr_ao=gamma_lr[np.ix_(atom, atom)] if lc else None,
                g1=g1, g1_lr=g1_lr, g1_ao=gamma_ao_gradients(g1, atom),
                g1_lr_ao=gamma_ao_gradients(g1_lr, atom) if lc else None, ds_rows=ds, dh_rows=dh,
                grad_s=expand_rows(ds, atom, n_atoms), grad_h0=expand_rows(dh, atom, n_atoms),
                q_ov=q_ov.reshape(n_atoms, o, v), q_oo=q_oo.reshape(n_atoms, o, o),
                q_vv=q_vv.reshape(n_atoms, v, v))


def tda_gradient(g, x_state, omega_state, exact=False, zvector_conv=1.0e-19, return_parts=False):
    """d omega / dR [3N] of one TDA state: tda_gradient_lc
    (excited_state.rs:259-487) when g carries gamma_lr, else
    tda_gradient_nolc (excited_state.rs:56-257: the same assembly with the
    Coulomb-only contractions h_a_nolr, helpers.rs:361-378 = the Hav terms
    without exchange, the factor-4 Z-vector of tda_zvector_no_lc, no f_lr).
    g: gradient_inputs(); x_state [o, v]; omega_state: its excitation energy.

    exact=False follows the reference.  exact=True changes the two LC
    response terms that central differences reject (1e-4 Ha/bohr on
    butadiene): the Z-vector solves with the ground-state orbital Hessian
    A + B (apb_operator) instead of the TDA A, and H+[Z] in W_ij carries both
    exchange patterns (Hav.wij).  Without LC both forms coincide."""
    orbe = np.asarray(g["orbe"])
    occ, virt = g["occ"], g["virt"]
    lc = g["gamma_lr"] is not None
    e_o, e_v = orbe[occ], orbe[virt]
    o, v = len(occ), len(virt)
    x = np.asarray(x_state, dtype=np.float64).reshape(o, v)
    ei, ea = np.diag(e_o), np.diag(e_v)
    # U, V, T                                                  :280-286
    u_ab = 2.0 * x.T @ x
    u_ij = 2.0 * x @ x.T
    v_ab = 2.0 * (ei @ x).T @ x
    v_ij = 2.0 * x @ ea @ x.T
    t_ab = x.T @ x
    t_ij = x @ x.T
    hav = Hav(g["q_ov"], g["q_oo"], g["q_vv"], g["gamma"], g["gamma_lr"])
    # Q                                                        :304-329
    g_ij = hav.tab(t_ab) - hav.tij(t_ij)
    q_ij = omega_state * u_ij - v_ij + g_ij
    q_ab = omega_state * u_ab + v_ab
    q_ia = x @ hav.qia_x(x).T + hav.qia_tab(t_ab) - hav.qia_tij(t_ij)
    q_ai = x.T @ hav.qai(x)
    r_ia = q_ai.T - q_ia                                       # :332
    # Z-vector                                                 :335-351
    omega_in = e_v[None, :] - e_o[:, None]
    eng = to.SystemEngine(g["q_ov"].reshape(-1, o * v), g["q_oo"].reshape(-1, o * o), g["q_vv"].reshape(-1, v * v),
                          g["gamma"], g["gamma_lr"], omega_in.ravel(), o, v)
    if exact and lc:
        saved = to.zvector_operator
        to.zvector_operator = apb_operator
        try:
            z, _ = to.zvector(eng, r_ia.reshape(-1, 1), conv=zvector_conv)
        finally:
            to.zvector_operator = saved
    else:
        z, _ = to.zvector(eng, r_ia.reshape(-1, 1), conv=zvector_conv)
    z_ia = z[:, 0].reshape(o, v)
    # W                                                        :354-392
    w_ij = q_ij + hav.wij(z_ia, exact=exact)
    w_ij[np.diag_indices(o)] /= 2.0
    w_ia = q_ai.T + ei @ z_ia
    w_ai = q_ai + (ei @ z_ia).T
    w_ab = q_ab.copy()
    w_ab[np.diag_indices(v)] /= 2.0
    w = np.block([[w_ij, w_ia], [w_ai, w_ab]])
    # AO quantities                                            :394-448
    s, grad_s = g["s"], g["grad_s"]
    diff_p = g["p"] - g["p0"]
    grad_h = g["grad_h0"] + f_v(diff_p, s, grad_s, g["g0_ao"], g["g1_ao"])
    if lc:
        grad_h = grad_h - 0.5 * ground_state_lr(g["p"], g["p0"], s, grad_s, g["g0_lr_ao"], g["g1_lr_ao"])
    orbs = np.asarray(g["orbs"])
    c_o, c_v = orbs[:, occ], orbs[:, virt]
    t_oo = c_o @ t_ij @ c_o.T
    t_vv = c_v @ t_ab @ c_v.T
    z_ao = c_o @ z_ia @ c_v.T
    c_ov = orbs[:, list(occ) + list(virt)]
    w_ao = c_ov @ np.triu(w) @ c_ov.T
    x_ao = c_o @ x @ c_v.T
    f = f_v(x_ao, s, grad_s, g["g0_ao"], g["g1_ao"])
    # assembly                                                 :450-486
    parts = dict(h=contract(grad_h, t_vv - t_oo + z_ao), s=-contract(grad_s, w_ao),
                 coul=2.0 * contract(f, x_ao), exch=np.zeros(3 * g["n_atoms"]))
    if lc:
        parts["exch"] = -contract(f_lr(x_ao.T, s, grad_s, g["g0_lr_ao"], g["g1_lr_ao"]), x_ao)
    grad = parts["h"] + parts["s"] + parts["coul"] + parts["exch"]
    if return_parts:
        return grad, dict(parts, z=z_ia, w=w, r=r_ia)
    return grad


def tda_gradient_lc(g, x_state, omega_state, **kw):
    """tda_gradient for an LC system (excited_state.rs:259-487)."""
    assert g["gamma_lr"] is not None
    return tda_gradient(g, x_state, omega_state, **kw)
