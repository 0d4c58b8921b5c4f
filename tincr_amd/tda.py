"""Host-side mirror of tincr's excited-state interface, backed by the HIP library.

Names, argument meaning and error behaviour follow the reference:

* ``trans_charges``      src/excited_states/transition_charges.rs:8-15
* ``orbe_differences``   src/excited_states/utils.rs:8-24
* ``TdaSystem``          the DavidsonEngine of System / Monomer
  (src/excited_states/solvers/mod.rs:20-34, :278-385) together with
  ``prepare_tda`` (tda/system.rs:10-52) and ``run_tda`` (tda/mod.rs:74-131);
  results land in ``properties`` under the reference's keys
  (``ci_eigenvalues``, ``ci_coefficients``, ``q_trans``, ``tr_dipoles``,
  ``oscillator_strengths``, ``q_ov``/``q_oo``/``q_vv``, ``omega``, ``homo``,
  ``lumo``).
* ``TdaCTPair``          PairChargeTransfer (tda/ct_pair.rs:12-180) with its
  DavidsonEngine (solvers/mod.rs:165-274)
* ``DavidsonError``      src/excited_states/solvers/davidson.rs:28-37 -- raised
  where the reference's callers ``unwrap()`` a non-converged Davidson.

All arithmetic runs on the GPU through libtincr_tda.so; there is no CPU path.
"""
import ctypes

import numpy as np

from . import _native as nat


class DavidsonError(RuntimeError):
    """Davidson Algorithm did not converge!"""


def _ranges(occ_indices, virt_indices):
    occ_indices = list(occ_indices)
    virt_indices = list(virt_indices)
    ob, oe = occ_indices[0], occ_indices[-1] + 1
    vb, ve = virt_indices[0], virt_indices[-1] + 1
    return ob, oe - ob, vb, ve - vb


_shared_ctx = None


def _ctx():
    global _shared_ctx
    if _shared_ctx is None:
        _shared_ctx = nat.Context(0, "fp64")
    return _shared_ctx


def trans_charges(n_atoms, orbs_per_atom, orbs, s, occ_indices, virt_indices, ctx=None):
    """(q_ov [N, o*v], q_oo [N, o*o], q_vv [N, v*v]) on the GPU.

    ``orbs_per_atom`` replaces the reference's ``atoms[].n_orbs``; occupied and
    virtual indices are used as contiguous ranges, as the reference does.
    """
    ctx = ctx or _ctx()
    ob, o, vb, v = _ranges(occ_indices, virt_indices)
    opa = np.ascontiguousarray(orbs_per_atom, dtype=np.int32)
    orbs = nat.f64(orbs)
    s = nat.f64(s)
    n_ao = orbs.shape[0]
    q_ov = np.empty((n_atoms, o * v))
    q_oo = np.empty((n_atoms, o * o))
    q_vv = np.empty((n_atoms, v * v))
    ctx.check(ctx.lib.tda_trans_charges(ctx.handle, n_atoms, n_ao, nat.iptr(opa), nat.dptr(orbs), nat.dptr(s),
                                        ob, o, vb, v, nat.dptr(q_ov), nat.dptr(q_oo), nat.dptr(q_vv)))
    return q_ov, q_oo, q_vv


def orbe_differences(occ_quant, virt_quant):
    occ = nat.f64(occ_quant)
    virt = nat.f64(virt_quant)
    out = np.empty(len(occ) * len(virt))
    lib = nat.load()
    rc = lib.tda_orbe_differences(nat.dptr(occ), len(occ), nat.dptr(virt), len(virt), nat.dptr(out))
    if rc != nat.TDA_OK:
        raise nat.TdaError(rc, "tda_orbe_differences failed")
    return out


class _TdaEngine:
    """DavidsonEngine + prepare/run shared by a System/Monomer and a CT pair.

    Subclasses load their inputs into ``self.ctx`` and set ``n_atoms`` (atoms
    of q_ov), ``n_atoms_oo`` / ``n_atoms_vv`` (atoms of q_oo / q_vv),
    ``n_occ``, ``n_virt``, ``occ_indices``, ``virt_indices`` and
    ``gammafunction_lc``."""

    def close(self):
        self.ctx.close()

    # -- prepare_tda (tda/system.rs:10-52) ---------------------------------
    def prepare_tda(self, fetch_charges=True):
        self.ctx.check(self.ctx.lib.tda_prepare(self.ctx.handle))
        N, o, v = self.n_atoms, self.n_occ, self.n_virt
        omega = np.empty(o * v)
        if fetch_charges:
            q_ov = np.empty((N, o * v))
            q_oo = np.empty((self.n_atoms_oo, o * o)) if self.gammafunction_lc else None
            q_vv = np.empty((self.n_atoms_vv, v * v)) if self.gammafunction_lc else None
            self.ctx.check(self.ctx.lib.tda_get_charges(self.ctx.handle, nat.dptr(q_ov), nat.dptr(q_oo),
                                                        nat.dptr(q_vv), nat.dptr(omega)))
            self.properties.update(q_ov=q_ov, q_oo=q_oo, q_vv=q_vv)
        else:
            self.ctx.check(self.ctx.lib.tda_get_charges(self.ctx.handle, None, None, None, nat.dptr(omega)))
        self.properties.update(omega=omega, homo=self.occ_indices[-1], lumo=self.virt_indices[0])

    def clear_charges(self):
        """Forget the prepared charges (Properties::reset, properties/mod.rs:23-63)."""
        self.ctx.check(self.ctx.lib.tda_clear_charges(self.ctx.handle))
        for key in ("q_ov", "q_oo", "q_vv", "omega"):
            self.properties.pop(key, None)

    # -- DavidsonEngine (solvers/mod.rs:20-34) --------------------------------
    def compute_products(self, x):
        """A x for every column of x [ov, k] (operator level, no cache)."""
        x = nat.f64(x)
        squeeze = x.ndim == 1
        if squeeze:
            x = x[:, None].copy()
        k = x.shape[1]
        ax = np.empty_like(x)
        self.ctx.check(self.ctx.lib.tda_compute_products(self.ctx.handle, nat.dptr(x), k, nat.dptr(ax)))
        return ax[:, 0] if squeeze else ax

    def zvector(self, rhs, max_iter=10000, conv=1.0e-19):
        """Z-vector solve A z = rhs for each column of rhs [ov, k] (or [ov]):
        tda_zvector_lc / tda_zvector_no_lc (gradients/helpers.rs:1130-1223,
        :1053-1128) on the device.  Returns z and the iterations per column."""
        b = nat.f64(rhs)
        squeeze = b.ndim == 1
        if squeeze:
            b = b[:, None].copy()
        k = b.shape[1]
        z = np.empty_like(b)
        its = np.zeros(k, dtype=np.int32)
        self.ctx.check(self.ctx.lib.tda_zvector(self.ctx.handle, nat.dptr(b), k, int(max_iter), float(conv),
                                                nat.dptr(z), its.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return (z[:, 0] if squeeze else z), its

    def precondition(self, r_k, w_k):
        r = nat.f64(r_k)
        out = np.empty_like(r)
        self.ctx.check(self.ctx.lib.tda_precondition(self.ctx.handle, nat.dptr(r), float(w_k), nat.dptr(out)))
        return out

    def get_size(self):
        return self.n_occ * self.n_virt

    # -- run_tda (tda/mod.rs:74-131) ----------------------------------------
    def run_tda(self, n_roots, max_iter, tolerance, copy_x=True, x_out=None):
        """Solve; returns (omega, X).  copy_x=False leaves the CI vectors on
        the device (X is None; tda_adjoint / adjoint() still use them).
        x_out: a C-contiguous float64 [ov, n_roots] array to receive X (e.g. a
        view of page-locked memory, which the device copies into directly)."""
        N, ov = self.n_atoms, self.n_occ * self.n_virt
        omega = np.empty(n_roots)
        if not copy_x:
            X = None
        elif x_out is not None:
            X = x_out
            if X.shape != (ov, n_roots) or X.dtype != np.float64 or not X.flags.c_contiguous:
                raise ValueError(f"x_out must be a C-contiguous float64 array of shape {(ov, n_roots)}")
        else:
            X = np.empty((ov, n_roots))
        q_trans = np.empty((N, n_roots))
        tr_dip = np.empty((3, n_roots))
        f = np.empty(n_roots)
        st = nat.TdaStats()
        rc = self.ctx.lib.tda_run(self.ctx.handle, n_roots, max_iter, float(tolerance), nat.dptr(omega),
                                  nat.dptr(X), nat.dptr(q_trans), nat.dptr(tr_dip), nat.dptr(f),
                                  nat.ctypes.byref(st))
        self.last_stats = st.as_dict()
        if rc == nat.TDA_ERR_NOT_CONVERGED:
            raise DavidsonError("Davidson Algorithm did not converge!")
        self.ctx.check(rc)
        self.properties.update(ci_eigenvalues=omega, ci_coefficients=X, q_trans=q_trans, tr_dipoles=tr_dip,
                               oscillator_strengths=f)
        return omega, X

    def tda_states(self, total_energy):
        """The TdaStates summary System::run_tda prints (tda/new_mod.rs:52-142)
        for the last run_tda with copy_x=True; total_energy: the ground-state
        total energy (Ha)."""
        from . import output
        p = self.properties
        x = p["ci_coefficients"]
        return output.tda_states_text(total_energy, p["ci_eigenvalues"],
                                      x.reshape(self.n_occ, self.n_virt, x.shape[1]),
                                      p["oscillator_strengths"], p["tr_dipoles"])

    def write_outputs(self, directory, total_energy=None):
        """The files System::run_tda writes (full_energies.npy,
        oscillator_strengths.npy; tda/mod.rs:118-119); returns the TdaStates
        text when total_energy is given."""
        from . import output
        p = self.properties
        output.write_tda_npy(directory, p["ci_eigenvalues"], p["oscillator_strengths"])
        return None if total_energy is None else self.tda_states(total_energy)

    def tdm(self, idx):
        """Transition density of state idx as [n_occ, n_virt] (properties/getter.rs:130-148)."""
        return self.properties["ci_coefficients"][:, idx].reshape(self.n_occ, self.n_virt)


class TdaSystem(_TdaEngine):
    """DavidsonEngine + prepare_tda/run_tda of one System or Monomer."""

    def __init__(self, n_atoms, orbs_per_atom, xyz, occ_indices, virt_indices, orbs, s, orbe, gamma,
                 gamma_lr=None, precision="fp64", device=0, hbm_budget_bytes=0, operator="ao"):
        self.ctx = nat.Context(device, precision, hbm_budget_bytes)
        self.ctx.set_operator(operator)
        self.n_atoms = self.n_atoms_oo = self.n_atoms_vv = n_atoms
        self.occ_indices = list(occ_indices)
        self.virt_indices = list(virt_indices)
        self.ob, self.n_occ, self.vb, self.n_virt = _ranges(occ_indices, virt_indices)
        self.opa = np.ascontiguousarray(orbs_per_atom, dtype=np.int32)
        self.xyz = nat.f64(np.asarray(xyz).reshape(n_atoms, 3))
        self._orbs = nat.f64(orbs)
        self._s = nat.f64(s)
        self._orbe = nat.f64(np.ravel(orbe))
        self._gamma = nat.f64(gamma)
        self._gamma_lr = nat.f64(gamma_lr)
        self.gammafunction_lc = gamma_lr is not None
        self.properties = {}
        self.last_stats = None
        n_ao = self._orbs.shape[0]
        self.ctx.check(self.ctx.lib.tda_set_system(
            self.ctx.handle, n_atoms, n_ao, nat.iptr(self.opa), nat.dptr(self._orbs), nat.dptr(self._s),
            nat.dptr(self._orbe), self.ob, self.n_occ, self.vb, self.n_virt, nat.dptr(self._gamma),
            nat.dptr(self._gamma_lr), nat.dptr(self.xyz)))

    # -- adjoint (Hellmann-Feynman) -----------------------------------------
    def adjoint(self, n_states, X=None, charges=False):
        """d omega_k / d(gamma, gamma_lr, q) for the first n_states CI vectors
        (X [ov, n_states], default: the last run_tda's).  Returns a dict with
        d_gamma, d_gamma_lr [n, N, N], c_qov [n, N] (d omega_k / d q_ov[A, ia]
        = c_qov[k, A] X[ia, k]) and, with charges=True, d_qoo [n, N, o*o] and
        d_qvv [n, N, v*v]."""
        N, o, v, n = self.n_atoms, self.n_occ, self.n_virt, n_states
        x = None if X is None else nat.f64(np.asarray(X)[:, :n])
        out = dict(d_gamma=np.empty((n, N, N)), c_qov=np.empty((n, N)))
        if self.gammafunction_lc:
            out["d_gamma_lr"] = np.empty((n, N, N))
            if charges:
                out["d_qoo"] = np.empty((n, N, o * o))
                out["d_qvv"] = np.empty((n, N, v * v))
        self.ctx.check(self.ctx.lib.tda_adjoint(
            self.ctx.handle, n, nat.dptr(x), nat.dptr(out["d_gamma"]), nat.dptr(out.get("d_gamma_lr")),
            nat.dptr(out["c_qov"]), nat.dptr(out.get("d_qoo")), nat.dptr(out.get("d_qvv"))))
        return out

    def nuclear_gradient(self, n_states, dgamma_dr, dgamma_lr_dr=None, X=None):
        """d omega_k / dR at frozen charges [n_states, 3N] from the gamma
        gradients [3N, N, N] in the layout of gamma_gradients_atomwise
        (gamma_approximation.rs:256-282: row i of d gamma / d R_i only)."""
        N, n = self.n_atoms, n_states
        x = None if X is None else nat.f64(np.asarray(X)[:, :n])
        g = np.empty((n, 3 * N))
        dg = nat.f64(np.asarray(dgamma_dr).reshape(3 * N, N, N))
        dglr = None if dgamma_lr_dr is None else nat.f64(np.asarray(dgamma_lr_dr).reshape(3 * N, N, N))
        self.ctx.check(self.ctx.lib.tda_adjoint_nuclear(self.ctx.handle, n, nat.dptr(x), nat.dptr(dg),
                                                        nat.dptr(dglr), nat.dptr(g)))
        return g


    # -- excited-state gradient -----------------------------------------------
    def tda_gradient(self, state, inputs, form="reference", omega=None, x=None, max_iter=10000, conv=1.0e-19,
                     return_z=False):
        """d omega / dR [3N] of state `state` (System::tda_gradient_lc,
        src/gradients/excited_state.rs:259-487, or tda_gradient_nolc :56-257
        without gamma_lr).  inputs: scc.excited_gradient_inputs() (P - P_ref,
        dS / dH0 rows, gamma gradients rows).  omega / x: the state's
        excitation energy and CI vector [n_occ * n_virt] (default: from the last
        run_tda).  form: "reference" or "exact" (the exact LC response terms,
        see tda_gradient in include/tincr_tda.h).  Returns grad, or (grad, z,
        Z-vector iterations) with return_z."""
        if omega is None:
            omega = float(self.properties["ci_eigenvalues"][state])
        if x is None:
            x = self.properties["ci_coefficients"][:, state]
        x = nat.f64(np.ravel(x))
        N, ov = self.n_atoms, self.n_occ * self.n_virt
        assert x.size == ov
        form_id = {"reference": 0, "exact": 1}[form]
        dp = nat.f64(inputs["dp"])
        ds = nat.f64(inputs["ds_rows"])
        dh = nat.f64(inputs["dh_rows"])
        dg = nat.f64(inputs["dgamma"])
        dglr = nat.f64(inputs["dgamma_lr"]) if self.gammafunction_lc else None
        grad = np.empty(3 * N)
        z = np.empty(ov) if return_z else None
        its = ctypes.c_int32(0)
        self.ctx.check(self.ctx.lib.tda_gradient(
            self.ctx.handle, form_id, float(omega), nat.dptr(x), nat.dptr(dp), nat.dptr(ds), nat.dptr(dh),
            nat.dptr(dg), nat.dptr(dglr), int(max_iter), float(conv), nat.dptr(grad), nat.dptr(z),
            ctypes.byref(its)))
        if return_z:
            return grad, z, its.value
        return grad

    def tda_gradient_lc(self, state, inputs, **kw):
        """tda_gradient of an LC system (excited_state.rs:259-487)."""
        if not self.gammafunction_lc:
            raise ValueError("tda_gradient_lc needs gamma_lr")
        return self.tda_gradient(state, inputs, **kw)

    def tda_gradient_nolc(self, state, inputs, **kw):
        """tda_gradient without long-range correction (excited_state.rs:56-257)."""
        if self.gammafunction_lc:
            raise ValueError("tda_gradient_nolc needs a system without gamma_lr")
        return self.tda_gradient(state, inputs, **kw)


class Fragment:
    """The monomer data a CT pair reads (fmo Monomer properties): atoms'
    AO counts, coordinates [n_atoms, 3] (bohr), occupied / virtual MO indices,
    orbitals, AO overlap, orbital energies and the monomer gamma."""

    def __init__(self, n_atoms, orbs_per_atom, xyz, occ_indices, virt_indices, orbs, s, orbe, gamma):
        self.n_atoms = n_atoms
        self.orbs_per_atom = np.ascontiguousarray(orbs_per_atom, dtype=np.int32)
        self.xyz = nat.f64(np.asarray(xyz).reshape(n_atoms, 3))
        self.occ_indices = list(occ_indices)
        self.virt_indices = list(virt_indices)
        self.orbs = nat.f64(orbs)
        self.s = nat.f64(s)
        self.orbe = nat.f64(np.ravel(orbe))
        self.gamma = nat.f64(gamma)


class TdaCTPair(_TdaEngine):
    """PairChargeTransfer H -> L: ``prepare_ct_tda`` (tda/ct_pair.rs:12-76),
    its DavidsonEngine (solvers/mod.rs:165-274) and ``run_ct_tda``
    (ct_pair.rs:125-180).  Holes in H's occupied MOs, particles in L's
    virtual MOs; q_ov over the H and L atoms (H first), q_oo of H, q_vv of L.

    ``s_hl`` is the H-L block of the AO overlap [n_ao_h, n_ao_l];
    ``gamma_hl`` / ``gamma_lr_hl`` the H-L blocks [n_atoms_h, n_atoms_l] of
    the full gamma and gamma_lr (``g0.slice(H atoms, L atoms)``, :24, :37).
    The pair always uses the per-atom-charges operator."""

    def __init__(self, h, l, s_hl, gamma_hl, gamma_lr_hl, precision="fp64", device=0, hbm_budget_bytes=0):
        self.ctx = nat.Context(device, precision, hbm_budget_bytes)
        self.h, self.l = h, l
        self.n_atoms = h.n_atoms + l.n_atoms
        self.n_atoms_oo, self.n_atoms_vv = h.n_atoms, l.n_atoms
        self.occ_indices = list(h.occ_indices)
        self.virt_indices = list(l.virt_indices)
        ob, self.n_occ, _, _ = _ranges(h.occ_indices, h.virt_indices)
        _, _, vb, self.n_virt = _ranges(l.occ_indices, l.virt_indices)
        self.xyz = nat.f64(np.concatenate([h.xyz, l.xyz]))
        self._s_hl = nat.f64(s_hl)
        self._gamma_hl = nat.f64(gamma_hl)
        self._gamma_lr_hl = nat.f64(gamma_lr_hl)
        self.gammafunction_lc = gamma_lr_hl is not None
        self.properties = {}
        self.last_stats = None
        self.ctx.check(self.ctx.lib.tda_set_ct_pair(
            self.ctx.handle, h.n_atoms, h.orbs.shape[0], nat.iptr(h.orbs_per_atom), nat.dptr(h.orbs),
            nat.dptr(h.s), nat.dptr(h.orbe), ob, self.n_occ, l.n_atoms, l.orbs.shape[0],
            nat.iptr(l.orbs_per_atom), nat.dptr(l.orbs), nat.dptr(l.s), nat.dptr(l.orbe), vb, self.n_virt,
            nat.dptr(self._s_hl), nat.dptr(h.gamma), nat.dptr(l.gamma), nat.dptr(self._gamma_hl),
            nat.dptr(self._gamma_lr_hl), nat.dptr(self.xyz)))

    def prepare_ct_tda(self, fetch_charges=True):
        self.prepare_tda(fetch_charges)

    def run_ct_tda(self, n_roots, max_iter, tolerance):
        return self.run_tda(n_roots, max_iter, tolerance)


class DenseEngine:
    """DavidsonEngine for a dense symmetric matrix (solvers/mod.rs:36-51):
    products h x, preconditioner r / (w - diag(h)) without a floor."""

    def __init__(self, h, device=0):
        self.h = nat.f64(h)
        assert self.h.ndim == 2 and self.h.shape[0] == self.h.shape[1]
        self.ctx = nat.Context(device, "fp64")
        self.ctx.check(self.ctx.lib.tda_set_dense(self.ctx.handle, self.h.shape[0], nat.dptr(self.h)))
        self.last_stats = None

    def close(self):
        self.ctx.close()

    def compute_products(self, x):
        x = nat.f64(x)
        squeeze = x.ndim == 1
        if squeeze:
            x = x[:, None].copy()
        ax = np.empty_like(x)
        self.ctx.check(self.ctx.lib.tda_compute_products(self.ctx.handle, nat.dptr(x), x.shape[1], nat.dptr(ax)))
        return ax[:, 0] if squeeze else ax

    def precondition(self, r_k, w_k):
        r = nat.f64(r_k)
        out = np.empty_like(r)
        self.ctx.check(self.ctx.lib.tda_precondition(self.ctx.handle, nat.dptr(r), float(w_k), nat.dptr(out)))
        return out

    def get_size(self):
        return self.h.shape[0]

    def davidson(self, n_roots, tolerance, max_iter):
        """Davidson::new(&mut h, initial_subspace(h.diag(), n_roots), n_roots,
        tolerance, max_iter) (basis.rs:500-501): (eigenvalues, eigenvectors
        [dim, n_roots]); raises DavidsonError if not converged."""
        dim = self.get_size()
        w = np.empty(n_roots)
        x = np.empty((dim, n_roots))
        st = nat.TdaStats()
        rc = self.ctx.lib.tda_run(self.ctx.handle, n_roots, max_iter, float(tolerance), nat.dptr(w), nat.dptr(x),
                                  None, None, None, nat.ctypes.byref(st))
        self.last_stats = st.as_dict()
        if rc == nat.TDA_ERR_NOT_CONVERGED:
            raise DavidsonError("Davidson Algorithm did not converge!")
        self.ctx.check(rc)
        return w, x


class TdaBatch:
    """Many fragments solved in one call: the FMO monomer loop
    (fmo/lcmo/basis.rs:462-465: ``prepare_tda`` + ``run_tda`` per monomer) and
    the CT-pair loop (basis.rs:172-238: ``run_ct_tda`` per pair).

    ``add_system`` takes ``TdaSystem``'s arguments, ``add_ct_pair`` those of
    ``TdaCTPair``; both take ``n_roots, max_iter, tolerance`` of the solve.
    ``run`` returns one dict per item (``ci_eigenvalues``, ``ci_coefficients``,
    ``q_trans``, ``tr_dipoles``, ``oscillator_strengths``, ``stats``,
    ``status``) and raises ``DavidsonError`` if an item did not converge
    (after every item has been solved).

    ``mode="grouped"`` (default): one Davidson for all fragments, every step
    one grouped launch per operation (charges-form operator; FP32 products
    with an FP64 verification unless ``precision="fp64"``).  ``mode="pool"``:
    worker threads with one context each solve whole fragments."""

    MODES = {"grouped": 0, "pool": 1}

    def __init__(self, precision="fp64", n_workers=8, device=0, hbm_budget_bytes=0, operator="ao", mode="grouped"):
        self.lib = nat.load()
        self.handle = nat.ctypes.c_void_p()
        rc = self.lib.tda_batch_create(device, n_workers, hbm_budget_bytes, nat.PREC[precision],
                                       nat.ctypes.byref(self.handle))
        if rc != nat.TDA_OK:
            raise nat.TdaError(rc, "tda_batch_create failed")
        self._check(self.lib.tda_batch_set_operator(self.handle, nat.OPERATOR[operator]))
        self._check(self.lib.tda_batch_set_mode(self.handle, self.MODES[mode]))
        self.items = []

    def _check(self, rc):
        if rc != nat.TDA_OK:
            raise nat.TdaError(rc, self.lib.tda_batch_last_error(self.handle).decode())

    def close(self):
        if self.handle:
            self.lib.tda_batch_destroy(self.handle)
            self.handle = nat.ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _outputs(self, n_atoms, ov, n_roots):
        return dict(ci_eigenvalues=np.empty(n_roots), ci_coefficients=np.empty((ov, n_roots)),
                    q_trans=np.empty((n_atoms, n_roots)), tr_dipoles=np.empty((3, n_roots)),
                    oscillator_strengths=np.empty(n_roots))

    def _add(self, idx, keep, out):
        if idx < 0:
            raise nat.TdaError(-idx, self.lib.tda_batch_last_error(self.handle).decode())
        self.items.append((keep, out))
        return idx

    def add_system(self, n_atoms, orbs_per_atom, xyz, occ_indices, virt_indices, orbs, s, orbe, gamma,
                   gamma_lr, n_roots, max_iter, tolerance):
        ob, o, vb, v = _ranges(occ_indices, virt_indices)
        keep = dict(opa=np.ascontiguousarray(orbs_per_atom, dtype=np.int32),
                    xyz=nat.f64(np.asarray(xyz).reshape(n_atoms, 3)), orbs=nat.f64(orbs), s=nat.f64(s),
                    orbe=nat.f64(np.ravel(orbe)), gamma=nat.f64(gamma), gamma_lr=nat.f64(gamma_lr))
        out = self._outputs(n_atoms, o * v, n_roots)
        idx = self.lib.tda_batch_add_system(
            self.handle, n_atoms, keep["orbs"].shape[0], nat.iptr(keep["opa"]), nat.dptr(keep["orbs"]),
            nat.dptr(keep["s"]), nat.dptr(keep["orbe"]), ob, o, vb, v, nat.dptr(keep["gamma"]),
            nat.dptr(keep["gamma_lr"]), nat.dptr(keep["xyz"]), n_roots, max_iter, float(tolerance),
            nat.dptr(out["ci_eigenvalues"]), nat.dptr(out["ci_coefficients"]), nat.dptr(out["q_trans"]),
            nat.dptr(out["tr_dipoles"]), nat.dptr(out["oscillator_strengths"]))
        return self._add(idx, keep, out)

    def add_ct_pair(self, h, l, s_hl, gamma_hl, gamma_lr_hl, n_roots, max_iter, tolerance):
        ob, o, _, _ = _ranges(h.occ_indices, h.virt_indices)
        _, _, vb, v = _ranges(l.occ_indices, l.virt_indices)
        keep = dict(h=h, l=l, s_hl=nat.f64(s_hl), gamma_hl=nat.f64(gamma_hl), gamma_lr_hl=nat.f64(gamma_lr_hl),
                    xyz=nat.f64(np.concatenate([h.xyz, l.xyz])))
        n_atoms = h.n_atoms + l.n_atoms
        out = self._outputs(n_atoms, o * v, n_roots)
        idx = self.lib.tda_batch_add_ct_pair(
            self.handle, h.n_atoms, h.orbs.shape[0], nat.iptr(h.orbs_per_atom), nat.dptr(h.orbs), nat.dptr(h.s),
            nat.dptr(h.orbe), ob, o, l.n_atoms, l.orbs.shape[0], nat.iptr(l.orbs_per_atom), nat.dptr(l.orbs),
            nat.dptr(l.s), nat.dptr(l.orbe), vb, v, nat.dptr(keep["s_hl"]), nat.dptr(h.gamma), nat.dptr(l.gamma),
            nat.dptr(keep["gamma_hl"]), nat.dptr(keep["gamma_lr_hl"]), nat.dptr(keep["xyz"]), n_roots, max_iter,
            float(tolerance), nat.dptr(out["ci_eigenvalues"]), nat.dptr(out["ci_coefficients"]),
            nat.dptr(out["q_trans"]), nat.dptr(out["tr_dipoles"]), nat.dptr(out["oscillator_strengths"]))
        return self._add(idx, keep, out)

    def run(self):
        n = len(self.items)
        status = np.zeros(n, dtype=np.int32)
        stats = (nat.TdaStats * max(n, 1))()
        rc = self.lib.tda_batch_run(self.handle, nat.iptr(status), stats)
        tot = nat.TdaStats()
        self._check(self.lib.tda_batch_last_totals(self.handle, nat.ctypes.byref(tot)))
        self.totals = tot.as_dict()
        results = []
        for i, (_, out) in enumerate(self.items):
            r = dict(out)
            r["stats"] = stats[i].as_dict()
            r["status"] = int(status[i])
            results.append(r)
        if rc == nat.TDA_ERR_NOT_CONVERGED:
            bad = [i for i in range(n) if status[i] == nat.TDA_ERR_NOT_CONVERGED]
            raise DavidsonError(f"Davidson Algorithm did not converge! (items {bad})")
        self._check(rc)
        return results
