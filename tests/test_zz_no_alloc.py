"""tda_trans_charges_noalloc (device pointers, fixed scratch array, no device
allocation) against tda_trans_charges (host arrays, library-owned device
memory) through the C ABI: the same kernels, so the same charges to the
last bits (1e-14 allowed for a different S C split); two calls in a row, the
second on a smaller system; size and count errors rejected."""
import ctypes

import numpy as np
import pytest

from tests._fixtures import molecule

pytestmark = pytest.mark.gpu


def _lib():
    from tincr_amd import _native as nat
    lib = nat.load()
    vp = ctypes.c_void_p
    i = ctypes.c_int
    lib.tda_trans_charges_noalloc.argtypes = [vp, i, i, vp, vp, vp, i, i, i, i, vp, vp, vp]
    lib.tda_trans_charges_noalloc.restype = ctypes.c_int
    hip = nat.hip_runtime()
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [vp]
    return nat, lib, hip


H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class DeviceArrays:
    """Test-side device buffers (hipMalloc by the test, not the library)."""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def download(self, p, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc(ctx, lib, hip, m, orbs_per_atom=None, n_ao=None):
    opa = np.asarray(m["orbs_per_atom"] if orbs_per_atom is None else orbs_per_atom, dtype=np.int32)
    n_at = len(opa)
    nao = m["orbs"].shape[0] if n_ao is None else n_ao
    occ, virt = m["occ_indices"], m["virt_indices"]
    o, v = len(occ), len(virt)
    dev = DeviceArrays(hip)
    try:
        d_opa = dev.upload(opa)
        d_c = dev.upload(np.asarray(m["orbs"], dtype=np.float64))
        d_s = dev.upload(np.asarray(m["s"], dtype=np.float64))
        shapes = [(n_at, a * b) for a, b in ((o, v), (o, o), (v, v))]
        outs = [dev.upload(np.full(sh, np.nan)) for sh in shapes]
        rc = lib.tda_trans_charges_noalloc(ctx.handle, n_at, nao, d_opa, d_c, d_s, occ[0], o, virt[0], v, *outs)
        got = [dev.download(p, sh) for p, sh in zip(outs, shapes)]
    finally:
        dev.free()
    return rc, got


@pytest.mark.parametrize("order", [("uracil", "h2o"), ("benzene", "ammonia")])
def test_noalloc_matches_trans_charges(order):
    from tincr_amd import tda
    nat, lib, hip = _lib()
    ctx = nat.Context()
    try:
        for name in order:            # the second system is the smaller one
            print(f"noalloc trans_charges {name}", flush=True)
            m = molecule(name)
            rc, got = _noalloc(ctx, lib, hip, m)
            assert rc == nat.TDA_OK, ctx.lib.tda_last_error(ctx.handle)
            ref = tda.trans_charges(m["n_atoms"], m["orbs_per_atom"], m["orbs"], m["s"], m["occ_indices"],
                                    m["virt_indices"])
            for a, b in zip(got, ref):
                assert a.shape == b.shape
                assert np.abs(a - b).max() <= 1e-14
    finally:
        ctx.close()


def test_noalloc_rejects_oversize_and_bad_counts():
    nat, lib, hip = _lib()
    ctx = nat.Context()
    try:
        m = molecule("h2o")
        print("noalloc oversize", flush=True)
        rc, _ = _noalloc(ctx, lib, hip, m, n_ao=4000)     # checked on the host before any work
        assert rc == nat.TDA_ERR_BAD_ARG
        print("noalloc bad counts", flush=True)
        bad = np.array(m["orbs_per_atom"], dtype=np.int32)
        bad[0] -= 1                                   # counts sum to n_ao - 1: offsets stay in bounds
        rc, _ = _noalloc(ctx, lib, hip, m, orbs_per_atom=bad)
        assert rc == nat.TDA_ERR_BAD_ARG
        rc, got = _noalloc(ctx, lib, hip, m)                # and the next good call is unaffected
        assert rc == nat.TDA_OK
        assert np.isfinite(got[0]).all()
    finally:
        ctx.close()
