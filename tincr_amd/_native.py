"""ctypes binding of libtincr_tda.so (include/tincr_tda.h).

The library must have been built (``python -m tincr_amd.build``); loading
fails loudly otherwise -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TINCR_TDA_LIB") or os.path.join(_HERE, "lib", "libtincr_tda.so")

TDA_OK = 0
TDA_ERR_NOT_CONVERGED = 1
TDA_ERR_BAD_ARG = 2
TDA_ERR_OOM = 3
TDA_ERR_HIP = 4
TDA_ERR_STATE = 5

PREC = {"fp64": 0, "fp32": 1, "bf16": 2, "fp16": 3}
OPERATOR = {"ao": 0, "charges": 1}

# every symbol include/tincr_tda.h declares
EXPORTS = [
    "tda_ctx_create", "tda_ctx_destroy", "tda_last_error", "tda_set_precision", "tda_set_operator",
    "tda_trans_charges", "tda_trans_charges_noalloc", "tda_orbe_differences", "tda_set_system", "tda_set_ct_pair", "tda_set_dense", "tda_prepare",
    "tda_clear_charges", "tda_get_charges", "tda_compute_products", "tda_precondition", "tda_run",
    "tda_adjoint", "tda_adjoint_nuclear", "tda_gradient", "tda_zvector",
    "tda_batch_create", "tda_batch_destroy", "tda_batch_last_error", "tda_batch_set_operator", "tda_batch_set_mode", "tda_batch_last_totals",
    "tda_batch_add_system", "tda_batch_add_ct_pair", "tda_batch_clear", "tda_batch_run",
]


class TdaStats(ctypes.Structure):
    _fields_ = [
        ("n_iter", ctypes.c_int32),
        ("n_iter_fp64", ctypes.c_int32),
        ("n_products", ctypes.c_int32),
        ("n_products_fp64", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("n_atom_blocks", ctypes.c_int32),
        ("n_iter_fp32", ctypes.c_int32),
        ("n_products_fp32", ctypes.c_int32),
        ("split_fp32", ctypes.c_int32),
        ("max_residual", ctypes.c_double),
        ("t_total_ms", ctypes.c_double),
        ("t_qbuild_ms", ctypes.c_double),
        ("t_products_ms", ctypes.c_double),
        ("exch_ms", (ctypes.c_double * 3) * 2),
        ("exch_launches", (ctypes.c_int32 * 3) * 2),
        ("exch_flops", (ctypes.c_double * 3) * 2),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if name.startswith("exch_"):
                v = [[v[i][j] for j in range(3)] for i in range(2)]
            out[name] = v
        return out


class TdaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first (python -m tincr_amd.build)")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    I32 = ctypes.POINTER(ctypes.c_int32)
    i = ctypes.c_int
    lib.tda_ctx_create.argtypes = [i, ctypes.c_size_t, i, ctypes.POINTER(P)]
    lib.tda_ctx_destroy.argtypes = [P]
    lib.tda_ctx_destroy.restype = None
    lib.tda_last_error.argtypes = [P]
    lib.tda_last_error.restype = ctypes.c_char_p
    lib.tda_set_precision.argtypes = [P, i]
    lib.tda_set_operator.argtypes = [P, i]
    lib.tda_trans_charges.argtypes = [P, i, i, I32, D, D, i, i, i, i, D, D, D]
    lib.tda_trans_charges_noalloc.argtypes = [P, i, i, P, P, P, i, i, i, i, P, P, P]
    lib.tda_orbe_differences.argtypes = [D, i, D, i, D]
    lib.tda_set_system.argtypes = [P, i, i, I32, D, D, D, i, i, i, i, D, D, D]
    lib.tda_set_ct_pair.argtypes = [P, i, i, I32, D, D, D, i, i, i, i, I32, D, D, D, i, i, D, D, D, D, D, D]
    lib.tda_set_dense.argtypes = [P, i, D]
    lib.tda_prepare.argtypes = [P]
    lib.tda_clear_charges.argtypes = [P]
    lib.tda_get_charges.argtypes = [P, D, D, D, D]
    lib.tda_compute_products.argtypes = [P, D, i, D]
    lib.tda_zvector.argtypes = [P, D, i, i, ctypes.c_double, D, ctypes.POINTER(ctypes.c_int32)]
    lib.tda_precondition.argtypes = [P, D, ctypes.c_double, D]
    lib.tda_run.argtypes = [P, i, i, ctypes.c_double, D, D, D, D, D, ctypes.POINTER(TdaStats)]
    lib.tda_adjoint.argtypes = [P, i, D, D, D, D, D, D]
    lib.tda_adjoint_nuclear.argtypes = [P, i, D, D, D, D]
    lib.tda_gradient.argtypes = [P, i, ctypes.c_double, D, D, D, D, D, D, i, ctypes.c_double, D, D,
                                 ctypes.POINTER(ctypes.c_int32)]
    lib.tda_batch_create.argtypes = [i, i, ctypes.c_size_t, i, ctypes.POINTER(P)]
    lib.tda_batch_destroy.argtypes = [P]
    lib.tda_batch_destroy.restype = None
    lib.tda_batch_last_error.argtypes = [P]
    lib.tda_batch_last_error.restype = ctypes.c_char_p
    lib.tda_batch_set_operator.argtypes = [P, i]
    lib.tda_batch_set_mode.argtypes = [P, i]
    lib.tda_batch_last_totals.argtypes = [P, ctypes.POINTER(TdaStats)]
    lib.tda_batch_add_system.argtypes = [P, i, i, I32, D, D, D, i, i, i, i, D, D, D, i, i, ctypes.c_double,
                                         D, D, D, D, D]
    lib.tda_batch_add_ct_pair.argtypes = [P, i, i, I32, D, D, D, i, i, i, i, I32, D, D, D, i, i, D, D, D, D, D, D,
                                          i, i, ctypes.c_double, D, D, D, D, D]
    lib.tda_batch_clear.argtypes = [P]
    lib.tda_batch_run.argtypes = [P, I32, ctypes.POINTER(TdaStats)]
    for name in EXPORTS:
        if name not in ("tda_ctx_destroy", "tda_last_error", "tda_batch_destroy", "tda_batch_last_error"):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


_hip = None


def hip_runtime():
    """The HIP runtime libtincr_tda.so is linked against (same loaded
    instance: libamdhip64.so.7 is resolved by soname after the library), for
    page-locked host buffers and device-wide synchronisation."""
    global _hip
    if _hip is None:
        load()
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
        h.hipHostMalloc.restype = ctypes.c_int
        h.hipHostFree.argtypes = [ctypes.c_void_p]
        h.hipHostFree.restype = ctypes.c_int
        h.hipDeviceSynchronize.restype = ctypes.c_int
        _hip = h
    return _hip


class PinnedArray:
    """A float64 numpy array in page-locked host memory (hipHostMalloc): the
    device copies into it directly.  `.array` is valid while this object
    lives."""

    def __init__(self, shape):
        h = hip_runtime()
        n = int(np.prod(shape))
        self.ptr = ctypes.c_void_p()
        rc = h.hipHostMalloc(ctypes.byref(self.ptr), max(n, 1) * 8, 0)
        if rc != 0:
            raise TdaError(TDA_ERR_OOM, f"hipHostMalloc of {n * 8} bytes failed ({rc})")
        buf = (ctypes.c_double * max(n, 1)).from_address(self.ptr.value)
        self.array = np.frombuffer(buf, dtype=np.float64, count=n).reshape(shape)

    def close(self):
        if self.ptr:
            self.array = None
            hip_runtime().hipHostFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_synchronize():
    rc = hip_runtime().hipDeviceSynchronize()
    if rc != 0:
        raise TdaError(TDA_ERR_HIP, f"hipDeviceSynchronize failed ({rc})")


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One device context (one HIP stream)."""

    def __init__(self, device=0, precision="fp64", hbm_budget_bytes=0):
        self.lib = load()
        self.handle = ctypes.c_void_p()
        rc = self.lib.tda_ctx_create(device, hbm_budget_bytes, PREC[precision], ctypes.byref(self.handle))
        if rc != TDA_OK:
            raise TdaError(rc, "tda_ctx_create failed")
        self.precision = precision

    def close(self):
        if self.handle:
            self.lib.tda_ctx_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc, allow=()):
        if rc != TDA_OK and rc not in allow:
            raise TdaError(rc, self.lib.tda_last_error(self.handle).decode())
        return rc

    def set_operator(self, op):
        self.check(self.lib.tda_set_operator(self.handle, OPERATOR[op]))

    def set_precision(self, precision):
        self.check(self.lib.tda_set_precision(self.handle, PREC[precision]))
        self.precision = precision


def load_debug():
    """Bind the kernel-level hooks of include/tincr_tda_debug.h."""
    lib = load()
    D = ctypes.POINTER(ctypes.c_double)
    i, i64 = ctypes.c_int, ctypes.c_int64
    lib.tda_debug_gemm.argtypes = [i, i, i, i64, i64, i64, D, D, D, ctypes.c_double, ctypes.c_double, i]
    lib.tda_debug_gemm.restype = ctypes.c_int
    lib.tda_debug_gemm_f32store.argtypes = [i, i, i64, i64, i64, D, D, D, ctypes.c_double, ctypes.c_double, i, i, i]
    lib.tda_debug_gemm_f32store.restype = ctypes.c_int
    lib.tda_debug_rowdot.argtypes = [i, i, i64, D, i, D, i, D, ctypes.c_double, ctypes.c_double]
    lib.tda_debug_rowdot.restype = ctypes.c_int
    I = ctypes.POINTER(ctypes.c_int)
    lib.tda_debug_gemm16.argtypes = [i, i, i64, i64, i64, i, i64, i64, D, D, ctypes.c_double, D, i, I, i, D, D, D, i,
                                     I]
    lib.tda_debug_gemm16.restype = ctypes.c_int
    lib.tda_debug_eigh.argtypes = [i, i, D, D, D]
    lib.tda_debug_eigh.restype = ctypes.c_int
    lib.tda_debug_eigh_timed.argtypes = [i, i, D, D, D, D]
    lib.tda_debug_eigh_timed.restype = ctypes.c_int
    lib.tda_debug_copy.argtypes = [ctypes.c_int64, i]
    lib.tda_debug_copy.restype = ctypes.c_int
    return lib


def debug_gemm(a, b, precision="fp64", a_kc=True, b_kc=False, alpha=1.0, beta=0.0, c=None, ksplit=0):
    """alpha * A @ B (+ beta * C) through the MFMA GEMM.  a: [M, K]; b: [K, N]."""
    lib = load_debug()
    m, k = a.shape
    n = b.shape[1]
    a_st = f64(a if a_kc else a.T)
    b_st = f64(b.T if b_kc else b)
    out = np.zeros((m, n)) if c is None else f64(c).copy()
    rc = lib.tda_debug_gemm(PREC[precision], int(a_kc), int(b_kc), m, n, k, dptr(a_st), dptr(b_st), dptr(out),
                            alpha, beta, ksplit)
    if rc != TDA_OK:
        raise TdaError(rc, "tda_debug_gemm failed")
    return out


def debug_gemm_f32store(a, b, a32=True, b32=True, a_kc=True, b_kc=False, alpha=1.0, beta=0.0, c=None, ksplit=0):
    """alpha * fl32(A) @ fl32(B) (+ beta * C) as an FP64 product whose flagged
    operands are stored as FP32 on the device.  a: [M, K]; b: [K, N]."""
    lib = load_debug()
    m, k = a.shape
    n = b.shape[1]
    a_st = f64(a if a_kc else a.T)
    b_st = f64(b.T if b_kc else b)
    out = np.zeros((m, n)) if c is None else f64(c).copy()
    rc = lib.tda_debug_gemm_f32store(int(a_kc), int(b_kc), m, n, k, dptr(a_st), dptr(b_st), dptr(out), alpha, beta,
                                     ksplit, int(a32), int(b32))
    if rc != TDA_OK:
        raise TdaError(rc, "tda_debug_gemm_f32store failed")
    return out


def debug_rowdot(a, b, a32=False, b32=False, alpha=1.0, beta=0.0, c=None):
    """alpha * A @ B^T (+ beta * C) through the basis stream kernel.  a: [M, K]; b: [N, K]."""
    lib = load_debug()
    m, k = a.shape
    n = b.shape[0]
    out = np.zeros((m, n)) if c is None else f64(c).copy()
    rc = lib.tda_debug_rowdot(m, n, k, dptr(f64(a)), int(a32), dptr(f64(b)), int(b32), dptr(out), alpha, beta)
    if rc != TDA_OK:
        raise TdaError(rc, "tda_debug_rowdot failed")
    return out


def debug_gemm16(a, b, precision="fp16", epi="store", alpha=1.0, h=None, hidx=None, diag=None, x=None, pad=0,
                 split=False, full_table=False):
    """The K-contiguous 16-bit GEMM: a [M, K], b [batch, N, K] -> [batch, M, N]
    (FP64 copy of the device result) and whether the dedicated kernel ran.
    Rows are stored with at least `pad` NaN elements past K (never read), the
    row stride rounded up to 8 elements (16 bytes)."""
    lib = load_debug()
    m, k = a.shape
    bt, n, _ = b.shape
    lda = ldb = (k + pad + 7) // 8 * 8
    a_st = np.full((m, lda), np.nan)
    a_st[:, :k] = a
    b_st = np.full((bt, n, ldb), np.nan)
    b_st[:, :, :k] = b
    out = np.zeros((bt, m, n))
    used = ctypes.c_int(-1)
    e = {"store": 0, "diag": 1, "hadamard": 2}[epi]
    hh = f64(h) if h is not None else None
    hi = np.ascontiguousarray(hidx, dtype=np.int32) if hidx is not None else None
    dg = f64(diag) if diag is not None else None
    xx = f64(x) if x is not None else None
    rc = lib.tda_debug_gemm16(PREC[precision], e, m, n, k, bt, lda, ldb, dptr(f64(a_st)), dptr(f64(b_st)), alpha,
                              dptr(hh) if hh is not None else None, hh.shape[0] if hh is not None else 0,
                              hi.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if hi is not None else None,
                              len(hi) if hi is not None else 0, dptr(dg) if dg is not None else None,
                              dptr(xx) if xx is not None else None, dptr(out), int(split) | (2 if full_table else 0),
                              ctypes.byref(used))
    if rc != TDA_OK:
        raise TdaError(rc, "tda_debug_gemm16 failed")
    return out, used.value == 1


def debug_eigh(h, n, timed=False):
    lib = load_debug()
    h = f64(h)
    m = h.shape[0]
    ev = np.empty(n)
    y = np.empty((n, m))
    ph = np.zeros(5)
    rc = lib.tda_debug_eigh_timed(m, n, dptr(h), dptr(ev), dptr(y), dptr(ph) if timed else None)
    if rc != TDA_OK:
        raise TdaError(rc, "tda_debug_eigh failed")
    return (ev, y, ph) if timed else (ev, y)


def debug_copy(nbytes, reps=1):
    """`reps` launches of calib_copy_kernel moving `nbytes` in and out (a
    known-traffic reference for counter runs)."""
    rc = load_debug().tda_debug_copy(int(nbytes), int(reps))
    if rc != TDA_OK:
        raise TdaError(rc, "tda_debug_copy failed")
