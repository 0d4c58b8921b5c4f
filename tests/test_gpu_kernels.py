"""Kernel-level GPU checks: MFMA GEMM (every precision / operand layout /
edge tiles / split-K) and the subspace eigensolver, against numpy."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _nat():
    from tincr_amd import _native as nat
    return nat


def _round(x, prec):
    import torch
    t = torch.from_numpy(x)
    if prec == "fp32":
        return t.float().double().numpy()
    if prec == "bf16":
        return t.float().bfloat16().double().numpy()
    if prec == "fp16":
        return t.float().half().double().numpy()
    return x


# relative tolerance of the product (inputs rounded as the device does; what is
# left is the accumulation order / accumulator precision).  BF16: the device
# and torch may round an exact halfway input differently (one BF16 ulp,
# 2^-8 |a_ik b_kj| in one term); the largest shapes hit a few such inputs.
TOL = {"fp64": 1e-13, "fp32": 1e-5, "bf16": 2.5e-5, "fp16": 1e-5}


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16", "fp16"])
@pytest.mark.parametrize("a_kc,b_kc", [(True, False), (True, True), (False, False), (False, True)])
@pytest.mark.parametrize("shape", [(16, 16, 64), (37, 53, 70), (130, 129, 257), (300, 20, 4100), (7, 600, 96),
                                   (600, 520, 700), (520, 600, 4200)])
def test_gemm(prec, a_kc, b_kc, shape):
    nat = _nat()
    print(f"gemm {prec} a_kc={a_kc} b_kc={b_kc} shape={shape}", flush=True)
    m, n, k = shape
    rng = np.random.default_rng(m * 1000 + n + k)
    a = rng.uniform(-1, 1, (m, k))
    b = rng.uniform(-1, 1, (k, n))
    c0 = rng.uniform(-1, 1, (m, n))
    ar, br = _round(a, prec), _round(b, prec)
    ref = 0.5 * ar @ br - 2.0 * c0
    out = nat.debug_gemm(a, b, prec, a_kc, b_kc, alpha=0.5, beta=-2.0, c=c0)
    scale = np.abs(ar) @ np.abs(br) + 1.0
    assert np.abs(out - ref).max() / scale.max() < TOL[prec]


@pytest.mark.parametrize("ksplit", [1, 3, 8])
def test_gemm_split_k(ksplit):
    nat = _nat()
    rng = np.random.default_rng(ksplit)
    a = rng.standard_normal((70, 3000))
    b = rng.standard_normal((3000, 45))
    out = nat.debug_gemm(a, b, "fp64", True, True, ksplit=ksplit)
    assert np.abs(out - a @ b).max() < 1e-11


def test_mfma_layout_identity():
    """16x16 identity times a matrix of distinct values: any lane-layout slip
    in the MFMA operand or result mapping moves values."""
    nat = _nat()
    for prec in ["fp64", "fp32", "bf16", "fp16"]:
        b = (np.arange(16 * 64).reshape(64, 16) % 61).astype(float)
        a = np.zeros((16, 64))
        a[np.arange(16), np.arange(16) * 3] = 1.0
        out = nat.debug_gemm(a, b, prec, True, False)
        assert np.array_equal(out, a @ b), prec


@pytest.mark.parametrize("m,n", [(1, 1), (2, 2), (5, 3), (40, 10), (150, 20), (300, 20), (301, 64)])
def test_eigh(m, n):
    nat = _nat()
    print(f"eigh m={m} n={n}", flush=True)
    rng = np.random.default_rng(m)
    x = rng.standard_normal((m, m))
    h = 0.5 * (x + x.T) + np.diag(np.linspace(0.1, 2.0, m))
    w_ref, v_ref = np.linalg.eigh(h)
    w, y = nat.debug_eigh(np.triu(h) + np.triu(h, 1).T * 0.0 + np.tril(rng.standard_normal((m, m)), -1), n)
    assert np.abs(w - w_ref[:n]).max() < 1e-12 * max(1.0, np.abs(w_ref).max())
    resid = h @ y.T - y.T * w[None, :]
    assert np.abs(resid).max() < 1e-11
    assert np.abs(y @ y.T - np.eye(n)).max() < 1e-11


def test_eigh_degenerate_clusters():
    """Exactly degenerate and tightly clustered eigenvalues (symmetric
    molecules give these): vectors must stay orthonormal."""
    nat = _nat()
    rng = np.random.default_rng(7)
    m = 60
    q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    ev = np.linspace(0.2, 1.5, m)
    ev[:6] = [0.3, 0.3, 0.3, 0.3 + 1e-9, 0.4, 0.4]
    h = (q * ev) @ q.T
    h = 0.5 * (h + h.T)
    w, y = nat.debug_eigh(h, 8)
    w_ref = np.sort(ev)[:8]
    assert np.abs(w - w_ref).max() < 1e-12
    assert np.abs(y @ y.T - np.eye(8)).max() < 1e-10
    assert np.abs(h @ y.T - y.T * w[None, :]).max() < 1e-10
