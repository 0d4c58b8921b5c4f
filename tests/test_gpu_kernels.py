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
                                   (300, 12, 4100), (13, 1000, 300), (5, 9, 9000),
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


@pytest.mark.parametrize("shape", [(128, 16, 4096), (37, 13, 1031), (100, 30, 500), (16, 3000, 128),
                                   (64, 64, 300), (300, 300, 64), (1, 5, 7)])
@pytest.mark.parametrize("a32,b32", [(True, True), (True, False), (False, True)])
@pytest.mark.parametrize("a_kc,b_kc", [(True, True), (True, False), (False, False), (False, True)])
def test_gemm_f32_stored_operands(shape, a32, b32, a_kc, b_kc):
    """FP64 products of FP32-stored operands (the FP32 Davidson basis): equal
    to the FP64 product of the FP32-rounded inputs to FP64 rounding, so a
    wrong widening index or vector width moves values by O(1)."""
    nat = _nat()
    m, n, k = shape
    print(f"gemm f32-stored a32={a32} b32={b32} a_kc={a_kc} b_kc={b_kc} shape={shape}", flush=True)
    rng = np.random.default_rng(m * 7919 + n * 31 + k)
    a = rng.standard_normal((m, k))
    b = rng.standard_normal((k, n))
    c0 = rng.standard_normal((m, n))
    ar = a.astype(np.float32).astype(np.float64) if a32 else a
    br = b.astype(np.float32).astype(np.float64) if b32 else b
    ref = -1.0 * ar @ br + 1.0 * c0
    out = nat.debug_gemm_f32store(a, b, a32, b32, a_kc, b_kc, alpha=-1.0, beta=1.0, c=c0)
    scale = (np.abs(ar) @ np.abs(br)).max() + 1.0
    assert np.abs(out - ref).max() / scale < 1e-14


@pytest.mark.parametrize("shape", [(1, 1, 7), (16, 16, 64), (37, 13, 1031), (128, 16, 20000), (70, 45, 3001),
                                   (200, 130, 2048)])
@pytest.mark.parametrize("a32,b32", [(False, False), (True, True), (True, False)])
def test_rowdot(shape, a32, b32):
    """Basis dot products V P^T (direct-from-HBM MFMA stream, slices summed in
    a second kernel) against numpy on the same rounded inputs, to FP64
    rounding; row / column / length tails and several slices."""
    nat = _nat()
    m, n, k = shape
    print(f"rowdot {shape} a32={a32} b32={b32}", flush=True)
    rng = np.random.default_rng(m * 131 + n * 7 + k)
    a, b, c0 = rng.standard_normal((m, k)), rng.standard_normal((n, k)), rng.standard_normal((m, n))
    ar = a.astype(np.float32).astype(float) if a32 else a
    br = b.astype(np.float32).astype(float) if b32 else b
    out = nat.debug_rowdot(a, b, a32, b32, alpha=-0.5, beta=2.0, c=c0)
    ref = -0.5 * ar @ br.T + 2.0 * c0
    scale = (np.abs(ar) @ np.abs(br).T).max() + 1.0
    assert np.abs(out - ref).max() / scale < 1e-14


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


@pytest.mark.parametrize("m,n", [(1, 1), (2, 2), (5, 3), (40, 10), (64, 8), (65, 16), (129, 20), (150, 20),
                                 (257, 30), (300, 20), (301, 64), (700, 32), (1024, 8)])
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


# --- the K-contiguous 16-bit GEMM (AO exchange products) -------------------

def _r16(x, prec):
    return _round(np.asarray(x, dtype=float), prec)


def _ref16(a, b, prec):
    """sum_k a[m, k] b[z, n, k] of the rounded inputs, FP64."""
    return np.einsum("mk,znk->zmn", _r16(a, prec), _r16(b, prec), optimize=True)


def _inputs16(rng, prec, *shapes):
    """Uniform(-1, 1) values already representable in the 16-bit type, so the
    device conversion is exact and only accumulation / output rounding remain."""
    return [_r16(rng.uniform(-1, 1, sh), prec) for sh in shapes]


# FP32 accumulation of up to ~500 exact products, relative to sum |a_k b_k|
ACC16 = 4e-5


# (M, N, K, batch): edge tiles in M / N, K not a multiple of 8 or of 64, K
# below one slice, and a batch large enough for the 256x256 tile (>= 512 tiles)
# (300, 290, 192, 2): K a multiple of 64 with edge tiles — the big tile then
# loads its K slices straight into LDS
SHAPES16 = [(16, 16, 64, 1), (37, 53, 70, 2), (130, 129, 5, 3), (200, 333, 501, 2), (257, 255, 130, 1),
            (520, 520, 72, 60), (300, 290, 192, 2)]


@pytest.fixture(params=["rule", "large"])
def tile16(request, monkeypatch):
    """The tile rule of small problems, or the 256x256 tile forced (the
    diagonal epilogue stays on 128x128)."""
    if request.param == "large":
        monkeypatch.setenv("TINCR_TDA_TILE16", "1")
    else:
        monkeypatch.delenv("TINCR_TDA_TILE16", raising=False)
    return request.param


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("shape", SHAPES16)
def test_gemm16_store(prec, shape, tile16):
    nat = _nat()
    m, n, k, bt = shape
    print(f"gemm16 store {prec} {shape} tile={tile16}", flush=True)
    rng = np.random.default_rng(m + 7 * n + 13 * k + bt)
    a, b = _inputs16(rng, prec, (m, k), (bt, n, k))
    out, used = nat.debug_gemm16(a, b, prec, "store", alpha=-0.25, pad=3 + (k % 5))
    assert used
    ref = -0.25 * _ref16(a, b, prec)
    # output rounded to the 16-bit type (half an ulp: 2^-11 / 2^-8 of |value|) plus the accumulation
    ulp = 2.0 ** (-11 if prec == "fp16" else -8)
    scale = 0.25 * np.einsum("mk,znk->zmn", np.abs(a), np.abs(b), optimize=True)
    assert np.all(np.abs(out - ref) <= ulp * (np.abs(ref) + ACC16 * scale) + ACC16 * scale + 1e-30)




@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("shape", [(37, 53, 70, 2), (200, 333, 501, 2), (520, 520, 72, 60), (300, 290, 192, 2)])
def test_gemm16_diag(prec, shape, tile16):
    nat = _nat()
    print(f"gemm16 diag {prec} {shape} tile={tile16}", flush=True)
    m, n, k, bt = shape
    rng = np.random.default_rng(k)
    a, b = _inputs16(rng, prec, (m, k), (bt, n, k))
    dg = rng.uniform(0.1, 1.5, (m, n))
    x = rng.uniform(-1, 1, (bt, m, n))
    out, used = nat.debug_gemm16(a, b, prec, "diag", alpha=1.0, diag=dg, x=x, pad=8)
    assert used
    ref = _ref16(a, b, prec) + dg[None] * x
    scale = np.einsum("mk,znk->zmn", np.abs(a), np.abs(b), optimize=True) + 1.0
    assert np.all(np.abs(out - ref) <= ACC16 * scale)


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("case", ["ao", "distinct", "identity"])
@pytest.mark.parametrize("shape", [(200, 333, 501, 2), (520, 520, 72, 60), (300, 290, 192, 2)])
def test_gemm16_hadamard(prec, case, shape, tile16):
    """H[r(m)][r(n)] with r = an AO->atom table applied modulo its length
    (runs of equal keys, a wrap inside tiles: LDS-staged table), distinct
    keys (table larger than the LDS: factors read from H) and no table."""
    nat = _nat()
    m, n, k, bt = shape
    rng = np.random.default_rng(11 * m + n)
    a, b = _inputs16(rng, prec, (m, k), (bt, n, k))
    if case == "ao":
        per_atom = rng.integers(1, 5, size=400)
        hidx = np.repeat(np.arange(400), per_atom)[:150]
        nh = int(hidx.max()) + 1
    elif case == "distinct":
        hidx = rng.permutation(max(m, n))
        nh = len(hidx)
    else:
        hidx, nh = None, max(m, n)
    h = rng.uniform(0.1, 0.5, (nh, nh))
    out, used = nat.debug_gemm16(a, b, prec, "hadamard", alpha=-0.25, h=h, hidx=hidx, pad=5)
    assert used
    if hidx is not None:
        # the precomputed FP32 factor table of the AO operator gives the same result
        out_t, used_t = nat.debug_gemm16(a, b, prec, "hadamard", alpha=-0.25, h=h, hidx=hidx, pad=5,
                                         full_table=True)
        assert used_t
        assert np.array_equal(out_t, out)
    r = (lambda x: hidx[x % len(hidx)]) if hidx is not None else (lambda x: x)
    hf = h[np.ix_(r(np.arange(m)), r(np.arange(n)))]
    ref = -0.25 * _ref16(a, b, prec) * hf[None]
    ulp = 2.0 ** (-11 if prec == "fp16" else -8)
    # H is applied in FP32 (one more rounding of the factor)
    scale = 0.25 * np.einsum("mk,znk->zmn", np.abs(a), np.abs(b), optimize=True) * hf[None]
    assert np.all(np.abs(out - ref) <= ulp * (np.abs(ref) + ACC16 * scale) + ACC16 * scale + 1e-30)


@pytest.mark.parametrize("epi", ["store", "diag", "hadamard"])
@pytest.mark.parametrize("shape", [(37, 53, 70, 2), (200, 333, 501, 2), (520, 520, 72, 60), (300, 290, 192, 2)])
def test_gemm16_split(epi, shape, tile16):
    """FP16 hi / lo split operands and output: about FP32 accuracy.  Plain
    FP16 inputs would be off by ~2^-11 per term (~2e-5 of sum |a b| here);
    the split must stay two orders below that."""
    nat = _nat()
    m, n, k, bt = shape
    print(f"gemm16 split {epi} {shape} tile={tile16}", flush=True)
    rng = np.random.default_rng(3 * m + n + k)
    a = rng.uniform(-1, 1, (m, k))
    b = rng.uniform(-1, 1, (bt, n, k))
    kw = {}
    ref = np.einsum("mk,znk->zmn", a, b, optimize=True)
    if epi == "diag":
        kw = {"diag": rng.uniform(0.1, 1.5, (m, n)), "x": rng.uniform(-1, 1, (bt, m, n))}
        ref = ref + kw["diag"][None] * kw["x"]
    elif epi == "hadamard":
        hidx = np.repeat(np.arange(60), 3)[:170]
        h = rng.uniform(0.1, 0.5, (60, 60))
        kw = {"h": h, "hidx": hidx}
        r = hidx[np.arange(max(m, n)) % len(hidx)]
        ref = ref * h[np.ix_(r[:m], r[:n])][None]
    out, used = nat.debug_gemm16(a, b, "fp16", epi, alpha=1.0, pad=8, split=True, **kw)
    assert used
    scale = np.einsum("mk,znk->zmn", np.abs(a), np.abs(b), optimize=True)
    err = np.abs(out - ref)
    assert err.max() <= 2e-6 * scale.max()
    assert np.all(err <= 2e-5 * scale + 1e-12)
