"""Kernel-level parity of the HIP data-movement and GEMM kernels (GPU).

Data movement (permute, sum_axis, trace_pair, diag_pair, select_axis) must be
bit-identical to the oracle's restatement of src/tensor.rs:453-669; the max
magnitude behind the power-of-two rescales exact against NumPy. GEMMs are
checked against float64/complex128 NumPy products with a precision-scaled
tolerance.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

DTYPES = [np.float32, np.float64, np.complex64, np.complex128]


def _rand(shape, dtype, rng):
    a = rng.standard_normal(shape)
    if np.issubdtype(np.dtype(dtype), np.complexfloating):
        a = a + 1j * rng.standard_normal(shape)
    return np.ascontiguousarray(a.astype(dtype))


def _gpu_op(op, a, args, out_shape):
    import torch
    from arctn_amd import _lib

    L = _lib.lib()
    code = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.complex64): 2,
            np.dtype(np.complex128): 3}[a.dtype]
    ta = torch.from_numpy(a).cuda()
    out = torch.empty(out_shape, dtype=ta.dtype, device="cuda")
    sh = (C.c_size_t * max(a.ndim, 1))(*a.shape)
    ar = (C.c_size_t * max(len(args), 1))(*args)
    _lib.check(L.atn_tensor_op(code, op, a.ndim, sh, ar, len(args), ta.data_ptr(), out.data_ptr(), None))
    torch.cuda.synchronize()
    return out.cpu().numpy()


PERM_CASES = [
    ((2,) * 10, None), ((2,) * 18, None), ((2,) * 24, None), ((3, 5, 2, 4), None), ((257, 3), (1, 0)),
    ((4096, 64), (1, 0)), ((64, 4096), (1, 0)), ((3, 3, 3, 3, 3, 3, 3), None), ((5, 7, 11, 13), (3, 1, 0, 2)),
    ((1, 2, 1, 3, 1), (4, 3, 2, 1, 0)), ((2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 4, 4), None), ((1024, 2, 1024), (2, 1, 0)),
]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", range(len(PERM_CASES)))
def test_permute_bit_exact(cuda, dtype, case):
    rng = np.random.default_rng(100 + case)
    shape, perm = PERM_CASES[case]
    if perm is None:
        perm = tuple(rng.permutation(len(shape)))
    a = _rand(shape, dtype, rng)
    want = O.tensor_op("permute", a, list(perm))
    assert np.array_equal(want, np.transpose(a, perm))  # oracle pinned to NumPy transpose
    got = _gpu_op(0, a, list(perm), want.shape)
    assert got.tobytes() == want.tobytes()


@pytest.mark.parametrize("dtype", DTYPES)
def test_reductions_and_gathers_bit_exact(cuda, dtype):
    rng = np.random.default_rng(7)
    a = _rand((3, 4, 5, 2), dtype, rng)
    for ax in range(4):
        want = O.tensor_op("sum_axis", a, ax)
        assert _gpu_op(1, a, [ax], want.shape).tobytes() == want.tobytes()
    for ax, i in [(0, 2), (1, 3), (3, 1), (2, 0)]:
        want = O.tensor_op("select_axis", a, ax, i)
        assert np.array_equal(want, np.take(a, i, axis=ax))
        assert _gpu_op(4, a, [ax, i], want.shape).tobytes() == want.tobytes()
    b = _rand((3, 2, 4, 2, 5), dtype, rng)
    want = O.tensor_op("trace_pair", b, 1, 3)
    assert _gpu_op(2, b, [1, 3], want.shape).tobytes() == want.tobytes()
    want = O.tensor_op("diag_pair", b, 1, 3)
    assert np.array_equal(want, np.moveaxis(np.diagonal(b, axis1=1, axis2=3), -1, 1))
    assert _gpu_op(3, b, [1, 3], want.shape).tobytes() == want.tobytes()


@pytest.mark.parametrize("dtype", DTYPES)
def test_max_magnitude(cuda, dtype):
    """Op 5: max over the elements of max(|re|, |im|), the value the strip's power-of-
    two rescales and the FP16 operand scales start from: exact against NumPy over
    sizes around the 16-byte vector loads and the grid, 16-byte aligned and
    unaligned starts, the max planted first and last, NaN components never winning."""
    import torch
    from arctn_amd import _lib

    L = _lib.lib()
    code = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.complex64): 2,
            np.dtype(np.complex128): 3}[np.dtype(dtype)]
    rng = np.random.default_rng(5)
    out = torch.empty(1, dtype=torch.float64, device="cuda")

    def gpu(t):
        sh = (C.c_size_t * 1)(t.numel())
        _lib.check(L.atn_tensor_op(code, 5, 1, sh, None, 0, t.data_ptr(), out.data_ptr(), None))
        torch.cuda.synchronize()
        return float(out.cpu())

    def want(a):
        r = a.view(a.real.dtype) if np.iscomplexobj(a) else a
        return float(np.nanmax(np.abs(r))) if r.size else 0.0

    for n in (1, 2, 3, 4, 5, 7, 8, 17, 1000, 4099, (1 << 20) + 3, (1 << 24) + 1):
        a = _rand((n + 1,), dtype, rng)
        for where in (0, n // 2, n):
            b = a.copy()
            b[where] = b[where] * 64
            t = torch.from_numpy(b).cuda()
            for lo in (0, 1):  # element 1 on: not 16-byte aligned
                assert gpu(t[lo:]) == want(b[lo:]), (n, where, lo)
    assert gpu(torch.zeros(33, dtype=torch.from_numpy(np.zeros(1, dtype)).dtype, device="cuda")) == 0.0
    a = _rand((4097,), dtype, rng)
    a[100] = np.nan
    a[4000] = 1000.0
    assert gpu(torch.from_numpy(a).cuda()) == 1000.0


GEMM_SHAPES = [(1, 1, 1, 1), (1, 3, 4, 5), (2, 17, 33, 9), (1, 128, 64, 128), (3, 130, 70, 129), (1, 256, 256, 256),
               (1, 1000, 48, 300), (1, 16, 4, 4096), (4, 64, 512, 64), (1, 8, 256, 1000), (2, 1000, 16, 8),
               (1, 3, 5, 100), (1, 200, 7, 3), (1, 32, 128, 2000),
               (1, 1, 100000, 1), (3, 2, 5000, 3), (1, 5, 70000, 7), (2, 4, 300, 16), (1, 3, 40000, 64),
               (2, 100, 30000, 40)]
# Unit roundoff of each arithmetic: FP32 / FP64 products; split products keep about
# 16 (BF16 hi + lo) or 22 (FP16 hi + lo) significant bits of every operand.
U32, U64, U_BF16X3, U_FP16X2 = 2.0 ** -24, 2.0 ** -53, 2.0 ** -16, 2.0 ** -21
NATIVE_KERNELS = (0, 1, 2, 3, 5, 6, 7)
# (dtype, precision code, kernel, unit roundoff): only the pairs that apply - split
# precision runs on the MFMA kernels (1, 4 pre-split, 8 A pre-split), 3M on the MFMA
# kernel for complex64, FP64 MFMA tiles have at most 64 rows (kernel 7 is 6 there)
GEMM_CASES = ([(dt, 0, k, U32) for dt in (np.float32, np.complex64) for k in NATIVE_KERNELS]
              + [(dt, 0, k, U64) for dt in (np.float64, np.complex128) for k in NATIVE_KERNELS if k != 7]
              + [(dt, 2, k, U_BF16X3) for dt in (np.float32, np.complex64) for k in (1, 4, 8)]
              + [(dt, 3, k, U_FP16X2) for dt in (np.float32, np.complex64) for k in (1, 4, 8)]
              + [(np.complex64, 1, 1, U32)]
              # split precisions with 3M complex products (kernel 8 with m > 32 runs the
              # 3M split kernel; the other kernels and float32 run the 2x2 form)
              + [(dt, p, k, u) for p, u in ((4, U_BF16X3), (5, U_FP16X2)) for dt in (np.float32, np.complex64)
                 for k in ((1, 4, 8) if dt == np.complex64 else (8,))])


def _magnitude_bound(a, b):
    """sum_p |a_ip| |b_pj| with |z| taken as |re| + |im|: bounds every real product
    that any of the 4M, 3M or split forms accumulates (|ar + ai| <= |ar| + |ai|)."""
    mag = lambda x: (np.abs(x.real) + np.abs(x.imag)).astype(np.float64)
    return np.matmul(mag(a), mag(b))


@pytest.mark.parametrize("case", range(len(GEMM_CASES)), ids=lambda c: "%s-p%d-k%d" % (
    np.dtype(GEMM_CASES[c][0]).name, GEMM_CASES[c][1], GEMM_CASES[c][2]))
def test_gemm_against_numpy(cuda, case):
    """Every GEMM kernel and arithmetic against complex128 NumPy products, each output
    element within 16 * u * sum_p |a_ip||b_pj| for the arithmetic's unit roundoff u
    (a bound independent of k: dropping a k tile, a term or a sign fails it)."""
    import torch
    from arctn_amd import _lib

    dtype, prec, kernel, u = GEMM_CASES[case]
    code = {np.float32: 0, np.float64: 1, np.complex64: 2, np.complex128: 3}[dtype]
    rng = np.random.default_rng(5)
    for bt, m, k, n in GEMM_SHAPES:
        a = _rand((bt, m, k), dtype, rng)
        b = _rand((bt, k, n), dtype, rng)
        want = np.matmul(a.astype(np.complex128), b.astype(np.complex128))
        bound = 16 * u * _magnitude_bound(a, b)
        ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
        tc = torch.empty((bt, m, n), dtype=ta.dtype, device="cuda")
        _lib.check(_lib.lib().atn_gemm(code, prec, kernel, bt, m, k, n, ta.data_ptr(), m * k, tb.data_ptr(), k * n,
                                       tc.data_ptr(), m * n, None))
        got = tc.cpu().numpy()
        err = np.abs(got - want)
        assert np.all(err <= bound), ((bt, m, k, n), float((err / bound).max()))


@pytest.mark.parametrize("dtype", [np.float32, np.complex64])
def test_gemm_broadcast_batch(cuda, dtype):
    """Stride-0 batch operand (the shared operand of a slice batch)."""
    import torch
    from arctn_amd import _lib

    code = {np.float32: 0, np.complex64: 2}[dtype]
    rng = np.random.default_rng(9)
    bt, m, k, n = 5, 96, 40, 72
    a = _rand((m, k), dtype, rng)
    b = _rand((bt, k, n), dtype, rng)
    want = np.matmul(a.astype(np.complex128)[None], b.astype(np.complex128))
    for kernel, prec in ((0, 0), (1, 0), (3, 0), (1, 2), (1, 3), (4, 2), (4, 3), (8, 2), (8, 3), (8, 4), (8, 5)):
        ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
        tc = torch.empty((bt, m, n), dtype=ta.dtype, device="cuda")
        _lib.check(_lib.lib().atn_gemm(code, prec, kernel, bt, m, k, n, ta.data_ptr(), 0, tb.data_ptr(), k * n,
                                       tc.data_ptr(), m * n, None))
        err = np.abs(tc.cpu().numpy() - want).max() / np.abs(want).max()
        assert err < 1e-5, (kernel, prec, err)
