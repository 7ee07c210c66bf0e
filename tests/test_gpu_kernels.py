"""Kernel-level parity of the HIP data-movement and GEMM kernels (GPU).

Data movement (permute, sum_axis, trace_pair, diag_pair, select_axis) must be
bit-identical to the oracle's restatement of src/tensor.rs:453-669. GEMMs are
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


GEMM_SHAPES = [(1, 1, 1, 1), (1, 3, 4, 5), (2, 17, 33, 9), (1, 128, 64, 128), (3, 130, 70, 129), (1, 256, 256, 256),
               (1, 1000, 48, 300), (1, 16, 4, 4096), (4, 64, 512, 64), (1, 8, 256, 1000), (2, 1000, 16, 8),
               (1, 3, 5, 100), (1, 200, 7, 3), (1, 32, 128, 2000),
               (1, 1, 100000, 1), (3, 2, 5000, 3), (1, 5, 70000, 7), (2, 4, 300, 16), (1, 3, 40000, 64),
               (2, 100, 30000, 40)]
# (dtype, precision code, per-product relative error bound)
PREC_CASES = [(np.float32, 0, 1e-6), (np.float64, 0, 1e-14), (np.complex64, 0, 1e-6), (np.complex128, 0, 1e-14),
              (np.float32, 2, 1e-5), (np.complex64, 2, 1e-5), (np.float32, 3, 1e-6), (np.complex64, 3, 1e-6)]


@pytest.mark.parametrize("case", range(len(PREC_CASES)))
@pytest.mark.parametrize("kernel", [0, 1, 2, 3, 4, 5, 6, 7, 8])
def test_gemm_against_numpy(cuda, case, kernel):
    """Simple (0), MFMA (1, native or split precision), skinny (2), split-K (3)
    pre-split MFMA (4) and A-pre-split MFMA (8) GEMMs vs complex128 NumPy."""
    import torch
    from arctn_amd import _lib

    dtype, prec, eps = PREC_CASES[case]
    if prec and kernel not in (1, 4, 8):
        pytest.skip("split precision applies to the MFMA kernels")
    if kernel == 7 and dtype in (np.float64, np.complex128):
        pytest.skip("FP64 MFMA tiles have at most 64 rows (7 runs as 6)")
    code = {np.float32: 0, np.float64: 1, np.complex64: 2, np.complex128: 3}[dtype]
    rng = np.random.default_rng(5)
    for bt, m, k, n in GEMM_SHAPES:
        a = _rand((bt, m, k), dtype, rng)
        b = _rand((bt, k, n), dtype, rng)
        want = np.matmul(a.astype(np.complex128), b.astype(np.complex128))
        ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
        tc = torch.empty((bt, m, n), dtype=ta.dtype, device="cuda")
        _lib.check(_lib.lib().atn_gemm(code, prec, kernel, bt, m, k, n, ta.data_ptr(), m * k, tb.data_ptr(), k * n,
                                       tc.data_ptr(), m * n, None))
        got = tc.cpu().numpy()
        scale = np.sqrt(k) * max(1.0, np.abs(want).max())
        err = np.abs(got - want).max() / scale
        assert err < 20 * eps, (bt, m, k, n, err)


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
    for kernel, prec in ((0, 0), (1, 0), (3, 0), (1, 2), (1, 3), (4, 2), (4, 3), (8, 2), (8, 3)):
        ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
        tc = torch.empty((bt, m, n), dtype=ta.dtype, device="cuda")
        _lib.check(_lib.lib().atn_gemm(code, prec, kernel, bt, m, k, n, ta.data_ptr(), 0, tb.data_ptr(), k * n,
                                       tc.data_ptr(), m * n, None))
        err = np.abs(tc.cpu().numpy() - want).max() / np.abs(want).max()
        assert err < 1e-5, (kernel, prec, err)
