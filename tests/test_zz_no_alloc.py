"""atn_gemm_noalloc, the allocation-free GEMM entry point, against atn_gemm.

Both entry points choose the same kernels with the same scratch sizes, so the
results must be bit-identical; every case runs twice in a row, the second time
on a smaller problem (the shared device scratch array keeps the first call's
contents). The kernels cover every scratch user: split-K partial sums, FP32
MFMA with k ranges, pre-split 16-bit operand halves (with and without k split
over workgroups) and FP16X2 operand maxima.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CODE = {np.float32: 0, np.float64: 1, np.complex64: 2, np.complex128: 3}
# (dtype, precision code, kernel code, first shape, second (smaller) shape); shape = (batch, m, k, n)
CASES = [
    (np.complex64, 0, 3, (1, 5, 70000, 7), (1, 3, 40000, 4)),
    (np.float32, 0, 3, (3, 2, 5000, 3), (2, 2, 300, 2)),
    (np.float32, 0, 1, (1, 64, 20000, 64), (1, 32, 3000, 48)),
    (np.complex128, 0, -1, (2, 100, 3000, 40), (1, 17, 33, 9)),
    (np.complex64, 3, 4, (2, 300, 256, 200), (1, 100, 64, 50)),
    (np.complex64, 2, 4, (2, 300, 256, 200), (1, 100, 64, 50)),
    (np.float32, 3, 8, (2, 500, 512, 300), (1, 130, 70, 129)),
    (np.complex64, 3, 1, (1, 256, 256, 256), (1, 64, 40, 72)),
    (np.complex64, 3, 4, (1, 64, 20000, 64), (1, 40, 9000, 30)),  # pre-split, k split over workgroups
]


def _rand(shape, dtype, rng):
    a = rng.standard_normal(shape)
    if np.issubdtype(np.dtype(dtype), np.complexfloating):
        a = a + 1j * rng.standard_normal(shape)
    return np.ascontiguousarray(a.astype(dtype))


def _run(fn, dtype, prec, kernel, a, b, shape):
    import torch
    from arctn_amd import _lib

    bt, m, k, n = shape
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    tc = torch.full((bt, m, n), float("nan"), dtype=ta.dtype, device="cuda")
    _lib.check(fn(CODE[dtype], prec, kernel, bt, m, k, n, ta.data_ptr(), m * k, tb.data_ptr(), k * n,
                  tc.data_ptr(), m * n, None))
    torch.cuda.synchronize()
    return tc.cpu().numpy()


@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_gemm_matches_gemm(cuda, case):
    from arctn_amd import _lib

    L = _lib.lib()
    dtype, prec, kernel, first, second = CASES[case]
    rng = np.random.default_rng(31 + case)
    for shape in (first, second):
        bt, m, k, n = shape
        print(f"noalloc case {case}: dtype {np.dtype(dtype).name} precision {prec} kernel {kernel} shape {shape}",
              flush=True)
        a, b = _rand((bt, m, k), dtype, rng), _rand((bt, k, n), dtype, rng)
        want = _run(L.atn_gemm, dtype, prec, kernel, a, b, shape)
        got = _run(L.atn_gemm_noalloc, dtype, prec, kernel, a, b, shape)
        assert got.tobytes() == want.tobytes(), (case, shape)
        exact = np.matmul(a.astype(np.complex128), b.astype(np.complex128))
        bound = np.matmul(np.abs(a).astype(np.float64), np.abs(b).astype(np.float64))
        assert np.all(np.abs(got - exact) <= 1e-4 * bound + 1e-30), (case, shape)


def test_noalloc_gemm_rejects_oversized_scratch(cuda):
    """A pre-split product whose 16-bit halves exceed the fixed scratch is
    refused before any launch (the buffers are real and large enough anyway)."""
    import torch
    from arctn_amd import _lib

    L = _lib.lib()
    bt, m, k, n = 1, 2048, 4096, 2048
    ta = torch.zeros((bt, m, k), dtype=torch.complex64, device="cuda")
    tb = torch.zeros((bt, k, n), dtype=torch.complex64, device="cuda")
    tc = torch.zeros((bt, m, n), dtype=torch.complex64, device="cuda")
    rc = L.atn_gemm_noalloc(2, 3, 4, bt, m, k, n, ta.data_ptr(), m * k, tb.data_ptr(), k * n, tc.data_ptr(), m * n,
                            None)
    assert rc != 0
    assert b"scratch" in L.atn_last_error()
    torch.cuda.synchronize()
