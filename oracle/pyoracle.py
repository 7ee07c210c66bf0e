"""ctypes wrapper of the oracle library (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module; the HIP executor never imports it.
"""

from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

DTYPE_CODE = {np.dtype(np.float32): 0, np.dtype(np.float64): 1,
              np.dtype(np.complex64): 2, np.dtype(np.complex128): 3}
MODES = {"compiled": 0, "stripped": 1, "direct": 2, "sliced": 3, "naive": 4}
OPS = {"permute": 0, "sum_axis": 1, "trace_pair": 2, "diag_pair": 3, "select_axis": 4}


class _Net(C.Structure):
    _fields_ = [
        ("n_inputs", C.c_size_t), ("input_lens", C.POINTER(C.c_size_t)), ("input_legs", C.POINTER(C.c_uint32)),
        ("n_output", C.c_size_t), ("output", C.POINTER(C.c_uint32)),
        ("n_dims", C.c_size_t), ("dim_legs", C.POINTER(C.c_uint32)), ("dim_sizes", C.POINTER(C.c_size_t)),
        ("n_steps", C.c_size_t), ("path", C.POINTER(C.c_size_t)),
        ("n_sliced", C.c_size_t), ("sliced", C.POINTER(C.c_uint32)),
    ]


class OracleError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OracleError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_last_error.restype = C.c_char_p
        L.orc_slice_parallel_chunks.restype = C.c_size_t
        L.orc_slice_parallel_chunks.argtypes = [C.c_size_t, C.c_size_t]
        L.orc_compile_describe.argtypes = [C.POINTER(_Net), C.POINTER(C.c_void_p)]
        L.orc_simulate_path.argtypes = [C.POINTER(_Net), C.POINTER(C.c_double)]
        L.orc_contract.argtypes = [C.POINTER(_Net), C.c_int, C.c_int, C.POINTER(C.c_void_p), C.c_void_p,
                                   C.POINTER(C.c_double), C.c_int]
        L.orc_contract_slice_range.argtypes = [C.POINTER(_Net), C.c_int, C.c_size_t, C.c_size_t,
                                               C.POINTER(C.c_void_p), C.c_void_p, C.c_int]
        L.orc_gradient.argtypes = [C.POINTER(_Net), C.c_int, C.POINTER(C.c_void_p), C.c_void_p, C.POINTER(C.c_void_p)]
        L.orc_tensor_op.argtypes = [C.c_int, C.c_int, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                    C.c_size_t, C.c_void_p, C.c_void_p]
        L.orc_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


class NetArgs:
    """Keeps the flat arrays alive while the struct points into them."""

    def __init__(self, net, path=(), sliced=()):
        self.lens = np.array([len(t) for t in net["inputs"]], dtype=np.uintp)
        self.legs = np.array([l for t in net["inputs"] for l in t], dtype=np.uint32)
        self.out = np.array(list(net["output"]), dtype=np.uint32)
        items = sorted(net["size_dict"].items())
        self.dl = np.array([k for k, _ in items], dtype=np.uint32)
        self.ds = np.array([v for _, v in items], dtype=np.uintp)
        self.path = np.array([x for p in path for x in p], dtype=np.uintp)
        self.sl = np.array(list(sliced), dtype=np.uint32)
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        self.s = _Net(len(self.lens), P(self.lens, C.c_size_t), P(self.legs, C.c_uint32),
                      len(self.out), P(self.out, C.c_uint32), len(self.dl), P(self.dl, C.c_uint32),
                      P(self.ds, C.c_size_t), len(path), P(self.path, C.c_size_t),
                      len(self.sl), P(self.sl, C.c_uint32))


def _ptrs(arrays):
    return (C.c_void_p * max(len(arrays), 1))(*[a.ctypes.data for a in arrays])


def _prep(arrays):
    # np.ascontiguousarray would turn a rank-0 array into shape (1,)
    arrays = [np.array(a, order="C", copy=True) if np.ndim(a) == 0 else np.ascontiguousarray(a) for a in arrays]
    dt = arrays[0].dtype
    for a in arrays:
        if a.dtype != dt:
            raise OracleError("mixed dtypes")
    return arrays, dt


def contract(net, arrays, path=(), mode="compiled", sliced=(), threads=1):
    """Returns the output array (and the log10 exponent for mode='stripped')."""
    arrays, dt = _prep(arrays)
    na = NetArgs(net, path, sliced)
    out = np.zeros([net["size_dict"][l] for l in net["output"]], dtype=dt)
    e = C.c_double(0.0)
    _check(lib().orc_contract(C.byref(na.s), DTYPE_CODE[dt], MODES[mode], _ptrs(arrays),
                              out.ctypes.data, C.byref(e), threads))
    return (out, e.value) if mode == "stripped" else out


def contract_slice_range(net, arrays, path, sliced, lo, hi, threads=1):
    """Sum of internal slices [lo, hi) of a sliced contraction."""
    arrays, dt = _prep(arrays)
    na = NetArgs(net, path, sliced)
    out = np.zeros([net["size_dict"][l] for l in net["output"]], dtype=dt)
    _check(lib().orc_contract_slice_range(C.byref(na.s), DTYPE_CODE[dt], lo, hi, _ptrs(arrays), out.ctypes.data,
                                          threads))
    return out


def gradient(net, arrays, path, grad_out):
    arrays, dt = _prep(arrays)
    na = NetArgs(net, path)
    g = np.ascontiguousarray(np.asarray(grad_out, dtype=dt))
    grads = [np.zeros_like(a) for a in arrays]
    _check(lib().orc_gradient(C.byref(na.s), DTYPE_CODE[dt], _ptrs(arrays), g.ctypes.data, _ptrs(grads)))
    return grads


def describe(net, path):
    na = NetArgs(net, path)
    p = C.c_void_p()
    _check(lib().orc_compile_describe(C.byref(na.s), C.byref(p)))
    try:
        return json.loads(C.string_at(p.value).decode())
    finally:
        lib().orc_free(p)


def simulate(net, path):
    na = NetArgs(net, path)
    v = (C.c_double * 6)()
    _check(lib().orc_simulate_path(C.byref(na.s), v))
    keys = ("log10_flops", "log2_max_size", "log2_max_contraction_size", "log2_total_size",
            "log2_read_write", "log2_peak_size")
    return dict(zip(keys, list(v)))


def slice_parallel_chunks(n_slices, out_numel):
    return int(lib().orc_slice_parallel_chunks(n_slices, out_numel))


def tensor_op(op, a, *arg):
    a = np.ascontiguousarray(a)
    shape = list(a.shape)
    args = list(arg[0]) if op == "permute" else list(arg)
    if op == "permute":
        oshape = [shape[x] for x in args]
    elif op == "sum_axis":
        oshape = shape[: args[0]] + shape[args[0] + 1:]
    elif op == "trace_pair":
        oshape = [s for i, s in enumerate(shape) if i not in args]
    elif op == "diag_pair":
        oshape = [s for i, s in enumerate(shape) if i != args[1]]
    else:
        oshape = shape[: args[0]] + shape[args[0] + 1:]
    out = np.zeros(oshape, dtype=a.dtype)
    sh = (C.c_size_t * max(len(shape), 1))(*shape)
    ar = (C.c_size_t * max(len(args), 1))(*args)
    _check(lib().orc_tensor_op(DTYPE_CODE[a.dtype], OPS[op], len(shape), sh, ar, len(args), a.ctypes.data,
                               out.ctypes.data))
    return out
