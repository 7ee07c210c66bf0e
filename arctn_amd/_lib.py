"""ctypes binding of the C ABI in include/arctn_hip.h (libarctn_hip.so).

This is the Python counterpart of the binding a host language would add over
the C ABI. The library is loaded from the package directory; importing any
GPU entry point without it raises instead of falling back to the CPU.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libarctn_hip.so")

ATN_F32, ATN_F64, ATN_C64, ATN_C128 = 0, 1, 2, 3
PREC = {"native": 0, "fp32": 0, "c3m": 1, "bf16x3": 2, "fp16x2": 3, "bf16x3_3m": 4, "fp16x2_3m": 5}


class NetworkDesc(C.Structure):
    _fields_ = [
        ("n_inputs", C.c_size_t), ("input_lens", C.POINTER(C.c_size_t)), ("input_legs", C.POINTER(C.c_uint32)),
        ("n_output", C.c_size_t), ("output", C.POINTER(C.c_uint32)),
        ("n_dims", C.c_size_t), ("dim_legs", C.POINTER(C.c_uint32)), ("dim_sizes", C.POINTER(C.c_size_t)),
        ("n_steps", C.c_size_t), ("ssa_path", C.POINTER(C.c_size_t)),
        ("n_sliced", C.c_size_t), ("sliced_legs", C.POINTER(C.c_uint32)),
    ]


class PlanInfo(C.Structure):
    _fields_ = [("n_inputs", C.c_size_t), ("n_steps", C.c_size_t), ("n_slices", C.c_size_t),
                ("n_blocks", C.c_size_t), ("n_varying_steps", C.c_size_t), ("log10_flops", C.c_double),
                ("log2_peak_size", C.c_double), ("macs_per_slice", C.c_double), ("mults_per_slice", C.c_double)]


class ExecOptions(C.Structure):
    _fields_ = [("dtype", C.c_int), ("precision", C.c_int), ("batched", C.c_int), ("use_graph", C.c_int),
                ("autotune", C.c_int), ("strip", C.c_int), ("max_batch", C.c_size_t),
                ("workspace_limit", C.c_size_t), ("slice_begin", C.c_size_t), ("slice_end", C.c_size_t),
                ("permute_b", C.c_int), ("gemm_kernel", C.c_int)]


class Report(C.Structure):
    _fields_ = [("elapsed_ms", C.c_double), ("flops", C.c_double), ("tflops", C.c_double),
                ("slices_per_s", C.c_double), ("log10_exponent", C.c_double), ("max_abs_result", C.c_double),
                ("n_slices", C.c_size_t), ("batch", C.c_size_t), ("n_launches", C.c_size_t),
                ("workspace_bytes", C.c_size_t)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("calls", C.c_size_t), ("launches", C.c_size_t), ("ms", C.c_double),
                ("flops", C.c_double), ("bytes", C.c_double)]


# name -> (restype, argtypes)
_VP = C.c_void_p
_SIGS = {
    "atn_abi_version": (C.c_int, []),
    "atn_last_error": (C.c_char_p, []),
    "atn_free": (None, [_VP]),
    "atn_plan_create": (C.c_int, [C.POINTER(NetworkDesc), C.POINTER(_VP)]),
    "atn_plan_destroy": (None, [_VP]),
    "atn_plan_stats": (C.c_int, [_VP, C.POINTER(PlanInfo)]),
    "atn_plan_describe": (C.c_int, [_VP, C.POINTER(_VP)]),
    "atn_executor_create": (C.c_int, [_VP, C.POINTER(ExecOptions), C.POINTER(_VP)]),
    "atn_executor_destroy": (None, [_VP]),
    "atn_executor_run": (C.c_int, [_VP, C.POINTER(_VP), _VP, _VP, C.POINTER(Report)]),
    "atn_executor_profile": (C.c_int, [_VP, C.POINTER(_VP), _VP, _VP, C.POINTER(KernelStat), C.c_size_t,
                                       C.POINTER(C.c_size_t)]),
    "atn_execute": (C.c_int, [_VP, C.POINTER(ExecOptions), C.POINTER(_VP), _VP, _VP, C.POINTER(Report)]),
    "atn_prep_forward": (C.c_int, [_VP, C.c_size_t, C.c_int, _VP, _VP, _VP]),
    "atn_prep_adjoint": (C.c_int, [_VP, C.c_size_t, C.c_int, _VP, _VP, _VP]),
    "atn_step_forward": (C.c_int, [_VP, C.c_size_t, C.c_int, C.c_int, _VP, _VP, _VP, _VP]),
    "atn_step_adjoint": (C.c_int, [_VP, C.c_size_t, C.c_int, _VP, _VP, _VP, _VP, _VP, _VP]),
    "atn_finalize_forward": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP]),
    "atn_finalize_adjoint": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP]),
    "atn_adjoint_create": (C.c_int, [_VP, C.c_int, C.c_int, C.c_size_t, C.POINTER(_VP)]),
    "atn_adjoint_destroy": (None, [_VP]),
    "atn_adjoint_forward": (C.c_int, [_VP, C.POINTER(_VP), _VP, _VP]),
    "atn_adjoint_backward": (C.c_int, [_VP, C.POINTER(_VP), _VP, C.POINTER(_VP), _VP]),
    "atn_tensor_op": (C.c_int, [C.c_int, C.c_int, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                C.c_size_t, _VP, _VP, _VP]),
    "atn_gemm": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, _VP,
                           C.c_size_t, _VP, C.c_size_t, _VP, C.c_size_t, _VP]),
    "atn_gemm_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                   _VP, C.c_size_t, _VP, C.c_size_t, _VP, C.c_size_t, _VP]),
    "atn_grouped_execute": (C.c_int, [C.POINTER(_VP), C.c_size_t, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(C.POINTER(_VP)), C.POINTER(_VP), C.POINTER(C.c_double), _VP,
                                      C.POINTER(Report)]),
    "atn_grouped_create": (C.c_int, [C.POINTER(_VP), C.c_size_t, C.c_int, C.c_int, C.c_int, C.POINTER(_VP)]),
    "atn_grouped_destroy": (None, [_VP]),
    "atn_grouped_run": (C.c_int, [_VP, C.POINTER(C.POINTER(_VP)), C.POINTER(_VP), C.POINTER(C.c_double), _VP,
                                  C.POINTER(Report)]),
}

_lib = None


class HipError(RuntimeError):
    pass


def lib():
    """Load libarctn_hip.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HipError(f"{LIB_PATH} is missing; build it with `python -m arctn_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.atn_abi_version() != 1:
            raise HipError("libarctn_hip.so ABI version mismatch")
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise ValueError(lib().atn_last_error().decode(errors="replace"))


class _Arrays:
    """Flat arrays backing a NetworkDesc (kept alive with it)."""

    def __init__(self, inputs, output, size_dict, ssa_path, sliced):
        import numpy as np

        self.lens = np.array([len(t) for t in inputs], dtype=np.uintp)
        self.legs = np.array([l for t in inputs for l in t], dtype=np.uint32)
        self.out = np.array(list(output), dtype=np.uint32)
        items = sorted(size_dict.items())
        self.dl = np.array([k for k, _ in items], dtype=np.uint32)
        self.ds = np.array([v for _, v in items], dtype=np.uintp)
        self.path = np.array([x for p in ssa_path for x in p], dtype=np.uintp)
        self.sl = np.array(list(sliced), dtype=np.uint32)
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        self.desc = NetworkDesc(len(self.lens), P(self.lens, C.c_size_t), P(self.legs, C.c_uint32),
                                len(self.out), P(self.out, C.c_uint32), len(self.dl), P(self.dl, C.c_uint32),
                                P(self.ds, C.c_size_t), len(ssa_path), P(self.path, C.c_size_t),
                                len(self.sl), P(self.sl, C.c_uint32))


def plan_create(inputs, output, size_dict, ssa_path, sliced=()):
    arr = _Arrays(inputs, output, size_dict, ssa_path, sliced)
    h = C.c_void_p()
    check(lib().atn_plan_create(C.byref(arr.desc), C.byref(h)))
    return h


def pointer_array(ptrs):
    return (C.c_void_p * max(len(ptrs), 1))(*ptrs)
