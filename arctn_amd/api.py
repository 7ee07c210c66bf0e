"""Python surface of the HIP executor, mirroring the reference's binding.

* ``CompiledContraction.compile(inputs, output, size_dict, *, ssa_path)``,
  ``.execute(arrays)``, ``.execute_stripped(arrays)``, ``.stats()`` mirror
  ``arctn_py.CompiledContraction`` (pybind/src/lib.rs:361-454).
* ``contract_sliced(inputs, output, size_dict, arrays, sliced_legs, path=None,
  ssa_path=None)`` mirrors ``arctn_py.contract_sliced`` (pybind/src/lib.rs:692-765).
* ``arctn_contract(..., backend="hip")`` mirrors the planning + execution entry
  point (pybind/python/arctn/__init__.py:903-1072) with the CPU planner of
  arctn_amd.planner and the GPU executor.

Arrays may be NumPy arrays (copied to the GPU; the result comes back as a
NumPy array, like the reference) or torch tensors on the GPU (used in place;
the result is a torch tensor on the same device). Errors are ValueError with
the reference's message prefixes. The GPU library is mandatory: without it
every entry point raises.
"""

from __future__ import annotations

import itertools
import math
import operator
import time

import numpy as np

from . import _lib
from . import planner as _planner

DTYPE_CODE = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.complex64): 2,
              np.dtype(np.complex128): 3}
DTYPE_NAME = {0: "float32", 1: "float64", 2: "complex64", 3: "complex128"}


def _torch():
    import torch

    return torch


def _torch_dtype(code):
    torch = _torch()
    return {0: torch.float32, 1: torch.float64, 2: torch.complex64, 3: torch.complex128}[code]


def _code_of(arr):
    torch = _torch() if not isinstance(arr, np.ndarray) else None
    if torch is not None and isinstance(arr, torch.Tensor):
        m = {torch.float32: 0, torch.float64: 1, torch.complex64: 2, torch.complex128: 3}
        if arr.dtype not in m:
            return None
        return m[arr.dtype]
    return DTYPE_CODE.get(np.asarray(arr).dtype)


def _build_net(inputs, output, size_dict):
    """pybind/src/lib.rs:609-637: legs are integers; unused size_dict entries rejected."""
    inputs = [[int(l) for l in t] for t in inputs]
    output = [int(l) for l in output]
    size_dict = {int(k): int(v) for k, v in dict(size_dict).items()}
    if not inputs:
        raise ValueError("inputs 为空 (inputs is empty)")
    used = {l for t in inputs for l in t} | set(output)
    extra = [k for k in size_dict if k not in used]
    if extra:
        raise ValueError(f"size_dict 含未在 inputs/output 使用的腿 {extra[0]} (unused size_dict leg)")
    return inputs, output, size_dict


def _device_inputs(net_inputs, size_dict, arrays, context):
    """Full dtype/shape preflight before any copy (pybind/src/lib.rs:241-330).

    Returns (device tensors, dtype code, return_numpy)."""
    torch = _torch()
    arrays = list(arrays)
    if len(arrays) != len(net_inputs):
        raise ValueError(f"{context}: arrays 个数 {len(arrays)} 与 inputs 个数 {len(net_inputs)} 不符")
    code = _code_of(arrays[0]) if arrays else None
    if code is None:
        raise ValueError(f"{context}: arrays 的 dtype 不支持；支持 float32 / float64 / complex64 / complex128")
    numpy_in = False
    for i, (a, legs) in enumerate(zip(arrays, net_inputs)):
        c = _code_of(a)
        if c != code:
            raise ValueError(f"{context}: arrays[{i}] 的 dtype 与首张量不一致（预期 {DTYPE_NAME[code]}）")
        shape = tuple(a.shape)
        want = tuple(size_dict[l] for l in legs)
        if shape != want:
            raise ValueError(f"{context}: arrays[{i}] shape {list(shape)} 与网络预期 {list(want)} 不一致（输入数组布局非法）")
        numpy_in = numpy_in or not isinstance(a, torch.Tensor)
    if not torch.cuda.is_available():
        raise _lib.HipError("no GPU device is available for the HIP executor")
    dev = None
    for a in arrays:
        if isinstance(a, torch.Tensor) and a.is_cuda:
            dev = a.device
            break
    dev = dev or torch.device("cuda", torch.cuda.current_device())
    out = []
    for a in arrays:
        if isinstance(a, torch.Tensor):
            # materialise lazy conjugate / negative views before handing raw pointers over
            t = a.resolve_conj().resolve_neg().to(dev)
        else:
            a = np.asarray(a)
            # np.ascontiguousarray would turn a rank-0 array into shape (1,)
            a = np.array(a, order="C", copy=True) if a.ndim == 0 else np.ascontiguousarray(a)
            t = torch.from_numpy(a).to(dev)
        out.append(t.contiguous())
    return out, code, numpy_in


class _Executor:
    def __init__(self, plan_handle, opts):
        self.h = _lib.C.c_void_p()
        _lib.check(_lib.lib().atn_executor_create(plan_handle, _lib.C.byref(opts), _lib.C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and _lib._lib is not None:
            _lib.lib().atn_executor_destroy(self.h)
            self.h = None


class Plan:
    """A validated plan (network + SSA path + sliced legs) and its executors."""

    def __init__(self, inputs, output, size_dict, ssa_path, sliced_legs=()):
        self.inputs, self.output, self.size_dict = _build_net(inputs, output, size_dict)
        self.ssa_path = [tuple(int(x) for x in p) for p in ssa_path]
        self.sliced_legs = [int(l) for l in sliced_legs]
        self._h = _lib.plan_create(self.inputs, self.output, self.size_dict, self.ssa_path, self.sliced_legs)
        self._executors = {}

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            self._executors.clear()
            _lib.lib().atn_plan_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def release(self):
        """Destroy cached executors (frees their device workspaces)."""
        self._executors.clear()

    def info(self):
        i = _lib.PlanInfo()
        _lib.check(_lib.lib().atn_plan_stats(self._h, _lib.C.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_}

    def describe(self):
        import json

        p = _lib.C.c_void_p()
        _lib.check(_lib.lib().atn_plan_describe(self._h, _lib.C.byref(p)))
        try:
            return json.loads(_lib.C.string_at(p.value).decode())
        finally:
            _lib.lib().atn_free(p)

    def executor(self, dtype_code, precision="native", batched=False, use_graph=False, autotune=False, strip=False,
                 max_batch=0, workspace_limit=0, slice_range=None, gather=True, gemm_kernel=None):
        """``slice_range=(begin, end)``: the batched form over global slices
        [begin, end) only (their partial sum), e.g. one device's share.
        ``gather=False`` permutes every GEMM B operand into [k, n] order
        instead of reading it through gather tables. ``gemm_kernel=k`` runs
        every GEMM step on kernel code k of ``atn_gemm`` (None: chosen per step)."""
        if precision not in _lib.PREC:
            raise ValueError(f"unknown precision {precision!r}; use one of {sorted(_lib.PREC)}")
        sb, se = (0, 0) if slice_range is None else (int(slice_range[0]), int(slice_range[1]))
        key = (dtype_code, precision, bool(batched), bool(use_graph), bool(autotune), bool(strip), max_batch,
               workspace_limit, sb, se, bool(gather), gemm_kernel)
        ex = self._executors.get(key)
        if ex is None:
            o = _lib.ExecOptions(dtype_code, _lib.PREC[precision], int(batched), int(use_graph), int(autotune),
                                 int(strip), max_batch, workspace_limit, sb, se, int(not gather),
                                 0 if gemm_kernel is None else int(gemm_kernel) + 1)
            ex = _Executor(self._h, o)
            self._executors[key] = ex
        return ex

    def check_device_args(self, dev_inputs, dtype_code, out=None):
        """Count, dtype, shape, contiguity and device of the device tensors (and of a
        caller-supplied output) before any pointer reaches the C ABI; the reference
        checks every operand before its GEMMs (src/compiled.rs:386-398)."""
        torch = _torch()
        if dtype_code not in DTYPE_NAME:
            raise ValueError(f"unsupported dtype code {dtype_code!r}")
        tdt = _torch_dtype(dtype_code)
        if len(dev_inputs) != len(self.inputs):
            raise ValueError(f"张量个数 {len(dev_inputs)} 与计划输入数 {len(self.inputs)} 不符")
        dev = None
        for i, (t, legs) in enumerate(zip(dev_inputs, self.inputs)):
            if not isinstance(t, torch.Tensor) or not t.is_cuda:
                raise ValueError(f"张量 {i} is not a GPU tensor")
            dev = dev or t.device
            if t.device != dev:
                raise ValueError(f"张量 {i} is on {t.device}, tensor 0 on {dev}")
            if t.dtype != tdt:
                raise ValueError(f"张量 {i} has dtype {t.dtype}, the run's dtype is {DTYPE_NAME[dtype_code]}")
            want = tuple(self.size_dict[l] for l in legs)
            if tuple(t.shape) != want:
                raise ValueError(f"张量 {i} shape {list(t.shape)} 与网络预期 {list(want)} 不一致")
            if not t.is_contiguous():
                raise ValueError(f"张量 {i} is not contiguous")
        if out is not None:
            want = tuple(self.size_dict[l] for l in self.output)
            if not isinstance(out, torch.Tensor) or not out.is_cuda or out.device != dev:
                raise ValueError("out must be a GPU tensor on the inputs' device")
            if out.dtype != tdt or tuple(out.shape) != want or not out.is_contiguous():
                raise ValueError(f"out must be a contiguous {DTYPE_NAME[dtype_code]} tensor of shape {list(want)}")

    def run(self, dev_inputs, dtype_code, out=None, report=False, stream=None, **opts):
        """Execute on device tensors; returns (output tensor, report dict or None)."""
        torch = _torch()
        self.check_device_args(dev_inputs, dtype_code, out)
        ex = self.executor(dtype_code, **opts)
        if out is None:
            shape = [self.size_dict[l] for l in self.output]
            out = torch.empty(shape, dtype=_torch_dtype(dtype_code), device=dev_inputs[0].device)
        ptrs = _lib.pointer_array([t.data_ptr() for t in dev_inputs])
        s = stream if stream is not None else torch.cuda.current_stream(dev_inputs[0].device)
        rep = _lib.Report()
        _lib.check(_lib.lib().atn_executor_run(ex.h, ptrs, out.data_ptr(), s.cuda_stream,
                                               _lib.C.byref(rep) if report else None))
        return out, (rep.as_dict() if report else None)


def profile(plan, dev_inputs, dtype_code, out=None, stream=None, **opts):
    """Per-kernel-kind device time and algorithmic work of one eager execution."""
    torch = _torch()
    plan.check_device_args(dev_inputs, dtype_code, out)
    ex = plan.executor(dtype_code, **opts)
    if out is None:
        shape = [plan.size_dict[l] for l in plan.output]
        out = torch.empty(shape, dtype=_torch_dtype(dtype_code), device=dev_inputs[0].device)
    ptrs = _lib.pointer_array([t.data_ptr() for t in dev_inputs])
    s = stream if stream is not None else torch.cuda.current_stream(dev_inputs[0].device)
    stats = (_lib.KernelStat * 32)()
    n = _lib.C.c_size_t(0)
    _lib.check(_lib.lib().atn_executor_profile(ex.h, ptrs, out.data_ptr(), s.cuda_stream, stats, 32, _lib.C.byref(n)))
    return {stats[i].name.decode(): {"calls": stats[i].calls, "launches": stats[i].launches, "ms": stats[i].ms,
                                     "flops": stats[i].flops, "bytes": stats[i].bytes} for i in range(n.value)}


def _finish(out, numpy_in):
    if numpy_in:
        return out.cpu().numpy()
    return out


class CompiledContraction:
    """A reusable exact contraction compiled from an SSA path (pybind/src/lib.rs:355-454).

    ``execute`` runs the unsliced plan with its tiny steps grouped: the steps of
    one level of the contraction tree that run on the one-thread-per-element
    kernel share one launch (the run of small absorptions that opens a circuit
    plan, simplify.rs:241-310, becomes a few launches), every other step runs as
    in the per-step path, with the same kernels and summation order.
    ``batched=False`` selects the per-step path itself; keyword options select
    the precision form (``precision=``) and the graph form.
    """

    def __init__(self, plan):
        self._plan = plan

    @staticmethod
    def compile(inputs, output, size_dict, *, ssa_path):
        try:
            return CompiledContraction(Plan(inputs, output, size_dict, ssa_path))
        except ValueError as e:
            raise ValueError(f"预编译路径非法: {e}") from None

    def _prepare(self, arrays):
        p = self._plan
        arrays = list(arrays)
        if len(arrays) != len(p.inputs):
            raise ValueError(f"arrays 个数 {len(arrays)} 与编译计划输入数 {len(p.inputs)} 不符")
        return _device_inputs(p.inputs, p.size_dict, arrays, "预编译收缩失败")

    def execute(self, arrays, **opts):
        dev, code, numpy_in = self._prepare(arrays)
        opts.setdefault("batched", True)  # unsliced: one "batch", tiny steps grouped per tree level
        try:
            out, _ = self._plan.run(dev, code, **opts)
        except ValueError as e:
            raise ValueError(f"预编译收缩失败: {e}") from None
        return _finish(out, numpy_in)

    def execute_stripped(self, arrays, **opts):
        """(mantissa, log10_exponent): result = mantissa * 10**log10_exponent,
        the pair of CompiledContraction::execute_stripped (compiled.rs:376-381, 412-430).

        The last step's result is divided by its exact max magnitude (f64 |z|,
        as the reference), so the mantissa before finalize has max magnitude 1,
        and the exponent is the sum of log10 of every factor. Intermediate steps
        are rescaled by powers of two only when their max leaves [2^-32, 2^32]
        (exact, and the same value as the reference's per-step division).
        """
        dev, code, numpy_in = self._prepare(arrays)
        try:
            out, rep = self._plan.run(dev, code, report=True, strip=True, **opts)
        except ValueError as e:
            raise ValueError(f"预编译收缩失败: {e}") from None
        return _finish(out, numpy_in), rep["log10_exponent"]

    def stats(self):
        i = self._plan.info()
        return {"path_format": "ssa-v1", "n_inputs": i["n_inputs"], "n_steps": i["n_steps"],
                "log10_flops": i["log10_flops"], "log2_peak_size": i["log2_peak_size"]}

    def describe(self):
        return self._plan.describe()

    @property
    def plan(self):
        return self._plan

    def __repr__(self):
        i = self._plan.info()
        return f"CompiledContraction(path_format='ssa-v1', n_inputs={i['n_inputs']}, n_steps={i['n_steps']})"


def contract_sliced(inputs, output, size_dict, arrays, sliced_legs, path=None, ssa_path=None, **opts):
    """Execute a supplied path over the requested slices (pybind/src/lib.rs:692-765)."""
    n = len(inputs)
    arrays = list(arrays)
    if len(arrays) != n:
        raise ValueError(f"arrays 个数 {len(arrays)} 与 inputs 个数 {n} 不符")
    if path is not None and ssa_path is not None:
        raise ValueError("path 与 ssa_path 互斥，只能给一个（path=linear 口径，ssa_path=SSA 口径）")
    inputs_i, output_i, sd = _build_net(inputs, output, size_dict)
    for leg in sliced_legs:
        if int(leg) not in sd:
            raise ValueError(f"切片腿 {leg} 没有出现在网络 size_dict 中")
    dev, code, numpy_in = _device_inputs(inputs_i, sd, arrays, "切片收缩失败")
    if ssa_path is not None:
        ssa = [tuple(p) for p in ssa_path]
    elif path is not None:
        ssa = _planner.linear_to_ssa([tuple(p) for p in path], n)
    else:
        raise ValueError("必须给 path=（linear 口径）或 ssa_path=（SSA 口径）之一。")
    try:
        plan = Plan(inputs_i, output_i, sd, ssa, sliced_legs)
        out, _ = plan.run(dev, code, **opts)
    except ValueError as e:
        raise ValueError(f"切片收缩失败: {e}") from None
    return _finish(out, numpy_in)


def _positive_dim(label, raw):
    if isinstance(raw, bool):
        raise ValueError(f"腿 {label!r} 的维度必须是正整数，不能是 bool")
    dim = operator.index(raw)
    if dim <= 0:
        raise ValueError(f"腿 {label!r} 的维度必须为正整数，得到 {dim}")
    return dim


def _to_int_labels(inputs, output, size_dict):
    """Hashable labels -> dense u32 leg ids, numbered by first occurrence over the
    inputs in order and then the output (the numbering and error messages of
    pybind/python/arctn/__init__.py:186-235). size_dict must size exactly the
    labels in use, each with a positive integer."""
    order = list(dict.fromkeys(itertools.chain(itertools.chain.from_iterable(inputs), output)))
    ids = {lab: i for i, lab in enumerate(order)}
    unsized = [lab for lab in order if lab not in size_dict]
    if unsized:
        raise ValueError(f"size_dict 缺少已使用腿的维度: {unsized!r}")
    stray = [lab for lab in size_dict if lab not in ids]
    if stray:
        raise ValueError(f"size_dict 含未在 inputs/output 使用的腿: {stray!r}")
    dims = {ids[lab]: _positive_dim(lab, size_dict[lab]) for lab in order}
    return [[ids[l] for l in t] for t in inputs], [ids[l] for l in output], dims, ids


def arctn_contract(inputs, output, size_dict, arrays, *, preset="greedy", seed=0, target_size=None,
                   slicing_mode="fixed", max_time=None, backend="hip", return_info=False,
                   allow_output_slicing=False, ntrials=1, **opts):
    """Plan on the CPU, optionally slice to ``target_size``, execute on the GPU.

    Mirrors pybind/python/arctn/__init__.py:903-1072 with ``backend="hip"``.
    Planning uses the reference's open greedy (``preset="greedy"``) or its
    random-greedy trials (``preset="rgreedy"``, ``ntrials``); the proprietary
    Light/Heavy engine is not available. Only ``slicing_mode="fixed"`` exists.
    """
    if backend != "hip":
        raise ValueError(f"backend 必须是 'hip'（this package is the HIP backend），收到 {backend!r}")
    if slicing_mode != "fixed":
        raise ValueError("only slicing_mode='fixed' is available")
    if target_size is not None:
        if isinstance(target_size, bool) or operator.index(target_size) <= 0:
            raise ValueError("target_size 必须大于 0")
    inputs_i, output_i, size_i, _ = _to_int_labels(inputs, output, size_dict)
    t0 = time.perf_counter()
    dev, code, numpy_in = _device_inputs(inputs_i, size_i, list(arrays), "切片收缩失败")
    prepare_s = time.perf_counter() - t0
    net = {"inputs": inputs_i, "output": output_i, "size_dict": size_i}
    t0 = time.perf_counter()
    if preset == "greedy":
        path = _planner.greedy(net)
        sliced = []
        if target_size is not None:
            sliced = _planner.find_slices_to_size(net, path, target_size, allow_output_slicing)
    elif preset == "rgreedy":
        best = _planner.random_greedy_sliced(net, ntrials, seed, target_size or (1 << 62), allow_output_slicing)
        path, sliced = best[0], best[1]
    else:
        raise ValueError(f"preset 只支持 'greedy' 或 'rgreedy'，收到 {preset!r}")
    if target_size is not None and sliced is None:
        raise ValueError(f"target_size 上限（log2={math.log2(target_size):.6g}）不可达")
    plan_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan = Plan(inputs_i, output_i, size_i, path, sliced or [])
    out, rep = plan.run(dev, code, report=True, **opts)
    exec_s = time.perf_counter() - t0
    result = _finish(out, numpy_in)
    if not return_info:
        return result
    info = {"path": [list(p) for p in path], "sliced": bool(sliced), "sliced_legs": list(sliced or []),
            "target_size": target_size, "allow_output_slicing": allow_output_slicing,
            "execution_backend": backend, "execution_implementation": "arctn_amd.hip_executor",
            "execution_input_prepare_wall_s": prepare_s, "planning_wall_s": plan_s,
            "execution_dispatch_wall_s": exec_s, "execution_device_ms": rep["elapsed_ms"],
            "execution_timing_scope": "hip_events", "report": rep}
    return result, info


class GroupedContraction:
    """Many independent unsliced plans executed together in lockstep grouped
    launches (one launch per kernel kind and step index over all plans).

    Each plan's result follows CompiledContraction.execute / execute_stripped
    (src/compiled.rs:369-441). The launch descriptors, the grouping into
    memory-bounded groups and the device arena are built once here;
    ``run`` replays them on new inputs.
    """

    def __init__(self, plans, dtype, strip=True, precision="native"):
        if precision not in _lib.PREC:
            raise ValueError(f"unknown precision {precision!r}; use one of {sorted(_lib.PREC)}")
        self.plans = list(plans)
        self.precision = precision
        if isinstance(dtype, int):
            self.code = dtype
        else:
            self.code = DTYPE_CODE.get(np.dtype(dtype))
            if self.code is None:
                raise ValueError(f"unsupported dtype {dtype!r}")
        self.strip = bool(strip)
        self._h = _lib.C.c_void_p()
        arr = (_lib._VP * len(self.plans))(*[p.handle for p in self.plans])
        _lib.check(_lib.lib().atn_grouped_create(arr, len(self.plans), self.code, _lib.PREC[precision], int(self.strip),
                                                 _lib.C.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib.lib().atn_grouped_destroy(self._h)
            self._h = None

    def run(self, dev_inputs, outs=None, report=False, stream=None):
        """dev_inputs[p]: device tensors of plan p (the executor's dtype).
        Returns (outputs, log10 exponents[, report])."""
        torch = _torch()
        n = len(self.plans)
        if len(dev_inputs) != n:
            raise ValueError(f"arrays 个数 {len(dev_inputs)} 与 plans 个数 {n} 不符")
        if n == 0:
            return ([], [], {}) if report else ([], [])
        dev = dev_inputs[0][0].device if dev_inputs[0] else torch.device("cuda", torch.cuda.current_device())
        if outs is None:
            outs = [torch.empty([p.size_dict[l] for l in p.output], dtype=_torch_dtype(self.code), device=dev)
                    for p in self.plans]
        keep = [_lib.pointer_array([t.data_ptr() for t in d]) for d in dev_inputs]
        in_arr = (_lib.C.POINTER(_lib._VP) * n)(*[_lib.C.cast(a, _lib.C.POINTER(_lib._VP)) for a in keep])
        out_arr = _lib.pointer_array([t.data_ptr() for t in outs])
        exps = (_lib.C.c_double * n)()
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        rep = _lib.Report()
        _lib.check(_lib.lib().atn_grouped_run(self._h, in_arr, out_arr, exps, s.cuda_stream, _lib.C.byref(rep)))
        if report:
            return outs, list(exps), rep.as_dict()
        return outs, list(exps)


def contract_grouped(plans, arrays, *, strip=True, precision="native", report=False, stream=None):
    """Contract many independent unsliced plans in lockstep grouped launches.

    ``plans`` is a list of unsliced ``Plan``; ``arrays[p]`` the input arrays of
    plan p (all of one dtype). With ``strip`` every step result is
    renormalised by a power of two and the result of plan p is
    ``outputs[p] * 10**exps[p]`` (execute_stripped). Returns ``(outputs, exps)``
    (plus the report dict when ``report``); outputs are torch tensors, or NumPy
    arrays for NumPy inputs.
    """
    if len(plans) != len(arrays):
        raise ValueError(f"plans 个数 {len(plans)} 与 arrays 个数 {len(arrays)} 不符")
    if not plans:
        return ([], [], {}) if report else ([], [])
    devs, code, numpy_in = [], None, False
    for p, arrs in zip(plans, arrays):
        d, c, n_in = _device_inputs(p.inputs, p.size_dict, arrs, "分组收缩失败")
        if code is not None and c != code:
            raise ValueError("分组收缩失败: 所有实例的 dtype 必须一致")
        code, numpy_in = c, numpy_in or n_in
        devs.append(d)
    try:
        g = GroupedContraction(plans, code, strip, precision)
        res = g.run(devs, report=report, stream=stream)
    except ValueError as e:
        raise ValueError(f"分组收缩失败: {e}") from None
    outs = [_finish(o, numpy_in) for o in res[0]]
    return (outs,) + tuple(res[1:])
