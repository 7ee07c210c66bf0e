"""Differentiable contraction on the GPU: a torch.autograd.Function over the
library's reverse-mode runner (atn_adjoint_* in include/arctn_hip.h).

The forward pass is one library call: every pair step of the compiled plan
runs level by level of the contraction tree (the small steps of a level in one
launch), the output is written and the checkpoint slots stay on the device.
The backward pass is one more call: the hand-written adjoint of every stage in
reverse order,

* pair step  C = perm_a(sum_a(A)) . perm_b(sum_b(B)) (batched GEMM):
  dA2 = dC . B2^H and dB2 = A2^H . dC (conjugate-transposed GEMMs), then the
  inverse permutations and a broadcast over every summed leg;
* input prep (trace / diagonal): the gradient is placed on the diagonal;
* finalize: inverse permutation and broadcast over residual summed legs,

recomputing the intermediates of each checkpoint segment first. The stage hook
of the reference (src/compiled.rs:7-10, 60-63, 143-153, 206-209) defines the
per-stage semantics; the reference itself has no gradient code.

Gradients follow PyTorch's complex convention (the incoming gradient is
dL/d(conj out)), so real losses of complex contractions differentiate through
torch's own autograd around this function.

Checkpointing: with ``checkpoint_every=k`` the steps, in tree-level order, form
segments of k; slots produced and consumed inside a segment are recomputed in
the backward pass, every other slot is kept from the forward pass.
"""

from __future__ import annotations

import torch

from . import _lib
from .api import Plan

_CODE = {torch.float32: 0, torch.float64: 1, torch.complex64: 2, torch.complex128: 3}


def _shapes(plan: Plan):
    """(input shapes, output shape) of an unsliced plan, cached on the plan."""
    got = plan.__dict__.get("_grad_shapes")
    if got is None:
        if plan.info()["n_slices"] != 1:
            raise ValueError("autograd contraction needs an unsliced plan")
        d = plan.describe()
        got = plan.__dict__["_grad_shapes"] = ([tuple(p["input_shape"]) for p in d["preps"]],
                                               tuple(d["finalize"]["output_shape"]))
    return got


def check_inputs(plan: Plan, inputs):
    """Count, dtype, device and shape of every input, checked before any pointer
    crosses the ABI (the reference checks every operand before its GEMMs,
    src/compiled.rs:386-398). Returns the dtype code."""
    shapes, _ = _shapes(plan)
    if len(inputs) != len(shapes):
        raise ValueError(f"张量个数 {len(inputs)} 与编译计划 {len(shapes)} 不符")
    code = _CODE.get(inputs[0].dtype) if inputs else None
    if code is None:
        raise ValueError(f"unsupported dtype {inputs[0].dtype if inputs else None}")
    dev = inputs[0].device
    for i, (t, want) in enumerate(zip(inputs, shapes)):
        if not isinstance(t, torch.Tensor):
            raise ValueError(f"张量 {i} is not a torch tensor")
        if _CODE.get(t.dtype) != code:
            raise ValueError(f"张量 {i} has dtype {t.dtype}, tensor 0 has {inputs[0].dtype}")
        if t.device != dev:
            raise ValueError(f"张量 {i} is on {t.device}, tensor 0 on {dev}")
        if tuple(t.shape) != want:
            raise ValueError(f"张量 {i} shape {list(t.shape)} does not match the plan's {list(want)}")
    if dev.type != "cuda":
        raise ValueError("the inputs must be GPU tensors")
    return code


class _Runner:
    """One atn_adjoint object: a plan's forward / backward for a dtype, precision
    and checkpoint spacing, with its device workspace."""

    def __init__(self, plan: Plan, code: int, precision: str, every: int):
        self.input_shapes, self.output_shape = _shapes(plan)
        self.code = code
        self.h = _lib.C.c_void_p()
        _lib.check(_lib.lib().atn_adjoint_create(plan.handle, code, _lib.PREC[precision], int(every),
                                                 _lib.C.byref(self.h)))
        self.generation = 0  # forward passes run so far

    def __del__(self):
        if getattr(self, "h", None) and _lib._lib is not None:
            _lib.lib().atn_adjoint_destroy(self.h)
            self.h = None

    def forward(self, inputs, stream):
        out = torch.empty(self.output_shape, dtype=inputs[0].dtype, device=inputs[0].device)
        ptrs = _lib.pointer_array([t.data_ptr() for t in inputs])
        _lib.check(_lib.lib().atn_adjoint_forward(self.h, ptrs, out.data_ptr(), stream))
        self.generation += 1
        return out

    def backward(self, inputs, gout, stream):
        # one buffer for all input gradients, handed out as views
        sizes = [t.numel() for t in inputs]
        flat = torch.empty(sum(sizes), dtype=gout.dtype, device=gout.device)
        grads, ptrs, off = [], [], 0
        es = flat.element_size()
        for t, n in zip(inputs, sizes):
            grads.append(flat[off:off + n].view(t.shape))
            ptrs.append(flat.data_ptr() + off * es)
            off += n
        _lib.check(_lib.lib().atn_adjoint_backward(self.h, _lib.pointer_array([t.data_ptr() for t in inputs]),
                                                   gout.data_ptr(), _lib.pointer_array(ptrs), stream))
        return grads


def _runner(plan, code, precision, every):
    cache = plan.__dict__.setdefault("_grad_runners", {})
    key = (code, precision, every)
    r = cache.get(key)
    if r is None:
        r = cache[key] = _Runner(plan, code, precision, every)
    return r


class ContractFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, precision, checkpoint_every, *inputs):
        code = check_inputs(plan, inputs)
        # materialise lazy conjugate / negative views before handing raw pointers over
        inputs = [t.resolve_conj().resolve_neg().contiguous() for t in inputs]
        r = _runner(plan, code, precision, int(checkpoint_every or 0))
        stream = torch.cuda.current_stream(inputs[0].device).cuda_stream
        out = r.forward(inputs, stream)
        ctx.runner, ctx.generation = r, r.generation
        ctx.save_for_backward(*inputs)
        return out

    @staticmethod
    def backward(ctx, gout):
        r = ctx.runner
        inputs = list(ctx.saved_tensors)
        stream = torch.cuda.current_stream(gout.device).cuda_stream
        if r.generation != ctx.generation:
            # another forward ran on this runner since: rebuild this one's checkpoints
            r.forward(inputs, stream)
            ctx.generation = r.generation
        gout = gout.resolve_conj().resolve_neg().contiguous()
        if tuple(gout.shape) != r.output_shape or _CODE.get(gout.dtype) != r.code:
            raise ValueError("output gradient shape or dtype does not match the contraction's output")
        grads = r.backward(inputs, gout, stream)
        return (None, None, None, *grads)


def contract(plan: Plan, *tensors, precision="native", checkpoint_every=None):
    """Differentiable contraction of device tensors with an unsliced plan."""
    if precision not in _lib.PREC:
        raise ValueError(f"unknown precision {precision!r}")
    if not isinstance(plan, Plan):
        raise ValueError("plan must be an arctn_amd.Plan")
    return ContractFunction.apply(plan, precision, checkpoint_every, *tensors)
