"""Differentiable contraction on the GPU: a torch.autograd.Function over the
executor's stage entry points.

The forward pass runs the compiled plan stage by stage (input preps, pair
steps in SSA order, finalize), the same sequence as CompiledContraction's
stage API (src/compiled.rs:7-10, 60-63, 143-153, 206-209). The backward pass
runs the hand-written adjoint of every stage in reverse order:

* pair step  C = perm_a(sum_a(A)) . perm_b(sum_b(B)) (batched GEMM):
  dA2 = dC . B2^H and dB2 = A2^H . dC (conjugate-transposed GEMMs), then the
  inverse permutations and a broadcast over every summed leg;
* input prep (trace / diagonal): the gradient is placed on the diagonal;
* finalize: inverse permutation and broadcast over residual summed legs.

Gradients follow PyTorch's complex convention (the incoming gradient is
dL/d(conj out)), so real losses of complex contractions differentiate through
torch's own autograd around this function.

Checkpointing: with ``checkpoint_every=k`` only the live SSA slots at every
k-th step boundary are kept; the backward pass recomputes each segment's
intermediates from its boundary before running the segment's adjoints.
"""

from __future__ import annotations

import torch

from . import _lib
from .api import Plan

_CODE = {torch.float32: 0, torch.float64: 1, torch.complex64: 2, torch.complex128: 3}


class _Stages:
    """Shapes and slot bookkeeping of a plan's per-slice stages."""

    def __init__(self, plan: Plan):
        d = plan.describe()
        self.n_inputs = d["n_inputs"]
        self.preps = d["preps"]
        self.steps = d["steps"]
        self.fin = d["finalize"]
        if plan.info()["n_slices"] != 1:
            raise ValueError("autograd contraction needs an unsliced plan")


def _stream(t):
    return torch.cuda.current_stream(t.device).cuda_stream


def _prep_fwd(plan, i, x):
    st = plan._stages.preps[i]
    if not st["traces"]:
        return x
    out = torch.empty(st["shape"], dtype=x.dtype, device=x.device)
    _lib.check(_lib.lib().atn_prep_forward(plan.handle, i, _CODE[x.dtype], x.data_ptr(), out.data_ptr(), _stream(x)))
    return out


def _step_fwd(plan, s, a, b, prec):
    st = plan._stages.steps[s]
    out = torch.empty(st["result_shape"], dtype=a.dtype, device=a.device)
    _lib.check(_lib.lib().atn_step_forward(plan.handle, s, _CODE[a.dtype], _lib.PREC[prec], a.data_ptr(),
                                           b.data_ptr(), out.data_ptr(), _stream(a)))
    return out


def _fin_fwd(plan, x):
    out = torch.empty(plan._stages.fin["output_shape"], dtype=x.dtype, device=x.device)
    _lib.check(_lib.lib().atn_finalize_forward(plan.handle, _CODE[x.dtype], x.data_ptr(), out.data_ptr(), _stream(x)))
    return out


class ContractFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, precision, checkpoint_every, *inputs):
        if not hasattr(plan, "_stages"):
            plan._stages = _Stages(plan)
        S = plan._stages
        # materialise lazy conjugate / negative views before handing raw pointers over
        inputs = [t.resolve_conj().resolve_neg().contiguous() for t in inputs]
        slots = {i: _prep_fwd(plan, i, t) for i, t in enumerate(inputs)}
        keep = checkpoint_every and checkpoint_every > 0
        boundaries = {0: dict(slots)} if keep else None
        saved = dict(slots) if not keep else None
        for s, st in enumerate(S.steps):
            if keep and s > 0 and s % checkpoint_every == 0:
                boundaries[s] = dict(slots)
            c = _step_fwd(plan, s, slots.pop(st["li"]), slots.pop(st["ri"]), precision)
            slots[st["out"]] = c
            if saved is not None:
                saved[st["out"]] = c
        final = slots[S.fin["slot"]]
        out = _fin_fwd(plan, final)
        ctx.plan, ctx.precision, ctx.every = plan, precision, checkpoint_every
        ctx.boundaries, ctx.saved = boundaries, saved
        ctx.final = final
        ctx.save_for_backward(*inputs)
        return out

    @staticmethod
    def backward(ctx, gout):
        plan, S = ctx.plan, ctx.plan._stages
        L = _lib.lib()
        inputs = ctx.saved_tensors
        gout = gout.contiguous()
        code = _CODE[gout.dtype]
        grads = {}
        gfin = torch.empty(S.fin["input_shape"], dtype=gout.dtype, device=gout.device)
        _lib.check(L.atn_finalize_adjoint(plan.handle, code, gout.data_ptr(), gfin.data_ptr(), _stream(gout)))
        grads[S.fin["slot"]] = gfin
        n_steps = len(S.steps)
        if ctx.saved is not None:
            segments = [(0, n_steps, ctx.saved)]
        else:
            starts = sorted(ctx.boundaries)
            segments = [(a, (starts[i + 1] if i + 1 < len(starts) else n_steps), None) for i, a in enumerate(starts)]
        for a, b, vals in reversed(segments):
            if vals is None:  # recompute the segment from its boundary
                live = dict(ctx.boundaries[a])
                vals = dict(live)
                for s in range(a, b):
                    st = S.steps[s]
                    c = _step_fwd(plan, s, live.pop(st["li"]), live.pop(st["ri"]), ctx.precision)
                    live[st["out"]] = c
                    vals[st["out"]] = c
            for s in range(b - 1, a - 1, -1):
                st = S.steps[s]
                g = grads.pop(st["out"], None)
                if g is None:
                    continue
                left, right = vals[st["li"]], vals[st["ri"]]
                gl = torch.empty(st["left_shape"], dtype=g.dtype, device=g.device)
                gr = torch.empty(st["right_shape"], dtype=g.dtype, device=g.device)
                _lib.check(L.atn_step_adjoint(plan.handle, s, code, left.data_ptr(), right.data_ptr(), g.data_ptr(),
                                              gl.data_ptr(), gr.data_ptr(), _stream(g)))
                grads[st["li"]] = gl
                grads[st["ri"]] = gr
        out = []
        for i, t in enumerate(inputs):
            g = grads.get(i)
            if g is None:
                out.append(torch.zeros_like(t))
                continue
            if S.preps[i]["traces"]:
                gi = torch.empty(S.preps[i]["input_shape"], dtype=g.dtype, device=g.device)
                _lib.check(L.atn_prep_adjoint(plan.handle, i, code, g.data_ptr(), gi.data_ptr(), _stream(g)))
                g = gi
            out.append(g)
        return (None, None, None, *out)


def contract(plan: Plan, *tensors, precision="native", checkpoint_every=None):
    """Differentiable contraction of device tensors with an unsliced plan."""
    if precision not in _lib.PREC:
        raise ValueError(f"unknown precision {precision!r}")
    return ContractFunction.apply(plan, precision, checkpoint_every, *tensors)
