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

    def forward_ptrs(self, ptrs, dtype, device, stream):
        """forward with the input pointers given as an array of addresses."""
        out = torch.empty(self.output_shape, dtype=dtype, device=device)
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


class _TableLayout:
    """Where each input of a plan comes from in a contract_table call: row
    ``index[i] >= 0`` of the table, or fixed tensor ``-index[i] - 1``. Element
    offsets of the table rows, and of every input's gradient in one flat buffer
    (table inputs first, in input order, then the fixed ones), computed once."""

    def __init__(self, plan, index, row_shape, n_rows, n_fixed):
        import numpy as np

        shapes, _ = _shapes(plan)
        if len(index) != len(shapes):
            raise ValueError(f"张量个数 {len(index)} 与编译计划 {len(shapes)} 不符")
        idx = np.asarray(index, dtype=np.int64)
        tab = idx >= 0
        if np.any(idx >= n_rows) or np.any(idx < -n_fixed):
            raise ValueError("index names a table row or fixed tensor that does not exist")
        for i in np.flatnonzero(tab):
            if shapes[i] != tuple(row_shape):
                raise ValueError(f"张量 {i} shape {list(row_shape)} does not match the plan's {list(shapes[i])}")
        row_numel = int(np.prod(row_shape, dtype=np.int64))
        self.index = idx
        self.table_inputs = np.flatnonzero(tab)
        self.rows = idx[tab]
        self.row_off = self.rows * row_numel  # element offset of each table input's row
        numel = np.array([int(np.prod(sh, dtype=np.int64)) for sh in shapes], dtype=np.int64)
        order = np.concatenate([np.flatnonzero(tab), np.flatnonzero(~tab)])
        goff = np.zeros(len(shapes), dtype=np.int64)
        goff[order] = np.concatenate([[0], np.cumsum(numel[order])[:-1]]) if len(order) else []
        self.grad_off = goff
        self.grad_numel = int(numel.sum())
        self.n_table_inputs = int(tab.sum())
        self.fixed_of = {int(i): int(-idx[i] - 1) for i in np.flatnonzero(~tab)}
        self.fixed_shapes = {int(i): shapes[i] for i in np.flatnonzero(~tab)}

    def rows_on(self, device):
        """The table rows of the table inputs as a device index tensor (kept)."""
        got = self.__dict__.setdefault("_rows_dev", {}).get(device)
        if got is None:
            got = self._rows_dev[device] = torch.from_numpy(self.rows).to(device)
        return got

    def pointers(self, table, fixed):
        """The input address array for this table and these fixed tensors."""
        import numpy as np

        es = table.element_size()
        addr = np.empty(len(self.index), dtype=np.uint64)
        addr[self.table_inputs] = (table.data_ptr() + self.row_off * es).astype(np.uint64)
        for i, j in self.fixed_of.items():
            addr[i] = fixed[j].data_ptr()
        return addr


class ContractTableFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, layout, precision, checkpoint_every, table, *fixed):
        r = _runner(plan, _CODE[table.dtype], precision, int(checkpoint_every or 0))
        addr = layout.pointers(table, fixed)
        stream = torch.cuda.current_stream(table.device).cuda_stream
        out = r.forward_ptrs(addr.ctypes.data_as(_lib.C.POINTER(_lib._VP)), table.dtype, table.device, stream)
        ctx.runner, ctx.generation, ctx.layout, ctx.addr = r, r.generation, layout, addr
        ctx.save_for_backward(table, *fixed)
        return out

    @staticmethod
    def backward(ctx, gout):
        import numpy as np

        r, layout, addr = ctx.runner, ctx.layout, ctx.addr
        table, *fixed = ctx.saved_tensors
        stream = torch.cuda.current_stream(gout.device).cuda_stream
        in_ptrs = addr.ctypes.data_as(_lib.C.POINTER(_lib._VP))
        if r.generation != ctx.generation:
            # another forward ran on this runner since: rebuild this one's checkpoints
            r.forward_ptrs(in_ptrs, table.dtype, table.device, stream)
            ctx.generation = r.generation
        gout = gout.resolve_conj().resolve_neg().contiguous()
        if tuple(gout.shape) != r.output_shape or _CODE.get(gout.dtype) != r.code:
            raise ValueError("output gradient shape or dtype does not match the contraction's output")
        flat = torch.empty(layout.grad_numel, dtype=gout.dtype, device=gout.device)
        gaddr = (flat.data_ptr() + layout.grad_off * flat.element_size()).astype(np.uint64)
        _lib.check(_lib.lib().atn_adjoint_backward(r.h, in_ptrs, gout.data_ptr(),
                                                   gaddr.ctypes.data_as(_lib.C.POINTER(_lib._VP)), stream))
        # the table inputs' gradients, in input order, added into their rows
        gin = flat[:layout.n_table_inputs * table[0].numel()].view(layout.n_table_inputs, *table.shape[1:])
        gtab = torch.zeros_like(table).index_add_(0, layout.rows_on(table.device), gin)
        return (None, None, None, None, gtab) + (None,) * len(fixed)


def contract_table(plan: Plan, table, index, fixed=(), precision="native", checkpoint_every=None):
    """Differentiable contraction whose inputs are rows of one parameter table.

    Input i of the plan is ``table[index[i]]`` when ``index[i] >= 0`` and the fixed
    tensor ``fixed[-index[i] - 1]`` otherwise (fixed tensors get no gradient). The
    result and the gradient of ``table`` are those of ``contract`` with those
    tensors; what is saved is the per-input work of many small tensors: one table
    and an address array cross the ABI, and the inputs' gradients come back as one
    buffer added into the table's rows (rows named by several inputs accumulate).
    """
    if precision not in _lib.PREC:
        raise ValueError(f"unknown precision {precision!r}")
    if not isinstance(plan, Plan):
        raise ValueError("plan must be an arctn_amd.Plan")
    if not isinstance(table, torch.Tensor) or table.dim() < 1:
        raise ValueError("table must be a torch tensor with a leading row axis")
    if _CODE.get(table.dtype) is None:
        raise ValueError(f"unsupported dtype {table.dtype}")
    key = (tuple(index), tuple(table.shape), len(fixed))
    cache = plan.__dict__.setdefault("_table_layouts", {})
    layout = cache.get(key)
    if layout is None:
        layout = cache[key] = _TableLayout(plan, index, tuple(table.shape[1:]), table.shape[0], len(fixed))
    for j, f in enumerate(fixed):
        if not isinstance(f, torch.Tensor) or f.dtype != table.dtype or f.device != table.device:
            raise ValueError(f"fixed tensor {j} is not a tensor of the table's dtype {table.dtype} on {table.device}")
    for i, j in layout.fixed_of.items():
        if tuple(fixed[j].shape) != layout.fixed_shapes[i]:
            raise ValueError(f"张量 {i} shape {list(fixed[j].shape)} does not match the plan's {list(layout.fixed_shapes[i])}")
    if table.device.type != "cuda":
        raise ValueError("the inputs must be GPU tensors")
    table = table.resolve_conj().resolve_neg().contiguous()
    fixed = [f.resolve_conj().resolve_neg().contiguous() for f in fixed]
    return ContractTableFunction.apply(plan, layout, precision, checkpoint_every, table, *fixed)


def contract(plan: Plan, *tensors, precision="native", checkpoint_every=None):
    """Differentiable contraction of device tensors with an unsliced plan."""
    if precision not in _lib.PREC:
        raise ValueError(f"unknown precision {precision!r}")
    if not isinstance(plan, Plan):
        raise ValueError("plan must be an arctn_amd.Plan")
    return ContractFunction.apply(plan, precision, checkpoint_every, *tensors)
