"""Differentiation form on the GPU: hand-written adjoint vs the oracle's
step-by-step complex128 gradient, checkpointing, and the 20-qubit brickwork
gradient (400 angles) vs central finite differences."""
import numpy as np
import pytest

from arctn_amd import Plan
from arctn_amd import planner as P
from arctn_amd import workloads as W
from oracle import pyoracle as O
from tests.golden import fixtures as F

pytestmark = pytest.mark.gpu


def _grads(plan, arrays, w, **kw):
    import torch
    from arctn_amd.autograd import contract

    xs = [torch.from_numpy(a).cuda().requires_grad_(True) for a in arrays]
    out = contract(plan, *xs, **kw)
    gs = torch.autograd.grad(out, xs, grad_outputs=torch.from_numpy(w).cuda())
    return out.detach().cpu().numpy(), [g.cpu().numpy() for g in gs]


@pytest.mark.parametrize("dtype", [np.complex128, np.float64])
def test_adjoint_matches_oracle_gradient(cuda, dtype):
    cases = []
    for seed in range(8):
        rng = np.random.default_rng(seed)
        n = int(rng.integers(2, 8))
        net = W.random_connected(n, int(rng.integers(0, n)), [2, 3, 4], int(rng.integers(0, 3)), seed)
        cases.append((net, P.greedy(net)))
    fx = F.load()["mixed_network"]
    cases.append((F.net_of(fx), [tuple(p) for p in fx["hand_path"]]))
    for net, path in cases:
        arrays = W.random_arrays(net, dtype, 1)
        plan = Plan(net["inputs"], net["output"], net["size_dict"], path)
        oshape = [net["size_dict"][l] for l in net["output"]]
        rng = np.random.default_rng(2)
        w = rng.standard_normal(oshape) + (1j * rng.standard_normal(oshape) if dtype == np.complex128 else 0)
        w = np.asarray(w, dtype=dtype)
        out, gs = _grads(plan, arrays, w)
        np.testing.assert_allclose(out, O.contract(net, arrays, path), rtol=1e-10, atol=1e-10)
        want = O.gradient(net, arrays, path, w)
        for g, r in zip(gs, want):
            assert np.abs(g - r).max() <= 1e-10 * max(1.0, np.abs(r).max())


def test_checkpointing_is_bit_identical(cuda):
    net, arrays, _ = W.random_circuit_amplitude(3, 3, 4, 1)
    path = P.greedy(net)
    plan = Plan(net["inputs"], net["output"], net["size_dict"], path)
    w = np.array(1.0 + 0.5j, dtype=np.complex64)
    _, g0 = _grads(plan, arrays, w)
    for k in (1, 3, 8, 1000):
        _, gk = _grads(plan, arrays, w, checkpoint_every=k)
        assert all(a.tobytes() == b.tobytes() for a, b in zip(g0, gk)), k


def test_brickwork_20q_gradient_vs_finite_differences(cuda):
    """<psi|Z0 Z1|psi>, 20 qubits, 10 layers of RY.RZ + CZ brickwork, 400 angles.

    complex64 forward + hand-written backward (checkpoint every 8 steps) vs
    central differences (h = 1e-3) of a complex128 forward: max abs diff
    <= 1e-4 * max(1, max |grad|); and the complex128 adjoint vs the CPU
    oracle's step-by-step gradient chained to the angles by torch on the CPU.
    """
    import torch
    from arctn_amd.autograd import contract

    nq, layers = 20, 10
    net, slots = W.brickwork_structure(nq, layers)
    path = P.greedy(net)
    plan = Plan(net["inputs"], net["output"], net["size_dict"], path)
    ang = W.brickwork_angles(nq, layers, 0)

    th = torch.tensor(ang, dtype=torch.float32, device="cuda", requires_grad=True)
    val = contract(plan, *W.brickwork_torch_tensors(slots, th), checkpoint_every=8).real
    (g32,) = torch.autograd.grad(val, th)
    g32 = g32.cpu().numpy()

    th64 = torch.tensor(ang, dtype=torch.float64, device="cuda", requires_grad=True)
    val64 = contract(plan, *W.brickwork_torch_tensors(slots, th64)).real
    (g64,) = torch.autograd.grad(val64, th64)
    g64 = g64.cpu().numpy()

    base = [t.detach() for t in W.brickwork_torch_tensors(slots, th64.detach())]
    h = 1e-3
    fd = np.zeros_like(ang)
    with torch.no_grad():
        for l in range(layers):
            for q in range(nq):
                for j in range(2):
                    vals = []
                    for sgn in (1.0, -1.0):
                        a2 = th64.detach().clone()
                        a2[l, q, j] += sgn * h
                        ts = W.brickwork_torch_tensors(slots, a2)
                        out, _ = plan.run(ts, 3)
                        vals.append(float(out.real))
                    fd[l, q, j] = (vals[0] - vals[1]) / (2 * h)
    scale = max(1.0, np.abs(fd).max())
    assert np.abs(g32 - fd).max() <= 1e-4 * scale, np.abs(g32 - fd).max()
    assert np.abs(g64 - fd).max() <= 1e-5 * scale

    # CPU step-by-step gradient (oracle, complex128) chained to the angles on the CPU
    thc = torch.tensor(ang, dtype=torch.float64, requires_grad=True)
    tsc = W.brickwork_torch_tensors(slots, thc)
    arrays = [t.detach().numpy() for t in tsc]
    g_in = O.gradient(net, arrays, path, np.array(1.0 + 0j))
    # d Re(out): the conjugate-convention gradient of Re(out) is grad_out = 1
    pairs = [(t, torch.from_numpy(g)) for t, g in zip(tsc, g_in) if t.requires_grad]
    torch.autograd.backward([t for t, _ in pairs], [g for _, g in pairs])
    assert np.abs(thc.grad.numpy() - g64).max() <= 1e-9 * scale


def _stage_gradient(plan, xs, gout):
    """Forward and reverse pass driven stage by stage through the stage ABI
    (atn_prep/step/finalize_forward and _adjoint), the hook a tape would use."""
    import torch
    from arctn_amd import _lib

    L, d = _lib.lib(), plan.describe()
    code = {torch.complex128: 3, torch.float64: 1, torch.complex64: 2, torch.float32: 0}[xs[0].dtype]
    s = torch.cuda.current_stream().cuda_stream
    slots = {}
    for i, (x, p) in enumerate(zip(xs, d["preps"])):
        y = torch.empty(p["shape"], dtype=x.dtype, device="cuda")
        _lib.check(L.atn_prep_forward(plan.handle, i, code, x.data_ptr(), y.data_ptr(), s))
        slots[i] = y
    for k, st in enumerate(d["steps"]):
        y = torch.empty(st["result_shape"], dtype=xs[0].dtype, device="cuda")
        _lib.check(L.atn_step_forward(plan.handle, k, code, 0, slots[st["li"]].data_ptr(), slots[st["ri"]].data_ptr(),
                                      y.data_ptr(), s))
        slots[st["out"]] = y
    fin = d["finalize"]
    out = torch.empty(fin["output_shape"], dtype=xs[0].dtype, device="cuda")
    _lib.check(L.atn_finalize_forward(plan.handle, code, slots[fin["slot"]].data_ptr(), out.data_ptr(), s))
    grads = {fin["slot"]: torch.empty(fin["input_shape"], dtype=xs[0].dtype, device="cuda")}
    _lib.check(L.atn_finalize_adjoint(plan.handle, code, gout.data_ptr(), grads[fin["slot"]].data_ptr(), s))
    for k in reversed(range(len(d["steps"]))):
        st = d["steps"][k]
        gl = torch.empty(st["left_shape"], dtype=xs[0].dtype, device="cuda")
        gr = torch.empty(st["right_shape"], dtype=xs[0].dtype, device="cuda")
        _lib.check(L.atn_step_adjoint(plan.handle, k, code, slots[st["li"]].data_ptr(), slots[st["ri"]].data_ptr(),
                                      grads[st["out"]].data_ptr(), gl.data_ptr(), gr.data_ptr(), s))
        grads[st["li"]], grads[st["ri"]] = gl, gr
    res = []
    for i, (x, p) in enumerate(zip(xs, d["preps"])):
        g = torch.empty(p["input_shape"], dtype=x.dtype, device="cuda")
        _lib.check(L.atn_prep_adjoint(plan.handle, i, code, grads[i].data_ptr(), g.data_ptr(), s))
        res.append(g)
    return out, res


@pytest.mark.parametrize("dtype", [np.complex128, np.float64])
def test_stage_abi_gradient_matches_runner(cuda, dtype):
    """The per-stage ABI (prep / step / finalize forward and adjoint, one call per
    stage) and the two-call runner give the same value and gradients, on random
    networks and on the fixture network with traces and diagonals."""
    import torch

    cases = []
    for seed in range(6):
        rng = np.random.default_rng(seed)
        n = int(rng.integers(2, 8))
        net = W.random_connected(n, int(rng.integers(0, n)), [2, 3, 4], int(rng.integers(0, 3)), seed)
        cases.append((net, P.greedy(net)))
    fx = F.load()["mixed_network"]
    cases.append((F.net_of(fx), [tuple(p) for p in fx["hand_path"]]))
    for net, path in cases:
        arrays = W.random_arrays(net, dtype, 3)
        plan = Plan(net["inputs"], net["output"], net["size_dict"], path)
        oshape = [net["size_dict"][l] for l in net["output"]]
        w = np.asarray(np.random.default_rng(4).standard_normal(oshape), dtype=dtype)
        out, gs = _grads(plan, arrays, w)
        xs = [torch.from_numpy(a).cuda() for a in arrays]
        out2, gs2 = _stage_gradient(plan, xs, torch.from_numpy(w).cuda())
        assert np.abs(out2.cpu().numpy() - out).max() <= 1e-12 * max(1.0, np.abs(out).max())
        for g, g2 in zip(gs, gs2):
            assert np.abs(g2.cpu().numpy() - g).max() <= 1e-12 * max(1.0, np.abs(g).max())


def test_runner_interleaved_forwards(cuda):
    """Two forward passes on one runner before their backward passes: each backward
    rebuilds its own checkpoints, so both gradients stay exact."""
    import torch
    from arctn_amd.autograd import contract

    net, arrays, _ = W.random_circuit_amplitude(3, 3, 4, 2, dtype=np.complex128)
    path = P.greedy(net)
    plan = Plan(net["inputs"], net["output"], net["size_dict"], path)
    xs1 = [torch.from_numpy(a).cuda().requires_grad_(True) for a in arrays]
    xs2 = [torch.from_numpy(a * 0.5).cuda().requires_grad_(True) for a in arrays]
    o1, o2 = contract(plan, *xs1, checkpoint_every=4), contract(plan, *xs2, checkpoint_every=4)
    g2 = torch.autograd.grad(o2, xs2, grad_outputs=torch.ones_like(o2))
    g1 = torch.autograd.grad(o1, xs1, grad_outputs=torch.ones_like(o1))
    w = np.array(1.0 + 0j)
    for xs, gs in ((arrays, g1), ([a * 0.5 for a in arrays], g2)):
        want = O.gradient(net, xs, path, w)
        for g, r in zip(gs, want):
            assert np.abs(g.cpu().numpy() - r).max() <= 1e-10 * max(1.0, np.abs(r).max())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_contract_table_matches_per_tensor_contract(cuda, dtype):
    """The gate-table form (autograd.contract_table: inputs as rows of one table
    plus fixed tensors) gives the value and angle gradients of the per-tensor form
    (autograd.contract) bit for bit on a 10-qubit 6-layer brickwork, with and
    without checkpointing; a row named by two inputs accumulates both gradients
    (T.T against torch's own matmul gradient)."""
    import torch
    from arctn_amd.autograd import contract, contract_table

    nq, layers = 10, 6
    net, slots = W.brickwork_structure(nq, layers)
    plan = Plan(net["inputs"], net["output"], net["size_dict"], P.greedy(net))
    th = torch.tensor(W.brickwork_angles(nq, layers, 1), dtype=torch.from_numpy(np.zeros(1, dtype)).dtype,
                      device="cuda", requires_grad=True)
    for every in (None, 4):
        v1 = contract(plan, *W.brickwork_torch_tensors(slots, th), checkpoint_every=every).real
        (g1,) = torch.autograd.grad(v1, th)
        table, index, fixed = W.brickwork_table(slots, th)
        v2 = contract_table(plan, table, index, fixed, checkpoint_every=every).real
        (g2,) = torch.autograd.grad(v2, th)
        assert float(v1) == float(v2)
        assert torch.equal(g1, g2), float((g1 - g2).abs().max())
    # one row, two inputs: d/dT sum(W * (T @ T)) = W T^T + T^T W
    mm = Plan([[0, 1], [1, 2]], [0, 2], {0: 2, 1: 2, 2: 2}, [(0, 1)])
    cdt = torch.complex128 if dtype == np.float64 else torch.complex64
    t = torch.randn(1, 2, 2, dtype=cdt, device="cuda", requires_grad=True)
    w = torch.randn(2, 2, dtype=cdt, device="cuda")
    (gt,) = torch.autograd.grad((contract_table(mm, t, [0, 0]) * w).sum().real, t)
    t2 = t.detach().clone().requires_grad_(True)
    (gr,) = torch.autograd.grad((torch.matmul(t2[0], t2[0]) * w).sum().real, t2)
    assert torch.allclose(gt, gr, rtol=1e-5 if cdt == torch.complex64 else 1e-12, atol=1e-6)


@pytest.mark.parametrize("precision", ["fp16x2", "bf16x3"])
def test_brickwork_20q_split_precision_gradient(cuda, precision):
    """The 20-qubit brickwork's 400 angle gradients with the large forward products in
    a split precision form (the benchmark picks the fastest form within 1e-4 of central
    differences): within 1e-5 of the native form's gradient relative to max |g|, and not
    bit-identical to it (the split products did run)."""
    import torch
    from arctn_amd.autograd import contract_table

    nq, layers = 20, 10
    net, slots = W.brickwork_structure(nq, layers)
    plan = Plan(net["inputs"], net["output"], net["size_dict"], P.greedy(net))
    th = torch.tensor(W.brickwork_angles(nq, layers, 0), dtype=torch.float32, device="cuda", requires_grad=True)
    grads = {}
    for prec in ("native", precision):
        table, index, fixed = W.brickwork_table(slots, th)
        val = contract_table(plan, table, index, fixed, precision=prec, checkpoint_every=8).real
        (grads[prec],) = torch.autograd.grad(val, th)
    g0, g1 = grads["native"], grads[precision]
    scale = float(g0.abs().max())
    assert float((g1 - g0).abs().max()) <= 1e-5 * scale
    assert not torch.equal(g0, g1)
