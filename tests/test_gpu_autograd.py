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
