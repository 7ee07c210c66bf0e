"""Grouped execution of uneven independent instances on the GPU.

Every instance of one call runs in lockstep grouped launches; each result must
match the oracle's float64 contraction of that instance (src/compiled.rs:383-441
restated). Float32 results are held to 1e-5 relative error, or, where the
scalar comes out of heavy cancellation, to the error the same plan makes in
float32 on the CPU (four times it): float32 arithmetic itself cannot do better.
"""
import numpy as np
import pytest

from arctn_amd import CompiledContraction, Plan, contract_grouped
from arctn_amd import planner as P
from arctn_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _plans(insts):
    return [Plan(net["inputs"], net["output"], net["size_dict"], path) for net, _, path in insts]


def _value(m, e):
    return np.asarray(m, dtype=np.float64) * 10.0 ** e


def _large_chain(dtype, seed):
    """Three tensors whose steps exceed the grouped kernels' size limits."""
    rng = np.random.default_rng(seed)
    net = {"inputs": [[0, 1, 2], [2, 3], [3, 1, 4]], "output": [0, 4],
           "size_dict": {0: 256, 1: 64, 2: 64, 3: 64, 4: 256}}
    arrays = [(rng.standard_normal([net["size_dict"][l] for l in t]) / 8).astype(dtype) for t in net["inputs"]]
    return net, arrays, [(0, 1), (2, 3)]


def test_grouped_regular_instances_float64(cuda):
    insts = W.uneven_instances(24, seed=3, n_lo=8, n_hi=40, ntrials=2, dtype=np.float64)
    for strip in (True, False):
        outs, exps = contract_grouped(_plans(insts), [a for _, a, _ in insts], strip=strip)
        for (net, arrays, path), m, e in zip(insts, outs, exps):
            ref = O.contract(net, arrays, path)
            got = _value(m, e)
            assert abs(got - ref) <= 1e-12 * abs(ref), (len(net["inputs"]), strip, got, ref)


def test_grouped_regular_instances_float32(cuda):
    insts = W.uneven_instances(32, seed=5, n_lo=8, n_hi=48, ntrials=2)
    outs, exps = contract_grouped(_plans(insts), [a for _, a, _ in insts], strip=True)
    worst = 0.0
    for (net, arrays, path), m, e in zip(insts, outs, exps):
        ref = float(O.contract(net, [a.astype(np.float64) for a in arrays], path))
        cpu32 = float(O.contract(net, arrays, path))
        err = abs(float(_value(m, e)) - ref)
        assert err <= max(1e-5 * abs(ref), 4 * abs(cpu32 - ref)), (len(net["inputs"]), err, ref, cpu32)
        worst = max(worst, err / abs(ref))
    print(f"worst float32 relative error over {len(insts)} instances: {worst:.2e}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.complex128])
def test_grouped_mixed_shapes_and_large_steps(cuda, dtype):
    """Open-legged networks (finalize permutes), a trace input, and a chain
    whose steps are launched on their own, all in one call."""
    insts = []
    for seed in range(6):
        net = W.random_connected(int(3 + seed % 5), seed % 3, [2, 3, 4], seed % 3, seed)
        arrays = W.random_arrays(net, dtype, seed)
        insts.append((net, arrays, P.greedy(net)))
    rng = np.random.default_rng(9)
    tr = {"inputs": [[0, 1, 0], [1, 2]], "output": [2], "size_dict": {0: 3, 1: 4, 2: 5}}
    insts.append((tr, W.random_arrays(tr, dtype, 11), [(0, 1)]))
    insts.append(_large_chain(dtype, 4))
    del rng
    cplx = np.issubdtype(np.dtype(dtype), np.complexfloating)
    tol = 1e-12 if np.dtype(dtype) in (np.dtype(np.float64), np.dtype(np.complex128)) else 2e-5
    wide = np.complex128 if cplx else np.float64
    for strip in (False, True):
        outs, exps = contract_grouped(_plans(insts), [a for _, a, _ in insts], strip=strip)
        for (net, arrays, path), m, e in zip(insts, outs, exps):
            ref = O.contract(net, [a.astype(wide) for a in arrays], path)
            got = np.asarray(m, dtype=wide) * 10.0 ** e
            assert got.shape == ref.shape
            scale = max(np.abs(ref).max(), 1e-300)
            assert np.abs(got - ref).max() <= tol * scale, (net["inputs"], strip)


def test_grouped_matches_single_plan_stripped(cuda):
    """The grouped result equals each plan's own execute_stripped value."""
    insts = W.uneven_instances(8, seed=7, n_lo=10, n_hi=30, ntrials=1, dtype=np.float64)
    outs, exps = contract_grouped(_plans(insts), [a for _, a, _ in insts])
    for (net, arrays, path), m, e in zip(insts, outs, exps):
        cc = CompiledContraction.compile(net["inputs"], net["output"], net["size_dict"], ssa_path=path)
        m1, e1 = cc.execute_stripped(arrays)
        assert abs(_value(m, e) - _value(m1, e1)) <= 1e-12 * abs(_value(m1, e1))
        # the pair itself: scalar mantissa of magnitude 1, exponent log10 of the value
        assert abs(abs(float(m)) - 1.0) < 1e-12 and abs(e - e1) < 1e-12


def test_grouped_rejects_sliced_plans_and_mixed_dtypes(cuda):
    net = {"inputs": [[0, 1], [1, 2]], "output": [0, 2], "size_dict": {0: 2, 1: 2, 2: 2}}
    a = [np.ones((2, 2)), np.ones((2, 2))]
    sliced = Plan(net["inputs"], net["output"], net["size_dict"], [(0, 1)], [1])
    with pytest.raises(ValueError, match="unsliced"):
        contract_grouped([sliced], [a])
    p = Plan(net["inputs"], net["output"], net["size_dict"], [(0, 1)])
    with pytest.raises(ValueError, match="dtype"):
        contract_grouped([p, p], [a, [x.astype(np.float32) for x in a]])
    assert contract_grouped([], []) == ([], [])


def test_grouped_executor_reused_with_new_inputs(cuda):
    """One GroupedContraction run on two input sets (new pointers) gives each set's own results."""
    import torch
    from arctn_amd import GroupedContraction

    insts = W.uneven_instances(10, seed=11, n_lo=8, n_hi=26, ntrials=1, dtype=np.float64)
    plans = _plans(insts)
    g = GroupedContraction(plans, np.float64, strip=True)
    for scale in (1.0, -0.5):
        dev = [[torch.from_numpy(x * scale).cuda() for x in a] for _, a, _ in insts]
        outs, exps = g.run(dev)
        for (net, arrays, path), m, e in zip(insts, outs, exps):
            ref = O.contract(net, [x * scale for x in arrays], path)
            assert abs(float(m.cpu()) * 10.0 ** e - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x2"])
def test_grouped_split_precision(cuda, precision):
    """Split products on the steps launched on their own (a chain of large
    steps), within 2e-5 of the float64 oracle; the small instances stay FP32."""
    insts = [_large_chain(np.float32, 4), _large_chain(np.float32, 5)]
    insts += W.uneven_instances(6, seed=13, n_lo=8, n_hi=20, ntrials=1)
    for strip in (False, True):
        outs, exps = contract_grouped(_plans(insts), [a for _, a, _ in insts], strip=strip, precision=precision)
        for (net, arrays, path), m, e in zip(insts, outs, exps):
            ref = O.contract(net, [a.astype(np.float64) for a in arrays], path)
            got = np.asarray(m, dtype=np.float64) * 10.0 ** e
            cpu32 = O.contract(net, arrays, path).astype(np.float64)
            scale = np.abs(ref).max()
            bound = max(2e-5 * scale, 4 * np.abs(cpu32 - ref).max())
            assert np.abs(got - ref).max() <= bound, (precision, strip, np.abs(got - ref).max() / scale)


@pytest.mark.parametrize("dtype", [np.float32, np.complex64])
@pytest.mark.parametrize("shape", ["stream", "mfma"])
def test_grouped_factored_b_matches_matmul(cuda, dtype, shape):
    """One product whose B has its contracted legs interleaved with its free legs in
    memory ([f1, k1, f2, k2, f3]), read in place through a factored column offset:
    by the streaming kernel (m = 4, n = 2^21) and by the FP32 MFMA kernel (m = k = 16,
    n = 2^26: int32 gather tables would pass 2^26 entries). Checked against torch's
    matmul of the permuted operands on the device (FP32 products, k <= 16)."""
    import torch
    from arctn_amd import GroupedContraction

    m, k1, k2, f = (4, 2, 2, (32, 64, 1024)) if shape == "stream" else (16, 4, 4, (64, 1024, 1024))
    net = {"inputs": [[0, 1, 2], [3, 1, 4, 2, 5]], "output": [0, 3, 4, 5],
           "size_dict": {0: m, 1: k1, 2: k2, 3: f[0], 4: f[1], 5: f[2]}}
    tdt = torch.complex64 if dtype == np.complex64 else torch.float32
    gen = torch.Generator(device="cuda").manual_seed(21)
    a = torch.randn((m, k1, k2), dtype=tdt, device="cuda", generator=gen)
    b = torch.randn((f[0], k1, f[1], k2, f[2]), dtype=tdt, device="cuda", generator=gen)
    plan = Plan(net["inputs"], net["output"], net["size_dict"], [(0, 1)])
    g = GroupedContraction([plan], dtype, strip=False)
    (got,), (e,) = g.run([[a, b]])
    assert e == 0.0
    ref = torch.matmul(a.reshape(m, k1 * k2), b.permute(1, 3, 0, 2, 4).contiguous().reshape(k1 * k2, -1))
    ref = ref.reshape(got.shape)
    err = float((got - ref).abs().max()) / float(ref.abs().max())
    print(f"factored B {shape} {np.dtype(dtype).name}: rel err {err:.2e}", flush=True)
    assert err <= 1e-5
    del g, got, ref, b
    torch.cuda.empty_cache()
