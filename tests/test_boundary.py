"""C-ABI boundary and host logic, without a GPU.

The library must load and export every entry point include/arctn_hip.h
declares; plan compilation (host-only) must reproduce the oracle's compiled
plan exactly; malformed plans must be rejected with the reference's error
behaviour before any device work.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from arctn_amd import Plan
from arctn_amd import _lib
from arctn_amd import planner as P
from arctn_amd import workloads as W
from oracle import pyoracle as O
from tests.golden import fixtures as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FX = F.load()


def _declared():
    text = open(os.path.join(ROOT, "include", "arctn_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(atn_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert L.atn_abi_version() == 1


def test_ctypes_signatures_cover_the_header():
    assert set(_declared()) == set(_lib._SIGS)


def _strip(d):
    d = dict(d)
    d.pop("varying", None)
    return d


def test_plan_describe_matches_oracle_compile():
    cases = [(F.net_of(FX["mixed_network"]), [tuple(p) for p in FX["mixed_network"]["hand_path"]]),
             (F.net_of(FX["two_node_geometry"]), [(0, 1)]),
             (F.net_of(FX["single_tensor"]), [])]
    for seed in range(20):
        rng = np.random.default_rng(seed)
        n = int(rng.integers(2, 9))
        net = W.random_connected(n, int(rng.integers(0, n)), [2, 3, 4], int(rng.integers(0, 3)), seed)
        cases.append((net, P.greedy(net)))
    net, _, _ = W.random_circuit_amplitude(4, 4, 8, 0)
    cases.append((net, P.greedy(net)))
    for net, path in cases:
        mine = Plan(net["inputs"], net["output"], net["size_dict"], path).describe()
        ref = O.describe(net, path)
        assert mine["n_inputs"] == ref["n_inputs"]
        assert mine["preps"] == ref["preps"]
        assert [_strip(s) for s in mine["steps"]] == ref["steps"]
        assert mine["finalize"] == ref["finalize"]


def test_known_geometry_and_stats():
    fx = FX["two_node_geometry"]
    p = Plan(fx["inputs"], fx["output"], F.net_of(fx)["size_dict"], fx["ssa_path"])
    st = p.describe()["steps"][0]
    assert st["gemm"] == fx["gemm_dims"] and st["result_legs"] == fx["result_legs"]
    fx = FX["liveness_peak"]
    info = Plan(fx["inputs"], fx["output"], F.net_of(fx)["size_dict"], fx["ssa_path"]).info()
    assert info["log2_peak_size"] == fx["log2_peak_size"]
    fx = FX["chain_costs"]
    info = Plan(fx["inputs"], fx["output"], F.net_of(fx)["size_dict"], fx["ssa_path"]).info()
    assert abs(info["log10_flops"] - np.log10(fx["flops"])) < 1e-12
    # A(BC): B.C is (m, k, n) = (4, 8, 3), A.(BC) is (2, 4, 3): 96 + 24 multiply-adds, which
    # is also the reference's scalar-multiplication count (no pre-summed legs here)
    assert info["macs_per_slice"] == 120 and info["mults_per_slice"] == fx["flops"]
    fx = FX["two_node_geometry"]
    info = Plan(fx["inputs"], fx["output"], F.net_of(fx)["size_dict"], fx["ssa_path"]).info()
    assert info["macs_per_slice"] == 1 * 3 * 4 * 5


def test_slice_bookkeeping_matches_reference_rules():
    fx = FX["mixed_ragged_chunks"]
    net = F.net_of(fx)
    d = Plan(net["inputs"], net["output"], net["size_dict"], fx["ssa_path"], fx["sliced"]).describe()
    # output leg 0 is an outer (block) leg, 1 is internal; chunking follows slice.rs:893-905
    assert d["outer"] == [0] and d["internal"] == [1] and d["n_blocks"] == 2 and d["n_slices"] == 514
    assert d["n_chunks"] == O.slice_parallel_chunks(257, 3) == 129
    # internal slicing removes the sliced legs from every per-slice operand
    fx = FX["shared_output_hyperedge"]
    net = F.net_of(fx)
    d = Plan(net["inputs"], net["output"], net["size_dict"], fx["ssa_path"], [4, 2]).describe()
    assert all(4 not in s["left_legs"] + s["right_legs"] for s in d["steps"])
    assert d["outer"] == [2] and d["internal"] == [4]


@pytest.mark.parametrize("bad,match", [
    (dict(sliced=[0, 0]), "repeated"),
    (dict(path=[(0, 0)]), "itself"),
    (dict(path=[(0, 5)]), "invalid"),
    (dict(path=[]), "incomplete"),
    (dict(size_dict={0: 2, 1: 0, 2: 2}), "dimension 0"),
])
def test_malformed_plans_are_rejected(bad, match):
    inputs, output = [[0, 1], [1, 2]], [0, 2]
    size_dict = bad.get("size_dict", {0: 2, 1: 2, 2: 2})
    path = bad.get("path", [(0, 1)])
    with pytest.raises(ValueError, match=match):
        Plan(inputs, output, size_dict, path, bad.get("sliced", ()))


def test_trace_leg_cannot_be_sliced():
    with pytest.raises(ValueError, match="同一输入张量中重复"):
        Plan([[0, 0]], [], {0: 2}, [], [0])


def test_unused_size_dict_leg_rejected():
    with pytest.raises(ValueError, match="size_dict"):
        Plan([[0], [0]], [], {0: 2, 1: 2}, [(0, 1)])


def test_executor_options_validated_before_device_work():
    L = _lib.lib()
    p = Plan([[0], [0]], [], {0: 2}, [(0, 1)], [0])
    o = _lib.ExecOptions(2, 0, 0, 0, 0, 1, 0, 0)  # strip on a sliced plan
    h = C.c_void_p()
    rc = L.atn_executor_create(p.handle, C.byref(o), C.byref(h))
    assert rc != 0 and b"unsliced" in L.atn_last_error()
    o = _lib.ExecOptions(1, 2, 0, 0, 0, 0, 0, 0)  # split precision on float64
    rc = L.atn_executor_create(p.handle, C.byref(o), C.byref(h))
    assert rc != 0 and b"split precision" in L.atn_last_error()
    o = _lib.ExecOptions(1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 10)  # GEMM kernel code 9 does not exist
    rc = L.atn_executor_create(p.handle, C.byref(o), C.byref(h))
    assert rc != 0 and b"GEMM kernel" in L.atn_last_error()


def test_linear_and_ssa_paths_round_trip():
    for seed in range(10):
        net = W.random_connected(6, 3, [2, 3], 1, seed)
        ssa = P.greedy(net)
        assert P.linear_to_ssa(P.ssa_to_linear(ssa, 6), 6) == [tuple(p) for p in ssa]


def test_planner_simulation_agrees_with_oracle():
    net, _, _ = W.random_circuit_amplitude(4, 4, 8, 0)
    path = P.greedy(net)
    st, _ = P.simulate_path_full(net, path)
    ref = O.simulate(net, path)
    assert abs(st.log10_flops - ref["log10_flops"]) < 1e-12
    assert abs(st.log2_peak_size - ref["log2_peak_size"]) < 1e-9


def test_slice_finder_meets_target():
    net, _, _ = W.random_circuit_amplitude(4, 4, 8, 0)
    path = P.greedy(net)
    for target in (2 ** 6, 2 ** 8, 2 ** 10):
        sl = P.find_slices_to_size(net, path, target)
        sd = dict(net["size_dict"])
        for l in sl:
            sd[l] = 1
        _, legs = P.simulate_path_full(dict(net, size_dict=sd), path)
        assert all(np.prod([sd[l] for l in ls]) <= target for ls in legs)


def test_slice_range_validated_before_device_work():
    L = _lib.lib()
    p = Plan([[0], [0]], [], {0: 4}, [(0, 1)], [0])
    h = C.c_void_p()
    o = _lib.ExecOptions(1, 0, 0, 0, 0, 0, 0, 0, 1, 3)  # plain form
    assert L.atn_executor_create(p.handle, C.byref(o), C.byref(h)) != 0
    assert b"batched" in L.atn_last_error()
    o = _lib.ExecOptions(1, 0, 1, 0, 0, 0, 0, 0, 2, 9)  # beyond the 4 slices
    assert L.atn_executor_create(p.handle, C.byref(o), C.byref(h)) != 0
    assert b"out of bounds" in L.atn_last_error()


def test_device_arguments_checked_before_the_abi():
    """Plan.run / profile and the autograd contraction check count, dtype, shape,
    contiguity and device of every tensor (and of a caller's output) before any
    pointer reaches the C ABI; these checks need no GPU."""
    import torch

    from arctn_amd import autograd

    net = {"inputs": [[0, 1], [1, 2]], "output": [0, 2], "size_dict": {0: 2, 1: 3, 2: 4}}
    plan = Plan(net["inputs"], net["output"], net["size_dict"], [(0, 1)])
    a, b = torch.zeros(2, 3, dtype=torch.complex64), torch.zeros(3, 4, dtype=torch.complex64)
    with pytest.raises(ValueError, match="张量个数"):
        plan.check_device_args([a], 2)
    with pytest.raises(ValueError, match="not a GPU tensor"):
        plan.check_device_args([a, b], 2)
    with pytest.raises(ValueError, match="张量个数"):
        autograd.check_inputs(plan, [a])
    with pytest.raises(ValueError, match="dtype"):
        autograd.check_inputs(plan, [a, b.to(torch.complex128)])
    with pytest.raises(ValueError, match="shape"):
        autograd.check_inputs(plan, [a, torch.zeros(4, 3, dtype=torch.complex64)])
    with pytest.raises(ValueError, match="GPU"):
        autograd.check_inputs(plan, [a, b])
    with pytest.raises(ValueError, match="unsupported dtype code"):
        plan.check_device_args([a, b], 7)


def test_contract_table_arguments_checked_before_the_abi():
    """contract_table checks the index against the plan (count, rows and fixed tensors
    that exist, row and fixed shapes), the fixed tensors' dtype and the device before
    any pointer reaches the C ABI; these checks need no GPU."""
    import torch

    from arctn_amd.autograd import contract_table

    net = {"inputs": [[0, 1], [1, 2], [2]], "output": [0], "size_dict": {0: 2, 1: 2, 2: 2}}
    plan = Plan(net["inputs"], net["output"], net["size_dict"], [(0, 1), (2, 3)])
    table = torch.zeros(3, 2, 2, dtype=torch.complex64)
    v = torch.zeros(2, dtype=torch.complex64)
    with pytest.raises(ValueError, match="张量个数"):
        contract_table(plan, table, [0, 1], [v])
    with pytest.raises(ValueError, match="does not exist"):
        contract_table(plan, table, [0, 3, -1], [v])
    with pytest.raises(ValueError, match="does not exist"):
        contract_table(plan, table, [0, 1, -2], [v])
    with pytest.raises(ValueError, match="shape"):
        contract_table(plan, table, [0, 1, 2], [v])  # a 2x2 row where the plan has a vector
    with pytest.raises(ValueError, match="shape"):
        contract_table(plan, table, [0, 1, -1], [torch.zeros(3, dtype=torch.complex64)])
    with pytest.raises(ValueError, match="dtype"):
        contract_table(plan, table, [0, 1, -1], [v.to(torch.complex128)])
    with pytest.raises(ValueError, match="GPU"):
        contract_table(plan, table, [0, 1, -1], [v])


def test_brickwork_gate_tensors_and_their_gradient():
    """The torch-built inputs of the doubled brickwork network equal the NumPy
    ones, and the gradient through the batched gate construction (one node hands
    out every 2x2 gate, ket and bra halves) matches finite differences
    (torch.autograd.gradcheck, float64, CPU)."""
    import torch

    net, slots = W.brickwork_structure(4, 3)
    ang = W.brickwork_angles(4, 3, 1)
    th = torch.tensor(ang, dtype=torch.float64, requires_grad=True)
    got = W.brickwork_torch_tensors(slots, th)
    for t, want in zip(got, W.brickwork_arrays(slots, ang)):
        np.testing.assert_array_equal(t.detach().numpy(), want)
    gen = torch.Generator().manual_seed(3)
    w = [torch.randn(t.shape, dtype=torch.complex128, generator=gen) for t in got]

    def f(a):
        return sum((wi * t).real.sum() for wi, t in zip(w, W.brickwork_torch_tensors(slots, a)))

    assert torch.autograd.gradcheck(f, (th,))
