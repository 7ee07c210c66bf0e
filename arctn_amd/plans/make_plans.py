"""Regenerate the cached plans (SSA path + sliced legs) of the circuit workloads.

Usage: python -m arctn_amd.plans.make_plans [name ...]

Each plan is computed once on the CPU (planning is not on the GPU path) and
stored as JSON next to this file; the workload network itself is rebuilt
from its seed by arctn_amd.workloads when the plan is loaded.
"""
import json
import os
import sys
import time

from arctn_amd import planner, workloads

HERE = os.path.dirname(os.path.abspath(__file__))

PLANS = {
    # 16-qubit 4x4 grid, 8 cycles, one amplitude, deterministic greedy, unsliced.
    "circuit_4x4_c8": dict(rows=4, cols=4, cycles=8, seed=0, open=0, target=None, trials=0),
    # 30-qubit 5x6 grid, 14 cycles, one amplitude, target 2^24 elements.
    "circuit_5x6_c14_t24": dict(rows=5, cols=6, cycles=14, seed=0, open=0, target=2**24, trials=40),
    # 36-qubit 6x6 grid, 12 cycles, 6 open qubits (64 amplitudes), target 2^26. All six open
    # legs are output-sliced (64 blocks, src/slice.rs:1070-1173); the path is searched on the
    # block network (open legs fixed) and the slice finder adds internal legs from there.
    "circuit_6x6_c12_o6_t26": dict(rows=6, cols=6, cycles=12, seed=0, open=6, target=2**26, trials=24,
                                   slice_open=True),
}
# 512 random 3-regular networks, n uniform over even [40, 120], dims {2, 3, 4},
# float32 N(0, 1) entries, scalar output; per-instance random-greedy path, seed = instance id.
UNEVEN = {
    "uneven_3regular_512": dict(n_inst=512, seed=0, n_lo=40, n_hi=120, dims=[2, 3, 4], trials=8),
}


def build_uneven(name):
    p = UNEVEN[name]
    t0 = time.time()
    nets = workloads.uneven_networks(p["n_inst"], p["seed"], p["n_lo"], p["n_hi"], p["dims"])
    paths, costs = [], []
    for i, (net, _) in enumerate(nets):
        path, st = planner.random_greedy(net, p["trials"], i)
        paths.append([list(x) for x in path])
        costs.append(st.log10_flops)
    doc = dict(name=name, workload=p, ssa_paths=paths, log10_mults=costs, planning_seconds=time.time() - t0)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(doc, f, separators=(",", ":"))
    print(name, "instances", len(paths), "planning s", round(time.time() - t0, 1), flush=True)


def load_uneven(name, dtype=None):
    """Returns (list of (net, arrays, ssa_path), plan_doc)."""
    import numpy as np

    with open(os.path.join(HERE, name + ".json")) as f:
        doc = json.load(f)
    p = doc["workload"]
    nets = workloads.uneven_networks(p["n_inst"], p["seed"], p["n_lo"], p["n_hi"], p["dims"],
                                     dtype or np.float32)
    return [(net, arrays, [tuple(x) for x in path]) for (net, arrays), path in zip(nets, doc["ssa_paths"])], doc


def build(name):
    if name in UNEVEN:
        return build_uneven(name)
    p = PLANS[name]
    net, _, _ = workloads.random_circuit_amplitude(
        p["rows"], p["cols"], p["cycles"], p["seed"], open_qubits=range(p["open"]))
    t0 = time.time()
    if p["target"] is None:
        path, sliced = planner.greedy(net), []
        cost = planner.sliced_cost_log10(net, path, sliced)
    elif p.get("slice_open"):
        block = {"inputs": net["inputs"], "output": [],
                 "size_dict": {l: (1 if l in net["output"] else d) for l, d in net["size_dict"].items()}}
        path, internal, _ = planner.random_greedy_sliced(block, p["trials"], p["seed"], p["target"])
        sliced = sorted(set(internal) | set(net["output"]))
        cost = planner.sliced_cost_log10(net, path, sliced)
    else:
        path, sliced, cost = planner.random_greedy_sliced(
            net, p["trials"], p["seed"], p["target"], allow_output_slicing=p["open"] > 0)
    doc = dict(name=name, workload=p, ssa_path=[list(x) for x in path], sliced_legs=sliced,
               log10_total_mults=cost, planning_seconds=time.time() - t0)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(doc, f)
    print(name, "sliced", len(sliced), "log10 total mults", round(cost, 3), flush=True)


def load(name):
    """Returns (net, arrays, plan_doc) with the network rebuilt from its seed."""
    with open(os.path.join(HERE, name + ".json")) as f:
        doc = json.load(f)
    p = doc["workload"]
    net, arrays, _ = workloads.random_circuit_amplitude(
        p["rows"], p["cols"], p["cycles"], p["seed"], open_qubits=range(p["open"]))
    return net, arrays, doc


if __name__ == "__main__":
    for n in sys.argv[1:] or list(PLANS) + list(UNEVEN):
        build(n)
