"""Writes tests/golden/circuit_fixtures.json: complex128 CPU-oracle values of the
sliced random-circuit workloads, for the GPU parity tests and the benchmark.

Usage: python -m tests.golden.make_circuit_fixtures [--threads 8] [name ...]

Workloads (networks rebuilt from their seeds, plans from arctn_amd/plans/):
  circuit_5x6_c14_t24      30-qubit 5x6 grid, 14 cycles, one amplitude, 256
                           internal slices: the whole amplitude (all slices)
                           and the partial sum over slices [0, 16).
  circuit_6x6_c12_o6_t26   36-qubit 6x6 grid, 12 cycles, six open qubits
                           output-sliced into 64 blocks of 256 internal slices:
                           for eight blocks, the partial sum over internal
                           slices [0, 16) of the block.

The complex64 inputs are promoted to complex128 and contracted by the oracle
(the CPU restatement of contract_network_sliced, src/slice.rs:960-1065, and
contract_slice_range over internal slices). For an output block, the open legs
are fixed to the block's values with np.take (block b = flat output index b,
last open leg fastest, src/slice.rs:1070-1173) and the block network's
internal slices are summed.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from arctn_amd.plans.make_plans import load  # noqa: E402
from oracle import pyoracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "circuit_fixtures.json")
C_BLOCKS = [0, 9, 18, 27, 36, 45, 54, 63]
C_SLICES = 16
B_SLICES = 16


def _c(z):
    z = complex(np.asarray(z).reshape(()))
    return [z.real, z.imag]


def block_network(net, arrays, block):
    """The closed network of output block `block`: every open leg fixed by np.take."""
    vals = np.unravel_index(block, [net["size_dict"][l] for l in net["output"]])
    fix = dict(zip(net["output"], (int(v) for v in vals)))
    inputs, out = [], []
    for legs, a in zip(net["inputs"], arrays):
        for ax in range(len(legs) - 1, -1, -1):
            if legs[ax] in fix:
                a = np.take(a, fix[legs[ax]], axis=ax)
        inputs.append([l for l in legs if l not in fix])
        out.append(np.ascontiguousarray(a))
    sd = {l: d for l, d in net["size_dict"].items() if l not in fix}
    return {"inputs": inputs, "output": [], "size_dict": sd}, out


def circuit_b(threads):
    net, arrays, doc = load("circuit_5x6_c14_t24")
    path = [tuple(p) for p in doc["ssa_path"]]
    a128 = [a.astype(np.complex128) for a in arrays]
    t0 = time.time()
    amp = O.contract(net, a128, path, "sliced", doc["sliced_legs"], threads=threads)
    t1 = time.time()
    part = O.contract_slice_range(net, a128, path, doc["sliced_legs"], 0, B_SLICES, threads=threads)
    print(f"circuit_5x6_c14_t24: amplitude {complex(amp):.12e} ({t1 - t0:.0f} s), "
          f"slices [0, {B_SLICES}) {complex(part):.12e}", flush=True)
    return {"amplitude": _c(amp), "partial_slices": [0, B_SLICES], "partial": _c(part),
            "n_slices": 256, "cpu_s": round(t1 - t0, 1), "threads": threads}


def circuit_c(threads):
    net, arrays, doc = load("circuit_6x6_c12_o6_t26")
    path = [tuple(p) for p in doc["ssa_path"]]
    internal = [l for l in doc["sliced_legs"] if l not in net["output"]]
    a128 = [a.astype(np.complex128) for a in arrays]
    blocks = {}
    for b in C_BLOCKS:
        t0 = time.time()
        bnet, barr = block_network(net, a128, b)
        v = O.contract_slice_range(bnet, barr, path, internal, 0, C_SLICES, threads=threads)
        blocks[str(b)] = _c(v)
        print(f"circuit_6x6_c12_o6_t26: block {b} slices [0, {C_SLICES}) {complex(v):.12e} "
              f"({time.time() - t0:.0f} s)", flush=True)
    n_int = 1
    for l in internal:
        n_int *= net["size_dict"][l]
    return {"blocks": blocks, "internal_slices": [0, C_SLICES], "n_internal": n_int, "n_blocks": 64}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("names", nargs="*")
    args = ap.parse_args()
    doc = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            doc = json.load(f)
    makers = {"circuit_5x6_c14_t24": circuit_b, "circuit_6x6_c12_o6_t26": circuit_c}
    for name in args.names or list(makers):
        doc[name] = makers[name](args.threads)
        with open(OUT + ".tmp", "w") as f:
            json.dump(doc, f, indent=1)
        os.replace(OUT + ".tmp", OUT)


if __name__ == "__main__":
    main()
