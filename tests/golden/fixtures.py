"""Loader helpers for reference_fixtures.json (inputs + expected values only)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    with open(os.path.join(HERE, "reference_fixtures.json")) as f:
        return json.load(f)


def net_of(fx):
    return {"inputs": fx["inputs"], "output": fx["output"], "size_dict": {int(k): v for k, v in fx["size_dict"].items()}}


def einsum_case(case):
    """(name, equation, net, sliced_legs, ssa_path) of an output-slicing case."""
    name, equation, dims, sliced, path = case
    terms, out = equation.split("->")
    terms = terms.split(",")
    labels = {lab: i for i, lab in enumerate(dict.fromkeys("".join(terms)))}
    net = {"inputs": [[labels[c] for c in t] for t in terms], "output": [labels[c] for c in out],
           "size_dict": {labels[c]: d for c, d in dims.items()}}
    return name, equation, net, [labels[c] for c in sliced], [tuple(p) for p in path]


def case_arrays(net, dtype, seed=871):
    """Standard-normal inputs (real and imaginary parts), like test_output_slicing.py:68-82."""
    rng = np.random.default_rng(seed)
    out = []
    for legs in net["inputs"]:
        shape = tuple(net["size_dict"][l] for l in legs)
        data = rng.standard_normal(shape)
        if np.issubdtype(np.dtype(dtype), np.complexfloating):
            data = data + 1j * rng.standard_normal(shape)
        out.append(np.array(data, dtype=dtype, order="C"))
    return out
