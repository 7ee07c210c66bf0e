"""Loader helpers for reference_fixtures.json (inputs + expected values only),
and the per-instance bar of the uneven-networks fixture."""
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


def uneven_bounds(doc):
    """Per-instance bar of uneven_3regular_512.json: max(1e-5, 16 s_i), s_i the
    root mean square of the relative errors of four CPU float32 contractions of
    instance i (the fixture's float32 run and three with rescaled inputs, each a
    fresh sample of where float32 rounding falls; the scalar of a random
    network comes out of heavy cancellation on some instances, where float32
    arithmetic itself is off by 1e-4 and more). A float32 contraction with its
    own summation order is one more sample of that error: under a normal model
    with s_i estimated from four samples it exceeds 16 s_i with probability
    about 1e-4 per instance, while a wrong step or a dropped term errs by orders
    of magnitude more."""
    out = []
    for v64, v32, rs in zip(doc["f64"], doc["f32"], doc["f32_rescaled"]):
        errs = [abs(x - v64) / abs(v64) for x in [v32] + list(rs)]
        out.append(max(1e-5, 16.0 * float(np.sqrt(np.mean(np.square(errs))))))
    return out
