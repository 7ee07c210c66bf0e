"""Writes tests/golden/uneven_3regular_512.json: the CPU oracle's value of every
instance of the 512-network uneven workload, in float64 and in float32.

Usage: python -m tests.golden.make_uneven_fixture [--threads 8]

Workload: 512 random 3-regular networks (vertex count uniform over even
[40, 120], per-edge dims uniform over {2, 3, 4}, N(0, 1) float32 entries,
scalar output), each with its cached random-greedy path (seed = instance id),
as arctn_amd.plans.make_plans.load_uneven builds them.

Per instance the oracle (CPU restatement of CompiledContraction::execute,
src/compiled.rs:369-441, and of contract_network_sliced, src/slice.rs:960-1065)
contracts the float32 inputs promoted to float64, and the float32 inputs as
they are. Instances whose live-set peak exceeds 2^25 elements are contracted
sliced (legs from the restated slice finder, src/slice.rs:254-288, to 2^25
elements, summed over slices with the reference's chunking) so that the
largest ones fit host memory; both dtypes use the same slicing. The float64
value is the reference value; the float32 value measures how much error the
reference's own float32 arithmetic makes on that instance (the scalar of a
random network comes out of heavy cancellation on some instances).

The scalars reach ~1e40, beyond float32's range, so the float32 run first
multiplies every input by the same power of two 2^-s, s = round(log2|v64| /
n_inputs), and multiplies the result back by 2^(s * n_inputs) in float64.
Power-of-two scaling is exact: away from overflow and subnormals it changes
no rounding, so the float32 error measured is the float32 arithmetic's own.

Three more float32 runs per instance ("f32_rescaled") contract the inputs
multiplied by c in {1.15, 1.3, 1.45} (each product x * c rounded to float32,
then the power-of-two shift) and divide the result by c^n_inputs in float64.
The exact value is unchanged, but every rounding in the contraction falls
differently: each run is a fresh sample of the float32 arithmetic's error on
that instance (plus one rounding of every input, the same size as one step's).
One sample alone says little about an instance's error scale, four give an
estimate of it (their root mean square) to within a small factor.

The file is rewritten after every instance, so an interrupted run resumes.
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from arctn_amd import planner  # noqa: E402
from arctn_amd.plans.make_plans import load_uneven  # noqa: E402
from oracle import pyoracle as O  # noqa: E402

NAME = "uneven_3regular_512"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), NAME + ".json")
SLICE_TARGET = 2 ** 25


def _values(net, arrays, path, sliced, threads):
    """(float64 value, float32 value, input shift) of one instance."""
    vals = []
    shift = 0
    for dt in (np.float64, np.float32):
        a = [np.ldexp(x.astype(dt), -shift).astype(dt) for x in arrays]
        if sliced:
            v = O.contract(net, a, path, "sliced", sliced=sliced, threads=threads)
        else:
            v = O.contract(net, a, path, "compiled")
        vals.append(float(np.ldexp(np.float64(v), shift * len(arrays))))
        if dt == np.float64:
            if not (np.isfinite(vals[0]) and vals[0] != 0.0):
                raise RuntimeError(f"float64 oracle value {vals[0]}")
            shift = int(round(np.log2(abs(vals[0])) / len(arrays)))
    return vals[0], vals[1], shift


RESCALE = (1.15, 1.3, 1.45)


def _rescaled(net, arrays, path, sliced, threads, v64):
    """float32 values of the instance with the inputs multiplied by each c."""
    out = []
    n = len(arrays)
    for c in RESCALE:
        shift = int(round((np.log2(abs(v64)) + n * np.log2(c)) / n))
        a = [np.ldexp((x.astype(np.float32) * np.float32(c)).astype(np.float32), -shift).astype(np.float32)
             for x in arrays]
        if sliced:
            v = O.contract(net, a, path, "sliced", sliced=sliced, threads=threads)
        else:
            v = O.contract(net, a, path, "compiled")
        out.append(float(np.ldexp(np.float64(v), shift * n) / np.float64(c) ** n))
    return out


def _rescaled_job(i, v64):
    inst, _ = load_uneven(NAME)
    net, arrays, path = inst[i]
    return i, _rescaled(net, arrays, path, [], 1, v64)


def _unsliced_job(i):
    inst, _ = load_uneven(NAME)
    net, arrays, path = inst[i]
    t0 = time.time()
    return i, _values(net, arrays, path, [], 1), time.time() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    args = ap.parse_args()
    inst, _ = load_uneven(NAME)
    n = len(inst)
    doc = {"workload": NAME, "slice_target": SLICE_TARGET, "f64": [None] * n, "f32": [None] * n,
           "f32_input_shift": [None] * n, "sliced_legs": [None] * n, "cpu_s": [None] * n,
           "f32_rescaled": [None] * n}
    if os.path.exists(OUT):
        with open(OUT) as f:
            old = json.load(f)
        if old.get("workload") == NAME and len(old["f64"]) == n:
            doc.update(old)
    cost, sliced = [], []
    for net, _, path in inst:
        st = O.simulate(net, path)
        cost.append(st["log10_flops"])
        big = st["log2_peak_size"] > np.log2(SLICE_TARGET)
        sliced.append((planner.find_slices_to_size(net, path, SLICE_TARGET) or []) if big else [])
    t_all = time.time()
    done = [sum(v is not None for v in doc["f64"])]

    def record(i, vals, dt):
        doc["f64"][i], doc["f32"][i], doc["f32_input_shift"][i] = vals
        doc["sliced_legs"][i] = sliced[i]
        doc["cpu_s"][i] = round(dt, 3)
        with open(OUT + ".tmp", "w") as f:
            json.dump(doc, f, separators=(",", ":"))
        os.replace(OUT + ".tmp", OUT)
        done[0] += 1
        rel = abs(vals[1] - vals[0]) / abs(vals[0])
        print(f"[{done[0]}/{n}] instance {i}: log10 mults {cost[i]:.2f}, {len(sliced[i])} sliced legs, "
              f"{dt:.1f} s, f32 rel err {rel:.2e}, total {time.time() - t_all:.0f} s", flush=True)

    todo = [i for i in sorted(range(n), key=lambda i: cost[i]) if doc["f64"][i] is None]
    # unsliced instances: one process per instance (the oracle runs one unsliced
    # contraction on one thread); sliced instances: one at a time over all threads
    pool_jobs = [i for i in todo if not sliced[i]]
    with cf.ProcessPoolExecutor(max_workers=args.threads) as ex:
        for fut in cf.as_completed([ex.submit(_unsliced_job, i) for i in pool_jobs]):
            i, vals, dt = fut.result()
            record(i, vals, dt)
    for i in todo:
        if sliced[i]:
            net, arrays, path = inst[i]
            t0 = time.time()
            record(i, _values(net, arrays, path, sliced[i], args.threads), time.time() - t0)

    def record_rescaled(i, vals):
        doc["f32_rescaled"][i] = vals
        with open(OUT + ".tmp", "w") as f:
            json.dump(doc, f, separators=(",", ":"))
        os.replace(OUT + ".tmp", OUT)
        v = doc["f64"][i]
        errs = [abs(doc["f32"][i] - v) / abs(v)] + [abs(x - v) / abs(v) for x in vals]
        print(f"instance {i}: f32 rel errs " + " ".join(f"{e:.1e}" for e in errs)
              + f", total {time.time() - t_all:.0f} s", flush=True)

    todo = [i for i in sorted(range(n), key=lambda i: cost[i]) if doc["f32_rescaled"][i] is None]
    with cf.ProcessPoolExecutor(max_workers=args.threads) as ex:
        futs = [ex.submit(_rescaled_job, i, doc["f64"][i]) for i in todo if not sliced[i]]
        for fut in cf.as_completed(futs):
            record_rescaled(*fut.result())
    for i in todo:
        if sliced[i]:
            net, arrays, path = inst[i]
            record_rescaled(i, _rescaled(net, arrays, path, sliced[i], args.threads, doc["f64"][i]))


if __name__ == "__main__":
    main()
