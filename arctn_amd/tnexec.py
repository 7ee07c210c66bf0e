"""`python -m arctn_amd.tnexec`: run a network file on the GPU, tnexec style.

Mirrors the execution half of src/bin/tnexec.rs (flags :100-380, run :560-640,
the EXEC line :729-775): the network comes from a `.net.json` file, the path and
slice set from a saved version 2 execution plan (``--load-path``) or from the
restated greedy planner (``--method greedy`` with ``--target-size``), the inputs
from a little-endian binary file (``--data``) or seeded random normals. The
contraction runs on the HIP executor. ``--check`` compares it on the CPU, in
double precision (f64 / complex128), with the reference's choice of check
(src/bin/tnexec.rs:23, 505-545, 646-683): NumPy's naive einsum for networks of
at most 8 tensors and 2^20 output elements ("naive"), else for sliced plans of
at most 2^12 slices and 2^22 output elements the unsliced contraction along the
plan's SSA path by NumPy pairwise einsums ("slice_oracle"), each only when its
live set fits 512 MiB; the bar is max |diff| <= 1e-6 x max(1, max |reference|).
One ``EXEC,...`` line on stdout; exit status 1 when the check fails, 2 on bad
input. ``--save-output`` writes the result tensor (little-endian, row-major).
"""

from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np

from . import planner as _planner
from .execution_plan import ExecutionPlan, PlanValidationError
from .netio import load_network_json, load_tensors_bin, save_tensors_bin

CHECK_RELATIVE_TOLERANCE = 1e-6
FULL_CHECK_BYTE_BUDGET = 512 << 20


def _einsum_pair(a, la, b, lb, keep):
    """np.einsum of two tensors by leg lists (legs renumbered into einsum's 52 labels)."""
    ids = {l: i for i, l in enumerate(dict.fromkeys(list(la) + list(lb) + list(keep)))}
    if len(ids) > 52:
        raise ValueError("step has more than 52 distinct legs")
    return np.einsum(a, [ids[l] for l in la], b, [ids[l] for l in lb], [ids[l] for l in keep], optimize=False)


def numpy_path_contract(net, arrays, ssa_path):
    """The unsliced contraction along an SSA path on the CPU with NumPy: each step
    keeps the legs still held by a live tensor or the output (hyperedges stay batch
    legs), then the result is summed / permuted to the output order."""
    slots = [(np.asarray(a), list(legs)) for a, legs in zip(arrays, net["inputs"])]
    alive = [True] * len(slots)
    out = list(net["output"])
    for i, j in ssa_path:
        (a, la), (b, lb) = slots[i], slots[j]
        alive[i] = alive[j] = False
        held = set(out)
        for k, (_, lk) in enumerate(slots):
            if alive[k]:
                held.update(lk)
        keep = [l for l in dict.fromkeys(la + lb) if l in held]
        slots.append((_einsum_pair(a, la, b, lb, keep), keep))
        alive.append(True)
    (x, lx), = [s for s, live in zip(slots, alive) if live]
    ids = {l: i for i, l in enumerate(dict.fromkeys(lx + out))}
    return np.einsum(x, [ids[l] for l in lx], [ids[l] for l in out], optimize=False)


def cpu_check(net, arrays, plan, result, is_complex):
    """(check string, failed) following src/bin/tnexec.rs:505-545, 646-683."""
    wide = np.complex128 if is_complex else np.float64
    n_slices = math.prod(net["size_dict"][l] for l in plan.sliced_legs)
    stats, _ = _planner.simulate_path_full(net, list(plan.ssa_path))
    live = 2.0 ** stats.log2_peak_size + result.size + 2 * sum(a.size for a in arrays)
    fits = live * np.dtype(wide).itemsize <= FULL_CHECK_BYTE_BUDGET
    ref, label = None, None
    if len(net["inputs"]) <= 8 and result.size <= (1 << 20) and fits:
        label = "naive"
        space = sum(math.log2(d) for d in net["size_dict"].values())
        if space <= 28:
            ids = {l: i for i, l in enumerate(sorted(net["size_dict"]))}
            ops = []
            for a, legs in zip(arrays, net["inputs"]):
                ops += [a.astype(wide), [ids[l] for l in legs]]
            ref = np.einsum(*ops, [ids[l] for l in net["output"]], optimize=False)
    elif plan.sliced_legs and result.size <= (1 << 22) and n_slices <= (1 << 12) and fits:
        label = "slice_oracle"
        ref = numpy_path_contract(net, [a.astype(wide) for a in arrays], plan.ssa_path)
    if ref is None:
        return "too_large_for_oracle", False
    ref = np.asarray(ref, dtype=wide)
    diff = float(np.max(np.abs(result.astype(wide) - ref))) if ref.size else 0.0
    scale = max(1.0, float(np.max(np.abs(ref))) if ref.size else 1.0)
    tol = CHECK_RELATIVE_TOLERANCE * scale
    failed = not (math.isfinite(diff) and math.isfinite(scale) and diff <= tol)
    return (f"{'FAIL' if failed else 'PASS'}({label}_max_abs_diff={diff:.2e};scale={scale:.2e};"
            f"tolerance={tol:.2e};oracle_dtype={'c64' if is_complex else 'f64'})"), failed


def _args(argv):
    ap = argparse.ArgumentParser(prog="python -m arctn_amd.tnexec")
    ap.add_argument("--net", required=True, help=".net.json network")
    ap.add_argument("--load-path", help="saved version 2 execution plan")
    ap.add_argument("--save-path", help="write the executed plan as a version 2 artifact")
    ap.add_argument("--method", choices=["greedy", "rgreedy"], help="plan with the restated planner")
    ap.add_argument("--trials", type=int, default=8, help="rgreedy trials")
    ap.add_argument("--target-size", type=int, help="max intermediate elements per slice (with --method)")
    ap.add_argument("--data", help="little-endian input tensors (net.inputs order, row-major)")
    ap.add_argument("--seed", type=int, default=0, help="seed of the random inputs when --data is absent")
    ap.add_argument("--single", action="store_true", help="f32 / c32 instead of f64 / c64")
    ap.add_argument("--complex", action="store_true", help="complex scalars")
    ap.add_argument("--precision", default="native", choices=["native", "fp32", "c3m", "bf16x3", "fp16x2", "bf16x3_3m", "fp16x2_3m"])
    ap.add_argument("--plain", action="store_true", help="per-slice executor instead of the batched one")
    ap.add_argument("--graph", action="store_true", help="batched form captured in a hipGraph")
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--check", action="store_true", help="compare with a double-precision CPU contraction (NumPy)")
    ap.add_argument("--save-output", help="write the result tensor (little-endian, row-major)")
    return ap.parse_args(argv)


def _dtype(a):
    name = {(True, False): "f32", (True, True): "c32", (False, False): "f64", (False, True): "c64"}
    key = name[(a.single, a.complex)]
    return key, {"f32": np.float32, "c32": np.complex64, "f64": np.float64, "c64": np.complex128}[key]


def main(argv=None):
    a = _args(sys.argv[1:] if argv is None else argv)
    if a.load_path and a.method:
        print("--load-path 已交付精确路径，不能再与 --method 或 --trials 并用", file=sys.stderr)
        return 2
    if not a.load_path and not a.method:
        print("需要 --load-path 或 --method", file=sys.stderr)
        return 2
    try:
        net, labels = load_network_json(a.net)
    except (OSError, ValueError) as e:
        print(f"载入网络 {a.net!r} 失败: {e}", file=sys.stderr)
        return 2
    try:
        if a.load_path:
            plan = ExecutionPlan.load(a.load_path)
            plan.validate(net["inputs"], net["output"], net["size_dict"])
            if plan.leg_labels != tuple(labels) and any(isinstance(x, str) for x in plan.leg_labels):
                raise PlanValidationError("plan network_leg_labels does not match the current network")
            method, verification = "load-path", "verified-v2"
        else:
            path = (_planner.greedy(net) if a.method == "greedy"
                    else _planner.random_greedy(net, a.trials, a.seed))
            sliced = (_planner.find_slices_to_size(net, path, a.target_size) or []) if a.target_size else []
            plan = ExecutionPlan.from_network(net["inputs"], net["output"], net["size_dict"], path, sliced,
                                              target_size=a.target_size, leg_labels=labels, source="arctn_amd.tnexec")
            method, verification = a.method, "planned"
    except (OSError, ValueError) as e:
        print(f"计划不可执行: {e}", file=sys.stderr)
        return 2
    if a.save_path:
        plan.save(a.save_path)
    dname, dtype = _dtype(a)
    if a.data:
        try:
            arrays = load_tensors_bin(a.data, net, dtype)
        except (OSError, ValueError) as e:
            print(f"读取 --data {a.data!r} 失败: {e}", file=sys.stderr)
            return 2
    else:
        rng = np.random.default_rng(a.seed)
        arrays = []
        for legs in net["inputs"]:
            shape = tuple(net["size_dict"][l] for l in legs)
            x = rng.standard_normal(shape)
            if a.complex:
                x = x + 1j * rng.standard_normal(shape)
            arrays.append(x.astype(dtype))

    import torch

    dev = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in arrays]
    opts = dict(precision=a.precision, batched=not a.plain, use_graph=a.graph and not a.plain, autotune=a.autotune)
    plan.execute(dev, **opts)  # plan build, autotune, graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = plan.execute(dev, **opts)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    result = out.cpu().numpy()

    stats, _ = _planner.simulate_path_full(net, list(plan.ssa_path))
    sd1 = dict(net["size_dict"])
    for leg in plan.sliced_legs:
        sd1[leg] = 1
    per, _ = _planner.simulate_path_full(dict(net, size_dict=sd1), list(plan.ssa_path))
    n_slices = math.prod(net["size_dict"][l] for l in plan.sliced_legs)
    log10_total = per.log10_flops + math.log10(n_slices)
    gflops = 10.0 ** log10_total / secs / 1e9 if secs > 0 else float("nan")

    check, failed = "skipped", False
    if a.check:
        check, failed = cpu_check(net, arrays, plan, result, a.complex)
    if a.save_output:
        save_tensors_bin(a.save_output, [result])

    target = plan.target_size
    print(f"EXEC,net={a.net},dtype={dname},method={method},backend=hip,precision={a.precision},"
          f"form={'plain' if a.plain else 'batched' + ('+graph' if opts['use_graph'] else '')},"
          f"log10_flops={stats.log10_flops:.4f},log2_peak={stats.log2_max_size:.4f},"
          f"log2_max_size={stats.log2_max_size:.4f},log2_max_contraction_size={stats.log2_max_contraction_size:.4f},"
          f"log2_peak_size={stats.log2_peak_size:.4f},dtype_itemsize={np.dtype(dtype).itemsize},"
          f"target_size={target if target is not None else 'null'},"
          f"execution_log10_flops_total={log10_total:.4f},per_slice_log2_max_size={per.log2_max_size:.4f},"
          f"per_slice_log2_max_contraction_size={per.log2_max_contraction_size:.4f},"
          f"per_slice_log2_peak_size={per.log2_peak_size:.4f},n_sliced_legs={len(plan.sliced_legs)},"
          f"n_slices={n_slices},plan_verification={verification},out_numel={result.size},t_exec_s={secs:.4f},"
          f"GFLOPS={gflops:.3f},check={check}", flush=True)
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
