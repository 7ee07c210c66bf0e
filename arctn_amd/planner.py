"""CPU plan tooling: greedy SSA path search, path simulation and slice finding.

Planning is not part of the GPU executor; it only produces the plan input
(SSA path + sliced legs) that the executor consumes, exactly as the
reference's planners do. These are restatements of the reference's open
planners so that the benchmark workloads get the same plans the reference
would compute with `greedy` + `find_slices_to_size`:

* ``simulate_path_full`` follows src/path.rs:239-434 (FLOPs = product of the
  union of operand dims per step, :358-361; live-set peak, :396-399).
* ``greedy`` follows src/paths/greedy.rs:263-320 (costmod 1, temperature 0,
  one branch, min-heap ordered by (cost, a, b), smallest-first outer
  products for disconnected components).
* ``find_slices_to_size`` follows src/slice.rs:254-288 with the leg choice of
  src/slice.rs:557-589 (most over-target intermediates, then largest dim,
  then smallest leg id) and the exact element-count feasibility test of
  src/slice.rs:46-96.

A network is a dict ``{"inputs": [[leg, ...], ...], "output": [leg, ...],
"size_dict": {leg: dim}}`` with integer legs.
"""

from __future__ import annotations

import heapq
import math
from dataclasses import dataclass


def _log2_dim(size_dict, leg):
    return math.log2(size_dict[leg])


def sorted_dedup(legs):
    return sorted(set(legs))


def validate_network(net):
    """Structural checks of src/network.rs:171-229 (subset needed by planning)."""
    inputs, output, sd = net["inputs"], net["output"], net["size_dict"]
    if not inputs:
        raise ValueError("network needs at least one input tensor")
    held = set()
    for i, legs in enumerate(inputs):
        counts = {}
        for leg in legs:
            if leg not in sd:
                raise ValueError(f"leg {leg} has no size_dict entry")
            counts[leg] = counts.get(leg, 0) + 1
            held.add(leg)
        for leg, c in counts.items():
            if c > 2:
                raise ValueError(f"leg {leg} appears {c} times in input {i}; at most 2 supported")
    if len(set(output)) != len(output):
        raise ValueError(f"output has repeated legs: {output}")
    for leg in output:
        if leg not in held:
            raise ValueError(f"output leg {leg} is not held by any input")
    for leg, d in sd.items():
        if d <= 0:
            raise ValueError(f"leg {leg} has dimension {d}")
    unused = sorted(l for l in sd if l not in held)
    if unused:
        raise ValueError(f"size_dict leg {unused[0]} is not used by any input")


@dataclass
class PathStats:
    log10_flops: float
    log2_max_size: float
    log2_max_contraction_size: float
    log2_total_size: float
    log2_read_write: float
    log2_peak_size: float


def _logaddexp2(a, b):
    # src/path.rs:32-42
    if a == -math.inf:
        return b
    if b == -math.inf:
        return a
    hi, lo = (a, b) if a > b else (b, a)
    return hi + math.log2(1.0 + 2.0 ** (lo - hi))


def simulate_path_full(net, path):
    """Replay an SSA path; returns (PathStats, step_result_legs). src/path.rs:239-434."""
    validate_network(net)
    inputs, output, sd = net["inputs"], net["output"], net["size_dict"]
    n = len(inputs)
    in_output = set(output)
    appearances = {}
    for term in inputs:
        for leg in term:
            appearances[leg] = appearances.get(leg, 0) + 1
    for leg in output:
        appearances[leg] = appearances.get(leg, 0) + 1
    legs = []
    refcount = {}
    contraction_slot = []
    rw_slot = []
    for t in inputs:
        d = sorted(t)
        leaf = 0.0
        i = 0
        while i < len(d):
            j = i + 1
            while j < len(d) and d[j] == d[i]:
                j += 1
            if j - i != appearances[d[i]]:
                leaf += _log2_dim(sd, d[i])
            i = j
        dd = sorted_dedup(t)
        for leg in dd:
            refcount[leg] = refcount.get(leg, 0) + 1
        legs.append(dd)
        contraction_slot.append(leaf)
        rw_slot.append(sum(_log2_dim(sd, l) for l in t))

    def size_of(ls):
        return 2.0 ** sum(_log2_dim(sd, l) for l in ls)

    log2_flops = -math.inf
    log2_max_size = sum(_log2_dim(sd, l) for l in output) if n == 1 else 0.0
    log2_max_contraction = 0.0
    log2_total = -math.inf
    log2_rw = -math.inf
    slot_size = [size_of(t) for t in inputs]
    alive_sum = sum(slot_size)
    peak = alive_sum
    step_legs = []
    for step, (a, b) in enumerate(path):
        if a == b:
            raise ValueError(f"step {step} contracts a tensor with itself")
        if a >= len(legs) or b >= len(legs) or legs[a] is None or legs[b] is None:
            raise ValueError(f"step {step} references an invalid or consumed tensor")
        la, lb = legs[a], legs[b]
        legs[a] = None
        legs[b] = None
        union = sorted(set(la) | set(lb))
        log2_flops = _logaddexp2(log2_flops, sum(_log2_dim(sd, l) for l in union))
        for l in la:
            refcount[l] -= 1
        for l in lb:
            refcount[l] -= 1
        result = [l for l in union if refcount[l] > 0 or l in in_output]
        for l in result:
            refcount[l] += 1
        size_log2 = sum(_log2_dim(sd, l) for l in result)
        log2_max_size = max(log2_max_size, size_log2)
        c = _logaddexp2(_logaddexp2(contraction_slot[a], contraction_slot[b]), size_log2)
        log2_max_contraction = max(log2_max_contraction, c)
        rw = _logaddexp2(_logaddexp2(rw_slot[a], rw_slot[b]), size_log2)
        log2_rw = _logaddexp2(log2_rw, rw)
        contraction_slot.append(size_log2)
        rw_slot.append(size_log2)
        log2_total = _logaddexp2(log2_total, size_log2)
        rsize = 2.0 ** size_log2
        peak = max(peak, alive_sum + rsize)
        alive_sum += rsize - slot_size[a] - slot_size[b]
        slot_size.append(rsize)
        step_legs.append(result)
        legs.append(result)
    alive = [i for i, l in enumerate(legs) if l is not None]
    if len(alive) != 1:
        raise ValueError(f"incomplete path: {len(alive)} tensors remain")
    fin = legs[alive[0]]
    out = sorted_dedup(output)
    if not path:
        if not all(l in fin for l in out):
            raise ValueError("final legs do not contain every output leg")
    elif fin != out:
        raise ValueError(f"final legs {fin} do not match output {out}")
    return (
        PathStats(
            log10_flops=log2_flops * math.log10(2.0),
            log2_max_size=log2_max_size,
            log2_max_contraction_size=log2_max_contraction,
            log2_total_size=log2_total,
            log2_read_write=log2_rw,
            log2_peak_size=math.log2(peak),
        ),
        step_legs,
    )


def _result_legs(la, lb, refcount, in_output):
    # src/path.rs:74-94: union of sorted legs kept if held elsewhere or open.
    sa, sb = set(la), set(lb)
    out = []
    for l in sorted(sa | sb):
        rc = refcount[l] - (l in sa) - (l in sb)
        if rc > 0 or l in in_output:
            out.append(l)
    return out


def greedy(net, costmod=1.0, temperature=0.0, nbranch=1, cost_fn=("mem",), rng=None):
    """Greedy SSA path (src/paths/greedy.rs:263-320).

    With the defaults this is the reference's deterministic ``greedy``
    (greedy.rs:307-320). ``temperature > 0`` samples among the first
    ``nbranch`` live candidates with Boltzmann weights (greedy.rs:203-257);
    ``cost_fn`` selects the variant of greedy.rs:17-29: ("mem",),
    ("skew", sk), ("ratio",), ("logratio",), ("hyperdeg", hd).
    """
    validate_network(net)
    inputs, output, sd = net["inputs"], net["output"], net["size_dict"]
    in_output = set(output)
    node_legs = []
    node_size = []
    refcount = {}
    holders = {}
    for i, t in enumerate(inputs):
        d = sorted_dedup(t)
        for l in d:
            refcount[l] = refcount.get(l, 0) + 1
            holders.setdefault(l, set()).add(i)
        node_legs.append(d)
        node_size.append(2.0 ** sum(_log2_dim(sd, l) for l in d))
    heap = []

    def legs_size(ls):
        return 2.0 ** sum(_log2_dim(sd, l) for l in ls)

    def push(a, b):
        res = _result_legs(node_legs[a], node_legs[b], refcount, in_output)
        rsize = legs_size(res)
        sa, sb = node_size[a], node_size[b]
        base = rsize - costmod * (sa + sb)
        kind = cost_fn[0]
        if kind == "mem":
            cost = base
        elif kind == "skew":
            cost = base + cost_fn[1] * abs(sa - sb)
        elif kind == "ratio":
            cost = base / (rsize + 1.0)
        elif kind == "logratio":
            cost = math.log2(rsize + 2.0) / math.log2(0.65 * (sa + sb) + 2.0)
        else:
            cost = base + cost_fn[1] * sum(refcount[l] for l in res)
        heapq.heappush(heap, (cost, a, b))

    seen = set()
    for l in sorted(holders):
        v = sorted(holders[l])
        for i in range(len(v)):
            for j in range(i + 1, len(v)):
                if (v[i], v[j]) not in seen:
                    seen.add((v[i], v[j]))
                    push(v[i], v[j])

    def alive(i):
        return i < len(node_legs) and node_legs[i] is not None

    def contract(a, b):
        res = _result_legs(node_legs[a], node_legs[b], refcount, in_output)
        la, lb = node_legs[a], node_legs[b]
        node_legs[a] = None
        node_legs[b] = None
        for l in la:
            refcount[l] -= 1
            holders[l].discard(a)
        for l in lb:
            refcount[l] -= 1
            holders[l].discard(b)
        new_id = len(node_legs)
        for l in res:
            refcount[l] += 1
            holders[l].add(new_id)
        node_legs.append(res)
        node_size.append(legs_size(res))
        nbrs = set()
        for l in res:
            for h in holders[l]:
                if h != new_id:
                    nbrs.add(h)
        for nb in sorted(nbrs):
            push(min(nb, new_id), max(nb, new_id))
        return new_id

    path = []
    while True:
        choices = []
        while len(choices) < max(nbranch, 1) and heap:
            c = heapq.heappop(heap)
            if alive(c[1]) and alive(c[2]):
                choices.append(c)
        if not choices:
            break
        idx = 0
        if temperature > 0.0 and len(choices) > 1:
            cmin = choices[0][0]
            t = temperature * max(abs(cmin), 1.0)
            w = [math.exp(-(c[0] - cmin) / t) for c in choices]
            total = sum(w)
            if math.isfinite(total) and total > 0.0:
                r = rng.uniform(0.0, total)
                for i, wi in enumerate(w):
                    if r < wi:
                        idx = i
                        break
                    r -= wi
        chosen = choices.pop(idx)
        for other in choices:
            heapq.heappush(heap, other)
        contract(chosen[1], chosen[2])
        path.append((chosen[1], chosen[2]))
    rest = [i for i in range(len(node_legs)) if alive(i)]
    while len(rest) > 1:
        rest.sort(key=lambda x: node_size[x])  # stable, like Rust sort_by
        a, b = rest[0], rest[1]
        a, b = min(a, b), max(a, b)
        new_id = contract(a, b)
        path.append((a, b))
        rest = rest[2:] + [new_id]
    return path


def _legs_feasible(sd, legs, target):
    # src/slice.rs:62-74: divide the remaining allowance, never overflow.
    remaining = target
    for l in legs:
        d = sd[l]
        if d > remaining:
            return False
        remaining //= d
    return True


def _slice_leg_allowed(net, leg, allow_output_slicing):
    # src/slice.rs:136-144
    sd = net["size_dict"]
    if sd.get(leg, 0) <= 0:
        return False
    if not allow_output_slicing and leg in net["output"]:
        return False
    if not any(leg in legs for legs in net["inputs"]):
        return False
    return not any(legs.count(leg) > 1 for legs in net["inputs"])


def find_slices_to_size(net, path, target_size, allow_output_slicing=False, presliced=()):
    """Greedy slice-leg selection (src/slice.rs:254-288, 557-589).

    ``presliced`` legs are sliced before the search starts (output-index
    slicing of a fixed set of open legs); the search adds legs from there.
    Returns the sorted list of sliced legs, or None when no executable leg
    makes further progress.
    """
    if target_size <= 0:
        return None
    sd0 = net["size_dict"]
    sliced = sorted(set(presliced))
    while True:
        cur_sd = dict(sd0)
        for l in sliced:
            cur_sd[l] = 1
        cur = {"inputs": net["inputs"], "output": net["output"], "size_dict": cur_sd}
        _, step_legs = simulate_path_full(cur, path)
        if len(net["inputs"]) == 1:
            feasible = _legs_feasible(cur_sd, net["output"], target_size)
            candidates = [net["output"]] if allow_output_slicing else []
        else:
            feasible = all(_legs_feasible(cur_sd, ls, target_size) for ls in step_legs)
            candidates = step_legs
        if feasible:
            return sliced
        score = {}
        for legs in candidates:
            if _legs_feasible(cur_sd, legs, target_size):
                continue
            for l in legs:
                d = math.log2(cur_sd[l])
                if d <= 0.0 or not _slice_leg_allowed(net, l, allow_output_slicing):
                    continue
                c, dd = score.get(l, (0, d))
                score[l] = (c + 1, dd)
        if not score:
            return None
        best = max(score.items(), key=lambda kv: (kv[1][0], kv[1][1], -kv[0]))[0]
        sliced.append(best)
        sliced.sort()


def ssa_to_linear(ssa_path, n):
    """pybind/src/lib.rs:138-151"""
    ids = list(range(n))
    out = []
    for k, (a, b) in enumerate(ssa_path):
        ia, ib = ids.index(a), ids.index(b)
        lo, hi = min(ia, ib), max(ia, ib)
        out.append((lo, hi))
        del ids[hi]
        del ids[lo]
        ids.append(n + k)
    return out


def linear_to_ssa(linear_path, n):
    """pybind/src/lib.rs:154-172"""
    ids = list(range(n))
    out = []
    for k, (a, b) in enumerate(linear_path):
        lo, hi = min(a, b), max(a, b)
        if hi >= len(ids) or lo == hi:
            raise ValueError(f"linear path index out of range or self contraction: ({a},{b})")
        out.append((ids[lo], ids[hi]))
        del ids[hi]
        del ids[lo]
        ids.append(n + k)
    return out


def _rgreedy_params(trial, rng):
    """Trial parameters in the ranges of src/paths/greedy.rs:325-360.

    Trials 0-2 are the deterministic anchors costmod 1, 4, 8; later trials
    draw costmod log-uniform in [0.1, 50] and temperature in [1e-3, 1]. The
    random stream is NumPy's, not the reference's ChaCha8, so trial paths are
    not bit-identical to the reference's random-greedy.
    """
    if trial <= 2:
        return (1.0, 4.0, 8.0)[trial], 0.0, ("mem",)
    cm = 10.0 ** rng.uniform(math.log10(0.1), math.log10(50.0))
    t = 10.0 ** rng.uniform(-3.0, 0.0)
    if trial % 5 == 3:
        mag = 10.0 ** rng.uniform(math.log10(0.05), math.log10(0.5))
        fn = ("skew", mag if rng.uniform() < 0.5 else -mag)
    elif trial % 5 == 4:
        if trial % 3 == 0:
            fn = ("ratio",)
        elif trial % 3 == 1:
            fn = ("logratio",)
        else:
            mag = 10.0 ** rng.uniform(math.log10(0.02), math.log10(0.5))
            fn = ("hyperdeg", mag if rng.uniform() < 0.5 else -mag)
    else:
        fn = ("mem",)
    return cm, t, fn


def sliced_cost_log10(net, path, sliced):
    """log10 of total scalar multiplications over all slices."""
    sd = dict(net["size_dict"])
    n_log2 = 0.0
    for l in sliced:
        n_log2 += math.log2(sd[l])
        sd[l] = 1
    st, _ = simulate_path_full({"inputs": net["inputs"], "output": net["output"], "size_dict": sd}, path)
    return st.log10_flops + n_log2 * math.log10(2.0)


def objective_score_log2(stats, flops_weight=1.0, read_write_weight=64.0):
    """The reference's fixed planner objective (objective.rs:38-41, 87-102):
    log2(flops_weight * flops + read_write_weight * read_write)."""
    log2_flops = stats.log10_flops / math.log10(2.0)
    return _logaddexp2(math.log2(flops_weight) + log2_flops, math.log2(read_write_weight) + stats.log2_read_write)


def random_greedy(net, ntrials, seed):
    """Best of ``ntrials`` random-greedy trials by the fixed objective
    (greedy.rs:325-412; ties keep the lower trial). Returns (path, stats)."""
    import numpy as np

    best = None
    for trial in range(max(ntrials, 1)):
        rng = np.random.default_rng([seed, trial])
        cm, t, fn = _rgreedy_params(trial, rng)
        path = greedy(net, costmod=cm, temperature=t, nbranch=8 if trial > 2 else 1, cost_fn=fn, rng=rng)
        st, _ = simulate_path_full(net, path)
        score = objective_score_log2(st)
        if best is None or score < best[2]:
            best = (path, st, score)
    return best[0], best[1]


def random_greedy_sliced(net, ntrials, seed, target_size, allow_output_slicing=False, presliced=()):
    """Random-greedy trials (greedy.rs:325-412, nbranch 8), each sliced to
    ``target_size`` with ``find_slices_to_size`` (starting from the
    ``presliced`` legs); the trial with the lowest total sliced FLOPs wins
    (ties: lower trial index). Returns ``(path, sliced, log10_total_flops)``.
    """
    import numpy as np

    best = None
    for trial in range(max(ntrials, 1)):
        rng = np.random.default_rng([seed, trial])
        cm, t, fn = _rgreedy_params(trial, rng)
        path = greedy(net, costmod=cm, temperature=t, nbranch=8 if trial > 2 else 1, cost_fn=fn, rng=rng)
        sliced = find_slices_to_size(net, path, target_size, allow_output_slicing, presliced)
        if sliced is None:
            continue
        cost = sliced_cost_log10(net, path, sliced)
        if best is None or cost < best[2]:
            best = (path, sliced, cost)
    return best
