"""Execution-plan v2 artifacts (schema "arctn-execution-plan") executed on the GPU.

A version 2 plan embeds the dense integer network, an SSA path and the exact
slice-leg set, so it runs without re-planning. Loading follows the validation of
src/execution_plan.rs:404-483 (`parse_v2`) and its helpers (:485-580 metrics,
:700-818 slice set and target declaration, :820-898 target and label table), and
the class mirrors pybind/python/arctn/_execution_plan.py:117-590
(`ArcTNExecutionPlan`: from_dict / load / save / to_dict / validate /
to_linear_path / compile / execute). Execution goes to the HIP executor
(`contract_sliced` / `CompiledContraction` of this package).

Errors raise `PlanValidationError` (a ValueError) before any device work, with the
reference's messages.
"""

from __future__ import annotations

import contextlib
import json
import math
import operator
import os
import secrets
from pathlib import Path
from types import MappingProxyType
from typing import NamedTuple, Optional

from . import planner as _planner

SCHEMA = "arctn-execution-plan"
SCHEMA_VERSION = 2
PATH_FORMAT = "ssa-v1"
MEMORY_CONSTRAINT = "max_intermediate_elements_per_slice"
# src/execution_plan.rs:28-31
_FEASIBILITY_TOL_LOG2 = 1e-9
_TARGET_DECL_TOL_LOG2 = 1e-12
_METRIC_ABS_TOL = 1e-9
_METRIC_REL_TOL = 1e-12
_METRIC_KEYS = ("log10_flops", "log2_max_size", "log2_max_contraction_size", "log2_total_size", "log2_read_write",
                "log2_peak_size", "log2_n_slices", "sliced_log10_flops_total", "sliced_log2_max_size",
                "sliced_log2_max_contraction_size", "sliced_log2_peak_size",
                "max_intermediate_log2_elements_per_slice")
_PLANNING_KEYS = ("preset", "method", "seed", "max_time", "rate_enabled", "planner_objective", "flops_weight",
                  "read_write_weight", "slicing_mode", "allow_output_slicing")


class PlanValidationError(ValueError):
    """An execution plan is incomplete or internally inconsistent."""


def network_canon(inputs, output, size_items):
    """Canonical network identity string (src/pathcache.rs:72-100):
    ``t:<legs of each input, ',' joined, ';' between inputs>|o:<output>|d:<leg=dim sorted by leg>``."""
    tensors = ";".join(",".join(str(l) for l in t) for t in inputs)
    out = ",".join(str(l) for l in output)
    dims = ",".join(f"{l}={d}" for l, d in sorted(size_items))
    return f"t:{tensors}|o:{out}|d:{dims}"


def _is_int(v):
    return isinstance(v, int) and not isinstance(v, bool)


def _legs_list(raw, what):
    if not isinstance(raw, list) or not all(_is_int(x) and x >= 0 for x in raw):
        raise PlanValidationError(f"{what} must be a list of non-negative integers")
    return list(raw)


def _network(root):
    """src/execution_plan.rs:61-105 (`ExecutionPlanNetwork::to_network`)."""
    raw = root.get("network")
    if raw is None:
        raise PlanValidationError("execution plan version 2 is missing network")
    if not isinstance(raw, dict) or not all(k in raw for k in ("inputs", "output", "size_dict")):
        raise PlanValidationError("execution plan network is invalid: inputs, output and size_dict are required")
    if not isinstance(raw["inputs"], list):
        raise PlanValidationError("execution plan network is invalid: inputs must be a list")
    inputs = [_legs_list(t, "execution plan network inputs") for t in raw["inputs"]]
    output = _legs_list(raw["output"], "execution plan network output")
    sd_raw = raw["size_dict"]
    if not isinstance(sd_raw, list) or not all(
            isinstance(e, list) and len(e) == 2 and _is_int(e[0]) and _is_int(e[1]) and e[0] >= 0 and e[1] >= 0
            for e in sd_raw):
        raise PlanValidationError("execution plan network is invalid: size_dict must be [leg, dimension] pairs")
    seen, nxt = set(), 0
    for leg in [l for t in inputs for l in t] + output:
        if leg not in seen:
            if leg != nxt:
                raise PlanValidationError("execution plan network legs must use dense first-occurrence order")
            seen.add(leg)
            nxt += 1
    if len(sd_raw) != len(seen) or any(leg != i for i, (leg, _) in enumerate(sd_raw)):
        raise PlanValidationError("execution plan network.size_dict must contain each dense leg once in order")
    size_dict = {leg: dim for leg, dim in sd_raw}
    net = {"inputs": inputs, "output": output, "size_dict": size_dict}
    try:
        _planner.validate_network(net)
    except ValueError as e:
        raise PlanValidationError(f"execution plan network is invalid: {e}") from None
    return net


def _label_table(net, labels):
    """src/execution_plan.rs:880-927 (`validate_label_table`, `unique_label_map`)."""
    if not isinstance(labels, list) or not all(isinstance(x, str) for x in labels):
        raise PlanValidationError("network_leg_labels is invalid: a list of strings is required")
    if len(labels) != len(net["size_dict"]):
        raise PlanValidationError("network_leg_labels length must match the embedded dense network")
    if len(set(labels)) != len(labels):
        raise PlanValidationError("network_leg_labels must be unique")


def _validate_slice_legs(net, legs):
    """src/slice.rs:146-172."""
    seen = set()
    for leg in legs:
        if leg in seen:
            raise ValueError(f"切片腿 {leg} 重复出现")
        seen.add(leg)
        dim = net["size_dict"].get(leg)
        if dim is None:
            raise ValueError(f"切片腿 {leg} 不在 size_dict 中")
        if dim == 0:
            raise ValueError(f"切片腿 {leg} 的维度为 0")
        if not any(leg in t for t in net["inputs"]):
            raise ValueError(f"切片腿 {leg} 未被任何输入张量持有")
        if any(t.count(leg) > 1 for t in net["inputs"]):
            raise ValueError(f"切片腿 {leg} 在同一输入张量中重复出现，暂不支持")


def _slice_set(root, net, labels):
    """src/execution_plan.rs:713-775 (`parse_slice_set`, explicit set required)."""
    raw = root.get("sliced")
    if raw is None:
        raise PlanValidationError("versioned execution plans must provide sliced; an unsliced plan uses empty legs")
    if not isinstance(raw, dict):
        raise PlanValidationError("sliced must be an object")
    if "legs" not in raw:
        raise PlanValidationError("versioned execution plans require sliced.legs")
    legs = _legs_list(raw["legs"], "sliced.legs")
    if "leg_labels" in raw:
        if labels is None:
            raise PlanValidationError("sliced.leg_labels cannot be checked without network_leg_labels")
        got = raw["leg_labels"]
        want = []
        for leg in legs:
            if leg >= len(labels):
                raise PlanValidationError(f"network_leg_labels has no entry for LegId {leg}")
            want.append(labels[leg])
        if got != want:
            raise PlanValidationError("sliced.legs and sliced.leg_labels do not identify the same legs")
    try:
        _validate_slice_legs(net, legs)
    except ValueError as e:
        raise PlanValidationError(f"invalid slice set: {e}") from None
    return legs


def _sliced_stats(net, path, sliced):
    """src/execution_plan.rs:852-863: path statistics with every sliced leg at dimension 1."""
    sd = dict(net["size_dict"])
    for leg in sliced:
        sd[leg] = 1
    return _planner.simulate_path_full(dict(net, size_dict=sd), path)


def _check_metrics(obj, where, expected):
    """src/execution_plan.rs:562-584 (`validate_metric_fields`)."""
    for name, want in expected:
        raw = obj.get(name)
        if raw is None:
            continue
        if not isinstance(raw, (int, float)) or isinstance(raw, bool) or not math.isfinite(raw):
            raise PlanValidationError(f"{where}.{name} must be a finite number")
        tol = max(_METRIC_ABS_TOL, _METRIC_REL_TOL * max(abs(want), abs(raw)))
        if not math.isfinite(want) or abs(raw - want) > tol:
            raise PlanValidationError(f"{where}.{name}={raw} does not match the recomputed value {want}")


def _metrics(root, net, path, stats, sliced):
    """src/execution_plan.rs:485-560 (`validate_reported_metrics`)."""
    path_metrics = [("log10_flops", stats.log10_flops), ("log2_max_size", stats.log2_max_size),
                    ("log2_max_contraction_size", stats.log2_max_contraction_size),
                    ("log2_total_size", stats.log2_total_size), ("log2_read_write", stats.log2_read_write),
                    ("log2_peak_size", stats.log2_peak_size)]
    _check_metrics(root, "execution plan", path_metrics)
    metrics = root.get("metrics")
    if metrics is not None and not isinstance(metrics, dict):
        raise PlanValidationError("execution plan metrics must be an object")
    if metrics is not None:
        _check_metrics(metrics, "execution plan metrics", path_metrics)
    per, _ = _sliced_stats(net, path, sliced)
    log2_n = sum(math.log2(net["size_dict"][l]) for l in sliced)
    total = per.log10_flops + log2_n * math.log10(2.0)
    canon = [("log2_n_slices", log2_n), ("sliced_log10_flops_total", total),
             ("sliced_log2_max_size", per.log2_max_size),
             ("sliced_log2_max_contraction_size", per.log2_max_contraction_size),
             ("sliced_log2_peak_size", per.log2_peak_size),
             ("max_intermediate_log2_elements_per_slice", per.log2_max_size)]
    _check_metrics(root, "execution plan", canon)
    if metrics is not None:
        _check_metrics(metrics, "execution plan metrics", canon)
    _check_metrics(root["sliced"], "execution plan sliced", [
        ("log2_n_slices", log2_n), ("per_slice_log10_flops", per.log10_flops),
        ("per_slice_log2_write", per.log2_total_size), ("per_slice_log2_read_write", per.log2_read_write),
        ("per_slice_log2_max_size", per.log2_max_size),
        ("per_slice_log2_max_contraction_size", per.log2_max_contraction_size),
        ("per_slice_log2_peak_size", per.log2_peak_size), ("log10_flops_total", total),
        ("log2_write_total", per.log2_total_size + log2_n), ("log2_read_write_total", per.log2_read_write + log2_n)])


def _target(root, net, path, sliced):
    """src/execution_plan.rs:777-850 (`parse_target_declaration` with the metric
    required, `validate_target`); returns the exact target or None."""
    exact = root.get("target_size")
    if exact is not None and not (_is_int(exact) and exact > 0):
        raise PlanValidationError("target_size must be a positive usize integer")
    log2 = root.get("memory_target_log2_elements")
    if log2 is not None and (not isinstance(log2, (int, float)) or isinstance(log2, bool)
                             or not math.isfinite(log2) or log2 < 0):
        raise PlanValidationError("memory_target_log2_elements must be a finite non-negative number")
    if exact is not None and log2 is not None:
        want = math.log2(exact)
        if abs(log2 - want) > _TARGET_DECL_TOL_LOG2:
            raise PlanValidationError(f"target_size={exact} is inconsistent with memory_target_log2_elements="
                                      f"{log2}; expected {want}")
    metric = root.get("memory_constraint_metric")
    if exact is None and (log2 is not None or metric is not None):
        raise PlanValidationError(
            "execution plan version 2 requires exact target_size for every memory target declaration")
    if exact is not None and log2 is None:
        raise PlanValidationError("execution plan version 2 target_size requires memory_target_log2_elements")
    if exact is not None and metric != MEMORY_CONSTRAINT:
        raise PlanValidationError(f"memory_constraint_metric must be {MEMORY_CONSTRAINT!r} when a target is declared")
    if exact is None:
        return None
    sd = dict(net["size_dict"])
    for leg in sliced:
        sd[leg] = 1
    cur = dict(net, size_dict=sd)
    _, step_legs = _planner.simulate_path_full(cur, path)
    # src/slice.rs:78-96: every produced intermediate (the output root of a unary network)
    groups = [net["output"]] if len(net["inputs"]) == 1 else step_legs
    if not all(_planner._legs_feasible(sd, legs, exact) for legs in groups):
        raise PlanValidationError(f"execution plan exceeds declared target_size={exact}")
    return exact


def parse_v2(artifact):
    """Validate one self-contained version 2 artifact (src/execution_plan.rs:233-248,
    250-268, 404-483). Returns (net, labels, ssa_path, sliced_legs, target_size)."""
    if not isinstance(artifact, dict):
        raise PlanValidationError("plan JSON root must be an object")
    schema, version = artifact.get("schema"), artifact.get("schema_version")
    if schema is None and version is None:
        raise PlanValidationError("schema-less execution plans are not self-contained")
    if schema is None or version is None:
        raise PlanValidationError("execution plan must contain both schema and schema_version")
    if schema != SCHEMA:
        raise PlanValidationError(f"unsupported execution plan schema: {json.dumps(schema)}")
    if not _is_int(version) or version < 0:
        raise PlanValidationError("execution plan schema_version must be an integer")
    if version != SCHEMA_VERSION:
        raise PlanValidationError(f"execution plan version {version} is not self-contained; version 2 is required")
    if artifact.get("path_format") != PATH_FORMAT:
        raise PlanValidationError(f"unsupported execution plan path_format: {json.dumps(artifact.get('path_format'))}")
    if "net" in artifact and not isinstance(artifact["net"], str):
        raise PlanValidationError("plan net field must be a string")
    net = _network(artifact)
    canon = artifact.get("network_canon")
    if not isinstance(canon, str):
        raise PlanValidationError("execution plan network_canon must be a string")
    if canon != network_canon(net["inputs"], net["output"], net["size_dict"].items()):
        raise PlanValidationError("execution plan network_canon does not match embedded network")
    labels = artifact.get("network_leg_labels")
    if labels is not None:
        _label_table(net, labels)
    raw_path = artifact.get("ssa_path")
    if raw_path is None:
        raise PlanValidationError("execution plan is missing ssa_path")
    if not isinstance(raw_path, list) or not all(
            isinstance(p, list) and len(p) == 2 and all(_is_int(x) and x >= 0 for x in p) for p in raw_path):
        raise PlanValidationError("ssa_path is invalid: a list of [i, j] pairs is required")
    path = [tuple(p) for p in raw_path]
    try:
        stats, _ = _planner.simulate_path_full(net, path)
    except ValueError as e:
        raise PlanValidationError(f"execution plan ssa_path is invalid: {e}") from None
    sliced = _slice_set(artifact, net, labels)
    _metrics(artifact, net, path, stats, sliced)
    target = _target(artifact, net, path, sliced)
    return net, labels, path, sliced, target


class _Parsed(NamedTuple):
    """The validated content of one artifact; `doc` is its canonical JSON text
    (immutable by construction: readers get fresh objects decoded from it)."""
    net: dict
    ssa_path: tuple
    sliced: tuple
    target: Optional[int]
    labels: tuple
    doc: str


# Flat top-level fields that a v2 artifact may carry outside its metadata
# objects, and the object each belongs to when that object is absent.
_METADATA_GROUPS = (("metrics", _METRIC_KEYS), ("planning", _PLANNING_KEYS),
                    ("provenance", ("arctn_version", "source")))
_BUILD = object()


def _group_metadata(artifact):
    for group, keys in _METADATA_GROUPS:
        if group not in artifact:
            artifact[group] = {k: artifact[k] for k in keys if k in artifact}
        elif not isinstance(artifact[group], dict):
            raise PlanValidationError(f"execution plan {group} must be an object")
    return artifact


def _same_dtype(dtypes):
    return len({str(d) for d in dtypes}) <= 1


class ExecutionPlan:
    """An immutable, validated version 2 plan; the surface of ArcTNExecutionPlan
    (pybind/python/arctn/_execution_plan.py:117-590): from_dict, load, to_dict,
    save, validate, to_linear_path, compile, execute."""

    __slots__ = ("_p",)
    __hash__ = None

    def __init__(self, token=None, parsed=None):
        if token is not _BUILD:
            raise TypeError("use ExecutionPlan.from_dict() or ExecutionPlan.load()")
        ExecutionPlan._p.__set__(self, parsed)

    def __setattr__(self, name, value):
        raise AttributeError("ExecutionPlan is immutable")

    @classmethod
    def from_dict(cls, artifact):
        """Validate and load one self-contained v2 artifact mapping."""
        if not isinstance(artifact, dict):
            raise PlanValidationError("execution plan root must be an object")
        try:  # a private JSON-clean copy: the caller's mapping is never kept or changed
            text = json.dumps(artifact, ensure_ascii=False, allow_nan=False)
        except (TypeError, ValueError) as e:
            raise PlanValidationError(f"invalid execution plan: {e}") from None
        artifact = json.loads(text)
        net, labels, path, sliced, target = parse_v2(artifact)
        artifact = _group_metadata(artifact)
        labels = tuple(labels) if labels is not None else tuple(sorted(net["size_dict"]))
        doc = json.dumps(artifact, ensure_ascii=False, allow_nan=False, sort_keys=True)
        return cls(_BUILD, _Parsed(net, tuple(path), tuple(sliced), target, labels, doc))

    @classmethod
    def load(cls, path):
        """Load and validate one self-contained v2 JSON plan file."""
        try:
            text = Path(os.fspath(path)).read_text(encoding="utf-8")
            artifact = json.loads(text)
        except (OSError, UnicodeError, json.JSONDecodeError) as e:
            raise PlanValidationError(f"cannot load execution plan: {e}") from None
        return cls.from_dict(artifact)

    @classmethod
    def from_network(cls, inputs, output, size_dict, ssa_path, sliced_legs=(), *, target_size=None,
                     leg_labels=None, source="arctn_amd"):
        """A v2 artifact for a dense integer network (legs in first-occurrence order)
        and a supplied SSA path and slice set, with recomputed metrics."""
        size_items = sorted((int(l), int(d)) for l, d in dict(size_dict).items())
        art = {"schema": SCHEMA, "schema_version": SCHEMA_VERSION, "path_format": PATH_FORMAT,
               "network_canon": network_canon(inputs, output, size_items),
               "network": {"inputs": [list(t) for t in inputs], "output": list(output),
                           "size_dict": [list(e) for e in size_items]},
               "ssa_path": [list(p) for p in ssa_path], "sliced": {"legs": list(sliced_legs)},
               "metrics": {}, "planning": {}, "provenance": {"source": source}}
        if leg_labels is not None:
            art["network_leg_labels"] = list(leg_labels)
            art["sliced"]["leg_labels"] = [leg_labels[l] for l in sliced_legs]
        if target_size is not None:
            art.update({"target_size": int(target_size), "memory_target_log2_elements": math.log2(target_size),
                        "memory_constraint_metric": MEMORY_CONSTRAINT})
        net = {"inputs": art["network"]["inputs"], "output": art["network"]["output"], "size_dict": dict(size_items)}
        stats, _ = _planner.simulate_path_full(net, [tuple(p) for p in ssa_path])
        per, _ = _sliced_stats(net, [tuple(p) for p in ssa_path], list(sliced_legs))
        log2_n = sum(math.log2(net["size_dict"][l]) for l in sliced_legs)
        art["metrics"] = {"log10_flops": stats.log10_flops, "log2_max_size": stats.log2_max_size,
                          "log2_peak_size": stats.log2_peak_size, "log2_n_slices": log2_n,
                          "sliced_log10_flops_total": per.log10_flops + log2_n * math.log10(2.0),
                          "sliced_log2_max_size": per.log2_max_size}
        return cls.from_dict(art)

    def _part(self, key):
        return MappingProxyType(json.loads(self._p.doc)[key])

    schema = property(lambda self: SCHEMA)
    schema_version = property(lambda self: SCHEMA_VERSION)
    path_format = property(lambda self: PATH_FORMAT)
    inputs = property(lambda self: tuple(tuple(t) for t in self._p.net["inputs"]))
    output = property(lambda self: tuple(self._p.net["output"]))
    size_dict = property(lambda self: MappingProxyType(dict(self._p.net["size_dict"])))
    ssa_path = property(lambda self: self._p.ssa_path)
    sliced_legs = property(lambda self: self._p.sliced)
    target_size = property(lambda self: self._p.target)
    metrics = property(lambda self: self._part("metrics"))
    planning = property(lambda self: self._part("planning"))
    provenance = property(lambda self: self._part("provenance"))
    leg_labels = property(lambda self: self._p.labels)
    network_canon = property(lambda self: json.loads(self._p.doc)["network_canon"])
    is_sliced = property(lambda self: bool(self._p.sliced))

    def to_dict(self):
        """The portable JSON-compatible v2 artifact mapping (a fresh copy)."""
        return json.loads(self._p.doc)

    def save(self, path):
        """Write the v2 JSON representation to `path`, replacing it atomically: the
        text goes to a new sibling file that is flushed to disk and then renamed
        over the destination, so readers see the old file or the whole new one."""
        dest = Path(os.fspath(path))
        body = json.dumps(self.to_dict(), ensure_ascii=False, allow_nan=False, indent=2, sort_keys=True) + "\n"
        partial = dest.with_name(f".{dest.name}.{os.getpid()}.{secrets.token_hex(4)}.partial")
        try:
            with open(partial, "x", encoding="utf-8") as f:
                f.write(body)
                f.flush()
                os.fsync(f.fileno())
            os.replace(partial, dest)
        except OSError as e:
            with contextlib.suppress(OSError):
                partial.unlink()
            raise PlanValidationError(f"cannot save execution plan: {e}") from None
        return dest

    def _check_caller_network(self, inputs, output, size_dict):
        mine = self._p.net
        same = ([list(t) for t in inputs] == mine["inputs"] and list(output) == mine["output"]
                and sorted(dict(size_dict).items()) == sorted(mine["size_dict"].items()))
        if not same:
            raise PlanValidationError("caller network does not match the embedded plan network")

    def _check_arrays(self, arrays):
        try:
            values = tuple(arrays)
        except TypeError:
            raise PlanValidationError("arrays must be a sequence") from None
        legs = self._p.net["inputs"]
        if len(values) != len(legs):
            raise PlanValidationError(f"arrays length {len(values)} does not match {len(legs)} plan inputs")
        for i, a in enumerate(values):
            if not (hasattr(a, "shape") and hasattr(a, "dtype")):
                raise PlanValidationError(f"arrays[{i}] must provide shape and dtype")
        for i, (a, t) in enumerate(zip(values, legs)):
            want = tuple(self._p.net["size_dict"][l] for l in t)
            got = tuple(operator.index(d) for d in a.shape)
            if got != want:
                raise PlanValidationError(f"arrays[{i}] shape {got} does not match {want}")
        if not _same_dtype(a.dtype for a in values):
            raise PlanValidationError("all arrays must use the same dtype")

    def validate(self, inputs=None, output=None, size_dict=None, *, arrays=None):
        """Check a caller network (dense labels, given all together) and the
        arrays' count, shapes and common dtype against the plan; raises
        PlanValidationError on the first mismatch."""
        given = [x is not None for x in (inputs, output, size_dict)]
        if any(given):
            if not all(given):
                raise TypeError("inputs, output, and size_dict must be provided together")
            self._check_caller_network(inputs, output, size_dict)
        if arrays is not None:
            self._check_arrays(arrays)

    def to_linear_path(self):
        """The opt_einsum-style linear path."""
        return tuple(_planner.ssa_to_linear(self._p.ssa_path, len(self._p.net["inputs"])))

    def compile(self):
        """A `CompiledContraction` of an unsliced plan for repeated execution."""
        if self._p.sliced:
            raise PlanValidationError("sliced plans cannot be compiled by CompiledContraction; "
                                      "call plan.execute(...)")
        from .api import CompiledContraction
        net = self._p.net
        return CompiledContraction.compile(net["inputs"], net["output"], net["size_dict"], ssa_path=self._p.ssa_path)

    def execute(self, arrays, *, return_info=False, **opts):
        """Run the plan on the GPU: internal slices summed, output blocks assembled,
        the full output in the plan's output order (_execution_plan.py:489-577).
        ``opts`` are the executor options of `contract_sliced` (precision, batched,
        use_graph, autotune, ...)."""
        import time

        from .api import contract_sliced
        arrays = tuple(arrays)
        self.validate(arrays=arrays)
        net, sliced = self._p.net, self._p.sliced
        t0 = time.perf_counter()
        if sliced:
            result = contract_sliced(net["inputs"], net["output"], net["size_dict"], arrays, list(sliced),
                                     ssa_path=self._p.ssa_path, **opts)
        else:
            result = self.compile().execute(arrays, **opts)
        if not return_info:
            return result
        return result, {"plan_schema": SCHEMA, "plan_schema_version": SCHEMA_VERSION, "execution_backend": "hip",
                        "execution_implementation": "arctn_amd.contract_sliced" if sliced
                        else "arctn_amd.CompiledContraction", "execution_wall_s": time.perf_counter() - t0,
                        "sliced_legs": list(sliced)}

    def __eq__(self, other):
        if not isinstance(other, ExecutionPlan):
            return NotImplemented
        return self._p.doc == other._p.doc

    def __repr__(self):
        p = self._p
        return (f"ExecutionPlan(n_inputs={len(p.net['inputs'])}, n_steps={len(p.ssa_path)}, "
                f"n_sliced_legs={len(p.sliced)}, target_size={p.target!r})")
