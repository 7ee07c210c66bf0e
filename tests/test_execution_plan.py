"""Execution-plan v2 artifacts: validation on the CPU (the cases of
src/execution_plan.rs:930-1153), GPU execution of saved plans against np.einsum
(pybind/tests/test_output_slicing.py:88-165)."""
import json
import math

import numpy as np
import pytest

from arctn_amd import ExecutionPlan, PlanValidationError
from arctn_amd import execution_plan as EP
from arctn_amd import planner as P
from tests.golden import fixtures as F

FX = F.load()
LABELS = ["a", "b", "c"]
TRI = {"inputs": [[0, 1], [0, 2], [1, 2]], "output": [], "size_dict": {0: 2, 1: 2, 2: 2}}
TRI_PATH = [(0, 1), (2, 3)]


def _tri(metrics=False, **kw):
    """The triangle record of src/execution_plan.rs:936-962 (metrics only on request,
    like the reference's record, so tests may change its slice set)."""
    rec = ExecutionPlan.from_network(TRI["inputs"], TRI["output"], TRI["size_dict"], TRI_PATH, leg_labels=LABELS,
                                     **kw).to_dict()
    if not metrics:
        rec["metrics"] = {}
    return rec


def _artifact(net, path, sliced):
    """The v2 form the reference's tests write by hand (test_output_slicing.py:88-108)."""
    sizes = sorted(net["size_dict"].items())
    return {"schema": "arctn-execution-plan", "schema_version": 2, "path_format": "ssa-v1",
            "network_canon": EP.network_canon(net["inputs"], net["output"], sizes),
            "network": {"inputs": net["inputs"], "output": net["output"], "size_dict": [list(s) for s in sizes]},
            "ssa_path": [list(p) for p in path], "sliced": {"legs": list(sliced)},
            "metrics": {}, "planning": {}, "provenance": {"source": "explicit-test-plan"}}


def test_network_canon_format():
    # src/pathcache.rs:72-100, and the identity string the reference's test builds
    assert EP.network_canon([[0, 1], [1, 2]], [2, 0], [(2, 4), (0, 2), (1, 3)]) == "t:0,1;1,2|o:2,0|d:0=2,1=3,2=4"
    assert EP.network_canon([[0]], [], [(0, 5)]) == "t:0|o:|d:0=5"


def test_version_two_round_trips_embedded_network(tmp_path):
    rec = _tri(metrics=True)
    assert rec["schema_version"] == 2 and rec["path_format"] == "ssa-v1"
    assert rec["network"]["size_dict"] == [[0, 2], [1, 2], [2, 2]]
    assert rec["sliced"]["leg_labels"] == []
    plan = ExecutionPlan.from_dict(rec)
    assert plan.inputs == ((0, 1), (0, 2), (1, 2)) and plan.output == ()
    assert plan.ssa_path == tuple(TRI_PATH) and plan.sliced_legs == ()
    assert plan.leg_labels == tuple(LABELS)
    stats, _ = P.simulate_path_full(TRI, TRI_PATH)
    assert plan.metrics["log10_flops"] == stats.log10_flops
    f = tmp_path / "plan.json"
    plan.save(f)
    assert ExecutionPlan.load(f) == plan
    assert json.loads(f.read_text())["network_canon"] == "t:0,1;0,2;1,2|o:|d:0=2,1=2,2=2"
    assert plan.to_linear_path() == tuple(P.ssa_to_linear(TRI_PATH, 3))


def test_version_two_rejects_tampered_path_metrics():
    rec = _tri(metrics=True)
    rec["metrics"]["log10_flops"] = -999.0
    with pytest.raises(PlanValidationError, match=r"metrics\.log10_flops.*recomputed value"):
        ExecutionPlan.from_dict(rec)
    rec = _tri()
    rec["log2_peak_size"] = 99.0  # top-level copies are checked too
    with pytest.raises(PlanValidationError, match="log2_peak_size"):
        ExecutionPlan.from_dict(rec)
    rec = _tri()
    rec["sliced"]["per_slice_log10_flops"] = 1.0
    with pytest.raises(PlanValidationError, match="sliced.per_slice_log10_flops"):
        ExecutionPlan.from_dict(rec)


def test_version_two_validates_slices_network_and_path_format():
    rec = _tri()
    rec["sliced"] = {"legs": [0], "leg_labels": ["a"]}
    ExecutionPlan.from_dict(rec)
    bad = json.loads(json.dumps(rec))
    bad["sliced"]["leg_labels"] = ["b"]
    with pytest.raises(PlanValidationError, match="do not identify the same legs"):
        ExecutionPlan.from_dict(bad)
    bad = json.loads(json.dumps(rec))
    bad["path_format"] = "linear"
    with pytest.raises(PlanValidationError, match="path_format"):
        ExecutionPlan.from_dict(bad)
    bad = json.loads(json.dumps(rec))
    bad["network"]["size_dict"][0][1] = 3
    with pytest.raises(PlanValidationError, match="network_canon"):
        ExecutionPlan.from_dict(bad)
    sparse = _artifact({"inputs": [[4]], "output": [4], "size_dict": {4: 2}}, [], [])
    with pytest.raises(PlanValidationError, match="dense first-occurrence"):
        ExecutionPlan.from_dict(sparse)
    for legs, msg in (([0, 0], "重复"), ([7], "不在 size_dict")):
        bad = _tri()
        bad["sliced"] = {"legs": legs}
        with pytest.raises(PlanValidationError, match=msg):
            ExecutionPlan.from_dict(bad)
    bad = _tri()
    bad["ssa_path"] = [[0, 1], [0, 2]]
    with pytest.raises(PlanValidationError, match="ssa_path is invalid"):
        ExecutionPlan.from_dict(bad)


def test_schema_versions_and_fields():
    rec = _tri()
    for key, val, msg in (("schema_version", 1, "not self-contained"), ("schema_version", 3, "version 2 is required"),
                          ("schema", "other", "unsupported execution plan schema")):
        bad = dict(rec, **{key: val})
        with pytest.raises(PlanValidationError, match=msg):
            ExecutionPlan.from_dict(bad)
    bad = {k: v for k, v in rec.items() if k not in ("schema", "schema_version")}
    with pytest.raises(PlanValidationError, match="schema-less"):
        ExecutionPlan.from_dict(bad)
    bad = {k: v for k, v in rec.items() if k != "schema_version"}
    with pytest.raises(PlanValidationError, match="both schema and schema_version"):
        ExecutionPlan.from_dict(bad)
    bad = {k: v for k, v in rec.items() if k != "sliced"}
    with pytest.raises(PlanValidationError, match="must provide sliced"):
        ExecutionPlan.from_dict(bad)
    bad = dict(rec, network_leg_labels=["a", "a", "c"])
    with pytest.raises(PlanValidationError, match="unique"):
        ExecutionPlan.from_dict(bad)
    with pytest.raises(TypeError):
        ExecutionPlan()
    plan = ExecutionPlan.from_dict(rec)
    with pytest.raises(AttributeError):
        plan._ssa_path = ()


def test_exact_target_is_checked_against_saved_slice_set():
    rec = _tri()
    log2_only = dict(rec, memory_target_log2_elements=1.0, memory_constraint_metric=EP.MEMORY_CONSTRAINT)
    with pytest.raises(PlanValidationError, match="exact target_size"):
        ExecutionPlan.from_dict(log2_only)
    with pytest.raises(PlanValidationError):
        ExecutionPlan.from_dict(dict(rec, memory_constraint_metric=EP.MEMORY_CONSTRAINT))
    with pytest.raises(PlanValidationError, match="requires memory_target_log2_elements"):
        ExecutionPlan.from_dict(dict(rec, target_size=8, memory_constraint_metric=EP.MEMORY_CONSTRAINT))
    rec.update(target_size=2, memory_target_log2_elements=1.0, memory_constraint_metric=EP.MEMORY_CONSTRAINT)
    with pytest.raises(PlanValidationError, match="exceeds declared target_size=2"):
        ExecutionPlan.from_dict(rec)
    rec["sliced"] = {"legs": [0, 1], "leg_labels": ["a", "b"]}
    assert ExecutionPlan.from_dict(rec).target_size == 2
    big = _tri()
    t = (1 << 53) + 1
    big.update(target_size=t, memory_target_log2_elements=math.log2(t), memory_constraint_metric=EP.MEMORY_CONSTRAINT)
    assert ExecutionPlan.from_dict(big).target_size == t


@pytest.mark.parametrize("case", range(len(FX["output_slicing_cases"]["cases"])))
def test_einsum_case_artifacts_load_and_round_trip(case, tmp_path):
    name, eq, net, sliced, path = F.einsum_case(FX["output_slicing_cases"]["cases"][case])
    plan = ExecutionPlan.from_dict(_artifact(net, path, sliced))
    assert plan.output == tuple(net["output"]) and plan.sliced_legs == tuple(sliced)
    assert plan.ssa_path == tuple(tuple(p) for p in path)
    f = tmp_path / "plan.json"
    plan.save(f)
    assert ExecutionPlan.load(f) == plan
    plan.validate(net["inputs"], net["output"], net["size_dict"], arrays=F.case_arrays(net, np.float64))
    with pytest.raises(PlanValidationError, match="arrays length"):
        plan.validate(arrays=[])


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(FX["output_slicing_cases"]["cases"])))
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.complex128])
def test_saved_plans_execute_on_gpu(cuda, case, dtype, tmp_path):
    """A saved v2 plan runs unchanged: plain and batched forms vs np.einsum at the
    reference's tolerances (2e-4 single, 1e-11 double)."""
    name, eq, net, sliced, path = F.einsum_case(FX["output_slicing_cases"]["cases"][case])
    f = tmp_path / "plan.json"
    ExecutionPlan.from_dict(_artifact(net, path, sliced)).save(f)
    plan = ExecutionPlan.load(f)
    arrays = F.case_arrays(net, dtype)
    ref = np.einsum(eq, *arrays, optimize=False)
    tol = 2e-4 if np.dtype(dtype) in (np.dtype(np.float32), np.dtype(np.complex64)) else 1e-11
    for opts in (dict(), dict(batched=True), dict(batched=True, use_graph=True)):
        got, info = plan.execute(arrays, return_info=True, **opts)
        assert got.shape == ref.shape and got.dtype == np.dtype(dtype)
        np.testing.assert_allclose(got, ref, rtol=tol, atol=tol, err_msg=f"{name} {opts}")
        assert info["sliced_legs"] == list(sliced)
