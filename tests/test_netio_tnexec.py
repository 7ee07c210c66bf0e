"""Network / tensor files of the CLIs (src/network.rs:75-132, 245-392) and the
tnexec-style runner (the cases of tests/tnexec_accuracy.rs:4-45)."""
import json

import numpy as np
import pytest

from arctn_amd import netio
from tests.golden import fixtures as F

FX = F.load()
TINY = {k: FX["demo_tiny6"][k] for k in ("inputs", "output", "size_dict")}


def test_labels_interned_in_first_occurrence_order():
    net, labels = netio.parse_network(dict(TINY, name="demo_tiny6"))
    assert labels == list("abcdefg")
    assert net["inputs"] == [[0, 1], [1, 2], [2, 3], [3, 4], [4, 5], [5, 0, 6]] and net["output"] == [6]
    assert net["size_dict"] == {0: 4, 1: 5, 2: 4, 3: 6, 4: 4, 5: 5, 6: 3} and net["name"] == "demo_tiny6"
    omeco, _ = netio.parse_network({"ixs": TINY["inputs"], "iy": TINY["output"], "sizes": TINY["size_dict"]})
    assert omeco["inputs"] == net["inputs"] and omeco["name"] == "unnamed"
    ints, lab = netio.parse_network({"inputs": [[7, 3], [3]], "output": [7], "size_dict": {"7": 2, "3": 5}})
    assert ints["inputs"] == [[0, 1], [1]] and lab == ["7", "3"]


@pytest.mark.parametrize("bad, msg", [
    ({"inputs": [["a"]], "output": [], "size_dict": {"a": 2}, "ixs": []}, "不能混用"),
    ({"inputs": [["a"]], "output": []}, "canonical schema 残缺"),
    ({"ixs": [["a"]]}, "OMECO schema 残缺"),
    ({}, "缺少完整的网络 schema"),
    ({"inputs": [["a", "b"]], "output": [], "size_dict": {"a": 2}}, "缺少 size_dict 维度"),
    ({"inputs": [["a"]], "output": [], "size_dict": {"a": 2, "z": 3}}, "未在"),
    ({"inputs": [[1, "1"]], "output": [], "size_dict": {"1": 2}}, "冲突"),
    ({"inputs": [["a", "a", "a"]], "output": [], "size_dict": {"a": 2}}, "at most 2"),
])
def test_network_json_errors(bad, msg):
    with pytest.raises(ValueError, match=msg):
        netio.parse_network(bad)


@pytest.mark.parametrize("name", ["f32", "f64", "c32", "c64"])
def test_tensors_bin_round_trip_and_length_check(name, tmp_path):
    dtype = netio.DTYPES[name]
    net, _ = netio.parse_network(TINY)
    rng = np.random.default_rng(3)
    arrays = [(rng.standard_normal(tuple(net["size_dict"][l] for l in legs)) * (1 + 1j if name[0] == "c" else 1))
              .astype(dtype) for legs in net["inputs"]]
    f = tmp_path / "x.bin"
    netio.save_tensors_bin(f, arrays)
    raw = f.read_bytes()
    assert len(raw) == sum(a.size for a in arrays) * np.dtype(dtype).itemsize
    if name[0] == "c":  # real part first, then imaginary, little endian
        half = np.dtype(dtype).itemsize // 2
        re, im = np.frombuffer(raw[:2 * half], dtype="<f4" if half == 4 else "<f8")
        assert re == arrays[0].reshape(-1)[0].real and im == arrays[0].reshape(-1)[0].imag
    back = netio.load_tensors_bin(f, net, dtype)
    assert all(b.dtype == np.dtype(dtype) and np.array_equal(a, b) for a, b in zip(arrays, back))
    f.write_bytes(raw[:-1])
    with pytest.raises(ValueError, match="字节数"):
        netio.load_tensors_bin(f, net, dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("flags, dtype", [(["--single"], "f32"), (["--single", "--complex"], "c32"), ([], "f64"),
                                          (["--complex"], "c64")])
@pytest.mark.parametrize("target", [None, "8"])
def test_tnexec_check_on_demo_network(cuda, flags, dtype, target, tmp_path, capsys):
    """tnexec on the demo network: its --check (NumPy in double precision on the CPU)
    passes, and the EXEC run's output file matches the oracle in f64 / complex128."""
    from arctn_amd import tnexec
    from oracle import pyoracle as O

    net_file = tmp_path / "demo_tiny6.net.json"
    net_file.write_text(json.dumps(dict(TINY, name="demo_tiny6")))
    plan_file = tmp_path / "plan.json"
    args = ["--net", str(net_file), "--method", "greedy", "--check", "--save-path", str(plan_file)] + flags
    if target:
        args += ["--target-size", target]
    assert tnexec.main(args) == 0
    line = capsys.readouterr().out.strip().splitlines()[-1]
    # six tensors: the reference checks against the naive einsum
    assert line.startswith("EXEC,") and f"dtype={dtype}," in line and "check=PASS(naive_" in line, line
    assert f"oracle_dtype={'c64' if 'c' in dtype else 'f64'}" in line
    if target:
        assert "n_sliced_legs=0," not in line
    # the saved plan and a data file run again through --load-path; the output is kept
    net, _ = netio.load_network_json(net_file)
    rng = np.random.default_rng(1)
    dt = {"f32": np.float32, "c32": np.complex64, "f64": np.float64, "c64": np.complex128}[dtype]
    arrays = [rng.standard_normal(tuple(net["size_dict"][l] for l in legs)).astype(dt) for legs in net["inputs"]]
    data, out_file = tmp_path / "x.bin", tmp_path / "out.bin"
    netio.save_tensors_bin(data, arrays)
    assert tnexec.main(["--net", str(net_file), "--load-path", str(plan_file), "--data", str(data), "--check",
                        "--graph", "--save-output", str(out_file)] + flags) == 0
    line = capsys.readouterr().out.strip().splitlines()[-1]
    assert "method=load-path" in line and "plan_verification=verified-v2" in line and "check=PASS" in line, line
    out_net = {"inputs": [net["output"]], "output": net["output"], "size_dict": net["size_dict"]}
    got = netio.load_tensors_bin(out_file, out_net, dt)[0]
    wide = np.complex128 if "c" in dtype else np.float64
    want = O.contract(net, [a.astype(wide) for a in arrays], [], "naive")
    tol = 1e-5 if dtype in ("f32", "c32") else 1e-12
    assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max())


@pytest.mark.gpu
def test_tnexec_slice_oracle_check_on_circuit(cuda, tmp_path, capsys):
    """A 4x4 random circuit (more than 8 tensors, sliced): tnexec's check is the unsliced
    double-precision CPU contraction along the plan's path; the output matches the
    complex128 oracle."""
    from arctn_amd import tnexec
    from arctn_amd import workloads as W
    from oracle import pyoracle as O

    net, arrays, _ = W.random_circuit_amplitude(4, 4, 8, 0, open_qubits=range(2))
    doc = {"inputs": net["inputs"], "output": net["output"], "size_dict": {str(k): v for k, v in net["size_dict"].items()}}
    net_file, data, out_file = tmp_path / "c.net.json", tmp_path / "c.bin", tmp_path / "o.bin"
    net_file.write_text(json.dumps(doc))
    loaded, _ = netio.load_network_json(net_file)
    netio.save_tensors_bin(data, arrays)
    assert tnexec.main(["--net", str(net_file), "--method", "greedy", "--target-size", "128", "--single", "--complex",
                        "--data", str(data), "--check", "--save-output", str(out_file)]) == 0
    line = capsys.readouterr().out.strip().splitlines()[-1]
    assert "check=PASS(slice_oracle_" in line and "n_sliced_legs=0," not in line, line
    out_net = {"inputs": [loaded["output"]], "output": loaded["output"], "size_dict": loaded["size_dict"]}
    got = netio.load_tensors_bin(out_file, out_net, np.complex64)[0]
    arr = netio.load_tensors_bin(data, loaded, np.complex64)
    want = O.contract(loaded, [a.astype(np.complex128) for a in arr], __import__("arctn_amd").planner.greedy(loaded),
                      "compiled")
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
