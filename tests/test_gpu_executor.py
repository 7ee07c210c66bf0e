"""Executor parity on the GPU: plain and batched paths against the oracle,
the reference's known answers and NumPy einsum."""
import numpy as np
import pytest

from arctn_amd import CompiledContraction, Plan, contract_sliced
from arctn_amd import planner as P
from arctn_amd import workloads as W
from oracle import pyoracle as O
from tests.golden import fixtures as F

pytestmark = pytest.mark.gpu

DTYPES = [np.float32, np.float64, np.complex64, np.complex128]
FX = F.load()
MODES = [dict(), dict(batched=True), dict(batched=True, use_graph=True), dict(batched=True, max_batch=3)]


def _vals(fx, dtype):
    v = fx["values"] if "values" in fx else None
    if np.issubdtype(np.dtype(dtype), np.complexfloating):
        return np.array([complex(a, b) for a, b in v], dtype=dtype)
    return np.array([a for a, _ in v], dtype=dtype)


@pytest.mark.parametrize("name", ["mpi_matrix_product", "mpi_vector_dot"])
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("mode", range(len(MODES)))
def test_mpi_fixtures_exact(cuda, name, dtype, mode):
    """tests/mpi_cli.rs:72-148: exact results under every slicing."""
    fx = FX[name]
    net = F.net_of(fx)
    vals = _vals(fx, dtype)
    half = len(vals) // 2
    shape = [net["size_dict"][l] for l in net["inputs"][0]]
    arrays = [vals[:half].reshape(shape), vals[half:].reshape(shape)]
    cplx = np.issubdtype(np.dtype(dtype), np.complexfloating)
    exp = np.array([complex(a, b) for a, b in fx["complex_expected"]] if cplx else fx["real_expected"], dtype=dtype)
    for sl in fx["slicings"]:
        got = contract_sliced(net["inputs"], net["output"], net["size_dict"], arrays, sl, ssa_path=fx["ssa_path"],
                              **MODES[mode])
        assert got.dtype == np.dtype(dtype)
        assert np.array_equal(got.reshape(-1), exp), (sl, got)


@pytest.mark.parametrize("mode", range(len(MODES)))
def test_unary_output_blocks_known_answer(cuda, mode):
    """tests/slice_output.rs:52-70."""
    fx = FX["unary_output_blocks"]
    net = F.net_of(fx)
    a = np.array([[[100 * x + 10 * y + z for z in range(4)] for y in range(3)] for x in range(2)], dtype=np.float64)
    for sl in fx["slicings"]:
        got = contract_sliced(net["inputs"], net["output"], net["size_dict"], [a], sl, ssa_path=[], **MODES[mode])
        assert got.shape == (4, 2)
        assert np.array_equal(got.reshape(-1), np.array(fx["expected"])), sl


def test_slice_validation_dot(cuda):
    fx = FX["slice_validation_dot"]
    net = F.net_of(fx)
    a, b = np.array(fx["a"]), np.array(fx["b"])
    for sl in fx["slicings"]:
        got = contract_sliced(net["inputs"], net["output"], net["size_dict"], [a, b], sl, ssa_path=fx["ssa_path"])
        assert got.reshape(()) == fx["expected"]


def test_scalar32_matmul(cuda):
    """tests/scalar32.rs:99-148 tolerances against naive einsum."""
    fx = FX["scalar32_matmul"]
    net = F.net_of(fx)
    cc = CompiledContraction.compile(net["inputs"], net["output"], net["size_dict"], ssa_path=fx["ssa_path"])
    ra, rb = np.array(fx["real_a"], np.float32).reshape(2, 3), np.array(fx["real_b"], np.float32).reshape(3, 2)
    want = O.contract(net, [ra, rb], mode="naive")
    assert np.abs(cc.execute([ra, rb]) - want).max() < 2e-6
    ca = np.array([complex(*v) for v in fx["complex_a"]], np.complex64).reshape(2, 3)
    cb = np.array([complex(*v) for v in fx["complex_b"]], np.complex64).reshape(3, 2)
    want = O.contract(net, [ca, cb], mode="naive")
    assert np.abs(cc.execute([ca, cb]) - want).max() < 1e-5


@pytest.mark.parametrize("case", range(len(FX["output_slicing_cases"]["cases"])))
@pytest.mark.parametrize("dtype", DTYPES)
def test_output_slicing_cases_vs_einsum(cuda, case, dtype):
    """pybind/tests/test_output_slicing.py:49-65, 113-143."""
    name, eq, net, sliced, path = F.einsum_case(FX["output_slicing_cases"]["cases"][case])
    arrays = F.case_arrays(net, dtype)
    ref = np.einsum(eq, *arrays, optimize=False)
    single = np.dtype(dtype) in (np.dtype(np.float32), np.dtype(np.complex64))
    tol = 2e-4 if single else 1e-11
    for mode in MODES:
        got = contract_sliced(net["inputs"], net["output"], net["size_dict"], arrays, sliced, ssa_path=path, **mode)
        assert got.shape == ref.shape and got.dtype == np.dtype(dtype)
        np.testing.assert_allclose(got, ref, rtol=tol, atol=tol, err_msg=f"{name} {mode}")


@pytest.mark.parametrize("case", range(len(FX["output_slicing_cases"]["cases"])))
@pytest.mark.parametrize("dtype", DTYPES)
def test_arctn_contract_string_labels_vs_einsum(cuda, case, dtype):
    """arctn_contract (pybind/python/arctn/__init__.py:903-1072) with backend="hip" on the
    einsum cases of test_output_slicing.py, labels as strings: planned on the CPU, sliced
    to a small target (output slicing allowed) and executed on the GPU, against np.einsum."""
    from arctn_amd import arctn_contract

    name, eq, _, _, _ = F.einsum_case(FX["output_slicing_cases"]["cases"][case])
    dims = FX["output_slicing_cases"]["cases"][case][2]
    terms, out = eq.split("->")
    inputs = [list(t) for t in terms.split(",")]
    net = {"inputs": inputs, "output": list(out), "size_dict": dict(dims)}
    rng = np.random.default_rng(case)
    arrays = []
    for t in inputs:
        x = rng.standard_normal(tuple(dims[c] for c in t))
        if np.issubdtype(np.dtype(dtype), np.complexfloating):
            x = x + 1j * rng.standard_normal(x.shape)
        arrays.append(x.astype(dtype))
    ref = np.einsum(eq, *arrays, optimize=False)
    single = np.dtype(dtype) in (np.dtype(np.float32), np.dtype(np.complex64))
    tol = 2e-4 if single else 1e-11
    for target, out_slicing, preset in ((None, False, "greedy"), (2, True, "greedy"), (2, True, "rgreedy")):
        got, info = arctn_contract(net["inputs"], net["output"], net["size_dict"], arrays, target_size=target,
                                   allow_output_slicing=out_slicing, preset=preset, ntrials=3, return_info=True)
        assert got.shape == ref.shape and got.dtype == np.dtype(dtype)
        np.testing.assert_allclose(got, ref, rtol=tol, atol=tol, err_msg=f"{name} target={target}")
        assert info["execution_backend"] == "hip"


def test_arctn_contract_circuit_vs_oracle(cuda):
    """arctn_contract on the 4x4 random circuit amplitude (string labels), unsliced and
    sliced, against the complex128 oracle."""
    from arctn_amd import arctn_contract

    net, arrays, _ = W.random_circuit_amplitude(4, 4, 8, 0)
    want = complex(O.contract(net, [a.astype(np.complex128) for a in arrays], P.greedy(net), "compiled"))
    name = lambda l: f"q{l}"
    inputs = [[name(l) for l in t] for t in net["inputs"]]
    sd = {name(l): d for l, d in net["size_dict"].items()}
    for target, preset in ((None, "greedy"), (2 ** 8, "greedy"), (2 ** 8, "rgreedy")):
        got, info = arctn_contract(inputs, [], sd, arrays, target_size=target, preset=preset, ntrials=4,
                                   return_info=True)
        assert abs(complex(got) - want) / abs(want) < 1e-5, (target, preset)
        assert bool(info["sliced"]) == (target is not None)


def _random_case(seed, dims=(2, 3, 4)):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 8))
    net = W.random_connected(n, int(rng.integers(0, n)), list(dims), int(rng.integers(0, 3)), seed)
    return net, P.greedy(net)


@pytest.mark.parametrize("dtype", DTYPES)
def test_random_networks_plain_and_batched_vs_oracle(cuda, dtype):
    """Random connected networks (correctness.rs:43-103 style), every leg sliced in turn."""
    single = np.dtype(dtype) in (np.dtype(np.float32), np.dtype(np.complex64))
    tol = 1e-4 if single else 1e-10
    for seed in range(12):
        net, path = _random_case(seed)
        arrays = W.random_arrays(net, dtype, seed)
        want = O.contract(net, [a.astype(np.complex128) for a in arrays], path, "naive")
        scale = max(1.0, np.abs(want).max())
        got = CompiledContraction.compile(net["inputs"], net["output"], net["size_dict"], ssa_path=path).execute(arrays)
        assert np.abs(got - want).max() / scale < tol, seed
        legs = sorted(net["size_dict"])
        for sl in ([legs[0]], legs[: min(3, len(legs))], legs[-2:]):
            for mode in MODES:
                got = contract_sliced(net["inputs"], net["output"], net["size_dict"], arrays, sl, ssa_path=path, **mode)
                assert np.abs(got - want).max() / scale < tol, (seed, sl, mode)


def test_mixed_network_traces_and_diagonals(cuda):
    """compiled_pair_steps.rs:68-122: diagonal, trace, dangling, batch and hyperedge legs."""
    fx = FX["mixed_network"]
    net = F.net_of(fx)
    for path in (P.greedy(net), [tuple(p) for p in fx["hand_path"]]):
        for dtype in (np.float64, np.complex128):
            arrays = W.random_arrays(net, dtype, 7)
            want = O.contract(net, arrays, path, "naive")
            got = CompiledContraction.compile(net["inputs"], net["output"], net["size_dict"], ssa_path=path).execute(arrays)
            assert np.abs(got - want).max() < 1e-9 * max(1, np.abs(want).max())


def test_plan_description_matches_oracle(cuda):
    fx = FX["mixed_network"]
    net = F.net_of(fx)
    path = [tuple(p) for p in fx["hand_path"]]
    mine = Plan(net["inputs"], net["output"], net["size_dict"], path).describe()
    ref = O.describe(net, path)
    for s_m, s_r in zip(mine["steps"], ref["steps"]):
        s_m = dict(s_m)
        s_m.pop("varying")
        assert s_m == s_r
    assert mine["finalize"] == ref["finalize"]


def test_circuit_4x4_amplitude_vs_complex128_oracle(cuda):
    """16-qubit 4x4 random circuit, 8 cycles: one amplitude, rel. err <= 1e-5."""
    net, arrays, _ = W.random_circuit_amplitude(4, 4, 8, 0)
    path = P.greedy(net)
    want = O.contract(net, [a.astype(np.complex128) for a in arrays], path, "compiled")
    cc = CompiledContraction.compile(net["inputs"], net["output"], net["size_dict"], ssa_path=path)
    for mode in MODES[:2]:
        got = cc.execute(arrays, **mode)
        assert abs(complex(got) - complex(want)) / abs(complex(want)) < 1e-5


@pytest.mark.parametrize("dtype", DTYPES)
def test_grouped_small_steps_equal_per_step_path(cuda, dtype):
    """CompiledContraction.execute groups the tiny steps of each tree level into one
    launch; the result is bit-identical to the per-step path (same kernels, same
    summation order): the 4x4 circuit and random networks."""
    cases = []
    for seed in range(8):
        net, path = _random_case(seed)
        cases.append((net, W.random_arrays(net, dtype, seed), path))
    if np.issubdtype(np.dtype(dtype), np.complexfloating):
        net, arrays, _ = W.random_circuit_amplitude(4, 4, 8, 0, dtype=dtype)
        cases.append((net, arrays, P.greedy(net)))
    for net, arrays, path in cases:
        cc = CompiledContraction.compile(net["inputs"], net["output"], net["size_dict"], ssa_path=path)
        grouped, per_step = cc.execute(arrays), cc.execute(arrays, batched=False)
        assert grouped.tobytes() == per_step.tobytes(), len(net["inputs"])


def test_stripped_reconstructs_plain(cuda):
    net, arrays, _ = W.random_circuit_amplitude(4, 4, 8, 0)
    path = P.greedy(net)
    cc = CompiledContraction.compile(net["inputs"], net["output"], net["size_dict"], ssa_path=path)
    plain = complex(cc.execute(arrays))
    mant, e = cc.execute_stripped(arrays)
    assert abs(complex(mant) * 10.0 ** e - plain) / abs(plain) < 1e-5
    fx = FX["strip_extreme"]
    net = F.net_of(fx)
    cc = CompiledContraction.compile(net["inputs"], net["output"], net["size_dict"], ssa_path=fx["ssa_path"])
    m, e = cc.execute_stripped([np.array(1e200), np.array(1.0)])
    assert e > 199 and abs(float(m) * 10.0 ** e - 1e200) / 1e200 < 1e-12
    m, e = cc.execute_stripped([np.array(1e200 - 1e200j), np.array(1.0 + 0j)])
    assert abs(complex(m) * 10.0 ** e - (1e200 - 1e200j)) / abs(1e200 - 1e200j) < 1e-12
    # the reference's pair (compiled.rs:412-430): |mantissa| = 1 at the exact max, exponent = log10 |z|
    assert abs(abs(complex(m)) - 1.0) < 1e-15 and abs(e - np.log10(abs(1e200 - 1e200j))) < 1e-12


@pytest.mark.parametrize("dtype", DTYPES)
def test_stripped_pair_matches_oracle(cuda, dtype):
    """execute_stripped returns the reference's (mantissa, log10_exponent) pair: the last
    step divided by its exact max magnitude, the exponent the sum of log10 of the factors
    (compiled.rs:376-381, 412-430), compared with the oracle's stripped mode."""
    double = np.dtype(dtype) in (np.dtype(np.float64), np.dtype(np.complex128))
    tol, span = (1e-12, 100) if double else (1e-5, 15)
    for seed in range(12):
        rng = np.random.default_rng(seed)
        n = int(rng.integers(3, 9))
        net = W.random_connected(n, int(rng.integers(0, n)), [2, 3, 4], int(rng.integers(0, 3)), seed)
        path = P.greedy(net)
        arrays = [a * 10.0 ** int(rng.integers(-span, span)) for a in W.random_arrays(net, dtype, seed)]
        arrays = [np.asarray(a, dtype=dtype) for a in arrays]
        cc = CompiledContraction.compile(net["inputs"], net["output"], net["size_dict"], ssa_path=path)
        m, e = cc.execute_stripped(arrays)
        wide = np.complex128 if np.iscomplexobj(arrays[0]) else np.float64
        m_ref, e_ref = O.contract(net, [a.astype(wide) for a in arrays], path, "stripped")
        assert m.shape == m_ref.shape and m.dtype == np.dtype(dtype)
        assert abs(e - e_ref) <= tol * max(1.0, abs(e_ref)), (seed, e, e_ref)
        assert np.abs(m - m_ref).max() <= 4 * tol * max(1.0, np.abs(m_ref).max()), seed
        if abs(np.abs(m_ref).max() - 1.0) < 1e-12:  # finalize summed no residual leg
            assert abs(np.abs(m).max() - 1.0) <= tol, seed


def test_sliced_circuit_plain_batched_agree(cuda):
    """A sliced 4x4 circuit: plain (T sums in reference chunk order) vs batched (FP64 sums)."""
    net, arrays, _ = W.random_circuit_amplitude(4, 4, 8, 0)
    path = P.greedy(net)
    sliced = P.find_slices_to_size(net, path, 2 ** 7)
    assert sliced and len(sliced) >= 3
    want = O.contract(net, [a.astype(np.complex128) for a in arrays], path, "compiled")
    outs = []
    for mode in MODES:
        got = contract_sliced(net["inputs"], net["output"], net["size_dict"], arrays, sliced, ssa_path=path, **mode)
        assert abs(complex(got) - complex(want)) / abs(complex(want)) < 1e-5, mode
        outs.append(complex(got))
    assert outs[1] == outs[2]  # graph replay == eager batched, bit for bit


@pytest.mark.parametrize("use_graph", [False, True])
def test_one_executor_over_new_inputs(cuda, use_graph):
    """One cached batched executor run on three sets of device tensors, each against
    the complex128 oracle: the grouped launches of slice-invariant steps read plan
    inputs through a device table of input pointers that follows every new set (and
    is left alone on a repeat of the same set)."""
    import torch

    net, arrays, _ = W.random_circuit_amplitude(4, 4, 8, 0)
    path = P.greedy(net)
    sliced = P.find_slices_to_size(net, path, 2 ** 7)
    plan = Plan(net["inputs"], net["output"], net["size_dict"], path, sliced)
    rng = np.random.default_rng(3)
    sets = [arrays] + [[(a + 0.3 * (rng.standard_normal(a.shape) + 1j * rng.standard_normal(a.shape))).astype(a.dtype)
                        for a in arrays] for _ in range(2)]
    for arrs in sets + sets[:1]:
        want = complex(O.contract(net, [a.astype(np.complex128) for a in arrs], path, "compiled"))
        dev = [torch.from_numpy(a).cuda() for a in arrs]
        for _ in range(2):  # a repeat of the same tensors keeps the table
            got, _ = plan.run(dev, 2, batched=True, use_graph=use_graph)
            assert abs(complex(got.cpu()) - want) / abs(want) < 1e-5


def test_slice_ranges_partial_sums(cuda):
    """Batched runs over global slice ranges return the partial sums of those
    slices (one device's share, tnmpi.rs:279-284): internal-only ranges match the
    oracle's slice-range contraction, and consecutive ranges add up to the whole."""
    import torch

    net, arrays, _ = W.random_circuit_amplitude(4, 4, 8, 0)
    path = P.greedy(net)
    sliced = P.find_slices_to_size(net, path, 2 ** 7)
    a128 = [a.astype(np.complex128) for a in arrays]
    plan = Plan(net["inputs"], net["output"], net["size_dict"], path, sliced)
    n = plan.info()["n_slices"]
    dev = [torch.from_numpy(a).cuda() for a in a128]
    cuts = [0, 1, n // 3, n // 2 + 1, n]
    total = 0
    for lo, hi in zip(cuts[:-1], cuts[1:]):
        for graph in (False, True):
            got, _ = plan.run(dev, 3, batched=True, use_graph=graph, slice_range=(lo, hi), max_batch=5)
            ref = O.contract_slice_range(net, a128, path, sliced, lo, hi)
            assert abs(complex(got.cpu()) - complex(ref)) <= 1e-12 * max(abs(complex(ref)), 1e-300), (lo, hi, graph)
        total += complex(got.cpu())
    full, _ = plan.run(dev, 3, batched=True)
    assert abs(total - complex(full.cpu())) <= 1e-12 * abs(complex(full.cpu()))
    # output-sliced plan: ranges may cross block boundaries
    fx = FX["mixed_ragged_chunks"]
    fnet = F.net_of(fx)
    arrs = [torch.from_numpy(a).cuda() for a in W.random_arrays(fnet, np.float64, 3)]
    fplan = Plan(fnet["inputs"], fnet["output"], fnet["size_dict"], fx["ssa_path"], fx["sliced"])
    nf = fplan.info()["n_slices"]
    whole, _ = fplan.run(arrs, 1, batched=True)
    parts = sum(fplan.run(arrs, 1, batched=True, slice_range=r)[0].cpu().numpy()
                for r in [(0, 100), (100, nf // 2 + 7), (nf // 2 + 7, nf)])
    np.testing.assert_allclose(parts, whole.cpu().numpy(), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("precision", ["native", "c3m", "bf16x3", "fp16x2", "bf16x3_3m", "fp16x2_3m"])
def test_autotuned_batched_graph_forms(cuda, precision):
    """The autotuner's per-step kernel / tile choices keep the result within 1e-5."""
    net, arrays, _ = W.random_circuit_amplitude(4, 4, 8, 0)
    path = P.greedy(net)
    sliced = P.find_slices_to_size(net, path, 2 ** 7)
    want = complex(O.contract(net, [a.astype(np.complex128) for a in arrays], path, "compiled"))
    for graph in (False, True):
        got = contract_sliced(net["inputs"], net["output"], net["size_dict"], arrays, sliced, ssa_path=path,
                              batched=True, use_graph=graph, autotune=True, precision=precision)
        assert abs(complex(got) - want) / abs(want) < 1e-5, graph


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.complex64])
def test_gather_tables_match_permuted_operands(cuda, dtype):
    """B operands read through gather tables (no permute) give the results of the
    permuted operands: random networks in every form, and a sliced circuit in every
    precision (against the complex128 oracle)."""
    import torch

    tol = 1e-12 if np.dtype(dtype) in (np.dtype(np.float64), np.dtype(np.complex128)) else 1e-5
    for seed in range(10):
        net, path = _random_case(seed)
        arrays = W.random_arrays(net, dtype, seed)
        want = O.contract(net, [a.astype(np.complex128) for a in arrays], path, "naive")
        scale = max(1.0, np.abs(want).max())
        plan = Plan(net["inputs"], net["output"], net["size_dict"], path, sorted(net["size_dict"])[:1])
        dev = [torch.from_numpy(a).cuda() for a in arrays]
        code = {np.dtype(np.float64): 1, np.dtype(np.complex128): 3, np.dtype(np.complex64): 2}[np.dtype(dtype)]
        for mode in MODES:
            g1, _ = plan.run(dev, code, gather=True, **mode)
            g0, _ = plan.run(dev, code, gather=False, **mode)
            g1, g0 = g1.cpu().numpy(), g0.cpu().numpy()
            assert np.abs(g1 - g0).max() <= tol * scale, (seed, mode)
            assert np.abs(g1 - want).max() <= max(tol, 1e-4 if tol > 1e-6 else tol) * scale * 10, (seed, mode)
    if dtype != np.complex64:
        return
    net, arrays, _ = W.random_circuit_amplitude(4, 4, 8, 0)
    path = P.greedy(net)
    sliced = P.find_slices_to_size(net, path, 2 ** 7)
    want = complex(O.contract(net, [a.astype(np.complex128) for a in arrays], path, "compiled"))
    plan = Plan(net["inputs"], net["output"], net["size_dict"], path, sliced)
    dev = [torch.from_numpy(a).cuda() for a in arrays]
    for prec in ("native", "c3m", "bf16x3", "fp16x2", "bf16x3_3m", "fp16x2_3m"):
        for gather in (True, False):
            got, _ = plan.run(dev, 2, batched=True, use_graph=True, precision=prec, gather=gather)
            assert abs(complex(got.cpu()) - want) / abs(want) < 1e-5, (prec, gather)


@pytest.mark.parametrize("dtype", [np.float32, np.complex64])
def test_forced_gemm_kernels_with_gather_tables(cuda, dtype):
    """Every GEMM kernel forced on every step, with B read through gather tables
    and permuted, in each precision: random networks (dims up to 8, so most steps
    leave partial tiles) and a sliced circuit against the complex128 oracle."""
    import torch

    code = {np.dtype(np.float32): 0, np.dtype(np.complex64): 2}[np.dtype(dtype)]
    cases = []
    for seed in range(6):
        net, path = _random_case(seed, dims=(3, 5, 8))
        cases.append((net, W.random_arrays(net, dtype, seed), path, sorted(net["size_dict"])[:1]))
    if dtype == np.complex64:
        net, arrays, _ = W.random_circuit_amplitude(5, 4, 10, 1)
        path = P.greedy(net)
        cases.append((net, arrays, path, P.find_slices_to_size(net, path, 2 ** 9)))
    for net, arrays, path, sliced in cases:
        want = O.contract(net, [a.astype(np.complex128) for a in arrays], path, "compiled")
        scale = max(np.abs(want).max(), 1e-30)
        plan = Plan(net["inputs"], net["output"], net["size_dict"], path, sliced)
        dev = [torch.from_numpy(a).cuda() for a in arrays]
        # BF16x3 drops products of 2^-16 relative size: forced on all ~300 steps of
        # the circuit (the heuristic keeps it to large steps) its error adds up
        forms = [("native", (0, 1, 2, 3, 5, 6, 7, 8), 1e-5), ("bf16x3", (1, 4, 8), 3e-4), ("fp16x2", (1, 4, 8), 1e-5)]
        if dtype == np.complex64:
            forms += [("c3m", (1, 5), 1e-5), ("bf16x3_3m", (8,), 3e-4), ("fp16x2_3m", (8,), 1e-5)]
        for prec, kernels, tol in forms:
            for kern in kernels:
                for gather in (True, False):
                    got, _ = plan.run(dev, code, batched=True, precision=prec, gather=gather, gemm_kernel=kern)
                    err = np.abs(got.cpu().numpy() - want).max() / scale
                    assert err < tol, (len(net["inputs"]), prec, kern, gather, err)
        plan.release()


@pytest.mark.parametrize("dtype", [np.float32, np.complex64])
def test_split_gemm_lanes_along_k_for_k_innermost_operands(cuda, dtype):
    """B whose contracted legs are its innermost legs (rows in runs of consecutive
    offsets, columns far apart): the B-staging split kernel reads it with lanes
    along k. Ragged sizes, every precision, against NumPy in complex128."""
    import torch

    code = {np.dtype(np.float32): 0, np.dtype(np.complex64): 2}[np.dtype(dtype)]
    rng = np.random.default_rng(11)
    for dims in ({0: 70, 1: 8, 2: 5, 3: 150}, {0: 130, 1: 4, 2: 2, 3: 97}, {0: 64, 1: 16, 2: 4, 3: 300}):
        net = {"inputs": [[0, 1, 2], [3, 1, 2]], "output": [0, 3], "size_dict": dims}
        arrays = [rng.standard_normal(tuple(dims[l] for l in t)).astype(dtype) for t in net["inputs"]]
        if np.issubdtype(np.dtype(dtype), np.complexfloating):
            arrays = [(a + 1j * rng.standard_normal(a.shape)).astype(dtype) for a in arrays]
        want = np.einsum("ikl,jkl->ij", *[a.astype(np.complex128) for a in arrays])
        plan = Plan(net["inputs"], net["output"], net["size_dict"], [(0, 1)])
        dev = [torch.from_numpy(a).cuda() for a in arrays]
        for prec, tol in (("native", 1e-5), ("bf16x3", 2e-4), ("fp16x2", 1e-5)):
            for kern in (8, 4, 1):
                got, _ = plan.run(dev, code, batched=True, precision=prec, gemm_kernel=kern)
                err = np.abs(got.cpu().numpy() - want).max() / np.abs(want).max()
                assert err < tol, (dims, prec, kern, err)
        plan.release()
