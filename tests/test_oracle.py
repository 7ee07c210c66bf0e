"""Pins the oracle (CPU restatement) before it is trusted as the checker.

Every oracle function is checked against something independent of it: the
reference's known answers (tests/golden), NumPy (transpose, sum, trace,
diagonal, take, einsum), brute force, closed forms, and central finite
differences for the adjoint.
"""
import itertools
import math

import numpy as np
import pytest

from arctn_amd import planner as P
from arctn_amd import workloads as W
from oracle import pyoracle as O
from tests.golden import fixtures as F

FX = F.load()
DTYPES = [np.float32, np.float64, np.complex64, np.complex128]


def _rand(shape, dtype, rng):
    a = rng.standard_normal(shape)
    if np.issubdtype(np.dtype(dtype), np.complexfloating):
        a = a + 1j * rng.standard_normal(shape)
    return np.ascontiguousarray(a.astype(dtype))


@pytest.mark.parametrize("dtype", DTYPES)
def test_tensor_ops_match_numpy(dtype):
    rng = np.random.default_rng(0)
    for shape in [(2, 3, 4), (5,), (2, 2, 2, 2, 2, 2), (3, 1, 4, 2)]:
        a = _rand(shape, dtype, rng)
        for perm in itertools.islice(itertools.permutations(range(len(shape))), 12):
            assert np.array_equal(O.tensor_op("permute", a, list(perm)), np.transpose(a, perm))
        for ax in range(len(shape)):
            s = O.tensor_op("sum_axis", a, ax)
            # sequential order: compare against an explicit left fold
            fold = np.zeros(s.shape, dtype=dtype)
            for i in range(shape[ax]):
                fold = fold + np.take(a, i, axis=ax)
            assert s.tobytes() == fold.tobytes()
            for i in range(shape[ax]):
                assert np.array_equal(O.tensor_op("select_axis", a, ax, i), np.take(a, i, axis=ax))
    b = _rand((3, 2, 4, 2, 5), dtype, rng)
    t = O.tensor_op("trace_pair", b, 1, 3)
    np.testing.assert_allclose(t, np.trace(b, axis1=1, axis2=3), rtol=1e-5)
    d = O.tensor_op("diag_pair", b, 1, 3)
    assert np.array_equal(d, np.moveaxis(np.diagonal(b, axis1=1, axis2=3), -1, 1))


def test_mpi_fixtures_all_dtypes_and_slicings():
    for name in ("mpi_matrix_product", "mpi_vector_dot"):
        fx = FX[name]
        net = F.net_of(fx)
        shape = [net["size_dict"][l] for l in net["inputs"][0]]
        for dtype in DTYPES:
            cplx = np.issubdtype(np.dtype(dtype), np.complexfloating)
            v = np.array([complex(a, b) if cplx else a for a, b in fx["values"]], dtype=dtype)
            h = len(v) // 2
            arrays = [v[:h].reshape(shape), v[h:].reshape(shape)]
            exp = np.array([complex(*c) for c in fx["complex_expected"]] if cplx else fx["real_expected"], dtype=dtype)
            for sl in fx["slicings"]:
                got = O.contract(net, arrays, fx["ssa_path"], "sliced", sl)
                assert np.array_equal(got.reshape(-1), exp), (name, dtype, sl)
            for mode in ("compiled", "direct", "naive"):
                assert np.array_equal(O.contract(net, arrays, fx["ssa_path"], mode).reshape(-1), exp)


def test_slice_validation_dot_and_rejections():
    fx = FX["slice_validation_dot"]
    net = F.net_of(fx)
    a, b = np.array(fx["a"]), np.array(fx["b"])
    for sl in fx["slicings"]:
        assert O.contract(net, [a, b], fx["ssa_path"], "sliced", sl).reshape(()) == fx["expected"]
    with pytest.raises(O.OracleError, match="repeated"):
        O.contract(net, [a, b], fx["ssa_path"], "sliced", [0, 0])
    trace = {"inputs": [[0, 0]], "output": [], "size_dict": {0: 2}}
    with pytest.raises(O.OracleError, match="repeated within one input"):
        O.contract(trace, [np.eye(2)], [], "sliced", [0])


def test_unary_output_blocks_known_answer():
    fx = FX["unary_output_blocks"]
    net = F.net_of(fx)
    a = np.array([[[100 * x + 10 * y + z for z in range(4)] for y in range(3)] for x in range(2)], dtype=np.float64)
    for sl in fx["slicings"]:
        got = O.contract(net, [a], [], "sliced", sl)
        assert got.shape == (4, 2) and np.array_equal(got.reshape(-1), fx["expected"])


def test_scalar32_matches_naive():
    fx = FX["scalar32_matmul"]
    net = F.net_of(fx)
    ra = np.array(fx["real_a"], np.float32).reshape(2, 3)
    rb = np.array(fx["real_b"], np.float32).reshape(3, 2)
    np.testing.assert_allclose(O.contract(net, [ra, rb], fx["ssa_path"]), ra.astype(np.float64) @ rb, atol=2e-6)
    ca = np.array([complex(*c) for c in fx["complex_a"]], np.complex64).reshape(2, 3)
    cb = np.array([complex(*c) for c in fx["complex_b"]], np.complex64).reshape(3, 2)
    np.testing.assert_allclose(O.contract(net, [ca, cb], fx["ssa_path"]), ca.astype(np.complex128) @ cb, atol=1e-5)


def test_geometry_liveness_and_costs():
    fx = FX["two_node_geometry"]
    d = O.describe(F.net_of(fx), fx["ssa_path"])
    assert d["steps"][0]["gemm"] == fx["gemm_dims"] and d["steps"][0]["result_legs"] == fx["result_legs"]
    assert all(not p["traces"] for p in d["preps"]) and d["finalize"]["sum"] == [] and d["finalize"]["perm"] == [0, 1]
    fx = FX["liveness_peak"]
    st = O.simulate(F.net_of(fx), fx["ssa_path"])
    assert st["log2_max_size"] == fx["log2_max_size"] and st["log2_peak_size"] == fx["log2_peak_size"]
    assert O.describe(F.net_of(fx), fx["ssa_path"])["log2_peak_size"] == fx["log2_peak_size"]
    fx = FX["chain_costs"]
    st = O.simulate(F.net_of(fx), fx["ssa_path"])
    assert abs(st["log10_flops"] - math.log10(fx["flops"])) < 1e-12
    assert abs(st["log2_read_write"] - math.log2(fx["read_write"])) < 1e-12
    # the Python planner's simulation agrees with the oracle's
    ps, _ = P.simulate_path_full(F.net_of(fx), [tuple(p) for p in fx["ssa_path"]])
    assert abs(ps.log10_flops - st["log10_flops"]) < 1e-12 and abs(ps.log2_peak_size - st["log2_peak_size"]) < 1e-12


def test_slot_invariants_of_compiled_plan():
    """compiled_pair_steps.rs:38-57: each step consumes two live slots and writes n_inputs + s."""
    fx = FX["mixed_network"]
    net = F.net_of(fx)
    for path in (P.greedy(net), [tuple(p) for p in fx["hand_path"]]):
        d = O.describe(net, path)
        live = {i: p["legs"] for i, p in enumerate(d["preps"])}
        for s, st in enumerate(d["steps"]):
            assert st["out"] == d["n_inputs"] + s and (st["li"], st["ri"]) == tuple(path[s])
            assert live.pop(st["li"]) == st["left_legs"] and live.pop(st["ri"]) == st["right_legs"]
            assert [net["size_dict"][l] for l in st["result_legs"]] == st["result_shape"]
            live[st["out"]] = st["result_legs"]
        assert list(live) == [d["finalize"]["slot"]]


def _rand_case(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 8))
    net = W.random_connected(n, int(rng.integers(0, n)), [2, 3, 4], int(rng.integers(0, 3)), seed)
    return net, P.greedy(net)


def test_compiled_equals_direct_bitwise_and_matches_einsum():
    """correctness.rs:43-103: compiled == direct bit for bit; both match naive/einsum."""
    for seed in range(25):
        net, path = _rand_case(seed)
        for dtype in (np.float64, np.complex128):
            arrays = W.random_arrays(net, dtype, seed)
            c = O.contract(net, arrays, path, "compiled")
            d = O.contract(net, arrays, path, "direct")
            assert c.tobytes() == d.tobytes()
            n = O.contract(net, arrays, path, "naive")
            sub = "".join(chr(97 + l) for l in range(len(net["size_dict"])))
            eq = ",".join("".join(sub[l] for l in t) for t in net["inputs"]) + "->" + "".join(sub[l] for l in net["output"])
            e = np.einsum(eq, *arrays, optimize=False)
            np.testing.assert_allclose(n, e, rtol=1e-9, atol=1e-9 * max(1, np.abs(e).max()))
            np.testing.assert_allclose(c, e, rtol=1e-9, atol=1e-9 * max(1, np.abs(e).max()))


def test_sliced_matches_unsliced_every_subset():
    """slice_output.rs:35-50: every subset of legs sliced, internal and output."""
    fx = FX["shared_output_hyperedge"]
    net = F.net_of(fx)
    labels = sorted(net["size_dict"])
    for dtype, tol in ((np.float64, 1e-11), (np.complex64, 1e-4)):
        arrays = W.random_arrays(net, dtype, 9817)
        want = O.contract(net, arrays, fx["ssa_path"], "naive")
        for mask in range(1 << len(labels)):
            sl = [l for i, l in enumerate(labels) if mask >> i & 1][::-1]
            got = O.contract(net, arrays, fx["ssa_path"], "sliced", sl)
            assert np.abs(got - want).max() < tol, sl


def _einsum_eq(net, legs_fixed=()):
    sub = {l: chr(97 + i) for i, l in enumerate(sorted(net["size_dict"]))}
    terms = ["".join(sub[l] for l in t if l not in legs_fixed) for t in net["inputs"]]
    return ",".join(terms) + "->" + "".join(sub[l] for l in net["output"])


def _fix_legs(net, arrays, legs, values):
    """Inputs with every sliced leg fixed by np.take (independent of the oracle's select_axis)."""
    out = []
    for t, a in zip(net["inputs"], arrays):
        for ax in range(len(t) - 1, -1, -1):
            if t[ax] in legs:
                a = np.take(a, values[legs.index(t[ax])], axis=ax)
        out.append(a)
    return out


def test_slice_range_pinned():
    """contract_slice_range (the sum over internal slices [lo, hi), src/slice.rs:1018-1064):
    the full range equals the sliced contraction bit for bit and the naive result; consecutive
    ranges add up to the whole; a one-slice range equals np.einsum of the network with the
    sliced legs fixed by np.take at np.unravel_index(s, dims) (last sliced leg fastest)."""
    checked = 0
    for seed in range(40):
        net, path = _rand_case(seed)
        internal = [l for l in sorted(net["size_dict"]) if l not in net["output"]
                    and not any(t.count(l) > 1 for t in net["inputs"])][:3]
        if len(internal) < 2:
            continue
        dims = [net["size_dict"][l] for l in internal]
        n = int(np.prod(dims))
        for dtype, tol in ((np.float64, 1e-10), (np.complex128, 1e-10)):
            arrays = W.random_arrays(net, dtype, seed)
            whole = O.contract_slice_range(net, arrays, path, internal, 0, n)
            assert whole.tobytes() == O.contract(net, arrays, path, "sliced", internal).tobytes()
            naive = O.contract(net, arrays, path, "naive")
            scale = max(1.0, np.abs(naive).max())
            assert np.abs(whole - naive).max() <= tol * scale
            a, b = n // 3, (2 * n) // 3 + 1
            parts = sum(O.contract_slice_range(net, arrays, path, internal, lo, hi)
                        for lo, hi in ((0, a), (a, b), (b, n)) if lo < hi)
            assert np.abs(parts - whole).max() <= tol * scale
            for s in (0, n // 2, n - 1):
                vals = [int(v) for v in np.unravel_index(s, dims)]
                want = np.einsum(_einsum_eq(net, internal), *_fix_legs(net, arrays, internal, vals), optimize=False)
                got = O.contract_slice_range(net, arrays, path, internal, s, s + 1)
                assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max()), (seed, s)
        checked += 1
    assert checked >= 10


def test_output_slicing_cases_vs_einsum():
    for case in FX["output_slicing_cases"]["cases"]:
        name, eq, net, sliced, path = F.einsum_case(case)
        for dtype in DTYPES:
            arrays = F.case_arrays(net, dtype)
            ref = np.einsum(eq, *arrays, optimize=False)
            tol = 2e-4 if np.dtype(dtype) in (np.dtype(np.float32), np.dtype(np.complex64)) else 1e-11
            np.testing.assert_allclose(O.contract(net, arrays, path, "sliced", sliced), ref, rtol=tol, atol=tol)


def test_slice_chunk_formula_closed_form():
    """slice_parallel.rs:97-136: chunks cover [0, n) contiguously, none empty, <= 256."""
    for n in list(range(1, 601)) + [1 << 20, 10 ** 9 + 7]:
        for out in (1, 3, 1 << 10, 1 << 22, 1 << 24):
            c = O.slice_parallel_chunks(n, out)
            assert 1 <= c <= 256 and c <= n
            chunk = -(-n // c)
            assert (c - 1) * chunk < n <= c * chunk  # last chunk non-empty
            if n > 1 and out <= (1 << 22) // 2:
                assert c >= min(n, 256, (1 << 22) // out) // 2
    assert O.slice_parallel_chunks(257, 3) == 129


def test_ragged_chunks_deterministic_across_threads():
    fx = FX["mixed_ragged_chunks"]
    net = F.net_of(fx)
    arrays = W.random_arrays(net, np.float64, 377)
    want = O.contract(net, arrays, fx["ssa_path"], "naive")
    bits = None
    for threads in (1, 2, 3, 8):
        got = O.contract(net, arrays, fx["ssa_path"], "sliced", fx["sliced"], threads=threads)
        assert np.abs(got - want).max() < 1e-10
        bits = bits or got.tobytes()
        assert got.tobytes() == bits


def test_stripped_reconstructs():
    fx = FX["strip_extreme"]
    net = F.net_of(fx)
    m, e = O.contract(net, [np.array(1e200), np.array(1.0)], fx["ssa_path"], "stripped")
    assert e > 199 and abs(abs(float(m)) - 1) < 1e-12 and abs(float(m) * 10 ** e - 1e200) / 1e200 < 1e-12
    z = 1e200 - 1e200j
    m, e = O.contract(net, [np.array(z), np.array(1 + 0j)], fx["ssa_path"], "stripped")
    assert abs(abs(complex(m)) - 1) < 1e-12 and abs(complex(m) * 10 ** e - z) / abs(z) < 1e-12
    for seed in range(5):
        net, path = _rand_case(seed)
        arrays = W.random_arrays(net, np.float64, seed)
        plain = O.contract(net, arrays, path)
        m, e = O.contract(net, arrays, path, "stripped")
        np.testing.assert_allclose(m * 10 ** e, plain, rtol=1e-9, atol=1e-9 * np.abs(plain).max())


def test_adjoint_matches_finite_differences():
    """Oracle gradient vs central differences of the real loss L = Re(sum(conj(w) * out))."""
    for seed in range(4):
        net, path = _rand_case(seed)
        arrays = W.random_arrays(net, np.complex128, seed)
        out = O.contract(net, arrays, path)
        rng = np.random.default_rng(seed)
        w = rng.standard_normal(out.shape) + 1j * rng.standard_normal(out.shape)
        # dL/d(conj out) for L = Re(sum conj(w) out) is w/2 in the conjugate Wirtinger
        # convention; PyTorch's grad convention passes w directly (grad = 2 dL/dconj z).
        grads = O.gradient(net, arrays, path, w)
        h = 1e-6
        for i, a in enumerate(arrays):
            flat = a.reshape(-1)
            for j in range(min(flat.size, 3)):
                for unit in (1.0, 1j):
                    ap = [x.copy() for x in arrays]
                    am = [x.copy() for x in arrays]
                    ap[i].reshape(-1)[j] += h * unit
                    am[i].reshape(-1)[j] -= h * unit
                    lp = np.real(np.sum(np.conj(w) * O.contract(net, ap, path)))
                    lm = np.real(np.sum(np.conj(w) * O.contract(net, am, path)))
                    fd = (lp - lm) / (2 * h)
                    g = grads[i].reshape(-1)[j]
                    an = g.real if unit == 1.0 else g.imag
                    assert abs(fd - an) < 1e-6 * max(1.0, abs(fd)), (seed, i, j, unit, fd, an)


def test_naive_rejects_large_space():
    net = {"inputs": [list(range(30))], "output": [], "size_dict": {l: 2 for l in range(30)}}
    with pytest.raises(O.OracleError):
        O.contract(net, [np.zeros([2] * 30, np.float32)], [], "naive")
