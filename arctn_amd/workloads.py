"""Synthetic tensor networks for the executor's workloads (seeded, deterministic).

Every generator returns ``(net, arrays)`` where ``net`` is a dict
``{"inputs", "output", "size_dict"}`` with integer legs (the layout of
src/network.rs:135-143) and ``arrays`` is a list of NumPy arrays in
``net["inputs"]`` order, row-major, axis ``i`` of input ``t`` carrying leg
``net["inputs"][t][i]``.

* ``random_circuit_amplitude``: grid random-circuit amplitude networks
  (Haar single-qubit layer + fSim(pi/2, pi/6) on one of four alternating
  nearest-neighbour edge patterns per cycle), kets |0>, bra a seeded
  bitstring, optionally some qubits left open as output legs.
* ``brickwork_expectation``: <psi| Z_0 Z_1 |psi> of a 1D RY.RZ + CZ
  brickwork circuit as a doubled bra-ket network, parameterised by angles.
* ``random_regular``: configuration-model d-regular graph networks without
  self-loops (the rule of src/network.rs:432-458), per-edge dims.
* ``random_connected``: small random connected networks with open legs (the
  shape of src/network.rs:492-540), for parity tests.
"""

from __future__ import annotations

import numpy as np


def haar_unitary_2x2(rng):
    z = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))[None, :]


def fsim(theta, phi):
    c, s = np.cos(theta), np.sin(theta)
    return np.array(
        [[1, 0, 0, 0], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [0, 0, 0, np.exp(-1j * phi)]],
        dtype=np.complex128,
    )


def grid_edge_pattern(rows, cols, pattern):
    """Edges of pattern 0..3: horizontal even/odd columns, vertical even/odd rows."""
    edges = []
    q = lambda r, c: r * cols + c
    if pattern in (0, 1):
        for r in range(rows):
            for c in range(pattern, cols - 1, 2):
                edges.append((q(r, c), q(r, c + 1)))
    else:
        p = pattern - 2
        for r in range(p, rows - 1, 2):
            for c in range(cols):
                edges.append((q(r, c), q(r + 1, c)))
    return edges


class _Builder:
    def __init__(self):
        self.inputs, self.arrays, self.size_dict = [], [], {}
        self.next_leg = 0

    def leg(self, dim=2):
        l = self.next_leg
        self.next_leg += 1
        self.size_dict[l] = dim
        return l

    def add(self, legs, array):
        assert tuple(array.shape) == tuple(self.size_dict[l] for l in legs)
        self.inputs.append(list(legs))
        self.arrays.append(array)


def random_circuit_amplitude(rows, cols, cycles, seed, open_qubits=(), dtype=np.complex64):
    """Random-circuit amplitude network <x| C |0...0> on a rows x cols grid.

    Gate tensors are stored as U[out, in] (single qubit) and
    G[out_a, out_b, in_a, in_b] (two qubit). Qubits in ``open_qubits`` keep
    their final leg open; output legs are ordered by qubit index.
    Returns ``(net, arrays, bitstring)``.
    """
    rng = np.random.default_rng(seed)
    bits = np.random.default_rng([seed, 1]).integers(0, 2, size=rows * cols)
    n = rows * cols
    b = _Builder()
    cur = []
    ket0 = np.array([1.0, 0.0], dtype=np.complex128)
    for _ in range(n):
        l = b.leg()
        b.add([l], ket0)
        cur.append(l)
    g2 = fsim(np.pi / 2, np.pi / 6).reshape(2, 2, 2, 2)
    for cyc in range(cycles):
        for qb in range(n):
            u = haar_unitary_2x2(rng)
            nl = b.leg()
            b.add([nl, cur[qb]], u)
            cur[qb] = nl
        for qa, qb in grid_edge_pattern(rows, cols, cyc % 4):
            na, nb = b.leg(), b.leg()
            b.add([na, nb, cur[qa], cur[qb]], g2)
            cur[qa], cur[qb] = na, nb
    output = []
    open_set = set(open_qubits)
    for qb in range(n):
        if qb in open_set:
            output.append(cur[qb])
        else:
            e = np.zeros(2, dtype=np.complex128)
            e[bits[qb]] = 1.0
            b.add([cur[qb]], e)
    net = {"inputs": b.inputs, "output": output, "size_dict": b.size_dict}
    arrays = [np.ascontiguousarray(a.astype(dtype)) for a in b.arrays]
    return net, arrays, bits


def ry(theta):
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def rz(phi):
    return np.array([[np.exp(-0.5j * phi), 0], [0, np.exp(0.5j * phi)]], dtype=np.complex128)


def brickwork_structure(n_qubits, layers):
    """Doubled bra-ket network structure for <psi|Z_0 Z_1|psi>.

    Returns ``(net, gate_slots)`` where ``gate_slots`` describes every input:
    ``("ket0",)``, ``("u", layer, qubit, conj)``, ``("cz", conj)``, ``("z",)``,
    ``("id",)``. The ket half uses U[out,in]; the bra half uses conj(U) with
    the same leg roles, and the two halves meet through Z (qubits 0, 1) or an
    identity on every other qubit.
    """
    b = _Builder()
    slots = []
    dummy = lambda legs: np.zeros([2] * len(legs), dtype=np.complex128)
    ends = []
    for conj in (False, True):
        cur = []
        for _ in range(n_qubits):
            l = b.leg()
            b.add([l], dummy([l]))
            slots.append(("ket0",))
            cur.append(l)
        for layer in range(layers):
            for q in range(n_qubits):
                nl = b.leg()
                b.add([nl, cur[q]], dummy([0, 0]))
                slots.append(("u", layer, q, conj))
                cur[q] = nl
            for q in range(layer % 2, n_qubits - 1, 2):
                na, nb = b.leg(), b.leg()
                b.add([na, nb, cur[q], cur[q + 1]], dummy([0, 0, 0, 0]))
                slots.append(("cz", conj))
                cur[q], cur[q + 1] = na, nb
        ends.append(cur)
    for q in range(n_qubits):
        b.add([ends[1][q], ends[0][q]], dummy([0, 0]))
        slots.append(("z",) if q < 2 else ("id",))
    net = {"inputs": b.inputs, "output": [], "size_dict": b.size_dict}
    return net, slots


def brickwork_angles(n_qubits, layers, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-np.pi, np.pi, size=(layers, n_qubits, 2))


def brickwork_arrays(slots, angles, dtype=np.complex128):
    """Input tensors of the doubled network for the given angles (NumPy)."""
    cz = np.diag([1, 1, 1, -1]).astype(np.complex128).reshape(2, 2, 2, 2)
    z = np.diag([1.0, -1.0]).astype(np.complex128)
    out = []
    for s in slots:
        if s[0] == "ket0":
            out.append(np.array([1.0, 0.0], dtype=np.complex128))
        elif s[0] == "u":
            _, layer, q, conj = s
            u = rz(angles[layer, q, 1]) @ ry(angles[layer, q, 0])
            out.append(np.conj(u) if conj else u)
        elif s[0] == "cz":
            out.append(cz)
        elif s[0] == "z":
            out.append(z)
        else:
            out.append(np.eye(2, dtype=np.complex128))
    return [np.ascontiguousarray(a.astype(dtype)) for a in out]


def random_regular(n, degree, dims, seed, dtype=np.float32):
    """d-regular multigraph network without self-loops, per-edge dims from ``dims``.

    Stub shuffling is resampled until no pair is a self-loop (the rule of
    src/network.rs:432-444). Entries are N(0, 1).
    """
    assert (n * degree) % 2 == 0
    rng = np.random.default_rng(seed)
    stubs = np.repeat(np.arange(n), degree)
    while True:
        rng.shuffle(stubs)
        pairs = stubs.reshape(-1, 2)
        if np.all(pairs[:, 0] != pairs[:, 1]):
            break
    inputs = [[] for _ in range(n)]
    size_dict = {}
    for leg, (a, b) in enumerate(pairs.tolist()):
        inputs[a].append(leg)
        inputs[b].append(leg)
        size_dict[leg] = int(dims[rng.integers(0, len(dims))])
    net = {"inputs": inputs, "output": [], "size_dict": size_dict}
    arrays = [
        np.ascontiguousarray(rng.standard_normal([size_dict[l] for l in legs]).astype(dtype))
        for legs in inputs
    ]
    return net, arrays


def random_connected(n, extra, dims, n_open, seed):
    """Small connected random network with open legs (shape of src/network.rs:492-540)."""
    rng = np.random.default_rng(seed)
    inputs = [[] for _ in range(n)]
    size_dict = {}
    nxt = [0]

    def add_edge(a, b):
        inputs[a].append(nxt[0])
        inputs[b].append(nxt[0])
        size_dict[nxt[0]] = int(dims[rng.integers(0, len(dims))])
        nxt[0] += 1

    for i in range(1, n):
        add_edge(i, int(rng.integers(0, i)))
    for _ in range(extra):
        a = int(rng.integers(0, n))
        b = int(rng.integers(0, n))
        while b == a:
            b = int(rng.integers(0, n))
        add_edge(a, b)
    output = []
    for _ in range(n_open):
        t = int(rng.integers(0, n))
        inputs[t].append(nxt[0])
        output.append(nxt[0])
        size_dict[nxt[0]] = int(dims[rng.integers(0, len(dims))])
        nxt[0] += 1
    return {"inputs": inputs, "output": output, "size_dict": size_dict}


def random_arrays(net, dtype, seed):
    """Uniform [-1, 1) entries (real and imaginary parts), like src/tensor.rs:117-119, 233-238."""
    rng = np.random.default_rng(seed)
    out = []
    for legs in net["inputs"]:
        shape = [net["size_dict"][l] for l in legs]
        a = rng.uniform(-1.0, 1.0, size=shape)
        if np.issubdtype(np.dtype(dtype), np.complexfloating):
            a = a + 1j * rng.uniform(-1.0, 1.0, size=shape)
        out.append(np.array(a, dtype=dtype, order="C"))
    return out


def _rows_function():
    import torch

    class Rows(torch.autograd.Function):
        """Rows ``base[rows[i]]`` of a stacked tensor as separate tensors (one
        gather); the backward adds all their gradients into one tensor of the
        base's shape with one stack and one index_add."""

        @staticmethod
        def forward(ctx, base, rows):
            ctx.base_shape = base.shape
            ctx.save_for_backward(rows)
            return tuple(base.index_select(0, rows).unbind(0))

        @staticmethod
        def backward(ctx, *grads):
            (rows,) = ctx.saved_tensors
            ref = next(g for g in grads if g is not None)
            g = torch.stack([x if x is not None else torch.zeros_like(ref) for x in grads])
            out = torch.zeros(ctx.base_shape, dtype=g.dtype, device=g.device)
            return out.index_add_(0, rows, g), None

    return Rows


_ROWS = []


def _rows(base, rows):
    """``Rows.apply`` (the Function class is made on first use: torch is
    imported lazily by this module)."""
    if not _ROWS:
        _ROWS.append(_rows_function())
    return _ROWS[0].apply(base, rows)


_BRICK_CONSTS = {}


def _brickwork_consts(dev, cdt):
    """ket |0>, CZ, Z and the identity as device tensors, built once per device and dtype."""
    import torch

    key = (str(dev), cdt)
    got = _BRICK_CONSTS.get(key)
    if got is None:
        got = _BRICK_CONSTS[key] = (
            torch.tensor([1.0, 0.0], dtype=cdt, device=dev),
            torch.diag(torch.tensor([1, 1, 1, -1], dtype=cdt, device=dev)).reshape(2, 2, 2, 2),
            torch.diag(torch.tensor([1.0, -1.0], dtype=cdt, device=dev)),
            torch.eye(2, dtype=cdt, device=dev))
    return got


def _brickwork_gates(angles):
    """Every single-qubit gate U = RZ(phi) . RY(theta) of ``angles`` (layers,
    n_qubits, 2) and its conjugate, as one [2 * layers * n_qubits, 2, 2] table
    (ket gates first, then bra gates), by a handful of batched torch operations."""
    import torch

    cdt = torch.complex128 if angles.dtype == torch.float64 else torch.complex64
    th, ph = angles[..., 0], angles[..., 1]
    c, s = torch.cos(0.5 * th).to(cdt), torch.sin(0.5 * th).to(cdt)
    e = torch.polar(torch.ones_like(ph), -0.5 * ph)  # exp(-i phi / 2)
    ec = e.conj()
    # RZ . RY = [[e c, -e s], [conj(e) s, conj(e) c]]
    u = torch.stack([torch.stack([e * c, -e * s], -1), torch.stack([ec * s, ec * c], -1)], -2)
    return torch.stack([u, u.conj()]).reshape(-1, 2, 2), cdt


def brickwork_table(slots, angles):
    """The doubled brickwork network's inputs for ``autograd.contract_table``:
    (gate table, index, fixed tensors). Input i is row index[i] of the table (the
    gate of its slot: ket or bra half, layer, qubit) or fixed tensor -index[i] - 1
    (|0>, CZ, Z, identity); gradients reach ``angles`` through the table."""
    table, cdt = _brickwork_gates(angles)
    layers, nq = angles.shape[0], angles.shape[1]
    fixed_of = {"ket0": -1, "cz": -2, "z": -3}
    index = [(sl[3] * layers + sl[1]) * nq + sl[2] if sl[0] == "u" else fixed_of.get(sl[0], -4) for sl in slots]
    return table, index, _brickwork_consts(angles.device, cdt)


def brickwork_torch_tensors(slots, angles):
    """Input tensors of the doubled brickwork network as torch tensors built
    from ``angles`` (shape (layers, n_qubits, 2): RY angle, RZ angle), so that
    gradients flow back to the angles through torch's autograd.

    Every single-qubit gate U = RZ(phi) . RY(theta) is built at once for all
    layers and qubits by a handful of batched torch operations, and the 2x2
    inputs are handed out by one ``Rows`` node (ket and bra halves alike), so
    the construction and its backward cost a few launches, not a few per gate
    (per-gate indexing made autograd build and add one zero-filled copy of the
    whole batch per gate in the backward pass)."""
    import torch

    dev = angles.device
    table, cdt = _brickwork_gates(angles)
    layers, nq = angles.shape[0], angles.shape[1]
    rows = [(sl[3] * layers + sl[1]) * nq + sl[2] for sl in slots if sl[0] == "u"]
    gates = iter(_rows(table, torch.tensor(rows, device=dev)))
    ket0, cz, z, eye = _brickwork_consts(dev, cdt)
    out = []
    for sl in slots:
        if sl[0] == "ket0":
            out.append(ket0)
        elif sl[0] == "u":
            out.append(next(gates))
        elif sl[0] == "cz":
            out.append(cz)
        elif sl[0] == "z":
            out.append(z)
        else:
            out.append(eye)
    return out


def uneven_networks(n_inst, seed=0, n_lo=40, n_hi=120, dims=(2, 3, 4), dtype=np.float32):
    """Independent random 3-regular networks of uneven sizes, each contracted to a scalar.

    Instance i: vertex count uniform over the even integers in [n_lo, n_hi]
    (3-regular needs an even count), multigraph without self-loops from
    ``random_regular`` with graph seed [seed, i], per-edge dims uniform over
    ``dims``, N(0, 1) entries. Returns a list of (net, arrays).
    """
    rng = np.random.default_rng([seed, 0x5EED])
    out = []
    for i in range(n_inst):
        n = 2 * int(rng.integers((n_lo + 1) // 2, n_hi // 2 + 1))
        out.append(random_regular(n, 3, list(dims), [seed, i], dtype))
    return out


def uneven_instances(n_inst, seed=0, n_lo=40, n_hi=120, dims=(2, 3, 4), ntrials=8, dtype=np.float32):
    """``uneven_networks`` plus, per instance i, the best of ``ntrials``
    random-greedy paths with seed = i (greedy.rs:375-412). Returns a list of
    (net, arrays, ssa_path)."""
    from . import planner as P

    out = []
    for i, (net, arrays) in enumerate(uneven_networks(n_inst, seed, n_lo, n_hi, dims, dtype)):
        path, _ = P.random_greedy(net, ntrials, i)
        out.append((net, arrays, path))
    return out
