"""Network and tensor files of the reference's CLIs.

`load_network_json` reads a `.net.json` network (src/network.rs:245-392: the
canonical ``inputs/output/size_dict`` schema or the OMECO ``ixs/iy/sizes`` one,
labels interned as consecutive leg ids in first-occurrence order over inputs then
output, integer and string labels kept apart); `load_tensors_bin` reads the
little-endian row-major input data (src/network.rs:75-132: tensors in
``net.inputs`` order, complex values real part first, the file length exactly
Σ numel × itemsize) and `save_tensors_bin` writes it.
"""

from __future__ import annotations

import json

import numpy as np

from . import planner as _planner

# CLI dtype names (src/bin/tnexec.rs:729-740: --single / --complex)
DTYPES = {"f32": np.float32, "f64": np.float64, "c32": np.complex64, "c64": np.complex128}


def parse_network(obj):
    """(net, labels) from a parsed `.net.json` object; net has ``name``, ``inputs``,
    ``output`` and ``size_dict`` with dense integer legs, labels[leg] its original label."""
    if not isinstance(obj, dict):
        raise ValueError("JSON 根节点必须是对象")
    canon, omeco = ("inputs", "output", "size_dict"), ("ixs", "iy", "sizes")
    nc, no = sum(k in obj for k in canon), sum(k in obj for k in omeco)
    if nc and no:
        raise ValueError("JSON 不能混用 canonical inputs/output/size_dict 与 OMECO ixs/iy/sizes schema")
    if nc == 3:
        ik, ok, sk = canon
    elif no == 3:
        ik, ok, sk = omeco
    elif nc:
        raise ValueError("canonical schema 残缺：缺少 " + ", ".join(k for k in canon if k not in obj))
    elif no:
        raise ValueError("OMECO schema 残缺：缺少 " + ", ".join(k for k in omeco if k not in obj))
    else:
        raise ValueError("缺少完整的网络 schema：需要 inputs/output/size_dict 或 ixs/iy/sizes")
    name = obj.get("name") if isinstance(obj.get("name"), str) else "unnamed"
    intern, is_int = {}, {}

    def leg_of(x):
        if isinstance(x, int) and not isinstance(x, bool) and x >= 0:
            key, integer = str(x), True
        elif isinstance(x, str):
            key, integer = x, False
        else:
            raise ValueError(f"无法解析腿标签: {json.dumps(x)}")
        if key in is_int and is_int[key] != integer:
            raise ValueError(f"不同 JSON 腿标签在正规化后冲突: 整数/字符串 {key!r}")
        is_int.setdefault(key, integer)
        return intern.setdefault(key, len(intern))

    if not isinstance(obj[ik], list):
        raise ValueError(f"{ik} 必须是数组")
    inputs = []
    for legs in obj[ik]:
        if not isinstance(legs, list):
            raise ValueError(f"{ik} 元素不是数组")
        inputs.append([leg_of(x) for x in legs])
    if not isinstance(obj[ok], list):
        raise ValueError(f"{ok} 必须是数组")
    output = [leg_of(x) for x in obj[ok]]
    if not isinstance(obj[sk], dict):
        raise ValueError(f"{sk} 必须是对象")
    size_dict = {}
    for k, v in obj[sk].items():
        if k not in intern:
            raise ValueError(f"{sk} 中的腿 {k} 未在 {ik}/{ok} 出现")
        if not isinstance(v, int) or isinstance(v, bool) or v < 0:
            raise ValueError(f"维度不是整数: {json.dumps(v)}")
        size_dict[intern[k]] = v
    for key, leg in intern.items():
        if leg not in size_dict:
            raise ValueError(f"腿 {key} 缺少 size_dict 维度")
    labels = [None] * len(intern)
    for key, leg in intern.items():
        labels[leg] = key
    net = {"name": name, "inputs": inputs, "output": output, "size_dict": size_dict}
    _planner.validate_network(net)
    return net, labels


def load_network_json(path):
    """(net, labels) of a `.net.json` file."""
    with open(path, "r", encoding="utf-8") as f:
        return parse_network(json.load(f))


def _shapes(net):
    return [tuple(net["size_dict"][l] for l in legs) for legs in net["inputs"]]


def load_tensors_bin(path, net, dtype):
    """Input tensors of `net` from a little-endian binary file (row-major, complex
    values as (re, im) pairs), in ``net.inputs`` order."""
    dt = np.dtype(dtype).newbyteorder("<")
    shapes = _shapes(net)
    numels = [int(np.prod(s, dtype=np.int64)) for s in shapes]
    data = np.fromfile(path, dtype=np.uint8)
    expect = sum(numels) * dt.itemsize
    if data.size != expect:
        raise ValueError(f"数据文件字节数 {data.size} 与网络期望 {expect} 不符（应为 Σ numel × {dt.itemsize} 字节的小端；"
                         f"检查 dtype 与文件是否匹配该网络）")
    out, off = [], 0
    for shape, n in zip(shapes, numels):
        a = np.frombuffer(data[off:off + n * dt.itemsize].tobytes(), dtype=dt).reshape(shape)
        out.append(a.astype(np.dtype(dtype), copy=True))
        off += n * dt.itemsize
    return out


def save_tensors_bin(path, arrays):
    """Write arrays (one dtype) as little-endian row-major data in order."""
    with open(path, "wb") as f:
        for a in arrays:
            a = np.ascontiguousarray(a)
            f.write(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes())
