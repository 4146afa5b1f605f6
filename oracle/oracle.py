"""ORACLE — test infrastructure only.

ctypes binding of oracle/librtg_oracle.so, the CPU (C++17/OpenMP) restatement of the
reference's triangle-counting, set-intersection and sorted-lookup path (rtg_oracle.cpp cites
the reference file:line each function follows). Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it, and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "librtg_oracle.so"


def build() -> Path:
    subprocess.run(["make", "-B", "-s", "-C", str(HERE)], check=True)
    return LIB


def _load():
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    lib.orc_relabel.restype = C.c_int32
    lib.orc_tc_preprocess.restype = C.c_int64
    lib.orc_symmetric_csr.restype = C.c_int64
    for f in ("orc_tc_count_1a2", "orc_tc_count_intersect", "orc_tc_count_2a1"):
        getattr(lib, f).restype = C.c_uint64
    return lib


_lib = _load()


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def max_threads() -> int:
    return _lib.orc_max_threads()


def relabel(pairs: np.ndarray) -> tuple[np.ndarray, int]:
    """Graph::ReadBinaryEdges relabel (tc/graph/graph.cpp:40-62). Returns (pairs, nodes)."""
    p = np.array(pairs, dtype=np.int32, copy=True).reshape(-1)
    nodes = _lib.orc_relabel(_p(p), C.c_int64(p.size // 2))
    return p, int(nodes)


def tc_preprocess(pairs: np.ndarray, nodes: int) -> tuple[np.ndarray, np.ndarray]:
    """TCPreprocessing + ConvertToCSR (graph.cpp:104-177, 71-102) on relabelled pairs."""
    p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
    off = np.zeros(nodes + 1, dtype=np.int32)
    adj = np.zeros(max(p.size // 2, 1), dtype=np.int32)
    n = C.c_int32()
    m = _lib.orc_tc_preprocess(_p(p), C.c_int64(p.size // 2), C.c_int32(nodes), _p(off), _p(adj), C.byref(n))
    return off[: n.value + 1].copy(), adj[:m].copy()


def csr_from_pairs(pairs: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """The full reference pipeline: relabel, then TCPreprocessing + ConvertToCSR."""
    p, nodes = relabel(pairs)
    return tc_preprocess(p, nodes)


def symmetric_csr(pairs: np.ndarray, nodes: int) -> tuple[np.ndarray, np.ndarray]:
    p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
    off = np.zeros(nodes + 1, dtype=np.int32)
    adj = np.zeros(max(p.size, 1), dtype=np.int32)
    m = _lib.orc_symmetric_csr(_p(p), C.c_int64(p.size // 2), C.c_int32(nodes), _p(off), _p(adj))
    return off, adj[:m].copy()


def _csr(off, adj):
    return np.ascontiguousarray(off, dtype=np.int32), np.ascontiguousarray(adj, dtype=np.int32)


def tc_count_1a2(off, adj, e_begin: int = 0, e_end: int = 0, threads: int = 0) -> int:
    off, adj = _csr(off, adj)
    return int(_lib.orc_tc_count_1a2(_p(off), _p(adj), C.c_int32(off.size - 1), C.c_int64(e_begin),
                                     C.c_int64(e_end), C.c_int(threads)))


def tc_count_intersect(off, adj, e_begin: int = 0, e_end: int = 0, threads: int = 0) -> int:
    off, adj = _csr(off, adj)
    return int(_lib.orc_tc_count_intersect(_p(off), _p(adj), C.c_int32(off.size - 1), C.c_int64(e_begin),
                                           C.c_int64(e_end), C.c_int(threads)))


def tc_count_2a1(off, adj, threads: int = 0) -> int:
    off, adj = _csr(off, adj)
    return int(_lib.orc_tc_count_2a1(_p(off), _p(adj), C.c_int32(off.size - 1), C.c_int(threads)))


def tc_per_edge(off, adj, form: int = 0) -> np.ndarray:
    off, adj = _csr(off, adj)
    out = np.zeros(adj.size, dtype=np.uint32)
    _lib.orc_tc_per_edge(_p(off), _p(adj), C.c_int32(off.size - 1), C.c_int(form), _p(out))
    return out


def tc_edge_model(off, adj) -> dict:
    off, adj = _csr(off, adj)
    out = np.zeros(4, dtype=np.uint64)
    _lib.orc_tc_edge_model(_p(off), _p(adj), C.c_int32(off.size - 1), _p(out))
    return {"wedges_v": int(out[0]), "wedges_u": int(out[1]), "alg_bytes": int(out[2]), "merge_bytes": int(out[3])}


def si_pairs(offsets, elements, pairs, threads: int = 0) -> np.ndarray:
    offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
    elements = np.ascontiguousarray(elements, dtype=np.uint32)
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1)
    out = np.zeros(pairs.size // 2, dtype=np.uint32)
    _lib.orc_si_pairs(_p(offsets), _p(elements) if elements.size else None, _p(pairs), C.c_int64(pairs.size // 2),
                      _p(out), C.c_int(threads))
    return out


def si_tile(set_offsets, elements, num_a: int, num_b: int, threads: int = 0) -> np.ndarray:
    set_offsets = np.ascontiguousarray(set_offsets, dtype=np.uint32)
    elements = np.ascontiguousarray(elements, dtype=np.uint32)
    out = np.zeros(num_a * num_b, dtype=np.uint32)
    _lib.orc_si_tile(_p(set_offsets), _p(elements) if elements.size else None, C.c_uint32(num_a), C.c_uint32(num_b),
                     _p(out), C.c_int(threads))
    return out


def lookup(data, keys, signed: bool = False) -> np.ndarray:
    dt = np.int32 if signed else np.uint32
    data = np.ascontiguousarray(data, dtype=dt)
    keys = np.ascontiguousarray(keys, dtype=dt)
    out = np.zeros(keys.size, dtype=np.int32)
    fn = _lib.orc_lookup_i32 if signed else _lib.orc_lookup_u32
    fn(_p(data) if data.size else None, C.c_int64(data.size), _p(keys), C.c_int64(keys.size), _p(out))
    return out
