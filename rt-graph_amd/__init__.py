"""rt-graph_amd: Python binding of librtg, the MI355X (gfx950) triangle-counting, sorted-set
intersection and sorted-lookup engine.

The binding is ctypes over the C ABI declared in include/rtg.h. It mirrors the reference's
host interface for this path:

* ``Graph`` — ``ReadBinaryEdges`` / ``TCPreprocessing`` / ``ConvertToCSR`` with ``adjs_``,
  ``offsets_`` and ``graph_info_`` (tc/graph/include/graph.h:10-30), the preprocessing run
  on the device;
* ``CountOnGPU(adjs, offsets, nodes, edges, device_id=0)`` (tc/bs_tc/bs_tc.h:5), which prints
  the reference's report line and also returns the exact 64-bit count;
* ``BinarySearchBasedIntersection`` (set_intersection/baseline/bs.cu:21-22) and
  ``BSAccessCountFunction`` (rt_search/simple_bs.h:5).

Importing fails with ImportError when librtg.so has not been built: there is no CPU
fallback. GPU entry points raise RTGError when no device is present.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import sys
from pathlib import Path

import numpy as np

__all__ = [
    "RTGError", "Context", "DeviceArray", "DeviceGraph", "TcPlan", "LookupIndex", "Graph", "GraphInfo",
    "CountOnGPU", "BinarySearchBasedIntersection", "BSAccessCountFunction", "Dataset",
    "read_dataset", "write_dataset", "TC_PLAIN_THREAD", "TC_PLAIN_WAVE", "TC_BATCHED", "TC_2A1",
    "SI_PLAIN", "SI_BATCHED", "SI_REFERENCE", "SI_MERGE", "SI_HASH", "SI_BITMAP", "SI_AUTO",
    "SI_METHODS", "gen_si_dataset", "gen_ba", "gen_rmat", "gen_powerlaw_degrees",
    "gen_config_model", "gen_zipf_lengths", "gen_sorted_lists", "gen_sorted_array",
    "gen_uniform_u32", "read_binary_pairs", "read_text_pairs", "read_mtx_pairs", "lib_path",
    "pairs_fingerprint", "csr_file_write", "csr_file_read",
]

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "librtg.so"

TC_PLAIN_THREAD, TC_PLAIN_WAVE, TC_BATCHED, TC_2A1 = 0, 1, 2, 3
SI_PLAIN, SI_BATCHED, SI_REFERENCE, SI_MERGE, SI_HASH, SI_BITMAP, SI_AUTO = 0, 1, 2, 3, 4, 5, 6
SI_METHODS = (SI_PLAIN, SI_BATCHED, SI_REFERENCE, SI_MERGE, SI_HASH, SI_BITMAP, SI_AUTO)
SI_STRATEGIES = {-1: "none", 0: "search", 1: "merge", 2: "hash", 3: "bitmap", 4: "sampled"}
PAIRS_RELABEL, PAIRS_ORIENT, PAIRS_DEVICE = 0x1, 0x2, 0x4
TC_MAX_BUCKETS = 8

STATUS = {0: "RTG_OK", 1: "RTG_EINVAL", 2: "RTG_EHIP", 3: "RTG_ENOMEM", 4: "RTG_EIO",
          5: "RTG_EUNSUPPORTED", 6: "RTG_EAGAIN", 7: "RTG_ECANCELED"}


def lib_path() -> Path:
    return _LIB_PATH


if not _LIB_PATH.exists():
    raise ImportError(f"{_LIB_PATH} is not built; run rt-graph_amd/build.py (hipcc, gfx950)")

_lib = C.CDLL(str(_LIB_PATH))
_lib.rtg_last_error.restype = C.c_char_p


class RTGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def _check(rc: int) -> None:
    if rc != 0:
        raise RTGError(rc, (_lib.rtg_last_error() or b"").decode(errors="replace"))


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return C.c_void_p(a.ptr)
    if isinstance(a, int):
        return C.c_void_p(a)
    assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
    return C.c_void_p(a.ctypes.data)


# ---- structures (layout of include/rtg.h) ---------------------------------------------------
class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("cu_count", C.c_int32),
                ("xcd_count", C.c_int32), ("lds_per_block", C.c_int64), ("hbm_bytes", C.c_int64),
                ("clock_khz", C.c_int32), ("mem_clock_khz", C.c_int32)]


class BuildStats(C.Structure):
    _fields_ = [("t_total_ms", C.c_double), ("t_relabel_ms", C.c_double), ("t_orient_ms", C.c_double),
                ("t_sort_ms", C.c_double), ("t_csr_ms", C.c_double), ("raw_pairs", C.c_int64),
                ("valid_pairs", C.c_int64), ("nodes_in", C.c_int32), ("sort_passes", C.c_int32)]


class TcOpts(C.Structure):
    _fields_ = [("method", C.c_int32), ("use_hip_graph", C.c_int32), ("e_begin", C.c_int64),
                ("e_end", C.c_int64), ("tune_cache", C.c_char_p), ("autotune", C.c_int32),
                ("per_edge", C.c_int32)]


class TcStats(C.Structure):
    _fields_ = [("t_count_ms", C.c_double), ("t_plan_ms", C.c_double), ("alg_bytes", C.c_uint64),
                ("merge_bytes", C.c_uint64), ("edges", C.c_int64), ("nbuckets", C.c_int32),
                ("strategy", C.c_int32 * TC_MAX_BUCKETS), ("bucket_items", C.c_int64 * TC_MAX_BUCKETS),
                ("bucket_ms", C.c_double * TC_MAX_BUCKETS), ("kernel_launches", C.c_int32),
                ("reserved", C.c_int32), ("staged_bytes", C.c_uint64)]

    def as_dict(self) -> dict:
        return {
            "t_count_ms": self.t_count_ms, "t_plan_ms": self.t_plan_ms, "alg_bytes": self.alg_bytes,
            "merge_bytes": self.merge_bytes, "edges": self.edges, "nbuckets": self.nbuckets,
            "strategy": list(self.strategy[: self.nbuckets]),
            "bucket_items": list(self.bucket_items[: self.nbuckets]),
            "kernel_launches": self.kernel_launches, "staged_bytes": self.staged_bytes,
        }


class SiStats(C.Structure):
    _fields_ = [("t_kernel_ms", C.c_double), ("t_plan_ms", C.c_double), ("alg_bytes", C.c_uint64),
                ("nbuckets", C.c_int32), ("strategy", C.c_int32),
                ("bucket_items", C.c_int64 * TC_MAX_BUCKETS)]

    def as_dict(self) -> dict:
        return {"t_kernel_ms": self.t_kernel_ms, "t_plan_ms": self.t_plan_ms, "alg_bytes": self.alg_bytes,
                "strategy": SI_STRATEGIES.get(self.strategy, str(self.strategy)),
                "units": {SI_STRATEGIES[q]: int(self.bucket_items[q]) for q in range(self.nbuckets)}}


# ---- context and device memory --------------------------------------------------------------
class Context:
    """An rtg_ctx: one device, one stream (rtg_create / rtg_destroy)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(_lib.rtg_create(C.c_int(device), C.byref(self._h)))
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            _lib.rtg_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        d = DeviceInfo()
        _check(_lib.rtg_device_query(self._h, C.byref(d)))
        return {"name": d.name.decode(), "arch": d.arch.decode(), "cu_count": d.cu_count,
                "xcd_count": d.xcd_count, "lds_per_block": d.lds_per_block, "hbm_bytes": d.hbm_bytes,
                "clock_khz": d.clock_khz, "mem_clock_khz": d.mem_clock_khz}

    def synchronize(self) -> None:
        _check(_lib.rtg_synchronize(self._h))

    def empty(self, shape, dtype) -> "DeviceArray":
        return DeviceArray(self, shape, dtype)

    def to_device(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        if a.nbytes:
            _check(_lib.rtg_memcpy_h2d(self._h, C.c_void_p(d.ptr), _ptr(a), C.c_size_t(a.nbytes)))
        return d


class DeviceArray:
    """Device (HBM) buffer owned by Python; freed when collected."""

    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(np.array(shape, dtype=np.int64)).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(_lib.rtg_malloc(ctx.handle, C.c_size_t(max(self.nbytes, 1)), C.byref(p)))
        self.ptr = p.value

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            _check(_lib.rtg_memcpy_d2h(self.ctx.handle, _ptr(out), C.c_void_p(self.ptr), C.c_size_t(self.nbytes)))
        return out

    def copy_from(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            _check(_lib.rtg_memcpy_h2d(self.ctx.handle, C.c_void_p(self.ptr), _ptr(a), C.c_size_t(a.nbytes)))

    def free(self) -> None:
        if self.ptr:
            _lib.rtg_free(self.ctx.handle, C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            if self.ptr and self.ctx._h:
                self.free()
        except Exception:
            pass


# ---- graphs and triangle counting -------------------------------------------------------------
class DeviceGraph:
    """A device-resident CSR (rtg_graph)."""

    def __init__(self, ctx: Context, handle: C.c_void_p, build_stats: BuildStats | None = None):
        self.ctx = ctx
        self._h = handle
        self.build_stats = build_stats
        n, m, mr = C.c_int32(), C.c_int64(), C.c_int32()
        _check(_lib.rtg_graph_info(self._h, C.byref(n), C.byref(m), C.byref(mr)))
        self.n, self.m, self.max_row = n.value, m.value, mr.value

    @classmethod
    def from_pairs(cls, ctx: Context, pairs, relabel: bool = True, orient: bool = True) -> "DeviceGraph":
        flags = (PAIRS_RELABEL if relabel else 0) | (PAIRS_ORIENT if orient else 0)
        if isinstance(pairs, DeviceArray):
            flags |= PAIRS_DEVICE
            npairs = pairs.nbytes // 8
        else:
            pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
            npairs = pairs.size // 2
        h = C.c_void_p()
        st = BuildStats()
        _check(_lib.rtg_graph_from_pairs(ctx.handle, _ptr(pairs), C.c_int64(npairs), C.c_uint32(flags),
                                         C.byref(h), C.byref(st)))
        return cls(ctx, h, st)

    @classmethod
    def from_csr(cls, ctx: Context, offsets: np.ndarray, adjs: np.ndarray) -> "DeviceGraph":
        offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        adjs = np.ascontiguousarray(adjs, dtype=np.int32)
        h = C.c_void_p()
        _check(_lib.rtg_graph_from_csr(ctx.handle, _ptr(offsets), _ptr(adjs), C.c_int32(offsets.size - 1),
                                       C.c_int64(adjs.size), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def load(cls, ctx: Context, path, key: int = 0) -> "DeviceGraph":
        """rtg_graph_load: a CSR cached by save(); RTGError(RTG_EIO) if the key, checksum or size
        does not match."""
        h = C.c_void_p()
        _check(_lib.rtg_graph_load(ctx.handle, os.fsencode(path), C.c_uint64(key), C.byref(h)))
        return cls(ctx, h)

    def save(self, path, key: int = 0) -> None:
        """rtg_graph_save: the CSR to a cache file tagged with `key` (e.g. pairs_fingerprint)."""
        _check(_lib.rtg_graph_save(self._h, os.fsencode(path), C.c_uint64(key)))

    def download(self) -> tuple[np.ndarray, np.ndarray]:
        off = np.empty(self.n + 1, dtype=np.int32)
        adj = np.empty(self.m, dtype=np.int32)
        _check(_lib.rtg_graph_download(self._h, _ptr(off), _ptr(adj)))
        return off, adj

    def ordered_csr(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """rtg_graph_ordered_csr: the degree-ordered copy batched counting plans count on (offsets,
        adjs) and inv (the graph's id of each copy vertex); built on first use."""
        off = np.empty(self.n + 1, dtype=np.int32)
        adj = np.empty(self.m, dtype=np.int32)
        inv = np.empty(self.n, dtype=np.int32)
        _check(_lib.rtg_graph_ordered_csr(self.ctx.handle, self._h, _ptr(off), _ptr(adj), _ptr(inv)))
        return off, adj, inv

    def device_ptrs(self) -> tuple[int, int]:
        o, a = C.c_void_p(), C.c_void_p()
        _check(_lib.rtg_graph_device_ptrs(self._h, C.byref(o), C.byref(a)))
        return o.value, a.value

    def tc_count(self, method: int = TC_BATCHED, e_begin: int = 0, e_end: int = 0,
                 use_hip_graph: bool = False) -> tuple[int, dict]:
        opts = TcOpts(method=method, use_hip_graph=int(use_hip_graph), e_begin=e_begin, e_end=e_end)
        cnt = C.c_uint64()
        st = TcStats()
        _check(_lib.rtg_tc_count(self.ctx.handle, self._h, C.byref(opts), C.byref(cnt), C.byref(st)))
        return cnt.value, st.as_dict()

    def tc_plan(self, method: int = TC_BATCHED, e_begin: int = 0, e_end: int = 0,
                use_hip_graph: bool = True, autotune: bool = False, tune_cache: str | None = None,
                per_edge: bool = False) -> "TcPlan":
        return TcPlan(self, method, e_begin, e_end, use_hip_graph, autotune, tune_cache, per_edge)

    def tc_per_edge(self, form: int = 0) -> np.ndarray:
        """rtg_tc_per_edge: form 0 |N+(u) & N+(v)| through a batched per-edge plan, form 2 the
        same by one wave per edge, form 1 the 1A2 attribution (edge a->c)."""
        out = DeviceArray(self.ctx, (self.m,), np.uint32)
        _check(_lib.rtg_tc_per_edge(self.ctx.handle, self._h, C.c_int32(form), C.c_void_p(out.ptr)))
        return out.numpy()

    def close(self) -> None:
        if self._h:
            _lib.rtg_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TcPlan:
    """rtg_tc_plan: buckets, work queue and captured launch sequence, reusable across runs."""

    def __init__(self, g: DeviceGraph, method: int, e_begin: int, e_end: int, use_hip_graph: bool,
                 autotune: bool = False, tune_cache: str | None = None, per_edge: bool = False):
        self.g = g
        self._h = C.c_void_p()
        self._cache = None if tune_cache is None else os.fsencode(tune_cache)
        opts = TcOpts(method=method, use_hip_graph=int(use_hip_graph), e_begin=e_begin, e_end=e_end,
                      tune_cache=self._cache, autotune=int(autotune), per_edge=int(per_edge))
        st = TcStats()
        _check(_lib.rtg_tc_plan_create(g.ctx.handle, g._h, C.byref(opts), C.byref(self._h), C.byref(st)))
        self.plan_stats = st.as_dict()

    def describe(self) -> dict:
        buf = C.create_string_buffer(1024)
        _check(_lib.rtg_tc_plan_describe(self._h, buf, C.c_size_t(1024)))
        out = {}
        for kv in buf.value.decode().split(","):
            k, _, v = kv.partition("=")
            out[k] = v
        return out

    def run(self) -> tuple[int, dict]:
        cnt = C.c_uint64()
        st = TcStats()
        _check(_lib.rtg_tc_plan_run(self.g.ctx.handle, self._h, C.byref(cnt), C.byref(st)))
        return cnt.value, st.as_dict()

    def run_per_edge(self, d_out: DeviceArray | None = None) -> tuple[DeviceArray, int, dict]:
        """rtg_tc_plan_run_per_edge (plan created with per_edge=True): d_out[e] = |N+(u) & N+(v)|
        for the plan's edge range; returns the device array, the range's total and the stats."""
        if d_out is None:
            d_out = DeviceArray(self.g.ctx, (max(self.g.m, 1),), np.uint32)
        cnt = C.c_uint64()
        st = TcStats()
        _check(_lib.rtg_tc_plan_run_per_edge(self.g.ctx.handle, self._h, C.c_void_p(d_out.ptr), C.byref(cnt),
                                             C.byref(st)))
        return d_out, cnt.value, st.as_dict()

    def close(self) -> None:
        if self._h:
            _lib.rtg_tc_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- reference-shaped host interface ----------------------------------------------------------
class GraphInfo:
    """graph_info_t (tc/utils/model.h:10-13)."""

    def __init__(self, nodes: int = 0, edges: int = 0):
        self.nodes, self.edges = nodes, edges

    def __repr__(self):
        return f"GraphInfo(nodes={self.nodes}, edges={self.edges})"


def read_binary_pairs(path: str | os.PathLike) -> np.ndarray:
    """Little-endian int32 (src, dst) pairs, npairs = file_size / 8 (graph.cpp:23-39)."""
    size = os.path.getsize(path)
    return np.fromfile(path, dtype="<i4", count=(size // 8) * 2).astype(np.int32, copy=False)


def _read_pairs(fn, path) -> np.ndarray:
    p = str(path).encode()
    n = C.c_int64()
    _check(fn(p, None, C.byref(n)))
    out = np.empty(2 * n.value, dtype=np.int32)
    _check(fn(p, _ptr(out) if out.size else None, C.byref(n)))
    return out[: 2 * n.value]


def pairs_fingerprint(pairs: np.ndarray, flags: int = PAIRS_RELABEL | PAIRS_ORIENT) -> int:
    """rtg_pairs_fingerprint: a 64-bit key for a pair list and its preprocessing flags."""
    pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
    out = C.c_uint64()
    _check(_lib.rtg_pairs_fingerprint(_ptr(pairs) if pairs.size else None, C.c_int64(pairs.size // 2),
                                      C.c_int32(flags), C.byref(out)))
    return out.value


def csr_file_write(path, offsets: np.ndarray, adjs: np.ndarray, key: int = 0, oriented: bool = True) -> None:
    """rtg_csr_file_write (host only): a CSR to a cache file."""
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    adjs = np.ascontiguousarray(adjs, dtype=np.int32)
    _check(_lib.rtg_csr_file_write(os.fsencode(path), C.c_uint64(key), C.c_int32(int(oriented)), _ptr(offsets),
                                   _ptr(adjs) if adjs.size else None, C.c_int32(offsets.size - 1),
                                   C.c_int64(adjs.size)))


def csr_file_read(path, key: int = 0) -> tuple[np.ndarray, np.ndarray, bool]:
    """rtg_csr_file_read (host only): (offsets, adjs, oriented) of a cache file."""
    p = os.fsencode(path)
    n, m, o = C.c_int32(), C.c_int64(), C.c_int32()
    _check(_lib.rtg_csr_file_read(p, C.c_uint64(key), None, None, C.byref(n), C.byref(m), C.byref(o)))
    off = np.empty(n.value + 1, dtype=np.int32)
    adj = np.empty(max(m.value, 1), dtype=np.int32)
    _check(_lib.rtg_csr_file_read(p, C.c_uint64(key), _ptr(off), _ptr(adj), C.byref(n), C.byref(m), C.byref(o)))
    return off, adj[: m.value], bool(o.value)


def read_text_pairs(path: str | os.PathLike) -> np.ndarray:
    """rtg_read_text_pairs: "src dst" per line; lines that do not start with two int32 values are
    skipped with a message (Graph::Txt2Bin, tc/graph/graph.cpp:179-214)."""
    return _read_pairs(_lib.rtg_read_text_pairs, path)


def read_mtx_pairs(path: str | os.PathLike) -> np.ndarray:
    """rtg_read_mtx_pairs: Matrix Market coordinate entries as 0-based pairs, symmetric files
    expanded (bfs/graph/graph.cpp:17-122)."""
    return _read_pairs(_lib.rtg_read_mtx_pairs, path)


class Graph:
    """Mirror of the reference's Graph (tc/graph/include/graph.h:10-30).

    ReadBinaryEdges reads the pair file; TCPreprocessing runs the relabel, raw-degree
    orientation, degree<=1 removal, sort and dedup on the device; ConvertToCSR copies the
    CSR into ``adjs_`` / ``offsets_`` (numpy int32) and sets ``graph_info_``.
    """

    def __init__(self, device_id: int = 0):
        self.device_id = device_id
        self.adjs_ = np.empty(0, dtype=np.int32)
        self.offsets_ = np.zeros(1, dtype=np.int32)
        self.graph_info_ = GraphInfo()
        self._pairs = None
        self._ctx = None
        self._dg = None

    def _context(self) -> Context:
        if self._ctx is None:
            self._ctx = Context(self.device_id)
        return self._ctx

    def ReadBinaryEdges(self, input_file_path) -> None:
        self._pairs = read_binary_pairs(input_file_path)
        print(f"The size of graph file: {os.path.getsize(input_file_path)} bytes")
        self.graph_info_.edges = self._pairs.size // 2

    def SetEdges(self, pairs: np.ndarray) -> None:
        """Pairs from memory instead of a file (same semantics as ReadBinaryEdges)."""
        self._pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
        self.graph_info_.edges = self._pairs.size // 2

    def TCPreprocessing(self) -> None:
        if self._pairs is None:
            raise RTGError(1, "TCPreprocessing: no edges read")
        self._dg = DeviceGraph.from_pairs(self._context(), self._pairs, relabel=True, orient=True)
        st = self._dg.build_stats
        print(f"Graph Info: nodes = {st.nodes_in}, edges = {self._pairs.size // 2}")
        print(f"TC Preprocessing Time = {st.t_total_ms:.6f} ms")

    def ConvertToCSR(self) -> None:
        if self._dg is None:
            raise RTGError(1, "ConvertToCSR: TCPreprocessing has not run")
        self.offsets_, self.adjs_ = self._dg.download()
        self.graph_info_ = GraphInfo(self._dg.n, self._dg.m)

    @property
    def device_graph(self) -> DeviceGraph | None:
        return self._dg

    @staticmethod
    def Txt2Bin(txt_file, bin_file) -> None:
        """Text edge list ("a b" per line) to binary int32 pairs (graph.cpp:179-214)."""
        pairs = read_text_pairs(txt_file)
        _check(_lib.rtg_write_binary_pairs(str(bin_file).encode(), _ptr(pairs) if pairs.size else None,
                                           C.c_int64(pairs.size // 2)))
        print("Edge list has been successfully converted to binary.")


def CountOnGPU(adjs, offsets, nodes: int, edges: int, device_id: int = 0,
               method: int = TC_BATCHED, quiet: bool = False) -> int:
    """tc/bs_tc/bs_tc.cu:232-354: copies the host CSR to the device, counts, prints
    ``BSTC: counting time = %f ms, counting result = %u`` and returns the exact count.
    The reference's counter is a 32-bit uint; the printed %u value is the low 32 bits."""
    offsets = np.ascontiguousarray(offsets[: nodes + 1], dtype=np.int32)
    adjs = np.ascontiguousarray(adjs[:edges], dtype=np.int32)
    with Context(device_id) as ctx:
        g = DeviceGraph.from_csr(ctx, offsets, adjs)
        cnt, st = g.tc_count(method=method)
        g.close()
    if not quiet:
        print(f"BSTC: counting time = {st['t_count_ms']:f} ms, counting result = {cnt & 0xFFFFFFFF}")
        if cnt >> 32:
            print(f"BSTC: exact 64-bit count = {cnt}")
    return cnt


# ---- set intersection ---------------------------------------------------------------------------
class Dataset:
    """set_intersection/utils/util.h:19-28: header (num_a, num_b, max_element, intersection),
    set_offsets[num_a+num_b+1] and elements (A sets first)."""

    def __init__(self, num_a: int, num_b: int, max_element: int, intersection: int,
                 set_offsets: np.ndarray, elements: np.ndarray):
        self.num_of_sets_a, self.num_of_sets_b = int(num_a), int(num_b)
        self.num_of_sets = self.num_of_sets_a + self.num_of_sets_b
        self.max_element, self.intersection = int(max_element), int(intersection)
        self.set_offsets = np.ascontiguousarray(set_offsets, dtype=np.uint32)
        self.elements = np.ascontiguousarray(elements, dtype=np.uint32)


def read_dataset(path) -> Dataset:
    """ReadData, set_intersection/utils/util.cpp:3-23."""
    with open(path, "rb") as f:
        num_a, num_b, max_el, inter = struct.unpack("<4I", f.read(16))
        offs = np.frombuffer(f.read(4 * (num_a + num_b + 1)), dtype="<u4").astype(np.uint32)
        elems = np.frombuffer(f.read(4 * int(offs[-1])), dtype="<u4").astype(np.uint32)
    return Dataset(num_a, num_b, max_el, inter, offs, elems)


def write_dataset(path, ds: Dataset) -> None:
    """Writer format of set_intersection/data_generator/data.cpp:344-373."""
    with open(path, "wb") as f:
        f.write(struct.pack("<4I", ds.num_of_sets_a, ds.num_of_sets_b, ds.max_element, ds.intersection))
        f.write(ds.set_offsets.astype("<u4").tobytes())
        f.write(ds.elements.astype("<u4").tobytes())


def si_pairs(ctx: Context, d_offsets: DeviceArray, d_elements: DeviceArray, nlists: int,
             d_pairs: DeviceArray, npairs: int, method: int = SI_BATCHED, d_out: DeviceArray | None = None):
    """rtg_si_pairs: d_out[k] = |list(pairs[2k]) & list(pairs[2k+1])| (device arrays)."""
    if d_out is None:
        d_out = DeviceArray(ctx, (npairs,), np.uint32)
    st = SiStats()
    _check(_lib.rtg_si_pairs(ctx.handle, _ptr(d_offsets), _ptr(d_elements), C.c_int64(nlists), _ptr(d_pairs),
                             C.c_int64(npairs), C.c_int32(method), _ptr(d_out), C.byref(st)))
    return d_out, st.as_dict()


SI_NOALLOC_MAX_PAIRS = 1 << 21  # RTG_SI_NOALLOC_MAX_PAIRS


def si_pairs_noalloc(ctx: Context, d_offsets: DeviceArray, d_elements: DeviceArray, nlists: int,
                     d_pairs: DeviceArray, npairs: int, method: int, d_out: DeviceArray) -> float:
    """rtg_si_pairs_noalloc: rtg_si_pairs without any device allocation (scratch in the
    library's fixed per-GPU array); d_out is the caller's. Returns the kernel time in ms."""
    ms = C.c_double()
    _check(_lib.rtg_si_pairs_noalloc(ctx.handle, _ptr(d_offsets), _ptr(d_elements), C.c_int64(nlists),
                                     _ptr(d_pairs), C.c_int64(npairs), C.c_int32(method), _ptr(d_out),
                                     C.byref(ms)))
    return ms.value


def si_tile(ctx: Context, d_offsets: DeviceArray, d_elements: DeviceArray, num_a: int, num_b: int,
            method: int = SI_BATCHED, d_out: DeviceArray | None = None):
    """rtg_si_tile: counts[a * num_b + b] = |A_a & B_b| for sets [0, num_a) and [num_a, num_a + num_b)."""
    if d_out is None:
        d_out = DeviceArray(ctx, (num_a * num_b,), np.uint32)
    st = SiStats()
    _check(_lib.rtg_si_tile(ctx.handle, _ptr(d_offsets), _ptr(d_elements), C.c_uint32(num_a), C.c_uint32(num_b),
                            C.c_int32(method), _ptr(d_out), C.byref(st)))
    return d_out, st.as_dict()


def BinarySearchBasedIntersection(dataset: Dataset, blocks: int = 0, block_size: int = 512,
                                  device_id: int = 0, method: int = SI_REFERENCE) -> np.ndarray:
    """set_intersection/baseline/bs.cu:91-135: counts[a * num_b + b] = |A_a & B_b|.
    ``blocks``/``block_size`` are accepted for signature parity; the launch shape is chosen
    by the library."""
    with Context(device_id) as ctx:
        d_off = ctx.to_device(dataset.set_offsets)
        d_el = ctx.to_device(dataset.elements if dataset.elements.size else np.zeros(1, np.uint32))
        out, st = si_tile(ctx, d_off, d_el, dataset.num_of_sets_a, dataset.num_of_sets_b, method)
        res = out.numpy()
        print(f"Binary Search based Set Intertsection: kernel time = {st['t_kernel_ms']} ms")
        for a in (out, d_el, d_off):
            a.free()
    return res


# ---- sorted lookup ------------------------------------------------------------------------------
def lookup(ctx: Context, d_data: DeviceArray, n: int, d_keys: DeviceArray, q: int,
           d_out: DeviceArray | None = None, signed: bool = False):
    if d_out is None:
        d_out = DeviceArray(ctx, (q,), np.int32)
    ms = C.c_double()
    fn = _lib.rtg_lookup_i32 if signed else _lib.rtg_lookup_u32
    _check(fn(ctx.handle, _ptr(d_data), C.c_int64(n), _ptr(d_keys), C.c_int64(q), _ptr(d_out), C.byref(ms)))
    return d_out, ms.value


class LookupIndex:
    """Search index built once over a device-resident ascending array (rtg_lookup_index);
    run() gives the same results as lookup(). The array must stay allocated and unchanged."""

    def __init__(self, ctx: Context, d_data: DeviceArray | int, n: int, signed: bool = False):
        self.ctx, self.n, self.signed = ctx, n, signed
        self._data = d_data  # keep the array alive while the index refers to it
        h = C.c_void_p()
        fn = _lib.rtg_lookup_index_create_i32 if signed else _lib.rtg_lookup_index_create_u32
        _check(fn(ctx.handle, _ptr(d_data) if n else None, C.c_int64(n), C.byref(h)))
        self._h = h

    @property
    def levels(self) -> int:
        lv = C.c_int32()
        _check(_lib.rtg_lookup_index_info(self._h, C.byref(lv), None))
        return lv.value

    @property
    def nbytes(self) -> int:
        b = C.c_int64()
        _check(_lib.rtg_lookup_index_info(self._h, None, C.byref(b)))
        return b.value

    def run(self, d_keys: DeviceArray, q: int, d_out: DeviceArray | None = None):
        if d_out is None:
            d_out = DeviceArray(self.ctx, (q,), np.int32)
        ms = C.c_double()
        _check(_lib.rtg_lookup_index_run(self.ctx.handle, self._h, _ptr(d_keys), C.c_int64(q), _ptr(d_out), C.byref(ms)))
        return d_out, ms.value

    def close(self) -> None:
        if self._h:
            _lib.rtg_lookup_index_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sort_device(ctx: Context, d_data: DeviceArray, n: int, signed: bool = False) -> float:
    """rtg_sort_u32 / rtg_sort_i32: ascending in-place device sort of n 32-bit keys; returns the
    device time in ms."""
    ms = C.c_double()
    fn = _lib.rtg_sort_i32 if signed else _lib.rtg_sort_u32
    _check(fn(ctx.handle, _ptr(d_data), C.c_int64(n), C.byref(ms)))
    return ms.value


def BSAccessCountFunction(data: np.ndarray, query: np.ndarray, device_id: int = 0) -> np.ndarray:
    """rt_search/simple_bs.cu:80-142: the unsorted ``data`` goes to the device, is sorted there
    (:99-108) and on the host in place (:92), and every query is looked up on the device.
    Returns, per query, the index of the first equal element of the sorted array, or -1, and
    prints the reference's lines."""
    import time
    d = np.ascontiguousarray(data, dtype=np.int32)
    qv = np.ascontiguousarray(query, dtype=np.int32)
    with Context(device_id) as ctx:
        dd, dq = ctx.to_device(d if d.size else np.zeros(1, np.int32)), ctx.to_device(qv if qv.size else np.zeros(1, np.int32))
        print(f"data size = {d.size}")
        t0 = time.process_time()
        data.sort()
        print(f"cpu sort time = {(time.process_time() - t0) * 1e3:.3f} ms")
        sort_ms = sort_device(ctx, dd, d.size, signed=True)
        out, ms = lookup(ctx, dd, d.size, dq, qv.size, signed=True)
        res = out.numpy()[: qv.size]
        print(f"Sorting time = {sort_ms:f} ms, Binary search time = {ms:f} ms, Total time = {sort_ms + ms:f} ms")
        for a in (out, dq, dd):
            a.free()
    return res


# ---- generators (host, seeded, counter-based) -----------------------------------------------------
def gen_ba(n: int, m: int, seed: int = 1) -> np.ndarray:
    out = np.empty(2 * m * max(n - m, 0), dtype=np.int32)
    k = C.c_int64()
    _check(_lib.rtg_gen_ba(C.c_int32(n), C.c_int32(m), C.c_uint64(seed), _ptr(out), C.byref(k)))
    return out[: 2 * k.value]


def gen_rmat(scale: int, nsamples: int, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1) -> np.ndarray:
    out = np.empty(2 * nsamples, dtype=np.int32)
    _check(_lib.rtg_gen_rmat(C.c_uint32(scale), C.c_int64(nsamples), C.c_double(a), C.c_double(b),
                             C.c_double(c), C.c_uint64(seed), _ptr(out)))
    return out


def gen_rmat_device(ctx: Context, scale: int, nsamples: int, a: float = 0.57, b: float = 0.19,
                    c: float = 0.19, seed: int = 1) -> DeviceArray:
    out = DeviceArray(ctx, (2 * nsamples,), np.int32)
    _check(_lib.rtg_gen_rmat_device(ctx.handle, C.c_uint32(scale), C.c_int64(nsamples), C.c_double(a),
                                    C.c_double(b), C.c_double(c), C.c_uint64(seed), C.c_void_p(out.ptr)))
    return out


def gen_sorted_array_device(ctx: Context, n: int, seed: int = 1) -> DeviceArray:
    out = DeviceArray(ctx, (n,), np.uint32)
    _check(_lib.rtg_gen_sorted_array_device(ctx.handle, C.c_int64(n), C.c_uint64(seed), C.c_void_p(out.ptr)))
    return out


def gen_config_model_device(ctx: Context, deg: np.ndarray, p_close: float, seed: int = 1) -> DeviceArray:
    """Device version of gen_config_model (same pairs, same order); returns a DeviceArray of
    2 * npairs int32 (over-allocated buffer trimmed by shape)."""
    deg = np.ascontiguousarray(deg, dtype=np.int32)
    total = int(deg.sum(dtype=np.int64))
    out = DeviceArray(ctx, (2 * (total // 2 + total),), np.int32)
    k = C.c_int64()
    _check(_lib.rtg_gen_config_model_device(ctx.handle, C.c_int32(deg.size), _ptr(deg), C.c_int64(total),
                                            C.c_double(p_close), C.c_uint64(seed), C.c_void_p(out.ptr), C.byref(k)))
    out.shape = (2 * k.value,)
    out.nbytes = 8 * k.value
    return out


def gen_powerlaw_degrees(n: int, gamma: float, kmax: float, mean: float, seed: int = 1) -> np.ndarray:
    deg = np.empty(n, dtype=np.int32)
    tot = C.c_int64()
    _check(_lib.rtg_gen_powerlaw_degrees(C.c_int32(n), C.c_double(gamma), C.c_double(kmax), C.c_double(mean),
                                         C.c_uint64(seed), _ptr(deg), C.byref(tot)))
    return deg


def gen_config_model(deg: np.ndarray, p_close: float, seed: int = 1) -> np.ndarray:
    deg = np.ascontiguousarray(deg, dtype=np.int32)
    total = int(deg.sum(dtype=np.int64))
    out = np.empty(2 * (total // 2 + total), dtype=np.int32)
    k = C.c_int64()
    _check(_lib.rtg_gen_config_model(C.c_int32(deg.size), _ptr(deg), C.c_int64(total), C.c_double(p_close),
                                     C.c_uint64(seed), _ptr(out), C.byref(k)))
    return out[: 2 * k.value]


def gen_zipf_lengths(nlists: int, s: float = 1.8, kmax: int = 65536, seed: int = 1) -> np.ndarray:
    out = np.empty(nlists, dtype=np.uint32)
    _check(_lib.rtg_gen_zipf_lengths(C.c_int64(nlists), C.c_double(s), C.c_uint32(kmax), C.c_uint64(seed),
                                     _ptr(out)))
    return out


def gen_sorted_lists(lengths: np.ndarray, value_range: int = 1 << 24, seed: int = 1):
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    offsets = np.zeros(lengths.size + 1, dtype=np.uint64)
    np.cumsum(lengths, out=offsets[1:])
    if offsets[-1] >= 1 << 32:
        raise RTGError(1, "gen_sorted_lists: more than 2^32 elements")
    offsets = offsets.astype(np.uint32)
    elems = np.empty(int(offsets[-1]), dtype=np.uint32)
    _check(_lib.rtg_gen_sorted_lists(C.c_int64(lengths.size), _ptr(offsets), C.c_uint32(value_range),
                                     C.c_uint64(seed), _ptr(elems)))
    return offsets, elems


def gen_si_dataset(num_a: int, num_b: int, length: int, skew_ratio: float = 1.0, selectivity: float = 0.1,
                   density: float = 0.1, seed: int = 1) -> Dataset:
    """The reference's uniform set-intersection dataset (set_intersection/data_generator/
    data.cpp:21-74, header fields as in its Writer, :344-373): every A set is the same sorted
    set of len_a = length distinct values below max_element = (len_a + len_b - len_c) / density
    (at most 2^31 - 1); every B set holds the len_c smallest values of that set plus
    len_b - len_c values outside it, so every A x B intersection has len_c elements, with
    len_b = length / skew_ratio and len_c = length * selectivity. Values come from numpy's
    PCG64 seeded with `seed` rather than from rand()."""
    len_a, len_b, len_c = int(length), int(length / skew_ratio), int(length * selectivity)
    if not (0 <= len_c <= min(len_a, len_b)):
        raise RTGError(1, "gen_si_dataset: selectivity leaves len_c outside [0, min(len_a, len_b)]")
    value_range = min(int((len_a + len_b - len_c) / density), (1 << 31) - 1)
    if value_range < len_a + len_b - len_c:
        raise RTGError(1, "gen_si_dataset: value range too small for the set lengths")
    rng = np.random.default_rng(seed)
    tmp = np.sort(rng.choice(value_range, size=len_a, replace=False)).astype(np.uint32)
    b_sets = []
    for _ in range(num_b):
        need, fresh = len_b - len_c, np.empty(0, np.uint32)
        while fresh.size < need:
            cand = rng.integers(0, value_range, size=2 * (need - fresh.size) + 16, dtype=np.uint32)
            cand = cand[~np.isin(cand, tmp)]
            fresh = np.unique(np.concatenate([fresh, cand]))
        fresh = rng.permutation(fresh)[:need]
        b_sets.append(np.sort(np.concatenate([tmp[:len_c], fresh])).astype(np.uint32))
    lengths = np.array([len_a] * num_a + [len_b] * num_b, dtype=np.uint64)
    offsets = np.zeros(lengths.size + 1, dtype=np.uint64)
    np.cumsum(lengths, out=offsets[1:])
    if offsets[-1] >= 1 << 32:
        raise RTGError(1, "gen_si_dataset: more than 2^32 elements")
    elements = np.concatenate([np.tile(tmp, num_a)] + b_sets).astype(np.uint32) if lengths.size else \
        np.zeros(0, np.uint32)
    return Dataset(num_a, num_b, value_range, len_c, offsets.astype(np.uint32), elements)


def gen_sorted_array(n: int, seed: int = 1) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    _check(_lib.rtg_gen_sorted_array(C.c_int64(n), C.c_uint64(seed), _ptr(out)))
    return out


def gen_uniform_u32(n: int, seed: int = 1, stream: int = 0x51554552) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    _check(_lib.rtg_gen_uniform_u32(C.c_int64(n), C.c_uint64(seed), C.c_uint64(stream), _ptr(out)))
    return out


# ---- query server ---------------------------------------------------------------------------------
REQ_PAIR, REQ_LOOKUP = 0, 1
REQ_PENDING, REQ_RUNNING, REQ_DONE, REQ_CANCELLED, REQ_TIMEOUT, REQ_FAILED = range(6)


class ServerOpts(C.Structure):
    _fields_ = [("slots", C.c_int32), ("max_batch_requests", C.c_int32), ("max_batch_pairs", C.c_int64),
                ("max_batch_keys", C.c_int64), ("mem_budget_bytes", C.c_int64), ("timeout_us", C.c_int64),
                ("poll_us", C.c_int32), ("null_executor", C.c_int32), ("null_latency_us", C.c_int32),
                ("copy_staging", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("admitted", C.c_int64), ("rejected", C.c_int64), ("completed", C.c_int64),
                ("cancelled", C.c_int64), ("failed", C.c_int64), ("batches", C.c_int64),
                ("max_batch_requests", C.c_int64), ("max_busy_slots", C.c_int64),
                ("mean_batch_requests", C.c_double), ("mean_us", C.c_double), ("p50_us", C.c_double),
                ("p99_us", C.c_double), ("p999_us", C.c_double), ("max_us", C.c_double),
                ("throughput_qps", C.c_double), ("offered_qps", C.c_double),
                ("mean_batch_device_us", C.c_double), ("mean_batch_pack_us", C.c_double),
                ("mean_batch_unpack_us", C.c_double), ("mean_batch_slot_us", C.c_double),
                ("max_batch_device_us", C.c_double), ("max_batch_slot_us", C.c_double),
                ("last_error", C.c_char * 128)]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "last_error"}
        d["last_error"] = self.last_error.decode(errors="replace")
        return d


class StreamSpec(C.Structure):
    _fields_ = [("rate_qps", C.c_double), ("nqueries", C.c_int64), ("warmup", C.c_int64),
                ("pair_frac", C.c_double), ("max_keys", C.c_int32), ("reserved", C.c_int32),
                ("seed", C.c_uint64)]


class Server:
    """rtg_server: continuous-batching query server over a symmetric device graph and a sorted
    device array. With null_executor=True it runs without a device (no results computed)."""

    def __init__(self, ctx: Context | None, graph: DeviceGraph | None, d_array: DeviceArray | None,
                 n_array: int = 0, **opts):
        self.ctx, self.graph, self.d_array = ctx, graph, d_array
        o = ServerOpts()
        for k, v in opts.items():
            setattr(o, k, int(v))
        self._h = C.c_void_p()
        _check(_lib.rtg_server_create(ctx.handle if ctx else None, graph._h if graph else None,
                                      C.c_void_p(d_array.ptr) if d_array else None, C.c_int64(n_array),
                                      C.byref(o), C.byref(self._h)))
        self._keep = {}

    def submit_pair(self, u: int, v: int) -> int:
        rid = C.c_uint64()
        _check(_lib.rtg_server_submit_pair(self._h, C.c_int32(u), C.c_int32(v), C.byref(rid)))
        return rid.value

    def submit_lookup(self, keys: np.ndarray) -> tuple[int, np.ndarray]:
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        out = np.full(keys.size, -3, dtype=np.int32)
        rid = C.c_uint64()
        _check(_lib.rtg_server_submit_lookup(self._h, _ptr(keys), C.c_int32(keys.size), _ptr(out), C.byref(rid)))
        self._keep[rid.value] = (keys, out)  # buffers must outlive the request
        return rid.value, out

    def poll(self, rid: int):
        st, pr, lat = C.c_int32(), C.c_uint32(), C.c_int64()
        _check(_lib.rtg_server_poll(self._h, C.c_uint64(rid), C.byref(st), C.byref(pr), C.byref(lat)))
        if st.value not in (REQ_PENDING, REQ_RUNNING):
            self._keep.pop(rid, None)
        return st.value, pr.value, lat.value

    def wait(self, rid: int, timeout_s: float = 30.0):
        import time
        t0 = time.perf_counter()
        while True:
            st, pr, lat = self.poll(rid)
            if st not in (REQ_PENDING, REQ_RUNNING):
                return st, pr, lat
            if time.perf_counter() - t0 > timeout_s:
                raise TimeoutError(f"request {rid} still {st}")
            time.sleep(20e-6)

    def cancel(self, rid: int) -> None:
        _check(_lib.rtg_server_cancel(self._h, C.c_uint64(rid)))

    def stats(self) -> dict:
        st = ServerStats()
        _check(_lib.rtg_server_get_stats(self._h, C.byref(st)))
        return st.as_dict()

    def reset_stats(self) -> None:
        _check(_lib.rtg_server_reset_stats(self._h))

    def replay(self, rate_qps: float, nqueries: int, warmup: int = 0, pair_frac: float = 0.7,
               max_keys: int = 4096, seed: int = 1, want_results: bool = False):
        spec = StreamSpec(rate_qps=rate_qps, nqueries=nqueries, warmup=warmup, pair_frac=pair_frac,
                          max_keys=max_keys, seed=seed)
        pr = np.zeros(nqueries, dtype=np.uint32) if want_results else None
        ls = np.zeros(nqueries, dtype=np.uint64) if want_results else None
        st = ServerStats()
        _check(_lib.rtg_server_replay(self._h, C.byref(spec), _ptr(pr), _ptr(ls), C.byref(st)))
        return st.as_dict(), pr, ls

    def close(self) -> None:
        if self._h:
            _lib.rtg_server_destroy(self._h)
            self._h = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_queries(graph: DeviceGraph | None, rate_qps: float, nqueries: int, pair_frac: float = 0.7,
                   max_keys: int = 4096, seed: int = 1) -> dict:
    """The replay's query list (rtg_stream_queries), for recomputing every answer."""
    spec = StreamSpec(rate_qps=rate_qps, nqueries=nqueries, warmup=0, pair_frac=pair_frac, max_keys=max_keys, seed=seed)
    pn = C.c_int64()
    _check(_lib.rtg_stream_queries(graph._h if graph else None, C.byref(spec), None, None, None, None, None, None,
                                   C.byref(pn)))
    kinds = np.zeros(nqueries, np.int32)
    a = np.zeros(nqueries, np.int32)
    b = np.zeros(nqueries, np.int32)
    off = np.zeros(nqueries, np.int64)
    nk = np.zeros(nqueries, np.int32)
    pool = np.zeros(pn.value, np.uint32)
    _check(_lib.rtg_stream_queries(graph._h if graph else None, C.byref(spec), _ptr(kinds), _ptr(a), _ptr(b),
                                   _ptr(off), _ptr(nk), _ptr(pool), C.byref(pn)))
    return {"kind": kinds, "u": a, "v": b, "key_off": off, "nkeys": nk, "pool": pool}
