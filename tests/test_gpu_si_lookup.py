"""GPU parity of sorted-set intersection (pairs and the reference's A x B tile mode) and of the
sorted-array lookup, against the oracle (exact integers)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_oracle, load_rtg
from test_oracle import reference_style_dataset

pytestmark = pytest.mark.gpu

rtg = load_rtg()
orc = load_oracle()
SI_METHODS = list(rtg.SI_METHODS)


def zipf_workload(nlists, seed):
    lens = rtg.gen_zipf_lengths(nlists, 1.8, 65536, seed=seed)
    offs, el = rtg.gen_sorted_lists(lens, 1 << 24, seed=seed)
    pairs = np.arange(nlists, dtype=np.uint32)  # pair k = lists (2k, 2k+1)
    return offs, el, pairs


@pytest.mark.parametrize("method", SI_METHODS)
def test_si_pairs_zipf(ctx, method):
    offs, el, pairs = zipf_workload(1 << 15, 7)
    exp = orc.si_pairs(offs, el, pairs)
    d_off, d_el, d_p = ctx.to_device(offs), ctx.to_device(el), ctx.to_device(pairs)
    out, st = rtg.si_pairs(ctx, d_off, d_el, offs.size - 1, d_p, pairs.size // 2, method)
    assert np.array_equal(out.numpy(), exp)


def test_si_pairs_full_zipf_2e20(ctx):
    """The set-intersection microbench at full size: 2^21 Zipf(1.8) lists of uniform values
    below 2^24 (about 59M elements) paired into 2^20 pairs; every method against the oracle."""
    print("[zipf 2^20 pairs] generating 2^21 lists", flush=True)
    offs, el, pairs = zipf_workload(1 << 21, 1)
    assert 5.0e7 < el.size < 7.0e7
    exp = orc.si_pairs(offs, el, pairs)
    d_off, d_el, d_p = ctx.to_device(offs), ctx.to_device(el), ctx.to_device(pairs)
    for method in SI_METHODS:
        print(f"[zipf 2^20 pairs] method {method}", flush=True)
        out, st = rtg.si_pairs(ctx, d_off, d_el, offs.size - 1, d_p, pairs.size // 2, method)
        assert np.array_equal(out.numpy(), exp), method


@pytest.mark.parametrize("method", SI_METHODS)
def test_si_pairs_edge_cases(ctx, method):
    rng = np.random.default_rng(3)
    big = np.sort(rng.choice(1 << 24, size=65536, replace=False)).astype(np.uint32)
    lists = [np.zeros(0, np.uint32), big, big[::3].copy(), big[:1].copy(), np.array([big[-1]], np.uint32),
             np.arange(100, dtype=np.uint32), np.arange(50, 150, dtype=np.uint32), big[1000:9000].copy()]
    offs = np.zeros(len(lists) + 1, np.uint32)
    np.cumsum([x.size for x in lists], out=offs[1:])
    el = np.concatenate(lists).astype(np.uint32)
    pairs = np.array([(a, b) for a in range(len(lists)) for b in range(len(lists))], np.uint32).reshape(-1)
    exp = orc.si_pairs(offs, el, pairs)
    out, _ = rtg.si_pairs(ctx, ctx.to_device(offs), ctx.to_device(el), len(lists), ctx.to_device(pairs),
                          pairs.size // 2, method)
    assert np.array_equal(out.numpy(), exp)


@pytest.mark.parametrize("method", SI_METHODS)
def test_si_pairs_ids_out_of_range(ctx, method):
    """A pair naming a list id >= nlists reads as an empty list on the device (the offsets are
    never read past nlists) and the call reports RTG_EINVAL; the other pairs are still counted."""
    offs, el = balanced_lists(np.random.default_rng(12))
    n = offs.size - 1
    pairs = np.array([(a, b) for a in range(n) for b in range(n)], np.uint32)
    exp = orc.si_pairs(offs, el, pairs.reshape(-1))
    bad_rows = [1, 7]
    pairs[bad_rows[0]] = (0, n)            # the first id past the end
    pairs[bad_rows[1]] = (0xFFFFFFFF, 2)   # the largest id
    exp[bad_rows] = 0
    d_off, d_el, d_p = ctx.to_device(offs), ctx.to_device(el), ctx.to_device(pairs.reshape(-1))
    d_out = rtg.DeviceArray(ctx, (pairs.shape[0],), np.uint32)
    with pytest.raises(rtg.RTGError) as e:
        rtg.si_pairs(ctx, d_off, d_el, n, d_p, pairs.shape[0], method, d_out=d_out)
    assert e.value.code == 1  # RTG_EINVAL
    assert np.array_equal(d_out.numpy(), exp)
    # the same lists and pairs with every id in range are accepted again
    good = pairs.copy()
    good[bad_rows] = (0, 1)
    rtg.si_pairs(ctx, d_off, d_el, n, ctx.to_device(good.reshape(-1)), good.shape[0], method, d_out=d_out)


@pytest.mark.parametrize("method", SI_METHODS)
@pytest.mark.parametrize("len_c", [0, 5, 300])
def test_si_tile_reference_invariant(ctx, method, len_c):
    rng = np.random.default_rng(len_c + 1)
    offs, el = reference_style_dataset(4, 6, 2000, 400, len_c, rng)
    out, _ = rtg.si_tile(ctx, ctx.to_device(offs), ctx.to_device(el), 4, 6, method)
    assert np.all(out.numpy() == len_c)


def balanced_lists(rng):
    """Long lists for which a merge reads fewer bytes than binary searches: lengths 65 to
    70,000 over dense and sparse value ranges, lists at the top of the uint32 range (one holding
    0xFFFFFFFF, the hash table's empty marker), and nested, shifted and identical lists."""
    specs = [(65, 200, 0), (1000, 1500, 0), (5000, 6000, 100), (40000, 50000, 0), (40000, 1 << 24, 0),
             (3000, 1 << 30, 12345), (20000, 25000, (1 << 32) - 25000), (20000, 1 << 20, (1 << 32) - (1 << 20)),
             (70000, 80000, 7)]
    lists = []
    for n, span, base in specs:
        v = np.sort(rng.choice(span, size=n, replace=False)).astype(np.uint64) + base
        lists.append(v.astype(np.uint32))
    lists.append(np.unique(np.concatenate([lists[6][::2], np.array([0xFFFFFFFF], np.uint32)])))
    lists.append(np.unique(np.concatenate([lists[7][1::3], np.array([0xFFFFFFFF], np.uint32)])))
    lists.append(lists[3][::7].copy())
    lists.append((lists[1].astype(np.uint64) + 1200).astype(np.uint32))
    lists.append(lists[8][5:60005].copy())
    offs = np.zeros(len(lists) + 1, np.uint32)
    np.cumsum([x.size for x in lists], out=offs[1:])
    return offs, np.concatenate(lists).astype(np.uint32)


@pytest.mark.parametrize("method", SI_METHODS)
def test_si_pairs_balanced_long(ctx, method):
    """Every ordered pair of balanced_lists, repeated so that each wave runs several
    consecutive units of one pair (the merge, hash and bitmap units then continue from the
    previous unit's end instead of searching for their start)."""
    offs, el = balanced_lists(np.random.default_rng(21))
    n = offs.size - 1
    base = np.array([(a, b) for a in range(n) for b in range(n)], np.uint32)
    pairs = np.tile(base, (60, 1)).reshape(-1)
    exp = orc.si_pairs(offs, el, pairs)
    out, st = rtg.si_pairs(ctx, ctx.to_device(offs), ctx.to_device(el), n, ctx.to_device(pairs), pairs.size // 2,
                           method)
    assert np.array_equal(out.numpy(), exp)
    kind = {rtg.SI_MERGE: "merge", rtg.SI_HASH: "hash", rtg.SI_BITMAP: "bitmap"}.get(method)
    if kind:
        assert st["strategy"] == kind and st["units"][kind] > 0 and st["units"]["search"] > 0
    if method == rtg.SI_BITMAP:
        assert st["units"]["merge"] > 0  # the sparse balanced lists


@pytest.mark.parametrize("method", SI_METHODS)
def test_si_tile_balanced_against_oracle(ctx, method):
    offs, el = balanced_lists(np.random.default_rng(5))
    n = offs.size - 1
    num_a = 6
    exp = orc.si_tile(offs, el, num_a, n - num_a)
    out, _ = rtg.si_tile(ctx, ctx.to_device(offs), ctx.to_device(el), num_a, n - num_a, method)
    assert np.array_equal(out.numpy(), exp)


@pytest.mark.parametrize("length,skew,sel,density", [(100000, 1.0, 0.1, 0.1), (10000, 0.2, 0.5, 0.01), (60000, 4.0, 0.2, 0.9),
                                                     (30000, 0.5, 0.0, 0.02)])
def test_si_tile_reference_generator(ctx, length, skew, sel, density):
    """Datasets shaped by the reference's generator (data.cpp:21-74; the first row is the shape
    its scripts generate, s_generator -l 100000 -r 1 -s 0.1 -d 0.1 with 1000 A sets and one B
    set): every A x B count is len_c, for every method."""
    num_a = 1000 if length == 100000 else 200
    num_b = 1 if length == 100000 else 3
    print(f"[si dataset] len {length} skew {skew} sel {sel} density {density}", flush=True)
    ds = rtg.gen_si_dataset(num_a, num_b, length, skew, sel, density, seed=3)
    d_off, d_el = ctx.to_device(ds.set_offsets), ctx.to_device(ds.elements)
    for method in SI_METHODS:
        print(f"[si dataset] method {method}", flush=True)
        out, st = rtg.si_tile(ctx, d_off, d_el, num_a, num_b, method)
        assert np.all(out.numpy() == ds.intersection), (method, st)


def test_si_auto_remembers_choice(ctx):
    offs, el = balanced_lists(np.random.default_rng(8))
    n = offs.size - 1
    pairs = np.tile(np.array([(a, b) for a in range(n) for b in range(n)], np.uint32), (20, 1)).reshape(-1)
    exp = orc.si_pairs(offs, el, pairs)
    d_off, d_el, d_p = ctx.to_device(offs), ctx.to_device(el), ctx.to_device(pairs)
    out1, st1 = rtg.si_pairs(ctx, d_off, d_el, n, d_p, pairs.size // 2, rtg.SI_AUTO)
    out2, st2 = rtg.si_pairs(ctx, d_off, d_el, n, d_p, pairs.size // 2, rtg.SI_AUTO)
    assert np.array_equal(out1.numpy(), exp) and np.array_equal(out2.numpy(), exp)
    assert st1["t_plan_ms"] > 0 and st2["t_plan_ms"] == 0
    assert st1["strategy"] == st2["strategy"] != "none"


def test_binary_search_based_intersection_dataset(tmp_path):
    rng = np.random.default_rng(9)
    offs, el = reference_style_dataset(2, 3, 500, 700, 120, rng)
    ds = rtg.Dataset(2, 3, int(el.max()), 120, offs, el)
    path = tmp_path / "x.s_data"
    rtg.write_dataset(path, ds)
    ds2 = rtg.read_dataset(path)
    res = rtg.BinarySearchBasedIntersection(ds2)
    assert res.tolist() == [120] * 6


@pytest.mark.parametrize("n", [0, 1, 2, 1000, 1 << 16, (1 << 20) + 3])
def test_lookup_u32(ctx, n):
    rng = np.random.default_rng(n)
    data = np.sort(rng.integers(0, 1 << 32, size=n, dtype=np.uint64)).astype(np.uint32)
    if n > 10:
        data[n // 2: n // 2 + 5] = data[n // 2]  # duplicates
    keys = np.concatenate([rng.integers(0, 1 << 32, size=5000, dtype=np.uint64).astype(np.uint32),
                           data[rng.integers(0, max(n, 1), size=5000)] if n else np.zeros(0, np.uint32),
                           np.array([0, 0xFFFFFFFF], np.uint32)])
    exp = orc.lookup(data, keys)
    d_data = ctx.to_device(data if n else np.zeros(1, np.uint32))
    out, _ = rtg.lookup(ctx, d_data, n, ctx.to_device(keys), keys.size)
    assert np.array_equal(out.numpy(), exp)


def test_lookup_signed_reference_interface():
    rng = np.random.default_rng(2)
    data = rng.integers(-1000, 1000, size=20000).astype(np.int32)
    q = rng.integers(-1100, 1100, size=30000).astype(np.int32)
    ref_sorted = np.sort(data)
    exp = orc.lookup(ref_sorted, q, signed=True)
    got = rtg.BSAccessCountFunction(data, q)
    assert np.array_equal(data, ref_sorted)  # sorted in place, as the reference does
    assert np.array_equal(got, exp)


# ---- search index (rtg_lookup_index) ---------------------------------------------------------
def _index_keys(data, rng, signed=False):
    dt = np.int32 if signed else np.uint32
    lo, hi = (-(1 << 31), 1 << 31) if signed else (0, 1 << 32)
    n = data.size
    parts = [rng.integers(lo, hi, size=4000, dtype=np.int64).astype(dt),
             np.array([lo, hi - 1, 0, 1], dtype=np.int64).astype(dt)]
    if n:
        parts.append(data[rng.integers(0, n, size=4000)])
        edge = np.array([data[0], data[-1]], dtype=np.int64)
        parts.append(np.concatenate([edge, edge - 1, edge + 1]).clip(lo, hi - 1).astype(dt))
    return np.concatenate(parts)


# sizes around the node (32) and LDS-top (8192) boundaries and every level count up to 3
INDEX_SIZES = [0, 1, 31, 32, 33, 8191, 8192, 8193, 8192 * 32, 8192 * 32 + 1, (1 << 20) + 3,
               8192 * 32 * 32 + 7]


@pytest.mark.parametrize("n", INDEX_SIZES)
def test_lookup_index_u32(ctx, n):
    rng = np.random.default_rng(1000 + n)
    data = np.sort(rng.integers(0, 1 << 32, size=n, dtype=np.uint64)).astype(np.uint32)
    if n > 200:
        data[n // 3: n // 3 + 100] = data[n // 3]  # a run of duplicates spanning node boundaries
    keys = _index_keys(data, rng)
    exp = orc.lookup(data, keys)
    d_data = ctx.to_device(data if n else np.zeros(1, np.uint32))
    d_keys = ctx.to_device(keys)
    with rtg.LookupIndex(ctx, d_data, n) as ix:
        assert ix.levels == (0 if n <= 8192 else 1 if n <= 8192 * 32 else 2 if n <= 8192 * 32 * 32 else 3)
        out, _ = ix.run(d_keys, keys.size)
        assert np.array_equal(out.numpy(), exp)
        # a second run on the same index, on fewer keys, into the same buffer
        out2, _ = ix.run(d_keys, keys.size // 2, out)
        assert np.array_equal(out2.numpy()[: keys.size // 2], exp[: keys.size // 2])


@pytest.mark.parametrize("n", [8192 * 32 + 5, 100000])
def test_lookup_index_max_value_tail(ctx, n):
    """Array ending in runs of 0xFFFFFFFF (the value the index pads with): first occurrences."""
    rng = np.random.default_rng(n)
    data = np.sort(rng.integers(0, 1 << 32, size=n, dtype=np.uint64)).astype(np.uint32)
    data[-70:] = 0xFFFFFFFF
    keys = np.concatenate([_index_keys(data, rng), np.full(3, 0xFFFFFFFF, np.uint32)])
    exp = orc.lookup(data, keys)
    assert exp[-1] == n - 70
    d_data = ctx.to_device(data)
    with rtg.LookupIndex(ctx, d_data, n) as ix:
        out, _ = ix.run(ctx.to_device(keys), keys.size)
        assert np.array_equal(out.numpy(), exp)


@pytest.mark.parametrize("n", [5000, 300007])
def test_lookup_index_i32(ctx, n):
    rng = np.random.default_rng(n)
    data = np.sort(rng.integers(-(1 << 31), 1 << 31, size=n, dtype=np.int64)).astype(np.int32)
    data[-3:] = np.iinfo(np.int32).max
    keys = _index_keys(data, rng, signed=True)
    exp = orc.lookup(data, keys, signed=True)
    d_data = ctx.to_device(data)
    with rtg.LookupIndex(ctx, d_data, n, signed=True) as ix:
        out, _ = ix.run(ctx.to_device(keys), keys.size)
        assert np.array_equal(out.numpy(), exp)
    # same answers as the stateless lookup
    out_s, _ = rtg.lookup(ctx, d_data, n, ctx.to_device(keys), keys.size, signed=True)
    assert np.array_equal(out_s.numpy(), exp)


def test_lookup_index_rejects_unaligned(ctx):
    d = ctx.to_device(np.arange(100000, dtype=np.uint32))
    with pytest.raises(rtg.RTGError):
        rtg.LookupIndex(ctx, d.ptr + 4, 99999)


def test_lookup_index_full_size(ctx):
    """2^28 sorted uint32 (1 GiB, the serving array): index vs the oracle on 2^20 keys and vs
    the stateless sampled lookup on 2^22 keys."""
    n = 1 << 28
    d_data = rtg.gen_sorted_array_device(ctx, n, seed=1)
    rng = np.random.default_rng(28)
    host = d_data.numpy()
    keys = np.concatenate([rng.integers(0, 1 << 32, size=1 << 19, dtype=np.uint64).astype(np.uint32),
                           host[rng.integers(0, n, size=1 << 19)]])
    exp = orc.lookup(host, keys)
    with rtg.LookupIndex(ctx, d_data, n) as ix:
        assert ix.levels == 3 and ix.nbytes < n // 7
        out, _ = ix.run(ctx.to_device(keys), keys.size)
        assert np.array_equal(out.numpy(), exp)
        big = rtg.gen_uniform_u32(1 << 22, seed=5)
        d_big = ctx.to_device(big)
        a, _ = ix.run(d_big, big.size)
        b, _ = rtg.lookup(ctx, d_data, n, d_big, big.size)
        assert np.array_equal(a.numpy(), b.numpy())


@pytest.mark.parametrize("n,q", [(300_000, 40_000), (300_000, 2_000), (5_000, 3)])
def test_lookup_stateless_both_forms(ctx, n, q):
    """rtg_lookup_u32 searches through a transient index when the keys are many (q >= n/128) or
    the array is small, and by the sampled kernel otherwise or when the array is not 16-byte
    aligned; every form matches the oracle."""
    rng = np.random.default_rng(n + q)
    data = np.sort(rng.integers(0, 1 << 32, size=n + 1, dtype=np.uint64)).astype(np.uint32)
    keys = np.concatenate([rng.integers(0, 1 << 32, size=q, dtype=np.uint64).astype(np.uint32),
                           data[rng.integers(0, n + 1, size=q)]])
    d_data, d_keys = ctx.to_device(data), ctx.to_device(keys)
    out, _ = rtg.lookup(ctx, d_data, n + 1, d_keys, keys.size)
    assert np.array_equal(out.numpy(), orc.lookup(data, keys))
    # the same array seen from its second element: 4-byte aligned only
    out2, _ = rtg.lookup(ctx, d_data.ptr + 4, n, d_keys, keys.size)
    assert np.array_equal(out2.numpy(), orc.lookup(data[1:], keys))


@pytest.mark.parametrize("method", SI_METHODS)
def test_set_intersection_tile_host_entry(method):
    """rtg_set_intersection_tile_host, the C form behind the compat header's baseline drop-ins
    (host arrays in, host counts out), for every method."""
    import ctypes as C
    rng = np.random.default_rng(11)
    offs, el = reference_style_dataset(5, 4, 3000, 2500, 700, rng)
    out = np.zeros(20, np.uint32)
    ms = C.c_double()
    rc = rtg._lib.rtg_set_intersection_tile_host(rtg._ptr(offs), rtg._ptr(el), C.c_uint32(5), C.c_uint32(4),
                                                 C.c_int32(method), C.c_int(0), rtg._ptr(out), C.byref(ms))
    assert rc == 0 and np.all(out == 700) and ms.value > 0


# ---- device sort of the lookup array (rtg_sort_i32 / rtg_sort_u32) ---------------------------
@pytest.mark.parametrize("n", [0, 1, 2, 63, 4095, 4096, 4097, 100_003, (1 << 20) + 3])
@pytest.mark.parametrize("signed", [True, False])
def test_device_sort_32(ctx, n, signed):
    rng = np.random.default_rng(n + signed)
    if signed:
        a = rng.integers(-(1 << 31), 1 << 31, size=n, dtype=np.int64).astype(np.int32)
        if n > 8:
            a[:4] = [np.iinfo(np.int32).min, np.iinfo(np.int32).max, -1, 0]
            a[n // 2: n // 2 + 3] = a[5]  # duplicates
    else:
        a = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
        if n > 8:
            a[:4] = [0, 0xFFFFFFFF, 0x80000000, 0x7FFFFFFF]
            a[n // 2: n // 2 + 3] = a[5]
    d = ctx.to_device(a if n else np.zeros(1, a.dtype))
    rtg.sort_device(ctx, d, n, signed=signed)
    assert np.array_equal(d.numpy()[:n], np.sort(a))


def test_device_sort_reference_lookup_shape(ctx):
    """The reference's lookup workload shape (rt_search/measure.cpp:32-51): int32 uniform in
    [0, 10N] with duplicates, unsorted; device-sorted, then every query answered as the
    host-sorted array's first occurrence."""
    n, q = 1 << 20, 1 << 16
    rng = np.random.default_rng(20)
    data = rng.integers(0, 10 * n + 1, size=n).astype(np.int32)
    query = rng.integers(0, 10 * n + 1, size=q).astype(np.int32)
    d_data, d_q = ctx.to_device(data), ctx.to_device(query)
    rtg.sort_device(ctx, d_data, n, signed=True)
    ref_sorted = np.sort(data)
    assert np.array_equal(d_data.numpy(), ref_sorted)
    out, _ = rtg.lookup(ctx, d_data, n, d_q, q, signed=True)
    assert np.array_equal(out.numpy(), orc.lookup(ref_sorted, query, signed=True))


def test_bs_access_count_function_c_abi(tmp_path):
    """The C drop-in BSAccessCountFunction (rt_search/simple_bs.h:5) called from a reference-style
    driver: results against the oracle, `data` sorted in place, and the reference's report lines
    (simple_bs.cu:88, :95, :137-139)."""
    import re
    import shutil
    import subprocess
    from conftest import ROOT
    cxx = shutil.which("g++") or shutil.which("clang++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    rng = np.random.default_rng(77)
    n, q = 50_000, 20_000
    data = rng.integers(-5000, 500_000, size=n).astype(np.int32)
    query = np.concatenate([rng.integers(-6000, 510_000, size=q - 2), [-5000, 499_999]]).astype(np.int32)
    (tmp_path / "data.bin").write_bytes(data.tobytes())
    (tmp_path / "query.bin").write_bytes(query.tobytes())
    src = tmp_path / "driver.cpp"
    src.write_text(r'''
#include <cstdio>
#include <vector>
#include "rtg_compat.h"
int main(int, char **argv) {
    std::vector<int> data(%d), query(%d), res(%d);
    FILE *f = std::fopen(argv[1], "rb"); if (std::fread(data.data(), 4, data.size(), f) != data.size()) return 3; std::fclose(f);
    f = std::fopen(argv[2], "rb"); if (std::fread(query.data(), 4, query.size(), f) != query.size()) return 3; std::fclose(f);
    BSAccessCountFunction(data.data(), (int)data.size(), query.data(), (int)query.size(), res.data());
    f = std::fopen(argv[3], "wb"); std::fwrite(res.data(), 4, res.size(), f); std::fwrite(data.data(), 4, data.size(), f); std::fclose(f);
    return 0;
}
''' % (n, q, q))
    lib = rtg.lib_path()
    exe = tmp_path / "driver"
    proc = subprocess.run([cxx, "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
                           f"-L{lib.parent}", "-lrtg", f"-Wl,-rpath,{lib.parent}"], capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr
    run = subprocess.run([str(exe), str(tmp_path / "data.bin"), str(tmp_path / "query.bin"), str(tmp_path / "out.bin")],
                         capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    lines = run.stdout.strip().splitlines()
    assert lines[0] == f"data size = {n}"
    assert re.fullmatch(r"cpu sort time = \d+\.\d{3} ms", lines[1])
    assert re.fullmatch(r"Sorting time = \d+\.\d{6} ms, Binary search time = \d+\.\d{6} ms, Total time = \d+\.\d{6} ms",
                        lines[2])
    out = np.frombuffer((tmp_path / "out.bin").read_bytes(), dtype=np.int32)
    res, sorted_data = out[:q], out[q:]
    ref_sorted = np.sort(data)
    assert np.array_equal(sorted_data, ref_sorted)
    assert np.array_equal(res, orc.lookup(ref_sorted, query, signed=True))
