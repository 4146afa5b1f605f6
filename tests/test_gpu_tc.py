"""GPU parity of the triangle-counting path: the device CSR build must be byte-identical to
the oracle's restatement of the reference preprocessing, and every counting method must
return the oracle's exact count (integer work: bit-exact, no tolerance)."""
from __future__ import annotations

import hashlib
import itertools
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import load_oracle, load_rtg

pytestmark = pytest.mark.gpu

rtg = load_rtg()
orc = load_oracle()
GOLD = json.loads((Path(__file__).parent / "golden" / "golden.json").read_text())
METHODS = [rtg.TC_PLAIN_THREAD, rtg.TC_PLAIN_WAVE, rtg.TC_BATCHED]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _graphs():
    rng = np.random.default_rng(5)
    g = {}
    g["hand"] = np.array([1, 2, 2, 3, 3, 1, 3, 4, 4, 5, 7, 3, 2, 1, 3, 3, 1, 2], dtype=np.int32)
    g["k40"] = np.array(list(itertools.combinations(range(40), 2)), dtype=np.int32).reshape(-1)
    g["wheel"] = np.array([(0, i) for i in range(1, 301)] + [(i, i % 300 + 1) for i in range(1, 301)],
                          dtype=np.int32).reshape(-1)
    e = rng.integers(0, 3000, size=(60000, 2)).astype(np.int32)  # duplicates and self-loops
    g["gnm_dups"] = e.reshape(-1)
    # self-loops count in the degree but leave no key; every pair here is a loop or touches a
    # degree-1 vertex, so the build has vertices (0 and 3) and no edge
    g["loops_only"] = np.array([0, 0, 0, 0, 1, 2, 3, 3, 3, 3, 3, 4], dtype=np.int32)
    g["ba_100000"] = rtg.gen_ba(100_000, 5, 1)
    # dense and nearly regular (degrees about 440): out-lists of 150-300 entries whose largest
    # entries fall anywhere in a partner's row, so span cuts land at every offset
    g["gnp_dense"] = rng.integers(0, 2000, size=(500_000, 2)).astype(np.int32).reshape(-1)
    g["rmat_s12"] = rtg.gen_rmat(12, 16 << 12, seed=1)
    g["rmat_s16"] = rtg.gen_rmat(16, 44 << 16, seed=1)
    deg = rtg.gen_powerlaw_degrees(20_000, 2.1, 2_000.0, 20.0, seed=1)
    g["plcm"] = rtg.gen_config_model(deg, 0.3, seed=1)
    # K_1200 (out-degrees up to 1199) plus a vertex x whose 2500 neighbours A_i each have 2501
    # leaves, so x keeps out-degree 2500 after the leaves are removed; A_i - A_i+1 closes
    # 2499 triangles through x.
    k = 1200
    clique = np.array(list(itertools.combinations(range(k), 2)), dtype=np.int32)
    x, a0, na, leaves = k, k + 1, 2500, 2501
    A = np.arange(a0, a0 + na, dtype=np.int32)
    star = np.stack([np.full(na, x, np.int32), A], 1)
    path = np.stack([A[:-1], A[1:]], 1)
    lf = np.stack([np.repeat(A, leaves), a0 + na + np.arange(na * leaves, dtype=np.int32)], 1)
    g["hub"] = np.concatenate([clique, star, path, lf]).astype(np.int32).reshape(-1)
    return g


GRAPHS = _graphs()
_ORACLE = {}


def oracle_csr(name):
    if name not in _ORACLE:
        off, adj = orc.csr_from_pairs(GRAPHS[name])
        _ORACLE[name] = (off, adj, orc.tc_count_intersect(off, adj))
    return _ORACLE[name]


@pytest.mark.parametrize("name", list(GRAPHS))
def test_csr_build_byte_identical(ctx, name):
    print(f"[csr] {name}", flush=True)
    off, adj, _ = oracle_csr(name)
    g = rtg.DeviceGraph.from_pairs(ctx, GRAPHS[name])
    doff, dadj = g.download()
    assert doff.dtype == np.int32 and dadj.dtype == np.int32
    assert np.array_equal(doff, off)
    assert np.array_equal(dadj, adj)
    assert g.max_row == (int(np.diff(off).max()) if off.size > 1 else 0)


@pytest.mark.parametrize("name", list(GRAPHS))
@pytest.mark.parametrize("method", METHODS)
def test_count_matches_oracle(ctx, name, method):
    print(f"[count] {name} method={method}", flush=True)
    off, adj, t = oracle_csr(name)
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    cnt, st = g.tc_count(method=method)
    assert cnt == t
    assert st["edges"] == adj.size


def test_golden_counts_ba_100000(ctx):
    e = GOLD["ba_n100000_m5_seed1"]
    g = rtg.DeviceGraph.from_pairs(ctx, rtg.gen_ba(100_000, 5, 1))
    off, adj = g.download()
    assert sha(off.astype("<i4")) == e["offsets_sha256"] and sha(adj.astype("<i4")) == e["adjs_sha256"]
    for m in METHODS + [rtg.TC_2A1]:
        assert g.tc_count(method=m)[0] == e["triangles"]


def test_two_hop_2a1(ctx):
    for name in ("hand", "k40", "rmat_s12", "plcm", "gnm_dups"):
        off, adj, t = oracle_csr(name)
        g = rtg.DeviceGraph.from_csr(ctx, off, adj)
        assert g.tc_count(method=rtg.TC_2A1)[0] == t, name


@pytest.mark.parametrize("name", ["rmat_s16", "plcm", "hub"])
def test_per_edge_counts(ctx, name):
    off, adj, _ = oracle_csr(name)
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    for form, oform in ((0, 0), (1, 1), (2, 0)):  # 0 batched plan and 2 wave per edge: a->b; 1: a->c
        assert np.array_equal(g.tc_per_edge(form), orc.tc_per_edge(off, adj, oform)), form


@pytest.mark.parametrize("name", list(GRAPHS))
def test_per_edge_plan(ctx, name):
    """Batched per-edge plans (every span's count stored by position, then gathered in edge
    order) on every test graph, with and without an edge range (entries outside it untouched);
    run twice to check that every run rewrites the output range."""
    off, adj, t = oracle_csr(name)
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    exp = orc.tc_per_edge(off, adj, 0)
    m = adj.size
    plan = g.tc_plan(method=rtg.TC_BATCHED, use_hip_graph=False, per_edge=True)
    for _ in range(2):
        out, cnt, st = plan.run_per_edge()
        assert cnt == t and np.array_equal(out.numpy()[:m], exp)
    assert plan.run()[0] == t  # the plan still counts
    plan.close()
    if m > 10:
        a, b = m // 3, m - m // 4
        d = rtg.DeviceArray(ctx, (m,), np.uint32)
        d.copy_from(np.full(m, 7, np.uint32))
        p2 = g.tc_plan(method=rtg.TC_BATCHED, use_hip_graph=False, e_begin=a, e_end=b, per_edge=True)
        _, cnt, _ = p2.run_per_edge(d)
        got = d.numpy()
        assert cnt == int(exp[a:b].sum())
        assert np.array_equal(got[a:b], exp[a:b]) and np.all(got[:a] == 7) and np.all(got[b:] == 7)
        p2.close()


def test_per_edge_plan_search_items(ctx, monkeypatch):
    """Per-edge counts from owners too long for the LDS table (search items): a cap of 64 slots
    sends every owner there."""
    for name in ("hub", "plcm"):
        off, adj, t = oracle_csr(name)
        g = rtg.DeviceGraph.from_csr(ctx, off, adj)
        monkeypatch.setenv("RTG_TC_TUNE", "threads=256,unroll=8,item=65536,wave_item=4096,long=128,grab=16,table_cap=64")
        plan = g.tc_plan(method=rtg.TC_BATCHED, use_hip_graph=False, per_edge=True)
        assert int(plan.describe()["search_items"]) > 0
        out, cnt, _ = plan.run_per_edge()
        assert cnt == t and np.array_equal(out.numpy()[:adj.size], orc.tc_per_edge(off, adj, 0)), name
        plan.close()


def test_per_edge_plan_rejects_count_plan(ctx):
    off, adj, _ = oracle_csr("rmat_s12")
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    plan = g.tc_plan(method=rtg.TC_BATCHED)
    with pytest.raises(rtg.RTGError):
        plan.run_per_edge()
    plan.close()


def test_edge_ranges_partition(ctx):
    off, adj, t = oracle_csr("rmat_s16")
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    m = adj.size
    cuts = [0, m // 5, m // 2, m - 7, m]
    for method in METHODS:
        parts = [g.tc_count(method=method, e_begin=a, e_end=b)[0] for a, b in zip(cuts, cuts[1:])]
        assert sum(parts) == t
        assert parts[1] == orc.tc_count_1a2(off, adj, cuts[1], cuts[2])


def test_hip_graph_replays(ctx):
    off, adj, t = oracle_csr("plcm")
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    plan = g.tc_plan(method=rtg.TC_BATCHED, use_hip_graph=True)
    for i in range(4):
        cnt, st = plan.run()
        assert cnt == t, i
        # eager work between replays
        assert g.tc_count(method=rtg.TC_PLAIN_WAVE)[0] == t
    plan.close()


def test_autotune_and_cache(ctx, tmp_path):
    off, adj, t = oracle_csr("plcm")
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    cache = tmp_path / "tune.json"
    p1 = g.tc_plan(method=rtg.TC_BATCHED, autotune=True, tune_cache=str(cache))
    assert p1.describe()["source"] == "autotune"
    assert p1.run()[0] == t
    assert cache.exists() and "gfx" in cache.read_text()
    p2 = g.tc_plan(method=rtg.TC_BATCHED, autotune=True, tune_cache=str(cache))
    d2 = p2.describe()
    assert d2["source"] == "cache" and d2["unroll"] == p1.describe()["unroll"]
    assert p2.run()[0] == t


def test_every_candidate_configuration(ctx, monkeypatch):
    off, adj, t = oracle_csr("hub")
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    for tune in ("threads=256,unroll=2,item=1000,wave_item=64,wave_max=0,long=1,grab=1",
                 "threads=512,unroll=4,item=50000,wave_item=100000,wave_max=2000,long=1000000,grab=64",
                 "threads=256,unroll=8,item=131072,wave_item=4096,wave_max=64,long=32,grab=16"):
        monkeypatch.setenv("RTG_TC_TUNE", tune)
        p = g.tc_plan(method=rtg.TC_BATCHED)
        assert p.describe()["source"] == "env"
        assert p.run()[0] == t, tune
        p.close()


def _zero_heavy_csr(seed=11, n=3000):
    """A sorted CSR (not an orientation: any CSR is valid input to the count) whose out-lists
    mostly hold vertex 0 and have ragged lengths 0..700, so that long partner spans end inside
    a round and the owners' tables hold 0: what a load past a span's end returns."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 701, size=n)
    short = rng.random(n) < 0.3
    lens[short] = rng.integers(0, 70, size=int(short.sum()))
    rows = []
    for u in range(n):
        row = rng.choice(np.arange(1, n), size=min(int(lens[u]), n - 1), replace=False)
        if u and rng.random() < 0.8:
            row = np.concatenate([[0], row[:-1]]) if row.size else np.array([0])
        rows.append(np.unique(row[row != u]).astype(np.int32))  # no self-loops
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum([r.size for r in rows])
    return off, np.concatenate(rows).astype(np.int32)


@pytest.mark.parametrize("unroll", [2, 4, 8])
def test_long_spans_with_vertex_zero(ctx, monkeypatch, unroll):
    """Long partner spans are read by loads bounded by the span, which return 0 past its end;
    the probes of those lanes are taken back when the owner's list holds 0. Every span long
    (long=1) or only some (long=128), counts and per-edge counts against the oracle."""
    off, adj = _zero_heavy_csr()
    assert (adj == 0).sum() > 1000
    t = orc.tc_count_intersect(off, adj)
    exp = orc.tc_per_edge(off, adj, 0)
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    for long_min in (1, 128):
        monkeypatch.setenv("RTG_TC_TUNE", f"threads=256,unroll={unroll},item=65536,wave_item=4096,long={long_min},grab=16")
        p = g.tc_plan(method=rtg.TC_BATCHED)
        assert p.run()[0] == t, long_min
        p.close()
        p = g.tc_plan(method=rtg.TC_BATCHED, use_hip_graph=False, per_edge=True)
        out, cnt, _ = p.run_per_edge()
        assert cnt == t and np.array_equal(out.numpy()[: adj.size], exp), long_min
        out.free()
        p.close()


@pytest.mark.parametrize("cap", [2048, 64])
def test_owners_beyond_the_lds_table(ctx, monkeypatch, cap):
    """Owners whose cuckoo table would exceed the LDS cap are binary-searched in HBM instead
    (search items); a cap of 64 slots sends every owner there."""
    for name in ("hub", "rmat_s16", "plcm"):
        off, adj, t = oracle_csr(name)
        g = rtg.DeviceGraph.from_csr(ctx, off, adj)
        monkeypatch.setenv("RTG_TC_TUNE", f"threads=256,unroll=8,item=65536,wave_item=4096,long=128,grab=16,table_cap={cap}")
        p = g.tc_plan(method=rtg.TC_BATCHED)
        d = p.describe()
        assert int(d["table_slots"]) <= cap, (name, d)
        if name == "hub" or cap == 64:  # out-degree 2500 needs 6272 slots
            assert int(d["search_items"]) > 0, (name, d)
        assert p.run()[0] == t, name
        p.close()


def test_symmetric_csr(ctx):
    p = GRAPHS["gnm_dups"]
    rp, nodes = orc.relabel(p)
    off, adj = orc.symmetric_csr(rp, nodes)
    g = rtg.DeviceGraph.from_pairs(ctx, p, relabel=True, orient=False)
    doff, dadj = g.download()
    assert np.array_equal(doff, off) and np.array_equal(dadj, adj)


def test_device_pairs_input_and_generator(ctx):
    host = rtg.gen_rmat(14, 20 << 14, seed=3)
    dev = rtg.gen_rmat_device(ctx, 14, 20 << 14, seed=3)
    assert np.array_equal(dev.numpy(), host)
    g1 = rtg.DeviceGraph.from_pairs(ctx, dev)
    g2 = rtg.DeviceGraph.from_pairs(ctx, host)
    a, b = g1.download(), g2.download()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_invalid_csr_rejected(ctx):
    with pytest.raises(rtg.RTGError):
        rtg.DeviceGraph.from_csr(ctx, np.array([0, 2, 2], np.int32), np.array([1, 1], np.int32))  # duplicate
    with pytest.raises(rtg.RTGError):
        rtg.DeviceGraph.from_csr(ctx, np.array([0, 1, 1], np.int32), np.array([5], np.int32))  # out of range
    with pytest.raises(rtg.RTGError):
        rtg.DeviceGraph.from_pairs(ctx, np.array([0, -3], np.int32))


def test_empty_graph(ctx):
    g = rtg.DeviceGraph.from_pairs(ctx, np.zeros(0, np.int32))
    assert g.n == 0 and g.m == 0
    for m in METHODS + [rtg.TC_2A1]:
        assert g.tc_count(method=m)[0] == 0


def test_reference_interface(capsys):
    off, adj, t = oracle_csr("rmat_s12")
    cnt = rtg.CountOnGPU(adj, off, off.size - 1, adj.size, 0)
    out = capsys.readouterr().out
    assert cnt == t
    assert f"counting result = {t & 0xFFFFFFFF}" in out and "BSTC: counting time = " in out
    G = rtg.Graph(0)
    G.SetEdges(GRAPHS["rmat_s12"])
    G.TCPreprocessing()
    G.ConvertToCSR()
    assert np.array_equal(G.offsets_, off) and np.array_equal(G.adjs_, adj)
    assert (G.graph_info_.nodes, G.graph_info_.edges) == (off.size - 1, adj.size)


def test_full_rmat_s21_ef44(ctx):
    """The benchmark graph (R-MAT scale 21, edge factor 44) against its golden count, CSR hashes
    and byte model."""
    e = GOLD["rmat_s21_ef44_seed1"]
    print("[full] generating R-MAT s21 ef44 on the device", flush=True)
    dev = rtg.gen_rmat_device(ctx, 21, 44 << 21, seed=1)
    g = rtg.DeviceGraph.from_pairs(ctx, dev)
    dev.free()
    assert (g.n, g.m, g.max_row) == (e["n"], e["m"], e["max_out_degree"])
    off, adj = g.download()
    assert sha(off.astype("<i4")) == e["offsets_sha256"] and sha(adj.astype("<i4")) == e["adjs_sha256"]
    print("[full] batched count", flush=True)
    cnt, st = g.tc_count(method=rtg.TC_BATCHED)
    assert cnt == e["triangles"]
    assert st["alg_bytes"] == e["alg_bytes"]
    print("[full] plain wave count", flush=True)
    assert g.tc_count(method=rtg.TC_PLAIN_WAVE)[0] == e["triangles"]


def test_full_plcm_n3000000(ctx):
    """The full power-law configuration model (n = 3,000,000, closing 0.3, the serving graph),
    generated and preprocessed on the device: CSR hashes, batched and plain counts and the
    byte model against the golden values."""
    e = GOLD.get("plcm_n3000000_seed1")
    if e is None:
        pytest.skip("golden entry not generated")
    deg = rtg.gen_powerlaw_degrees(3_000_000, 2.1, 30_000.0, 50.0, seed=1)
    assert sha(deg.astype("<i4")) == e["degrees_sha256"]
    print("[full] generating the 3M-vertex configuration model on the device", flush=True)
    dev = rtg.gen_config_model_device(ctx, deg, 0.3, seed=1)
    g = rtg.DeviceGraph.from_pairs(ctx, dev)
    dev.free()
    assert (g.n, g.m, g.max_row) == (e["n"], e["m"], e["max_out_degree"])
    off, adj = g.download()
    assert sha(off.astype("<i4")) == e["offsets_sha256"] and sha(adj.astype("<i4")) == e["adjs_sha256"]
    print("[full] batched count", flush=True)
    cnt, st = g.tc_count(method=rtg.TC_BATCHED)
    assert cnt == e["triangles"]
    assert st["alg_bytes"] == e["alg_bytes"]
    print("[full] plain wave count", flush=True)
    assert g.tc_count(method=rtg.TC_PLAIN_WAVE)[0] == e["triangles"]


def test_bs_tc_command_line(tmp_path):
    """The reference's driver contract, bs_tc <graph_file> <device_id> (tc/bs_tc/main.cu:6-20),
    on the same graph as binary pairs, a text edge list and a Matrix Market file."""
    import subprocess
    cli = Path(rtg.lib_path()).with_name("bs_tc")
    assert cli.exists(), "rt-graph_amd/bs_tc is not built"
    pairs = GRAPHS["rmat_s12"]
    off, adj, t = oracle_csr("rmat_s12")
    p = pairs.reshape(-1, 2)
    pairs.astype("<i4").tofile(tmp_path / "g.bin")
    (tmp_path / "g.txt").write_text("".join(f"{a} {b}\n" for a, b in p))
    ids = p.astype(np.int64) + 1
    n = int(ids.max())
    (tmp_path / "g.mtx").write_text(f"%%MatrixMarket matrix coordinate pattern general\n{n} {n} {len(ids)}\n"
                                    + "".join(f"{a} {b}\n" for a, b in ids))
    for name in ("g.bin", "g.txt", "g.mtx"):
        r = subprocess.run([str(cli), str(tmp_path / name), "0"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (name, r.stdout, r.stderr)
        assert f"counting result = {t & 0xFFFFFFFF}" in r.stdout, (name, r.stdout)


def test_csr_cache_save_load(ctx, tmp_path):
    """A preprocessed graph saved to the CSR cache and loaded back counts the same; the
    symmetric (unoriented) form keeps its flag; a stale key is refused."""
    pairs = GRAPHS["rmat_s12"]
    key = rtg.pairs_fingerprint(pairs)
    g = rtg.DeviceGraph.from_pairs(ctx, pairs)
    t = g.tc_count()[0]
    g.save(tmp_path / "g.csr", key)
    g2 = rtg.DeviceGraph.load(ctx, tmp_path / "g.csr", key)
    o1, a1 = g.download()
    o2, a2 = g2.download()
    assert np.array_equal(o1, o2) and np.array_equal(a1, a2) and g2.tc_count()[0] == t
    with pytest.raises(rtg.RTGError):
        rtg.DeviceGraph.load(ctx, tmp_path / "g.csr", key ^ 1)
    s = rtg.DeviceGraph.from_pairs(ctx, pairs, orient=False)
    s.save(tmp_path / "s.csr", 5)
    off, adj, oriented = rtg.csr_file_read(tmp_path / "s.csr", 5)
    assert not oriented
    s2 = rtg.DeviceGraph.load(ctx, tmp_path / "s.csr", 5)
    with pytest.raises(rtg.RTGError):  # counting needs a degree-oriented graph
        s2.tc_count()


def _banded_csr(seed=3, n=20_000):
    """An oriented-style CSR with locality: vertex u's out-list lies in (u, u + 4000], some
    owners hold long runs (up to 3000 consecutive ids), so that bitmaps over the owners' value
    ranges fit LDS and every strategy has owners it applies to in every bucket."""
    rng = np.random.default_rng(seed)
    rows = []
    for u in range(n):
        hi = min(n, u + 4001)
        if hi <= u + 1:
            rows.append(np.zeros(0, np.int32))
            continue
        if u % 97 == 0:
            k = min(hi - u - 1, int(rng.integers(1000, 3000)))
        elif u % 13 == 0:
            k = min(hi - u - 1, int(rng.integers(150, 600)))
        else:
            k = min(hi - u - 1, int(rng.integers(0, 60)))
        rows.append(np.sort(rng.choice(np.arange(u + 1, hi), size=k, replace=False)).astype(np.int32))
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum([r.size for r in rows])
    return off, np.concatenate(rows).astype(np.int32)


STRATEGY_TUNES = ["strat=hhhs", "strat=bbbb", "strat=mmmm", "strat=ssss", "strat=bmsm", "strat=mbhb"]


@pytest.mark.parametrize("strat", STRATEGY_TUNES)
@pytest.mark.parametrize("graph", ["banded", "hub", "rmat_s16", "zero_heavy", "ba_100000"])
def test_bucket_strategies(ctx, monkeypatch, strat, graph):
    """Every per-bucket strategy (h: LDS cuckoo hash, b: LDS bitmap over the owner's value
    range, m: the owner's sorted list in LDS merged with each partner list, s: binary search in
    HBM), forced for all buckets and mixed: counts and per-edge counts against the oracle, with
    long partner spans (long=1, loads past a span's end read 0) and without."""
    if graph == "banded":
        off, adj = _banded_csr()
    elif graph == "zero_heavy":
        off, adj = _zero_heavy_csr()
    else:
        off, adj, _ = oracle_csr(graph)
    t = orc.tc_count_intersect(off, adj)
    exp = orc.tc_per_edge(off, adj, 0)
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    for long_min in (1, 128):
        monkeypatch.setenv("RTG_TC_TUNE", f"threads=256,unroll=4,item=65536,wave_item=4096,long={long_min},grab=16,{strat}")
        p = g.tc_plan(method=rtg.TC_BATCHED)
        d = p.describe()
        assert d["bucket_strategy"].count("/") == 3 and d["strat"] == strat.split(",")[0].split("=")[1]
        if graph == "banded" and strat in ("strat=bbbb", "strat=mmmm"):
            assert d["all_strategies"] == "1", d  # the LDS bitmap / sorted list is really used
        assert p.run()[0] == t, (long_min, d)
        p.close()
        p = g.tc_plan(method=rtg.TC_BATCHED, use_hip_graph=False, per_edge=True)
        out, cnt, _ = p.run_per_edge()
        assert cnt == t and np.array_equal(out.numpy()[: adj.size], exp), (long_min, strat)
        out.free()
        p.close()


def test_autotuned_bucket_strategies(ctx, tmp_path):
    """The autotuner picks a strategy per owner bucket and the cache records it; a plan from the
    cache uses the same strategies and gives the oracle's count."""
    off, adj = _banded_csr()
    t = orc.tc_count_intersect(off, adj)
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    cache = tmp_path / "tune.json"
    p1 = g.tc_plan(method=rtg.TC_BATCHED, autotune=True, tune_cache=str(cache))
    d1 = p1.describe()
    assert d1["source"] == "autotune" and len(d1["strat"]) == 4
    assert p1.run()[0] == t
    assert "strat=" in cache.read_text()
    p2 = g.tc_plan(method=rtg.TC_BATCHED, autotune=True, tune_cache=str(cache))
    d2 = p2.describe()
    assert d2["source"] == "cache" and d2["bucket_strategy"] == d1["bucket_strategy"]
    assert p2.run()[0] == t
    st = p2.plan_stats
    assert st["nbuckets"] == 4 and len(st["strategy"]) == 4


# On the copy a count runs up to three kernels: block bitmaps, small owners' wave items (hashed,
# also under a bitmap strategy that cannot fit a wave's share) and the general kernel for the
# rest; "bhbb" and "bmbs" put items in all three.
RENUMBER_TUNES = ["strat=hhhs", "strat=bbbs", "strat=hbbs", "strat=mmms", "strat=ssss", "strat=bhbb",
                  "strat=hbbs,unroll=8", "strat=bbbs,threads=512,unroll=2", "strat=hbbs,wave_max=256,unroll=8,grab=32",
                  "strat=bhbb,threads=512,unroll=8", "strat=bmbs,wave_max=16"]


@pytest.mark.parametrize("tune", RENUMBER_TUNES)
def test_renumbered_counts(ctx, monkeypatch, tune):
    """Batched counting plans on the graph's degree-ordered copy (RTG_TC_RENUMBER=1 forces it
    below the default size): every fixture graph, every strategy mix, long spans of 1 and 128
    elements (16-bit and 32-bit spans, loads past the end), against the oracle's count."""
    monkeypatch.setenv("RTG_TC_RENUMBER", "1")
    for name in GRAPHS:
        _, _, t = oracle_csr(name)
        g = rtg.DeviceGraph.from_pairs(ctx, GRAPHS[name])
        if g.m == 0:
            continue
        for long_min in (1, 128):
            monkeypatch.setenv("RTG_TC_TUNE", f"threads=256,unroll=4,item=65536,wave_item=1024,long={long_min},grab=16,{tune}")
            p = g.tc_plan(method=rtg.TC_BATCHED)
            d = p.describe()
            assert d["renumbered"] == "1", (name, d)
            assert p.run()[0] == t, (name, tune, long_min)
            p.close()
        g.close()


def test_renumbering_default_and_opt_out(ctx, monkeypatch):
    """Graphs below 2^20 edges keep their own CSR by default; RTG_TC_RENUMBER=0 keeps it for any
    graph; edge-range plans never renumber; the copy is built once per graph and shared by its
    later plans (a per-edge plan adds the copy's edge map)."""
    off, adj, t = oracle_csr("ba_100000")  # about 500k oriented edges
    assert adj.size < (1 << 20)
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    d0 = g.tc_plan(method=rtg.TC_BATCHED).describe()
    assert d0["renumbered"] == "0" and int(d0["streamed16"]) == 0
    monkeypatch.setenv("RTG_TC_RENUMBER", "1")
    p1 = g.tc_plan(method=rtg.TC_BATCHED)
    p2 = g.tc_plan(method=rtg.TC_BATCHED)
    d1, d2 = p1.describe(), p2.describe()
    assert d1["renumbered"] == "1" and d2["renumbered"] == "1"
    assert float(d1["renumber_ms"]) > 0 and float(d2["renumber_ms"]) == 0  # built once
    assert int(d1["streamed16"]) > 0  # rows of the 65,536 highest-degree vertices' neighbours fit 16 bits
    assert int(d1["streamed16"]) <= int(d1["streamed"])
    # without cuts both CSRs stream the same elements (the shorter list of each edge); no partner
    # list here is longer than 16 elements, the shortest the planner cuts
    uncut = int(np.minimum(np.diff(off)[np.repeat(np.arange(off.size - 1), np.diff(off))], np.diff(off)[adj]).sum())
    assert int(d0["streamed"]) == uncut and int(d1["streamed"]) == uncut
    assert p1.run()[0] == t and p2.run()[0] == t
    pe = g.tc_plan(method=rtg.TC_BATCHED, per_edge=True)
    dpe = pe.describe()
    assert dpe["renumbered"] == "1" and float(dpe["renumber_ms"]) > 0  # the copy's edge map, built now
    half = g.tc_plan(method=rtg.TC_BATCHED, e_begin=0, e_end=g.m // 2)
    assert half.describe()["renumbered"] == "0"
    monkeypatch.setenv("RTG_TC_RENUMBER", "0")
    assert g.tc_plan(method=rtg.TC_BATCHED).describe()["renumbered"] == "0"


def _degree_ordered(off, adj):
    """The degree-ordered copy by its definition: vertices renamed by descending out + in degree
    (ties by id), each row's entries renamed and sorted; inv[i] = the old id of new vertex i."""
    n = off.size - 1
    dplus = np.diff(off).astype(np.int64)
    row = np.repeat(np.arange(n), dplus)
    tot = dplus + np.bincount(adj, minlength=n)
    inv = np.lexsort((np.arange(n), -tot))
    rank = np.empty(n, np.int64)
    rank[inv] = np.arange(n)
    nrow, nadj = rank[row], rank[adj]
    o = np.lexsort((nadj, nrow))
    noff = np.zeros(n + 1, np.int64)
    np.add.at(noff, nrow + 1, 1)
    return np.cumsum(noff), nadj[o], inv


@pytest.mark.parametrize("name", ["hand", "k40", "wheel", "gnm_dups", "ba_100000", "rmat_s12", "rmat_s16", "plcm",
                                  "hub", "gnp_dense"])
def test_ordered_csr_matches_definition(ctx, monkeypatch, name):
    """The degree-ordered copy, built by sorting every row where its new row starts (rows of up
    to 64 entries across one wave's lanes, up to 512 in a wave's LDS buffer, longer ones in a
    block's: "hub" has rows of 1,199 and 2,500 entries) and by the key sort over all edges
    (RTG_RENUMBER_KEY_SORT=1, the path of graphs with rows beyond 4,096), against its
    definition restated in numpy."""
    off, adj, _ = oracle_csr(name)
    eoff, eadj, einv = _degree_ordered(off, adj)
    for key_sort in ("0", "1"):
        monkeypatch.setenv("RTG_RENUMBER_KEY_SORT", key_sort)
        g = rtg.DeviceGraph.from_csr(ctx, off, adj)
        o, a, inv = g.ordered_csr()
        assert np.array_equal(o, eoff), (name, key_sort)
        assert np.array_equal(a, eadj), (name, key_sort)
        assert np.array_equal(inv, einv), (name, key_sort)
        g.close()


def test_span_cuts_streamed(ctx, monkeypatch):
    """The planner's span cuts on the degree-ordered copy, by a plain scan: every edge with both
    out-lists non-empty streams its partner's row (the shorter list; ties to the smaller id's
    row being the owner's), and a row longer than 16 entries whose last entry exceeds the
    owner's largest entry hi ends before its first entry at a multiple-of-16 position above hi.
    On a dense graph the cuts land at every offset of a row; the streamed total must match
    exactly, and the count the oracle's."""
    monkeypatch.setenv("RTG_TC_RENUMBER", "1")
    off, adj, t = oracle_csr("gnp_dense")
    noff, nadj, _ = _degree_ordered(off, adj)
    n = noff.size - 1
    nd = np.diff(noff)
    vmax = np.full(n, -1, np.int64)
    vmax[nd > 0] = nadj[noff[1:][nd > 0] - 1]
    nrow = np.repeat(np.arange(n), nd)
    du, dv = nd[nrow], nd[nadj]
    work = np.minimum(du, dv) > 0
    own_u = (du > dv) | ((du == dv) & (nrow < nadj))
    owner = np.where(own_u, nrow, nadj)[work]
    partner = np.where(own_u, nadj, nrow)[work]
    p0, p1, hi = noff[partner], noff[partner + 1], vmax[owner]
    cut = p1.copy()
    need = np.nonzero((p1 - p0 > 16) & (vmax[partner] > hi))[0]
    assert need.size > 1000
    for k in need:
        pos = (p0[k] + 15) // 16 * 16
        while pos < p1[k] and nadj[pos] <= hi[k]:
            pos += 16
        if pos < p1[k]:
            cut[k] = pos
    expect = int((cut - p0).sum())
    assert expect < int((p1 - p0).sum())
    g = rtg.DeviceGraph.from_csr(ctx, off, adj)
    p = g.tc_plan(method=rtg.TC_BATCHED)
    d = p.describe()
    assert d["renumbered"] == "1"
    assert int(d["streamed"]) == expect
    assert int(d["streamed16"]) == expect  # 2,000 vertices: every row fits 16 bits
    assert p.run()[0] == t
    p.close()
    g.close()


@pytest.mark.parametrize("tune", ["strat=hbbs", "strat=hhhs", "strat=bmbs", "strat=hbbs,unroll=4,long=1",
                                  "strat=bhbb,wave_max=256,unroll=8,grab=32"])
def test_renumbered_per_edge(ctx, monkeypatch, tune):
    """Per-edge plans on the degree-ordered copy: every span's count is written at the graph's
    own edge through the copy's edge map; every fixture graph against the oracle's per-edge
    counts, twice (the output is cleared between runs)."""
    monkeypatch.setenv("RTG_TC_RENUMBER", "1")
    monkeypatch.setenv("RTG_TC_TUNE", f"threads=256,unroll=8,item=65536,wave_item=1024,long=128,grab=16,{tune}")
    for name in GRAPHS:
        off, adj, t = oracle_csr(name)
        if adj.size == 0:
            continue
        g = rtg.DeviceGraph.from_csr(ctx, off, adj)
        exp = orc.tc_per_edge(off, adj, 0)
        plan = g.tc_plan(method=rtg.TC_BATCHED, use_hip_graph=False, per_edge=True)
        assert plan.describe()["renumbered"] == "1"
        for _ in range(2):
            out, cnt, _ = plan.run_per_edge()
            assert cnt == t and np.array_equal(out.numpy()[:adj.size], exp), (name, tune)
        assert plan.run()[0] == t
        plan.close()
        g.close()
