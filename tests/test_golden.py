"""Golden vectors (CPU only): the seeded generators reproduce their recorded SHA-256, and the
oracle reproduces the recorded CSR hashes and triangle counts, which were cross-checked
against scipy and networkx when the fixtures were written (tests/golden/make_golden.py)."""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import load_oracle, load_rtg

GOLD = json.loads((Path(__file__).parent / "golden" / "golden.json").read_text())
rtg = load_rtg()
orc = load_oracle()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def check_graph(entry, pairs):
    assert pairs.size // 2 == entry["pairs"]
    assert sha(pairs.astype("<i4")) == entry["pairs_sha256"]
    off, adj = orc.csr_from_pairs(pairs)
    assert (off.size - 1, adj.size) == (entry["n"], entry["m"])
    assert sha(off.astype("<i4")) == entry["offsets_sha256"]
    assert sha(adj.astype("<i4")) == entry["adjs_sha256"]
    t = orc.tc_count_intersect(off, adj)
    assert t == entry["triangles"] == entry.get("triangles_scipy", entry.get("triangles_1a2"))
    if "triangles_networkx" in entry:
        assert t == entry["triangles_networkx"]
    assert orc.tc_edge_model(off, adj)["alg_bytes"] == entry["alg_bytes"]


def test_ba_100000():
    check_graph(GOLD["ba_n100000_m5_seed1"], rtg.gen_ba(100_000, 5, 1))


def test_rmat_small():
    check_graph(GOLD["rmat_s12_ef16_seed1"], rtg.gen_rmat(12, 16 << 12, seed=1))


def test_rmat_s16():
    check_graph(GOLD["rmat_s16_ef44_seed1"], rtg.gen_rmat(16, 44 << 16, seed=1))


def test_plcm_small():
    e = GOLD["plcm_n20000_seed1"]
    deg = rtg.gen_powerlaw_degrees(20_000, 2.1, 2_000.0, 20.0, seed=1)
    assert sha(deg.astype("<i4")) == e["degrees_sha256"]
    check_graph(e, rtg.gen_config_model(deg, 0.3, seed=1))


def test_zipf_lists():
    e = GOLD["zipf_lists_2e16_seed1"]
    lens = rtg.gen_zipf_lengths(1 << 16, 1.8, 65536, seed=1)
    assert sha(lens.astype("<u4")) == e["lengths_sha256"]
    offs, el = rtg.gen_sorted_lists(lens, 1 << 24, seed=1)
    assert sha(el.astype("<u4")) == e["elements_sha256"]
    # every list sorted, unique, in range
    d = np.diff(el.astype(np.int64))
    starts = offs[1:-1].astype(np.int64)
    mask = np.ones(d.size, bool)
    mask[starts[(starts > 0) & (starts <= d.size)] - 1] = False
    assert np.all(d[mask] > 0) and el.max() < (1 << 24)


def test_zipf_distribution():
    lens = rtg.gen_zipf_lengths(1 << 20, 1.8, 65536, seed=3)
    from scipy.special import zeta
    p1 = 1.0 / zeta(1.8)
    assert abs(np.mean(lens == 1) - p1) < 0.003
    assert lens.min() >= 1 and lens.max() <= 65536


def test_sorted_array():
    arr = rtg.gen_sorted_array(1 << 16, seed=1)
    assert sha(arr.astype("<u4")) == GOLD["sorted_array_2e16_seed1"]["sha256"]
    assert np.all(np.diff(arr.astype(np.int64)) >= 0)


def test_ba_structure():
    p = rtg.gen_ba(5000, 4, seed=9).reshape(-1, 2)
    assert p.shape[0] == 4 * (5000 - 4)
    assert np.all(p[:, 0] != p[:, 1])
    key = np.minimum(p[:, 0], p[:, 1]).astype(np.int64) * 5000 + np.maximum(p[:, 0], p[:, 1])
    assert np.unique(key).size == key.size
    assert not np.array_equal(p, rtg.gen_ba(5000, 4, seed=10).reshape(-1, 2))


def test_rmat_range_and_determinism():
    a = rtg.gen_rmat(10, 5000, seed=2)
    assert a.min() >= 0 and a.max() < 1024
    assert np.array_equal(a, rtg.gen_rmat(10, 5000, seed=2))
