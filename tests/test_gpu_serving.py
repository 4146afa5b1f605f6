"""GPU query server: every pair and lookup result equals the oracle and equals the result of the
same request submitted alone; the replayed Poisson stream's answers match the oracle; the
device generators reproduce the host generators."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_oracle, load_rtg

pytestmark = pytest.mark.gpu

rtg = load_rtg()
orc = load_oracle()


@pytest.fixture(scope="module")
def world(ctx):
    deg = rtg.gen_powerlaw_degrees(20_000, 2.1, 2_000.0, 20.0, seed=1)
    pairs = rtg.gen_config_model(deg, 0.3, seed=1)
    g = rtg.DeviceGraph.from_pairs(ctx, pairs, relabel=False, orient=False)
    off, adj = g.download()
    arr = rtg.gen_sorted_array(1 << 20, seed=2)
    d_arr = ctx.to_device(arr)
    return g, off, adj, arr, d_arr


def oracle_pairs(off, adj, uv):
    return orc.si_pairs(off.astype(np.uint32), adj.astype(np.uint32), np.asarray(uv, np.uint32).reshape(-1))


@pytest.mark.parametrize("staging", [0, 1], ids=["zero_copy", "copy_staging"])
def test_mix_equals_alone_and_oracle(ctx, world, staging):
    g, off, adj, arr, d_arr = world
    rng = np.random.default_rng(0)
    n = g.n
    deg = np.diff(off)
    hubs = np.argsort(deg)[-50:]
    uv = [(int(rng.integers(n)), int(rng.integers(n))) for _ in range(300)]
    uv += [(int(hubs[i % 50]), int(hubs[(i * 7 + 3) % 50])) for i in range(100)]  # long lists
    lookups = [np.concatenate([rng.integers(0, 1 << 32, size=int(k), dtype=np.uint64).astype(np.uint32),
                               arr[rng.integers(0, arr.size, size=int(k))]]) for k in rng.integers(1, 2048, size=60)]
    exp_pairs = oracle_pairs(off, adj, uv)
    exp_lk = [orc.lookup(arr, k) for k in lookups]
    srv = rtg.Server(ctx, g, d_arr, arr.size, slots=4, copy_staging=staging)
    # the mix: everything in flight together, interleaved
    pid = [srv.submit_pair(u, v) for (u, v) in uv]
    lid = [srv.submit_lookup(k) for k in lookups]
    got_pairs = [srv.wait(r)[1] for r in pid]
    for (r, out), e in zip(lid, exp_lk):
        assert srv.wait(r)[0] == rtg.REQ_DONE
        assert np.array_equal(out, e)
    assert np.array_equal(np.array(got_pairs, np.uint32), exp_pairs)
    st = srv.stats()
    assert st["completed"] == len(uv) + len(lookups) and st["failed"] == 0
    # alone: one request at a time
    for i in range(0, len(uv), 37):
        st_, pr, _ = srv.wait(srv.submit_pair(*uv[i]))
        assert st_ == rtg.REQ_DONE and pr == got_pairs[i]
    for i in range(0, len(lookups), 11):
        r, out = srv.submit_lookup(lookups[i])
        srv.wait(r)
        assert np.array_equal(out, exp_lk[i])
    srv.close()


def test_small_batches_split_on_probe_items(ctx, world):
    """With 2 pairs per batch the item capacity (16 per pair plus the longest row) splits hub
    pairs into many batches; results still equal the oracle."""
    g, off, adj, arr, d_arr = world
    deg = np.diff(off)
    hubs = np.argsort(deg)[-40:]
    uv = [(int(hubs[i]), int(hubs[(i * 3 + 1) % 40])) for i in range(40)] + [(0, 0), (1, 1)]
    exp = oracle_pairs(off, adj, uv)
    srv = rtg.Server(ctx, g, d_arr, arr.size, slots=2, max_batch_pairs=2)
    ids = [srv.submit_pair(u, v) for (u, v) in uv]
    got = np.array([srv.wait(r)[1] for r in ids], np.uint32)
    st = srv.stats()
    srv.close()
    assert np.array_equal(got, exp)
    assert st["batches"] >= len(uv) // 2 and st["max_batch_requests"] <= 2


def test_replay_stream_matches_oracle(ctx, world):
    g, off, adj, arr, d_arr = world
    srv = rtg.Server(ctx, g, d_arr, arr.size, slots=4)
    nq = 20_000
    st, pr, ls = srv.replay(rate_qps=200_000, nqueries=nq, warmup=0, pair_frac=0.7, max_keys=512, seed=5,
                            want_results=True)
    srv.close()
    assert st["completed"] == nq and st["rejected"] == 0 and st["failed"] == 0
    q = rtg.stream_queries(g, 200_000, nq, pair_frac=0.7, max_keys=512, seed=5)
    isp = q["kind"] == rtg.REQ_PAIR
    exp = oracle_pairs(off, adj, np.stack([q["u"][isp], q["v"][isp]], 1))
    assert np.array_equal(pr[isp], exp)
    for i in np.flatnonzero(~isp)[:400]:
        keys = q["pool"][q["key_off"][i]: q["key_off"][i] + q["nkeys"][i]]
        assert int(ls[i]) == int((orc.lookup(arr, keys).astype(np.int64) + 2).sum())


def test_device_generators_match_host(ctx):
    host = rtg.gen_sorted_array(1 << 18, seed=4)
    dev = rtg.gen_sorted_array_device(ctx, 1 << 18, seed=4)
    assert np.array_equal(dev.numpy(), host)
    deg = rtg.gen_powerlaw_degrees(30_000, 2.1, 3_000.0, 15.0, seed=6)
    h = rtg.gen_config_model(deg, 0.3, seed=6)
    d = rtg.gen_config_model_device(ctx, deg, 0.3, seed=6)
    assert d.shape[0] == h.size
    assert np.array_equal(d.numpy(), h)


def test_replay_full_size_matches_oracle(ctx):
    """The serving workload at full size: the power-law configuration-model graph (n = 3,000,000,
    exponent 2.1, degrees up to 30,000, triangle closing 0.3, seed 1) as a symmetric CSR and 2^28
    sorted uint32 values; 2^20 Poisson queries at 2e6 q/s (70% degree-biased pairs, 30% lookups of
    U{1..4096} keys) replayed through the server. Every pair answer is checked against the oracle
    (hub pairs with lists tens of thousands long included), and 3,000 lookup batches."""
    print("[serving full size] building the 3M-node graph and the 2^28 array", flush=True)
    deg = rtg.gen_powerlaw_degrees(3_000_000, 2.1, 30_000.0, 50.0, seed=1)
    d_pairs = rtg.gen_config_model_device(ctx, deg, 0.3, seed=1)
    g = rtg.DeviceGraph.from_pairs(ctx, d_pairs, relabel=False, orient=False)
    d_pairs.free()
    d_arr = rtg.gen_sorted_array_device(ctx, 1 << 28, seed=1)
    nq = 1 << 20
    print("[serving full size] replaying 2^20 queries", flush=True)
    srv = rtg.Server(ctx, g, d_arr, 1 << 28, slots=2)
    st, pr, ls = srv.replay(rate_qps=2.0e6, nqueries=nq, warmup=0, pair_frac=0.7, max_keys=4096, seed=1,
                            want_results=True)
    srv.close()
    assert st["completed"] == nq and st["rejected"] == 0 and st["failed"] == 0
    print("[serving full size] checking against the oracle", flush=True)
    q = rtg.stream_queries(g, 2.0e6, nq, pair_frac=0.7, max_keys=4096, seed=1)
    off, adj = g.download()
    isp = q["kind"] == rtg.REQ_PAIR
    uv = np.stack([q["u"][isp], q["v"][isp]], 1)
    dmin = np.minimum(np.diff(off)[uv[:, 0]], np.diff(off)[uv[:, 1]])
    assert dmin.max() > 10_000  # hub pairs are part of the checked stream
    exp = oracle_pairs(off, adj, uv)
    assert np.array_equal(pr[isp], exp)
    del off, adj
    arr = d_arr.numpy()
    rng = np.random.default_rng(3)
    lk = np.flatnonzero(~isp)
    for i in rng.choice(lk, size=3000, replace=False):
        keys = q["pool"][q["key_off"][i]: q["key_off"][i] + q["nkeys"][i]]
        assert int(ls[i]) == int((orc.lookup(arr, keys).astype(np.int64) + 2).sum())
    g.close()
    d_arr.free()
