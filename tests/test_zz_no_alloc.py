"""The allocation-free set-intersection entry point (rtg_si_pairs_noalloc) against the allocating
one (rtg_si_pairs) and the oracle: every fixed method, two calls in a row (the largest input the
scratch array takes, then a smaller one), the out-of-range flag cleared between calls, and the
limits it refuses."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_oracle, load_rtg

pytestmark = pytest.mark.gpu

rtg = load_rtg()
orc = load_oracle()
NOALLOC_METHODS = [rtg.SI_PLAIN, rtg.SI_BATCHED, rtg.SI_REFERENCE, rtg.SI_MERGE, rtg.SI_HASH, rtg.SI_BITMAP]


def workload():
    """Zipf(1.8) lists plus balanced long lists (so the merge, hash and bitmap units run), and
    random pairs over all of them."""
    lens = rtg.gen_zipf_lengths(1 << 14, 1.8, 65536, seed=5)
    offs, el = rtg.gen_sorted_lists(lens, 1 << 24, seed=5)
    rng = np.random.default_rng(21)
    extra = [np.sort(rng.choice(span, size=n, replace=False)).astype(np.uint32)
             for n, span in [(3000, 5000), (20000, 40000), (60000, 1 << 20), (5000, 9000)]]
    el = np.concatenate([el] + extra).astype(np.uint32)
    offs = np.concatenate([offs, offs[-1] + np.cumsum([x.size for x in extra])]).astype(np.uint32)
    return offs, el


@pytest.fixture(scope="module")
def data(ctx):
    offs, el = workload()
    nl = offs.size - 1
    rng = np.random.default_rng(8)
    big = rng.integers(0, nl, size=2 * rtg.SI_NOALLOC_MAX_PAIRS, dtype=np.int64).astype(np.uint32)
    big[: 2 * 64] = np.repeat(np.arange(nl - 4, nl, dtype=np.uint32), 32)  # balanced long pairs
    small = big[2 * 1000: 2 * 1300].copy()
    d = {"offs": offs, "el": el, "nl": nl, "big": big, "small": small,
         "d_off": ctx.to_device(offs), "d_el": ctx.to_device(el),
         "d_big": ctx.to_device(big), "d_small": ctx.to_device(small)}
    d["exp_small"] = orc.si_pairs(offs, el, small)
    return d


@pytest.mark.parametrize("method", NOALLOC_METHODS)
def test_noalloc_matches_allocating_entry_point(ctx, data, method):
    nb, ns = data["big"].size // 2, data["small"].size // 2
    assert nb == rtg.SI_NOALLOC_MAX_PAIRS
    print(f"[noalloc] method {method}: {nb} pairs, then {ns}", flush=True)
    ref_big, _ = rtg.si_pairs(ctx, data["d_off"], data["d_el"], data["nl"], data["d_big"], nb, method)
    ref_small, _ = rtg.si_pairs(ctx, data["d_off"], data["d_el"], data["nl"], data["d_small"], ns, method)
    out = rtg.DeviceArray(ctx, (nb,), np.uint32)
    rtg.si_pairs_noalloc(ctx, data["d_off"], data["d_el"], data["nl"], data["d_big"], nb, method, out)
    assert np.array_equal(out.numpy(), ref_big.numpy())
    out2 = rtg.DeviceArray(ctx, (ns,), np.uint32)
    rtg.si_pairs_noalloc(ctx, data["d_off"], data["d_el"], data["nl"], data["d_small"], ns, method, out2)
    got = out2.numpy()
    assert np.array_equal(got, ref_small.numpy())
    assert np.array_equal(got, data["exp_small"])


def test_noalloc_flag_cleared_between_calls(ctx, data):
    """A call with an out-of-range id fails with RTG_EINVAL (that pair counts 0); the next call
    on valid pairs succeeds, so the flag left in the scratch array was cleared."""
    bad = data["small"].copy()
    bad[0] = data["nl"]
    exp = data["exp_small"].copy()
    exp[0] = 0
    out = rtg.DeviceArray(ctx, (bad.size // 2,), np.uint32)
    with pytest.raises(rtg.RTGError) as e:
        rtg.si_pairs_noalloc(ctx, data["d_off"], data["d_el"], data["nl"], ctx.to_device(bad), bad.size // 2,
                             rtg.SI_BATCHED, out)
    assert e.value.code == 1
    assert np.array_equal(out.numpy(), exp)
    rtg.si_pairs_noalloc(ctx, data["d_off"], data["d_el"], data["nl"], data["d_small"], bad.size // 2,
                         rtg.SI_BATCHED, out)
    assert np.array_equal(out.numpy(), data["exp_small"])


def test_noalloc_refuses_beyond_limits(ctx, data):
    out = rtg.DeviceArray(ctx, (16,), np.uint32)
    with pytest.raises(rtg.RTGError) as e:  # refused on the host, before any launch
        rtg.si_pairs_noalloc(ctx, data["d_off"], data["d_el"], data["nl"], data["d_big"],
                             rtg.SI_NOALLOC_MAX_PAIRS + 1, rtg.SI_BATCHED, out)
    assert e.value.code == 1 and "RTG_SI_NOALLOC_MAX_PAIRS" in str(e.value)
    with pytest.raises(rtg.RTGError) as e:
        rtg.si_pairs_noalloc(ctx, data["d_off"], data["d_el"], data["nl"], data["d_small"], 4, rtg.SI_AUTO, out)
    assert e.value.code == 1
