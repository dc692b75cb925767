"""The narrowphase entry point that allocates nothing
(b2g_collide_pairs_noalloc) against b2g_collide_pairs, through the C ABI.

Its staging lives in one device array that keeps the last call's contents,
so two calls run back to back, the second on fewer pairs than the first:
each result must equal b2g_collide_pairs' on the same pairs bit for bit (and
so the reference's golden manifolds). Sizes beyond the array are refused.
"""
import numpy as np
import pytest

from util import golden, load_pkg, manifold_fields_equal

pytestmark = pytest.mark.gpu


def _pairs(g, sl):
    return g["shape_a"][sl], g["xf_a"][sl], g["shape_b"][sl], g["xf_b"][sl]


def test_gpu_collide_pairs_noalloc_matches_collide_pairs_on_two_calls(engine):
    g = golden("narrowphase.npz")
    n = len(g["shape_a"])
    assert n <= engine.NOALLOC_MAX_PAIRS
    # the second call is smaller and covers other pairs, so stale rows of the
    # first call would show if the entry point read past its own n
    for label, sl in (("all pairs", slice(0, n)), ("smaller second call", slice(n // 2 + 7, n // 2 + 7 + 301))):
        args = _pairs(g, sl)
        want = engine.collide_pairs(*args)
        got = engine.collide_pairs(*args, noalloc=True)
        print(f"{label}: {len(got)} pairs, touching {(got['point_count'] > 0).sum()}", flush=True)
        assert len(got) == len(want)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), label
        assert manifold_fields_equal(g["manifold"][sl], got).all(), label


def test_gpu_collide_pairs_noalloc_refuses_more_than_its_maximum(engine):
    g = golden("narrowphase.npz")
    m = engine.NOALLOC_MAX_PAIRS + 1
    idx = np.arange(m) % len(g["shape_a"])
    with pytest.raises(engine.B2GError, match="at most"):
        engine.collide_pairs(g["shape_a"][idx], g["xf_a"][idx], g["shape_b"][idx], g["xf_b"][idx], noalloc=True)
    # and the maximum itself runs
    idx = idx[:-1]
    got = engine.collide_pairs(g["shape_a"][idx], g["xf_a"][idx], g["shape_b"][idx], g["xf_b"][idx], noalloc=True)
    want = engine.collide_pairs(g["shape_a"][idx], g["xf_a"][idx], g["shape_b"][idx], g["xf_b"][idx])
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
