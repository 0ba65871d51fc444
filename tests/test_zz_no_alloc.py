"""bn_first_hit_noalloc (static scratch, no device allocation) against
bn_first_hit, through the C ABI: bit-identical instance, primitive and t, on
two calls in a row (the second smaller, so it runs over the first call's
leftover scratch), and a clear capacity error beyond 2^20 pixels."""
import numpy as np
import pytest

from barnacle_amd import configs, native, scene

pytestmark = pytest.mark.gpu


def _same(a, b):
    ai, ap, at = a
    bi, bp, bt = b
    assert np.array_equal(ai, bi)
    assert np.array_equal(ap, bp)
    assert np.array_equal(at.view(np.int32), bt.view(np.int32))


def test_noalloc_matches_first_hit_twice():
    ctx = native.Context(0)
    try:
        sc = scene.load_scene(configs.config1(width=64, height=64, spp=2))
        ds = ctx.scene(sc)
        big = sc.pt_params(width=512, height=384)
        small = sc.pt_params(width=97, height=61)
        ref_big = ds.first_hit(big, 1)
        ref_small = ds.first_hit(small, 0)
        got_big = ds.first_hit_noalloc(big, 1)
        got_small = ds.first_hit_noalloc(small, 0)
        print(f"noalloc: 512x384 hits {int(np.sum(got_big[0] >= 0))}, 97x61 hits {int(np.sum(got_small[0] >= 0))}",
              flush=True)
        _same(got_big, ref_big)
        _same(got_small, ref_small)
        assert np.sum(ref_big[0] >= 0) > 0.5 * ref_big[0].size
    finally:
        ctx.close()


def test_noalloc_second_context_and_cornell():
    ctx = native.Context(0)
    try:
        sc = scene.load_scene(configs.config0())
        ds = ctx.scene(sc)
        p = sc.pt_params()
        _same(ds.first_hit_noalloc(p, 3), ds.first_hit(p, 3))
        p1 = sc.pt_params(width=1, height=1)
        _same(ds.first_hit_noalloc(p1, 0), ds.first_hit(p1, 0))
    finally:
        ctx.close()


def test_noalloc_rejects_films_beyond_its_scratch():
    ctx = native.Context(0)
    try:
        sc = scene.load_scene(configs.config0())
        ds = ctx.scene(sc)
        p = sc.pt_params(width=1025, height=1024)
        with pytest.raises(native.BarnacleError) as e:
            ds.first_hit_noalloc(p, 0)
        assert e.value.code == -5 and "2^20" in str(e.value)
        # the largest film it takes still matches
        p = sc.pt_params(width=1024, height=1024)
        _same(ds.first_hit_noalloc(p, 0), ds.first_hit(p, 0))
    finally:
        ctx.close()
