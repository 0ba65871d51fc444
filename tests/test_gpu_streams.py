"""The two-stream schedule of compact path-tracing rounds: the shadow rays of
iteration i traced on a second stream beside iteration i + 1's closest-hit
traversal.  Each path adds its radiance terms in the reference's order
whatever the overlap, so films are bit-identical with the overlap on and off,
and so are the exact work counters of a counting render.  Batches run their
large rounds over one scene with the same schedule.  The same holds for the
order in which shading writes a workgroup's outputs (grouped by ray-direction
octant or in thread order)."""
import numpy as np
import pytest

from barnacle_amd import configs, scene

pytestmark = pytest.mark.gpu


def with_overlap(ctx, on, fn):
    ctx.set_option("overlap_shadow", 1 if on else 0)
    try:
        return fn()
    finally:
        ctx.set_option("overlap_shadow", 1)


@pytest.mark.parametrize("name", ["config0", "config1", "spheres", "batch7"])
def test_overlap_keeps_films_and_counters(ctx, name):
    info = {"config0": lambda: configs.config0(),
            "config1": lambda: configs.config1(width=160, height=120, spp=8, freq=24),
            "spheres": lambda: configs.sphere_scene(96, 64, 8, 8),
            "batch7": lambda: configs.batch_scene(7, max_res=128, max_spp=8)}[name]()
    sc = scene.load_scene(info)
    ds = ctx.scene(sc)
    p = sc.pt_params()
    print(f"overlap on/off: {name}", flush=True)
    a, sa = with_overlap(ctx, False, lambda: ds.render_pt(p, stats=True))
    b, sb = with_overlap(ctx, True, lambda: ds.render_pt(p, stats=True))
    c = with_overlap(ctx, True, lambda: ds.render_pt(p))
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    assert np.array_equal(a.view(np.int32), c.view(np.int32))
    assert sa == sb
    assert sa["rays_shadow"] > 0


def test_overlap_over_several_rounds(ctx):
    """A frame split into rounds (window smaller than the frame's paths):
    every round joins its streams before the next begins."""
    sc = scene.load_scene(configs.config1(width=128, height=128, spp=6, freq=16))
    ds = ctx.scene(sc)
    p = sc.pt_params()
    ref = with_overlap(ctx, False, lambda: ds.render_pt(p))
    ctx.set_option("window_paths", 40000)
    try:
        got = with_overlap(ctx, True, lambda: ds.render_pt(p))
    finally:
        ctx.set_option("window_paths", 0)
    assert np.array_equal(ref.view(np.int32), got.view(np.int32))


def test_batch_round_over_one_scene_uses_two_streams(ctx):
    """An item of 2^20 paths or more gets a round over its own scene, run
    eagerly with the two-stream schedule; the small item of another scene
    gets a graph-replayed round.  Both equal their plain renders."""
    big_sc = scene.load_scene(configs.config1(width=256, height=256, spp=16, freq=16))
    small_sc = scene.load_scene(configs.config0())
    big = (ctx.scene(big_sc), big_sc.pt_params())
    small = (ctx.scene(small_sc), small_sc.pt_params(width=32, height=32, spp=2))
    assert big[1].width * big[1].height * big[1].spp >= 1 << 20
    films, st = ctx.render_batch([big, small])
    assert st["scene_rounds"] >= 1
    for (ds, p), f in zip([big, small], films):
        assert np.array_equal(f.view(np.int32), with_overlap(ctx, False, lambda: ds.render_pt(p)).view(np.int32))


@pytest.mark.parametrize("name", ["config0", "config1", "spheres", "batch7"])
def test_octant_order_keeps_films_and_counters(ctx, name):
    """Shading may write a workgroup's continuing paths and shadow rays in
    any order (here grouped by ray-direction octant): every path adds its
    terms in its own order, so films and exact counters do not change."""
    info = {"config0": lambda: configs.config0(),
            "config1": lambda: configs.config1(width=160, height=120, spp=8, freq=24),
            "spheres": lambda: configs.sphere_scene(96, 64, 8, 8),
            "batch7": lambda: configs.batch_scene(7, max_res=128, max_spp=8)}[name]()
    sc = scene.load_scene(info)
    ds = ctx.scene(sc)
    p = sc.pt_params()
    print(f"octant order on/off: {name}", flush=True)
    try:
        ctx.set_option("octant_order", 0)
        a, sa = ds.render_pt(p, stats=True)
        ctx.set_option("octant_order", 1)
        b, sb = ds.render_pt(p, stats=True)
        c = with_overlap(ctx, False, lambda: ds.render_pt(p))
    finally:
        ctx.set_option("octant_order", 1)
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    assert np.array_equal(a.view(np.int32), c.view(np.int32))
    assert sa == sb


def test_octant_order_in_batches(ctx):
    """Batched rounds (graph-replayed, mixed scenes) with and without the
    octant grouping give the plain renders' films."""
    infos = [configs.batch_scene(s, max_res=64, max_spp=4, tri_range=(300, 20000)) for s in (3, 11, 19)]
    items = []
    for info in infos:
        sc = scene.load_scene(info)
        items.append((ctx.scene(sc), sc.pt_params()))
    try:
        ctx.set_option("octant_order", 0)
        ref = [ds.render_pt(p) for ds, p in items]
        got0, _ = ctx.render_batch(items)
        ctx.set_option("octant_order", 1)
        got1, _ = ctx.render_batch(items)
    finally:
        ctx.set_option("octant_order", 1)
    for r, a, b in zip(ref, got0, got1):
        assert np.array_equal(r.view(np.int32), a.view(np.int32))
        assert np.array_equal(r.view(np.int32), b.view(np.int32))


@pytest.mark.parametrize("name", ["config1", "spheres", "batch7"])
def test_dark_shadow_rays_skipped_keep_films(ctx, name):
    """A shadow ray whose NEE term is exactly zero (an operand zero, the other
    finite) adds nothing, so not casting it keeps the film bit for bit.  Mode
    2 counts the rays traced: the same extension rays, fewer or equal shadow
    rays than the reference casts (mode 1's counting frame)."""
    info = {"config1": lambda: configs.config1(width=160, height=120, spp=8, freq=24),
            "spheres": lambda: configs.sphere_scene(96, 64, 8, 8),
            "batch7": lambda: configs.batch_scene(7, max_res=128, max_spp=8)}[name]()
    sc = scene.load_scene(info)
    ds = ctx.scene(sc)
    p = sc.pt_params()
    print(f"skip_dark modes: {name}", flush=True)
    try:
        ctx.set_option("skip_dark", 0)
        a = ds.render_pt(p)
        ctx.set_option("skip_dark", 1)
        b, s_ref = ds.render_pt(p, stats=True)
        c = ds.render_pt(p)
        ctx.set_option("skip_dark", 2)
        d, s_tr = ds.render_pt(p, stats=True)
    finally:
        ctx.set_option("skip_dark", 1)
    for f in (b, c, d):
        assert np.array_equal(a.view(np.int32), f.view(np.int32))
    assert s_tr["rays_closest"] == s_ref["rays_closest"]
    assert s_tr["nodes_closest"] == s_ref["nodes_closest"]
    assert 0 < s_tr["rays_shadow"] <= s_ref["rays_shadow"]


def test_frame_left_on_device(ctx):
    """bn_render_pt with a NULL film renders into the context's device film
    without a host copy; the frames before and after are unaffected."""
    sc = scene.load_scene(configs.config1(width=96, height=64, spp=4, freq=16))
    ds = ctx.scene(sc)
    p = sc.pt_params()
    ref = ds.render_pt(p)
    assert ds.render_pt(p, host=False) is None
    _, st = ds.render_pt(p, stats=True, host=False)
    assert st["rays_closest"] > 0
    assert np.array_equal(ref.view(np.int32), ds.render_pt(p).view(np.int32))
