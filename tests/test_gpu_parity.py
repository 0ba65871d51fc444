"""GPU (HIP, through the C ABI) against the CPU oracle on the same scenes.

Bars: first-hit (instance, primitive, t) bit-exact; films bit-exact except
for rare transcendental rounding differences, and within rel-RMSE 1e-3 in any
case (the tolerance stated for the float path).
"""
import numpy as np
import pytest

from barnacle_amd import configs, scene

pytestmark = pytest.mark.gpu


def rel_rmse(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.fixture(scope="module")
def c0():
    return scene.load_scene(configs.config0())


@pytest.fixture(scope="module")
def c1_small():
    return scene.load_scene(configs.config1(width=128, height=128, spp=2))


def test_first_hit_config0(ctx, orc, c0):
    p = c0.pt_params()
    gi, gp, gt = ctx.scene(c0).first_hit(p, 0)
    oi, op, ot = orc.scene(c0).first_hit(p, 0)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gp, op)
    assert np.array_equal(gt.view(np.int32), ot.view(np.int32))


def test_first_hit_config1(ctx, orc, c1_small):
    p = c1_small.pt_params(width=256, height=256)
    for k in (0, 1):
        gi, gp, gt = ctx.scene(c1_small).first_hit(p, k)
        oi, op, ot = orc.scene(c1_small).first_hit(p, k)
        assert np.array_equal(gi, oi)
        assert np.array_equal(gp, op)
        assert np.array_equal(gt.view(np.int32), ot.view(np.int32))


def test_pt_config0_film(ctx, orc, c0):
    p = c0.pt_params()
    g = ctx.scene(c0).render_pt(p)
    o, _ = orc.scene(c0).render(p)
    assert np.all(np.isfinite(g))
    assert rel_rmse(g, o) <= 1e-3
    assert np.mean(g == o) > 0.999


def test_pt_config1_film(ctx, orc, c1_small):
    p = c1_small.pt_params()
    g = ctx.scene(c1_small).render_pt(p)
    o, _ = orc.scene(c1_small).render(p)
    assert rel_rmse(g, o) <= 1e-3
    assert np.mean(g == o) > 0.999


def test_trace_counters_match_oracle(ctx, orc, c0):
    p = c0.pt_params(width=64, height=64, spp=2)
    _, gs = ctx.scene(c0).render_pt(p, stats=True)
    _, os_ = orc.scene(c0).render(p)
    for k in ("rays_closest", "rays_shadow", "nodes_closest", "tris_closest", "insts_closest",
              "nodes_shadow", "tris_shadow", "insts_shadow", "vertices", "paths"):
        assert gs[k] == os_[k], k


def test_normal_view_config1(ctx, orc, c1_small):
    """NormalIntegrator (Extensions/Integrator/Normal.fs:10-17): one bounce."""
    p = c1_small.pt_params(spp=3)
    g = ctx.scene(c1_small).render_normal(p)
    o, _ = orc.scene(c1_small).render(p, integrator=0)
    assert np.mean(g == o) > 0.999
    assert rel_rmse(g, o) <= 1e-5


@pytest.mark.parametrize("depth", [0, 1, 3])
def test_pt_shallow_depths(ctx, orc, c0, depth):
    """max-depth 0 renders black; 1 is emission only; paths of different
    depths share the slot pool."""
    p = c0.pt_params(width=40, height=36, spp=3, max_depth=depth)
    g = ctx.scene(c0).render_pt(p)
    o, _ = orc.scene(c0).render(p)
    if depth == 0:
        assert not np.any(g)
    assert np.mean(g == o) > 0.999
    assert rel_rmse(g, o) <= 1e-3 if np.any(o) else not np.any(g)


def nonfinite_scene(vertices):
    info = configs.cornell_box(48, 48, 2, 4)
    info["primitives"].append({"type": "mesh", "vertices": np.array(vertices, np.float32).reshape(-1),
                               "indices": np.arange(len(vertices))})
    info["instances"].append({"primitive": len(info["primitives"]) - 1, "material": 0})
    info["nodes"][0]["instances"].append(len(info["instances"]) - 1)
    return scene.load_scene(info)


def test_nonfinite_geometry_matches_oracle(ctx, orc):
    """A triangle with an infinite vertex: its bounds, and the TLAS bounds
    above it, are infinite, so the scene traverses with the exact
    NaN-propagating slab test (Ray.fs:26-39) everywhere, and about half of
    the camera rays miss, as in the reference.  First hits, counters and film
    as the oracle's."""
    sc = nonfinite_scene([[100, 100, -300], [200, 100, -300], [np.inf, 200, -300]])
    p = sc.pt_params()
    ds, os_ = ctx.scene(sc), orc.scene(sc)
    gi, gp, gt = ds.first_hit(p, 0)
    oi, op, ot = os_.first_hit(p, 0)
    assert np.array_equal(gi, oi) and np.array_equal(gp, op)
    assert np.array_equal(gt.view(np.int32), ot.view(np.int32))
    assert (oi < 0).sum() > 500 and (oi >= 0).sum() > 500
    g, gst = ds.render_pt(p, stats=True)
    o, ost = os_.render(p)
    for k in ("nodes_closest", "tris_closest", "insts_closest", "nodes_shadow", "tris_shadow"):
        assert gst[k] == ost[k], k
    assert np.array_equal(np.isfinite(g), np.isfinite(o))
    ok = np.all(np.isfinite(o), axis=1)
    assert rel_rmse(g[ok], o[ok]) <= 1e-3


def test_nan_bounds_fail_the_build_like_the_reference(ctx, orc):
    """A NaN vertex makes a TLAS bin index negative (Util/BVH.fs:163-174),
    where the reference's Span indexer throws: the scene build fails with an
    error on both sides, and the context stays usable."""
    from barnacle_amd import native
    sc = nonfinite_scene([[300, 300, -250], [400, 300, -250], [350, np.nan, -250]])
    with pytest.raises(native.BarnacleError, match="bin index out of range"):
        ctx.scene(sc)
    with pytest.raises(RuntimeError, match="bin index out of range"):
        orc.scene(sc)
    ok = scene.load_scene(configs.cornell_box(16, 16, 1, 2))
    assert np.isfinite(ctx.scene(ok).render_pt(ok.pt_params())).all()


def test_singular_instance_transform_matches_oracle(ctx, orc):
    """A box flattened to zero height: Matrix4x4.Invert of its transform is
    singular, so its world-to-object map is not finite and rays entering it
    take the exact path.  First hits, counters and film as the oracle's."""
    info = configs.cornell_box(48, 48, 2, 4, boxes=True)
    info["transforms"][0]["keyframes"][0]["scale"] = [82.5, 0.0, 82.5]
    sc = scene.load_scene(info)
    p = sc.pt_params()
    ds, os_ = ctx.scene(sc), orc.scene(sc)
    assert not np.isfinite(os_.export()["w2o"]).all()
    gi, gp, gt = ds.first_hit(p, 0)
    oi, op, ot = os_.first_hit(p, 0)
    assert np.array_equal(gi, oi) and np.array_equal(gp, op)
    assert np.array_equal(gt.view(np.int32), ot.view(np.int32))
    g, gst = ds.render_pt(p, stats=True)
    o, ost = os_.render(p)
    for k in ("nodes_closest", "tris_closest", "insts_closest", "nodes_shadow", "tris_shadow"):
        assert gst[k] == ost[k], k
    assert np.array_equal(np.isfinite(g), np.isfinite(o))
    ok = np.all(np.isfinite(o), axis=1)
    assert rel_rmse(g[ok], o[ok]) <= 1e-3


@pytest.mark.parametrize("w,h,spp", [(37, 23, 1), (1, 1, 3), (13, 1, 2), (9, 9, 0)])
def test_ragged_and_empty_films(ctx, orc, w, h, spp):
    """Films whose sides are not multiples of the 8x8 path tiles, a single
    pixel, a single row, and zero samples per pixel (an all-zero film, as
    ProgressiveIntegrator leaves it): the oracle's film."""
    sc = scene.load_scene(configs.config0())
    p = sc.pt_params(width=w, height=h, spp=spp)
    g = ctx.scene(sc).render_pt(p)
    o, _ = orc.scene(sc).render(p)
    assert g.shape == (w * h, 3)
    assert np.array_equal(g.view(np.int32), o.view(np.int32))
    if spp == 0:
        assert not g.any()
