"""Analytic spheres and the thin-lens camera, GPU against the oracle.

The scene (configs.sphere_scene) has the features of the reference's own
asset that configs 0-4 leave out: SpherePrimitive instances under uniform and
non-uniform scale (Sphere.fs:13-77), a sphere light sampled beside the quad
light (Sphere.fs:90-126), a rotated mirror cube, and a thin-lens camera with a
finite aperture (Camera/ThinLens.fs:12-35).  Bars as for the other scenes:
first hits bit-exact, traversal counters equal, films within rel-RMSE 1e-3
and almost all bit-equal, PSSMLT chain trajectories exact, batched films
equal to plain ones."""
import numpy as np
import pytest

from barnacle_amd import configs, scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sph():
    return scene.load_scene(configs.sphere_scene(64, 64, 4, 8))


def test_first_hit_spheres_thin_lens(ctx, orc, sph):
    p = sph.pt_params()
    for k in (0, 3):
        gi, gp, gt = ctx.scene(sph).first_hit(p, k)
        oi, op, ot = orc.scene(sph).first_hit(p, k)
        assert np.array_equal(gi, oi)
        assert np.array_equal(gp, op)
        assert np.array_equal(gt.view(np.int32), ot.view(np.int32))
    # the metal, glass and emitting spheres are all seen
    ex = orc.scene(sph).export()
    spheres = [k for k, i in enumerate(ex["inst_order"]) if sph.primitives[sph.instances[i][0]]["kind"] == 1]
    assert len(spheres) == 3 and np.isin(spheres, gi).all()


def test_pt_spheres_film_and_counters(ctx, orc, sph):
    p = sph.pt_params()
    g, gst = ctx.scene(sph).render_pt(p, stats=True)
    o, ost = orc.scene(sph).render(p)
    for k in ("rays_closest", "nodes_closest", "tris_closest", "insts_closest",
              "rays_shadow", "nodes_shadow", "tris_shadow", "insts_shadow"):
        assert gst[k] == ost[k], k
    assert np.array_equal(np.isfinite(g), np.isfinite(o))
    ok = np.all(np.isfinite(o), axis=1)
    rel = np.sqrt(np.sum((g[ok] - o[ok]) ** 2) / np.sum(o[ok] ** 2))
    assert rel <= 1e-3
    assert np.mean(g == o) > 0.99


def test_pssmlt_spheres(ctx, orc, sph):
    info = configs.sphere_scene(32, 32, 8, 8)
    info["integrator"] = {"type": "pssmlt", "spp": 8, "max-depth": 8, "rr-depth": 5, "n-bootstrap": 2048,
                          "n-chains": 256, "mutation-strategy": "Gaussian", "mutation-sigma": 1.0 / 64.0,
                          "large-step-prob": 0.3}
    sc = scene.load_scene(info)
    p = sc.mlt_params()
    ds = ctx.scene(sc)
    _, gst = ds.render_pssmlt(p)
    gboot, gacc = ds.pssmlt_last(p)
    _, ost, oboot, oacc = orc.scene(sc).render_pssmlt(p)
    assert np.array_equal(gboot.view(np.int32), oboot.view(np.int32))
    assert gst["b"] == ost["b"]
    assert np.array_equal(gacc, oacc)


def test_batch_with_spheres_equals_plain(ctx, sph):
    c0 = scene.load_scene(configs.cornell_box(32, 32, 2, 8))
    jobs = [(ctx.scene(sph), sph.pt_params()), (ctx.scene(c0), c0.pt_params()),
            (ctx.scene(sph), sph.pt_params(frame_id=1))]
    films, _ = ctx.render_batch(jobs)
    for (ds, p), f in zip(jobs, films):
        assert np.array_equal(f.view(np.int32), ds.render_pt(p).view(np.int32))


def test_nested_node_transforms_first_hit(ctx, orc):
    """A sphere two node levels down (transforms composed local * parent):
    the GPU scene build places it where the oracle does, whose placement
    tests/test_closed_form.py pins to the closed form."""
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    from test_closed_form import nested_sphere_scene
    sc = scene.load_scene(nested_sphere_scene(48, 48))
    p = sc.pt_params()
    gi, gp, gt = ctx.scene(sc).first_hit(p, 0)
    oi, op, ot = orc.scene(sc).first_hit(p, 0)
    assert np.array_equal(gi, oi) and np.array_equal(gp, op)
    assert np.array_equal(gt.view(np.int32), ot.view(np.int32))
    assert (gi >= 0).sum() > 100


def test_normal_view_spheres(ctx, orc, sph):
    """NormalIntegrator (Extensions/Integrator/Normal.fs:10-17) on the sphere
    scene: sphere normals (Sphere.fs:60-77) through instance transforms."""
    p = sph.pt_params(spp=2)
    g = ctx.scene(sph).render_normal(p)
    o, _ = orc.scene(sph).render(p, integrator=0)
    assert np.mean(g == o) > 0.999
    assert np.sqrt(np.sum((g - o) ** 2) / np.sum(o ** 2)) <= 1e-5


def test_closed_form_scenes_on_the_gpu(ctx, orc):
    """The closed-form scenes that pin the oracle, rendered on the GPU: inside
    an emitting Lambertian sphere the film is Le (sum_{k<D} a^k + a^D / 2)
    (no roulette), and the sphere-light floor point matches the oracle's
    film bit for bit at equal samples (its expectation is rho Le (r/d)^2)."""
    import test_oracle_pins as pins
    a, le, depth = 0.6, 1.5, 8
    sc = scene.load_scene(pins.sphere_furnace_scene(a, le, depth, rr_depth=100, w=16, h=16, spp=4))
    g = ctx.scene(sc).render_pt(sc.pt_params())
    o, _ = orc.scene(sc).render(sc.pt_params(), threads=4)
    assert np.allclose(g, pins.furnace_radiance(a, le, depth), rtol=2e-5)
    assert np.array_equal(g, o)
    sc = scene.load_scene(pins.sphere_light_scene())
    info = pins.sphere_light_scene()
    info["integrator"]["spp"] = 4096
    sc = scene.load_scene(info)
    p = sc.pt_params()
    g = ctx.scene(sc).render_pt(p)
    o, _ = orc.scene(sc).render(p, threads=4)
    assert np.allclose(g, o, rtol=1e-5)
    assert g[0, 0] == pytest.approx(0.5 * 5.0 * 0.01, rel=0.05)
