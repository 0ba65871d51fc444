"""The reference's own scene and mesh files end to end on the GPU, and the
remaining integrators of its CLI.

* Asset/cbox.json (tests/golden/cbox.json): PSSMLT at a reduced film and
  chain count against the oracle (bootstrap luminances, B and per-chain
  accept counts bit-exact, NaN masks equal); path tracing of the same scene
  against the oracle; at the file's own film size the ACES-mapped 8-bit
  images of path tracing and PSSMLT against the reference's two sample
  renders (mean RGB within 0.01, 32 x 32-pixel block means within 0.05
  outside the metal sphere: a weak pin, the samples' settings and runs are
  not recorded, and their sphere differs, see SPHERE).
* DirectIntegrator (Direct.fs:10-40) against the oracle.
* Asset/stanford-bunny.obj as config 1's mesh: first hits bit-exact, film
  against the oracle.
* The -i/-o command line: the PNG it writes is Film.PostProcess of the film
  it renders, which equals bn_render_pt's.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from barnacle_amd import configs, film, scene

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CBOX = os.path.join(GOLDEN, "cbox.json")


def rel_rmse(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def cbox_info(**over):
    info = json.load(open(CBOX, encoding="utf-8-sig"))
    for k, v in over.items():
        sect, key = k.split(".")
        info[sect][key] = v
    return info


def test_cbox_pssmlt_reduced_matches_oracle(ctx, orc):
    info = cbox_info(**{"film.width": 256, "film.height": 192, "integrator.n-bootstrap": 1 << 14,
                        "integrator.n-chains": 1024, "integrator.spp": 1})
    sc = scene.load_scene(info, base_dir=GOLDEN)
    p = sc.mlt_params()
    print("cbox pssmlt 256x192, 2^14 bootstrap, 1024 chains", flush=True)
    ds = ctx.scene(sc)
    g, gs = ds.render_pssmlt(p)
    gboot, gacc = ds.pssmlt_last(p)
    o, os_, oboot, oacc = orc.scene(sc).render_pssmlt(p, threads=8)
    assert np.array_equal(gboot.view(np.int32), oboot.view(np.int32))
    assert gs["b"] == os_["b"] and gs["b"] > 0
    assert np.array_equal(gacc, oacc)
    assert gs["accepted"] == os_["accepted"] and gs["proposed"] == os_["proposed"]
    nan_g, nan_o = ~np.isfinite(g).all(1), ~np.isfinite(o).all(1)
    assert np.array_equal(nan_g, nan_o)
    fin = ~nan_o
    assert rel_rmse(g[fin], o[fin]) <= 1e-5


def test_cbox_path_tracing_matches_oracle(ctx, orc):
    info = cbox_info(**{"film.width": 128, "film.height": 96, "integrator.type": "path-tracing",
                        "integrator.spp": 4})
    sc = scene.load_scene(info, base_dir=GOLDEN)
    p = sc.pt_params()
    g = ctx.scene(sc).render_pt(p)
    o, _ = orc.scene(sc).render(p, threads=8)
    assert np.array_equal(~np.isfinite(g), ~np.isfinite(o))
    fin = np.isfinite(o).all(1)
    assert rel_rmse(g[fin], o[fin]) <= 1e-3 and np.mean(g == o) > 0.99
    gi, gp, gt = ctx.scene(sc).first_hit(p, 2)
    oi, op, ot = orc.scene(sc).first_hit(p, 2)
    assert np.array_equal(gi, oi) and np.array_equal(gp, op) and np.array_equal(gt.view(np.int32), ot.view(np.int32))


@pytest.mark.parametrize("name", ["config0", "two_lights", "spheres", "cbox"])
def test_direct_integrator_matches_oracle(ctx, orc, name):
    if name == "config0":
        sc = scene.load_scene(configs.cornell_box(64, 64, 8, 8))
    elif name == "two_lights":
        import test_closed_form as cf
        sc = scene.load_scene(cf.two_light_scene(spp=4096, integrator="direct"))
    elif name == "spheres":
        sc = scene.load_scene(configs.sphere_scene(48, 48, 4, 8))
    else:
        sc = scene.load_scene(cbox_info(**{"film.width": 96, "film.height": 72, "integrator.type": "direct",
                                           "integrator.spp": 4}), base_dir=GOLDEN)
    p = sc.pt_params()
    g = ctx.scene(sc).render_direct(p)
    o, _ = orc.scene(sc).render(p, integrator=1, threads=8)
    assert np.array_equal(~np.isfinite(g), ~np.isfinite(o))
    fin = np.isfinite(o).all(1)
    assert np.abs(o[fin]).sum() > 0
    assert rel_rmse(g[fin], o[fin]) <= 1e-3
    assert np.mean(g == o) > 0.99


def test_bunny_config1_variant(ctx, orc):
    sc = scene.load_scene(configs.config1_bunny(os.path.join(GOLDEN, "stanford-bunny.obj"), 128, 128, 2, 16))
    assert sc.n_triangles == 69451 + 24
    p = sc.pt_params()
    ds = ctx.scene(sc)
    for k in (0, 1):
        gi, gp, gt = ds.first_hit(p, k)
        oi, op, ot = orc.scene(sc).first_hit(p, k)
        assert np.array_equal(gi, oi) and np.array_equal(gp, op)
        assert np.array_equal(gt.view(np.int32), ot.view(np.int32))
    bunny = len(sc.instances) - 1
    ex = orc.scene(sc).export()
    k_bunny = int(np.nonzero(ex["inst_order"] == bunny)[0][0])
    assert np.mean(gi == k_bunny) > 0.02                  # the bunny is in view
    g = ds.render_pt(p)
    o, _ = orc.scene(sc).render(p, threads=8)
    assert rel_rmse(g, o) <= 1e-3


def test_cli_renders_png_and_pfm(tmp_path):
    info = cbox_info(**{"film.width": 64, "film.height": 48, "integrator.type": "path-tracing",
                        "integrator.spp": 2})
    path = tmp_path / "cbox.json"
    path.write_text(json.dumps(info))
    png, pfm = str(tmp_path / "out.png"), str(tmp_path / "out.pfm")
    r = subprocess.run([sys.executable, "-m", "barnacle_amd", "-i", str(path), "-o", png, "--pfm", pfm],
                       cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert f"Loaded scene from {path}" in r.stdout and "Render time:" in r.stdout
    lin = film.read_pfm(pfm)
    _, want = film.postprocess(lin, "aces")
    assert np.array_equal(film.read_png_rgba(png).reshape(-1, 4), want)
    from barnacle_amd import native
    c = native.Context(0)
    try:
        sc = scene.load_scene(str(path))
        direct = c.scene(sc).render_pt(sc.pt_params())
    finally:
        c.close()
    assert np.array_equal(lin.view(np.int32), direct.view(np.int32))


# Blocks of the 32 x 24 thumbnail that hold the metal sphere (pbr, roughness
# 0.4, metallic 0.9).  There the reference's sample renders differ from every
# render of cbox.json as written: swept on the GPU, the sphere's blocks match
# a GGX alpha near 0.3 (roughness ~0.55) within 0.03 but alpha = 0.16
# (roughness 0.4, PBR.fs:11) only within 0.14, while the rest of the image
# agrees within 0.03 at alpha = 0.16.  The samples were evidently made with
# another material setting or PBR revision; the reference keeps no history.
SPHERE = (slice(13, 21), slice(7, 14))


def tone_mapped(g):
    _, rgba = film.postprocess(g, "aces")
    return rgba[:, :3].reshape(768, 1024, 3).astype(np.float64) / 255.0


def thumb_diff(img, ref):
    d = np.abs(img.reshape(24, 32, 32, 32, 3).mean(axis=(1, 3)) - np.array(ref)).max(-1)
    outside = d.copy()
    outside[SPHERE] = 0.0
    return d, outside


def render_cbox(ctx, over):
    info = cbox_info(**over)
    sc = scene.load_scene(info, base_dir=GOLDEN)
    ds = ctx.scene(sc)
    if sc.integrator["type"] == "pssmlt":
        return ds.render_pssmlt(sc.mlt_params())[0]
    return ds.render_pt(sc.pt_params())


def test_cbox_path_tracing_full_size_against_sample_render(ctx):
    """cbox.json path traced at its own film (1024 x 768, thin lens, ACES),
    256 spp, tone mapped and quantised as Film.Save does, against "Sample -
    Path Tracing.png": mean RGB within 0.01, 32 x 32-pixel block means within
    0.05 outside the metal sphere (see SPHERE)."""
    gold = json.load(open(os.path.join(GOLDEN, "cbox_samples.json")))["path_tracing"]
    g = render_cbox(ctx, {"integrator.type": "path-tracing", "integrator.spp": 256})
    img = tone_mapped(g)
    mean = img.mean(axis=(0, 1))
    d, outside = thumb_diff(img, gold["thumbnail"])
    print(f"cbox PT 256 spp: mean RGB {mean.round(4)} vs sample {gold['mean_rgb']}; block diff outside the "
          f"sphere {outside.max():.4f}, on it {d[SPHERE].max():.4f}", flush=True)
    assert np.all(np.abs(mean - np.array(gold["mean_rgb"])) <= 0.01)
    assert outside.max() <= 0.05


def test_cbox_pssmlt_against_sample_render(ctx):
    """PSSMLT on cbox.json.  At the file's own settings (16 mutations per
    pixel in 1024 chains of 12,288) the linear film's mean equals the path
    traced one within 1%, but the image is grainy and the concave ACES curve
    pulls the tone-mapped mean down (0.537 against the sample's 0.558), so
    that run is held to 0.03.  With 2^14 chains of 3,072 mutations (64 per
    pixel) it meets the sample like the path tracer: mean RGB within 0.01,
    block means within 0.05 outside the sphere."""
    gold = json.load(open(os.path.join(GOLDEN, "cbox_samples.json")))["pssmlt"]
    ref = np.array(gold["mean_rgb"])
    print("cbox.json PSSMLT at its own settings", flush=True)
    g = render_cbox(ctx, {})
    pt = render_cbox(ctx, {"integrator.type": "path-tracing", "integrator.spp": 256})
    fin = np.isfinite(g).all(1)
    assert fin.mean() > 0.999
    lin, lin_pt = g[fin].mean(0), pt[np.isfinite(pt).all(1)].mean(0)
    assert np.all(np.abs(lin / lin_pt - 1) <= 0.01), (lin, lin_pt)
    assert np.all(np.abs(tone_mapped(g).mean(axis=(0, 1)) - ref) <= 0.03)
    g = render_cbox(ctx, {"integrator.n-chains": 1 << 14, "integrator.spp": 64})
    img = tone_mapped(g)
    mean = img.mean(axis=(0, 1))
    d, outside = thumb_diff(img, gold["thumbnail"])
    print(f"cbox PSSMLT 2^14 chains, 64 spp: mean RGB {mean.round(4)} vs sample {ref}; block diff outside "
          f"the sphere {outside.max():.4f}, on it {d[SPHERE].max():.4f}", flush=True)
    assert np.all(np.abs(mean - ref) <= 0.01)
    assert outside.max() <= 0.05


def test_keyframe_animation_matches_oracle(ctx, orc):
    """Scenes whose transforms interpolate between key frames at t = 0: the
    scene build's instance matrices equal the oracle's bit for bit, and so do
    the first hits of the animated sphere."""
    import test_keyframes as kf
    cases = [
        [{"time": 1.0, "scale": [4.0, 4.0, 4.0], "rotation": [0.0, 1.0, 0.0], "translation": [30.0, 4.0, 5.0]},
         {"time": -1.0, "scale": [2.0, 2.0, 2.0], "rotation": [0.0, 0.4, 0.0], "translation": [10.0, 2.0, 3.0]}],
        [{"time": 0.0, "scale": [2.0, 2.0, 2.0], "rotation": [0.3, 0.0, 0.0], "translation": [5.0, 6.0, 7.0]},
         {"time": 1.0, "scale": [9.0, 9.0, 9.0], "translation": [1.0, 1.0, 1.0]}],
        [{"time": -2.0, "rotation": [0.0, 3.0, 0.0], "translation": [8.0, 9.0, 10.0]},
         {"time": 2.0, "rotation": [0.0, -3.0, 0.0], "translation": [8.0, 9.0, 10.0]}],
    ]
    for kfs in cases:
        info = kf.animated_scene(kfs)
        info["film"]["width"] = info["film"]["height"] = 64
        sc = scene.load_scene(info)
        ds = ctx.scene(sc)
        g, o = ds.export(), orc.scene(sc).export()
        assert np.array_equal(g["inst_order"], o["inst_order"])
        assert np.array_equal(g["o2w"].view(np.int32), o["o2w"].view(np.int32))
        assert np.array_equal(g["w2o"].view(np.int32), o["w2o"].view(np.int32))
        p = sc.pt_params()
        gi, gp, gt = ds.first_hit(p, 0)
        oi, op, ot = orc.scene(sc).first_hit(p, 0)
        assert np.array_equal(gi, oi) and np.array_equal(gp, op)
        assert np.array_equal(gt.view(np.int32), ot.view(np.int32))
        k = int(np.nonzero(o["inst_order"] == 0)[0][0])
        assert np.sum(gi == k) >= 4                           # the animated sphere is in view
