"""Film output (Film.PostProcess / Film.Save, Base/Film.fs:21-30, 55-66), the
PNG and PFM writers, and the -i/-o command line (Program.fs:6-27) without a
GPU: the tone curves at points where they are known in closed form, the byte
conversion, file round trips, and the CLI failing loudly when it cannot
render."""
import os

import numpy as np
import pytest

from barnacle_amd import cli, film, native, scene

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_tone_curves_at_known_points():
    x = np.array([[0.0, 1.0, 0.5], [4.0, 1e-3, 1e6]], np.float32)
    aces, _ = film.postprocess(x, "aces")
    # ACES fit of Film.fs:24-26: f(0) = 0, f(1) = (2.51 + 0.03) / (2.43 + 0.59 + 0.14),
    # f(x) -> 2.51 / 2.43 > 1 for large x, clamped to 1
    assert aces[0, 0] == 0.0
    assert aces[0, 1] == pytest.approx(2.54 / 3.16, rel=1e-6)
    assert aces[0, 2] == pytest.approx(0.5 * (2.51 * 0.5 + 0.03) / (0.5 * (2.43 * 0.5 + 0.59) + 0.14), rel=1e-6)
    assert aces[1, 2] == 1.0
    gam, _ = film.postprocess(x, "gamma")
    assert gam[0, 2] == pytest.approx(0.5 ** (1 / 2.2), rel=1e-6)
    assert gam[1, 1] == pytest.approx(1e-3 ** (1 / 2.2), rel=1e-5)
    ident, _ = film.postprocess(x, "identity")
    assert np.array_equal(ident, np.clip(x, 0, 1))


def test_byte_conversion_and_nan():
    """Rgba32(Vector3): v * 255 + 0.5 truncated; a NaN pixel (a PSSMLT NaN
    chain) becomes 0 under .NET 9's saturating conversion; alpha is 255."""
    v = np.array([[0.5 / 255 - 1e-6, 0.5 / 255 + 1e-6, 1.0], [np.nan, -3.0, 254.5 / 255]], np.float32)
    _, b = film.postprocess(v, "identity")
    assert b.tolist() == [[0, 1, 255, 255], [0, 0, 255, 255]]
    rgb, _ = film.postprocess(v, "identity")
    assert np.isnan(rgb[1, 0])


def test_unknown_tone_mapping():
    with pytest.raises(ValueError):
        film.postprocess(np.zeros((1, 3), np.float32), "filmic")


def test_png_and_pfm_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    w, h = 7, 5
    f = rng.uniform(0, 3, (w * h, 3)).astype(np.float32)
    png, pfm = str(tmp_path / "a.png"), str(tmp_path / "a.pfm")
    film.save(png, f, w, h, "aces")
    film.save(pfm, f, w, h, "aces")
    _, want = film.postprocess(f, "aces")
    got = film.read_png_rgba(png)
    assert got.shape == (h, w, 4)
    assert np.array_equal(got.reshape(-1, 4), want)
    assert np.array_equal(film.read_pfm(pfm), f)
    try:
        from PIL import Image
    except ImportError:
        return
    assert np.array_equal(np.asarray(Image.open(png)).reshape(-1, 4), want)


def test_cli_requires_input_and_output():
    with pytest.raises(SystemExit):
        cli.main(["-i", os.path.join(GOLDEN, "cbox.json")])


def test_cli_fails_loudly_without_the_gpu_library(tmp_path, monkeypatch):
    """The CLI renders only through the HIP library: with no device (or no
    library) it raises instead of falling back to anything else."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises((native.BarnacleError, RuntimeError)):
        cli.main(["-i", os.path.join(GOLDEN, "cbox.json"), "-o", str(tmp_path / "x.png")])
    assert not (tmp_path / "x.png").exists()


def test_cbox_fixture_loads_with_the_reference_semantics():
    """Asset/cbox.json (the reference's only scene): PSSMLT with the loader's
    defaults (Loader.fs:185-204), thin lens with push-forward 140 and a camera
    aspect (1.3333) other than the film's, a 60-radian rotation and a (12, 18,
    12) scale on the cube, a zero-albedo material, ACES film."""
    sc = scene.load_scene(os.path.join(GOLDEN, "cbox.json"))
    it = sc.integrator
    assert it["type"] == "pssmlt" and it["spp"] == 16 and it["max_depth"] == 8 and it["rr_depth"] == 5
    assert it["n_bootstrap"] == 4 * 1024 * 1024 and it["n_chains"] == 1024 and it["strategy"] == 0
    assert it["sigma"] == np.float32(1e-2) and it["large_step_prob"] == np.float32(0.5)
    cam = sc.camera
    assert cam["kind"] == 1 and cam["aperture"] == 4.0 and cam["focus_distance"] == 210.0
    assert cam["push_forward"] == 140.0 and cam["aspect"] == np.float32(1.3333) and cam["fov"] == 25.0
    assert sc.film == (1024, 768, "aces")
    cube = sc.transforms[2]
    assert cube["rotation"][1] == 60.0 and [float(v) for v in cube["scale"]] == [12.0, 18.0, 12.0]
    assert any(m["albedo"] == [0.0, 0.0, 0.0] for m in sc.materials)
    kinds = [p["kind"] for p in sc.primitives]
    assert kinds.count(1) == 1 and len(sc.instances) == 9


def test_bunny_obj_fixture():
    """Asset/stanford-bunny.obj through the OBJ reader (Mesh.fs:246-279)."""
    v, ix = scene.load_obj(os.path.join(GOLDEN, "stanford-bunny.obj"))
    assert v.shape == (35947, 3) and ix.size == 3 * 69451
    assert ix.min() == 0 and ix.max() == len(v) - 1
    lo, hi = v.min(0), v.max(0)
    assert np.allclose(lo, [-0.0947, 0.0330, -0.0619], atol=1e-3)
    assert np.allclose(hi, [0.0610, 0.1873, 0.0588], atol=1e-3)


def test_scene_uri_mesh_loads_the_obj(tmp_path):
    """A "mesh" primitive with "uri" is read relative to the scene file."""
    import json
    import shutil
    shutil.copy(os.path.join(GOLDEN, "stanford-bunny.obj"), tmp_path / "bunny.obj")
    info = json.load(open(os.path.join(GOLDEN, "cbox.json"), encoding="utf-8-sig"))
    info["primitives"].append({"type": "mesh", "uri": "bunny.obj"})
    info["instances"].append({"primitive": len(info["primitives"]) - 1, "material": 4})
    info["nodes"][0]["instances"].append(len(info["instances"]) - 1)
    path = tmp_path / "scene.json"
    path.write_text(json.dumps(info))
    sc = scene.load_scene(str(path))
    assert sc.primitives[-1]["n_indices"] == 3 * 69451 and sc.n_triangles == 69451 + 6 * 2 + 2 + 12
