"""Path-replay adjoint on the GPU (bn_backward_pt) against the FP64 oracle
adjoint of the same definition and against central finite differences with
common random numbers of GPU renders (h=1e-3, Russian roulette off so that
the estimator is smooth in albedo and emission)."""
import numpy as np
import pytest

from barnacle_amd import configs, scene

pytestmark = pytest.mark.gpu


def make(max_depth=6, rr_depth=3, spp=4, res=24, patches=16, freq=6):
    info = configs.config3(width=res, height=res, spp=spp, max_depth=max_depth, freq=freq, n_patches=patches)
    info["integrator"]["rr-depth"] = rr_depth
    return info, scene.load_scene(info)


def film_adjoint(film, seed=5):
    d = np.random.default_rng(seed).normal(size=film.shape).astype(np.float32)
    d[~np.isfinite(film).all(axis=1)] = 0.0
    return d


@pytest.mark.parametrize("rr_depth", [3, 6])
def test_adjoint_matches_fp64_oracle(ctx, orc, rr_depth):
    info, sc = make(rr_depth=rr_depth)
    p = sc.pt_params()
    ds = ctx.scene(sc)
    film = ds.render_pt(p)
    d = film_adjoint(film)
    gm, gl = ds.backward_pt(p, d)
    om, ol = orc.scene(sc).adjoint_pt(p, d)
    for g, o in ((gm[:, :3], om[:, :3]), (gm[:, 3], om[:, 3]), (gl, ol)):
        scale = np.abs(o).max()
        assert scale > 0
        assert np.allclose(g, o, rtol=1e-2, atol=2e-3 * scale)
        assert np.linalg.norm(g - o) <= 5e-3 * np.linalg.norm(o)


def test_adjoint_matches_gpu_fd(ctx):
    info, sc = make(max_depth=4, rr_depth=4)
    p = sc.pt_params()
    ds = ctx.scene(sc)
    (mat, light), _, _ = configs.config3_params(sc, info)
    film = ds.render_pt(p)
    finite = np.isfinite(film).all(axis=1)
    d = film_adjoint(film)
    gm, gl = ds.backward_pt(p, d)
    fm, fl = info["x-patches"]["first_material"], info["x-patches"]["first_light"]
    h = 1e-3
    checks = [("mat", 0, 0), ("mat", 2, 2), ("mat", fm + 1, 0), ("mat", fm + 4, 1), ("light", 0, 1),
              ("light", fl + 2, 0), ("light", fl + 6, 2)]
    try:
        for which, k, c in checks:
            vals = []
            for sgn in (1, -1):
                m2, l2 = mat.copy(), light.copy()
                (m2 if which == "mat" else l2)[k, c] += sgn * h
                ds.set_params(m2, l2)
                f = ds.render_pt(p)
                vals.append(float(np.sum(np.where(finite[:, None], f, 0).astype(np.float64) * d)))
            fd = (vals[0] - vals[1]) / (2 * h)
            ad = gm[k, c] if which == "mat" else gl[k, c]
            assert abs(fd - ad) <= 0.01 * abs(fd) + 1e-5, (which, k, c, fd, ad)
    finally:
        ds.set_params(mat, light)


def test_adjoint_4096_patches_window_matches_oracle(ctx, orc):
    """Full 4096-patch inverse scene at a small film: the adjoint of a 16x16
    window of pixels equals the oracle's over the same window."""
    info = configs.config3(width=96, height=96, spp=2, max_depth=8)
    sc = scene.load_scene(info)
    p = sc.pt_params()
    ds = ctx.scene(sc)
    film = ds.render_pt(p)
    d = film_adjoint(film)
    x0, y0 = 40, 30
    mask = np.zeros((p.height, p.width), bool)
    mask[y0:y0 + 16, x0:x0 + 16] = True
    d[~mask[::-1].reshape(-1)] = 0.0                       # film rows are stored bottom-up
    gm, gl = ds.backward_pt(p, d)
    om, ol = orc.scene(sc).adjoint_pt(p, d, window=(x0, y0, x0 + 16, y0 + 16))
    for g, o in ((gm, om), (gl, ol)):
        scale = np.abs(o).max()
        assert scale > 0
        assert np.linalg.norm(g - o) <= 5e-3 * np.linalg.norm(o)


def test_autograd_function_and_optimisation(ctx):
    """RenderPT.apply's gradient equals bn_backward_pt's, and Adam on the
    masked L2 loss lowers it."""
    torch = pytest.importorskip("torch")
    from barnacle_amd import diff
    info, sc = make(max_depth=4, rr_depth=3, spp=4, res=24)
    p = sc.pt_params()
    ds = ctx.scene(sc)
    (tm, tl), (m0, l0), _ = configs.config3_params(sc, info)
    ds.set_params(tm, tl)
    target = torch.from_numpy(ds.render_pt(p))
    mat = torch.tensor(m0, requires_grad=True)
    light = torch.tensor(l0, requires_grad=True)
    film = diff.RenderPT.apply(mat, light, ds, p)
    loss = diff.masked_l2(film, target)
    loss.backward()
    ok = torch.isfinite(film).all(dim=1) & torch.isfinite(target).all(dim=1)
    d = torch.where(ok[:, None], 2.0 * (film - target) / (int(ok.sum()) * 3), torch.zeros_like(film))
    gm, gl = ds.backward_pt(p, d.detach().numpy().astype(np.float32))
    assert np.allclose(mat.grad.numpy(), gm, rtol=1e-3, atol=1e-6 * np.abs(gm).max())
    assert np.allclose(light.grad.numpy(), gl, rtol=1e-3, atol=1e-6 * np.abs(gl).max())
    _, hist = diff.optimise(ds, info, sc, p, steps=8, lr=5e-2, target=target)
    assert hist[-1] < hist[0]


def test_adjoint_sphere_scene_matches_fp64_oracle(ctx, orc):
    """Gradients through analytic spheres (metal GGX, glass), a mirror cube,
    a sphere light beside the quad light and the thin-lens camera."""
    sc = scene.load_scene(configs.sphere_scene(32, 32, 4, 6))
    p = sc.pt_params()
    ds = ctx.scene(sc)
    film = ds.render_pt(p)
    d = film_adjoint(film)
    gm, gl = ds.backward_pt(p, d)
    om, ol = orc.scene(sc).adjoint_pt(p, d)
    for g, o in ((gm[:, :3], om[:, :3]), (gm[:, 3], om[:, 3]), (gl, ol)):
        scale = np.abs(o).max()
        assert scale > 0
        assert np.allclose(g, o, rtol=1e-2, atol=2e-3 * scale)
        assert np.linalg.norm(g - o) <= 5e-3 * np.linalg.norm(o)


def test_adjoint_lambertian_box_and_zero_loss(ctx, orc):
    """The plain Cornell box (three Lambertian materials, one quad light):
    gradients as the FP64 oracle's; a zero film adjoint gives exactly zero
    gradients."""
    sc = scene.load_scene(configs.cornell_box(24, 24, 4, 5))
    p = sc.pt_params()
    ds = ctx.scene(sc)
    film = ds.render_pt(p)
    d = film_adjoint(film)
    gm, gl = ds.backward_pt(p, d)
    om, ol = orc.scene(sc).adjoint_pt(p, d)
    for g, o in ((gm[:, :3], om[:, :3]), (gl, ol)):
        scale = np.abs(o).max()
        assert scale > 0
        assert np.allclose(g, o, rtol=1e-2, atol=2e-3 * scale)
    zm, zl = ds.backward_pt(p, np.zeros_like(d))
    assert not zm.any() and not zl.any()


def test_adjoint_uncached_and_multi_round_match_cached(ctx, orc):
    """backward_pt's other two paths: with no forward pass of the same frame
    cached (after a normal view, which invalidates the cache, the forward
    rounds run inside backward_pt), and with rounds smaller than the frame
    (window_paths below the frame's paths: several forward + replay rounds).
    Both equal the cached single-round gradients up to float atomics order,
    and the FP64 oracle adjoint."""
    info, sc = make(max_depth=5, rr_depth=3, spp=8, res=32)
    p = sc.pt_params()
    ds = ctx.scene(sc)
    film = ds.render_pt(p)
    d = film_adjoint(film)
    cached = ds.backward_pt(p, d)
    ds.render_normal(p)                                     # invalidates the cached forward
    uncached = ds.backward_pt(p, d)
    om, ol = orc.scene(sc).adjoint_pt(p, d)
    # 32x32 pixels = 1024 paths per sample: 2048-path rounds give 4 rounds of 2 samples
    ctx.set_option("window_paths", 2048)
    try:
        multi = ds.backward_pt(p, d)
        film_multi = ds.render_pt(p)
    finally:
        ctx.set_option("window_paths", 0)
    assert np.array_equal(film_multi.view(np.int32), film.view(np.int32))   # rounds do not change the film
    for g in (uncached, multi):
        for a, b in zip(g, cached):
            scale = np.abs(b).max()
            assert np.allclose(a, b, rtol=1e-4, atol=1e-5 * scale)
    for gm, gl in (uncached, multi):
        for g, o in ((gm[:, :3], om[:, :3]), (gm[:, 3], om[:, 3]), (gl, ol)):
            assert np.linalg.norm(g - o) <= 5e-3 * np.linalg.norm(o)
