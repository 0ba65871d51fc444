"""Oracle PSSMLT (Extensions/Integrator/PSSMLT.fs) against independent facts:
the Giles erfinv polynomial against scipy.special.erfinv, the degenerate
alias table's uniform choice, and the normalisation of the estimator (the
mean of the PSSMLT image equals the mean of a converged path-traced image of
the same scene: B, the effective sample count and the large-step weighting
all enter that mean)."""
import numpy as np
import pytest
from scipy.special import erfinv

from barnacle_amd import configs, scene


def test_erfinv_matches_scipy(orc):
    xs = np.linspace(-0.99998, 0.99998, 2001).astype(np.float32)
    got = np.array([orc.erfinv(x) for x in xs], np.float64)
    ref = erfinv(np.clip(xs.astype(np.float64), -0.99999, 0.99999))
    assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3)) < 2e-6


def test_alias_table_is_degenerate_uniform(orc):
    w = np.array([1.0, 3.0, 0.5, 2.5], np.float32)
    al, pr, pd = orc.alias_build(w)
    assert np.array_equal(al, np.arange(4)) and np.all(pr == 1.0)
    assert np.allclose(pd, w / w.sum())
    for u, want in [(0.1, 0), (0.3, 1), (0.6, 2), (0.99, 3)]:
        assert orc.alias_sample((al, pr, pd), np.float32(u))[0] == want
    # Negative weights (possible for PSSMLT luminances) push w/sum above 1 and
    # the pairing loop runs.  Worked by hand from AliasTable.fs:14-51 for
    # w = [3, -2]: sum 1, prob [3, -2]; over = [0], under = [1]; pairing sets
    # prob[0] = 3 + (-2) - 1 = 0, alias[1] = 0, pushes 0 to under; the final
    # pass sets alias[0] = 0, prob[0] = 1.  prob[1] stays -2.
    al2, pr2, pd2 = orc.alias_build(np.array([3.0, -2.0], np.float32))
    assert list(al2) == [0, 0] and list(pr2) == [1.0, -2.0] and list(pd2) == [3.0, -2.0]
    assert orc.alias_sample((al2, pr2, pd2), np.float32(0.75)) == (0, np.float32(3.0))


def closed_box(W=32, spp=8):
    """Cornell box closed by a front wall with the camera inside: no path
    escapes, so no chain starts at zero luminance (the reference's uniform
    chain start, AliasTable.fs:52-62, then carries no NaN or start-up bias
    worth measuring)."""
    info = configs.cornell_box(width=W, height=W, spp=spp, max_depth=6)
    S = 555.0
    info["primitives"].append(configs._quad((0, 0, 0), (0, S, 0), (S, S, 0), (S, 0, 0)))
    info["instances"].append({"primitive": len(info["primitives"]) - 1, "material": 0})
    info["nodes"][0]["instances"].append(len(info["instances"]) - 1)
    for t in info["transforms"]:
        kf = t["keyframes"][0]
        if kf.get("translation") == [277.5, 277.5, 800.0]:
            kf["translation"] = [277.5, 277.5, -20.0]
    info["camera"]["fov"] = 60.0
    return scene.load_scene(info)


@pytest.mark.parametrize("strategy,chains,pl", [(0, 128, 0.3), (0, 2048, 0.3), (1, 256, 0.5), (0, 128, 1.0)])
def test_pssmlt_mean_matches_path_tracing(orc, strategy, chains, pl):
    sc = closed_box()
    osc = orc.scene(sc)
    pt, _ = osc.render(sc.pt_params(spp=512))
    sc.integrator.update(n_bootstrap=1 << 14, n_chains=chains, strategy=strategy, large_step_prob=np.float32(pl))
    mp = sc.mlt_params(spp=128)
    film, st, boot, acc = osc.render_pssmlt(mp)
    assert st["proposed"] == st["mutations_per_chain"] * mp.n_chains
    assert st["mutations_per_chain"] == (128 * 32 * 32 + mp.n_chains - 1) // mp.n_chains
    assert st["b"] == np.float32(np.cumsum(boot, dtype=np.float32)[-1] / np.float32(len(boot)))
    assert 0 < st["accepted"] < st["proposed"]
    assert acc.sum() == st["accepted"]
    assert np.all(np.isfinite(film))
    assert np.allclose(film.mean(axis=0), pt.mean(axis=0), rtol=0.025), (film.mean(axis=0), pt.mean(axis=0))
