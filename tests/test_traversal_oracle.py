"""Pins the oracle's camera, scene graph and ray traversal without trusting it.

* Camera (Pinhole.fs:12-27, Camera.fs:12-20): every sample ray is compared with
  the closed form of a pinhole camera evaluated in float64 from the scene JSON
  (vertical fov in degrees, aspect from the camera, lower-left origin, -Z view,
  row-vector camera-to-world).
* Scene graph (Base/Scene.fs:43-51, Loader.fs:20-41): every instance's
  object-to-world matrix is recomposed in float64 as S*Rx*Ry*Rz*T times the
  parent's, in depth-first order, with System.Numerics rotation matrices.
* Traversal (Extensions/Aggregate/BVH.fs:37-58, Mesh.fs:50-82,217-242,
  Primitive.fs:111-129): the BVH first hit of every camera ray equals a brute
  force over every triangle of every instance, with the same float32
  Moller-Trumbore arithmetic written here in numpy; only exact t ties may name
  a different triangle.
"""
import math

import numpy as np
import pytest

from barnacle_amd import configs, scene

F = np.float32


def rot(axis, a):
    c, s = math.cos(a), math.sin(a)
    m = np.eye(4)
    if axis == 0:
        m[1, 1], m[1, 2], m[2, 1], m[2, 2] = c, s, -s, c
    elif axis == 1:
        m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, -s, s, c
    else:
        m[0, 0], m[0, 1], m[1, 0], m[1, 1] = c, s, -s, c
    return m


def local_matrix(t):
    m = np.eye(4)
    if t["flags"] & 1:
        m = m @ np.diag([float(t["scale"][0]), float(t["scale"][1]), float(t["scale"][2]), 1.0])
    if t["flags"] & 2:
        m = m @ rot(0, float(t["rotation"][0])) @ rot(1, float(t["rotation"][1])) @ rot(2, float(t["rotation"][2]))
    if t["flags"] & 4:
        tm = np.eye(4)
        tm[3, :3] = [float(x) for x in t["translation"]]
        m = m @ tm
    return m


def world_matrices(sc):
    """Depth-first Node.Traverse: (instance index -> o2w, camera o2w)."""
    o2w, cam = {}, None

    def visit(ni, parent):
        nonlocal cam
        tr, has_cam, io, n_i, co, n_c = sc.nodes[ni]
        m = (local_matrix(sc.transforms[tr]) if tr >= 0 else np.eye(4)) @ parent
        if has_cam:
            cam = m
        for k in range(n_i):
            o2w[sc.node_refs[io + k]] = m
        for k in range(n_c):
            visit(sc.node_refs[co + k], m)

    visit(sc.root, np.eye(4))
    return o2w, cam


def lcg_floats(x, y, s, n):
    """Hash.XXHash32Three seed then n Sampler draws (pinned by tests/test_rng.py)."""
    import xxhash
    seed = (s - 8) & 0xFFFFFFFF
    st = xxhash.xxh32_intdigest(np.array([x, y], "<u4").tobytes(), seed=seed)
    out = []
    for _ in range(n):
        st = (0x269EC3 + st * 0x343FD) & 0xFFFFFFFF
        out.append(float(np.array([(st >> 9) | 0x3F800000], np.uint32).view(np.float32)[0]) - 1.0)
    return out


@pytest.fixture(scope="module", params=["config0", "mesh"])
def case(request, orc):
    if request.param == "config0":
        sc = scene.load_scene(configs.config0())
        p = sc.pt_params(width=40, height=32)
    else:
        sc = scene.load_scene(configs.config1(width=36, height=36, spp=1, freq=7))
        p = sc.pt_params()
    return sc, p, orc.scene(sc)


def test_instance_transforms_match_scene_graph(case):
    sc, _, osc = case
    ex = osc.export()
    o2w, _ = world_matrices(sc)
    for i, desc in enumerate(ex["inst_order"]):
        want = o2w[int(desc)]
        got = ex["o2w"][i].reshape(4, 4).astype(np.float64)
        assert np.allclose(got, want, rtol=1e-5, atol=1e-4 * max(1.0, np.abs(want).max())), (i, desc)
        inv = ex["w2o"][i].reshape(4, 4).astype(np.float64)
        assert np.allclose(got @ inv, np.eye(4), atol=1e-4)


def test_camera_rays_match_pinhole_closed_form(case, orc):
    sc, p, osc = case
    cam = sc.camera
    assert cam["kind"] == 0                                  # pinhole
    _, c2w = world_matrices(sc)
    vh = 2.0 * math.tan(float(cam["fov"]) * math.pi / 360.0)
    vw = float(cam["aspect"]) * vh
    rng = np.random.default_rng(3)
    for _ in range(40):
        x, y, k = int(rng.integers(p.width)), int(rng.integers(p.height)), int(rng.integers(4))
        o, d = osc.camera_ray(p, x, y, k)
        u = lcg_floats(x, y, p.frame_id * p.spp + k, 2)
        pc = np.array([-vw / 2 + (x + u[0]) * vw / p.width, -vh / 2 + (y + u[1]) * vh / p.height, -1.0])
        pc /= np.linalg.norm(pc)
        dw = pc @ c2w[:3, :3]
        dw /= np.linalg.norm(dw)
        ow = c2w[3, :3] + float(cam["push_forward"]) * dw
        # float32 evaluation vs float64 closed form; one pixel spans ~1e-2 here
        assert np.allclose(d, dw, atol=2e-5), (x, y, k)
        assert np.allclose(o, ow, rtol=1e-6, atol=1e-4)


def xform_point(m, v):
    return ((v[..., 0:1] * m[0, :3] + v[..., 1:2] * m[1, :3]) + v[..., 2:3] * m[2, :3]) + m[3, :3]


def brute_first_hit(sc, osc, o, d):
    """Closest hit over every triangle of every instance, float32 throughout."""
    ex = osc.export()
    n = len(o)
    best_t = np.full(n, np.inf, F)
    best = np.full((n, 2), -1, np.int64)
    ties = np.zeros(n, bool)
    blas = {}
    for i, desc in enumerate(ex["inst_order"]):
        prim = sc.instances[int(desc)][0]
        pr = sc.primitives[prim]
        if prim not in blas:
            tri = osc.export_blas(prim)["tri"].reshape(-1, 3)
            v = sc.vertices.reshape(-1, 3)[pr["vert_offset"]:pr["vert_offset"] + pr["n_verts"]].astype(F)
            blas[prim] = (v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]])
        p0, p1, p2 = blas[prim]
        w2o = ex["w2o"][i].reshape(4, 4).astype(F)
        oo = xform_point(w2o, o)                            # Ray.Transform (Ray.fs:19-22)
        dd = xform_point(w2o, d) - w2o[3, :3]
        e0, e1 = p1 - p0, p2 - p0                            # (T, 3)

        def cross(a, b):
            return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                             a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                             a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)

        def dot(a, b):
            return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]

        D = dd[:, None, :]
        q = cross(D, e1[None])
        det = dot(e0[None], q)
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            inv = F(1.0) / det
            s = oo[:, None, :] - p0[None]
            u = inv * dot(s, q)
            r = cross(s, e0[None])
            v = inv * dot(D, r)
            t = inv * dot(e1[None], r)
            ok = (np.abs(det) >= np.finfo(F).smallest_subnormal) & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1)
            ok &= t > np.finfo(F).smallest_subnormal
        t = np.where(ok, t, np.inf).astype(F)
        k = np.argmin(t, axis=1)
        tk = t[np.arange(n), k]
        nk = (t == tk[:, None]).sum(axis=1)
        closer = tk < best_t
        same = (tk == best_t) & np.isfinite(tk)
        ties |= same | (closer & (nk > 1))
        ties &= ~(closer & (nk == 1))
        best_t = np.where(closer, tk, best_t)
        best[closer] = np.stack([np.full(n, i), k], 1)[closer]
    return best, best_t, ties


def test_first_hit_equals_brute_force(case):
    sc, p, osc = case
    inst, prim, t = osc.first_hit(p, 0)
    rays = [osc.camera_ray(p, x, y, 0) for y in range(p.height) for x in range(p.width)]
    o = np.array([r[0] for r in rays], F)
    d = np.array([r[1] for r in rays], F)
    best, best_t, ties = brute_first_hit(sc, osc, o, d)
    hit = np.isfinite(best_t)
    assert hit.mean() > 0.5
    assert np.array_equal(inst >= 0, hit)
    assert np.array_equal(t[hit].view(np.int32), best_t[hit].view(np.int32))
    same = (inst[hit] == best[hit, 0]) & (prim[hit] == best[hit, 1])
    assert np.all(same | ties[hit])
    assert ties.mean() < 0.02


def test_triangle_hits_against_float64_linear_solve(case):
    """Moller-Trumbore (Mesh.fs:50-82) pinned by a different method: for
    every camera ray that hits a triangle, solving o' + t d' = p0 + u e0 +
    v e1 in float64 by LU (numpy.linalg.solve) for the object-space ray gives
    the oracle's t within float32 rounding, with (u, v) inside the triangle;
    and no triangle of the scene has a float64 root clearly nearer.  A
    swapped edge, a wrong sign or the wrong cross product moves t or the
    barycentrics."""
    sc, p, osc = case
    inst, prim, t = osc.first_hit(p, 0)
    ex = osc.export()
    rays = [osc.camera_ray(p, x, y, 0) for y in range(p.height) for x in range(p.width)]
    o = np.array([r[0] for r in rays], np.float64)
    d = np.array([r[1] for r in rays], np.float64)
    tris = {}
    nearest = np.full(len(o), np.inf)
    for i, desc in enumerate(ex["inst_order"]):
        pid = sc.instances[int(desc)][0]
        pr = sc.primitives[pid]
        if pr["kind"] != 0:
            continue
        if pid not in tris:
            tri = osc.export_blas(pid)["tri"].reshape(-1, 3)
            v = sc.vertices.reshape(-1, 3)[pr["vert_offset"]:pr["vert_offset"] + pr["n_verts"]].astype(np.float64)
            tris[pid] = (v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]])
        p0, p1, p2 = tris[pid]
        # the object-space ray as Ray.Transform forms it in float32 (its
        # direction quirk included, pinned by the scene-graph tests above)
        w2o = ex["w2o"][i].reshape(4, 4).astype(F)
        oo = xform_point(w2o, o.astype(F)).astype(np.float64)
        dd = (xform_point(w2o, d.astype(F)) - w2o[3, :3]).astype(np.float64)
        # every ray against every triangle: A [t, u, v] = o' - p0, A = [-d', e0, e1]
        e0, e1 = p1 - p0, p2 - p0
        A = np.empty((len(o), len(p0), 3, 3))
        A[..., 0] = -dd[:, None, :]
        A[..., 1] = e0[None]
        A[..., 2] = e1[None]
        b = oo[:, None, :] - p0[None]
        ok = np.abs(np.linalg.det(A)) > 1e-12
        sol = np.full(b.shape, np.nan)
        sol[ok] = np.linalg.solve(A[ok], b[ok][..., None])[..., 0]
        tt, uu, vv = sol[..., 0], sol[..., 1], sol[..., 2]
        inside = ok & (uu >= 1e-6) & (vv >= 1e-6) & (uu + vv <= 1 - 1e-6) & (tt > 1e-3)
        nearest = np.minimum(nearest, np.where(inside, tt, np.inf).min(axis=1))
        for k in np.nonzero(inst == i)[0]:
            j = prim[k]
            assert ok[k, j]
            assert abs(tt[k, j] - t[k]) <= 2e-5 * t[k] + 1e-4, (k, tt[k, j], t[k])
            assert uu[k, j] >= -1e-5 and vv[k, j] >= -1e-5 and uu[k, j] + vv[k, j] <= 1 + 1e-5
    hit = inst >= 0
    assert hit.mean() > 0.5
    assert np.all(t[hit] <= nearest[hit] * (1 + 2e-5) + 1e-4)
    assert np.all(np.isinf(nearest[~hit]))
