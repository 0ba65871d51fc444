"""The pre-built upload path of the F# host (bn_upload_blas / bn_upload_scene)
against the scene build (bn_scene_build), through the C ABI.

A scene built by bn_scene_build is exported (TLAS, instance records, every
BLAS with its alias pdfs, camera-to-world) and handed back through the upload
entry points, as a host that builds its BVHs itself with the reference code
would hand them over.  Films, normal views and first hits must be
bit-identical.  Also: the alias pdfs are used verbatim, invalid arguments
fail with BN_E_INVALID, released BLAS cannot be referenced."""
import ctypes as C

import numpy as np
import pytest

from barnacle_amd import abi, configs, native, scene

pytestmark = pytest.mark.gpu


def reupload(ctx, sc, ds, pdf_scale=None):
    ex = ds.export()
    insts = ds.export_instances()
    ids = {}
    for rec in insts:
        if rec.blas < 0 or rec.blas in ids:
            continue
        prim = rec.blas
        b = ds.export_blas(prim)
        pr = sc.primitives[prim]
        v = sc.vertices.reshape(-1, 3)[pr["vert_offset"]:pr["vert_offset"] + pr["n_verts"]]
        pdf = b["alias_pdf"] if pdf_scale is None else b["alias_pdf"] * np.float32(pdf_scale)
        ids[prim] = ctx.upload_blas(b["nodes"], v, b["tri"].reshape(-1, 3), pdf)
    for rec in insts:
        if rec.blas >= 0:
            rec.blas = ids[rec.blas]
    return ctx.upload_scene(sc, ex["tlas"], insts, ex["c2w"]), ids


CASES = {
    "config0": lambda: configs.config0(),
    "config1": lambda: configs.config1(width=96, height=80, spp=4, freq=24),
    "spheres": lambda: configs.sphere_scene(64, 48, 4, 8),
    "config3": lambda: configs.config3(width=64, height=64, spp=2, freq=16, n_patches=256),
}


@pytest.mark.parametrize("name", list(CASES))
def test_upload_round_trip_is_bit_identical(ctx, name):
    sc = scene.load_scene(CASES[name]())
    print(f"upload round trip: {name}", flush=True)
    built = ctx.scene(sc)
    up, _ = reupload(ctx, sc, built)
    p = sc.pt_params()
    a, b = built.render_pt(p), up.render_pt(p)
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    assert np.isfinite(a).mean() > 0.4 and np.abs(a[np.isfinite(a)]).sum() > 0
    for k in (0, 1):
        fa, fb = built.first_hit(p, k), up.first_hit(p, k)
        for x, y in zip(fa, fb):
            assert np.array_equal(np.asarray(x).view(np.int32), np.asarray(y).view(np.int32))
    na, nb = built.render_normal(p), up.render_normal(p)
    assert np.array_equal(na, nb)
    up.release()
    built.release()


def test_uploaded_alias_pdfs_are_used_verbatim(ctx):
    """The light triangle pdfs a host passes are its AliasTable.Table[i].pdf
    (Mesh.fs:181-186, read at :290 and :300-304): scaled by 1 + 2^-10 they
    must change the film (a table rebuilt from them would renormalise the
    scale away)."""
    sc = scene.load_scene(configs.config0())
    built = ctx.scene(sc)
    p = sc.pt_params(width=64, height=64, spp=2)
    same, _ = reupload(ctx, sc, built)
    scaled, _ = reupload(ctx, sc, built, pdf_scale=1.0 + 2.0 ** -10)
    a, b, c = built.render_pt(p), same.render_pt(p), scaled.render_pt(p)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_upload_rejects_bad_arguments(ctx):
    sc = scene.load_scene(configs.config0())
    built = ctx.scene(sc)
    ex = built.export()
    insts = built.export_instances()
    lib = ctx.lib
    b = built.export_blas(insts[0].blas)
    pr = sc.primitives[insts[0].blas]
    v = np.ascontiguousarray(sc.vertices.reshape(-1, 3)[pr["vert_offset"]:pr["vert_offset"] + pr["n_verts"]],
                             np.float32)
    blas = ctx.upload_blas(b["nodes"], v, b["tri"].reshape(-1, 3), b["alias_pdf"])
    from barnacle_amd.bvh_util import numpy_to_nodes
    tl = numpy_to_nodes(ex["tlas"])
    recs = (abi.InstanceUpload * len(insts))(*insts)
    for r in recs:
        r.blas = blas if r.blas >= 0 else -1
    desc, keep = sc.desc()
    c2w = np.ascontiguousarray(ex["c2w"], np.float32)
    out = C.c_int32()

    def upload(n_mat=desc.n_materials, mats=desc.materials, n_light=desc.n_lights, lights=desc.lights):
        return lib.bn_upload_scene(ctx.h, tl, len(ex["tlas"]), recs, len(insts), mats, n_mat, lights, n_light,
                                   C.byref(desc.camera), c2w.ctypes.data_as(C.POINTER(C.c_float)), C.byref(out))

    assert upload() == 0
    assert upload(n_mat=-1) == abi.BN_E_INVALID
    assert upload(n_light=-2) == abi.BN_E_INVALID
    assert upload(mats=None) == abi.BN_E_INVALID
    assert upload(lights=None) == abi.BN_E_INVALID
    assert b"bad arguments" in lib.bn_last_error()
    # negative vertex count, null vertices
    nd = numpy_to_nodes(b["nodes"])
    tri = np.ascontiguousarray(b["tri"], np.int32)
    ob = C.c_int32()
    assert lib.bn_upload_blas(ctx.h, nd, len(b["nodes"]), v.ctypes.data_as(C.POINTER(C.c_float)), -3,
                              tri.ctypes.data_as(C.POINTER(C.c_int32)), len(tri) // 3, None,
                              C.byref(ob)) == abi.BN_E_INVALID
    assert lib.bn_upload_blas(ctx.h, nd, len(b["nodes"]), None, len(v),
                              tri.ctypes.data_as(C.POINTER(C.c_int32)), len(tri) // 3, None,
                              C.byref(ob)) == abi.BN_E_INVALID
    # a released BLAS can neither be referenced nor released again
    ctx.release_blas(blas)
    assert upload() == abi.BN_E_INVALID
    assert b"unknown BLAS" in lib.bn_last_error()
    with pytest.raises(native.BarnacleError):
        ctx.release_blas(blas)
    del keep
    # the context stays usable
    p = sc.pt_params(width=32, height=32, spp=1)
    assert np.isfinite(built.render_pt(p)).all()


def test_scene_release_frees_its_blas_and_keeps_others(ctx):
    """bn_scene_release drops the BLAS its build made; another scene of the
    same description renders on, bit-identical to a fresh build."""
    sc = scene.load_scene(configs.config1(width=48, height=48, spp=2, freq=12))
    p = sc.pt_params()
    a = ctx.scene(sc)
    b = ctx.scene(sc)
    fa = a.render_pt(p)
    a.release()
    with pytest.raises(native.BarnacleError):
        a.render_pt(p)
    assert np.array_equal(fa, b.render_pt(p))
    c = ctx.scene(sc)
    assert np.array_equal(fa, c.render_pt(p))


def test_pbr_given_alpha_matches_pbr_given_roughness(ctx):
    """A host holding constructed PBRMaterial objects has Alpha =
    Max(roughness^2, 1e-3) (PBR.fs:11), not roughness: uploaded as
    BN_MAT_PBR_ALPHA with that alpha, the scene renders bit-identically to
    the BN_MAT_PBR scene of the same roughness."""
    sc = scene.load_scene(CASES["config1"]())
    built = ctx.scene(sc)
    p = sc.pt_params()
    ref = built.render_pt(p)
    mats = sc.materials
    alt = [dict(m) for m in mats]
    for m in alt:
        if m["kind"] == 3:
            r = np.float32(m["roughness"])
            m["kind"] = 4
            m["roughness"] = max(np.float32(r * r), np.float32(1e-3))
    sc.materials = alt
    try:
        up, _ = reupload(ctx, sc, built)
    finally:
        sc.materials = mats
    assert any(m["kind"] == 4 for m in alt)
    assert np.array_equal(up.render_pt(p).view(np.int32), ref.view(np.int32))
