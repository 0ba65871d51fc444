"""ctypes binding of libbarnacle_hip.so (the C ABI of include/barnacle_hip.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the library has not been built, and every call raises BarnacleError with
the library's message on a non-zero status (the reference raises F#
exceptions at the same points).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbarnacle_hip.so")

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)


class BarnacleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_LIB = None

_SIGS = {
    "bn_create": ([C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "bn_destroy": ([C.c_void_p], None),
    "bn_last_error": ([], C.c_char_p),
    "bn_abi_version": ([], C.c_int),
    "bn_struct_sizes": ([i32p, C.c_int32], C.c_int),
    "bn_bvh_build": ([f32p, C.c_int32, C.POINTER(abi.BVHNode), i32p, i32p], C.c_int),
    "bn_bvh_build_gpu": ([C.c_void_p, f32p, C.c_int32, C.POINTER(abi.BVHNode), i32p, i32p], C.c_int),
    "bn_scene_build": ([C.c_void_p, C.POINTER(abi.SceneDesc), i32p], C.c_int),
    "bn_scene_release": ([C.c_void_p, C.c_int32], C.c_int),
    "bn_scene_export": ([C.c_void_p, C.c_int32, C.POINTER(abi.BVHNode), i32p, i32p, i32p, f32p, f32p,
                         i32p, i32p, f32p], C.c_int),
    "bn_scene_export_blas": ([C.c_void_p, C.c_int32, C.c_int32, C.POINTER(abi.BVHNode), i32p, i32p, i32p,
                              f32p], C.c_int),
    "bn_upload_blas": ([C.c_void_p, C.POINTER(abi.BVHNode), C.c_int32, f32p, C.c_int32, i32p, C.c_int32,
                        f32p, i32p], C.c_int),
    "bn_upload_scene": ([C.c_void_p, C.POINTER(abi.BVHNode), C.c_int32, C.POINTER(abi.InstanceUpload),
                         C.c_int32, C.POINTER(abi.Material), C.c_int32, C.POINTER(abi.Light), C.c_int32,
                         C.POINTER(abi.Camera), f32p, i32p], C.c_int),
    "bn_release_blas": ([C.c_void_p, C.c_int32], C.c_int),
    "bn_scene_export_instances": ([C.c_void_p, C.c_int32, C.POINTER(abi.InstanceUpload), i32p], C.c_int),
    "bn_film_postprocess": ([f32p, C.c_int64, C.c_int32, f32p, C.POINTER(C.c_uint8)], C.c_int),
    "bn_render_pt": ([C.c_void_p, C.c_int32, C.POINTER(abi.PTParams), f32p, C.POINTER(abi.TraceStats)], C.c_int),
    "bn_render_normal": ([C.c_void_p, C.c_int32, C.POINTER(abi.PTParams), f32p], C.c_int),
    "bn_render_direct": ([C.c_void_p, C.c_int32, C.POINTER(abi.PTParams), f32p], C.c_int),
    "bn_first_hit": ([C.c_void_p, C.c_int32, C.POINTER(abi.PTParams), C.c_int32, i32p, i32p, f32p], C.c_int),
    "bn_first_hit_noalloc": ([C.c_void_p, C.c_int32, C.POINTER(abi.PTParams), C.c_int32, i32p, i32p, f32p],
                             C.c_int),
    "bn_set_option": ([C.c_void_p, C.c_char_p, C.c_int64], C.c_int),
    "bn_kernel_times": ([C.c_void_p, C.c_void_p, i32p], C.c_int),
    "bn_reset_kernel_times": ([C.c_void_p], C.c_int),
    "bn_render_pssmlt": ([C.c_void_p, C.c_int32, C.POINTER(abi.MLTParams), f32p, C.POINTER(abi.MLTStats)],
                         C.c_int),
    "bn_render_batch": ([C.c_void_p, C.POINTER(abi.BatchItem), C.c_int32, C.POINTER(f32p),
                         C.POINTER(abi.BatchStats)], C.c_int),
    "bn_autotune": ([C.c_void_p, C.POINTER(abi.BatchItem), C.c_int32, C.c_int32, C.POINTER(abi.TuneResult)],
                    C.c_int),
    "bn_set_params": ([C.c_void_p, C.c_int32, f32p, C.c_int32, f32p, C.c_int32], C.c_int),
    "bn_backward_pt": ([C.c_void_p, C.c_int32, C.POINTER(abi.PTParams), f32p, f32p, f32p], C.c_int),
    "bn_pssmlt_last": ([C.c_void_p, f32p, C.c_int32, i32p, C.c_int32], C.c_int),
}


def load(path: str = LIB_PATH):
    """Load the HIP library (idempotent).  Raises if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"libbarnacle_hip.so is not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = lib
    return lib


def check(rc: int):
    if rc != 0:
        raise BarnacleError(rc, load().bn_last_error().decode())


def _p(a, t):
    return a.ctypes.data_as(t)


def bvh_build(bounds):
    """BVHNode.Build (Util/BVH.fs:239-247) on the host: (nodes, perm)."""
    lib = load()
    b = np.ascontiguousarray(bounds, np.float32).reshape(-1, 6)
    n = len(b)
    nodes = (abi.BVHNode * (2 * n + 1))()
    nn = C.c_int32(2 * n + 1)
    perm = np.zeros(n, np.int32)
    check(lib.bn_bvh_build(_p(b, f32p), n, nodes, C.byref(nn), _p(perm, i32p)))
    from .bvh_util import nodes_to_numpy
    return nodes_to_numpy(nodes, nn.value), perm


class Context:
    """One bn_ctx (device memory, stream) on one HIP device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.bn_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.bn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, key: str, value: int):
        check(self.lib.bn_set_option(self.h, key.encode(), int(value)))

    def bvh_build(self, bounds):
        """BVHNode.Build (Util/BVH.fs:128-247) on this context's GPU: (nodes, perm)."""
        b = np.ascontiguousarray(bounds, np.float32).reshape(-1, 6)
        n = len(b)
        nodes = (abi.BVHNode * (2 * n + 1))()
        nn = C.c_int32(2 * n + 1)
        perm = np.zeros(n, np.int32)
        check(self.lib.bn_bvh_build_gpu(self.h, _p(b, f32p), n, nodes, C.byref(nn), _p(perm, i32p)))
        from .bvh_util import nodes_to_numpy
        return nodes_to_numpy(nodes, nn.value), perm

    def kernel_times(self):
        """{kernel name: (launches, total ms)} accumulated since the last reset."""
        buf = (abi.KernelTime * 64)()
        n = C.c_int32(64)
        check(self.lib.bn_kernel_times(self.h, buf, C.byref(n)))
        return {buf[i].name.decode(): (int(buf[i].launches), float(buf[i].total_ms)) for i in range(n.value)}

    def reset_kernel_times(self):
        check(self.lib.bn_reset_kernel_times(self.h))

    def scene(self, sc) -> "DeviceScene":
        return DeviceScene(self, sc)

    def upload_blas(self, nodes, vertices, tri_indices, alias_pdf=None) -> int:
        """bn_upload_blas: a BLAS built by the caller (nodes in the 32-byte
        Util/BVH.fs layout, vertices (n, 3), triangle vertex indices in BVH
        order (t, 3), AliasTable pdfs per triangle or None)."""
        from .bvh_util import numpy_to_nodes
        nd = numpy_to_nodes(nodes)
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1)
        t = np.ascontiguousarray(tri_indices, np.int32).reshape(-1)
        pdf = None if alias_pdf is None else np.ascontiguousarray(alias_pdf, np.float32)
        out = C.c_int32()
        check(self.lib.bn_upload_blas(self.h, nd, len(nodes), _p(v, f32p), len(v) // 3, _p(t, i32p), len(t) // 3,
                                      None if pdf is None else _p(pdf, f32p), C.byref(out)))
        return out.value

    def release_blas(self, blas: int):
        check(self.lib.bn_release_blas(self.h, blas))

    def upload_scene(self, sc, tlas, instances, camera_to_world) -> "DeviceScene":
        """bn_upload_scene: TLAS nodes, bn_instance_upload records (TLAS
        order), the materials, lights and camera of scene `sc`; returns a
        DeviceScene over the uploaded id."""
        from .bvh_util import numpy_to_nodes
        tl = numpy_to_nodes(tlas)
        n = len(instances)
        ins = (abi.InstanceUpload * n)(*instances)
        desc, keep = sc.desc()
        c2w = np.ascontiguousarray(camera_to_world, np.float32)
        out = C.c_int32()
        check(self.lib.bn_upload_scene(self.h, tl, len(tlas), ins, n, desc.materials, desc.n_materials,
                                       desc.lights, desc.n_lights, C.byref(desc.camera), _p(c2w, f32p),
                                       C.byref(out)))
        del keep
        return DeviceScene(self, sc, scene_id=out.value)

    def render_batch(self, jobs):
        """jobs: list of (DeviceScene, PTParams).  One batched launch sequence;
        returns (films, {"rounds", "graph"})."""
        n = len(jobs)
        items = (abi.BatchItem * n)()
        films = [np.zeros((p.width * p.height, 3), np.float32) for _, p in jobs]
        ptrs = (f32p * n)()
        for k, (ds, p) in enumerate(jobs):
            items[k] = abi.BatchItem(ds.id, p)
            ptrs[k] = _p(films[k], f32p)
        st = abi.BatchStats()
        check(self.lib.bn_render_batch(self.h, items, n, ptrs, C.byref(st)))
        return films, {"rounds": st.rounds, "graph": st.graph, "scene_rounds": st.scene_rounds}


    def autotune(self, jobs, batched: bool = False):
        """Tune launch parameters on a workload of (DeviceScene, PTParams);
        the chosen setting stays on the context.  Returns it as a dict."""
        n = len(jobs)
        items = (abi.BatchItem * n)()
        for k, (ds, p) in enumerate(jobs):
            items[k] = abi.BatchItem(ds.id, p)
        r = abi.TuneResult()
        check(self.lib.bn_autotune(self.h, items, n, abi.BN_TUNE_BATCH if batched else 0, C.byref(r)))
        return {n_: getattr(r, n_) for n_, _ in abi.TuneResult._fields_}


class DeviceScene:
    """A scene built on the host (BLAS/TLAS) and uploaded to the device."""

    def __init__(self, ctx: Context, sc, scene_id: int | None = None):
        self.ctx, self.sc = ctx, sc
        if scene_id is None:
            desc, keep = sc.desc()
            sid = C.c_int32()
            check(ctx.lib.bn_scene_build(ctx.h, C.byref(desc), C.byref(sid)))
            scene_id = sid.value
        self.id = scene_id
        self.param_gen = 0

    def release(self):
        check(self.ctx.lib.bn_scene_release(self.ctx.h, self.id))

    def export_instances(self):
        """bn_instance_upload records of the built scene (TLAS order)."""
        n = C.c_int32(max(1, len(self.sc.instances)))
        out = (abi.InstanceUpload * n.value)()
        check(self.ctx.lib.bn_scene_export_instances(self.ctx.h, self.id, out, C.byref(n)))
        return [out[i] for i in range(n.value)]

    def render_pt(self, params: abi.PTParams, stats: bool = False, out=None, host: bool = True):
        """Film in Film.Pixels order, shape (H*W, 3); optional traversal counters.
        out: a C-contiguous float32 (H*W, 3) array to render into (as the
        reference renders into its existing Film.Pixels).  host=False: the
        film stays in device memory (nothing is copied; None is returned in
        its place)."""
        n = params.width * params.height
        if not host:
            st = abi.TraceStats() if stats else None
            check(self.ctx.lib.bn_render_pt(self.ctx.h, self.id, C.byref(params), None,
                                            C.byref(st) if st is not None else None))
            if stats:
                return None, {n: int(getattr(st, n)) for n, _ in abi.TraceStats._fields_}
            return None
        if out is None:
            film = np.zeros((n, 3), np.float32)
        else:
            if out.dtype != np.float32 or out.shape != (n, 3) or not out.flags.c_contiguous:
                raise ValueError("out must be a C-contiguous float32 array of shape (width*height, 3)")
            film = out
        st = abi.TraceStats() if stats else None
        check(self.ctx.lib.bn_render_pt(self.ctx.h, self.id, C.byref(params), _p(film, f32p),
                                        C.byref(st) if st is not None else None))
        if stats:
            return film, {n: int(getattr(st, n)) for n, _ in abi.TraceStats._fields_}
        return film

    def render_normal(self, params: abi.PTParams):
        film = np.zeros((params.width * params.height, 3), np.float32)
        check(self.ctx.lib.bn_render_normal(self.ctx.h, self.id, C.byref(params), _p(film, f32p)))
        return film

    def render_direct(self, params: abi.PTParams):
        """DirectIntegrator (Direct.fs:10-40): Film.Pixels order, (H*W, 3)."""
        film = np.zeros((params.width * params.height, 3), np.float32)
        check(self.ctx.lib.bn_render_direct(self.ctx.h, self.id, C.byref(params), _p(film, f32p)))
        return film

    def first_hit(self, params: abi.PTParams, sample_id: int = 0):
        n = params.width * params.height
        inst, prim, t = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.float32)
        check(self.ctx.lib.bn_first_hit(self.ctx.h, self.id, C.byref(params), sample_id, _p(inst, i32p),
                                        _p(prim, i32p), _p(t, f32p)))
        return inst, prim, t

    def first_hit_noalloc(self, params: abi.PTParams, sample_id: int = 0):
        """first_hit through the allocation-free entry point (static scratch)."""
        n = params.width * params.height
        inst, prim, t = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.float32)
        check(self.ctx.lib.bn_first_hit_noalloc(self.ctx.h, self.id, C.byref(params), sample_id, _p(inst, i32p),
                                                _p(prim, i32p), _p(t, f32p)))
        return inst, prim, t

    def render_pssmlt(self, params: abi.MLTParams):
        film = np.zeros((params.width * params.height, 3), np.float32)
        st = abi.MLTStats()
        check(self.ctx.lib.bn_render_pssmlt(self.ctx.h, self.id, C.byref(params), _p(film, f32p), C.byref(st)))
        return film, {"accepted": st.accepted, "proposed": st.proposed, "b": np.float32(st.b),
                      "mutations_per_chain": st.mutations_per_chain}

    def pssmlt_last(self, params: abi.MLTParams):
        """(bootstrap luminances, per-chain accept counts) of the last PSSMLT render."""
        boot = np.zeros(params.n_bootstrap, np.float32)
        acc = np.zeros(params.n_chains, np.int32)
        check(self.ctx.lib.bn_pssmlt_last(self.ctx.h, _p(boot, f32p), params.n_bootstrap, _p(acc, i32p),
                                          params.n_chains))
        return boot, acc

    def set_params(self, mat_params=None, light_emission=None):
        """Albedo rgb + roughness per material ((n_mat, 4)) and emission per
        light ((n_light, 3)); None keeps the current values."""
        mp = None if mat_params is None else np.ascontiguousarray(mat_params, np.float32).reshape(-1, 4)
        le = None if light_emission is None else np.ascontiguousarray(light_emission, np.float32).reshape(-1, 3)
        check(self.ctx.lib.bn_set_params(self.ctx.h, self.id, None if mp is None else _p(mp, f32p),
                                         0 if mp is None else len(mp), None if le is None else _p(le, f32p),
                                         0 if le is None else len(le)))
        self.param_gen += 1

    def backward_pt(self, params: abi.PTParams, d_film):
        """Gradients of sum(d_film * film): ((n_mat, 4), (n_light, 3))."""
        df = np.ascontiguousarray(d_film, np.float32).reshape(-1, 3)
        assert len(df) == params.width * params.height
        dm = np.zeros((max(1, len(self.sc.materials)), 4), np.float32)
        dl = np.zeros((max(1, len(self.sc.lights)), 3), np.float32)
        check(self.ctx.lib.bn_backward_pt(self.ctx.h, self.id, C.byref(params), _p(df, f32p), _p(dm, f32p),
                                          _p(dl, f32p)))
        return dm[:len(self.sc.materials)], dl[:len(self.sc.lights)]

    def export(self):
        cap = 4 * max(1, len(self.sc.instances)) * 64 + 16
        lib = self.ctx.lib
        tl = (abi.BVHNode * cap)()
        ntl, ni, nl = C.c_int32(cap), C.c_int32(cap), C.c_int32(cap)
        order = np.zeros(cap, np.int32)
        o2w, w2o = np.zeros((cap, 16), np.float32), np.zeros((cap, 16), np.float32)
        lights = np.zeros(cap, np.int32)
        c2w = np.zeros(16, np.float32)
        check(lib.bn_scene_export(self.ctx.h, self.id, tl, C.byref(ntl), _p(order, i32p), C.byref(ni),
                                  _p(o2w, f32p), _p(w2o, f32p), _p(lights, i32p), C.byref(nl), _p(c2w, f32p)))
        from .bvh_util import nodes_to_numpy
        n = ni.value
        return {"tlas": nodes_to_numpy(tl, ntl.value), "inst_order": order[:n], "o2w": o2w[:n], "w2o": w2o[:n],
                "lights": lights[:nl.value], "c2w": c2w}

    def export_blas(self, prim: int):
        p = self.sc.primitives[prim]
        nt = p["n_indices"] // 3
        nodes = (abi.BVHNode * (2 * nt + 1))()
        nn, ntr = C.c_int32(2 * nt + 1), C.c_int32(nt)
        tri = np.zeros(3 * nt, np.int32)
        pdf = np.zeros(nt, np.float32)
        check(self.ctx.lib.bn_scene_export_blas(self.ctx.h, self.id, prim, nodes, C.byref(nn), _p(tri, i32p),
                                                C.byref(ntr), _p(pdf, f32p)))
        from .bvh_util import nodes_to_numpy
        return {"nodes": nodes_to_numpy(nodes, nn.value), "tri": tri, "alias_pdf": pdf}
