"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the barnacle_amd package.  The oracle
takes the same flat scene description as the product (see oracle.h); the
descriptor structs are marshalled with the ctypes classes of barnacle_amd.abi,
whose layout equals the oracle's (tests/test_abi.py checks both libraries).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from barnacle_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)


class OracleStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("rays_closest", "rays_shadow", "nodes_closest", "tris_closest",
                                          "insts_closest", "nodes_shadow", "tris_shadow", "insts_shadow",
                                          "vertices", "paths")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


def build():
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    def __init__(self, path: str = LIB_PATH):
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        L = self.lib = C.CDLL(path)
        L.orc_last_error.restype = C.c_char_p
        L.orc_xxhash32_two.restype = C.c_uint32
        L.orc_xxhash32_two.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_xxhash32_three.restype = C.c_uint32
        L.orc_xxhash32_three.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_lcg.restype = C.c_float
        L.orc_lcg.argtypes = [u32p]
        L.orc_erfinv.restype = C.c_float
        L.orc_erfinv.argtypes = [C.c_float]
        L.orc_alias_sample.argtypes = [i32p, f32p, f32p, C.c_int32, C.c_float, i32p, f32p]
        L.orc_scene_create.argtypes = [C.POINTER(abi.SceneDesc), C.POINTER(C.c_void_p)]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_render.argtypes = [C.c_void_p, C.c_int32, C.POINTER(abi.PTParams), C.c_int32, f32p,
                                 C.POINTER(OracleStats)] + [C.c_int32] * 4
        L.orc_trace_sample.argtypes = [C.c_void_p, C.POINTER(abi.PTParams), C.c_int32, C.c_int32, C.c_int32, f32p]
        L.orc_first_hit.argtypes = [C.c_void_p, C.POINTER(abi.PTParams), C.c_int32, i32p, i32p, f32p]
        L.orc_camera_ray.argtypes = [C.c_void_p, C.POINTER(abi.PTParams), C.c_int32, C.c_int32, C.c_int32, f32p]
        f64p = C.POINTER(C.c_double)
        L.orc_set_params.argtypes = [C.c_void_p, f32p, C.c_int32, f32p, C.c_int32]
        L.orc_adjoint_pt.argtypes = [C.c_void_p, C.POINTER(abi.PTParams), f32p, C.c_int32, C.c_int32, C.c_int32,
                                     C.c_int32, C.c_int32, f64p, f64p]
        L.orc_pbr_grad_d.argtypes = [f64p, C.c_double, C.c_double, f64p, f64p, f64p, f64p, f64p]
        L.orc_render_pssmlt.argtypes = [C.c_void_p, C.POINTER(abi.MLTParams), C.c_int32, f32p,
                                        C.POINTER(abi.MLTStats), C.c_int32, C.c_int32, f32p, i32p]
        L.orc_scene_export.argtypes = [C.c_void_p, C.POINTER(abi.BVHNode), i32p, i32p, i32p, f32p, f32p,
                                       i32p, i32p, f32p]
        L.orc_scene_export_blas.argtypes = [C.c_void_p, C.c_int32, C.POINTER(abi.BVHNode), i32p, i32p, i32p, f32p]
        L.orc_bvh_build.argtypes = [f32p, C.c_int32, C.POINTER(abi.BVHNode), i32p, i32p]
        L.orc_material_sample.argtypes = [C.POINTER(abi.Material), f32p, C.c_float, f32p, f32p]
        L.orc_material_eval.argtypes = [C.POINTER(abi.Material), f32p, f32p, f32p]
        L.orc_material_eval_n.argtypes = [C.POINTER(abi.Material), C.c_int32, f32p, f32p, f32p]
        L.orc_camera_ray_u.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, f32p, f32p, f32p]

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError("oracle: " + self.lib.orc_last_error().decode())

    # -- Hash.fs / Sampler.fs / AliasTable.fs ---------------------------
    def xxhash2(self, x, y):
        return self.lib.orc_xxhash32_two(x, y)

    def xxhash3(self, x, y, z):
        return self.lib.orc_xxhash32_three(x, y, z)

    def lcg(self, state: int, n: int):
        s = C.c_uint32(state)
        out = [self.lib.orc_lcg(C.byref(s)) for _ in range(n)]
        return np.array(out, np.float32), s.value

    def pbr_grad(self, albedo, alpha, metallic, wo, wi):
        """PBRMaterial.Eval in double: (f + [pdf], d f / d albedo, d f / d alpha + [d pdf / d alpha])."""
        f64p = C.POINTER(C.c_double)
        a = np.ascontiguousarray(albedo, np.float64)
        o = np.ascontiguousarray(wo, np.float64)
        i = np.ascontiguousarray(wi, np.float64)
        f, da, dal = np.zeros(4), np.zeros(3), np.zeros(4)
        self._check(self.lib.orc_pbr_grad_d(a.ctypes.data_as(f64p), float(alpha), float(metallic),
                                            o.ctypes.data_as(f64p), i.ctypes.data_as(f64p), f.ctypes.data_as(f64p),
                                            da.ctypes.data_as(f64p), dal.ctypes.data_as(f64p)))
        return f, da, dal

    def erfinv(self, x):
        return np.float32(self.lib.orc_erfinv(float(x)))

    def alias_build(self, w):
        w = np.ascontiguousarray(w, np.float32)
        n = len(w)
        al, pr, pd = np.zeros(n, np.int32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        self._check(self.lib.orc_alias_build(_ptr(w, f32p), n, _ptr(al, i32p), _ptr(pr, f32p), _ptr(pd, f32p)))
        return al, pr, pd

    def alias_sample(self, table, u):
        al, pr, pd = table
        i, p = C.c_int32(), C.c_float()
        self.lib.orc_alias_sample(_ptr(al, i32p), _ptr(pr, f32p), _ptr(pd, f32p), len(al), u, C.byref(i), C.byref(p))
        return i.value, np.float32(p.value)

    def bvh_build(self, bounds):
        b = np.ascontiguousarray(bounds, np.float32).reshape(-1, 6)
        n = len(b)
        nodes = (abi.BVHNode * (2 * n + 1))()
        nn = C.c_int32(2 * n + 1)
        perm = np.zeros(n, np.int32)
        self._check(self.lib.orc_bvh_build(_ptr(b, f32p), n, nodes, C.byref(nn), _ptr(perm, i32p)))
        return nodes_to_numpy(nodes, nn.value), perm

    def material_sample(self, mat: abi.Material, wo, u_lobe, u2):
        out = np.zeros(7, np.float32)
        wo = np.asarray(wo, np.float32)
        u2 = np.asarray(u2, np.float32)
        self.lib.orc_material_sample(C.byref(mat), _ptr(wo, f32p), u_lobe, _ptr(u2, f32p), _ptr(out, f32p))
        return out

    def material_eval(self, mat: abi.Material, wo, wi):
        out = np.zeros(4, np.float32)
        wo = np.asarray(wo, np.float32)
        wi = np.asarray(wi, np.float32)
        self.lib.orc_material_eval(C.byref(mat), _ptr(wo, f32p), _ptr(wi, f32p), _ptr(out, f32p))
        return out

    def material_eval_n(self, mat: abi.Material, wo, wi):
        """Eval over arrays of directions (n, 3): (n, 4) = bsdf rgb, pdf."""
        wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
        wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
        out = np.zeros((len(wo), 4), np.float32)
        self.lib.orc_material_eval_n(C.byref(mat), len(wo), _ptr(wo, f32p), _ptr(wi, f32p), _ptr(out, f32p))
        return out

    def scene(self, sc) -> "OracleScene":
        return OracleScene(self, sc)


def nodes_to_numpy(nodes, n):
    """BVH nodes as a numpy structured array (the 32-byte Util/BVH.fs layout)."""
    dt = np.dtype([("pmin", np.float32, 3), ("pmax", np.float32, 3), ("right_or_offset", np.int32),
                   ("is_leaf", np.uint8), ("split_axis", np.int8), ("count", np.int8), ("mask", np.uint8)])
    raw = C.string_at(C.addressof(nodes), 32 * n)
    return np.frombuffer(raw, dt).copy()


class OracleScene:
    def __init__(self, orc: Oracle, sc):
        self.o, self.sc = orc, sc
        desc, self._keep = sc.desc()
        h = C.c_void_p()
        orc._check(orc.lib.orc_scene_create(C.byref(desc), C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            self.o.lib.orc_scene_destroy(self.h)
        except Exception:
            pass

    def render(self, params: abi.PTParams, integrator: int = 2, threads: int = 0, window=None):
        """Film (Film.Pixels order, (H*W,3)) and traversal counters."""
        threads = threads or os.cpu_count() or 1
        film = np.zeros((params.width * params.height, 3), np.float32)
        st = OracleStats()
        x0, y0, x1, y1 = window if window else (0, 0, 0, 0)
        self.o._check(self.o.lib.orc_render(self.h, integrator, C.byref(params), threads, _ptr(film, f32p),
                                            C.byref(st), x0, y0, x1, y1))
        return film, st.as_dict()

    def trace_sample(self, params, x, y, k):
        out = np.zeros(3, np.float32)
        self.o._check(self.o.lib.orc_trace_sample(self.h, C.byref(params), x, y, k, _ptr(out, f32p)))
        return out

    def set_params(self, mat_params=None, light_emission=None):
        mp = None if mat_params is None else np.ascontiguousarray(mat_params, np.float32).reshape(-1, 4)
        le = None if light_emission is None else np.ascontiguousarray(light_emission, np.float32).reshape(-1, 3)
        self.o._check(self.o.lib.orc_set_params(self.h, None if mp is None else _ptr(mp, f32p),
                                                0 if mp is None else len(mp),
                                                None if le is None else _ptr(le, f32p), 0 if le is None else len(le)))

    def adjoint_pt(self, params, d_film, threads: int = 0, window=None):
        """FP64 gradients of sum(d_film * film): ((n_mat, 4), (n_light, 3))."""
        threads = threads or os.cpu_count() or 1
        df = np.ascontiguousarray(d_film, np.float32).reshape(-1)
        dm = np.zeros((max(1, len(self.sc.materials)), 4), np.float64)
        dl = np.zeros((max(1, len(self.sc.lights)), 3), np.float64)
        x0, y0, x1, y1 = window if window else (0, 0, 0, 0)
        f64p = C.POINTER(C.c_double)
        self.o._check(self.o.lib.orc_adjoint_pt(self.h, C.byref(params), _ptr(df, f32p), threads, x0, y0, x1, y1,
                                                dm.ctypes.data_as(f64p), dl.ctypes.data_as(f64p)))
        return dm[:len(self.sc.materials)], dl[:len(self.sc.lights)]

    def camera_ray(self, params, x, y, sample_id=0):
        out = np.zeros(6, np.float32)
        self.o._check(self.o.lib.orc_camera_ray(self.h, C.byref(params), x, y, sample_id, _ptr(out, f32p)))
        return out[:3], out[3:]

    def camera_ray_u(self, width, height, x, y, u_pix, u_lens):
        """Camera ray of pixel (x, y) for explicit uPixel / uLens draws."""
        out = np.zeros(6, np.float32)
        up = np.ascontiguousarray(u_pix, np.float32)
        ul = np.ascontiguousarray(u_lens, np.float32)
        self.o._check(self.o.lib.orc_camera_ray_u(self.h, width, height, x, y, _ptr(up, f32p), _ptr(ul, f32p),
                                                  _ptr(out, f32p)))
        return out[:3], out[3:]

    def first_hit(self, params, sample_id=0):
        n = params.width * params.height
        inst, prim, t = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.float32)
        self.o._check(self.o.lib.orc_first_hit(self.h, C.byref(params), sample_id, _ptr(inst, i32p),
                                               _ptr(prim, i32p), _ptr(t, f32p)))
        return inst, prim, t

    def render_pssmlt(self, params: abi.MLTParams, threads: int = 0, chains=None):
        threads = threads or os.cpu_count() or 1
        film = np.zeros((params.width * params.height, 3), np.float32)
        st = abi.MLTStats()
        boot = np.zeros(params.n_bootstrap, np.float32)
        acc = np.zeros(params.n_chains, np.int32)
        c0, c1 = chains if chains else (0, 0)
        self.o._check(self.o.lib.orc_render_pssmlt(self.h, C.byref(params), threads, _ptr(film, f32p),
                                                   C.byref(st), c0, c1, _ptr(boot, f32p), _ptr(acc, i32p)))
        return film, {"accepted": st.accepted, "proposed": st.proposed, "b": np.float32(st.b),
                      "mutations_per_chain": st.mutations_per_chain}, boot, acc

    def export(self):
        cap = 4 * max(1, len(self.sc.instances)) * 64 + 16
        tl = (abi.BVHNode * cap)()
        ntl, ni, nl = C.c_int32(cap), C.c_int32(cap), C.c_int32(cap)
        order = np.zeros(cap, np.int32)
        o2w, w2o = np.zeros((cap, 16), np.float32), np.zeros((cap, 16), np.float32)
        lights = np.zeros(cap, np.int32)
        c2w = np.zeros(16, np.float32)
        self.o._check(self.o.lib.orc_scene_export(self.h, tl, C.byref(ntl), _ptr(order, i32p), C.byref(ni),
                                                  _ptr(o2w, f32p), _ptr(w2o, f32p), _ptr(lights, i32p),
                                                  C.byref(nl), _ptr(c2w, f32p)))
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
        self.o._check(self.o.lib.orc_scene_export_blas(self.h, prim, nodes, C.byref(nn), _ptr(tri, i32p),
                                                       C.byref(ntr), _ptr(pdf, f32p)))
        return {"nodes": nodes_to_numpy(nodes, nn.value), "tri": tri, "alias_pdf": pdf}
