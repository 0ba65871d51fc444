"""Scene front end: the reference's scene JSON schema to the flat bn_scene_desc.

Restates the defaults and construction rules of Extensions/Scene/Loader.fs:14-281
(keys are case-sensitive, rotations are radians, S*Rx*Ry*Rz*T per key frame)
and the OBJ reader of Extensions/Primitive/Mesh.fs:246-279.  The matrices,
BVHs and instance order are derived by the native host code (bn_scene_build),
not here.

A scene is given either as a path to a JSON file or as a dict in the same
schema; in a dict, "vertices"/"indices" of a mesh may be numpy arrays.
Extension keys (absent from the reference schema) are documented where read:
  integrator."mutation-sigma"  Gaussian sigma of PSSMLT (reference hard-codes
                               1e-2 at Loader.fs:195,198).
"""
from __future__ import annotations

import ctypes as C
import json
import os
from fractions import Fraction

import numpy as np

from . import abi

# Builtin meshes of Extensions/Primitive/Mesh.fs:126-155 (the data must be the
# same for the BLAS build to reproduce the reference's node order).
QUAD_VERTICES = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
QUAD_INDICES = np.array([0, 1, 2, 0, 2, 3], np.int32)
CUBE_VERTICES = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                          [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], np.float32)
CUBE_INDICES = np.array([0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4,
                         3, 7, 6, 3, 6, 2, 0, 4, 7, 0, 7, 3, 1, 2, 6, 1, 6, 5], np.int32)

MATERIAL_KINDS = {"lambertian": 0, "mirror": 1, "dielectric": 2, "pbr": 3}
INTEGRATOR_KINDS = ("normal", "direct", "path-tracing", "pssmlt")
TONE_MAPPINGS = ("aces", "gamma", "identity")


class SceneError(ValueError):
    """Raised where Loader.fs raises (failwith) on a malformed scene."""


def f32_exact(text: str) -> np.float32:
    """Decimal text to the nearest float32, rounding once (ties to even).

    System.Text.Json parses float32 fields directly; going through a double
    first can round twice.  Candidates around the double rounding are compared
    exactly with rational arithmetic."""
    q = Fraction(text)
    d = np.float32(float(q))
    best = None
    for c in (np.nextafter(d, np.float32(-np.inf)), d, np.nextafter(d, np.float32(np.inf))):
        if not np.isfinite(c):
            continue
        err = abs(Fraction(float(c)) - q)
        if best is None or err < best[0] or (err == best[0] and (int(c.view(np.uint32)) & 1) == 0):
            best = (err, c)
    return np.float32(best[1]) if best is not None else d


def _f(v, default):
    return np.float32(default) if v is None else np.float32(v)


def load_obj(path: str):
    """MeshPrimitive.Load (Mesh.fs:246-279): 'v x y z' and fan-triangulated
    'f a b c ...' records; a negative index counts back from the vertices read
    so far, a positive one is 1-based; '/...' suffixes are ignored."""
    verts, idx = [], []
    with open(path, "r", encoding="utf-8-sig") as fh:
        for raw in fh:
            line = raw.strip()
            if line.startswith("v "):
                tok = line[2:].split()
                verts.append([f32_exact(t) for t in tok[:3]])
            elif line.startswith("f "):
                tok = line[2:].split()

                def ix(t):
                    i = int(t.split("/", 1)[0])
                    return len(verts) + i if i < 0 else i - 1
                i0, i1 = ix(tok[0]), ix(tok[1])
                for t in tok[2:]:
                    i2 = ix(t)
                    idx += [i0, i1, i2]
                    i1 = i2
    return np.asarray(verts, np.float32).reshape(-1, 3), np.asarray(idx, np.int32)


class Scene:
    """Parsed scene: flat arrays ready for bn_scene_build / the oracle."""

    def __init__(self):
        self.vertices = np.zeros((0,), np.float32)
        self.indices = np.zeros((0,), np.int32)
        self.primitives, self.materials, self.lights = [], [], []
        self.instances, self.transforms, self.nodes = [], [], []
        self.node_refs = []
        self.root = 0
        self.time = np.float32(0.0)           # Scene.Render(t): the reference renders t = 0
        self.camera = None
        self.integrator = {}
        self.film = (0, 0, "aces")

    # -- parameters of the render calls --------------------------------
    def pt_params(self, **over) -> abi.PTParams:
        it = dict(self.integrator, **over)
        w, h = over.get("width", self.film[0]), over.get("height", self.film[1])
        return abi.PTParams(w, h, it["spp"], it["max_depth"], it["rr_depth"], it.get("frame_id", 0))

    def mlt_params(self, **over) -> abi.MLTParams:
        it = dict(self.integrator, **over)
        w, h = over.get("width", self.film[0]), over.get("height", self.film[1])
        return abi.MLTParams(w, h, it["spp"], it["max_depth"], it["rr_depth"], it.get("frame_id", 0),
                             it["n_bootstrap"], it["n_chains"], it["strategy"], it["sigma"],
                             it["eps_min"], it["eps_max"], it["large_step_prob"])

    @property
    def n_triangles(self) -> int:
        return sum(p["n_indices"] // 3 for p in self.primitives if p["kind"] == 0)

    # -- ctypes view -----------------------------------------------------
    def desc(self):
        """Return (bn_scene_desc, keepalive).  Keep `keepalive` referenced for
        as long as the descriptor is used."""
        keep = []

        def arr(ctype, rows, make):
            a = (ctype * max(1, len(rows)))()
            for i, r in enumerate(rows):
                a[i] = make(r)
            keep.append(a)
            return a, len(rows)

        v = np.ascontiguousarray(self.vertices, np.float32).reshape(-1)
        ix = np.ascontiguousarray(self.indices, np.int32).reshape(-1)
        keep += [v, ix]
        prims, n_p = arr(abi.Primitive, self.primitives, lambda p: abi.Primitive(
            p["kind"], p["vert_offset"], p["n_verts"], p["index_offset"], p["n_indices"], p["radius"]))
        mats, n_m = arr(abi.Material, self.materials, lambda m: abi.Material(
            m["kind"], (C.c_float * 3)(*m["albedo"]), m["ior"], m["roughness"], m["metallic"]))
        lights, n_l = arr(abi.Light, self.lights, lambda l: abi.Light(
            0, (C.c_float * 3)(*l["emission"]), int(l["two_sided"])))
        insts, n_i = arr(abi.Instance, self.instances, lambda i: abi.Instance(*i))
        xfs, n_x = arr(abi.Transform, self.transforms, lambda t: abi.Transform(
            t["flags"], (C.c_float * 3)(*t["scale"]), (C.c_float * 3)(*t["rotation"]),
            (C.c_float * 3)(*t["translation"]), t["time"], t.get("n_keys", 0)))
        nodes, n_n = arr(abi.Node, self.nodes, lambda n: abi.Node(*n))
        refs = np.asarray(self.node_refs if self.node_refs else [0], np.int32)
        keep.append(refs)
        cam = self.camera
        d = abi.SceneDesc(
            v.ctypes.data_as(C.POINTER(C.c_float)), v.size,
            ix.ctypes.data_as(C.POINTER(C.c_int32)), ix.size,
            prims, n_p, mats, n_m, lights, n_l, insts, n_i, xfs, n_x, nodes, n_n,
            refs.ctypes.data_as(C.POINTER(C.c_int32)), len(self.node_refs), self.root,
            abi.Camera(cam["kind"], cam["fov"], cam["aspect"], cam["aperture"],
                       cam["focus_distance"], cam["push_forward"]), float(self.time))
        return d, keep


def _keyframe(kf) -> dict:
    """KeyFrameInfo.ToKeyFrame (Loader.fs:20-41): absent parts are identity."""
    flags, s, r, t = 0, [1, 1, 1], [0, 0, 0], [0, 0, 0]
    if kf.get("scale") is not None:
        flags |= 1
        s = kf["scale"]
    if kf.get("rotation") is not None:
        flags |= 2
        r = kf["rotation"]
    if kf.get("translation") is not None:
        flags |= 4
        t = kf["translation"]
    return {"time": _f(kf.get("time"), 0.0), "flags": flags,
            "scale": [np.float32(x) for x in s[:3]],
            "rotation": [np.float32(x) for x in r[:3]],
            "translation": [np.float32(x) for x in t[:3]]}


def _keyframes(keyframes) -> list:
    """TransformInfo.ToTransform (Loader.fs:44-47): the key frames sorted by
    time.  Array.sortBy is not stable, so the reference orders equal times
    arbitrarily; here they keep their order in the file.  The native scene
    build evaluates them with Transform.Eval (Base/Scene.fs:27-36),
    interpolating between key frames (:13-19)."""
    if not keyframes:
        raise SceneError("KeyFrames cannot be empty")
    ks = sorted((_keyframe(k) for k in keyframes), key=lambda k: k["time"])
    ks[0]["n_keys"] = len(ks)
    return ks


def load_scene(src, base_dir: str | None = None) -> Scene:
    """Scene.Load (Loader.fs:277-281) followed by SceneInfo.ToScene (:253-272)."""
    if isinstance(src, (str, os.PathLike)):
        base_dir = base_dir or os.path.dirname(os.path.abspath(src))
        with open(src, "r", encoding="utf-8-sig") as fh:
            info = json.load(fh, parse_float=f32_exact)
    else:
        info = src
    base_dir = base_dir or os.getcwd()
    sc = Scene()

    verts, idxs, vcount, icount = [], [], 0, 0
    for p in info.get("primitives") or []:
        ty = p.get("type")
        if ty == "sphere":
            sc.primitives.append(dict(kind=1, vert_offset=0, n_verts=0, index_offset=0,
                                      n_indices=0, radius=_f(p.get("radius"), 1.0)))
            continue
        if ty == "quad":
            v, ix = QUAD_VERTICES, QUAD_INDICES
        elif ty == "cube":
            v, ix = CUBE_VERTICES, CUBE_INDICES
        elif ty == "mesh":
            if p.get("uri") is not None:
                uri = p["uri"] if os.path.isabs(p["uri"]) else os.path.join(base_dir, p["uri"])
                v, ix = load_obj(uri)
            elif p.get("vertices") is not None and p.get("indices") is not None:
                v = np.asarray(p["vertices"], np.float32).reshape(-1, 3)
                ix = np.asarray(p["indices"], np.int32).reshape(-1)
            else:
                raise SceneError("Invalid mesh primitive")
        else:
            raise SceneError(f"Unknown primitive type: {ty}")
        if ix.size % 3 != 0:
            raise SceneError("mesh index count is not a multiple of 3")
        if ix.size and (ix.min() < 0 or ix.max() >= len(v)):
            raise SceneError("mesh index out of range")
        sc.primitives.append(dict(kind=0, vert_offset=vcount, n_verts=len(v), index_offset=icount,
                                  n_indices=int(ix.size), radius=np.float32(0)))
        verts.append(np.asarray(v, np.float32).reshape(-1))
        idxs.append(np.asarray(ix, np.int32))
        vcount += len(v)
        icount += ix.size
    sc.vertices = np.concatenate(verts) if verts else np.zeros((0,), np.float32)
    sc.indices = np.concatenate(idxs) if idxs else np.zeros((0,), np.int32)

    for m in info.get("materials") or []:
        ty = m.get("type")
        if ty not in MATERIAL_KINDS:
            raise SceneError(f"Unknown material type: {ty}")
        alb = m.get("albedo")
        sc.materials.append(dict(kind=MATERIAL_KINDS[ty],
                                 albedo=[np.float32(x) for x in alb[:3]] if alb is not None else [np.float32(1)] * 3,
                                 ior=_f(m.get("ior"), 1.5), roughness=_f(m.get("roughness"), 1.0),
                                 metallic=_f(m.get("metallic"), 0.0)))
    for l in info.get("lights") or []:
        if l.get("type") != "diffuse":
            raise SceneError(f"Unknown light type: {l.get('type')}")
        em = l.get("emission")
        ts = l.get("two-sided")
        sc.lights.append(dict(emission=[np.float32(x) for x in em[:3]] if em is not None else [np.float32(0)] * 3,
                              two_sided=True if ts is None else bool(ts)))
    n_prim, n_mat, n_light = len(sc.primitives), len(sc.materials), len(sc.lights)
    for i in info.get("instances") or []:
        p = i["primitive"]
        m = -1 if i.get("material") is None else int(i["material"])
        l = -1 if i.get("light") is None else int(i["light"])
        if not (0 <= p < n_prim) or m >= n_mat or l >= n_light or m < -1 or l < -1:
            raise SceneError("instance references an undefined primitive/material/light")
        sc.instances.append((int(p), m, l))
    first_key = []
    for t in info.get("transforms") or []:
        first_key.append(len(sc.transforms))
        sc.transforms += _keyframes(t.get("keyframes") or [])

    nodes = info.get("nodes") or []
    for n in nodes:
        inst = list(n.get("instances") or [])
        ch = list(n.get("children") or [])
        if any(not (0 <= k < len(sc.instances)) for k in inst) or any(not (0 <= k < len(nodes)) for k in ch):
            raise SceneError("node references an undefined instance/node")
        tr = -1 if n.get("transform") is None else int(n["transform"])
        if tr >= len(first_key) or tr < -1:
            raise SceneError("node references an undefined transform")
        tr = first_key[tr] if tr >= 0 else -1
        io = len(sc.node_refs)
        sc.node_refs += inst
        co = len(sc.node_refs)
        sc.node_refs += ch
        sc.nodes.append((tr, int(bool(n.get("has-camera") or False)), io, len(inst), co, len(ch)))
    sc.root = int(info.get("root") or 0)
    if not (0 <= sc.root < len(nodes)):
        raise SceneError("root node out of range")

    it = info.get("integrator") or {}
    ity = it.get("type")
    if ity not in INTEGRATOR_KINDS:
        raise SceneError(f"Unknown integrator type: {ity}")
    ms = it.get("mutation-strategy")
    if ms not in (None, "Gaussian", "Kelemen"):
        raise SceneError(f"Unknown mutation strategy: {ms}")
    sc.integrator = dict(
        type=ity, spp=int(it["spp"]) if it.get("spp") is not None else 1,
        max_depth=int(it["max-depth"]) if it.get("max-depth") is not None else 8,
        rr_depth=int(it["rr-depth"]) if it.get("rr-depth") is not None else 5,
        n_bootstrap=int(it["n-bootstrap"]) if it.get("n-bootstrap") is not None else 4 * 1024 * 1024,
        n_chains=int(it["n-chains"]) if it.get("n-chains") is not None else 1024,
        strategy=1 if ms == "Kelemen" else 0,
        sigma=_f(it.get("mutation-sigma"), 1e-2),
        eps_min=np.float32(1.0) / np.float32(1024.0), eps_max=np.float32(1.0) / np.float32(16.0),
        large_step_prob=_f(it.get("large-step-prob"), 0.5))

    cam = info.get("camera") or {}
    cty = cam.get("type")
    if cty not in ("pinhole", "thin-lens"):
        raise SceneError(f"Unknown camera type: {cty}")
    sc.camera = dict(kind=0 if cty == "pinhole" else 1, fov=_f(cam.get("fov"), 45.0),
                     aspect=_f(cam.get("aspect-ratio"), 1.0), aperture=_f(cam.get("aperture"), 0.0),
                     focus_distance=_f(cam.get("focus-distance"), 1.0),
                     push_forward=_f(cam.get("push-forward"), 0.0))
    film = info.get("film") or {}
    if film.get("tone-mapping") not in TONE_MAPPINGS:
        raise SceneError(f"Unknown tone mapping type: {film.get('tone-mapping')}")
    sc.film = (int(film["width"]), int(film["height"]), film["tone-mapping"])
    return sc


def to_json(info: dict) -> str:
    """Serialise a scene dict (numpy arrays allowed) in the reference schema."""
    def conv(o):
        if isinstance(o, np.ndarray):
            return [conv(x) for x in o.reshape(-1).tolist()] if o.dtype.kind == "f" else o.reshape(-1).tolist()
        if isinstance(o, (np.floating,)):
            return float(o)
        if isinstance(o, (np.integer,)):
            return int(o)
        if isinstance(o, dict):
            return {k: conv(v) for k, v in o.items()}
        if isinstance(o, (list, tuple)):
            return [conv(x) for x in o]
        return o
    return json.dumps(conv(info))
