"""Seeded synthetic scenes in the reference's JSON schema.

None of these scenes ships with the reference (its only scene is
Asset/cbox.json); they are authored here from the workload descriptions:

  cornell_box()        classic Cornell box: 5 axis-aligned wall quads and two
                       rotated boxes (34 triangles) plus a 2-triangle area light
                       of radiance (17,12,4); Lambertian red/green/white;
                       256x256, fov 39.3, 16 spp, depth 8.
  displaced_sphere()   geodesic icosphere of frequency 59 (20*59^2 = 69,620
                       triangles; no icosphere has exactly 69,451 faces, which is
                       the face count of the reference's bunny asset) displaced
                       radially by 4-octave value noise (amplitude 0.08).
  config1()            Cornell box + displaced sphere with one GGX (pbr)
                       material drawn from seed 1; 1024x1024, 64 spp, depth 16.
  config1_bunny()      config1 with the reference's bunny OBJ as the mesh.
  caustic_scene()      Cornell box + dielectric sphere (IOR 1.5) and a GGX metal
                       sphere (roughness 0.1) of 20,480 triangles each, PSSMLT.

Geometry is generated in float64 and rounded once to float32.  The camera
looks down -Z (Extensions/Camera/Pinhole.fs:21), so the box spans z in
[-555, 0] and the camera sits at z = +800 with no rotation.
"""
from __future__ import annotations

import numpy as np

WHITE = [0.73, 0.73, 0.73]
RED = [0.63, 0.065, 0.05]
GREEN = [0.14, 0.45, 0.09]


def _quad(p0, p1, p2, p3):
    v = np.array([p0, p1, p2, p3], np.float64).astype(np.float32)
    return {"type": "mesh", "vertices": v.reshape(-1), "indices": np.array([0, 1, 2, 0, 2, 3], np.int32)}


def icosphere(freq: int):
    """Geodesic sphere: every icosahedron face split into freq^2 triangles,
    vertices projected to the unit sphere, shared vertices merged, faces wound
    so that cross(p1-p0, p2-p0) points outward."""
    t = (1.0 + 5 ** 0.5) / 2.0
    base = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                     [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    faces = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
                      (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
                      (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)])
    ii, jj = np.meshgrid(np.arange(freq + 1), np.arange(freq + 1), indexing="ij")
    keep = ii + jj <= freq
    ii, jj = ii[keep], jj[keep]
    lid = -np.ones((freq + 1, freq + 1), np.int64)
    lid[ii, jj] = np.arange(len(ii))
    A, B, Cc = base[faces[:, 0]], base[faces[:, 1]], base[faces[:, 2]]
    P = (A[:, None, :] + (B - A)[:, None, :] * (ii / freq)[None, :, None] +
         (Cc - A)[:, None, :] * (jj / freq)[None, :, None]).reshape(-1, 3)
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    keys = np.round(P * 1e9).astype(np.int64)
    _, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)                      # vertices numbered by first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    V = P[first[order]]
    gid = rank[inv.reshape(-1)].reshape(20, -1)
    i, j = np.meshgrid(np.arange(freq), np.arange(freq), indexing="ij")
    up = (i + j) <= freq - 1
    dn = (i + j) < freq - 1
    t_up = np.stack([lid[i[up], j[up]], lid[i[up] + 1, j[up]], lid[i[up], j[up] + 1]], 1)
    t_dn = np.stack([lid[i[dn] + 1, j[dn]], lid[i[dn] + 1, j[dn] + 1], lid[i[dn], j[dn] + 1]], 1)
    local = np.concatenate([t_up, t_dn])
    T = np.concatenate([gid[f][local] for f in range(20)])
    n = np.cross(V[T[:, 1]] - V[T[:, 0]], V[T[:, 2]] - V[T[:, 0]])
    flip = np.einsum("ij,ij->i", n, V[T].mean(axis=1)) < 0
    T[flip] = T[flip][:, [0, 2, 1]]
    return V, T.astype(np.int32)


def value_noise(p, seed: int, octaves: int = 4):
    """Sum of `octaves` trilinear value-noise layers (frequency x2, amplitude
    /2 per octave) over a 256-entry seeded lattice, normalised to [-1, 1]."""
    rng = np.random.RandomState(seed)
    perm = rng.permutation(256)
    vals = rng.uniform(-1.0, 1.0, 256)

    def lattice(ix, iy, iz):
        return vals[perm[(perm[(perm[ix & 255] + iy) & 255] + iz) & 255]]

    total, amp, freq, norm = np.zeros(len(p)), 1.0, 2.0, 0.0
    for _ in range(octaves):
        q = p * freq + 17.0
        i = np.floor(q).astype(np.int64)
        f = q - i
        s = f * f * (3 - 2 * f)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = (s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1]) * \
                        (s[:, 2] if dz else 1 - s[:, 2])
                    acc += w * lattice(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz)
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def displaced_sphere(freq: int = 59, amplitude: float = 0.08, seed: int = 1):
    V, T = icosphere(freq)
    V = V * (1.0 + amplitude * value_noise(V, seed))[:, None]
    return {"type": "mesh", "vertices": V.astype(np.float32).reshape(-1), "indices": T.reshape(-1)}


def plain_sphere(freq: int = 32):
    V, T = icosphere(freq)
    return {"type": "mesh", "vertices": V.astype(np.float32).reshape(-1), "indices": T.reshape(-1)}


def cornell_box(width=256, height=256, spp=16, max_depth=8, boxes=True):
    S = 555.0
    prims = [
        _quad((0, 0, 0), (S, 0, 0), (S, 0, -S), (0, 0, -S)),          # floor
        _quad((0, S, 0), (0, S, -S), (S, S, -S), (S, S, 0)),          # ceiling
        _quad((0, 0, -S), (S, 0, -S), (S, S, -S), (0, S, -S)),        # back
        _quad((0, 0, 0), (0, 0, -S), (0, S, -S), (0, S, 0)),          # left (red)
        _quad((S, 0, 0), (S, S, 0), (S, S, -S), (S, 0, -S)),          # right (green)
        _quad((213, 554, -227), (343, 554, -227), (343, 554, -332), (213, 554, -332)),  # light
    ]
    mats = [{"type": "lambertian", "albedo": WHITE}, {"type": "lambertian", "albedo": RED},
            {"type": "lambertian", "albedo": GREEN}]
    insts = [{"primitive": 0, "material": 0}, {"primitive": 1, "material": 0},
             {"primitive": 2, "material": 0}, {"primitive": 3, "material": 1},
             {"primitive": 4, "material": 2}, {"primitive": 5, "light": 0}]
    nodes = [{"instances": list(range(6)), "children": []}]
    transforms = []
    if boxes:
        prims.append({"type": "cube"})
        transforms += [
            {"keyframes": [{"time": 0.0, "scale": [82.5, 82.5, 82.5], "rotation": [0.0, 0.3141593, 0.0],
                            "translation": [185.5, 82.5, -168.5]}]},
            {"keyframes": [{"time": 0.0, "scale": [82.5, 165.0, 82.5], "rotation": [0.0, -0.296706, 0.0],
                            "translation": [368.5, 165.0, -351.5]}]}]
        insts += [{"primitive": 6, "material": 0}, {"primitive": 6, "material": 0}]
        nodes += [{"instances": [6], "transform": 0}, {"instances": [7], "transform": 1}]
    transforms.append({"keyframes": [{"time": 0.0, "translation": [277.5, 277.5, 800.0]}]})
    nodes.append({"transform": len(transforms) - 1, "has-camera": True})
    nodes[0]["children"] = list(range(1, len(nodes)))
    return {
        "root": 0, "nodes": nodes, "instances": insts, "transforms": transforms, "primitives": prims,
        "materials": mats, "lights": [{"type": "diffuse", "emission": [17.0, 12.0, 4.0]}],
        "integrator": {"type": "path-tracing", "spp": spp, "max-depth": max_depth, "rr-depth": 5},
        "camera": {"type": "pinhole", "fov": 39.3, "aspect-ratio": 1.0},
        "film": {"width": width, "height": height, "tone-mapping": "aces"},
    }


def _add_mesh_instance(info, prim, material, scale, translation):
    info["primitives"].append(prim)
    info["materials"].append(material)
    info["instances"].append({"primitive": len(info["primitives"]) - 1,
                              "material": len(info["materials"]) - 1})
    info["transforms"].append({"keyframes": [{"time": 0.0, "scale": [scale] * 3,
                                              "translation": list(translation)}]})
    info["nodes"].append({"instances": [len(info["instances"]) - 1],
                          "transform": len(info["transforms"]) - 1})
    info["nodes"][0]["children"].append(len(info["nodes"]) - 1)


def config0():
    return cornell_box()


def config1(width=1024, height=1024, spp=64, max_depth=16, freq=59):
    info = cornell_box(width, height, spp, max_depth)
    rng = np.random.default_rng(1)
    base = [float(np.float32(x)) for x in rng.uniform(0.2, 0.9, 3)]
    metallic = float(rng.integers(0, 2))
    rough = float(np.float32(rng.uniform(0.05, 0.8)))
    _add_mesh_instance(info, displaced_sphere(freq, 0.08, 1),
                       {"type": "pbr", "albedo": base, "metallic": metallic, "roughness": rough},
                       90.0, (185.5, 300.0, -168.5))
    return info


def config1_bunny(obj_path: str, width=1024, height=1024, spp=64, max_depth=16):
    """Config 1 with the reference's own mesh asset in place of the displaced
    sphere: Asset/stanford-bunny.obj (69,451 triangles, read by the OBJ
    reader of Mesh.fs:246-279 through the scene's "uri"), scaled by 1200 and
    set on top of the short box, with config 1's GGX material."""
    import os
    info = cornell_box(width, height, spp, max_depth)
    rng = np.random.default_rng(1)
    base = [float(np.float32(x)) for x in rng.uniform(0.2, 0.9, 3)]
    metallic = float(rng.integers(0, 2))
    rough = float(np.float32(rng.uniform(0.05, 0.8)))
    _add_mesh_instance(info, {"type": "mesh", "uri": os.path.abspath(obj_path)},
                       {"type": "pbr", "albedo": base, "metallic": metallic, "roughness": rough},
                       1200.0, (205.9, 125.4, -166.7))
    return info


def _morton3(q):
    """Interleave three 10-bit integer coordinates (Morton order)."""
    def spread(v):
        v = v.astype(np.uint64) & 0x3FF
        v = (v | (v << 16)) & 0x030000FF
        v = (v | (v << 8)) & 0x0300F00F
        v = (v | (v << 4)) & 0x030C30C3
        v = (v | (v << 2)) & 0x09249249
        return v
    return spread(q[:, 0]) | (spread(q[:, 1]) << 1) | (spread(q[:, 2]) << 2)


def split_mesh(mesh, n_patches):
    """Cut a mesh into n_patches spatially compact meshes of nearly equal
    triangle count (Morton order of the triangle centroids)."""
    V = np.asarray(mesh["vertices"], np.float32).reshape(-1, 3)
    T = np.asarray(mesh["indices"], np.int64).reshape(-1, 3)
    c = V[T].mean(axis=1)
    lo, hi = c.min(axis=0), c.max(axis=0)
    q = np.floor((c - lo) / np.maximum(hi - lo, 1e-12) * 1023.0).astype(np.int64)
    order = np.argsort(_morton3(q), kind="stable")
    out = []
    for part in np.array_split(order, n_patches):
        tri = T[part]
        used, inv = np.unique(tri.reshape(-1), return_inverse=True)
        out.append({"type": "mesh", "vertices": V[used].reshape(-1), "indices": inv.reshape(-1).astype(np.int64)})
    return out


def emitter_wall(n_side: int, emission, z: float = -555.0, size: float = 555.0):
    """The box's back wall as n_side x n_side light-only quads (two triangles
    each, facing +z), quad k = (row k // n_side, column k % n_side) with
    emission[k]."""
    prims, lights = [], []
    step = size / n_side
    for k in range(n_side * n_side):
        r, c = divmod(k, n_side)
        x0, x1 = c * step, (c + 1) * step
        y0, y1 = r * step, (r + 1) * step
        prims.append(_quad((x0, y0, z), (x1, y0, z), (x1, y1, z), (x0, y1, z)))
        lights.append({"type": "diffuse", "emission": [float(v) for v in emission[k]]})
    return prims, lights


def config3(width=512, height=512, spp=32, max_depth=8, freq=59, n_patches=4096, seed=2):
    """Inverse-rendering scene on config 1's geometry: n_patches parameter
    sets, each an RGB albedo, a roughness and an RGB emission.  The displaced
    sphere is cut into n_patches GGX patch instances (albedo, roughness;
    metallic drawn 0 or 1 per patch and held fixed), and the box's back wall
    into n_patches light-only tiles (emission; sqrt(n_patches) a side), so the
    uniform light sampler covers every tile.  Parameter set k is patch k's
    material and tile k's light.  Emission sits on light-only instances
    because the reference spawns rays without an origin offset: a surface
    that both reflects and emits can hit itself at a distance that rounds to
    zero, and the light pdf of that hit is 0/0 (a NaN radiance that covered
    47% of the pixels when the patches carried their own emission).
    Parameters are the ground truth drawn from `seed`; config3_params gives
    them and the start point."""
    n_side = int(round(n_patches ** 0.5))
    if n_side * n_side != n_patches:
        raise ValueError("n_patches must be a square")
    info = cornell_box(width, height, spp, max_depth)
    rng = np.random.default_rng(seed)
    alb = rng.uniform(0.2, 0.9, (n_patches, 3)).astype(np.float32)
    rough = rng.uniform(0.05, 0.8, n_patches).astype(np.float32)
    metal = rng.integers(0, 2, n_patches).astype(np.float32)
    em = rng.uniform(0.0, 0.2, (n_patches, 3)).astype(np.float32)
    # the back wall (primitive 2) gives way to the tiles
    back = next(k for k, i in enumerate(info["instances"]) if i["primitive"] == 2)
    info["nodes"][0]["instances"].remove(back)
    node = {"instances": [], "transform": len(info["transforms"])}
    info["transforms"].append({"keyframes": [{"time": 0.0, "scale": [90.0] * 3,
                                              "translation": [185.5, 300.0, -168.5]}]})
    first_material = len(info["materials"])
    for k, prim in enumerate(split_mesh(displaced_sphere(freq, 0.08, 1), n_patches)):
        info["primitives"].append(prim)
        info["materials"].append({"type": "pbr", "albedo": [float(x) for x in alb[k]], "metallic": float(metal[k]),
                                  "roughness": float(rough[k])})
        info["instances"].append({"primitive": len(info["primitives"]) - 1,
                                  "material": len(info["materials"]) - 1})
        node["instances"].append(len(info["instances"]) - 1)
    info["nodes"].append(node)
    info["nodes"][0]["children"].append(len(info["nodes"]) - 1)
    first_light = len(info["lights"])
    prims, lights = emitter_wall(n_side, em)
    for prim, light in zip(prims, lights):
        info["primitives"].append(prim)
        info["lights"].append(light)
        info["instances"].append({"primitive": len(info["primitives"]) - 1, "light": len(info["lights"]) - 1})
        info["nodes"][0]["instances"].append(len(info["instances"]) - 1)
    info["x-patches"] = {"first_material": first_material, "first_light": first_light, "count": n_patches}
    return info


def config3_params(sc, info):
    """(truth, start, mask): parameter arrays (material (n_mat, 4) albedo rgb +
    roughness, light (n_light, 3) emission) of the scene as drawn, the
    start point (patch albedo 0.5, roughness 0.5, tile emission 0) and the
    mask of the parameters being optimised (patches and tiles)."""
    mat = np.array([[*m["albedo"], m["roughness"]] for m in sc.materials], np.float32)
    light = np.array([l["emission"] for l in sc.lights], np.float32)
    p = info["x-patches"]
    fm, fl, n = p["first_material"], p["first_light"], p["count"]
    mat0, light0 = mat.copy(), light.copy()
    mat0[fm:fm + n] = [0.5, 0.5, 0.5, 0.5]
    light0[fl:fl + n] = 0.0
    mmask = np.zeros(mat.shape, bool)
    mmask[fm:fm + n] = True
    lmask = np.zeros(light.shape, bool)
    lmask[fl:fl + n] = True
    return (mat, light), (mat0, light0), (mmask, lmask)


def caustic_scene(width=1024, height=1024, spp=256, max_depth=16, freq=32,
                  n_bootstrap=1 << 20, n_chains=1 << 20, sigma=1.0 / 64.0, large_step=0.3):
    info = cornell_box(width, height, spp, max_depth, boxes=False)
    _add_mesh_instance(info, plain_sphere(freq), {"type": "dielectric", "ior": 1.5},
                       110.0, (380.0, 110.0, -200.0))
    _add_mesh_instance(info, plain_sphere(freq), {"type": "pbr", "albedo": [0.9, 0.9, 0.9],
                                                  "metallic": 1.0, "roughness": 0.1},
                       110.0, (160.0, 110.0, -380.0))
    info["integrator"] = {"type": "pssmlt", "spp": spp, "max-depth": max_depth, "rr-depth": 5,
                          "n-bootstrap": n_bootstrap, "n-chains": n_chains,
                          "mutation-strategy": "Gaussian", "mutation-sigma": sigma,
                          "large-step-prob": large_step}
    return info


def sphere_scene(width=64, height=64, spp=4, max_depth=8):
    """The scene features of the reference's own asset that configs 0-4 leave
    out: analytic spheres (SpherePrimitive; uniform and non-uniform scale),
    a sphere light (SphereInstance.Sample / EvalPDF) beside the quad light, a
    rotated mirror cube, a panel of the builtin quad and a thin-lens camera
    with a finite aperture."""
    info = cornell_box(width, height, spp, max_depth, boxes=False)
    prims, mats, insts, tfs, nodes = (info[k] for k in ("primitives", "materials", "instances", "transforms", "nodes"))
    cam_node = nodes.pop()                           # keep the camera node last
    prims.append({"type": "sphere", "radius": 1.0})
    prims.append({"type": "cube"})
    prims.append({"type": "quad"})
    sphere, cube, quad = len(prims) - 3, len(prims) - 2, len(prims) - 1
    mats += [{"type": "pbr", "albedo": [0.9, 0.85, 0.8], "metallic": 1.0, "roughness": 0.2},
             {"type": "dielectric", "ior": 1.5},
             {"type": "mirror", "albedo": [0.95, 0.95, 0.95]}]
    info["lights"].append({"type": "diffuse", "emission": [4.0, 4.0, 8.0]})
    placed = [({"primitive": sphere, "material": 3}, {"scale": [90.0, 90.0, 90.0], "translation": [150.0, 90.0, -200.0]}),
              ({"primitive": sphere, "material": 4}, {"scale": [70.0, 110.0, 70.0], "translation": [400.0, 110.0, -300.0]}),
              ({"primitive": sphere, "light": 1}, {"scale": [30.0, 30.0, 30.0], "translation": [280.0, 420.0, -150.0]}),
              ({"primitive": cube, "material": 5}, {"scale": [50.0, 50.0, 50.0], "rotation": [0.3, 0.5, 0.0],
                                                     "translation": [300.0, 50.0, -100.0]}),
              ({"primitive": quad, "material": 0}, {"scale": [80.0, 1.0, 60.0], "rotation": [1.5707964, 0.0, 0.2],
                                                     "translation": [120.0, 330.0, -500.0]})]
    for inst, tf in placed:
        insts.append(inst)
        tfs.append({"keyframes": [dict(time=0.0, **tf)]})
        nodes.append({"instances": [len(insts) - 1], "transform": len(tfs) - 1})
    nodes.append(cam_node)
    nodes[0]["children"] = list(range(1, len(nodes)))
    info["camera"] = {"type": "thin-lens", "fov": 39.3, "aspect-ratio": 1.0, "aperture": 12.0,
                      "focus-distance": 1000.0, "push-forward": 0.0}
    return info


def batch_scene(seed: int, max_res: int = 1024, max_spp: int = 64, tri_range=(1000, 500000),
                max_instances: int = 256):
    """One scene of the 32-scene batch: a ground quad, a pool of at most 8
    displaced icospheres (log-uniform triangle counts) instanced 1-256 times
    with random affine transforms (scale, rotation in radians, translation),
    a random mix of lambertian / mirror / dielectric (IOR 1.5) / pbr materials
    and 1-8 quad area lights.  Resolution from {256, 512, 768, 1024}^2 (capped
    at max_res), spp 4-64 (capped at max_spp), depth 4-16."""
    rng = np.random.default_rng(2024 + seed)
    res = int(rng.choice([256, 512, 768, 1024]))
    res = min(res, max_res)
    spp = int(min(rng.integers(4, 65), max_spp))
    depth = int(rng.integers(4, 17))
    n_blas = int(rng.integers(1, 9))
    lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
    prims, mats = [], []
    for b in range(n_blas):
        tris = float(np.exp(rng.uniform(lo, hi)))
        freq = max(2, int(round(np.sqrt(tris / 20.0))))
        prims.append(displaced_sphere(freq, 0.08, int(rng.integers(0, 1 << 30))))
    kinds = ["lambertian", "mirror", "dielectric", "pbr"]
    for m in range(8):
        k = kinds[int(rng.integers(0, 4))]
        alb = [float(np.float32(x)) for x in rng.uniform(0.2, 0.9, 3)]
        if k == "pbr":
            mats.append({"type": k, "albedo": alb, "metallic": float(rng.integers(0, 2)),
                         "roughness": float(np.float32(rng.uniform(0.05, 0.8)))})
        elif k == "dielectric":
            mats.append({"type": k, "albedo": alb, "ior": 1.5})
        else:
            mats.append({"type": k, "albedo": alb})
    mats.append({"type": "lambertian", "albedo": WHITE})              # ground
    n_inst = int(np.clip(np.round(np.exp(rng.uniform(0, np.log(max_instances)))), 1, max_instances))
    R = 100.0
    insts, transforms, node_children, nodes = [], [], [], [{"instances": [], "children": []}]
    for i in range(n_inst):
        s_ = float(np.float32(min(0.6 * R, rng.uniform(0.2, 0.5) * 2 * R / n_inst ** (1 / 3))))
        sc = [s_ * float(np.float32(x)) for x in rng.uniform(0.7, 1.3, 3)]
        rot = [float(np.float32(x)) for x in rng.uniform(-np.pi, np.pi, 3)]
        tr = [float(np.float32(x)) for x in (rng.uniform(-R, R), s_ + rng.uniform(0.0, 2 * R), rng.uniform(-R, R))]
        transforms.append({"keyframes": [{"time": 0.0, "scale": sc, "rotation": rot, "translation": tr}]})
        insts.append({"primitive": int(rng.integers(0, n_blas)), "material": int(rng.integers(0, 8))})
        nodes.append({"instances": [len(insts) - 1], "transform": len(transforms) - 1})
        node_children.append(len(nodes) - 1)
    G = 4 * R
    prims.append(_quad((-G, 0, -G), (G, 0, -G), (G, 0, G), (-G, 0, G)))
    insts.append({"primitive": len(prims) - 1, "material": 8})
    nodes[0]["instances"].append(len(insts) - 1)
    lights = []
    n_lights = int(rng.integers(1, 9))
    for l in range(n_lights):
        cx, cz = rng.uniform(-R, R, 2)
        h = float(rng.uniform(2.2 * R, 3.0 * R))
        w = float(rng.uniform(10, 40))
        prims.append(_quad((cx - w, h, cz - w), (cx + w, h, cz - w), (cx + w, h, cz + w), (cx - w, h, cz + w)))
        e = [float(np.float32(x)) for x in rng.uniform(5, 20, 3)]
        lights.append({"type": "diffuse", "emission": e})
        insts.append({"primitive": len(prims) - 1, "light": l})
        nodes[0]["instances"].append(len(insts) - 1)
    transforms.append({"keyframes": [{"time": 0.0, "rotation": [-0.35, 0.0, 0.0],
                                      "translation": [0.0, 1.6 * R, 4.2 * R]}]})
    nodes.append({"transform": len(transforms) - 1, "has-camera": True})
    node_children.append(len(nodes) - 1)
    nodes[0]["children"] = node_children
    return {"root": 0, "nodes": nodes, "instances": insts, "transforms": transforms, "primitives": prims,
            "materials": mats, "lights": lights,
            "integrator": {"type": "path-tracing", "spp": spp, "max-depth": depth, "rr-depth": 5},
            "camera": {"type": "pinhole", "fov": 50.0, "aspect-ratio": 1.0},
            "film": {"width": res, "height": res, "tone-mapping": "aces"}}


def config2(n: int = 32, **kw):
    """The 32-scene batch (seeds 2024..2055)."""
    return [batch_scene(i, **kw) for i in range(n)]
