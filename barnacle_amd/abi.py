"""ctypes mirror of include/barnacle_hip.h.

The struct layouts here are checked against the compiled library with
bn_struct_sizes() (tests/test_abi.py).  Field meanings follow the reference
scene schema, Extensions/Scene/Loader.fs:14-251.
"""
import ctypes as C

i32, i64, f32, u8, i8 = C.c_int32, C.c_int64, C.c_float, C.c_uint8, C.c_int8


class Primitive(C.Structure):
    _fields_ = [("kind", i32), ("vert_offset", i32), ("n_verts", i32),
                ("index_offset", i32), ("n_indices", i32), ("radius", f32)]


class Material(C.Structure):
    _fields_ = [("kind", i32), ("albedo", f32 * 3), ("ior", f32),
                ("roughness", f32), ("metallic", f32)]


class Light(C.Structure):
    _fields_ = [("kind", i32), ("emission", f32 * 3), ("two_sided", i32)]


class Instance(C.Structure):
    _fields_ = [("primitive", i32), ("material", i32), ("light", i32)]


class Transform(C.Structure):
    _fields_ = [("flags", i32), ("scale", f32 * 3), ("rotation", f32 * 3),
                ("translation", f32 * 3), ("time", f32), ("n_keys", i32)]


class Node(C.Structure):
    _fields_ = [("transform", i32), ("has_camera", i32), ("inst_offset", i32),
                ("n_inst", i32), ("child_offset", i32), ("n_child", i32)]


class Camera(C.Structure):
    _fields_ = [("kind", i32), ("fov", f32), ("aspect", f32), ("aperture", f32),
                ("focus_distance", f32), ("push_forward", f32)]


class SceneDesc(C.Structure):
    _fields_ = [("vertices", C.POINTER(f32)), ("n_vertex_floats", i64),
                ("indices", C.POINTER(i32)), ("n_indices", i64),
                ("primitives", C.POINTER(Primitive)), ("n_primitives", i32),
                ("materials", C.POINTER(Material)), ("n_materials", i32),
                ("lights", C.POINTER(Light)), ("n_lights", i32),
                ("instances", C.POINTER(Instance)), ("n_instances", i32),
                ("transforms", C.POINTER(Transform)), ("n_transforms", i32),
                ("nodes", C.POINTER(Node)), ("n_nodes", i32),
                ("node_refs", C.POINTER(i32)), ("n_node_refs", i32),
                ("root", i32), ("camera", Camera), ("time", f32)]


class PTParams(C.Structure):
    _fields_ = [("width", i32), ("height", i32), ("spp", i32),
                ("max_depth", i32), ("rr_depth", i32), ("frame_id", i32)]


class MLTParams(C.Structure):
    _fields_ = [("width", i32), ("height", i32), ("spp", i32),
                ("max_depth", i32), ("rr_depth", i32), ("frame_id", i32),
                ("n_bootstrap", i32), ("n_chains", i32), ("strategy", i32),
                ("sigma", f32), ("eps_min", f32), ("eps_max", f32),
                ("large_step_prob", f32)]


class MLTStats(C.Structure):
    _fields_ = [("accepted", i64), ("proposed", i64), ("b", f32),
                ("mutations_per_chain", i32)]


class TraceStats(C.Structure):
    _fields_ = [("rays_closest", i64), ("rays_shadow", i64),
                ("nodes_closest", i64), ("tris_closest", i64), ("insts_closest", i64),
                ("nodes_shadow", i64), ("tris_shadow", i64), ("insts_shadow", i64),
                ("vertices", i64), ("paths", i64)]


class BVHNode(C.Structure):
    _fields_ = [("pmin", f32 * 3), ("pmax", f32 * 3), ("right_or_offset", i32),
                ("is_leaf", u8), ("split_axis", i8), ("count", i8), ("mask", u8)]


class InstanceUpload(C.Structure):
    _fields_ = [("o2w", f32 * 16), ("w2o", f32 * 16), ("bounds", f32 * 6),
                ("blas", i32), ("sphere_radius", f32), ("material", i32),
                ("light", i32)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", i64), ("total_ms", C.c_double)]


class BatchItem(C.Structure):
    _fields_ = [("scene", i32), ("params", PTParams)]


class BatchStats(C.Structure):
    _fields_ = [("rounds", i32), ("graph", i32), ("scene_rounds", i32)]


class TuneResult(C.Structure):
    _fields_ = [("trace_blocks_per_cu", i32), ("refill", i32), ("stack_k", i32), ("trials", i32),
                ("chunk_slots", i64), ("baseline_ms", C.c_float), ("best_ms", C.c_float)]


BN_TUNE_BATCH = 1

# Declaration order of include/barnacle_hip.h, as bn_struct_sizes() reports it.
STRUCT_ORDER = [Primitive, Material, Light, Instance, Transform, Node, Camera,
                SceneDesc, PTParams, MLTParams, MLTStats, TraceStats, BVHNode,
                InstanceUpload, BatchItem, BatchStats, KernelTime, TuneResult]

BN_OK, BN_E_INVALID, BN_E_NO_LIGHT, BN_E_DEVICE, BN_E_UNSUPPORTED, BN_E_CAPACITY = 0, -1, -2, -3, -4, -5
