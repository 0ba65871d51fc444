/*
 * barnacle_hip.h -- C ABI of the MI355X wavefront path tracer.
 *
 * This is the drop-in boundary for the hot path of the Barnacle renderer
 * (LeonKang130/Barnacle).  The reference's plugin seam is the F# abstract class
 * IntegratorBase (Base/Integrator.fs:8-11), whose Render(camera, film,
 * aggregate, lightSampler) is called once per frame by Scene.Render
 * (Extensions/Scene/Render.fs:16).  A GPU integrator registered under a new
 * "type" string in IntegratorInfo.ToIntegrator (Extensions/Scene/Loader.fs:185-204)
 * binds the functions below through P/Invoke; the Python ctypes layer in
 * barnacle_amd/abi.py binds exactly the same symbols.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative BN_E* code on failure;
 *    bn_last_error() then holds a message (thread-local, valid until the next
 *    call on that thread).  Nothing throws or aborts across the ABI; the
 *    reference raises F# exceptions at the same places (failwith).
 *  - Inputs are caller-owned host arrays borrowed for the duration of the call;
 *    the library deep-copies what it keeps into device memory.  Outputs go to
 *    caller-allocated host buffers.  All calls are blocking.
 *  - One bn_ctx per host thread.  A context owns all device memory, streams and
 *    graphs it creates.
 *  - All structs are plain int32/float records (4-byte alignment, no padding).
 */
#ifndef BARNACLE_HIP_H
#define BARNACLE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
#define BN_OK              0
#define BN_E_INVALID      -1   /* bad argument / malformed scene            */
#define BN_E_NO_LIGHT     -2   /* LightSamplerBase: no light primitives     */
#define BN_E_DEVICE       -3   /* HIP runtime error                         */
#define BN_E_UNSUPPORTED  -4   /* valid scene using a feature not built     */
#define BN_E_CAPACITY     -5   /* size beyond a documented maximum          */

/* ------------------------------------------------------ scene description
 * A flat restatement of the reference's scene JSON schema
 * (Extensions/Scene/Loader.fs:14-251).  The JSON front end
 * (barnacle_amd/scene.py) resolves defaults exactly as Loader.fs does and
 * expands the builtin "quad"/"cube" meshes (Extensions/Primitive/Mesh.fs:126-155).
 */
#define BN_PRIM_MESH    0
#define BN_PRIM_SPHERE  1

typedef struct bn_primitive {          /* PrimitiveInfo (Loader.fs:92-116)   */
    int32_t kind;                      /* BN_PRIM_*                          */
    int32_t vert_offset;               /* first vertex in vertices[] (xyz)   */
    int32_t n_verts;
    int32_t index_offset;              /* first index in indices[]           */
    int32_t n_indices;                 /* 3 per triangle                     */
    float   radius;                    /* sphere only                        */
} bn_primitive;

#define BN_MAT_LAMBERTIAN  0
#define BN_MAT_MIRROR      1
#define BN_MAT_DIELECTRIC  2
#define BN_MAT_PBR         3
/* PBRMaterial handed over as built: `roughness` holds PBRMaterial.Alpha
 * (= Max(roughness^2, 1e-3), PBR.fs:11) itself, which is what a host that
 * holds the constructed material object has; metallic is clamped as for
 * BN_MAT_PBR.  Upload path only (bn_upload_scene). */
#define BN_MAT_PBR_ALPHA   4

typedef struct bn_material {           /* MaterialInfo (Loader.fs:50-73)     */
    int32_t kind;                      /* BN_MAT_*                           */
    float   albedo[3];
    float   ior;
    float   roughness;
    float   metallic;
} bn_material;

typedef struct bn_light {              /* LightInfo (Loader.fs:75-90)        */
    int32_t kind;                      /* 0 = diffuse (the only type)        */
    float   emission[3];
    int32_t two_sided;
} bn_light;

typedef struct bn_instance {           /* PrimitiveInstanceInfo (:118-145)   */
    int32_t primitive;
    int32_t material;                  /* -1 = none                          */
    int32_t light;                     /* -1 = none                          */
} bn_instance;

#define BN_XF_SCALE        1
#define BN_XF_ROTATION     2
#define BN_XF_TRANSLATION  4

/* One key frame (KeyFrameInfo, Loader.fs:14-41).  A transform is a run of
 * n_keys consecutive records sorted by time (Loader.fs:47); nodes reference
 * the run's first record, which carries n_keys (0 counts as 1).  The scene is
 * evaluated at bn_scene_desc.time by Transform.Eval (Base/Scene.fs:27-36),
 * interpolating between key frames as KeyFrame.Interpolate does (:13-19). */
typedef struct bn_transform {
    int32_t flags;                     /* BN_XF_* present components         */
    float   scale[3];                  /* (Loader.fs:20-41): S*Rx*Ry*Rz*T    */
    float   rotation[3];               /* radians                            */
    float   translation[3];
    float   time;                      /* key frame time, default 0          */
    int32_t n_keys;                    /* key frames in this run (first)     */
} bn_transform;

typedef struct bn_node {               /* NodeInfo (Loader.fs:147-168)       */
    int32_t transform;                 /* -1 = identity                      */
    int32_t has_camera;
    int32_t inst_offset;               /* into node_refs[]                   */
    int32_t n_inst;
    int32_t child_offset;              /* into node_refs[]                   */
    int32_t n_child;
} bn_node;

#define BN_CAM_PINHOLE    0
#define BN_CAM_THIN_LENS  1

typedef struct bn_camera {             /* CameraInfo (Loader.fs:206-224)     */
    int32_t kind;
    float   fov;                       /* vertical, degrees                  */
    float   aspect;                    /* from the camera, not the film      */
    float   aperture;
    float   focus_distance;
    float   push_forward;
} bn_camera;

typedef struct bn_scene_desc {
    const float*        vertices;     int64_t n_vertex_floats;
    const int32_t*      indices;      int64_t n_indices;
    const bn_primitive* primitives;   int32_t n_primitives;
    const bn_material*  materials;    int32_t n_materials;
    const bn_light*     lights;       int32_t n_lights;
    const bn_instance*  instances;    int32_t n_instances;
    const bn_transform* transforms;   int32_t n_transforms;
    const bn_node*      nodes;        int32_t n_nodes;
    const int32_t*      node_refs;    int32_t n_node_refs;
    int32_t             root;
    bn_camera           camera;
    float               time;         /* Scene.Render(t): the reference renders t = 0 */
} bn_scene_desc;

/* ------------------------------------------------------ render parameters */
typedef struct bn_pt_params {          /* PathTracingIntegrator + Film       */
    int32_t width, height;             /* FilmInfo (Loader.fs:226-239)       */
    int32_t spp;                       /* "spp"       default 1              */
    int32_t max_depth;                 /* "max-depth" default 8              */
    int32_t rr_depth;                  /* "rr-depth"  default 5              */
    int32_t frame_id;                  /* ProgressiveIntegrator.FrameId      */
} bn_pt_params;

#define BN_MLT_GAUSSIAN  0
#define BN_MLT_KELEMEN   1

typedef struct bn_mlt_params {         /* PSSMLTIntegrator (PSSMLT.fs:153)   */
    int32_t width, height;
    int32_t spp;                       /* mutations per pixel                */
    int32_t max_depth, rr_depth, frame_id;
    int32_t n_bootstrap;               /* default 4*1024*1024                */
    int32_t n_chains;                  /* default 1024                       */
    int32_t strategy;                  /* BN_MLT_*                           */
    float   sigma;                     /* Gaussian sigma (reference: 1e-2)   */
    float   eps_min, eps_max;          /* Kelemen (reference: 1/1024, 1/16)  */
    float   large_step_prob;           /* default 0.5                        */
} bn_mlt_params;

typedef struct bn_mlt_stats {          /* PSSMLT.fs:412-414 prints these     */
    int64_t accepted;
    int64_t proposed;
    float   b;                         /* mean bootstrap luminance           */
    int32_t mutations_per_chain;
} bn_mlt_stats;

/* Per-launch counters of the traversal, for the algorithmic-bytes roofline. */
typedef struct bn_trace_stats {
    int64_t rays_closest, rays_shadow;
    int64_t nodes_closest, tris_closest, insts_closest;
    int64_t nodes_shadow,  tris_shadow,  insts_shadow;
    int64_t vertices;                  /* path vertices shaded               */
    int64_t paths;
} bn_trace_stats;

/* Reference BVHNode, 32 bytes, explicit layout of Util/BVH.fs:52-73. */
typedef struct bn_bvh_node {
    float   pmin[3];
    float   pmax[3];
    int32_t right_or_offset;           /* RightChild | InstanceOffset @24    */
    uint8_t is_leaf;                   /* @28                                */
    int8_t  split_axis;                /* @29                                */
    int8_t  count;                     /* @30                                */
    uint8_t mask;                      /* @31 VisibilityMask (unused)        */
} bn_bvh_node;

/* ----------------------------------------------------------- the context */
typedef struct bn_ctx bn_ctx;

/* No reference counterpart (the reference has no device); the F# integrator
 * creates one context in its constructor. */
int         bn_create(int device, bn_ctx** out);
void        bn_destroy(bn_ctx* ctx);
const char* bn_last_error(void);
int         bn_abi_version(void);
/* sizeof of every struct above, in declaration order, for binding checks. */
int         bn_struct_sizes(int32_t* out, int32_t n);

/* BVHNode.Build (Util/BVH.fs:239-247): binned-SAH build over item bounds
 * (6 floats per item: pmin xyz, pmax xyz).  Writes up to 2*n nodes in the
 * reference's pre-order and the permutation perm[i] = original index of the
 * item now at position i.  Host only; the F# host calls its own BVHNode.Build
 * and uses bn_upload_* instead. */
int bn_bvh_build(const float* item_bounds, int32_t n_items,
                 bn_bvh_node* out_nodes, int32_t* out_n_nodes, int32_t* out_perm);

/* BVHNode.Build (Util/BVH.fs:128-247) on the context's GPU: the same nodes,
 * node order and permutation as bn_bvh_build, bit for bit, for finite item
 * bounds (BN_E_INVALID otherwise; bn_bvh_build takes any).  bn_scene_build
 * builds the BLAS of every mesh with finite triangle bounds this way unless
 * the option "bvh_gpu" is 0. */
int bn_bvh_build_gpu(bn_ctx* ctx, const float* item_bounds, int32_t n_items,
                     bn_bvh_node* out_nodes, int32_t* out_n_nodes, int32_t* out_perm);

/* Scene construction as Scene.Load + Scene.Render(0) perform it before the
 * timed region (Loader.fs:253-272, Base/Scene.fs:38-51, Render.fs:11-14):
 * every mesh BLAS is built (Mesh.fs:159-169), the node graph is traversed DFS
 * with world = local * parent, the TLAS is built over the instance array
 * (Extensions/Aggregate/BVH.fs:9) and the light list is taken in TLAS order
 * (Base/LightSampler.fs:7).  Everything is uploaded to the device. */
int bn_scene_build(bn_ctx* ctx, const bn_scene_desc* desc, int32_t* out_scene);
int bn_scene_release(bn_ctx* ctx, int32_t scene);

/* Read back what bn_scene_build derived, for parity checks against an
 * independent build.  n_* are in/out: capacity in, count out. */
int bn_scene_export(bn_ctx* ctx, int32_t scene,
                    bn_bvh_node* tlas, int32_t* n_tlas,
                    int32_t* inst_order, int32_t* n_inst,
                    float* inst_o2w /* 16 per instance */,
                    float* inst_w2o /* 16 per instance */,
                    int32_t* light_insts, int32_t* n_lights,
                    float* camera_to_world /* 16 */);
int bn_scene_export_blas(bn_ctx* ctx, int32_t scene, int32_t primitive,
                         bn_bvh_node* nodes, int32_t* n_nodes,
                         int32_t* tri_indices /* 3 per tri, BVH order */, int32_t* n_tris,
                         float* alias_pdf /* per tri */);

/* Pre-built upload path for the F# host, which builds the BVHs itself with
 * the reference code.  BLAS: MeshPrimitive.BVHNodes / TriangleIndices /
 * Vertices / AliasTable pdfs (Mesh.fs:157-186).  Scene: the TLAS nodes and the
 * permuted instance array of BVHAggregate (Extensions/Aggregate/BVH.fs:9) with
 * ObjectToWorld / WorldToObject / Bounds (Base/Primitive.fs:103-105) as
 * row-major 4x4 (System.Numerics layout: translation in M41..M43). */
typedef struct bn_instance_upload {
    float   o2w[16];
    float   w2o[16];
    float   bounds[6];                 /* world AABB pmin xyz, pmax xyz      */
    int32_t blas;                      /* id from bn_upload_blas, or -1      */
    float   sphere_radius;             /* used when blas == -1               */
    int32_t material;                  /* index into materials, -1 = none    */
    int32_t light;                     /* index into lights,    -1 = none    */
} bn_instance_upload;

/* alias_pdf: AliasTable.Table[i].pdf per triangle in BVH order, kept
 * verbatim (NULL: derived from the triangle areas as Mesh.fs:181-186 does). */
int bn_upload_blas(bn_ctx* ctx, const bn_bvh_node* nodes, int32_t n_nodes,
                   const float* vertices, int32_t n_verts,
                   const int32_t* tri_indices, int32_t n_tris,
                   const float* alias_pdf, int32_t* out_blas);
int bn_upload_scene(bn_ctx* ctx, const bn_bvh_node* tlas, int32_t n_tlas,
                    const bn_instance_upload* instances, int32_t n_instances,
                    const bn_material* materials, int32_t n_materials,
                    const bn_light* lights, int32_t n_lights,
                    const bn_camera* camera, const float* camera_to_world,
                    int32_t* out_scene);
/* Frees the host copy of a BLAS from bn_upload_blas once no further scene
 * will reference it (scenes keep their own device copies).  BLAS made by
 * bn_scene_build are released with their scene by bn_scene_release. */
int bn_release_blas(bn_ctx* ctx, int32_t blas);

/* The instance records bn_scene_build derived (TLAS order), in the form
 * bn_upload_scene takes them: blas = the primitive index of the scene
 * description (-1 for an analytic sphere, with its radius), world bounds as
 * PrimitiveInstance.UpdateTransform computes them (Base/Primitive.fs:131-138).
 * n: capacity in, count out. */
int bn_scene_export_instances(bn_ctx* ctx, int32_t scene, bn_instance_upload* out, int32_t* n);

/* PathTracingIntegrator.Render (Extensions/Integrator/PathTracing.fs:14-81
 * through ProgressiveIntegrator.RenderTile/Render, Base/Integrator.fs:22-55):
 * writes film_rgb[W*H*3] in Film.Pixels order, row (H-y-1) first
 * (Base/Film.fs:41-46).  film_rgb may be NULL: the frame is rendered into
 * the context's device film and not copied to the host.  stats may be NULL;
 * when given, a counting pass of the traversal fills it (that pass is slower
 * and not what is timed). */
int bn_render_pt(bn_ctx* ctx, int32_t scene, const bn_pt_params* params,
                 float* film_rgb, bn_trace_stats* stats);

/* Film.PostProcess (Base/Film.fs:21-30): identity, ACES or gamma-2.2 tone
 * curve per component, clamped to [0, 1] (out_rgb, 3 floats per pixel), and
 * the 8-bit RGBA pixels Film.Save writes (Film.fs:55-66; out_rgba8, 4 bytes per
 * pixel, v * 255 + 0.5 truncated, alpha 255, NaN -> 0), in Film.Pixels order
 * (top row first).  Host only; either output may be NULL. */
#define BN_TONE_IDENTITY 0
#define BN_TONE_ACES     1
#define BN_TONE_GAMMA    2
int bn_film_postprocess(const float* film_rgb, int64_t n_pixels, int32_t tone_mapping,
                        float* out_rgb, uint8_t* out_rgba8);

/* DirectIntegrator (Extensions/Integrator/Direct.fs:10-40): emission seen
 * directly, plus one light sample and one BSDF sample per camera hit, each
 * divided by the light sampler's pdf (no MIS); max_depth and rr_depth are
 * ignored, as the reference's DirectIntegrator(spp) has neither. */
int bn_render_direct(bn_ctx* ctx, int32_t scene, const bn_pt_params* params,
                     float* film_rgb);

/* NormalIntegrator (Extensions/Integrator/Normal.fs:10-17), the first-hit view. */
int bn_render_normal(bn_ctx* ctx, int32_t scene, const bn_pt_params* params,
                     float* film_rgb);

/* First hit of sample `sample_id` of every pixel: the closest-hit query of
 * BVHAggregate.Intersect (Extensions/Aggregate/BVH.fs:37-58).  Outputs are in
 * pixel order y*W+x: TLAS-order instance index (-1 = miss), primitive index
 * in BVH order (the Mesh.fs:232 tag; sphere: 0 near root, 1 far root) and t. */
int bn_first_hit(bn_ctx* ctx, int32_t scene, const bn_pt_params* params,
                 int32_t sample_id, int32_t* inst, int32_t* prim, float* t);

/* bn_first_hit without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and outputs as bn_first_hit, for films of at most 2^20
 * pixels (BN_E_CAPACITY beyond).  Its scratch -- fetch counter, result
 * staging, traversal-stack spill, 40 MiB in all -- is one static device
 * array per GPU, set up by bn_create and held until the process ends; every
 * call on a GPU shares it, so calls that overlap on one GPU (from several
 * contexts or threads) race.  The scene must already be resident. */
int bn_first_hit_noalloc(bn_ctx* ctx, int32_t scene, const bn_pt_params* params,
                         int32_t sample_id, int32_t* inst, int32_t* prim, float* t);

/* Batched rendering (no reference counterpart: the reference renders one
 * scene per Scene.Render).  Each item is one PathTracingIntegrator frame of
 * a scene from bn_scene_build / bn_upload_scene; all items advance through
 * one launch sequence per round (scenes, resolutions, spp and depths may
 * differ), with a block of iterations captured once as a HIP graph and
 * replayed.  An item with at least an eighth of a round of paths gets rounds
 * over its own scene (with that scene's other items); smaller items share
 * rounds.  films[k] receives item k's film exactly as bn_render_pt would
 * write it (bit-identical); NULL entries are skipped. */
typedef struct bn_batch_item {
    int32_t      scene;
    bn_pt_params params;
} bn_batch_item;
typedef struct bn_batch_stats {
    int32_t rounds;
    int32_t graph;                     /* 1 when the rounds replayed a graph */
    int32_t scene_rounds;              /* rounds whose items all share one scene: their
                                          traversals bind it once per launch */
} bn_batch_stats;
int bn_render_batch(bn_ctx* ctx, const bn_batch_item* items, int32_t n, float* const* films,
                    bn_batch_stats* stats);

/* Tunables and instrumentation (no reference counterpart).  Keys:
 *   "timing"               1: record a HIP event pair around every kernel
 *                          launch of the render calls, on the context's stream
 *   "chunk_slots"          path slots in flight (default 64M)
 *   "window_paths"         paths per round, radiance kept by path id (default 0:
 *                          as many as chunk_slots, every path in flight at once)
 *   "trace_blocks_per_cu"  persistent blocks per CU of the trace kernels
 *   "refill"               idle lanes of a wave that trigger a refill (1..64, default 40)
 *   "mlt_chunk"            PSSMLT chains (bootstrap paths) in flight per pass
 *   "stack_k"              LDS window of the traversal stack: 16 (default) or 32
 *   "graph"                1 (default): replay batched rounds as a HIP graph
 *   "batch_pack"           0 (default): pack rounds item-major, rounds over one
 *                          scene for big items; 1: fair share
 *   "material_sort"        1: bin the shading queue by material kind; 0: queue
 *                          order, which keeps path-state access coalesced;
 *                          -1 (default): bins for gradient rounds only, whose
 *                          atomics gain from material-coherent waves
 *   "bvh_gpu"              1 (default): bn_scene_build builds mesh BLAS on the
 *                          GPU (bn_bvh_build_gpu); 0: on the host
 *   "bvh_small"            largest subtree (triangles, 128..512, default 512)
 *                          that one workgroup of the GPU build finishes alone
 *   "overlap_shadow"       1 (default): path-tracing frames trace each bounce's
 *                          shadow rays on a second stream beside the next
 *                          bounce's closest-hit traversal; 0: one stream
 *   "octant_order"         1 (default): shading writes each workgroup's
 *                          continuing paths and shadow rays grouped by the
 *                          octant of their ray direction; 0: in thread order.
 *                          Films and counters are the same either way.
 *   "skip_dark"            path-tracing frames cast no shadow ray whose NEE
 *                          term is exactly zero (an operand zero, the other
 *                          finite: the ray could add nothing).  0: every ray;
 *                          1 (default): in plain frames, while counting
 *                          frames (bn_trace_stats) cast and count the
 *                          reference's rays; 2: counting frames count the
 *                          rays traced.  Films are the same in every mode. */
int bn_set_option(bn_ctx* ctx, const char* key, int64_t value);

typedef struct bn_kernel_time {
    char    name[32];
    int64_t launches;
    double  total_ms;                  /* sum of event-measured durations    */
} bn_kernel_time;
/* Per-kernel totals accumulated since the last reset (n: capacity in,
 * count out). */
int bn_kernel_times(bn_ctx* ctx, bn_kernel_time* out, int32_t* n);
int bn_reset_kernel_times(bn_ctx* ctx);

/* Launch-parameter autotuner (no reference counterpart).  Times the given
 * workload -- plain renders of every item in turn, or with BN_TUNE_BATCH one
 * batched render -- over trace_blocks_per_cu, refill, stack_k and
 * chunk_slots by coordinate descent, and leaves the fastest setting on the
 * context.  Images do not depend on these parameters. */
#define BN_TUNE_BATCH 1
typedef struct bn_tune_result {
    int32_t trace_blocks_per_cu, refill, stack_k, trials;
    int64_t chunk_slots;
    float baseline_ms, best_ms;   /* workload time at the entry / chosen setting */
} bn_tune_result;
int bn_autotune(bn_ctx* ctx, const bn_batch_item* items, int32_t n, int32_t flags, bn_tune_result* out);

/* Differentiable rendering (no reference counterpart; the adjoint of
 * PathTracingIntegrator.Li, PathTracing.fs:14-81).  bn_set_params replaces the
 * albedo and roughness of every material (4 floats each: albedo rgb,
 * roughness; MaterialInfo, Loader.fs:50-73) and the emission of every light
 * (3 floats each; LightInfo, Loader.fs:75-90) of a scene; either array may be
 * NULL to keep it, and its count must equal the scene's.  bn_render_pt is the
 * forward pass.  bn_backward_pt returns the gradient of sum(d_film * film)
 * for the frame bn_render_pt renders with the same parameters, by path-replay
 * backpropagation: every path is re-traced from its seed, nothing per vertex
 * is stored.  Derivatives are taken at fixed sampled directions, light
 * samples and Russian-roulette outcomes, with the importance factors 1/pdf
 * held constant and the MIS weights differentiated (unbiased for the
 * derivative of the expected image, visibility changes excluded).  Pixels
 * whose d_film is zero cost nothing; a path whose radiance is not finite
 * contributes nothing.  d_mat: 4 per material (d albedo rgb, d roughness);
 * d_light: 3 per light (d emission rgb); either may be NULL. */
int bn_set_params(bn_ctx* ctx, int32_t scene, const float* mat_params, int32_t n_mat,
                  const float* light_emission, int32_t n_light);
int bn_backward_pt(bn_ctx* ctx, int32_t scene, const bn_pt_params* params, const float* d_film,
                   float* d_mat, float* d_light);

/* PSSMLTIntegrator.Render (Extensions/Integrator/PSSMLT.fs:379-414).  The film
 * is cleared first (Film.Clear, Render.fs:11) and accumulated with device
 * atomics (the reference's Film.Accumulate is an unsynchronised add). */
int bn_render_pssmlt(bn_ctx* ctx, int32_t scene, const bn_mlt_params* params,
                     float* film_rgb, bn_mlt_stats* stats);

/* Per-path detail of the last bn_render_pssmlt: the bootstrap luminances
 * (BootstrapWeights, PSSMLT.fs:163) and the accepted-mutation count of every
 * chain (the per-chain terms of AcceptedMutationCount, :376).  Either output
 * may be NULL; sizes must equal n_bootstrap / n_chains of that call. */
int bn_pssmlt_last(bn_ctx* ctx, float* bootstrap, int32_t n_bootstrap,
                   int32_t* chain_accepts, int32_t n_chains);

#ifdef __cplusplus
}
#endif
#endif /* BARNACLE_HIP_H */
