/*
 * oracle.h -- CPU restatement of the Barnacle renderer's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  The product (barnacle_amd) never links or calls it.
 *
 * It restates, in plain C++ compiled with -ffp-contract=off, the F# files
 * Util/Hash.fs, Base/Sampler.fs, Util/AliasTable.fs, Util/BVH.fs, Base/Ray.fs,
 * Base/Primitive.fs, Base/Camera.fs, Base/Light.fs, Base/Film.fs,
 * Base/Integrator.fs, Base/Scene.fs, Extensions/Camera/{Pinhole,ThinLens}.fs,
 * Extensions/Primitive/{Mesh,Sphere}.fs, Extensions/Aggregate/BVH.fs,
 * Extensions/Material/{Lambertian,Mirror,Dielectric,PBR}.fs, Extensions/LightSampler/Uniform.fs and
 * Extensions/Integrator/{PathTracing,PSSMLT,Normal,Direct}.fs
 * (each function in oracle.cpp cites the lines it follows).
 *
 * The scene description structs have the same layout as bn_scene_desc in
 * include/barnacle_hip.h (checked by tests/test_abi.py) but are declared
 * independently so that the oracle depends on nothing in the product.
 */
#ifndef BARNACLE_ORACLE_H
#define BARNACLE_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct { int32_t kind, vert_offset, n_verts, index_offset, n_indices; float radius; } orc_primitive;
typedef struct { int32_t kind; float albedo[3]; float ior, roughness, metallic; } orc_material;
typedef struct { int32_t kind; float emission[3]; int32_t two_sided; } orc_light;
typedef struct { int32_t primitive, material, light; } orc_instance;
typedef struct { int32_t flags; float scale[3], rotation[3], translation[3]; float time; int32_t n_keys; } orc_transform;
typedef struct { int32_t transform, has_camera, inst_offset, n_inst, child_offset, n_child; } orc_node;
typedef struct { int32_t kind; float fov, aspect, aperture, focus_distance, push_forward; } orc_camera;
typedef struct {
    const float* vertices;            int64_t n_vertex_floats;
    const int32_t* indices;           int64_t n_indices;
    const orc_primitive* primitives;  int32_t n_primitives;
    const orc_material* materials;    int32_t n_materials;
    const orc_light* lights;          int32_t n_lights;
    const orc_instance* instances;    int32_t n_instances;
    const orc_transform* transforms;  int32_t n_transforms;
    const orc_node* nodes;            int32_t n_nodes;
    const int32_t* node_refs;         int32_t n_node_refs;
    int32_t root;
    orc_camera camera;
    float time;
} orc_scene_desc;

typedef struct { int32_t width, height, spp, max_depth, rr_depth, frame_id; } orc_pt_params;
typedef struct {
    int32_t width, height, spp, max_depth, rr_depth, frame_id, n_bootstrap, n_chains, strategy;
    float sigma, eps_min, eps_max, large_step_prob;
} orc_mlt_params;
typedef struct {
    int64_t rays_closest, rays_shadow;
    int64_t nodes_closest, tris_closest, insts_closest;
    int64_t nodes_shadow, tris_shadow, insts_shadow;
    int64_t vertices, paths;
} orc_stats;
typedef struct { int64_t accepted, proposed; float b; int32_t mutations_per_chain; } orc_mlt_stats;
typedef struct {
    float pmin[3], pmax[3];
    int32_t right_or_offset;
    uint8_t is_leaf; int8_t split_axis; int8_t count; uint8_t mask;
} orc_bvh_node;

typedef struct orc_scene orc_scene;

const char* orc_last_error(void);
int  orc_struct_sizes(int32_t* out, int32_t n);

/* Hash.fs / Sampler.fs */
uint32_t orc_xxhash32_two(uint32_t x, uint32_t y);
uint32_t orc_xxhash32_three(uint32_t x, uint32_t y, uint32_t z);
float    orc_lcg(uint32_t* state);
/* ErfInv of the Gaussian mutation (PSSMLT.fs:69-94). */
float    orc_erfinv(float x);

/* AliasTable.fs:14-62 */
int orc_alias_build(const float* w, int32_t n, int32_t* alias, float* prob, float* pdf);
int orc_alias_sample(const int32_t* alias, const float* prob, const float* pdf, int32_t n,
                     float u, int32_t* idx, float* out_pdf);

/* BVHNode.Build (Util/BVH.fs:239-247) over item bounds (6 floats each). */
int orc_bvh_build(const float* bounds, int32_t n, orc_bvh_node* nodes, int32_t* n_nodes, int32_t* perm);

/* Material.Sample / Eval in the local frame (Extensions/Material/{Lambertian,Mirror,Dielectric,PBR}.fs).
 * out = {bsdf.x, bsdf.y, bsdf.z, pdf, wi.x, wi.y, wi.z} */
int orc_material_sample(const orc_material* m, const float* wo, float u_lobe, const float* u2, float* out7);
int orc_material_eval(const orc_material* m, const float* wo, const float* wi, float* out4);
int orc_material_eval_n(const orc_material* m, int32_t n, const float* wo, const float* wi, float* out4);

int  orc_scene_create(const orc_scene_desc* d, orc_scene** out);
void orc_scene_destroy(orc_scene* s);
int  orc_scene_export(orc_scene* s, orc_bvh_node* tlas, int32_t* n_tlas, int32_t* inst_order,
                      int32_t* n_inst, float* o2w, float* w2o, int32_t* light_insts,
                      int32_t* n_lights, float* camera_to_world);
int  orc_scene_export_blas(orc_scene* s, int32_t primitive, orc_bvh_node* nodes, int32_t* n_nodes,
                           int32_t* tri_indices, int32_t* n_tris, float* alias_pdf);

/* integrator: 0 normal, 1 direct, 2 path tracing.  Renders the pixel window
 * [x0,x1)x[y0,y1) (all pixels when x1 <= x0) into film_rgb (W*H*3, Film.Pixels
 * order), using n_threads worker threads over 16x16 tiles. */
int orc_render(orc_scene* s, int32_t integrator, const orc_pt_params* p, int32_t n_threads,
               float* film_rgb, orc_stats* stats, int32_t x0, int32_t y0, int32_t x1, int32_t y1);

/* Radiance of one camera sample (Integrator.fs:34-41) of pixel (x,y). */
int orc_trace_sample(orc_scene* s, const orc_pt_params* p, int32_t x, int32_t y, int32_t sample_id,
                     float* rgb);

/* Differentiable rendering: replace material albedo + roughness (4 floats
 * each) and light emission (3 floats each); either array may be NULL. */
int orc_set_params(orc_scene* s, const float* mat_params, int32_t n_mat, const float* light_emission,
                   int32_t n_light);
/* FP64 adjoint of the path-traced frame: gradients of sum(d_film * film) with
 * respect to material albedo rgb + roughness (4 doubles each) and light
 * emission (3 doubles each), over the pixel window (empty window: the whole
 * film), with sampled directions, light samples, Russian roulette, pdfs and
 * MIS weights held fixed (the definition the HIP adjoint implements). */
int orc_adjoint_pt(orc_scene* s, const orc_pt_params* p, const float* d_film, int32_t n_threads, int32_t x0,
                   int32_t y0, int32_t x1, int32_t y1, double* d_mat, double* d_light);
/* PBRMaterial.Eval (PBR.fs:37-49) in double and its analytic derivatives,
 * for checking them numerically: f[0..2] = f, f[3] = pdf; d_albedo[c] =
 * d f_c / d albedo_c; d_alpha[0..2] = d f / d alpha, d_alpha[3] = d pdf / d alpha. */
int orc_pbr_grad_d(const double* albedo, double alpha, double metallic, const double* wo, const double* wi,
                   double* f, double* d_albedo, double* d_alpha);

/* Camera ray of sample sample_id of pixel (x,y) (Integrator.fs:35-39,
 * Camera.fs:12-20): origin in out[0..2], direction in out[3..5]. */
int orc_camera_ray(orc_scene* s, const orc_pt_params* p, int32_t x, int32_t y, int32_t sample_id, float* out6);
int orc_camera_ray_u(orc_scene* s, int32_t W, int32_t H, int32_t x, int32_t y, const float* u_pix,
                     const float* u_lens, float* out6);

/* First hit of sample sample_id of every pixel (pixel order y*W+x). */
int orc_first_hit(orc_scene* s, const orc_pt_params* p, int32_t sample_id,
                  int32_t* inst, int32_t* prim, float* t);

/* PSSMLTIntegrator.Render (PSSMLT.fs:379-414).  chain_end <= chain_begin means
 * all chains.  Bootstrap weights are written to `bootstrap` when non-NULL
 * (n_bootstrap floats) and per-chain accept counts to `chain_accepts` when
 * non-NULL (n_chains ints, zero outside the rendered chain range). */
int orc_render_pssmlt(orc_scene* s, const orc_mlt_params* p, int32_t n_threads, float* film_rgb,
                      orc_mlt_stats* stats, int32_t chain_begin, int32_t chain_end,
                      float* bootstrap, int32_t* chain_accepts);

#ifdef __cplusplus
}
#endif
#endif
