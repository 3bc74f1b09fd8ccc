/* C ABI of the MI355X-native Cavalry hot path (libcavalry_amd.so).
 *
 * Plain pointers and sizes only. Every entry point returns a cav_status; the
 * message of the last failure on a context is available from cav_last_error.
 * Nothing calls exit(). One HIP stream per context; entry points are
 * synchronous with respect to the host unless their name ends in _async.
 *
 * Geometry conventions:
 *   triangles : float[n][9]  v0.xyz v1.xyz v2.xyz (a triangle soup, the
 *               reference's Triangle{mesh, index} resolved to its vertices)
 *   rays      : float[r][6]  origin.xyz direction.xyz
 *   bounds    : float[6]     min.xyz max.xyz
 * "_device" variants take pointers to device (HBM) memory the caller owns;
 * the others take host memory and copy it in and out.
 */
#ifndef CAVALRY_AMD_H
#define CAVALRY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CAV_OK = 0,
    CAV_ERR_INVALID = 1,   /* bad argument (null pointer, n < 2, bad morton_bits ...) */
    CAV_ERR_HIP = 2,       /* a HIP runtime call failed */
    CAV_ERR_OOM = 3,       /* device allocation failed */
    CAV_ERR_OVERFLOW = 4,  /* a traversal exceeded the 64-entry stack (results partial) */
    CAV_ERR_PARSE = 5,     /* malformed scene description */
    CAV_ERR_BUSY = 6,      /* serving: no capacity for the request right now */
    CAV_ERR_NOT_FOUND = 7  /* serving: unknown request id */
} cav_status;

typedef struct cav_ctx cav_ctx;
typedef struct cav_bvh cav_bvh;

/* ---- context ---------------------------------------------------------- */
int cav_ctx_create(int device, cav_ctx** out);
int cav_ctx_destroy(cav_ctx* ctx);
const char* cav_last_error(const cav_ctx* ctx);
/* the context's hipStream_t, for callers that record events on it */
void* cav_ctx_stream(cav_ctx* ctx);
int cav_ctx_synchronize(cav_ctx* ctx);
/* library version string */
const char* cav_version(void);

/* ---- LBVH build -------------------------------------------------------- */
/* Per-stage device time of the last build (HIP events on the context stream)
 * and the SAH cost of the root before and after treelet restructuring. */
typedef struct {
    float ms_bounds, ms_morton, ms_sort, ms_karras, ms_refit, ms_treelet, ms_flatten, ms_total;
    float root_cost_lbvh, root_cost_final;
    int n, morton_bits, opt_rounds;
} cav_bvh_stats;

/* Replaces static BVH BVH::build(Triangle*, int, const AABB&, int)
 * (BVH/BVH.h:19, BVH/BVH.cu:317-370) together with optimizeBVH
 * (BVH/Optimization.h:4, BVH/Optimization.cu:509-525): Morton codes
 * (morton_bits 30 or 63; BVH/BVH.cu:14-27 selects 63), stable radix sort,
 * Karras radix tree, bottom-up refit, then opt_rounds treelet rounds.
 * scene_bounds may be NULL: the union of the triangle boxes is used, as
 * Scene::prepareForRender computes it (Core/Scene.cu:131-136).
 * tri_mesh (may be NULL = all 0) gives each triangle's mesh id for the
 * source/target exclusion of shadow rays (BVH/BVH.h:139). n >= 2. */
int cav_bvh_build(cav_ctx* ctx, const float* triangles, int n, const float* scene_bounds,
                  const int32_t* tri_mesh, int morton_bits, int opt_rounds, cav_bvh** out);
int cav_bvh_build_device(cav_ctx* ctx, const float* triangles_dev, int n,
                         const float* scene_bounds, const int32_t* tri_mesh_dev,
                         int morton_bits, int opt_rounds, cav_bvh** out);
/* Rebuilds an existing BVH in place from device triangles of the same count,
 * reusing its device buffers (no allocation). */
int cav_bvh_rebuild_device(cav_ctx* ctx, cav_bvh* bvh, const float* triangles_dev,
                           const float* scene_bounds, int morton_bits, int opt_rounds);
int cav_bvh_destroy(cav_bvh* bvh);
int cav_bvh_stats_get(const cav_bvh* bvh, cav_bvh_stats* out);
/* Device time of every kernel (or copy) of the last build, in launch order,
 * from HIP events recorded on the context stream between launches. Writes up
 * to max entries (names are static strings) and returns the total count. */
int cav_bvh_kernel_times(const cav_bvh* bvh, const char** names, float* ms, int max);
/* Node arrays in the reference's unified id space (BVHRestructureNode,
 * BVH/BVHNode.h:48-66; internal i at i, leaves from N-1 before
 * restructuring): 2n-1 entries each. prim = original triangle index of a leaf
 * or -1; left/right = -1 for leaves; the root's parent is 0. Any pointer may
 * be NULL. sorted_codes (u64[n]) / order (int32[n]) are the sorted Morton
 * codes and the triangle index at each sorted position. */
int cav_bvh_export(cav_ctx* ctx, const cav_bvh* bvh, float* box6, float* surface_area,
                   float* cost, int32_t* left, int32_t* right, int32_t* parent, int32_t* prim,
                   uint64_t* sorted_codes, int32_t* order);

/* ---- traversal --------------------------------------------------------- */
/* Replaces BVH::intersect (BVH/BVH.h:23-107) as used by
 * SceneHandle::intersect (Core/Scene.h:42-51), geometry only: the closest
 * hit's original triangle index (-1 on a miss), distance and barycentric u, v.
 * t, u, v of a miss are 0. u, v may be NULL. */
int cav_trace_closest(cav_ctx* ctx, const cav_bvh* bvh, const float* rays, int64_t nrays,
                      int32_t* hit_prim, float* hit_t, float* hit_u, float* hit_v);
int cav_trace_closest_device(cav_ctx* ctx, const cav_bvh* bvh, const float* rays_dev,
                             int64_t nrays, int32_t* hit_prim_dev, float* hit_t_dev,
                             float* hit_u_dev, float* hit_v_dev);
/* Replaces BVH::testVisibility (BVH/BVH.h:109-189) as used by
 * SceneHandle::testVisibility (Core/Scene.h:53-56): visible[i] = 1 when no
 * triangle blocks ray i before limit[i] (+inf = no limit, the reference's
 * useDistanceLimit = false). Triangles of mesh src_mesh[i] or dst_mesh[i]
 * are ignored (arrays may be NULL = -1). */
int cav_trace_occluded(cav_ctx* ctx, const cav_bvh* bvh, const float* rays, const float* limit,
                       const int32_t* src_mesh, const int32_t* dst_mesh, int64_t nrays,
                       uint8_t* visible);
int cav_trace_occluded_device(cav_ctx* ctx, const cav_bvh* bvh, const float* rays_dev,
                              const float* limit_dev, const int32_t* src_mesh_dev,
                              const int32_t* dst_mesh_dev, int64_t nrays, uint8_t* visible_dev);
/* cav_trace_closest without any device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, results and errors as cav_trace_closest (host buffers), staged
 * through a static device buffer of CAV_NOALLOC_MAX_RAYS rays that each GPU
 * holds until the process ends; larger batches are refused with
 * CAV_ERR_INVALID. Calls on one GPU share that buffer, so calls that overlap
 * on one GPU (several contexts or threads) race. */
#define CAV_NOALLOC_MAX_RAYS 262144
int cav_trace_closest_noalloc(cav_ctx* ctx, const cav_bvh* bvh, const float* rays, int64_t nrays,
                              int32_t* hit_prim, float* hit_t, float* hit_u, float* hit_v);
/* Launch without synchronising (device pointers); errors surface at the next
 * synchronising call. Used by timing harnesses. */
int cav_trace_closest_async(cav_ctx* ctx, const cav_bvh* bvh, const float* rays_dev,
                            int64_t nrays, int32_t* hit_prim_dev, float* hit_t_dev);
int cav_trace_occluded_async(cav_ctx* ctx, const cav_bvh* bvh, const float* rays_dev,
                             const float* limit_dev, int64_t nrays, uint8_t* visible_dev);

/* Traversal health of a BVH's standalone launches (cav_trace_*): out[0] and
 * out[1] = closest-hit / any-hit rays cut short by the reference's 64-entry
 * stack (BVH/BVH.h:41,122; results partial), out[2] and out[3] = the deepest
 * stack any closest-hit / any-hit ray reached. The _async variants only add
 * to these; the synchronous variants clear out[0] / out[1] before their
 * launch. reset != 0 clears all four after reading. No counterpart in the
 * reference, which writes past its stack. */
int cav_trace_counters(cav_ctx* ctx, const cav_bvh* bvh, uint32_t* out, int32_t reset);
/* Work of a BVH's standalone launches (out holds 6 words): out[0] / out[1] =
 * 64-byte traversal records (node visits and triangle tests, one record
 * each) fetched by every closest-hit / any-hit launch since the last reset;
 * out[2] = rays of the last any-hit launch finished by a whole wave (long
 * rays still in flight once every ray was handed out), out[3] = those of
 * them the wave could not hold and one lane traced again from the root;
 * out[4] = rays of the last closest-hit launch traced again by a whole wave
 * from the root, out[5] = those of them one lane traced once more in the
 * reference's order (two hits within a hair of each other, or a stack that
 * could overflow). reset != 0 clears out[0..1] after reading. No counterpart
 * in the reference; it measures the algorithmic traffic of BVH::intersect /
 * testVisibility (BVH/BVH.h:23-189). */
int cav_trace_fetches(cav_ctx* ctx, const cav_bvh* bvh, uint64_t* out, int32_t reset);
/* The deepest leaf of the BVH, in edges from the root: no traversal can push
 * more stack entries than this, so a depth of at most 64 rules out overflow
 * (BVH/BVH.h:41,122). Computed on the device on each call. */
int cav_bvh_max_depth(cav_ctx* ctx, const cav_bvh* bvh, int32_t* depth);

/* ---- scenes and rendering ---------------------------------------------- */
enum { CAV_MAT_MATTE = 0, CAV_MAT_MIRROR = 1, CAV_MAT_SUBSTRATE = 2, CAV_MAT_METAL = 3, CAV_MAT_GLASS = 4,
       CAV_MAT_PLASTIC = 5 };
/* path (Integrators/PathTracingIntegrator.cu), rlpath
 * (Integrators/ReinforcementLearningPathTracing.cu), directlighting
 * (Integrators/DirectLightingIntegrator.cu: one bounce, NEE without MIS) */
enum { CAV_INTEGRATOR_PATH = 0, CAV_INTEGRATOR_RLPATH = 1, CAV_INTEGRATOR_DIRECT = 2 };
/* cav_render_params.flags. SORT_BY_MATERIAL shades each bounce's hits in
 * the order of the reference's sortLightingQueue: by material for path and
 * directlighting (Integrators/PathTracingIntegrator.cu:50-92), by material
 * and then Q entry for rlpath (ReinforcementLearningPathTracing.cu:226-272,
 * stable): the film does not change, only the speed. DETERMINISTIC makes rlpath's table
 * updates (the Q proposals and light statistics that the reference adds with
 * float atomics, ReinforcementLearningPathTracing.cu:453-456,576-585) sum in
 * sample order, so the film is the same on every run; slower (one sort of
 * the updates per bounce). Other integrators ignore it. FILM_F64 (cav_render
 * only) adds each pixel's film sums in binary64 instead of the reference's
 * binary32 (Core/Film.h), in the same order, and rounds them to binary32
 * once: for high-spp reference images. */
enum { CAV_RENDER_SORT_BY_MATERIAL = 1, CAV_RENDER_DETERMINISTIC = 2, CAV_RENDER_FILM_F64 = 4 };

/* One mesh: triangles [first_tri, first_tri + num_tris) of the scene's soup.
 * material follows MaterialType (Core/Material.h:10-12): 0 matte (kd),
 * 1 mirror (kr[0]), 2 substrate (kd, ks, GGX alphas roughness and
 * roughness_v; FresnelBlendBSDF), 3 metal (ks = Schlick F0, GGX alphas;
 * MicrofacetBSDF of colour 1), 4 glass (kr, kt, ior; perfectly specular),
 * 5 plastic (kd, ks, roughness = GGX alpha). roughness_v = 0 means the same
 * alpha on both axes. texture: index into the scene's textures of a Kd
 * texture (matte, plastic, substrate), or -1 for none. is_light marks a
 * DiffuseAreaLight of radiance `emission` (Lights/DiffuseAreaLight.h).
 * sphere = centre and radius when the triangles tessellate a sphere shape
 * (MeshShapeType::Sphere, ShapeMesh/Sphere.cu:84-88): as a light it is
 * sampled within the cone it subtends (TriangleMesh::sampleSphere,
 * Core/TriangleMesh.h:155-211); radius 0 = a plain triangle mesh. */
typedef struct {
    int32_t first_tri, num_tris;
    int32_t material;
    int32_t texture;
    float kd[3], ks[3], kr[3], kt[3];
    float roughness, ior;
    float emission[3];
    int32_t is_light, reverse_orientation;
    float roughness_v;
    float sphere[4];
} cav_mesh;

/* Perspective camera: column-major camera-to-world matrix (glm layout), film
 * depth for unit-size pixels, resolution (Cameras/PerspectiveCamera.h). */
typedef struct {
    float cam_to_world[16];
    float depth;
    int32_t width, height;
} cav_camera;

/* LookAt + fov (degrees, of the shorter image axis) -> camera, as the
 * reference's parser and PerspectiveCamera constructor compute it
 * (SceneLoading/Parsing.cu:196-204,257; Cameras/PerspectiveCamera.cu:6-20). */
int cav_camera_look_at(const float* eye3, const float* center3, const float* up3,
                       float fov_degrees, int width, int height, cav_camera* out);
/* Camera-to-world matrix (column-major, as glm) + fov -> camera: the tail of
 * PerspectiveCamera::createFromParams (Cameras/PerspectiveCamera.cu:22-26). */
int cav_camera_from_matrix(const float* cam_to_world16, float fov_degrees, int width, int height,
                           cav_camera* out);

typedef struct cav_scene cav_scene;

/* Replaces Scene::prepareForRender (Core/Scene.cu:117-151): uploads the
 * triangle soup (float[n][9], meshes as contiguous ranges covering it) and
 * the mesh table, computes mesh areas (TriangleMesh::computeArea) and builds
 * the BVH with bvh_opt_rounds treelet rounds (the reference's default is 3,
 * or 10 when spp >= 64; resolve it before calling). tri_uv (float[n][6]) may
 * be NULL. At least one mesh must be an area light. */
int cav_scene_create(cav_ctx* ctx, const float* triangles, int n, const cav_mesh* meshes,
                     int nmeshes, const float* tri_uv, int bvh_opt_rounds, int morton_bits,
                     cav_scene** out);
/* An RGBA8 image (Texture2D, Core/Texture.h:12-125): width * height texels,
 * row-major, row 0 at texture coordinate v = 0 (the bottom row of an image
 * file, which the reference flips on load), 1 to 65536 texels per side.
 * Read with wrap addressing and bilinear filtering as value / 255. */
typedef struct {
    const uint8_t* rgba;
    int32_t width, height;
} cav_texture;

/* An environment-type light (EnvironmentMap, Core/Light.h:43-122): radiance
 * `color`, times the texel of `texture` (-1: none) at
 * (phi / 2pi, theta / pi) of the direction rotated by world_to_light
 * (row-major 3x3, the upper block of the inverse light-to-world transform).
 * Sampled as a light by a uniform direction on the sphere (pdf 1 / 4pi) and
 * an unlimited shadow ray. A LightSource the reference does not know becomes
 * such a light with color 0 (Lights/LightObject.h:124-126). */
typedef struct {
    float color[3];
    int32_t texture;
    float world_to_light[9];
} cav_env_light;

/* The same with every scene input in one descriptor, including per-vertex
 * shading normals: Triangle::fillIntersectionInformation and
 * TriangleMesh::getNormal (Core/Triangle.h:94-136, Core/TriangleMesh.h:76-104)
 * shade a mesh that has them with normalize(n0 (1-u-v) + u n1 + v n2).
 * Textures: a mesh whose texture field is k >= 0 reads textures[k] at the
 * interpolated tri_uv (all zero when tri_uv is NULL); the texel multiplies a
 * matte colour (Materials/MatteMaterial.h:23-33) or replaces a plastic's
 * diffuse colour (Materials/Plastic.h:38-50). */
typedef struct {
    const float* tris;            /* float[ntris][9] */
    int32_t ntris;
    const cav_mesh* meshes;
    int32_t nmeshes;
    const float* tri_uv;          /* float[ntris][6] or NULL */
    const float* tri_normals;     /* float[ntris][9] per-vertex normals, or NULL */
    const uint8_t* mesh_normals;  /* per mesh 1 = shade with tri_normals (NULL: every mesh) */
    int32_t bvh_opt_rounds;
    int32_t morton_bits;
    const cav_texture* textures;  /* ntextures images, or NULL */
    int32_t ntextures;
    /* Lights beyond the area-light meshes. light_order lists every light in
     * the order the scene defines them (the order randInt / the learnt
     * distribution index): a mesh id >= 0 for an area light, -1 - k for
     * env_lights[k]; NULL means area lights in mesh order, then env_lights.
     * env_background (-1: none; ignored when nenv is 0) is the light that
     * rays which miss every triangle see (Scene::environmentMapIndex). */
    const cav_env_light* env_lights;
    int32_t nenv;
    const int32_t* light_order;
    int32_t nlight_order;
    int32_t env_background;
} cav_scene_desc;
int cav_scene_create_desc(cav_ctx* ctx, const cav_scene_desc* desc, cav_scene** out);
/* Frees the scene. Refused with CAV_ERR_BUSY while a server still holds it:
 * a server holds a scene from the submit of a request on it until that
 * request's final status, and for as long as it keeps a ray pool for it
 * (until cav_serve_close). Close servers before the scenes they render. */
int cav_scene_destroy(cav_scene* scene);
/* the scene's BVH (owned by the scene) */
const cav_bvh* cav_scene_bvh(const cav_scene* scene);

typedef struct {
    int32_t integrator;  /* CAV_INTEGRATOR_PATH, _RLPATH or _DIRECT */
    int32_t spp;         /* samples per pixel */
    int32_t max_depth;   /* bounces (the scene's "integer maxdepth"); ignored by _DIRECT */
    int32_t flags;       /* CAV_RENDER_* bits */
    /* rlpath only. Q entries per axis: 0 = 32, the reference's Q_TABLE_SIZE
     * (Integrators/ReinforcementLearningPathTracing.cu:151); at most 64. */
    int32_t q_grid;
    /* rlpath only. Scale of the learning rate n/(T+n) (QEntry::alpha, :44-50):
     * 0 or 1 = the reference's rule. */
    float alpha_scale;
    /* Samples per pixel per round (Sampler::genAllCameraSamples, sppBySpp
     * order, Core/Sampler.cu:77-96): 0 = as many as free device memory holds
     * (Samplers/DecideSampleCount.cu:7-23). Rounds do not change a path film;
     * rlpath keeps learning across them as the reference does. */
    int32_t spp_per_round;
} cav_render_params;

typedef struct {
    double ms_total;            /* device time of the render (HIP events) */
    int64_t closest_rays;       /* camera + continuation rays traced */
    int64_t shadow_rays;        /* shadow rays traced */
    double mrays_per_s;         /* (closest + shadow) / ms_total */
    int32_t rounds;             /* spp rounds the render took */
} cav_render_stats;

/* Replaces Renderer::render -> Integrator::render (Core/Renderer.cu:3-6,
 * Integrators/PathTracingIntegrator.cu:318-425): renders all spp samples of
 * every pixel and writes the box-filtered film as per-pixel sums (rgb,
 * float[h][w][3]) and weights (float[h][w]); the image is rgb / weight. */
int cav_render(cav_ctx* ctx, cav_scene* scene, const cav_camera* camera,
               const cav_render_params* params, float* film_rgb, float* film_weight,
               cav_render_stats* stats);

/* ---------------------------------------------------------------- PBRT front end */

/* The CLI overrides of main.cu:25-97 (cxxopts --output, --integrator, --spp,
 * --bvh-opt-rounds), applied while parsing as SceneLoading/Parsing.cu does. */
typedef struct {
    const char* output;      /* Film "filename" (NULL or "": keep the file's) */
    const char* integrator;  /* Integrator name (NULL or "": keep the file's) */
    int32_t spp;             /* Sampler "pixelsamples" (0: keep the file's) */
    int32_t bvh_opt_rounds;  /* -1: 3, or 10 when spp >= 64 (Core/Scene.cu:141-146) */
    int32_t morton_bits;     /* 0: 63, the reference's compiled default (BVH/BVH.cu:14) */
} cav_pbrt_overrides;

/* What RenderSetup holds after parsing (Core/RenderSetup.h), resolved. */
typedef struct {
    cav_camera camera;
    int32_t width, height;
    int32_t integrator;      /* CAV_INTEGRATOR_*; unknown names give path, depth 8 */
    int32_t max_depth;
    int32_t spp;
    int32_t bvh_opt_rounds;
    int32_t morton_bits;
    int32_t warnings;        /* unsupported features replaced or ignored */
    int64_t ntris;
    int32_t nmeshes, nlights;
    char output[1024];       /* output file name ("output.png" by default) */
} cav_render_setup;

typedef struct cav_pbrt cav_pbrt;

/* Replaces runLexing + runParsing (SceneLoading/Lexing.cu:8-50,
 * SceneLoading/Parsing.cu:391-405) on the host: the scene as a triangle soup
 * and mesh table (the input of cav_scene_create) plus the render setup.
 * CAV_ERR_PARSE with the message in err on failure. No device work. */
int cav_pbrt_parse(const char* path, const cav_pbrt_overrides* overrides, cav_pbrt** out, char* err,
                   int err_len);
/* Views into a parsed scene (owned by it): float[ntris][9], cav_mesh[nmeshes],
 * the setup and the warnings, one per line. */
int cav_pbrt_get(const cav_pbrt* p, const float** tris, const cav_mesh** meshes, cav_render_setup* setup,
                 const char** messages);
/* Per-vertex normals of a parsed scene (float[ntris][9] and one flag per
 * mesh), or NULLs when no mesh has any ("normal N", PLY nx/ny/nz, spheres). */
int cav_pbrt_get_normals(const cav_pbrt* p, const float** tri_normals, const uint8_t** mesh_normals);
/* Per-vertex texture coordinates (float[ntris][6], NULL when no shape gave
 * "uv" or PLY u/v) and the textures that cav_mesh::texture indexes
 * ("Texture" directives, PNG / TGA files decoded as Texture2D receives them:
 * rows bottom-up, sRGB curve removed). Owned by the parsed scene. */
int cav_pbrt_get_textures(const cav_pbrt* p, const float** tri_uv, const cav_texture** textures,
                          int32_t* ntextures);
/* The light list of a parsed scene (cav_scene_desc's env_lights,
 * light_order, env_background): "infinite" LightSources with their colour,
 * map and rotation; other LightSource / AreaLightSource types as black
 * lights; all in directive order. Owned by the parsed scene. */
int cav_pbrt_get_lights(const cav_pbrt* p, const cav_env_light** env_lights, int32_t* nenv,
                        const int32_t** light_order, int32_t* nlight_order, int32_t* env_background);
void cav_pbrt_destroy(cav_pbrt* p);
/* Replaces readRenderSetup + Scene::prepareForRender (main.cu:10-15):
 * parse, then cav_scene_create with the setup's BVH rounds and Morton width. */
int cav_scene_create_from_pbrt(cav_ctx* ctx, const char* path, const cav_pbrt_overrides* overrides,
                               cav_scene** out, cav_render_setup* setup);

/* ---------------------------------------------------------------- launch shapes */

/* Launch shape of the persistent traversal kernels, per context: workgroups
 * per CU (0: as many as fit), rays a wave takes from the dispenser at a time
 * (64..65536), traversal-stack entries kept in LDS (32 or 48; the rest of
 * the reference's 64 in scratch) and threads per traversal workgroup (64,
 * 128 or 256; 0: 256); for WAVEFRONT also the threads per workgroup of the
 * per-sample shading kernels (path shading, the RL integrator's Q-table
 * sampling; 64, 128 or 256; 0: 256), and the steps a ray still in flight
 * once every ray has been handed out must have taken to be finished by a
 * whole wave instead of its lane (1..1<<30; 0: 1024; 1<<30: never, and the
 * cooperative launches are skipped), and what one any-hit step fetches
 * (0 or CAV_ANY_STEP_NODE: one node; CAV_ANY_STEP_PAIR: a pair of sibling
 * nodes sharing a 128-byte line, both tested), and whether any-hit rays
 * visit boxes that start beyond their limit (1: every box the ray's line
 * crosses, the reference's visiting order; 0, the default: a box starting
 * more than 2^-10 (relative) beyond the limit is skipped, since no triangle
 * in it can block the ray unless Moller-Trumbore's t is off by more than
 * that, which takes a ray within about 1e-4 rad of a sliver's plane).
 * Each ray's closest-hit node order is the reference's, the any-hit answer
 * does not depend on the order the reachable leaves are tested in, and each
 * sample's shading is its own, so results do not depend on the shape. STANDALONE applies to cav_trace_*, WAVEFRONT to
 * cav_render and to servers opened on the context afterwards. No
 * counterpart in the reference (its launch shapes are compile-time, 256
 * threads: Utils/GpuCommons.h:63); cavalry_amd/tune.py times the candidates
 * and keeps the fastest in a JSON cache. */
typedef struct {
    int32_t workgroups_per_cu;
    int32_t ray_chunk;
    int32_t lds_stack;
    int32_t workgroup_size;
    int32_t shade_workgroup_size;
    int32_t handoff_steps;
    int32_t any_hit_step;
    int32_t any_hit_whole_line;
} cav_trace_launch;
enum { CAV_TRACE_STANDALONE = 0, CAV_TRACE_WAVEFRONT = 1 };
enum { CAV_ANY_STEP_NODE = 1, CAV_ANY_STEP_PAIR = 2 };
int cav_trace_set_launch(cav_ctx* ctx, int32_t which, const cav_trace_launch* launch);
int cav_trace_get_launch(cav_ctx* ctx, int32_t which, cav_trace_launch* out);

/* ---------------------------------------------------------------- serving */

/* No counterpart in the reference (it renders one scene per process): a
 * render server over resident scenes with continuous batching. All requests
 * on one scene share a persistent ray pool: camera rays of new requests join
 * at every bounce step, and tasks of many requests, integrators and depths
 * share each launch. path and directlighting requests are admitted in chunks
 * of samples; each film is bit-identical to the same request rendered alone
 * by cav_render. An rlpath request (whose learning couples the samples of a
 * bounce) is admitted one round at a time, spp_per_round samples of every
 * pixel (0: as many as chunk_samples holds, at least 1), with its own Q table
 * and light statistics; its film is cav_render's with the same
 * spp_per_round (bit for bit with CAV_RENDER_DETERMINISTIC; at most 32
 * deterministic rlpath requests per pool at once). The bounce sequence of a
 * pool can be replayed from graphs built once per pool (use_graphs).
 * Admission is bounded by the pool's task slots and by a device arena for
 * per-request sample, film and table buffers; the server allocates nothing
 * while serving. All functions are thread-safe. */
typedef struct cav_server cav_server;
typedef struct {
    int64_t arena_bytes;       /* device arena for admitted requests (0: 16 GiB); 20 B/sample + 160 B/pixel */
    int64_t host_arena_bytes;  /* pinned staging for finished films (0: 1 GiB); 16 B/pixel */
    int64_t pool_tasks;        /* ray-pool capacity per scene (0: 4M tasks, ~170 B each) */
    int64_t chunk_samples;     /* admission quantum per request per step (0: pool_tasks / 4) */
    int32_t max_steps_ahead;   /* bounce steps the host may enqueue ahead of the device (0: 4) */
    int32_t use_graphs;        /* replay each pool's bounce sequences from hipGraphs built once per pool */
    int32_t rl_workers;        /* unused (rlpath requests share the pools) */
} cav_serve_config;
enum { CAV_SERVE_QUEUED = 0, CAV_SERVE_RUNNING = 1, CAV_SERVE_DONE = 2, CAV_SERVE_CANCELLED = 3,
       CAV_SERVE_FAILED = 4 };
typedef struct {
    int32_t status;                   /* CAV_SERVE_* */
    int32_t error;                    /* cav_status when FAILED */
    double submit_ms, start_ms, done_ms;  /* server clock (ms since cav_serve_open); -1 if not reached */
    int64_t closest_rays, shadow_rays;    /* 0: requests share their pool's ray tallies (cav_serve_stats) */
} cav_serve_result;
typedef struct {
    int64_t pool_closest_rays, pool_shadow_rays, rl_closest_rays, rl_shadow_rays;
    int64_t steps;                    /* bounce steps over all pools */
    int32_t pools;
    int32_t graph_pools;              /* pools that replayed their bounce steps from graphs */
} cav_serve_stats;

int cav_serve_open(cav_ctx* ctx, const cav_serve_config* cfg, cav_server** out);
/* Queues a render of `scene` (created by cav_scene_create on the same
 * device) with the camera and parameters of cav_render. The film buffers
 * (float[h][w][3] sums and float[h][w] weights) must stay valid until poll
 * returns a final status for the request or cav_serve_close returns (close
 * finishes every admitted request and writes its film). */
int cav_serve_submit(cav_server* s, cav_scene* scene, const cav_camera* camera, const cav_render_params* params,
                     float* film_rgb, float* film_weight, int64_t* id);
/* Status of a request; wait != 0 blocks until it is final. Returning a final
 * status (DONE, CANCELLED, FAILED) releases the id. */
int cav_serve_poll(cav_server* s, int64_t id, int32_t wait, cav_serve_result* out);
/* Cancels a request: queued ones are dropped, running ones finish without
 * writing the caller's film. */
int cav_serve_cancel(cav_server* s, int64_t id);
/* Drains outstanding work, then frees the server; stats may be NULL. */
int cav_serve_close(cav_server* s, cav_serve_stats* stats);

/* ---------------------------------------------------------------- film output */

/* Replaces Film::readCurrentResult + apply + writeColorAt (Core/Film.cu:20-46,
 * Core/Color.h:53-81): per pixel rgb / weight, clamped to [0, 1], sRGB gamma,
 * times 255 truncated to u8. Host only; out is uint8[h][w][3]. */
int cav_film_to_srgb8(const float* film_rgb, const float* film_weight, int width, int height,
                      uint8_t* out);
/* Replaces RenderResult::saveToPNG (Core/RenderResult.cu:9-12): 8-bit RGB PNG,
 * stored (uncompressed) deflate blocks. */
int cav_write_png(const char* path, const uint8_t* rgb8, int width, int height);
/* No reference counterpart (the reference writes 8-bit PNG only): the film as
 * little-endian colour PFM of rgb / weight (weight may be NULL: rgb as is). */
int cav_write_pfm(const char* path, const float* film_rgb, const float* film_weight, int width,
                  int height);

#ifdef __cplusplus
}
#endif
#endif
