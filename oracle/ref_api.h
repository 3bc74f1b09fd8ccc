// CPU oracle: C entry points loaded by the tests and bench.py (ctypes).
//
// TEST INFRASTRUCTURE ONLY. The product never links or calls these.
#pragma once
#include <cstdint>

// One mesh of a scene; triangles [first_tri, first_tri + num_tris) of the
// soup belong to it. Same layout as cav_mesh in include/cavalry.h.
typedef struct {
    int32_t first_tri, num_tris;
    int32_t material;  // 0 matte, 1 mirror, 2 substrate, 3 metal, 4 glass, 5 plastic (Core/Material.h:10-12)
    int32_t texture;   // -1: none
    float kd[3], ks[3], kr[3], kt[3];
    float roughness, ior;
    float emission[3];
    int32_t is_light, reverse_orientation;
    float roughness_v;  // 0: same as roughness
    float sphere[4];    // centre, radius of an analytic sphere shape; radius 0: none
} cav_mesh_ref;

extern "C" {
void ref_scene_bounds(const float* tris, int n, float* out6);
void ref_morton(const float* tris, int n, const float* scene6, int bits, uint64_t* codes);
int ref_lbvh_build(const float* tris, int n, int bits, const float* scene6_or_null,
                   uint64_t* sorted_codes, int32_t* order, float* box, float* sa, float* cost,
                   int32_t* left, int32_t* right, int32_t* parent, int32_t* prim);
int ref_treelet_optimize(int n, int rounds, float* box, float* sa, float* cost, int32_t* left,
                         int32_t* right, int32_t* parent, int32_t* prim, int64_t* treelets_done);
int ref_trace_closest(const float* box, const int32_t* left, const int32_t* right,
                      const int32_t* prim, const float* tris, const float* rays, int64_t nrays,
                      int32_t* hit_prim, float* hit_t, float* hit_u, float* hit_v,
                      int32_t* visits, int32_t* max_stack);
int ref_trace_occluded(const float* box, const int32_t* left, const int32_t* right,
                       const int32_t* prim, const float* tris, const int32_t* mesh,
                       const float* rays, const float* limit, const int32_t* src_mesh,
                       const int32_t* dst_mesh, int64_t nrays, uint8_t* visible);

// render oracles (ref_render.cpp)
// An environment-type light; same layout as cav_env_light in include/cavalry.h.
typedef struct {
    float color[3];
    int32_t texture;          // -1: none
    float world_to_light[9];  // row-major
} ref_env_light;

// Inputs beyond the triangles: per-vertex normals, texture coordinates, the
// Kd textures that cav_mesh_ref::texture indexes, and lights besides the
// area-light meshes (any pointer may be null).
typedef struct {
    const float* tri_normals;     // float[ntris][9]
    const uint8_t* mesh_normals;  // per mesh 1 = shade with tri_normals (null: every mesh)
    const float* tri_uv;          // float[ntris][6]
    const uint8_t* texels;        // RGBA8 texels of every texture, concatenated
    const int64_t* tex_offset;    // per texture: first texel
    const int32_t* tex_size;      // per texture: width, height
    const ref_env_light* env;     // nenv environment-type lights
    int32_t nenv;
    const int32_t* light_order;   // the light list (mesh id >= 0, or -1 - k for env[k]); null: area lights then env
    int32_t nlight_order;
    int32_t env_background;       // env light that missing rays see, -1: none
} ref_surface;
void ref_tex_fetch(const uint8_t* rgba, int w, int h, float u, float v, float* out3);
void ref_env_eval(const ref_surface* surf, int e, const float* dirs, int n, float* out);
void ref_bsdf_eval(const cav_mesh_ref* m, const float* wi, const float* wo, int n, float* f_out, float* pdf_out);
void ref_bsdf_sample(const cav_mesh_ref* m, const float* u2, const float* wo, int n, float* wi_out, float* f_out,
                     float* pdf_out);
void ref_halton_table(uint32_t* primes, uint32_t* start, uint16_t* perm, int perm_cap);
float ref_halton(uint64_t index, int dim);
void ref_mesh_areas(const float* tris, const cav_mesh_ref* meshes, int nmeshes, float* area);
int ref_render_path(const float* tris, int ntris, const int32_t* tri_mesh, const cav_mesh_ref* meshes,
                    int nmeshes, const float* box, const int32_t* left, const int32_t* right,
                    const int32_t* prim, const ref_surface* surf, const float* cam16, float cam_depth, int width, int height,
                    int spp, int max_depth, float* film_rgb, float* film_w, int64_t* rays_traced);
int ref_render_path_crop(const float* tris, int ntris, const int32_t* tri_mesh, const cav_mesh_ref* meshes,
                         int nmeshes, const float* box, const int32_t* left, const int32_t* right,
                         const int32_t* prim, const ref_surface* surf, const float* cam16, float cam_depth,
                         int width, int height, int spp, int max_depth, const int32_t* crop, float* film_rgb,
                         float* film_w, float* samples_out, int64_t* rays_traced);
int ref_render_direct(const float* tris, int ntris, const int32_t* tri_mesh, const cav_mesh_ref* meshes,
                      int nmeshes, const float* box, const int32_t* left, const int32_t* right,
                      const int32_t* prim, const ref_surface* surf, const float* cam16, float cam_depth, int width, int height,
                      int spp, float* film_rgb, float* film_w, int64_t* rays_traced);
int ref_render_rlpath(const float* tris, int ntris, const int32_t* tri_mesh,
                      const cav_mesh_ref* meshes, int nmeshes, const float* box,
                      const int32_t* left, const int32_t* right, const int32_t* prim, const ref_surface* surf,
                      const float* scene6, int q_grid, float alpha_scale, const float* cam16,
                      float cam_depth, int width, int height, int spp, int max_depth,
                      int spp_per_round, float* film_rgb, float* film_w, int64_t* rays_traced);
int ref_rl_dir_to_cell(const float* dir3);
void ref_rl_cell_dir(float rx, float ry, int cell, float* out3);
void ref_rl_finish_q(float* Q, float* newQ, int32_t* cnt, int32_t* total, int nent, float alpha_scale);
void ref_rl_finish_light(float* ldist, float* lres, int32_t* lcount, float* lnew, int32_t* lnewc,
                         int nent, int L);
}
