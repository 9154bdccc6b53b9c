/*
 * specrel_scene.h — host-side scene producer: the caller side of the drop-in boundary.
 *
 * In the reference these steps run in Rust before `GpuRenderer::new`/`update_scene`:
 * mesh and worldline sampling (src/mesh4d.rs:452-1087), worldtube extrusion
 * (mesh4d.rs:640-854), the 8-bin 4D SAH `Bvh4D::build` (src/bvh4d.rs:122-427), and packing
 * into the GPU records (`PackedSceneGpuData::from_scene`, src/gpu.rs:247-289). This library
 * restates them in C++ so the renderer can be fed and measured without a Rust toolchain, and
 * adds the synthetic workload generators.
 *
 * Errors: functions return 0 on success and a negative code otherwise; the message is
 * available from specrel_scene_last_error() (per thread).
 */
#ifndef SPECREL_SCENE_H
#define SPECREL_SCENE_H

#include "specrel_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct specrel_host_scene specrel_host_scene;
typedef struct specrel_builder specrel_builder;

/* Mesh kinds for specrel_builder_add_tube. */
enum {
    SPECREL_MESH_CUBE = 0,       /* params: half extents hx, hy, hz       (Mesh3D::cube, mesh4d.rs:454) */
    SPECREL_MESH_UV_SPHERE = 1,  /* params: radius, lat_steps, lon_steps  (Mesh3D::uv_sphere, 508) */
    SPECREL_MESH_GROUND = 2,     /* params: half_x, half_z, y, divs_x, divs_z (ground_grid, 566) */
    SPECREL_MESH_ICOSPHERE = 3   /* params: radius, subdivisions (320 triangles at 2) */
};
/* Worldline kinds for specrel_builder_add_tube / specrel_worldline_sample. */
enum {
    SPECREL_WL_INERTIAL = 0,     /* d: pos x,y,z, beta x,y,z, t_min, t_max, steps (mesh4d.rs:912) */
    SPECREL_WL_HYPERBOLIC = 1,   /* d: centre x,y,z, dir x,y,z, accel a, t_turn, max a|t-t_turn|, t_min, t_max, steps */
    SPECREL_WL_TWIN = 2,         /* d: t_offset, y, z, t_min, t_max, steps (scene.rs:124-138, 333-341) */
    SPECREL_WL_CIRCLE = 3,       /* d: radius, beta, t_min, t_max, steps (orbit in the xz plane) */
    SPECREL_WL_BORN_RAMP = 4,    /* d: beta_max, t_min, t_max, steps (mesh4d.rs:1254-1267 profile) */
    SPECREL_WL_OSCILLATE = 5,    /* d: cx, amp, omega, y, z, t_min, t_max, steps: x = cx + amp sin(omega t)
                                    (scene.rs:553-563) */
    SPECREL_WL_ORBIT = 6         /* d: cx, cy, cz, radius, omega, t_ref, phase, wobble, t_min, t_max, steps:
                                    angle = omega (t - t_ref) + phase, (cx + r cos, cy + wobble sin 2angle,
                                    cz + r sin) (scene.rs:567-578, 683-700) */
};

typedef struct specrel_object {
    uint32_t prim_first, prim_count;
    float center[4]; /* (t, x, y, z) */
    uint32_t material, pad;
} specrel_object;

typedef struct specrel_camera {    /* Camera4D (renderer.rs:45-54) */
    float position[4];
    float e[4][4];
    float proper_time;
    float fov_y_rad;
    float pad[2];
} specrel_camera;

typedef struct specrel_render_config {  /* RenderConfig (renderer.rs:160-168) */
    uint32_t width, height;
    uint32_t mode, shadow_mode, slice_mode;
    float prophecy_delta;
} specrel_render_config;

typedef struct specrel_flash {     /* FlashEvent (renderer.rs:185-190) */
    float at[4];
    float width, intensity;
    float color[3];
    float pad;
} specrel_flash;

typedef struct specrel_dynamics {  /* FrameDynamics (renderer.rs:205-218) */
    uint32_t has_frontier;
    float frontier_origin[4];
    float frontier_proper_time;
    float frontier_lag;
    uint32_t num_flashes;
    const specrel_flash *flashes;
    uint32_t num_remote;
    const float (*remote_origins)[4];
} specrel_dynamics;

/* Synthetic workload parameters (generators are seeded SplitMix64). */
typedef struct specrel_gen_params {
    uint64_t seed;
    uint32_t n_objects;          /* large scene: objects */
    uint32_t slices;             /* worldline samples - 1 */
    uint32_t n_stars;
    uint32_t n_clocks;           /* light-clock worldtubes (accelerated scene) */
    float icosphere_fraction;    /* share of icospheres among objects (rest are cubes) */
    float accel_fraction;        /* share of objects with constant proper acceleration */
    float accel_min, accel_max;
    float t_min, t_max;          /* worldtube span in lab time */
    uint32_t first_object;       /* subset: skip this many objects of the sequence */
    uint32_t pad;
} specrel_gen_params;

const char *specrel_scene_last_error(void);

/* Defaults for a named generator ("cube", "large", "accel_clocks"). */
int specrel_gen_defaults(const char *name, specrel_gen_params *out);
/* Named scenes: "cube" (2304-tet cube), "large" (icosphere field + star sphere),
 * "accel_clocks" (Born-rigid accelerated bodies + light clocks), the reference's five demo scenes
 * "twin" (TwinParadox, scene.rs:283-379), "barn" (BarnAndPole, 383-455), "penrose"
 * (TerrellPenroseStreet, 460-571), "train" (EinsteinTrainSync, 576-643), "carousel"
 * (RelativisticCarousel, 648-719), and "decision_age" (gpu.rs:1156-1185). p may be NULL. */
int specrel_scene_generate(const char *name, const specrel_gen_params *p, specrel_host_scene **out);
/* Suggested observer for a named scene (config camera / DemoScene::initial_camera_pos). */
int specrel_scene_default_camera(const char *name, const specrel_gen_params *p, specrel_camera *out);

int specrel_scene_get_arrays(const specrel_host_scene *s, specrel_scene_arrays *out);
/* One entry per worldtube added (generators and specrel_builder_add_tube): its primitives are
 * input indices [prim_first, prim_first + prim_count) (scene order[] maps leaf slots to them);
 * center is its worldline sample nearest to lab time 0. Returns the object count; copies at
 * most cap entries to out (which may be NULL). */
uint32_t specrel_scene_objects(const specrel_host_scene *s, specrel_object *out, uint32_t cap);
/* Star list as generated (temperature, intensity per star) for tests of pack_stars. */
int specrel_scene_star_params(const specrel_host_scene *s, float *temp_intensity, uint32_t n);
void specrel_scene_free(specrel_host_scene *s);

/* Incremental construction. */
specrel_builder *specrel_builder_new(void);
void specrel_builder_free(specrel_builder *b);
int specrel_builder_add_tube(specrel_builder *b, uint32_t mesh_kind, const float *mesh_params,
                             uint32_t wl_kind, const double *wl_params, float sync_t, uint32_t do_sync,
                             uint32_t material, const float *decided_at, uint32_t n_decided);
/* Simplex4D::new on explicit vertices (mesh4d.rs:312-371); returns 1 if added, 0 if degenerate. */
int specrel_builder_add_simplex(specrel_builder *b, const float verts[16], const float uv[8],
                                const float tri_bary[12], const float tau[4], const float grad[8],
                                const float uv_world_scale[2], const float normal[4],
                                const float four_vel[4], const float decided_at[4], uint32_t material);
/* PointStar::moving (star.rs:85-93) or PointStar::along (97-112) when n_events > 1. */
int specrel_builder_add_star(specrel_builder *b, const float *pos_or_events, uint32_t n_events,
                             const float beta[3], float temperature, float intensity);
int specrel_builder_set_light(specrel_builder *b, float x, float y, float z);
int specrel_builder_set_materials(specrel_builder *b, const specrel_material *m, uint32_t n);
uint32_t specrel_builder_num_simplices(const specrel_builder *b);
/* Bvh4D::build + packing; the builder stays valid (may be finished again). */
int specrel_builder_finish(specrel_builder *b, specrel_host_scene **out);

/* Worldline sampling (for physics checks). out holds steps+1 samples. */
int specrel_worldline_sample(uint32_t wl_kind, const double *wl_params, specrel_wl_sample *out,
                             uint32_t cap, uint32_t *n_out);
int specrel_sync_proper_time(specrel_wl_sample *s, uint32_t n, float t_sync);
/* Extrusion of one tube into raw records (input order): geoms, attrs; returns count. */
int specrel_extrude(uint32_t mesh_kind, const float *mesh_params, const specrel_wl_sample *s,
                    uint32_t n_samples, uint32_t material, specrel_geom *geoms, specrel_attr *attrs,
                    uint32_t cap, uint32_t *n_out);
/* extend_worldtube_simplices (mesh4d.rs:721-854) for n tubes on the host, the records of tube i
 * following those of tube i-1: the CPU counterpart of specrel_extrude_tubes (include/specrel.h).
 * boxes (may be NULL) receive each record's padded bounds (min t,x,y,z, max t,x,y,z); counts
 * (may be NULL) the records per tube. Fails with the reference's message when a Born-rigid tube
 * crosses its Rindler horizon, or when more than cap records would be written. */
int specrel_extrude_tubes_host(const specrel_tube *tubes, uint32_t n, specrel_geom *geoms, specrel_attr *attrs,
                               float *boxes, uint64_t cap, uint64_t *n_out, uint64_t *counts);

/* Deferred extrusion: the named scene's generator runs with every tube recorded instead of
 * extruded; specrel_builder_pending_tubes hands out their descriptors (valid while the builder
 * lives), and the records come back through specrel_builder_set_tube_records (for example from
 * specrel_extrude_tubes on the GPU) or specrel_builder_extrude_pending (host), in tube order.
 * specrel_builder_finish then builds the BVH as usual; the result equals specrel_scene_generate's. */
int specrel_scene_generate_deferred(const char *name, const specrel_gen_params *p, specrel_builder **out);
uint32_t specrel_builder_pending_tubes(const specrel_builder *b, specrel_tube *out, uint32_t cap);
int specrel_builder_set_tube_records(specrel_builder *b, const specrel_geom *geoms, const specrel_attr *attrs,
                                     const float *boxes, uint64_t n, const uint64_t *counts);
int specrel_builder_extrude_pending(specrel_builder *b);
/* Bvh4D::build (bvh4d.rs:122-427) on the host over n records' padded bounds (8 floats each: min
 * t,x,y,z, max t,x,y,z): nodes (capacity 2n - 1) in preorder, the leaf order (n) and the node count;
 * the CPU counterpart of specrel_build_bvh (include/specrel.h). */
int specrel_build_bvh_host(const float *boxes, uint64_t n, specrel_node *nodes, uint32_t *order, uint32_t *n_nodes);
/* specrel_builder_finish with a tree built elsewhere (specrel_build_bvh over the builder's records,
 * in their order): the records are put in the tree's leaf order, nothing is rebuilt. */
int specrel_builder_finish_with_tree(specrel_builder *b, const specrel_node *nodes, uint32_t n_nodes,
                                     const uint32_t *order, specrel_host_scene **out);

/* Tetrad operations (minkowski.rs:424-596); e is e[0..4] as (t,x,y,z). */
void specrel_tetrad_rest(float e[16]);
void specrel_tetrad_from_beta(float e[16], const float beta[3]);
void specrel_tetrad_boost_local(float e[16], const float d_rapidity[3]);
void specrel_tetrad_rotate_local(float e[16], float yaw, float pitch, float roll);
void specrel_tetrad_boost_to(float e[16], const float u[4]);

/* GpuUniforms::build equivalent (gpu.rs:147-212, with renderer.rs:334-355). */
int specrel_frame_build(const specrel_camera *cam, const specrel_render_config *cfg,
                        const specrel_dynamics *dyn, specrel_frame *out);
/* The reference's 14 common materials (scene.rs:181-275), packed. Returns the count. */
uint32_t specrel_common_materials(specrel_material *out, uint32_t cap);
/* Tables uploaded next to the scene: gamma LUT (renderer.rs:1104-1114) + glyph rows
 * (gpu.rs:223-229). out holds SPECREL_LUT_SIZE + 14*7 words. */
void specrel_build_tables(uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
