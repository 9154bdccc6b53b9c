/*
 * oracle.h — CPU restatement of the reference renderer, used ONLY as the test checker and as
 * the CPU baseline in bench.py. Nothing in the product links or calls it.
 *
 * It consumes the same packed scene arrays and per-view frame block as the HIP library
 * (include/specrel_types.h) and follows src/renderer.rs, src/bvh4d.rs, src/mesh4d.rs,
 * src/star.rs and src/bloom.rs operation by operation (see oracle.c for per-function
 * citations). Compile with -ffp-contract=off and without fast-math.
 */
#ifndef SPECREL_ORACLE_H
#define SPECREL_ORACLE_H

#include <stdint.h>
#include "../include/specrel_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_counts {
    uint64_t rays, node_pops, box_tests, prim_tests, hits;
    uint64_t shadow_rays, shadow_node_pops, shadow_box_tests, shadow_prim_tests;
    uint64_t star_pairs; /* star x pixel pairs evaluated (all stars, every pixel) */
    /* oracle_render only: distinct interior nodes whose children were tested, distinct primitive
     * records tested and distinct hit primitives shaded (each counted once per frame) */
    uint64_t distinct_nodes, distinct_prims, distinct_hits;
} oracle_counts;

/* Full frame: HDR shading, bloom, tone map (render_frame, renderer.rs:327-412).
 * Optional outputs may be NULL. hdr_out receives the pre-bloom HDR image (3 floats/pixel). */
int oracle_render(const specrel_scene_arrays *s, const specrel_frame *f, int nthreads, uint32_t *rgb,
                  uint32_t *hit_index, float *lambda, float *tau, float *hdr_out, oracle_counts *counts);
/* Shade only rows row0, row0+step, ... (nrows of them) into hdr_rows (3 floats/pixel, the
 * pre-bloom image) and, when not NULL, their hit index / lambda / tau (w per row); the
 * CPU-baseline sample. */
int oracle_render_rows(const specrel_scene_arrays *s, const specrel_frame *f, int nthreads, uint32_t row0,
                       uint32_t step, uint32_t nrows, float *hdr_rows, uint32_t *hit_rows, float *lambda_rows,
                       float *tau_rows, oracle_counts *counts);
void oracle_bloom(float *hdr3, uint32_t w, uint32_t h);
void oracle_pack(const float *hdr3, uint32_t n, uint32_t *out);
void oracle_gamma_lut(uint8_t out[SPECREL_LUT_SIZE]);

int oracle_intersect(const specrel_geom *g, const float o[4], const float d[4], float tmin, float tmax,
                     float out_lambda_event_bary[9]);
/* Closest hit through the BVH (bvh4d.rs:445-520). Returns the leaf slot or -1. */
int64_t oracle_trace(const specrel_scene_arrays *s, const float o[4], const float d[4], float tmin, float tmax,
                     float out_lambda_event_bary[9], oracle_counts *counts);
/* Linear scan over every leaf slot in order: the smallest lambda (first slot on ties). */
int64_t oracle_trace_brute(const specrel_scene_arrays *s, const float o[4], const float d[4], float tmin,
                           float tmax, float out_lambda_event_bary[9]);
int oracle_occluded(const specrel_scene_arrays *s, const float o[4], const float d[4], float tmin, float tmax);
int oracle_occluded_brute(const specrel_scene_arrays *s, const float o[4], const float d[4], float tmin,
                          float tmax);

void oracle_star_light(const float *star_words, uint32_t num_stars, const float origin[4], const float dir[4],
                       const float e0[4], float cone, float width, float spread, float lambda_max, float out[3]);
int oracle_cone_event(const float *star_words, uint32_t num_stars, uint32_t star, const float origin[4], const float e0[4], float cone,
                      float d_out[4], float u_out[4]);
void oracle_blackbody(float temperature, float out[3]);
float oracle_color_norm(float temperature, float intensity);
int oracle_digital_tau(float u, float v, float proper_time);
float oracle_decision_age(const float event[4], const float origin[4], const float u[4]);
void oracle_shade_sky(const float dir[4], float out[3]);
void oracle_haze(const float base[3], const float event[4], const float *origins, uint32_t n, float out[3]);
/* The forward state of every pixel that the FP64 adjoint reference (adjoint_ref.cpp) takes as
 * given: the hit's leaf slot (-1 on a miss), the albedo without the wireframe blend, the checker
 * cell colour and grid-line factor, the shadow decision, the shaded colour before haze and stars,
 * and the pixel's unit local direction. Shaded past-light-cone views only (else -1). */
typedef struct oracle_pixel_state {
    int64_t slot;
    float nw_albedo[3];
    float col0[3];
    float dark;
    int32_t pattern;
    int32_t in_shadow;
    float color[3];
    float local_dir[3];
    uint32_t material;
    uint32_t flags; /* ORACLE_PX_*: the f32 forward's branch at each kink of the shading */
    int32_t mtb_k;  /* which triangle barycentric is the smallest (0..2) */
} oracle_pixel_state;
enum {
    ORACLE_PX_WIRE = 1,       /* inside the wireframe band, blend not clamped */
    ORACLE_PX_CL = 2,         /* light cosine > 0 */
    ORACLE_PX_CV = 4,         /* view cosine > 0 */
    ORACLE_PX_RL_FREE = 8,    /* light rest-frame cosine neither clamped nor degenerate */
    ORACLE_PX_RV_FREE = 16,   /* view rest-frame cosine neither clamped nor degenerate */
    ORACLE_PX_HZ_FREE = 32,   /* sky horizon blend not clamped */
    ORACLE_PX_DIST_FREE = 64, /* light distance above its 1e-4 floor */
    ORACLE_PX_INSIDE = 128    /* the hit lies inside its tetrahedron (every barycentric > 0) */
};
int oracle_pixel_states(const specrel_scene_arrays *s, const specrel_frame *f, int nthreads, oracle_pixel_state *out);

/* Test entries: shading of a hand-built hit record and its intermediate values. */
typedef struct oracle_record {
    float lambda;
    float event[4];
    float tet_bary[4];
    float tri_bary[3];
    float uv[2];
    float proper_time;
    float tau_uv_grad[2];
    float uv_world_scale[2];
    float normal[4];
    float four_velocity[4];
    float decided_at;
    uint32_t material;
} oracle_record;
typedef struct oracle_shade_parts {
    float albedo[3];   /* sample_surface_albedo */
    float lighting;    /* compute_illumination */
    float cos_light, cos_view, t_emit, dist;
    int32_t light_shadowed;
    float flash[3];    /* compute_flash_illumination */
    float lit[3];      /* before the distance fog */
} oracle_shade_parts;
int oracle_shade_record(const specrel_scene_arrays *s, const specrel_frame *f, const float dir[4],
                        const oracle_record *r, float out[3], oracle_shade_parts *parts);
void oracle_pixel_ray(const specrel_frame *f, uint32_t px, uint32_t py, float origin[4], float dir[4]);
void oracle_sky_gradient(const float dir[4], float out[3]);
void oracle_distance_fog(const float color[3], const float dir[4], float lambda, float out[3]);
void oracle_math_vec(int fn, const float *a, const float *b, float *out, uint64_t n);

/* f32 helpers as the oracle defines them (checked against high-precision values). */
float oracle_expf(float x);
float oracle_sinf(float x);
float oracle_cosf(float x);
float oracle_atan2f(float y, float x);
float oracle_acosf(float x);
float oracle_rem_euclid(float x, float m);

#ifdef __cplusplus
}
#endif
#endif
