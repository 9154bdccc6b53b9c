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
} oracle_counts;

/* Full frame: HDR shading, bloom, tone map (render_frame, renderer.rs:327-412).
 * Optional outputs may be NULL. hdr_out receives the pre-bloom HDR image (3 floats/pixel). */
int oracle_render(const specrel_scene_arrays *s, const specrel_frame *f, int nthreads, uint32_t *rgb,
                  uint32_t *hit_index, float *lambda, float *tau, float *hdr_out, oracle_counts *counts);
/* Shade only rows row0, row0+step, ... (nrows of them) into hdr_rows (3 floats/pixel);
 * the CPU-baseline sample. */
int oracle_render_rows(const specrel_scene_arrays *s, const specrel_frame *f, int nthreads, uint32_t row0,
                       uint32_t step, uint32_t nrows, float *hdr_rows, oracle_counts *counts);
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
