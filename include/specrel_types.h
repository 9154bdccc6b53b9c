/*
 * specrel_types.h — byte layouts shared across the drop-in boundary.
 *
 * The scene records are the reference's packed GPU records, byte for byte, so a
 * Rust caller can hand over what `PackedSceneGpuData::from_scene` already builds:
 *   specrel_node     <- GpuNode     (src/gpu.rs:42-64)     48 B
 *   specrel_geom     <- GpuGeom     (src/gpu.rs:66-74)     80 B, leaf order
 *   specrel_attr     <- GpuAttr     (src/gpu.rs:76-92, packed at 358-381) 192 B, leaf order
 *   specrel_material <- GpuMaterial (src/gpu.rs:94-100, packed at 255-275) 48 B
 *   star words       <- pack_stars  (src/gpu.rs:102-119)   float4 headers + 3 float4 per segment
 * plus the leaf-slot -> input-order map `order` (src/bvh4d.rs:72-76), which the reference
 * keeps on the host; the library uses it to report the hit tetrahedron's input index.
 *
 * specrel_frame replaces `GpuUniforms` (src/gpu.rs:121-212). It carries the same fields, but
 * keeps each flash's colour and intensity apart: the CPU renderer multiplies them in a
 * different order (src/renderer.rs:685) from the WGSL port's premultiplied colour
 * (gpu.rs:164), and the CPU order is the one reproduced bit for bit.
 *
 * All vectors are (t, x, y, z) for 4-vectors and (x, y, z, _) for colours.
 */
#ifndef SPECREL_TYPES_H
#define SPECREL_TYPES_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPECREL_MAX_FLASHES 16 /* renderer.rs:179 */
#define SPECREL_MAX_HAZE 5     /* renderer.rs:224 */
#define SPECREL_BLOOM_LEVELS 6 /* bloom.rs:23 */
#define SPECREL_LUT_SIZE 4096  /* renderer.rs:1102 */
#define SPECREL_NUM_GLYPHS 14  /* gpu.rs:33-35: '0'-'9', '+', '-', '.', 'S' */
#define SPECREL_NO_HIT 0xFFFFFFFFu

/* RenderMode (renderer.rs:11), ShadowMode (128), ViewMode/SliceMode (69), SurfacePattern
 * (mesh4d.rs:9); numbering as GpuUniforms::build (gpu.rs:181-197). */
enum { SPECREL_MODE_SHADED = 0, SPECREL_MODE_ISOCHRONES = 1, SPECREL_MODE_INSPECTOR = 2,
       SPECREL_MODE_DECISION_AGE = 3 };
enum { SPECREL_SHADOW_OFF = 0, SPECREL_SHADOW_NULL = 1, SPECREL_SHADOW_CAMERA_SLICE = 2 };
enum { SPECREL_VIEW_PAST_CONE = 0, SPECREL_VIEW_SIMULTANEITY = 1, SPECREL_VIEW_FRONTIER = 2,
       SPECREL_VIEW_PROPHECY = 3 };
enum { SPECREL_PATTERN_CHECKER = 0, SPECREL_PATTERN_CLOCK = 1, SPECREL_PATTERN_LIGHT_CLOCK = 2 };

typedef struct specrel_node {      /* GpuNode, gpu.rs:44-51 */
    float min[4];
    float max[4];
    uint32_t right_or_first;       /* leaf: first leaf slot; interior: right child */
    uint32_t count;                /* > 0 for leaves */
    uint32_t left;                 /* interior: left child */
    uint32_t pad;
} specrel_node;

typedef struct specrel_geom {      /* GpuGeom, gpu.rs:68-74 (SimplexGeom4D, mesh4d.rs:153-159) */
    float v0[4];
    float n[4];
    float b1[4];
    float b2[4];
    float b3[4];
} specrel_geom;

typedef struct specrel_attr {      /* GpuAttr, gpu.rs:79-92: one row per component, 4 vertices */
    float uv_u[4];
    float uv_v[4];
    float tb_x[4];
    float tb_y[4];
    float tb_z[4];
    float tau[4];
    float grad_u[4];
    float grad_v[4];
    float normal[4];
    float four_vel[4];
    float misc[4];                 /* uv_world_scale.0, .1, bits(material_id), 0 */
    float decided_at[4];
} specrel_attr;

typedef struct specrel_material {  /* GpuMaterial, gpu.rs:96-100 */
    float base[4];                 /* base rgb, emission */
    float checker[4];              /* checker rgb, grid_scale */
    float params[4];               /* wireframe_width, bits(pattern), cavity_length, 0 */
} specrel_material;

/* The arrays `specrel_upload_scene` copies to the device (PackedSceneGpuData, gpu.rs:232-244). */
typedef struct specrel_wl_sample {  /* WorldlineSample (mesh4d.rs:615-630) */
    float center[4];
    float proper_time;
    float phi[3];
    float proper_accel_local[3];
    float pad;
    float e[4][4];
} specrel_wl_sample;

/* One worldtube to extrude: a triangle mesh in its rest frame swept along a sampled worldline
 * (extend_worldtube_simplices, mesh4d.rs:721-854; the mesh as Mesh3D, mesh4d.rs:437-451). */
typedef struct specrel_tube {
    const float *vertices;          /* 3 per vertex (x, y, z), rest frame */
    const uint32_t *tri_index;      /* 3 per triangle */
    const float *tri_uv;            /* 6 per triangle: (u, v) of each corner */
    const float *tri_normal;        /* 3 per triangle: outward rest-frame normal */
    const specrel_wl_sample *samples;
    const float *decided_at;        /* one per sample, or NULL: DECIDED_AT_STATIC (-1e9) */
    uint32_t n_vertices, n_triangles, n_samples, material;
} specrel_tube;

typedef struct specrel_scene_arrays {
    const specrel_node *nodes;
    uint32_t num_nodes;
    uint32_t num_prims;
    const specrel_geom *geoms;     /* num_prims, leaf order */
    const specrel_attr *attrs;     /* num_prims, leaf order */
    const uint32_t *order;         /* num_prims: leaf slot -> input simplex index */
    const specrel_material *materials;
    uint32_t num_materials;
    uint32_t num_stars;
    const float *star_words;       /* pack_stars: 4 * num_star_words floats */
    uint32_t num_star_words;       /* float4 count: num_stars headers + 3 per segment */
    uint32_t pad;
    float light_pos[4];            /* (0, x, y, z) as gpu.rs:286 */
} specrel_scene_arrays;

typedef struct specrel_frame {     /* per-view inputs; replaces GpuUniforms (gpu.rs:123-144) */
    float cam_pos[4];              /* Camera4D::position */
    float e[4][4];                 /* Camera4D::tetrad.e[0..4] */
    float half_wh[2];              /* aspect*tan(fov/2), tan(fov/2) (renderer.rs:336-338) */
    uint32_t size[2];              /* width, height */
    uint32_t mode, shadow_mode, slice_mode, num_flashes;
    float frontier_origin[4];      /* frontier_or_camera (renderer.rs:266) */
    float frontier_tau;            /* its proper time */
    float frontier_view_offset;    /* FRONTIER_RAY_OFFSET + frontier_lag (renderer.rs:235) */
    float prophecy_delta;
    uint32_t num_haze;
    float haze_origin[SPECREL_MAX_HAZE][4];      /* haze_origins (renderer.rs:253) */
    float flash_at[SPECREL_MAX_FLASHES][4];
    float flash_color_width[SPECREL_MAX_FLASHES][4]; /* colour rgb (unscaled), width */
    float flash_intensity[SPECREL_MAX_FLASHES];
} specrel_frame;

/* GpuUniforms (gpu.rs:120-144) byte for byte, 768 B: the reference's own per-view block as
 * GpuUniforms::build makes it (gpu.rs:146-212), for specrel_render_uniforms. */
typedef struct specrel_gpu_uniforms {
    float cam_pos[4];
    float e[4][4];
    float light_pos[4];       /* the scene's (PackedSceneGpuData::light_pos) */
    float half_wh[2];
    uint32_t size[2];
    uint32_t mode, shadow_mode, slice_mode;
    uint32_t num_nodes;       /* the scene's */
    float frontier_origin[4];
    float frontier_view_offset;
    uint32_t num_flashes;
    float frontier_tau;
    float prophecy_delta;
    uint32_t num_haze;
    uint32_t num_stars;       /* the scene's */
    uint32_t pad[2];
    float haze_origin[SPECREL_MAX_HAZE][4];
    float flash_at[SPECREL_MAX_FLASHES][4];
    float flash_col_width[SPECREL_MAX_FLASHES][4]; /* colour x intensity, width */
} specrel_gpu_uniforms;

/* One wave of a render's tiled primary-ray kernel (SPECREL_OPT_WAVE_CLOCK, specrel_last_wave_clock). */
typedef struct specrel_wave_clock {
    uint64_t start, end;  /* device clock ticks at the wave's first and last instruction */
    uint32_t tile;        /* the 16x16 tile (index over all views' tiles) */
    uint32_t quarter;     /* which quarter of it */
    uint32_t hw_id;       /* the wave's hardware location (HW_ID register: wave, SIMD, CU, shader array, engine) */
    uint32_t xcc;         /* the accelerator complex die it ran on */
} specrel_wave_clock;

/* Traversal counters of one render (the instrumented variant), summed over all rays. */
typedef struct specrel_stats {
    uint64_t rays;               /* primary rays */
    uint64_t node_pops;          /* closest-hit stack pops that reached a node */
    uint64_t box_tests;          /* child boxes tested, closest-hit */
    uint64_t prim_tests;         /* intersect_fast calls, closest-hit */
    uint64_t hits;               /* primary rays that hit */
    uint64_t shadow_rays;        /* any-hit queries (light + flashes) */
    uint64_t shadow_node_pops;
    uint64_t shadow_box_tests;
    uint64_t shadow_prim_tests;
    uint64_t star_tests;         /* per-pixel star kernel tests after tile culling */
    uint64_t distinct_nodes;     /* interior nodes whose child boxes some ray tested (each once) */
    uint64_t distinct_prims;     /* primitive records some ray tested (each once) */
    uint64_t distinct_hits;      /* primitives whose attributes some hit read (each once) */
    uint64_t reserved[3];
} specrel_stats;

#ifdef __cplusplus
}
#endif

#if defined(__cplusplus)
static_assert(sizeof(specrel_node) == 48, "GpuNode is 48 B");
static_assert(sizeof(specrel_geom) == 80, "GpuGeom is 80 B");
static_assert(sizeof(specrel_attr) == 192, "GpuAttr is 192 B");
static_assert(sizeof(specrel_material) == 48, "GpuMaterial is 48 B");
static_assert(offsetof(specrel_attr, misc) == 160, "GpuAttr.misc offset");
static_assert(offsetof(specrel_attr, decided_at) == 176, "GpuAttr.decided_at offset");
static_assert(offsetof(specrel_frame, haze_origin) == 144, "frame layout");
static_assert(sizeof(specrel_frame) == 800, "frame layout");
static_assert(sizeof(specrel_scene_arrays) == 88 && offsetof(specrel_scene_arrays, light_pos) == 72, "arrays");
static_assert(sizeof(specrel_stats) == 128, "stats layout");
static_assert(sizeof(specrel_wave_clock) == 32, "wave clock layout");
static_assert(sizeof(specrel_gpu_uniforms) == 768 && offsetof(specrel_gpu_uniforms, num_nodes) == 124 &&
                  offsetof(specrel_gpu_uniforms, haze_origin) == 176 &&
                  offsetof(specrel_gpu_uniforms, flash_at) == 256 &&
                  offsetof(specrel_gpu_uniforms, flash_col_width) == 512,
              "GpuUniforms layout");
#else
_Static_assert(sizeof(specrel_node) == 48, "GpuNode is 48 B");
_Static_assert(sizeof(specrel_geom) == 80, "GpuGeom is 80 B");
_Static_assert(sizeof(specrel_attr) == 192, "GpuAttr is 192 B");
_Static_assert(sizeof(specrel_material) == 48, "GpuMaterial is 48 B");
_Static_assert(sizeof(specrel_frame) == 800, "frame layout");
#endif

#endif
