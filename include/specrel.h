/*
 * specrel.h — C ABI of the MI355X (gfx950) renderer: the drop-in for `GpuRenderer`
 * (src/gpu.rs:432-901). Plain pointers and sizes only; callable from C, C++, Python (ctypes)
 * or a Rust `extern "C"` block.
 *
 * Semantics follow the reference:
 *   - every call is synchronous unless its name says otherwise (render waits like
 *     gpu.rs:871-876);
 *   - a context is owned by one host thread at a time (GpuRenderer methods take &mut self);
 *     separate contexts may be used concurrently;
 *   - errors are return codes, never aborts: where the reference panics (framebuffer size
 *     mismatch, gpu.rs:808) the library returns SPECREL_EINVAL and sets the message;
 *   - scene arrays are copied to device memory at upload; the caller keeps ownership.
 * There is no CPU path inside this library: without a usable gfx950 device, specrel_create
 * fails (SPECREL_ENODEV) and the caller selects its own CPU renderer, as main.rs:275-296 does
 * when GpuRenderer::new returns Err.
 */
#ifndef SPECREL_H
#define SPECREL_H

#include "specrel_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SPECREL_OK 0
#define SPECREL_ENODEV (-1)  /* no HIP device / not gfx950 */
#define SPECREL_EINVAL (-2)  /* bad argument, size mismatch, malformed scene */
#define SPECREL_ENOMEM (-3)  /* device allocation failed */
#define SPECREL_EDEVICE (-4) /* a HIP runtime call failed */

typedef struct specrel_ctx specrel_ctx;

/* Where the last render's time went (GpuTimings, gpu.rs:408-420), in milliseconds. Kernel
 * times come from HIP events on the context's stream. */
typedef struct specrel_timings {
    float submit_ms;      /* host time to enqueue everything */
    float wait_ms;        /* host wait for the device */
    float kernel_ms;      /* first kernel start to last kernel end, device clock (SPECREL_OPT_STAGE_EVENTS 2) */
    float readback_ms;    /* device->host copies */
    float star_prep_ms;   /* per-view star pass + star binning */
    float trace_ms;       /* primary rays: closest hit per pixel */
    float shade_ms;       /* shading, shadow rays, haze, stars */
    float bloom_ms;       /* pyramid down + up passes */
    float composite_ms;   /* composite + tone map + pack */
    float adjoint_ms;     /* the last specrel_adjoint's kernels */
    uint32_t shadow_path; /* where the last render traced its shadow queries: SPECREL_SHADOWS_* */
    float extrude_ms;     /* the last specrel_extrude_tubes' kernels */
    uint32_t trace_block_waves; /* waves per block the last render's primary-ray kernel ran with */
    uint32_t trace_variant;     /* traversal the last render's primary-ray kernel ran: SPECREL_TRACE_* */
} specrel_timings;

/* Traversal variants of the primary-ray kernel (specrel_timings.trace_variant); every one visits
 * the reference's sequence of nodes and leaves, so results do not depend on it. */
enum {
    SPECREL_TRACE_PLAIN = 0,  /* one round of loads per step */
    SPECREL_TRACE_MERGED = 1, /* SPECREL_OPT_MERGED_LOADS */
    SPECREL_TRACE_WIDE4 = 2,  /* SPECREL_OPT_NODE_WIDTH 4 */
    SPECREL_TRACE_LDS_TOP = 3 /* SPECREL_OPT_LDS_TOP */
};

enum {
    SPECREL_SHADOWS_NONE = 0,     /* no shaded view with shadows on */
    SPECREL_SHADOWS_IN_SHADE = 1, /* by the shading kernel, per pixel */
    SPECREL_SHADOWS_QUEUED = 2,   /* queued by the shading kernel, traced by a dense any-hit pass */
    SPECREL_SHADOWS_IN_TRACE = 3  /* by the primary-ray kernel, right after the pixel's closest hit */
};

/* Options for specrel_set_option. No counterpart in the reference: launch choices of this backend. */
enum {
    SPECREL_OPT_STATS = 1,      /* 1: run the instrumented trace variant and fill specrel_stats */
    SPECREL_OPT_STAR_CULL = 2,  /* 1 (default): per-tile star culling; 0: every star per pixel */
    SPECREL_OPT_STACK = 3,      /* traversal stack: 8, 10, 12 or 16 (default) positions per lane in LDS, or
                                   -4, -8, -12, -16: that many top entries in registers */
    SPECREL_OPT_PERSISTENT = 4, /* 1: persistent primary-ray kernel with per-lane refill; 0 (default): one tile per block */
    SPECREL_OPT_SHADOW_QUEUE = 5, /* 1: queue shadow rays and trace them densely; 0 (default): trace them in the shading lane */
    SPECREL_OPT_GRAPH = 6,        /* replay a captured HIP graph of the launches: 0 never, 1 (default) for
                                     batches, 2 for every render. Graph renders report kernel_ms only. */
    SPECREL_OPT_MERGED_LOADS = 7, /* 1: a step of the primary-ray kernel (LDS stacks) loads the node
                                     pairs of a wave's lanes at interior nodes and the first three
                                     primitives of its lanes at leaves in one round, waited for once;
                                     0 (default): each kind waits for its own loads. Same visiting order. */
    SPECREL_OPT_SHADOW_IN_TRACE = 8, /* 1 (default): a hit's shadow queries are traced by the primary-ray
                                       kernel right after its ray (LDS stacks, tiled rays, no queue);
                                       0: by the shading pass */
    SPECREL_OPT_TRACE_WAVES = 9, /* 1 (default), 2 or 4: waves per block of the tiled primary-ray kernel
                                    with an LDS stack, each wave one quarter of a 16x16 tile; the
                                    persistent kernel and register stacks run blocks of 4
                                    (specrel_timings.trace_block_waves reports what ran) */
    SPECREL_OPT_STAR_OVERLAP = 10, /* 1: the star pass runs on a second stream beside the primary
                                     rays; 0 (default): after them on the same stream */
    SPECREL_OPT_WAVE_SHAPE = 11,  /* pixels of a 16x16 tile one wave of the tiled primary-ray kernel
                                     takes: 0 (default) an 8x8 quarter, 1 a 16x4 band, 2 a 4x16 band */
    SPECREL_OPT_TILE_ORDER = 12,  /* 1 (default): when a render has more tiles than four per CU, the
                                     tiled primary-ray kernel starts the tiles that took longest in the
                                     previous render of the same view sizes first (row order after a
                                     change of sizes); 0: row order. Results do not depend on it. */
    SPECREL_OPT_LEAF_POSTPONE = 13, /* 1 (default): a closest-hit ray of the primary-ray kernel (LDS
                                     stacks, not merged loads) that reaches a leaf keeps it pending
                                     and traverses on to its next leaf; the pending leaves are tested
                                     in order once 3/8 of the wave's tracing lanes hold one; a ray
                                     whose stack would pass 62 entries is traced again testing every
                                     leaf when reached. 2..62: the same with that many entries as the
                                     limit (to exercise the second trace). 0: every leaf is tested
                                     when reached. Results do not depend on it; instrumented renders
                                     (SPECREL_OPT_STATS) test every leaf when reached, so their
                                     counters are the reference's. */
    SPECREL_OPT_ADJOINT_MASK = 14, /* which pixels specrel_adjoint[_batch] counts as valid: 0 (default)
                                     every hit inside its tetrahedron whose neighbours are on the same
                                     object (silhouettes masked); 1 also no facet, texture or shadow
                                     edge among the neighbours (the set finite differences resolve) */
    SPECREL_OPT_WAVE_CLOCK = 15,  /* 1: each wave of the tiled primary-ray kernel records when it
                                     started and ended (the device's constant-rate clock) and where it
                                     ran; read with specrel_last_wave_clock. Renders do not replay
                                     graphs meanwhile. Results never depend on it */
    SPECREL_OPT_NODE_WIDTH = 16,  /* BVH node fetch width of the postponed-leaf traversal (tiled primary
                                     rays, LDS stack of 8 or 16): 2 (default) one child box pair per
                                     step; 4: a step also fetches both children's box pairs, so a
                                     continued child's boxes are in registers at its step. Same
                                     visiting order. */
    SPECREL_OPT_LDS_TOP = 17,     /* 0 (default), or 1..8: each block of the tiled primary-ray kernel
                                     (LDS stack of 8 or 16) copies the interior nodes of that many top
                                     levels of the tree into LDS and reads them there (takes
                                     precedence over SPECREL_OPT_NODE_WIDTH). Same visiting order. */
    SPECREL_OPT_FUSED_SHADE = 18, /* tiled primary rays with an LDS stack of 8 or 16 and their shadow
                                     queries traced there shade their own pixels (the star pass runs
                                     first): 1 (default) for renders of one view, 2 for every render,
                                     0 never (a shading pass of its own after them). Instrumented
                                     renders, queued or overlapped passes use the separate pass.
                                     Results do not depend on it. */
    SPECREL_OPT_STAGE_EVENTS = 19 /* device events of a directly launched render (specrel_timings):
                                     2 (default) at every stage boundary (every *_ms field), 1 only
                                     around the primary-ray kernel (trace_ms; the other device
                                     fields read 0), 0 none. Each event idles the device for a few
                                     microseconds between the kernels it separates. */
};

/* Per-view outputs of specrel_render_batch; every pointer may be NULL (not copied). len is the
 * element count of each non-NULL buffer and must be the view's width*height (else
 * SPECREL_EINVAL before any work, where the reference asserts, gpu.rs:808). */
typedef struct specrel_view_out {
    uint32_t *rgb;
    uint32_t *hit;
    float *lambda;
    float *tau;
    uint64_t len;
} specrel_view_out;

/* Launch configuration chosen by specrel_autotune (it is also applied to the context). */
typedef struct specrel_tuning {
    int32_t stack;        /* SPECREL_OPT_STACK */
    int32_t persistent;   /* SPECREL_OPT_PERSISTENT */
    int32_t shadow_queue; /* SPECREL_OPT_SHADOW_QUEUE */
    int32_t graph;        /* SPECREL_OPT_GRAPH */
    uint32_t candidates;  /* configurations timed */
    float best_ms;        /* median wall time per render of the batch with the choice */
    float default_ms;     /* the same with the options the context had before */
    int32_t shadow_in_trace; /* SPECREL_OPT_SHADOW_IN_TRACE */
    int32_t wave_shape;      /* SPECREL_OPT_WAVE_SHAPE */
    int32_t leaf_postpone;   /* SPECREL_OPT_LEAF_POSTPONE */
    int32_t trace_waves;     /* SPECREL_OPT_TRACE_WAVES */
    int32_t node_width;      /* SPECREL_OPT_NODE_WIDTH */
    int32_t lds_top;         /* SPECREL_OPT_LDS_TOP */
    int32_t fused_shade;     /* SPECREL_OPT_FUSED_SHADE */
} specrel_tuning;

/* GpuRenderer::new (gpu.rs:472-620) without the scene upload. */
int specrel_create(int device_ordinal, specrel_ctx **out);
void specrel_destroy(specrel_ctx *ctx);
/* Message for the last failing call on ctx (or of specrel_create when ctx is NULL). */
const char *specrel_last_error(const specrel_ctx *ctx);
/* GpuRenderer::adapter_name (gpu.rs:734-736). */
const char *specrel_device_name(const specrel_ctx *ctx);
int specrel_set_option(specrel_ctx *ctx, int option, int value);

/* GpuRenderer::upload_scene (gpu.rs:622-699): copy (and re-pack) one scene into slot
 * scene_idx, replacing what was there. The arrays are PackedSceneGpuData's plus `order`. */
int specrel_upload_scene(specrel_ctx *ctx, uint32_t scene_idx, const specrel_scene_arrays *scene);
/* The same from a scene file (include/specrel_file.h) written by any producer, in place of the
 * arrays PackedSceneGpuData::from_scene (gpu.rs:247-289) hands GpuRenderer::new in process (and
 * web.rs:431-479 hands WebGPU). The file is mapped read-only and copied to the device; verify != 0
 * checks the section checksums first. SPECREL_EINVAL names what is wrong with a bad file. */
int specrel_upload_scene_file(specrel_ctx *ctx, uint32_t scene_idx, const char *path, int verify);
/* extend_worldtube_simplices (mesh4d.rs:721-854) for n tubes on the device: one thread per prism
 * (a triangle swept over one worldline step), its three tetrahedra built by Simplex4D::new
 * (mesh4d.rs:312-362) and packed as GpuGeom / GpuAttr (gpu.rs:68-92, 358-381), degenerate ones
 * dropped; the records of tube i follow those of tube i-1, bit-identical to the scene library's
 * specrel_extrude_tubes_host. *n_out receives the record count and counts[i] (may be NULL) the
 * records of tube i. geoms / attrs / boxes (padded bounds: min t,x,y,z, max t,x,y,z; any may be
 * NULL) are host arrays of cap records; with all three NULL the records stay on the device and only
 * the counts come back. SPECREL_EINVAL for a tube crossing its Rindler horizon, a bad index or too
 * small a cap. specrel_timings.extrude_ms holds the kernels' time. */
int specrel_extrude_tubes(specrel_ctx *ctx, const specrel_tube *tubes, uint32_t n_tubes, specrel_geom *geoms,
                          specrel_attr *attrs, float *boxes, uint64_t cap, uint64_t *n_out, uint64_t *counts);

/* Bvh4D::build (bvh4d.rs:122-427) on the device, emitting the same tree as the scene library's
 * host build (specrel_builder_finish) bit for bit: nodes (capacity 2n - 1) in preorder and the leaf
 * order (n), for n records whose padded bounds are boxes (8 floats per record: min t,x,y,z, max
 * t,x,y,z; host memory), or, with boxes NULL, the records of the last specrel_extrude_tubes still
 * on the device. *n_nodes receives the node count; stats (may be NULL) what the build did. */
typedef struct specrel_bvh_stats {
    float ms;                     /* wall time of the build, device work and host decisions */
    uint32_t levels;              /* breadth-first levels of large segments */
    uint32_t large_segments;      /* segments partitioned level-parallel */
    uint32_t small_subtrees;      /* subtrees of at most 256 records, one per thread */
    uint32_t fallback_segments;   /* partitions that fell back to the centroid sort */
    uint32_t jump_passes;         /* pointer-jumping passes of the partitions */
    uint32_t reserved[2];
} specrel_bvh_stats;
int specrel_build_bvh(specrel_ctx *ctx, const float *boxes, uint64_t n, specrel_node *nodes, uint32_t *order,
                      uint32_t *n_nodes, specrel_bvh_stats *stats);

/* A range [first, first + count) of node or primitive (leaf-order) indices. */
typedef struct specrel_range {
    uint32_t first, count;
} specrel_range;

/* GpuRenderer::update_scene (gpu.rs:704-731) with the dirty ranges of PackedSceneGpuData::update
 * (gpu.rs:297-344): `scene` is the slot's scene after an in-place update (same node, primitive,
 * material and star counts). The bounds of the nodes in node_ranges and the records, attributes
 * and order entries of the primitives in prim_ranges are copied to the device; materials, stars
 * and the light are copied whole. When a listed node's children or leaf range differ from the
 * uploaded tree, or a count differs, everything is uploaded again instead, as the reference does
 * when the ranges do not apply; *out_full (may be NULL) tells which happened. */
int specrel_patch_scene(specrel_ctx *ctx, uint32_t scene_idx, const specrel_scene_arrays *scene,
                        const specrel_range *node_ranges, uint32_t n_node_ranges, const specrel_range *prim_ranges,
                        uint32_t n_prim_ranges, int *out_full);
/* BvhVersion (bvh4d.rs:47-61): which layout of a BVH's arrays a copy was made from (numbered from 1;
 * 0 means none) and how many in-place edits that layout has had since. */
typedef struct specrel_bvh_version {
    uint64_t layout;
    uint64_t edit;
} specrel_bvh_version;

/* GpuRenderer::upload_scene (gpu.rs:622-699) with the version of the BVH the arrays were packed
 * from. As the reference does for a scene edited in place (edit > 0, gpu.rs:639-656), the node,
 * primitive and star arrays get room to grow by half, so later edits can be patched in. */
int specrel_upload_scene_versioned(specrel_ctx *ctx, uint32_t scene_idx, const specrel_scene_arrays *scene,
                                   const specrel_bvh_version *version);

/* GpuRenderer::update_scene (gpu.rs:701-731): `scene` is the slot's scene after the BVH edit that
 * took it to `version` (Bvh4D::begin_edit .. write_part / write_nodes, bvh4d.rs:181-215), and
 * node_ranges / prim_ranges are the ranges that edit wrote (Bvh4D::dirty). When the slot holds
 * the version just before (same layout, edit - 1: `follows`, gpu.rs:341-343) and the grown arrays
 * fit its buffers, only what changed is written: the device rows (child boxes and entries) of the
 * listed interior nodes and of their parents, and of reached nodes that the previous tree did not
 * reach or whose links differ from the device's; the primitives' records, attributes and order
 * entries of prim_ranges; the stars, materials and light whole. The new tree is validated first.
 * Otherwise everything is uploaded afresh; *out_patched (may be NULL) tells which happened. As in
 * the reference, a node whose box changed must be listed. A scene file
 * (specrel_upload_scene_file) carries its version too. */
int specrel_update_scene(specrel_ctx *ctx, uint32_t scene_idx, const specrel_scene_arrays *scene,
                         const specrel_bvh_version *version, const specrel_range *node_ranges, uint32_t n_node_ranges,
                         const specrel_range *prim_ranges, uint32_t n_prim_ranges, int *out_patched);

/* Number of scene slots in use (highest uploaded index + 1). */
uint32_t specrel_num_scenes(const specrel_ctx *ctx);

/* GpuRenderer::render (gpu.rs:798-900): render one view of scenes[scene_idx] and wait.
 * out_rgb receives width*height 0x00RRGGBB pixels (required). The optional per-pixel outputs
 * (NULL to skip) are the hit tetrahedron's input-order index (SPECREL_NO_HIT on a miss), the
 * hit's ray parameter lambda (+inf on a miss) and its interpolated proper time (0 on a miss).
 * len is the element count of each output buffer: it must equal width*height of the frame, or
 * the call returns SPECREL_EINVAL before any work (the reference's
 * `assert_eq!(framebuffer.len(), w * h)`, gpu.rs:808). The outputs are copied behind the render
 * on its stream into pinned memory and from there into the caller's buffers. */
int specrel_render(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frame, uint32_t *out_rgb,
                   uint32_t *out_hit, float *out_lambda, float *out_tau, uint64_t len);

/* The same render, leaving every output in device memory (no readback) and returning once
 * the work is enqueued; with want_hits the hit/lambda/tau buffers are written too.
 * specrel_sync waits; specrel_download copies the resident outputs of the last render. The
 * frame is copied before the call returns (to pinned staging memory, or into the launch's
 * arguments), so the caller may change it at once and enqueue further renders; each render's
 * outputs replace the previous one's. */
/* GpuRenderer::render (gpu.rs:798-900) fed the reference's own per-view block: `u` is the
 * GpuUniforms that GpuUniforms::build made (gpu.rs:146-212), passed unchanged, so a Rust caller
 * hands over what it already builds. Its scene fields (num_nodes, num_stars, light_pos) must be
 * the uploaded scene's, else SPECREL_EINVAL; `len` as for specrel_render. A flash arrives as
 * colour x intensity and is rendered as that colour at intensity 1, which multiplies it as the
 * shader does (render.wgsl:820); the CPU renderer scales the colour by intensity times the
 * falloff instead (renderer.rs:685), so with flashes the flash term is rounded as the shader's,
 * not bit for bit as the CPU renderer's. Without flashes it is specrel_render's image. */
int specrel_render_uniforms(specrel_ctx *ctx, uint32_t scene_idx, const specrel_gpu_uniforms *u, uint32_t *out_rgb,
                            uint64_t len);

int specrel_render_resident(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frame, int want_hits);

/* No counterpart in the reference. For callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc -- the same one-view
 * render as specrel_render_resident, writing the caller's device buffers (d_rgb: width*height
 * 0x00RRGGBB pixels, required; d_hit / d_lambda / d_tau: all three or none, as specrel_render's
 * optional outputs; len: their element count, which must be width*height) and allocating nothing:
 * every intermediate buffer lives in a fixed scratch area of SPECREL_NOALLOC_SCRATCH_BYTES that the
 * library holds on each GPU until the process ends (a 1280x800 view of a scene with 10,000 stars
 * needs about 35 MB). A view needing more returns SPECREL_EINVAL, as do SPECREL_OPT_STATS and
 * SPECREL_OPT_SHADOW_QUEUE (they need buffers of their own). Tiles are traced in row order.
 * Returns once enqueued on the context's stream; specrel_sync waits. Every call on one GPU, from
 * any context, uses that GPU's one scratch area: calls that overlap on a GPU race. The downloads
 * and the adjoint refuse to run after it (its outputs are the caller's). Images are bit-identical
 * to specrel_render's. */
#define SPECREL_NOALLOC_SCRATCH_BYTES (40u << 20)
int specrel_render_noalloc(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frame, uint32_t *d_rgb,
                           uint32_t *d_hit, float *d_lambda, float *d_tau, uint64_t len);
int specrel_sync(specrel_ctx *ctx);
/* len: element count of each non-NULL buffer, which must be the view's width*height. */
int specrel_download(specrel_ctx *ctx, uint32_t *out_rgb, uint32_t *out_hit, float *out_lambda, float *out_tau,
                     uint64_t len);

/* The linear HDR radiance of view `view` of the last render before bloom and tone mapping (the
 * image the adjoint's loss is defined on): width*height*3 floats, RGB per pixel; len (floats)
 * must be width*height*3. */
int specrel_download_linear(specrel_ctx *ctx, uint32_t view, float *out_rgb3, uint64_t len);

/* Batched observers: n_views (1..4096) frames of any sizes and modes rendered over one scene
 * with one launch per stage (each kernel's grid spans the tiles of every view), replayed from a
 * captured graph (SPECREL_OPT_GRAPH). Every view's outputs are bit-identical to specrel_render of
 * its frame alone. No counterpart in the reference, whose callers loop over
 * GpuRenderer::render (gpu.rs:798); `outs` (n_views entries) may be NULL to leave everything
 * resident. Synchronous, like specrel_render. */
int specrel_render_batch(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frames, uint32_t n_views,
                         const specrel_view_out *outs);
/* The same, returning once enqueued; specrel_download_view copies one view afterwards (view 0
 * of a single render is what specrel_download copies). */
int specrel_render_batch_resident(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frames,
                                  uint32_t n_views, int want_hits);
int specrel_download_view(specrel_ctx *ctx, uint32_t view, uint32_t *out_rgb, uint32_t *out_hit, float *out_lambda,
                          float *out_tau, uint64_t len);
/* Times every launch configuration (stack depth 16/8 in LDS, persistent or tiled primary rays,
 * shadow rays in the shading pass, queued or in the primary-ray kernel, graph replay or direct
 * launches, the wave shape of tiled primary rays), then the fastest with leaf postponement
 * switched, then the fastest with each other block size of tiled primary rays, on this batch,
 * `iters` renders each after one warm-up, and keeps the fastest by median wall time. */
int specrel_autotune(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frames, uint32_t n_views,
                     uint32_t iters, specrel_tuning *out);

/* Differentiation of the last render (one view, RelativisticShaded, PastLightCone, null or no
 * shadows, no flashes or haze). No counterpart in the reference, which has no gradients.
 *
 * Loss: L = sum over valid pixels of |I - target|^2, with I the linear HDR radiance before bloom
 * (3 floats per pixel; target likewise, width*height*3 floats). Gradients of L with respect to
 * the observer event, the observer tetrad legs e_a (as passed in the frame) and, per object, the
 * rapidity vector of a Lorentz boost of the object's whole worldtube about `center`: its
 * primitives' records, normals and four-velocities are transformed by that boost, which at zero
 * is the derivative with respect to the object's velocity there.
 *
 * Each pixel is recomputed from the render's per-pixel hit (the checkpoint): intersection,
 * interpolation, shading and its shadow ray, then differentiated by hand. Valid pixels hit inside
 * their tetrahedron (every barycentric > 0), carry no starlight and have 8 neighbours inside the
 * image on the same object, none of them a miss or starlight: their shading is differentiable in
 * every parameter, and the gradient is its derivative. The others are counted per class and left
 * out: misses, starlight, the image border, silhouettes (a neighbour off the object) and edges (a
 * hit on its tetrahedron's boundary). With SPECREL_OPT_ADJOINT_MASK = 1, a neighbour with another
 * albedo (without the wireframe blend), shadow state or facet (normal and four-velocity, constant
 * per tetrahedron) makes an edge too: the set a finite step can cross nothing in. mask_in (width*height bytes, nonzero =
 * valid; may be NULL) replaces that test, so a loss can be evaluated on a fixed pixel set;
 * mask_out (may be NULL) receives the mask used. n_pixels is the pixel count the target and the
 * masks hold and must equal the last render's (SPECREL_EINVAL otherwise). d_objects receives 3
 * doubles per object. */
typedef struct specrel_adj_object {
    uint32_t prim_first, prim_count; /* input-order primitive range; objects sorted and disjoint */
    float center[4];                 /* event the boost is about (t, x, y, z) */
} specrel_adj_object;

typedef struct specrel_adjoint_out {
    double loss;
    double d_event[4];     /* dL / d observer event (t, x, y, z) */
    double d_tetrad[4][4]; /* dL / d e_a (t, x, y, z) */
    uint64_t valid, miss, starlight, border, silhouette, edge; /* pixels per class */
} specrel_adjoint_out;

int specrel_adjoint(specrel_ctx *ctx, const float *target, const uint8_t *mask_in, uint8_t *mask_out,
                    uint64_t n_pixels, const specrel_adj_object *objects, uint32_t n_objects,
                    specrel_adjoint_out *out, double *d_objects);
/* The same for every view of the last render (a batch of uneven views, each with its own
 * observer): one launch per pass over all views' tiles, one gradient per view. targets, masks_in
 * and masks_out hold the views' pixels one after another (view 0 first, width*height each, 3
 * floats per pixel for targets), n_pixels of them in all; outs receives n_views results and
 * d_objects n_views * 3 * n_objects doubles (view-major). */
int specrel_adjoint_batch(specrel_ctx *ctx, const float *targets, const uint8_t *masks_in, uint8_t *masks_out,
                          uint64_t n_pixels, const specrel_adj_object *objects, uint32_t n_objects,
                          specrel_adjoint_out *outs, double *d_objects);

int specrel_last_timings(const specrel_ctx *ctx, specrel_timings *out);
/* Counters of the last render made with SPECREL_OPT_STATS on. */
int specrel_last_stats(const specrel_ctx *ctx, specrel_stats *out);

/* No counterpart in the reference: a profiling aid. The records of the last render's waves, when
 * it ran with SPECREL_OPT_WAVE_CLOCK and the tiled primary-ray kernel: copies min(cap, waves) of
 * them to out (may be NULL to ask the count), *n_waves = waves (0 when none were recorded),
 * *clock_khz = the clock's rate. Waits for the render. */
int specrel_last_wave_clock(specrel_ctx *ctx, specrel_wave_clock *out, uint64_t cap, uint64_t *n_waves,
                            uint32_t *clock_khz);

#ifdef __cplusplus
}
static_assert(sizeof(specrel_timings) == 56 && offsetof(specrel_timings, trace_variant) == 52, "timings layout");
static_assert(sizeof(specrel_tuning) == 56 && offsetof(specrel_tuning, fused_shade) == 52, "tuning layout");
#endif
#endif
