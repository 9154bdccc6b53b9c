/*
 * specrel_file.h — versioned on-disk scene file: the packed GPU arrays of one scene
 * (PackedSceneGpuData, src/gpu.rs:232-289, the same byte layout the browser build hands to
 * WebGPU through the web_scene_*_ptr accessors, src/web.rs:431-479) plus the leaf order and the
 * BVH version, so a producer in another process or language (the Dynamic-Mode arena, the web
 * build, a Rust host) can feed the renderer without linking the scene library.
 *
 * Layout (little-endian, every offset from the start of the file):
 *   [0, 128)      specrel_file_header
 *   [128, ...)    num_sections x specrel_file_section (32 B each)
 *   sections      each starts on a 4096-byte boundary, elements packed, zero padding between
 *
 * Section kinds and element sizes:
 *   NODES 48 B (GpuNode), GEOMS 80 B (GpuGeom, leaf order), ATTRS 192 B (GpuAttr, leaf order),
 *   ORDER 4 B (leaf slot -> input simplex index, Bvh4D::order, bvh4d.rs:72-76),
 *   MATERIALS 48 B (GpuMaterial), STARS 16 B (pack_stars words, gpu.rs:102-119),
 *   OBJECTS 32 B (specrel_object: input-index range and centre per worldtube; optional).
 * NODES, GEOMS, ATTRS, ORDER, MATERIALS and STARS are required (possibly empty); their counts
 * must equal the header's.
 *
 * Checksum of a section: CRC-32C (Castagnoli, reflected, init and final xor 0xFFFFFFFF) of each
 * 1 MiB block of its bytes (the last block shorter), then CRC-32C of the concatenated
 * little-endian block CRCs. An empty section's checksum is the CRC-32C of zero bytes, 0.
 *
 * Readers accept format version 1 and reject files whose magic, byte-order mark, header size,
 * section table, bounds, element sizes or (when asked to verify) checksums do not match.
 */
#ifndef SPECREL_FILE_H
#define SPECREL_FILE_H

#include <stdint.h>

#include "specrel_types.h"

#ifdef __cplusplus
extern "C" {
#endif

struct specrel_object; /* specrel_scene.h */

#define SPECREL_FILE_MAGIC "SPECRL4D"
#define SPECREL_FILE_VERSION 1u
#define SPECREL_FILE_BOM 0x01020304u
#define SPECREL_FILE_ALIGN 4096u
#define SPECREL_FILE_BLOCK (1u << 20)

enum {
    SPECREL_SECTION_NODES = 1,
    SPECREL_SECTION_GEOMS = 2,
    SPECREL_SECTION_ATTRS = 3,
    SPECREL_SECTION_ORDER = 4,
    SPECREL_SECTION_MATERIALS = 5,
    SPECREL_SECTION_STARS = 6,
    SPECREL_SECTION_OBJECTS = 7
};

typedef struct specrel_file_header {
    char magic[8];            /* SPECREL_FILE_MAGIC, not NUL-terminated */
    uint32_t version;         /* SPECREL_FILE_VERSION */
    uint32_t byte_order;      /* SPECREL_FILE_BOM as written by the producer */
    uint32_t header_bytes;    /* 128 */
    uint32_t num_sections;
    uint32_t num_nodes, num_prims, num_materials, num_stars, num_star_words, num_objects;
    float light_pos[4];       /* (0, x, y, z), gpu.rs:286 */
    uint64_t bvh_layout;      /* BvhVersion (bvh4d.rs:47-61) of the packed tree */
    uint64_t bvh_edit;
    uint64_t file_bytes;      /* total size, for truncation checks */
    uint32_t flags;           /* bit 0: section checksums are filled in */
    uint32_t reserved[9];
} specrel_file_header;

typedef struct specrel_file_section {
    uint32_t kind;            /* SPECREL_SECTION_* */
    uint32_t elem_bytes;
    uint64_t count;
    uint64_t offset;          /* multiple of SPECREL_FILE_ALIGN */
    uint32_t checksum;        /* see above; 0 when the header's flags bit 0 is clear */
    uint32_t reserved;
} specrel_file_section;

typedef struct specrel_file_info {
    uint32_t version, flags;
    uint64_t bvh_layout, bvh_edit;
    uint64_t file_bytes;
    uint32_t num_objects, reserved;
} specrel_file_info;

typedef struct specrel_scene_file specrel_scene_file;

/* Writes `s` (and, when n_objects > 0, the object table) to `path`, with checksums.
 * bvh_layout / bvh_edit are stored as given. Returns 0, or a negative code with the message in
 * specrel_scene_last_error(). */
int specrel_scene_file_write(const char *path, const specrel_scene_arrays *s, const struct specrel_object *objects,
                             uint32_t n_objects, uint64_t bvh_layout, uint64_t bvh_edit);
/* Maps `path` read-only and fills `arrays` with pointers into the mapping (valid until
 * specrel_scene_file_close). verify != 0 checks every section checksum. */
int specrel_scene_file_open(const char *path, int verify, specrel_scene_file **out, specrel_scene_arrays *arrays,
                            specrel_file_info *info);
/* The object table of an open file (NULL when it has none). */
const struct specrel_object *specrel_scene_file_objects(const specrel_scene_file *f, uint32_t *count);
void specrel_scene_file_close(specrel_scene_file *f);

#ifdef __cplusplus
}
#endif

#if defined(__cplusplus)
static_assert(sizeof(specrel_file_header) == 128, "file header is 128 B");
static_assert(sizeof(specrel_file_section) == 32, "section entry is 32 B");
#else
_Static_assert(sizeof(specrel_file_header) == 128, "file header is 128 B");
_Static_assert(sizeof(specrel_file_section) == 32, "section entry is 32 B");
#endif

#endif
