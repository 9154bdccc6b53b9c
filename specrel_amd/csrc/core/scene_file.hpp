// Reader and writer of the versioned scene file (include/specrel_file.h), shared by the scene
// library (producer side) and the renderer (specrel_upload_scene_file). Host-only.
//
// The reader maps the file read-only and hands out pointers into the mapping; checksums are
// computed over 1 MiB blocks in parallel (slicing-by-8 CRC-32C, one block per task).
#pragma once
#include <fcntl.h>
#include <stdint.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "specrel_file.h"

namespace sr {
namespace file {

// CRC-32C (Castagnoli polynomial 0x1EDC6F41, reflected 0x82F63B78), slicing-by-8 tables.
struct Crc32c {
    uint32_t t[8][256];
    Crc32c() {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0x82F63B78u & (0u - (c & 1u)));
            t[0][i] = c;
        }
        for (uint32_t i = 0; i < 256; ++i)
            for (int s = 1; s < 8; ++s) t[s][i] = (t[s - 1][i] >> 8) ^ t[0][t[s - 1][i] & 0xFF];
    }
    uint32_t operator()(const uint8_t *p, size_t n, uint32_t crc = 0) const {
        uint32_t c = ~crc;
        while (n && ((uintptr_t)p & 7)) {
            c = (c >> 8) ^ t[0][(c ^ *p++) & 0xFF];
            --n;
        }
        while (n >= 8) {
            uint32_t lo, hi;
            std::memcpy(&lo, p, 4);
            std::memcpy(&hi, p + 4, 4);
            lo ^= c;
            c = t[7][lo & 0xFF] ^ t[6][(lo >> 8) & 0xFF] ^ t[5][(lo >> 16) & 0xFF] ^ t[4][lo >> 24] ^
                t[3][hi & 0xFF] ^ t[2][(hi >> 8) & 0xFF] ^ t[1][(hi >> 16) & 0xFF] ^ t[0][hi >> 24];
            p += 8;
            n -= 8;
        }
        while (n--) c = (c >> 8) ^ t[0][(c ^ *p++) & 0xFF];
        return ~c;
    }
};

inline const Crc32c &crc32c() {
    static const Crc32c table;
    return table;
}

// Section checksum: CRC-32C of every SPECREL_FILE_BLOCK block, then of the block CRCs.
inline uint32_t section_checksum(const uint8_t *p, uint64_t bytes) {
    if (bytes == 0) return 0;
    const uint64_t nb = (bytes + SPECREL_FILE_BLOCK - 1) / SPECREL_FILE_BLOCK;
    std::vector<uint32_t> blocks(nb);
    std::atomic<uint64_t> next{0};
    auto work = [&] {
        for (uint64_t b; (b = next.fetch_add(1)) < nb;) {
            const uint64_t off = b * SPECREL_FILE_BLOCK;
            blocks[b] = crc32c()(p + off, (size_t)std::min<uint64_t>(SPECREL_FILE_BLOCK, bytes - off));
        }
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nt = (unsigned)std::min<uint64_t>(hw, nb);
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < nt; ++i) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    std::vector<uint8_t> le(4 * nb);
    for (uint64_t b = 0; b < nb; ++b)
        for (int k = 0; k < 4; ++k) le[4 * b + k] = (uint8_t)(blocks[b] >> (8 * k));
    return crc32c()(le.data(), le.size());
}

struct Error {
    std::string msg;
};

inline uint64_t align_up(uint64_t x) { return (x + SPECREL_FILE_ALIGN - 1) / SPECREL_FILE_ALIGN * SPECREL_FILE_ALIGN; }

struct SectionIn {
    uint32_t kind, elem;
    const void *data;
    uint64_t count;
};

// Writes the file; throws Error.
inline void write(const char *path, const specrel_scene_arrays &s, const void *objects, uint32_t n_objects,
                  uint64_t layout, uint64_t edit) {
    if (!path) throw Error{"path is NULL"};
    if (s.num_nodes && !s.nodes) throw Error{"nodes is NULL"};
    if (s.num_prims && (!s.geoms || !s.attrs || !s.order)) throw Error{"prim arrays are NULL"};
    if (s.num_materials && !s.materials) throw Error{"materials is NULL"};
    if (s.num_star_words && !s.star_words) throw Error{"star words are NULL"};
    if (n_objects && !objects) throw Error{"objects is NULL"};
    std::vector<SectionIn> secs = {
        {SPECREL_SECTION_NODES, 48, s.nodes, s.num_nodes},
        {SPECREL_SECTION_GEOMS, 80, s.geoms, s.num_prims},
        {SPECREL_SECTION_ATTRS, 192, s.attrs, s.num_prims},
        {SPECREL_SECTION_ORDER, 4, s.order, s.num_prims},
        {SPECREL_SECTION_MATERIALS, 48, s.materials, s.num_materials},
        {SPECREL_SECTION_STARS, 16, s.star_words, s.num_star_words},
    };
    if (n_objects) secs.push_back({SPECREL_SECTION_OBJECTS, 32, objects, n_objects});
    specrel_file_header h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, SPECREL_FILE_MAGIC, 8);
    h.version = SPECREL_FILE_VERSION;
    h.byte_order = SPECREL_FILE_BOM;
    h.header_bytes = sizeof(specrel_file_header);
    h.num_sections = (uint32_t)secs.size();
    h.num_nodes = s.num_nodes;
    h.num_prims = s.num_prims;
    h.num_materials = s.num_materials;
    h.num_stars = s.num_stars;
    h.num_star_words = s.num_star_words;
    h.num_objects = n_objects;
    std::memcpy(h.light_pos, s.light_pos, sizeof h.light_pos);
    h.bvh_layout = layout;
    h.bvh_edit = edit;
    h.flags = 1u;
    std::vector<specrel_file_section> table(secs.size());
    uint64_t off = align_up(sizeof h + secs.size() * sizeof(specrel_file_section));
    for (size_t i = 0; i < secs.size(); ++i) {
        specrel_file_section &e = table[i];
        std::memset(&e, 0, sizeof e);
        e.kind = secs[i].kind;
        e.elem_bytes = secs[i].elem;
        e.count = secs[i].count;
        e.offset = off;
        e.checksum = section_checksum((const uint8_t *)secs[i].data, e.count * e.elem_bytes);
        off = align_up(off + e.count * e.elem_bytes);
    }
    h.file_bytes = off;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) throw Error{std::string("cannot create ") + path + ": " + std::strerror(errno)};
    bool ok = std::fwrite(&h, sizeof h, 1, fp) == 1 &&
              std::fwrite(table.data(), sizeof(specrel_file_section), table.size(), fp) == table.size();
    uint64_t pos = sizeof h + table.size() * sizeof(specrel_file_section);
    static const char zeros[SPECREL_FILE_ALIGN] = {};
    for (size_t i = 0; ok && i < secs.size(); ++i) {
        const uint64_t pad = table[i].offset - pos;
        ok = pad == 0 || std::fwrite(zeros, 1, (size_t)pad, fp) == pad;
        const uint64_t bytes = table[i].count * table[i].elem_bytes;
        const uint8_t *p = (const uint8_t *)secs[i].data;
        for (uint64_t done = 0; ok && done < bytes;) {
            const size_t chunk = (size_t)std::min<uint64_t>(bytes - done, 64ull << 20);
            ok = std::fwrite(p + done, 1, chunk, fp) == chunk;
            done += chunk;
        }
        pos = table[i].offset + bytes;
    }
    if (ok && h.file_bytes > pos) ok = std::fwrite(zeros, 1, (size_t)(h.file_bytes - pos), fp) == h.file_bytes - pos;
    if (std::fclose(fp) != 0) ok = false;
    if (!ok) {
        std::remove(path);
        throw Error{std::string("write failed: ") + path};
    }
}

// A mapped, validated file.
struct Mapped {
    void *base = nullptr;
    size_t bytes = 0;
    specrel_file_header h{};
    const void *objects = nullptr;
    ~Mapped() {
        if (base) munmap(base, bytes);
    }
};

// Maps and validates `path`, fills `a` with pointers into the mapping; throws Error.
inline void open(const char *path, bool verify, Mapped &m, specrel_scene_arrays &a) {
    if (!path) throw Error{"path is NULL"};
    const int fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (fd < 0) throw Error{std::string("cannot open ") + path + ": " + std::strerror(errno)};
    struct stat st;
    if (fstat(fd, &st) != 0) {
        ::close(fd);
        throw Error{std::string("cannot stat ") + path};
    }
    const uint64_t size = (uint64_t)st.st_size;
    if (size < sizeof(specrel_file_header)) {
        ::close(fd);
        throw Error{std::string(path) + ": too short for a scene file header"};
    }
    void *p = mmap(nullptr, (size_t)size, PROT_READ, MAP_PRIVATE, fd, 0);
    ::close(fd);
    if (p == MAP_FAILED) throw Error{std::string("cannot map ") + path + ": " + std::strerror(errno)};
    m.base = p;
    m.bytes = (size_t)size;
    madvise(p, (size_t)size, MADV_SEQUENTIAL);
    const uint8_t *b = (const uint8_t *)p;
    specrel_file_header &h = m.h;
    std::memcpy(&h, b, sizeof h);
    const std::string where = std::string(path) + ": ";
    if (std::memcmp(h.magic, SPECREL_FILE_MAGIC, 8) != 0) throw Error{where + "not a scene file (bad magic)"};
    if (h.byte_order != SPECREL_FILE_BOM) throw Error{where + "byte order mark mismatch (written big-endian?)"};
    if (h.version != SPECREL_FILE_VERSION)
        throw Error{where + "format version " + std::to_string(h.version) + " is not supported (this reader: " +
                    std::to_string(SPECREL_FILE_VERSION) + ")"};
    if (h.header_bytes != sizeof(specrel_file_header)) throw Error{where + "unexpected header size"};
    if (h.file_bytes != size)
        throw Error{where + "file is " + std::to_string(size) + " bytes, header says " + std::to_string(h.file_bytes) +
                    " (truncated or padded)"};
    if (h.num_sections == 0 || h.num_sections > 64) throw Error{where + "bad section count"};
    const uint64_t table_end = sizeof h + (uint64_t)h.num_sections * sizeof(specrel_file_section);
    if (table_end > size) throw Error{where + "section table past the end of the file"};
    std::vector<specrel_file_section> table(h.num_sections);
    std::memcpy(table.data(), b + sizeof h, table.size() * sizeof(specrel_file_section));
    const void *ptr[8] = {};
    bool seen[8] = {};
    for (const specrel_file_section &e : table) {
        static const uint32_t elem[8] = {0, 48, 80, 192, 4, 48, 16, 32};
        if (e.kind < 1 || e.kind > 7) continue;  // unknown kinds are skipped (forward compatible)
        if (seen[e.kind]) throw Error{where + "section " + std::to_string(e.kind) + " appears twice"};
        seen[e.kind] = true;
        if (e.elem_bytes != elem[e.kind])
            throw Error{where + "section " + std::to_string(e.kind) + " has " + std::to_string(e.elem_bytes) +
                        "-byte elements, expected " + std::to_string(elem[e.kind])};
        if (e.offset % SPECREL_FILE_ALIGN || e.offset < table_end || e.count > size / e.elem_bytes ||
            e.offset > size - e.count * e.elem_bytes)
            throw Error{where + "section " + std::to_string(e.kind) + " lies outside the file"};
        const uint64_t want = e.kind == SPECREL_SECTION_NODES       ? h.num_nodes
                              : e.kind == SPECREL_SECTION_MATERIALS ? h.num_materials
                              : e.kind == SPECREL_SECTION_STARS     ? h.num_star_words
                              : e.kind == SPECREL_SECTION_OBJECTS   ? h.num_objects
                                                                    : h.num_prims;
        if (e.count != want) throw Error{where + "section " + std::to_string(e.kind) + " count disagrees with the header"};
        ptr[e.kind] = b + e.offset;
    }
    for (int k = 1; k <= 6; ++k)
        if (!seen[k]) throw Error{where + "required section " + std::to_string(k) + " is missing"};
    if (h.num_objects && !seen[SPECREL_SECTION_OBJECTS]) throw Error{where + "object section is missing"};
    if (verify && (h.flags & 1u))
        for (const specrel_file_section &e : table) {
            if (e.kind < 1 || e.kind > 7) continue;
            const uint32_t c = section_checksum(b + e.offset, e.count * e.elem_bytes);
            if (c != e.checksum) throw Error{where + "checksum mismatch in section " + std::to_string(e.kind)};
        }
    std::memset(&a, 0, sizeof a);
    a.nodes = (const specrel_node *)ptr[SPECREL_SECTION_NODES];
    a.num_nodes = h.num_nodes;
    a.num_prims = h.num_prims;
    a.geoms = (const specrel_geom *)ptr[SPECREL_SECTION_GEOMS];
    a.attrs = (const specrel_attr *)ptr[SPECREL_SECTION_ATTRS];
    a.order = (const uint32_t *)ptr[SPECREL_SECTION_ORDER];
    a.materials = (const specrel_material *)ptr[SPECREL_SECTION_MATERIALS];
    a.num_materials = h.num_materials;
    a.num_stars = h.num_stars;
    a.star_words = (const float *)ptr[SPECREL_SECTION_STARS];
    a.num_star_words = h.num_star_words;
    std::memcpy(a.light_pos, h.light_pos, sizeof a.light_pos);
    m.objects = ptr[SPECREL_SECTION_OBJECTS];
}

}  // namespace file
}  // namespace sr
