"""The versioned scene file (include/specrel_file.h): round trips, the checksum definition
(pinned by the CRC-32C check value), the layout, and every rejection a bad file gets. CPU only;
the GPU upload of a file is in test_gpu_scene_file.py."""
import ctypes as C
import struct

import numpy as np
import pytest

from specrel_amd import _ffi as F
from specrel_amd import scene as S
from oracle import pyoracle as O
from tests.helpers import frame_for


def crc32c(data, crc=0):
    """Bitwise CRC-32C (Castagnoli, reflected polynomial 0x82F63B78)."""
    table = _table()
    c = crc ^ 0xFFFFFFFF
    for b in bytes(data):
        c = (c >> 8) ^ table[(c ^ b) & 0xFF]
    return c ^ 0xFFFFFFFF


_T = None


def _table():
    global _T
    if _T is None:
        _T = []
        for i in range(256):
            c = i
            for _ in range(8):
                c = (c >> 1) ^ (0x82F63B78 if c & 1 else 0)
            _T.append(c)
    return _T


def section_checksum(data):
    data = bytes(data)
    if not data:
        return 0
    blocks = [crc32c(data[i:i + (1 << 20)]) for i in range(0, len(data), 1 << 20)]
    return crc32c(b"".join(struct.pack("<I", c) for c in blocks))


def test_crc32c_check_value():
    # the catalogued check value of CRC-32C over "123456789"
    assert crc32c(b"123456789") == 0xE3069283


@pytest.fixture(scope="module")
def twin():
    return S.Scene.generate("twin")


def arrays_of(sc):
    return {"nodes": sc.nodes(), "geoms": sc.geoms(), "attrs": sc.attrs(), "order": sc.order(),
            "stars": sc.star_words()}


def read_table(path):
    raw = open(path, "rb").read()
    h = F.FileHeader.from_buffer_copy(raw[:128])
    secs = [F.FileSection.from_buffer_copy(raw[128 + 32 * i:160 + 32 * i]) for i in range(h.num_sections)]
    return raw, h, secs


def test_round_trip_is_byte_identical(twin, tmp_path):
    p = tmp_path / "twin.srs"
    twin.save(p, bvh_layout=7, bvh_edit=3)
    f = S.SceneFile(p)
    for k, a in arrays_of(twin).items():
        b = arrays_of(f)[k]
        assert a.shape == b.shape and a.tobytes() == b.tobytes(), k
    assert f.arrays.num_materials == twin.arrays.num_materials and f.arrays.num_stars == twin.arrays.num_stars
    ma = C.string_at(twin.arrays.materials, 48 * twin.arrays.num_materials)
    assert C.string_at(f.arrays.materials, 48 * f.arrays.num_materials) == ma
    assert list(f.arrays.light_pos) == list(twin.arrays.light_pos)
    assert f.objects() == twin.objects() and len(f.objects()) > 0
    assert (f.info.version, f.info.bvh_layout, f.info.bvh_edit) == (1, 7, 3)
    # saving the loaded file gives the same bytes
    q = tmp_path / "again.srs"
    f.save(q)
    assert open(p, "rb").read() == open(q, "rb").read()
    f.close()


def test_layout_and_checksums_follow_the_format(tmp_path):
    sc = S.Scene.generate("large", S.gen_params("large", n_objects=2, slices=8, n_stars=50))
    p = tmp_path / "large.srs"
    sc.save(p)
    raw, h, secs = read_table(p)
    assert h.magic == b"SPECRL4D" and h.version == 1 and h.byte_order == 0x01020304 and h.header_bytes == 128
    assert h.file_bytes == len(raw) and h.flags & 1
    assert [s.kind for s in secs] == [1, 2, 3, 4, 5, 6, 7]
    assert [s.elem_bytes for s in secs] == [48, 80, 192, 4, 48, 16, 32]
    assert [s.count for s in secs] == [h.num_nodes, h.num_prims, h.num_prims, h.num_prims, h.num_materials,
                                       h.num_star_words, h.num_objects]
    ends = 128 + 32 * len(secs)
    for s in secs:
        assert s.offset % 4096 == 0 and s.offset >= ends
        ends = s.offset + s.count * s.elem_bytes
    attrs = secs[2]
    assert attrs.count * 192 > (2 << 20)  # several 1 MiB blocks
    for s in secs:
        assert s.checksum == section_checksum(raw[s.offset:s.offset + s.count * s.elem_bytes]), s.kind
    assert raw[secs[1].offset:secs[1].offset + 80 * h.num_prims] == sc.geoms().tobytes()


def test_oracle_renders_the_loaded_file_identically(twin, tmp_path):
    p = tmp_path / "twin.srs"
    twin.save(p)
    f = S.SceneFile(p)
    fr = frame_for("twin", 96, 60)
    a, b = O.render(twin, fr), O.render(f, fr)
    assert (a["rgb"] == b["rgb"]).all() and (a["hit"] == b["hit"]).all()


def patched(path, tmp_path, edit):
    raw = bytearray(open(path, "rb").read())
    edit(raw)
    q = tmp_path / "bad.srs"
    q.write_bytes(bytes(raw))
    return q


@pytest.mark.parametrize("what,edit,message", [
    ("magic", lambda r: r.__setitem__(slice(0, 8), b"NOTSCENE"), "bad magic"),
    ("version", lambda r: r.__setitem__(slice(8, 12), struct.pack("<I", 2)), "format version 2"),
    ("byte order", lambda r: r.__setitem__(slice(12, 16), struct.pack(">I", 0x01020304)), "byte order"),
    ("truncated", lambda r: r.__delitem__(slice(len(r) - 4096, len(r))), "truncated"),
    ("elem size", lambda r: r.__setitem__(slice(128 + 32 + 4, 128 + 32 + 8), struct.pack("<I", 64)), "64-byte"),
    ("missing", lambda r: r.__setitem__(slice(128 + 32, 128 + 36), struct.pack("<I", 99)), "required section 2"),
    ("count", lambda r: r.__setitem__(slice(128 + 32 * 3 + 8, 128 + 32 * 3 + 16), struct.pack("<Q", 5)),
     "count disagrees"),
    ("bounds", lambda r: r.__setitem__(slice(128 + 32 * 2 + 16, 128 + 32 * 2 + 24), struct.pack("<Q", 1 << 40)),
     "outside the file"),
])
def test_bad_files_are_rejected(twin, tmp_path, what, edit, message):
    p = tmp_path / "twin.srs"
    twin.save(p)
    q = patched(p, tmp_path, edit)
    with pytest.raises(RuntimeError, match=message):
        S.SceneFile(q)


def test_checksums_catch_a_flipped_byte(twin, tmp_path):
    p = tmp_path / "twin.srs"
    twin.save(p)
    _, h, secs = read_table(p)
    geoms = secs[1]

    def flip(r):
        r[geoms.offset + 100] ^= 0x40
    q = patched(p, tmp_path, flip)
    with pytest.raises(RuntimeError, match="checksum mismatch in section 2"):
        S.SceneFile(q)
    S.SceneFile(q, verify=False).close()  # the layout is still valid
    with pytest.raises(RuntimeError, match="cannot open"):
        S.SceneFile(tmp_path / "absent.srs")
