"""Host-side readers on malformed input (no GPU calls): the LTWS weight
container shared by the warm-start network and the super-resolution
upscaler (csrc/ltws.c, format of simulation/src/superres.c:43-101) and the
OBJ reader (csrc/lbm_model.c, record rules of
simulation/obj-file-loader/lib/model_loader.c:7-142).  These are the inputs
the sanitizer build (csrc/sanitize.mk) runs against."""
import ctypes
import struct

import numpy as np
import pytest

from lattice_fluid_dynamics_amd import lbm
from lattice_fluid_dynamics_amd import model as M

MAX_DIMS = 8
OK, CANNOT_OPEN, BAD_MAGIC, TRUNCATED = 0, -1, -2, -3


class Ltws(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint32), ("tensors", ctypes.c_uint32),
                ("ndim", ctypes.POINTER(ctypes.c_uint32)), ("dims", ctypes.POINTER(ctypes.c_uint32 * MAX_DIMS)),
                ("offset", ctypes.POINTER(ctypes.c_long)), ("total", ctypes.c_long),
                ("data", ctypes.POINTER(ctypes.c_float))]


@pytest.fixture(scope="module")
def L():
    lib = lbm.load_library()
    lib.lbm_ltws_read.restype = ctypes.c_int
    lib.lbm_ltws_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(Ltws)]
    lib.lbm_ltws_free.argtypes = [ctypes.POINTER(Ltws)]
    lib.lbm_ltws_write.restype = ctypes.c_int
    lib.lbm_ltws_write.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                   ctypes.POINTER(ctypes.c_uint32 * MAX_DIMS), ctypes.POINTER(ctypes.c_float)]
    return lib


def read(L, path):
    t = Ltws()
    rc = L.lbm_ltws_read(str(path).encode(), ctypes.byref(t))
    out = None
    if rc == OK:
        shapes = [tuple(t.dims[p][d] for d in range(t.ndim[p])) for p in range(t.tensors)]
        data = np.ctypeslib.as_array(t.data, (t.total,)).copy() if t.total else np.zeros(0, np.float32)
        out = (t.version, shapes, [t.offset[p] for p in range(t.tensors)], data)
        L.lbm_ltws_free(ctypes.byref(t))
    else:
        assert not t.data and not t.ndim and t.total == 0  # nothing is handed out on failure
    return rc, out


def ltws_bytes(shapes, data, magic=0x4C545753, version=1, count=None):
    b = struct.pack("<3I", magic, version, len(shapes) if count is None else count)
    o = 0
    for s in shapes:
        n = int(np.prod(s)) if s else 1
        b += struct.pack(f"<{1 + len(s)}I", len(s), *s) + np.asarray(data[o:o + n], "<f4").tobytes()
        o += n
    return b


def test_writer_and_reader_agree_with_the_byte_format(L, tmp_path):
    shapes = [(3, 2), (4,), (2, 1, 3)]
    data = np.arange(6 + 4 + 6, dtype=np.float32) * 0.5 - 3
    p = tmp_path / "w.ltws"
    nd = (ctypes.c_uint32 * 3)(*[len(s) for s in shapes])
    dims = ((ctypes.c_uint32 * MAX_DIMS) * 3)()
    for i, s in enumerate(shapes):
        for d, v in enumerate(s):
            dims[i][d] = v
    assert L.lbm_ltws_write(str(p).encode(), 3, nd, dims, data.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == OK
    assert p.read_bytes() == ltws_bytes(shapes, data)
    rc, (version, got_shapes, offsets, got) = read(L, p)
    assert rc == OK and version == 1 and got_shapes == shapes and offsets == [0, 6, 10]
    assert np.array_equal(got, data)


def test_every_truncation_is_rejected(L, tmp_path):
    shapes = [(2, 3), (5,)]
    full = ltws_bytes(shapes, np.arange(11, dtype=np.float32))
    p = tmp_path / "t.ltws"
    for cut in range(len(full)):
        p.write_bytes(full[:cut])
        rc, _ = read(L, p)
        assert rc == (BAD_MAGIC if cut < 4 else TRUNCATED), cut
    p.write_bytes(full)
    assert read(L, p)[0] == OK


def test_foreign_and_oversized_headers_are_rejected(L, tmp_path):
    p = tmp_path / "h.ltws"
    assert read(L, tmp_path / "missing.ltws")[0] == CANNOT_OPEN
    p.write_bytes(ltws_bytes([(2,)], [1, 2], magic=0x4C545754))
    assert read(L, p)[0] == BAD_MAGIC
    # sizes whose product overflows 64 bits, or exceeds the 2^28-float limit
    p.write_bytes(struct.pack("<3I", 0x4C545753, 1, 1) + struct.pack("<9I", 8, *([0xFFFFFFFF] * 8)))
    assert read(L, p)[0] == TRUNCATED
    p.write_bytes(struct.pack("<3I", 0x4C545753, 1, 1) + struct.pack("<3I", 2, 1 << 14, 1 << 15))
    assert read(L, p)[0] == TRUNCATED
    # more dimensions than the container allows, and a huge tensor count
    p.write_bytes(struct.pack("<3I", 0x4C545753, 1, 1) + struct.pack("<10I", 9, *([1] * 9)))
    assert read(L, p)[0] == TRUNCATED
    p.write_bytes(struct.pack("<3I", 0x4C545753, 1, 0xFFFFFFFF))
    assert read(L, p)[0] == TRUNCATED
    # a claimed size far beyond the bytes present
    p.write_bytes(struct.pack("<3I", 0x4C545753, 1, 1) + struct.pack("<2I", 1, 1 << 27) + b"\0" * 64)
    assert read(L, p)[0] == TRUNCATED


def test_degenerate_but_valid_files(L, tmp_path):
    p = tmp_path / "d.ltws"
    p.write_bytes(struct.pack("<3I", 0x4C545753, 7, 0))  # no tensors; the version is the caller's concern
    rc, (version, shapes, _, data) = read(L, p)
    assert rc == OK and version == 7 and shapes == [] and data.size == 0
    p.write_bytes(ltws_bytes([(0, 5), (), (3,)], [9, 8, 7, 6]))  # an empty tensor and a scalar
    rc, (_, shapes, offsets, data) = read(L, p)
    assert rc == OK and shapes == [(0, 5), (), (3,)] and offsets == [0, 0, 1]
    assert data.tolist() == [9, 8, 7, 6]


def test_obj_reader_on_hostile_lines(tmp_path):
    long_x = "v " + "1" * 2000 + " 2 3\n"  # longer than both line readers
    long_tok = "f " + "1" * 100 + " 2 3\n"  # a token longer than the 63-character token buffer
    text = ("v 0 0 0\nv 1 0 0\nv 0 1 0\n" + long_x + long_tok + "f\n" + "v\n" + "\n" * 3 + "f 1\n"
            + "f a/b/c x y\n" + "f 1/2/3 2//5 3/1\n" + "\0garbage\n" + "f " + " ".join(["1"] * 300) + "\n")
    p = tmp_path / "hostile.obj"
    p.write_bytes(text.encode())
    m = M.LoadedModel(str(p))
    v, f = m.vertices(), m.faces()
    assert v.shape[0] == m.m.vertexCount and f.shape[0] == m.m.faceCount
    assert np.array_equal(v[:3], [[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    assert [1, 2, 3] in f.tolist()  # "f 1/2/3 2//5 3/1" keeps the leading integers
    tris = m.triangles(1.0, 0.0, 0.0, 0.0, 0.0)
    assert tris.shape[1] == 12 and np.isfinite(tris).all()
    # only faces whose three references name existing vertices become triangles
    nv = m.m.vertexCount
    assert tris.shape[0] == sum(all(1 <= r <= nv for r in face) for face in f.tolist())
