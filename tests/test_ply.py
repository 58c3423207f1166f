"""PLY input (PLYPointCloudLoader.swift) through the C ABI — host code, no GPU.

Pinned by the reference's own PLY case (LayoutTests.swift:44-67), by closed
forms of every scalar type's conversion, by ASCII/binary agreement on the same
values, and by each error of PLYPointCloudLoaderError."""
import struct
import zlib

import numpy as np
import pytest

from _pkg import orc, scr

_FMT = {"char": "b", "uchar": "B", "short": "h", "ushort": "H", "int": "i", "uint": "I",
        "float": "f", "double": "d", "int8": "b", "uint8": "B", "int16": "h", "uint16": "H",
        "int32": "i", "uint32": "I", "float32": "f", "float64": "d"}
_MAX = {"char": 127, "int8": 127, "uchar": 255, "uint8": 255, "short": 32767, "int16": 32767,
        "ushort": 65535, "uint16": 65535, "int": 2147483647, "int32": 2147483647,
        "uint": 4294967295, "uint32": 4294967295}


def _ply(fmt, props, rows, head_extra=(), eol="\n", count=None):
    head = ["ply", f"format {fmt} 1.0", *head_extra, f"element vertex {len(rows) if count is None else count}"]
    head += [f"property {t} {n}" for t, n in props]
    head.append("end_header")
    data = eol.join(head).encode() + eol.encode()
    if fmt == "ascii":
        def text(t, v):
            return repr(float(np.float32(v))) if t in ("float", "float32", "double", "float64") else str(int(v))
        data += eol.join(" ".join(text(t, v) for (t, _), v in zip(props, r)) for r in rows).encode()
    else:
        for r in rows:
            data += b"".join(struct.pack("<" + _FMT[t], v) for (t, _), v in zip(props, r))
    return data


def _expect_channel(t, v):
    f = np.float32(v)
    if t in _MAX:
        f = np.float32(f / np.float32(_MAX[t]))
    return np.float32(min(max(f, np.float32(0)), np.float32(1))) if f == f else np.float32(0)


def test_reference_ascii_case_parses_and_packs():
    # LayoutTests.swift:44-67, verbatim input
    ply = ("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
           "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n"
           "0 0 0 255 0 0\n1 0 0 0 255 0\n0 1 0 0 0 255").encode()
    pos, col = scr.PLYPointCloudLoader.parse(ply)
    assert np.array_equal(pos, [[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0]])
    assert np.array_equal(col, [[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1]])
    packed = orc.pack(pos[:, :3], col, pointsPerBatch=2)
    assert packed.pointCount == 3 and packed.batchCount == 2
    assert set(int(c) & 0x00FFFFFF for c in packed.colors) == {0x000000FF, 0x0000FF00, 0x00FF0000}
    assert packed.levels.shape[0] == 3


@pytest.mark.parametrize("ctype", ["char", "uchar", "short", "ushort", "int", "uint", "float", "double",
                                   "int8", "uint8", "int16", "uint16", "int32", "uint32", "float32", "float64"])
def test_binary_scalar_types_closed_forms(ctype):
    rng = np.random.default_rng(zlib.crc32(ctype.encode()))
    if ctype in _MAX:
        info = np.iinfo({"b": np.int8, "B": np.uint8, "h": np.int16, "H": np.uint16,
                         "i": np.int32, "I": np.uint32}[_FMT[ctype]])
        vals = [info.min, info.max, 0, 1] + [int(v) for v in rng.integers(info.min, info.max, 12, endpoint=True)]
    else:
        vals = [0.0, 1.0, -0.5, 1.5, 0.25, 1e-30, float("inf"), float("nan")] + list(rng.normal(0.5, 1.0, 8))
    # the scalar type serves as a coordinate (x) and as a colour channel (red)
    props = [(ctype, "x"), ("float", "y"), ("double", "z"), (ctype, "red")]
    rows = [(v, 2.0, 3.0, v) for v in vals]
    pos, col = scr.PLYPointCloudLoader.parse(_ply("binary_little_endian", props, rows))
    for i, v in enumerate(vals):
        x = np.float32(v)
        assert pos[i, 0] == x or (x != x and pos[i, 0] != pos[i, 0]), (v, pos[i, 0])
        assert pos[i, 1] == 2.0 and pos[i, 2] == 3.0 and pos[i, 3] == 0.0
        assert col[i, 0] == _expect_channel(ctype, v), (ctype, v, col[i, 0])
        assert np.array_equal(col[i, 1:], [1, 1, 1])


def test_ascii_and_binary_agree():
    rng = np.random.default_rng(11)
    n = 500
    props = [("float", "x"), ("float", "y"), ("float", "z"), ("uchar", "red"), ("ushort", "green"),
             ("float", "blue"), ("int", "intensity")]
    rows = [(float(a), float(b), float(c), int(r), int(g), float(bl), int(it)) for a, b, c, r, g, bl, it in zip(
        rng.normal(0, 100, n).astype(np.float32), rng.normal(0, 100, n).astype(np.float32),
        rng.normal(0, 100, n).astype(np.float32), rng.integers(0, 256, n), rng.integers(0, 65536, n),
        rng.random(n).astype(np.float32), rng.integers(-1000, 1000, n))]
    pa, ca = scr.PLYPointCloudLoader.parse(_ply("ascii", props, rows))
    pb, cb = scr.PLYPointCloudLoader.parse(_ply("binary_little_endian", props, rows))
    assert pa.shape == (n, 4) and np.array_equal(pa, pb) and np.array_equal(ca, cb)
    assert np.array_equal(pa[:, 0], np.array([r[0] for r in rows], np.float32))
    assert np.array_equal(ca[:, 1], np.array([np.float32(r[4]) / np.float32(65535) for r in rows], np.float32))


def test_header_rules_aliases_and_ignored_elements():
    # CRLF, comments and obj_info, runs of spaces, property aliases, an
    # ignored property, a later duplicate role winning, a face element after
    # the vertices, blank body lines and no "ply" magic line
    text = ("format  ascii 1.0\r\ncomment made by hand\r\nobj_info x\r\nelement vertex 2\r\n"
            "property double x\r\nproperty double y\r\nproperty double z\r\nproperty uchar r\r\n"
            "property uchar diffuse_green\r\nproperty float alpha\r\nproperty uchar red\r\n"
            "element face 1\r\nproperty list uchar int vertex_indices\r\nend_header\r\n"
            "1.5  2.5 3.5 10 51 0.5 255\r\n\r\n-1 -2 -3 0 255 0.25 0\r\n3 0 1 2\r\n")
    pos, col = scr.PLYPointCloudLoader.parse(text.encode())
    assert np.array_equal(pos, [[1.5, 2.5, 3.5, 0], [-1, -2, -3, 0]])
    assert np.array_equal(col, np.array([[1, np.float32(51) / np.float32(255), 1, 1], [0, 1, 1, 1]], np.float32))


def test_binary_body_follows_header_newline_rules():
    # end_header followed by CR LF: the body starts after both
    props = [("float", "x"), ("float", "y"), ("float", "z")]
    data = _ply("binary_little_endian", props, [(1.0, 2.0, 3.0)], eol="\r\n")
    pos, _ = scr.PLYPointCloudLoader.parse(data)
    assert np.array_equal(pos, [[1, 2, 3, 0]])


def test_empty_vertex_element():
    pos, col = scr.PLYPointCloudLoader.parse(_ply("ascii", [("float", "x"), ("float", "y"), ("float", "z")], []))
    assert pos.shape == (0, 4) and col.shape == (0, 4)


def test_load_reads_a_file(tmp_path):
    props = [("float", "x"), ("float", "y"), ("float", "z"), ("uchar", "blue")]
    f = tmp_path / "cloud.ply"
    f.write_bytes(_ply("binary_little_endian", props, [(1.0, 2.0, 3.0, 51), (4.0, 5.0, 6.0, 255)]))
    pos, col = scr.PLYPointCloudLoader.load(f)
    assert np.array_equal(pos[:, :3], [[1, 2, 3], [4, 5, 6]])
    assert np.array_equal(col[:, 2], np.array([np.float32(51) / np.float32(255), 1], np.float32))


_XYZ = [("float", "x"), ("float", "y"), ("float", "z")]


@pytest.mark.parametrize("data,code,text", [
    (b"", -11, "not a valid PLY"),
    (b"ply\nformat ascii 1.0\nelement vertex 1\n", -11, "not a valid PLY"),
    (b"ply\nformat binary_big_endian 1.0\nend_header\n", -12, "Unsupported PLY format: binary_big_endian."),
    (b"ply\nelement vertex 1\nproperty float x\nend_header\n", -11, "not a valid PLY"),
    (b"ply\nformat ascii\nelement vertex 1\nend_header\n", -11, "not a valid PLY"),
    (b"ply\nformat ascii 1.0\nelement face 1\nend_header\n", -13, "does not contain a vertex element"),
    (b"ply\nformat ascii 1.0\nelement vertex x1\nend_header\n", -11, "not a valid PLY"),
    (b"ply\nformat ascii 1.0\nelement vertex -1\nend_header\n", -11, "negative vertex count"),
    (b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nproperty float y\nend_header\n0 0\n",
     -14, "missing required property 'x, y, z'"),
    (b"ply\nformat ascii 1.0\nelement vertex 1\nproperty list uchar int vertex_indices\nend_header\n",
     -15, "'list uchar int vertex_indices'"),
    (b"ply\nformat ascii 1.0\nelement vertex 1\nproperty int64 x\nend_header\n", -15, "property 'int64'"),
    (b"ply\ncomment \xff\nformat ascii 1.0\nelement vertex 0\nproperty float x\nproperty float y\n"
     b"property float z\nend_header\n", -11, "not a valid PLY"),
])
def test_header_errors(data, code, text):
    with pytest.raises(scr.PLYPointCloudLoaderError) as e:
        scr.PLYPointCloudLoader.parse(data)
    assert e.value.code == code and text in str(e.value), str(e.value)


def test_body_errors():
    with pytest.raises(scr.PLYPointCloudLoaderError) as e:   # fewer lines than vertices
        scr.PLYPointCloudLoader.parse(_ply("ascii", _XYZ, [(0, 0, 0)], count=2))
    assert e.value.code == -16
    bad = _ply("ascii", _XYZ, [(0, 0, 0)]).replace(b"0.0 0.0 0.0", b"0.0 0.0")
    with pytest.raises(scr.PLYPointCloudLoaderError) as e:   # fewer values than properties
        scr.PLYPointCloudLoader.parse(bad)
    assert e.value.code == -16
    bad = _ply("ascii", _XYZ, [(0, 0, 0), (1, 1, 1)]).replace(b"1.0 1.0 1.0", b"1.0 1.0x 1.0")
    with pytest.raises(scr.PLYPointCloudLoaderError) as e:   # not a float
        scr.PLYPointCloudLoader.parse(bad)
    assert e.value.code == -17 and "property 'y' on vertex line 2" in str(e.value)
    with pytest.raises(scr.PLYPointCloudLoaderError) as e:   # body not UTF-8
        scr.PLYPointCloudLoader.parse(_ply("ascii", _XYZ, []) + b"\xc3\x28 0 0\n")
    assert e.value.code == -16
    data = _ply("binary_little_endian", _XYZ, [(1.0, 2.0, 3.0), (4.0, 5.0, 6.0)])
    with pytest.raises(scr.PLYPointCloudLoaderError) as e:   # one byte short
        scr.PLYPointCloudLoader.parse(data[:-1])
    assert e.value.code == -16
    scr.PLYPointCloudLoader.parse(data)                      # the whole body parses


def test_ascii_float_spellings():
    # Float(String) reads decimal, exponent and hex forms, inf and nan
    rows = b"1e2 -0x1p-2 .5\ninf -inf nan\n"
    head = b"ply\nformat ascii 1.0\nelement vertex 2\nproperty float x\nproperty float y\nproperty float z\nend_header\n"
    pos, _ = scr.PLYPointCloudLoader.parse(head + rows)
    assert pos[0, 0] == 100 and pos[0, 1] == -0.25 and pos[0, 2] == 0.5
    assert pos[1, 0] == np.inf and pos[1, 1] == -np.inf and np.isnan(pos[1, 2])
    with pytest.raises(scr.PLYPointCloudLoaderError):
        scr.PLYPointCloudLoader.parse(head + b"1\t2 3\n4 5 6\n")


_SCALARS = ["char", "uchar", "short", "ushort", "int", "uint", "float", "double"]


def _value(t, u):
    # a value of type t from u in [0, 1)
    if t in _MAX:
        info = np.iinfo({"b": np.int8, "B": np.uint8, "h": np.int16, "H": np.uint16,
                         "i": np.int32, "I": np.uint32}[_FMT[t]])
        return int(info.min + int(u * (int(info.max) - int(info.min))))
    return float(np.float32((u - 0.5) * 1000.0))


try:
    from hypothesis import given, settings, strategies as st
except ImportError:  # pragma: no cover
    given = None

if given is not None:
    @settings(max_examples=60, deadline=None)
    @given(st.lists(st.sampled_from(_SCALARS), min_size=0, max_size=4),
           st.lists(st.floats(0, 0.999), min_size=0, max_size=40), st.integers(0, 3), st.data())
    def test_random_layouts_parse_alike_and_truncation_is_reported(extra, us, pos, data):
        names = ["red", "green", "blue", "intensity"]
        props = [("float", "x"), ("float", "y"), ("float", "z")]
        for k, t in enumerate(extra):
            props.insert(min(pos, len(props)), (t, names[k]))
        n = len(us) // max(1, len(props))
        rows = [tuple(_value(t, us[(i * len(props) + k) % max(1, len(us))]) for k, (t, _) in enumerate(props))
                for i in range(n)]
        pa, ca = scr.PLYPointCloudLoader.parse(_ply("ascii", props, rows))
        binary = _ply("binary_little_endian", props, rows)
        pb, cb = scr.PLYPointCloudLoader.parse(binary)
        assert pa.shape == (n, 4) and np.array_equal(pa, pb) and np.array_equal(ca, cb)
        if n:
            cut = data.draw(st.integers(1, len(binary) - binary.index(b"end_header\n") - 11))
            with pytest.raises(scr.PLYPointCloudLoaderError) as e:
                scr.PLYPointCloudLoader.parse(binary[:-cut])
            assert e.value.code == -16
