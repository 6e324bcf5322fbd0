"""The PLY container and the cloud decoded from it (host code in libvsim.so,
no device calls), on the cases of apps/simulator/test/unit/splats.test.js:85-158
and against an independent numpy decode of the same bytes."""
import math

import numpy as np
import pytest

from vsim import ingest

NP_TYPES = {"char": "<i1", "int8": "<i1", "uchar": "<u1", "uint8": "<u1", "short": "<i2", "int16": "<i2",
            "ushort": "<u2", "uint16": "<u2", "int": "<i4", "int32": "<i4", "uint": "<u4", "uint32": "<u4",
            "float": "<f4", "float32": "<f4", "double": "<f8", "float64": "<f8"}


def build_ply(columns, fmt=None, extra_header=(), truncate=0, element="vertex", newline="\n", tail=b""):
    """A binary little-endian PLY in memory: columns = [(name, type, values)]."""
    count = len(columns[0][2])
    header = newline.join(["ply", fmt or "format binary_little_endian 1.0", f"element {element} {count}"] +
                          [f"property {t} {n}" for n, t, _ in columns] + list(extra_header) + ["end_header", ""])
    dtype = np.dtype([(f"c{i}", NP_TYPES[t]) for i, (_, t, _) in enumerate(columns)])
    rec = np.zeros(count, dtype)
    for i, (_, _, v) in enumerate(columns):
        rec[f"c{i}"] = v
    body = rec.tobytes() + tail
    data = header.encode() + body
    return data[:len(data) - truncate] if truncate else data


PROPS = [("x", "float", [1, 2, 3]), ("flag", "uchar", [7, 8, 9]), ("y", "double", [-1.5, 0, 1.5]),
         ("code", "ushort", [1000, 2000, 3000])]


def test_reads_every_scalar_type_into_floats():
    count, read = ingest.read_ply(build_ply(PROPS))
    assert count == 3
    assert list(read) == ["x", "flag", "y", "code"]
    assert read["x"].tolist() == [1, 2, 3] and read["flag"].tolist() == [7, 8, 9]
    assert read["y"].tolist() == [-1.5, 0, 1.5] and read["code"].tolist() == [1000, 2000, 3000]
    # every spelling the format allows, signed values and f32 rounding of doubles
    cols = [("a", "char", [-128, 127]), ("b", "int8", [-5, 5]), ("c", "short", [-32768, 32767]),
            ("d", "int16", [-2, 2]), ("e", "uint16", [65535, 0]), ("f", "int", [-2 ** 31, 2 ** 31 - 1]),
            ("g", "int32", [-7, 7]), ("h", "uint", [2 ** 32 - 1, 16777217]), ("i", "uint32", [1, 2]),
            ("j", "float32", [0.1, -0.25]), ("k", "float64", [0.1, 1e300]), ("l", "uint8", [255, 0])]
    count, read = ingest.read_ply(build_ply(cols))
    for name, t, v in cols:
        with np.errstate(over="ignore"):
            want = np.asarray(v, NP_TYPES[t]).astype(np.float64).astype(np.float32)
        assert np.array_equal(read[name], want), name
    assert read["k"][1] == np.inf and read["h"][1] == np.float32(16777217)


def test_strides_past_properties_it_was_not_asked_for():
    count, read = ingest.read_ply(build_ply(PROPS), ["y"])
    assert list(read) == ["y"] and read["y"].tolist() == [-1.5, 0, 1.5]
    _, read = ingest.read_ply(build_ply(PROPS), ["code", "x"])
    assert read["code"].tolist() == [1000, 2000, 3000] and read["x"].tolist() == [1, 2, 3]


def test_rejects_what_it_cannot_honestly_read():
    with pytest.raises(ValueError, match="binary_little_endian"):
        ingest.read_ply(build_ply(PROPS, fmt="format ascii 1.0"))
    with pytest.raises(ValueError, match="list property"):
        ingest.read_ply(build_ply(PROPS, extra_header=["property list uchar int vertex_index"]))
    with pytest.raises(ValueError, match="truncated"):
        ingest.read_ply(build_ply(PROPS, truncate=4))
    with pytest.raises(ValueError, match="no property nope"):
        ingest.read_ply(build_ply(PROPS), ["nope"])
    with pytest.raises(ValueError, match="not a PLY"):
        ingest.read_ply(b"not a ply at all")
    with pytest.raises(ValueError, match="unknown PLY property type"):
        ingest.read_ply(build_ply(PROPS, extra_header=["property half w"]))
    with pytest.raises(ValueError, match="no usable element"):
        ingest.read_ply(b"ply\nformat binary_little_endian 1.0\nend_header\n")
    # the reference's exact wording
    with pytest.raises(ValueError) as e:
        ingest.read_ply(build_ply(PROPS, truncate=4))
    assert str(e.value) == (f"PLY is truncated: 3 records of 15 bytes need {len(build_ply(PROPS))} bytes, "
                            f"file has {len(build_ply(PROPS)) - 4}")


def test_header_details():
    # CRLF line ends, only the first element read, a later duplicate name wins.
    data = build_ply(PROPS, newline="\r\n")
    assert ingest.read_ply(data)[1]["y"].tolist() == [-1.5, 0, 1.5]
    two = build_ply(PROPS, extra_header=["element face 4", "property uchar q"])
    count, read = ingest.read_ply(two)
    assert count == 3 and "q" not in read
    dup = build_ply([("x", "float", [1, 2]), ("x", "double", [5, 6])])
    assert ingest.read_ply(dup, ["x"])[1]["x"].tolist() == [5, 6]
    # bytes after the body are ignored
    assert ingest.read_ply(build_ply(PROPS, tail=b"junk"))[0] == 3
    assert ingest.ply_properties(build_ply(PROPS)) == (3, ["x", "flag", "y", "code"])


SH_C0 = 0.28209479177387814


def cloud_ply(values, order=None):
    names = order or ingest.REQUIRED_PROPERTIES
    return build_ply([(n, "float", values[n]) for n in names])


def test_cloud_undoes_the_file_encoding():
    v = {n: np.zeros(2) for n in ingest.REQUIRED_PROPERTIES}
    v["x"][:] = [1, -1]
    v["scale_0"][:] = [math.log(0.5), 0]
    v["opacity"][:] = [0, 100]
    v["f_dc_0"][:] = [1 / SH_C0, 0]
    v["rot_0"][:] = [-2, 0]   # unnormalised, negative real part
    c = ingest.cloud_from_ply(cloud_ply(v))
    assert c.count == 2
    assert c.positions[:, 0].tolist() == [1, -1]
    assert abs(c.scales[0, 0] - 0.5) < 1e-6 and abs(c.scales[1, 0] - 1) < 1e-6
    assert abs(c.opacity[0] - 0.5) < 1e-6 and abs(c.opacity[1] - 1) < 1e-6
    assert abs(c.colors[0, 0] - 1.5) < 1e-6 and abs(c.colors[1, 0] - 0.5) < 1e-6
    assert abs(c.rotations[0, 3] - 1) < 1e-6 and abs(np.linalg.norm(c.rotations[0].astype(np.float64)) - 1) < 1e-6
    assert c.rotations[1].tolist() == [0, 0, 0, 1]   # a zero quaternion is the identity


def test_cloud_matches_an_independent_decode():
    rng = np.random.default_rng(7)
    n = 5000
    v = {n_: rng.normal(size=n).astype(np.float32) for n_ in ingest.REQUIRED_PROPERTIES}
    v["opacity"] = (rng.normal(size=n) * 8).astype(np.float32)
    v["rot_0"][:7] = 0
    v["rot_1"][:7] = 0
    v["rot_2"][:7] = 0
    v["rot_3"][:7] = 0
    # the file's property order is its own: shuffle it
    order = list(ingest.REQUIRED_PROPERTIES)
    rng.shuffle(order)
    c = ingest.cloud_from_ply(cloud_ply(v, order))
    f = {k: a.astype(np.float64) for k, a in v.items()}
    assert np.array_equal(c.positions, np.stack([v["x"], v["y"], v["z"]], 1))
    scales = np.exp(np.stack([f["scale_0"], f["scale_1"], f["scale_2"]], 1)).astype(np.float32)
    o = f["opacity"]
    t = np.exp(np.minimum(o, 0))
    opacity = np.where(o >= 0, 1 / (1 + np.exp(-np.maximum(o, 0))), t / (1 + t)).astype(np.float32)
    colors = (np.stack([f["f_dc_0"], f["f_dc_1"], f["f_dc_2"]], 1) * SH_C0 + 0.5).astype(np.float32)
    q = np.stack([f["rot_1"], f["rot_2"], f["rot_3"], f["rot_0"]], 1)
    length = np.sqrt((q * q).sum(1))
    scale = np.where(length > 0, np.where(q[:, 3] < 0, -1.0, 1.0) / np.where(length > 0, length, 1), 0)
    rot = q * scale[:, None]
    rot[:, 3] = np.where(length > 0, rot[:, 3], 1)
    rot = rot.astype(np.float32)
    # transcendentals may differ in the last bit of a double between libms
    for got, want in ((c.scales, scales), (c.opacity, opacity)):
        assert np.max(np.abs(got - want) / np.abs(want)) <= 1.2e-7
        assert np.mean(got == want) > 0.999
    assert np.array_equal(c.colors, colors)
    assert np.array_equal(c.rotations, rot)
    assert (c.rotations[:, 3] >= 0).all()
    assert c.rotations[:7].tolist() == [[0, 0, 0, 1]] * 7


def test_cloud_requires_the_fourteen_properties():
    v = {n: np.zeros(1) for n in ingest.REQUIRED_PROPERTIES if n != "rot_3"}
    with pytest.raises(ValueError, match="has no property rot_3"):
        ingest.cloud_from_ply(build_ply([(n, "float", a) for n, a in v.items()]))


def test_a_header_count_past_the_file_is_truncation():
    # 2^60 records of 15 bytes: the byte count wraps a 64-bit integer.
    data = build_ply(PROPS).replace(b"element vertex 3", b"element vertex 1152921504606846976")
    with pytest.raises(ValueError, match="truncated"):
        ingest.read_ply(data)


def test_cloud_from_columns_equals_cloud_from_ply():
    rng = np.random.default_rng(9)
    v = {n: rng.normal(size=40).astype(np.float32) for n in ingest.REQUIRED_PROPERTIES}
    a = ingest.cloud_from_ply(cloud_ply(v))
    b = ingest.cloud_from_columns(40, v)
    for name in ("positions", "rotations", "scales", "opacity", "colors"):
        assert np.array_equal(getattr(a, name), getattr(b, name))
