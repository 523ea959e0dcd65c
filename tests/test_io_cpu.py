"""Image files, landmarks and the TRE through the C ABI (host code of
libskm.so; no GPU needed).

Files written by the library are re-read by an independent parser here
(VTK legacy: the header lines and big-endian float data; MetaImage: key =
value lines and little-endian raw data), and hand-made files in the formats
ITK writes (VTK legacy BINARY/ASCII STRUCTURED_POINTS, MetaImage .mhd/.mha)
are read back through the library.  The TRE is checked against
oracle/warp_oracle.py.
"""
import os

import numpy as np
import pytest

import warp_oracle as wo
from skmimageregistration_amd import Image, landmark_tre, read_image, read_landmarks, write_image
from skmimageregistration_amd._capi import SKM_ERROR_IO, SKM_ERROR_NOT_SUPPORTED, SkmError


CASES = [
    ("scalar3d", (4, 5, 6), 3, None),
    ("vector3d", (4, 5, 6), 3, (3,)),
    ("tensor3d", (3, 4, 5), 3, (3, 3)),
    ("scalar2d", (7, 9), 2, None),
    ("vector2d", (7, 9), 2, (2,)),
]


def _image(shape, dim, trailing, seed):
    rng = np.random.default_rng(seed)
    data = rng.normal(size=shape + (trailing or ())).astype(np.float32)
    origin = tuple(rng.uniform(-10, 10, size=dim))
    spacing = tuple(rng.uniform(0.3, 2.5, size=dim))
    return Image(data, origin, spacing, dim=dim)


@pytest.mark.parametrize("ext", [".vtk", ".mhd", ".mha"])
@pytest.mark.parametrize("name,shape,dim,trailing", CASES)
def test_round_trip(tmp_path, ext, name, shape, dim, trailing):
    img = _image(shape, dim, trailing, hash(name) % 1000)
    path = str(tmp_path / (name + ext))
    write_image(path, img)
    back = read_image(path)
    assert back.dim == dim
    assert back.origin == img.origin and back.spacing == img.spacing
    assert back.data.shape == img.data.shape
    np.testing.assert_array_equal(back.data, img.data)


def _parse_vtk(path):
    raw = open(path, "rb").read()
    lines, pos = [], 0
    while True:
        e = raw.index(b"\n", pos)
        lines.append(raw[pos:e].decode())
        pos = e + 1
        if lines[-1].startswith(("LOOKUP_TABLE", "VECTORS")):
            break
    return lines, raw[pos:]


def test_vtk_file_layout(tmp_path):
    img = _image((4, 5, 6), 3, (3,), 7)
    path = str(tmp_path / "df.vtk")
    write_image(path, img)
    lines, data = _parse_vtk(path)
    assert lines[0] == "# vtk DataFile Version 3.0"
    assert lines[2] == "BINARY" and lines[3] == "DATASET STRUCTURED_POINTS"
    assert lines[4].split() == ["DIMENSIONS", "6", "5", "4"]
    assert [float(v) for v in lines[5].split()[1:]] == list(img.spacing)
    assert [float(v) for v in lines[6].split()[1:]] == list(img.origin)
    assert lines[7] == "POINT_DATA 120" and lines[8] == "VECTORS vectors float"
    vals = np.frombuffer(data[: 120 * 3 * 4], dtype=">f4").reshape(4, 5, 6, 3)
    np.testing.assert_array_equal(vals, img.data)
    # scalar image: SCALARS + LOOKUP_TABLE, 2D written with nz = 1
    img2 = _image((7, 9), 2, None, 8)
    path2 = str(tmp_path / "s.vtk")
    write_image(path2, img2)
    lines, data = _parse_vtk(path2)
    assert lines[4].split() == ["DIMENSIONS", "9", "7", "1"]
    assert lines[8] == "SCALARS scalars float 1" and lines[9] == "LOOKUP_TABLE default"
    np.testing.assert_array_equal(np.frombuffer(data[: 63 * 4], dtype=">f4").reshape(7, 9), img2.data)


def test_metaimage_file_layout(tmp_path):
    img = _image((3, 4, 5), 3, None, 9)
    path = str(tmp_path / "im.mhd")
    write_image(path, img)
    keys = dict(line.split(" = ", 1) for line in open(path).read().splitlines())
    assert keys["NDims"] == "3" and keys["DimSize"] == "5 4 3" and keys["ElementType"] == "MET_FLOAT"
    assert keys["BinaryDataByteOrderMSB"] == "False" and keys["ElementDataFile"] == "im.raw"
    assert [float(v) for v in keys["Offset"].split()] == list(img.origin)
    raw = np.fromfile(str(tmp_path / "im.raw"), dtype="<f4").reshape(3, 4, 5)
    np.testing.assert_array_equal(raw, img.data)


@pytest.mark.parametrize("vtype,np_type", [("unsigned_char", ">u1"), ("short", ">i2"), ("double", ">f8"),
                                           ("int", ">i4"), ("unsigned_short", ">u2")])
def test_hand_made_binary_vtk(tmp_path, vtype, np_type):
    rng = np.random.default_rng(11)
    vals = rng.integers(0, 120, size=(2, 3, 4)).astype(np_type)
    head = ("# vtk DataFile Version 3.0\nmade by hand\nBINARY\nDATASET STRUCTURED_POINTS\n"
            "DIMENSIONS 4 3 2\nSPACING 0.5 1 2\nORIGIN 1 2 3\nPOINT_DATA 24\n"
            f"SCALARS intensity {vtype}\nLOOKUP_TABLE default\n").encode()
    path = str(tmp_path / "h.vtk")
    open(path, "wb").write(head + vals.tobytes() + b"\n")
    img = read_image(path)
    assert img.dim == 3 and img.origin == (1.0, 2.0, 3.0) and img.spacing == (0.5, 1.0, 2.0)
    np.testing.assert_array_equal(img.data, vals.astype(np.float32))


def test_hand_made_ascii_vtk_vectors(tmp_path):
    vals = np.arange(2 * 2 * 3 * 3, dtype=np.float64).reshape(2, 2, 3, 3) * 0.25 - 1.0
    body = "\n".join(" ".join(f"{v:g}" for v in row) for row in vals.reshape(-1, 3))
    text = ("# vtk DataFile Version 2.0\nfield\nASCII\n\nDATASET STRUCTURED_POINTS\nDIMENSIONS 3 2 2\n"
            "ORIGIN 0 0 0\nSPACING 1 1 1\nPOINT_DATA 12\nVECTORS df double\n" + body + "\n")
    path = str(tmp_path / "a.vtk")
    open(path, "w").write(text)
    img = read_image(path)
    assert img.data.shape == (2, 2, 3, 3)
    np.testing.assert_array_equal(img.data, vals.astype(np.float32))


def test_hand_made_metaimage(tmp_path):
    vals = (np.arange(5 * 6).reshape(5, 6) - 7).astype(">i2")
    (tmp_path / "data.bin").write_bytes(b"\x00" * 16 + vals.tobytes())
    (tmp_path / "m.mhd").write_text("ObjectType = Image\nNDims = 2\nBinaryData = True\nBinaryDataByteOrderMSB = True\n"
                                    "TransformMatrix = 1 0 0 1\nOffset = -1.5 2\nElementSpacing = 0.25 4\n"
                                    "DimSize = 6 5\nHeaderSize = 16\nElementType = MET_SHORT\n"
                                    "ElementDataFile = data.bin\n")
    img = read_image(str(tmp_path / "m.mhd"))
    assert img.dim == 2 and img.origin == (-1.5, 2.0) and img.spacing == (0.25, 4.0)
    np.testing.assert_array_equal(img.data, vals.astype(np.float32))
    # .mha, two channels of doubles after the header
    v2 = np.linspace(-1, 1, 2 * 3 * 4 * 2).reshape(2, 3, 4, 2)
    (tmp_path / "m.mha").write_bytes(b"NDims = 3\nDimSize = 4 3 2\nElementNumberOfChannels = 2\n"
                                     b"ElementType = MET_DOUBLE\nElementDataFile = LOCAL\n" + v2.astype("<f8").tobytes())
    img = read_image(str(tmp_path / "m.mha"))
    np.testing.assert_array_equal(img.data, v2.astype(np.float32))


def test_unsupported_and_broken_files(tmp_path):
    with pytest.raises(SkmError) as e:
        read_image(str(tmp_path / "x.nii"))
    assert e.value.status == SKM_ERROR_NOT_SUPPORTED
    (tmp_path / "c.mha").write_text("NDims = 2\nDimSize = 2 2\nCompressedData = True\nElementType = MET_FLOAT\n"
                                    "ElementDataFile = LOCAL\n")
    with pytest.raises(SkmError) as e:
        read_image(str(tmp_path / "c.mha"))
    assert e.value.status == SKM_ERROR_NOT_SUPPORTED
    (tmp_path / "r.mha").write_text("NDims = 2\nDimSize = 2 2\nTransformMatrix = 0 1 1 0\nElementType = MET_FLOAT\n"
                                    "ElementDataFile = LOCAL\n")
    with pytest.raises(SkmError) as e:
        read_image(str(tmp_path / "r.mha"))
    assert e.value.status == SKM_ERROR_NOT_SUPPORTED
    head = b"# vtk DataFile Version 3.0\nt\nBINARY\nDATASET STRUCTURED_POINTS\nDIMENSIONS 4 4 4\nPOINT_DATA 64\n" \
           b"SCALARS s float 1\nLOOKUP_TABLE default\n"
    (tmp_path / "t.vtk").write_bytes(head + b"\x00" * 100)
    with pytest.raises(SkmError) as e:
        read_image(str(tmp_path / "t.vtk"))
    assert e.value.status == SKM_ERROR_IO
    with pytest.raises(SkmError):
        read_image(str(tmp_path / "missing.vtk"))


def test_landmarks(tmp_path):
    path = str(tmp_path / "lm.vtk")
    open(path, "w").write("# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET POLYDATA\nPOINTS 3 float\n"
                          "1 2 3 4 5 6\n7.5 8 9\n")
    p3 = read_landmarks(path, 3)
    np.testing.assert_array_equal(p3, [[1, 2, 3], [4, 5, 6], [7.5, 8, 9]])
    np.testing.assert_array_equal(read_landmarks(path, 2), [[1, 2], [4, 5], [7.5, 8]])
    open(path, "w").write("POINTS 3 float\n1 2 3\n")
    with pytest.warns(UserWarning):
        assert len(read_landmarks(path, 3)) == 1


def test_tre_matches_oracle():
    rng = np.random.default_rng(5)
    shape = (6, 7, 8)
    origin, spacing = (1.0, -2.0, 0.5), (1.5, 1.0, 0.75)
    df = rng.normal(size=shape + (3,)).astype(np.float32)
    m = rng.uniform(-3, 12, size=(40, 3))
    f = rng.uniform(-3, 12, size=(40, 3))
    got = landmark_tre(Image(df, origin, spacing, dim=3), f, m)
    ref = wo.tre(df, (origin, spacing), f, m)
    assert got == pytest.approx(ref, rel=1e-12)
    assert landmark_tre(Image(df, origin, spacing, dim=3), np.zeros((0, 3)), np.zeros((0, 3))) == -1.0
