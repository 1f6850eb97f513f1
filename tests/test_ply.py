"""The .ply format without a GPU: the oracle (oracle/ply.py) pinned by a hand-assembled
file, the written positions of known values, the header's exact text and the
loader's errors; and the C ABI's header handling (sizes and every loader error
come back before the device is touched)."""
import ctypes
import math
import struct

import numpy as np
import pytest

import diorama_amd
from oracle import oracle as oc
from oracle import ply


def test_header_text_and_size():
    h = ply.header(1)
    assert h.startswith(b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float x\n")
    assert h.endswith(b"property float rot_2\nproperty float rot_3\nend_header\n")
    # 66 lines: ply, format, element, 62 properties, end_header. The 62 property
    # lines are "property float " (15 bytes) + 470 bytes of names.
    assert len(h) == 3 + 31 + 16 + 62 * 15 + 470 + 10 + 66 == 1526
    assert len(ply.PROPERTIES) == 62 and len(set(ply.PROPERTIES)) == 62


def test_written_columns_of_known_values():
    params = np.arange(16, dtype=np.float32)[None] + 100
    sh = np.arange(45, dtype=np.float32)[None] + 1000   # sh[c * 3 + channel], coefficient-major
    data = ply.write(params, sh)
    row = np.frombuffer(data, "<f4", offset=len(ply.header(1)))
    assert row.size == 62
    assert list(row[0:9]) == [100, 101, 102, 0, 0, 0, 112, 113, 114]
    # f_rest_0 = red c1, f_rest_1 = red c2, f_rest_14 = red c15, f_rest_15 = green c1, f_rest_44 = blue c15
    assert row[9] == 1000 and row[10] == 1003 and row[23] == 1042 and row[24] == 1001 and row[53] == 1044
    assert list(row[54:62]) == [103, 104, 105, 106, 108, 109, 110, 111]


def hand_file(extra_header=b"", rows=None, truncate=0):
    """Two vertices in another column order, degree 1, an unused property, a face
    element after the vertices and a header that ends off a 4-byte boundary."""
    names = (["red", "rot_3", "rot_0", "rot_2", "rot_1", "opacity", "scale_2", "scale_1", "scale_0"]
             + [f"f_rest_{k}" for k in (4, 0, 8, 1, 5, 2, 6, 3, 7)] + ["z", "y", "x", "f_dc_2", "f_dc_0", "f_dc_1"])
    text = b"ply\nformat binary_little_endian 1.0\ncomment by hand\nelement vertex 2\n"
    text += b"".join(b"property float " + n.encode() + b"\n" for n in names)
    text += extra_header + b"element face 0\nproperty list uchar int vertex_indices\nend_header\n"
    if rows is None:
        v0 = {"red": 7, "rot_0": 1, "rot_1": 0.5, "rot_2": -0.5, "rot_3": 2, "opacity": 0, "scale_0": 0,
              "scale_1": math.log(2), "scale_2": -math.log(4), "x": 1, "y": 2, "z": 3, "f_dc_0": 0, "f_dc_1": 1,
              "f_dc_2": -1}
        v0.update({f"f_rest_{k}": 10 * k for k in range(9)})
        v1 = dict(v0, opacity=-1000.0, x=-4, scale_0=-2)
        rows = [v0, v1]
    body = b"".join(struct.pack("<" + "f" * len(names), *[r[n] for n in names]) for r in rows)
    data = text + body
    return data[:len(data) - truncate] if truncate else data


def test_oracle_reads_a_hand_assembled_file():
    data = hand_file()
    assert (data.index(b"end_header\n") + 11) % 4 != 0   # the body does not start 4-byte aligned
    out = ply.parse(data)
    assert out["count"] == 2 and out["sh_degree"] == 1
    g, sh = out["gaussians"], out["sh_rest"]
    f = np.float32
    want0 = [1, 2, 3, 0.5, 1, 2, 0.25, 0, 1, 0.5, -0.5, 2, 0.5, f(0.5 + 0.28209479177387814),
             f(0.5 - 0.28209479177387814), 0]
    assert np.allclose(g[0], want0, rtol=0, atol=2e-7), g[0]
    assert g[1, 3] == 0.0 and g[1, 0] == -4 and g[1, 4] == f(math.exp(-2))
    # coefficient-major c1 rgb, c2 rgb, c3 rgb from the file's f_rest_{channel * 3 + c} = 10 (channel * 3 + c)
    assert list(sh[0, :9]) == [0, 30, 60, 10, 40, 70, 20, 50, 80] and not sh[0, 9:].any()


def test_round_trip_as_the_reference_checks_it():
    """src/checks/training.ts:122-136: activate(raw) to 1e-6, the SH exactly, degree 3."""
    rng = np.random.default_rng(3)
    n = 500
    params = rng.normal(0, 1, (n, 16)).astype(np.float32)
    params[:, 4:7] = rng.normal(-4, 0.5, (n, 3))
    params[:, [7, 15]] = rng.integers(0, 3000, (n, 2))   # the trainer's stamps in the pads are not written
    sh = rng.normal(0, 0.1, (n, 45)).astype(np.float32)
    out = ply.parse(ply.write(params, sh))
    want = oc.activate(params.astype(np.float64)).reshape(n, 16)
    assert out["count"] == n and out["sh_degree"] == 3
    assert np.abs(out["gaussians"] - want).max() < 1e-6
    assert np.array_equal(out["sh_rest"], sh)


ERRORS = [
    (b"hello", "No .ply header found."),
    (b"ply\ncomment " + b"x" * 70000 + b"\nend_header\n", "No .ply header found."),
    (b"plx\nformat binary_little_endian 1.0\nend_header\n", "Not a .ply file."),
    (b"ply\nformat ascii 1.0\nelement vertex 0\nend_header\n", "Only binary little-endian .ply files are supported."),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty double x\nend_header\n",
     "Property x is double; only float properties are supported."),
    (b"ply\nformat binary_little_endian 1.0\nelement face 0\nelement vertex 0\nproperty float x\nend_header\n",
     "Expected the .ply to start with its vertex element."),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 0\nproperty float x\nproperty float y\n"
     b"property float z\nend_header\n", 'The .ply has no "f_dc_0" property. Is it a 3D Gaussian Splatting file?'),
    (hand_file(extra_header=b"property float f_rest_9\n"), "Unexpected number of spherical-harmonic coefficients (10)."),
    (hand_file(truncate=1), "The .ply file is shorter than its header says."),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex many\nend_header\n",
     'The .ply vertex count "many" is not a whole number.'),
]


@pytest.mark.parametrize("data,message", ERRORS)
def test_loader_errors(data, message):
    with pytest.raises(ply.PlyError) as e:
        ply.parse(data)
    assert str(e.value) == message


@pytest.mark.parametrize("data,message", ERRORS)
def test_abi_loader_errors_before_the_device(data, message):
    L = diorama_amd.lib()
    g, s, n, d = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_uint32(), ctypes.c_uint32()
    status = L.dio_ply_parse(None, data, len(data), ctypes.byref(g), ctypes.byref(s), ctypes.byref(n), ctypes.byref(d))
    assert status == -1 and L.dio_last_error().decode() == message


def test_abi_checks_a_good_header_then_needs_a_context():
    L = diorama_amd.lib()
    data = hand_file()
    g, s, n, d = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_uint32(), ctypes.c_uint32()
    status = L.dio_ply_parse(None, data, len(data), ctypes.byref(g), ctypes.byref(s), ctypes.byref(n), ctypes.byref(d))
    assert status == -1 and L.dio_last_error().decode() == "ctx is null"


def test_abi_file_size_without_a_device():
    L = diorama_amd.lib()
    for count in (0, 1, 12345):
        length = ctypes.c_size_t()
        assert L.dio_ply_write(None, None, count, None, 0, ctypes.byref(length)) == 0
        assert length.value == len(ply.header(count)) + 248 * count
