"""CPU restatement of the 3DGS .ply writer and loader, for the tests only.

TEST INFRASTRUCTURE: only the tests load this module, as the checker. It follows
writeGaussianPly (src/plyWriter.ts:8-43) and parseGaussianPly with its header
search (src/plyLoader.ts:21-110) in numpy: the writer's byte layout exactly, the
loader's activations in float64 rounded once to float32 (as Math.exp results
stored into a Float32Array are). It is pinned by tests/test_oracle_ply.py: a
hand-assembled file in another column order and degree, the written column
positions of known coefficients, the header text, every loader error, and the
reference's own round-trip check (src/checks/training.ts:122-136).
"""
import re

import numpy as np

SH_C0 = 0.28209479177387814
SH_REST = 45

# src/plyWriter.ts:8-12
PROPERTIES = (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
              + [f"f_rest_{k}" for k in range(SH_REST)]
              + ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])


class PlyError(ValueError):
    pass


def header(count):
    """src/plyWriter.ts:16-23."""
    lines = ["ply", "format binary_little_endian 1.0", f"element vertex {count}"]
    lines += [f"property float {name}" for name in PROPERTIES] + ["end_header", ""]
    return "\n".join(lines).encode()


def write(params, sh):
    """src/plyWriter.ts:14-43: raw params (n x 16) and SH (n x 45, coefficient-major)
    to the file's bytes."""
    p = np.asarray(params, np.float32).reshape(-1, 16)
    s = np.asarray(sh, np.float32).reshape(-1, SH_REST)
    n = len(p)
    body = np.zeros((n, len(PROPERTIES)), np.float32)
    body[:, 0:3] = p[:, 0:3]                 # positions; normals stay 0
    body[:, 6:9] = p[:, 12:15]               # f_dc
    for channel in range(3):                 # the file is channel-major (plyWriter.ts:32-36)
        for c in range(SH_REST // 3):
            body[:, 9 + channel * 15 + c] = s[:, c * 3 + channel]
    body[:, 9 + SH_REST] = p[:, 3]           # opacity logit
    body[:, 10 + SH_REST:13 + SH_REST] = p[:, 4:7]  # log scales
    body[:, 13 + SH_REST:17 + SH_REST] = p[:, 8:12]  # quaternion
    return header(n) + body.astype("<f4").tobytes()


def _find(data, needle, limit):
    """src/plyLoader.ts:100-110: the needle must lie wholly inside the first `limit` bytes."""
    return data.find(needle, 0, min(len(data), limit))


def parse(data):
    """src/plyLoader.ts:21-98: {"count", "gaussians" (n x 16 float32, activated),
    "sh_rest" (n x 45 float32), "sh_degree"}; PlyError with the loader's messages."""
    data = bytes(data)
    marker = b"end_header\n"
    at = _find(data, marker, 64 * 1024)
    if at < 0:
        raise PlyError("No .ply header found.")
    header_end = at + len(marker)
    text = data[:header_end].decode("utf-8", errors="replace")
    if not text.startswith("ply\n"):
        raise PlyError("Not a .ply file.")
    if not any(line == "format binary_little_endian 1.0" for line in re.split(r"[\r\n]", text)):
        raise PlyError("Only binary little-endian .ply files are supported.")
    elements, properties, count = [], [], None
    for line in text.split("\n"):
        words = line.split()
        if not words:
            continue
        if words[0] == "element":
            elements.append(words[1] if len(words) > 1 else "")
            if elements[-1] == "vertex":
                n = words[2] if len(words) > 2 else ""
                if not n.isdigit():
                    raise PlyError(f'The .ply vertex count "{n}" is not a whole number.')
                count = int(n)
        elif words[0] == "property" and elements and elements[-1] == "vertex":
            kind = words[1] if len(words) > 1 else ""
            name = words[2] if len(words) > 2 else ""
            if kind != "float":
                raise PlyError(f"Property {name} is {kind}; only float properties are supported.")
            properties.append(name)
    if not elements or elements[0] != "vertex":
        raise PlyError("Expected the .ply to start with its vertex element.")
    if count is None or count > 0xFFFFFFFF:
        raise PlyError("The .ply vertex count is out of range.")

    def column(name):
        if name not in properties:
            raise PlyError(f'The .ply has no "{name}" property. Is it a 3D Gaussian Splatting file?')
        return properties.index(name)

    position = [column(n) for n in ("x", "y", "z")]
    dc = [column(f"f_dc_{k}") for k in range(3)]
    opacity = column("opacity")
    scale = [column(f"scale_{k}") for k in range(3)]
    rotation = [column(f"rot_{k}") for k in range(4)]
    rest_count = sum(1 for n in properties if n.startswith("f_rest_"))
    per_channel = rest_count / 3
    degree = round((per_channel + 1) ** 0.5) - 1
    if (degree + 1) ** 2 - 1 != per_channel or degree > 3:
        raise PlyError(f"Unexpected number of spherical-harmonic coefficients ({rest_count}).")
    per_channel = int(per_channel)
    rest = [column(f"f_rest_{k}") for k in range(rest_count)]
    stride = len(properties)
    if header_end + count * stride * 4 > len(data):
        raise PlyError("The .ply file is shorter than its header says.")
    rows = np.frombuffer(data, dtype="<f4", count=count * stride, offset=header_end).reshape(count, stride)
    d = rows.astype(np.float64)
    g = np.zeros((count, 16), np.float32)
    g[:, 0:3] = rows[:, position]
    with np.errstate(over="ignore"):
        g[:, 3] = (1.0 / (1.0 + np.exp(-d[:, opacity]))).astype(np.float32)
        g[:, 4:7] = np.exp(d[:, scale]).astype(np.float32)
    g[:, 8:12] = rows[:, rotation]
    g[:, 12:15] = (0.5 + SH_C0 * d[:, dc]).astype(np.float32)
    sh = np.zeros((count, SH_REST), np.float32)
    for c in range(per_channel):
        for channel in range(3):
            sh[:, c * 3 + channel] = rows[:, rest[channel * per_channel + c]]
    return {"count": count, "gaussians": g, "sh_rest": sh, "sh_degree": degree}
