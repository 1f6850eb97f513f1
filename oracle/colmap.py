"""CPU restatement of the COLMAP reader and the training-camera conventions, for
the tests only.

TEST INFRASTRUCTURE: only the tests load this module, as the checker. It follows
parseColmap (src/colmap.ts:18-129), cameraView and verticalFov
(src/sceneCameras.ts:49-67) and perspective (src/mat4.ts:39-48) in plain Python.
It is pinned by tests/test_colmap.py: hand-assembled binary files with known
poses, the pose invariants (R orthonormal, R c + t = 0), and a pinhole
projection worked by hand through the view and projection matrices.
"""
import math
import struct


class ColmapError(ValueError):
    pass


class _Reader:
    def __init__(self, data, what):
        self.data, self.at, self.what = bytes(data), 0, what

    def take(self, fmt):
        size = struct.calcsize(fmt)
        if self.at + size > len(self.data):
            raise ColmapError(f"The COLMAP file {self.what} is shorter than its contents say.")
        out = struct.unpack_from(fmt, self.data, self.at)
        self.at += size
        return out if len(out) > 1 else out[0]

    def skip(self, n):
        if self.at + n > len(self.data):
            raise ColmapError(f"The COLMAP file {self.what} is shorter than its contents say.")
        self.at += n

    def text(self):
        end = self.data.find(b"\0", self.at)
        if end < 0:
            raise ColmapError(f"The COLMAP file {self.what} is shorter than its contents say.")
        out = self.data[self.at:end].decode()
        self.at = end + 1
        return out


def rotation(q):
    """src/colmap.ts:123-129: R of the (w, x, y, z) quaternion."""
    w, x, y, z = q
    return [[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
            [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
            [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]]


def parse(cameras_bin, images_bin, points_bin):
    """src/colmap.ts:18-121: {"cameras": [...sorted by name], "positions": [[x, y, z]], "colors": [[r, g, b]]}."""
    rc = _Reader(cameras_bin, "cameras.bin")
    intrinsics = {}
    for _ in range(rc.take("<Q")):
        cid, model, width, height = rc.take("<iiQQ")
        if model == 0:
            f = rc.take("<d")
            rc.skip(16)
            intrinsics[cid] = (width, height, f, f)
        elif model == 1:
            fx, fy = rc.take("<dd")
            rc.skip(16)
            intrinsics[cid] = (width, height, fx, fy)
        else:
            raise ColmapError(f"Camera {cid} uses COLMAP model {model}; only undistorted pinhole cameras are supported.")
    ri = _Reader(images_bin, "images.bin")
    cameras = []
    for _ in range(ri.take("<Q")):
        ri.take("<i")
        q = ri.take("<dddd")
        t = ri.take("<ddd")
        cid = ri.take("<i")
        name = ri.text()
        ri.skip(ri.take("<Q") * 24)
        if cid not in intrinsics:
            raise ColmapError(f"Photo {name} uses camera {cid}, which isn't in cameras.bin.")
        width, height, fx, fy = intrinsics[cid]
        r = rotation(q)
        position = [-(r[0][c] * t[0] + r[1][c] * t[1] + r[2][c] * t[2]) for c in range(3)]
        cameras.append({"name": name, "width": width, "height": height, "fx": fx, "fy": fy, "position": position,
                        "right": r[0], "down": r[1], "forward": r[2]})
    cameras.sort(key=lambda c: c["name"].encode())
    rp = _Reader(points_bin, "points3D.bin")
    positions, colors = [], []
    for _ in range(rp.take("<Q")):
        rp.take("<Q")
        positions.append(list(rp.take("<ddd")))
        colors.append(list(rp.take("<BBB")))
        rp.take("<d")
        rp.skip(rp.take("<Q") * 8)
    return {"cameras": cameras, "positions": positions, "colors": colors}


def camera_view(c):
    """src/sceneCameras.ts:54-67, column-major."""
    up = [-v for v in c["down"]]
    back = [-v for v in c["forward"]]
    dot = lambda a: sum(a[k] * c["position"][k] for k in range(3))
    right = c["right"]
    return [right[0], up[0], back[0], 0, right[1], up[1], back[1], 0, right[2], up[2], back[2], 0,
            -dot(right), -dot(up), -dot(back), 1]


def perspective(fov_y, aspect, near, far):
    """src/mat4.ts:39-48 (clip depth 0..1)."""
    f = 1 / math.tan(fov_y / 2)
    depth = 1 / (near - far)
    return [f / aspect, 0, 0, 0, 0, f, 0, 0, 0, 0, far * depth, -1, 0, 0, near * far * depth, 0]


def training_camera(c, width, height):
    """src/loadDataset.ts:80-86."""
    fov = 2 * math.atan(c["height"] / (2 * c["fy"]))
    return {"view": camera_view(c), "proj": perspective(fov, width / height, 0.01, 1000), "viewport": [width, height]}
