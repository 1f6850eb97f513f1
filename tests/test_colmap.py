"""COLMAP reconstructions without a GPU: hand-assembled cameras.bin, images.bin and
points3D.bin with known poses read through the C ABI and the oracle
(oracle/colmap.py), the pose invariants, the reader's errors, the training camera
checked by a pinhole projection worked by hand, and loadDataset's split."""
import math
import struct

import numpy as np
import pytest

import diorama_amd
from oracle import colmap


def cameras_bin(cams):
    out = struct.pack("<Q", len(cams))
    for cid, model, w, h, params in cams:
        out += struct.pack("<iiQQ", cid, model, w, h) + struct.pack("<" + "d" * len(params), *params)
    return out


def images_bin(images):
    out = struct.pack("<Q", len(images))
    for iid, q, t, cid, name, points2d in images:
        out += struct.pack("<i", iid) + struct.pack("<dddd", *q) + struct.pack("<ddd", *t) + struct.pack("<i", cid)
        out += name.encode() + b"\0" + struct.pack("<Q", points2d)
        out += b"".join(struct.pack("<ddq", 1.5 * k, 2.5 * k, k) for k in range(points2d))
    return out


def points_bin(points):
    out = struct.pack("<Q", len(points))
    for pid, xyz, rgb, track in points:
        out += struct.pack("<Q", pid) + struct.pack("<ddd", *xyz) + struct.pack("<BBB", *rgb) + struct.pack("<d", 0.5)
        out += struct.pack("<Q", track) + b"".join(struct.pack("<ii", k, k + 1) for k in range(track))
    return out


def unit(q):
    n = math.sqrt(sum(v * v for v in q))
    return [v / n for v in q]


def scene_files(seed=0, photos=10, points=25):
    rng = np.random.default_rng(seed)
    cams = [(1, 0, 640, 480, [500.0, 320.0, 240.0]), (7, 1, 800, 600, [600.0, 620.0, 400.0, 300.0])]
    names = [f"img_{k:02d}.jpg" for k in rng.permutation(photos)]
    images = [(k + 1, unit(rng.normal(size=4)), list(rng.normal(size=3) * 3), 1 if k % 2 else 7, names[k], k % 3)
              for k in range(photos)]
    pts = [(k, list(rng.normal(size=3)), list(rng.integers(0, 256, 3)), k % 4) for k in range(points)]
    return cameras_bin(cams), images_bin(images), points_bin(pts), images, pts


def test_reads_a_hand_assembled_reconstruction():
    cb, ib, pb, images, pts = scene_files()
    got = diorama_amd.parse_colmap(cb, ib, pb)
    want = colmap.parse(cb, ib, pb)
    assert [c["name"] for c in got["cameras"]] == sorted(im[4] for im in images)
    assert [c["name"] for c in got["cameras"]] == [c["name"] for c in want["cameras"]]
    for g, w in zip(got["cameras"], want["cameras"]):
        assert (g["width"], g["height"], g["fx"], g["fy"]) == (w["width"], w["height"], w["fx"], w["fy"])
        for key in ("position", "right", "down", "forward"):
            assert np.allclose(g[key], w[key], rtol=0, atol=1e-12)
    assert np.array_equal(got["positions"], np.array(want["positions"], np.float32))
    assert np.array_equal(got["colors"], np.array(want["colors"], np.uint8))
    # SIMPLE_PINHOLE camera 1: fx = fy = 500 at 640 x 480; PINHOLE camera 7: 600, 620 at 800 x 600
    by_name = {im[4]: im for im in images}
    for c in got["cameras"]:
        cid = by_name[c["name"]][3]
        assert (c["width"], c["height"], c["fx"], c["fy"]) == ((640, 480, 500.0, 500.0) if cid == 1 else
                                                               (800, 600, 600.0, 620.0))


def test_pose_invariants():
    cb, ib, pb, images, _ = scene_files(seed=3)
    by_name = {im[4]: im for im in images}
    for c in diorama_amd.parse_colmap(cb, ib, pb)["cameras"]:
        r = np.array([c["right"], c["down"], c["forward"]])
        assert np.allclose(r @ r.T, np.eye(3), atol=1e-12) and abs(np.linalg.det(r) - 1) < 1e-12
        t = np.array(by_name[c["name"]][2])
        # the centre is where the world-to-camera map x_c = R x_w + t sends to zero
        assert np.allclose(r @ np.array(c["position"]) + t, 0, atol=1e-12)


def test_training_camera_projects_like_a_pinhole():
    cb, ib, pb, _, _ = scene_files(seed=5)
    cams = diorama_amd.parse_colmap(cb, ib, pb)["cameras"]
    for c in cams[:4]:
        w, h = c["width"], c["height"]
        got = diorama_amd.training_camera_data(c, (w, h))
        want = colmap.training_camera(c, w, h)
        assert np.allclose(got["view"], np.float32(want["view"]), rtol=0, atol=1e-6)
        assert np.allclose(got["proj"], np.float32(want["proj"]), rtol=1e-6, atol=0)
        # a point 4 units ahead, 0.3 right and 0.2 down in the camera's own frame lands on
        # pixel (w/2 + fy x/z, h/2 + fy y/z): the renderer keeps the centre as principal
        # point and takes the horizontal field from the aspect ratio
        p = np.array(c["position"]) + 4 * np.array(c["forward"]) + 0.3 * np.array(c["right"]) + 0.2 * np.array(c["down"])
        view = np.array(got["view"], np.float64).reshape(4, 4).T
        proj = np.array(got["proj"], np.float64).reshape(4, 4).T
        clip = proj @ view @ np.append(p, 1.0)
        ndc = clip[:3] / clip[3]
        px, py = (ndc[0] + 1) / 2 * w, (1 - ndc[1]) / 2 * h
        assert abs(px - (w / 2 + c["fy"] * 0.3 / 4)) < 1e-3 and abs(py - (h / 2 + c["fy"] * 0.2 / 4)) < 1e-3
        assert 0 < ndc[2] < 1


ERRORS = [
    (cameras_bin([(3, 2, 10, 10, [1.0, 2.0, 3.0, 4.0])]), "Camera 3 uses COLMAP model 2; only undistorted pinhole cameras are supported."),
    ("missing", "Photo a.jpg uses camera 9, which isn't in cameras.bin."),
    ("short", "The COLMAP file points3D.bin is shorter than its contents say."),
]


@pytest.mark.parametrize("case,message", ERRORS)
def test_reader_errors(case, message):
    cb, ib, pb, _, _ = scene_files(seed=1)
    if isinstance(case, bytes):
        cb = case
    elif case == "missing":
        ib = images_bin([(1, [1, 0, 0, 0], [0, 0, 0], 9, "a.jpg", 0)])
    else:
        pb = pb[:-5]
    with pytest.raises(diorama_amd.DioramaError, match="^" + message.replace(".", r"\.").replace("(", r"\(") + "$"):
        diorama_amd.parse_colmap(cb, ib, pb)
    with pytest.raises(colmap.ColmapError) as e:
        colmap.parse(cb, ib, pb)
    assert str(e.value) == message


def test_counts_without_outputs_and_the_split():
    cb, ib, pb, _, _ = scene_files(seed=2, photos=17, points=9)
    got = diorama_amd.parse_colmap(cb, ib, pb)
    assert len(got["cameras"]) == 17 and got["positions"].shape == (9, 3)
    split = diorama_amd.split_dataset(got["cameras"])
    names = [c["name"] for c in got["cameras"]]
    assert [c["name"] for c in split["test"]] == [names[0], names[8], names[16]]
    assert len(split["train"]) == 14
    centers = np.array([c["position"] for c in split["train"]])
    assert split["extent"] == pytest.approx(1.1 * np.linalg.norm(centers - centers.mean(0), axis=1).max())
