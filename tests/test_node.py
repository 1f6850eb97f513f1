"""The Node boundary (diorama_amd/node): the N-API addon over the C ABI and the JS
classes with the reference's operator names. tests/node_parity.js does the work;
this file applies the bars."""
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ADDON = os.path.join(os.path.dirname(HERE), "diorama_amd", "node", "diorama.node")
SCRIPT = os.path.join(HERE, "node_parity.js")

needs_node = pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(ADDON),
                                reason="node or the built addon is absent")


def run(mode, timeout):
    r = subprocess.run(["node", SCRIPT, mode], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


@needs_node
def test_node_module_loads_without_gpu():
    out = run("cpu", 120)
    assert out["version"].startswith("diorama-mi355x")
    # DEFAULT_SETTINGS, src/trainer.ts:65-82, as float32 values.
    want = {"positionLrSteps": 30000, "positionLr": [0.00016, 0.0000016], "opacityLr": 0.025, "scaleLr": 0.005,
            "rotationLr": 0.001, "colorLr": 0.0025, "shLr": 0.0025 / 20, "shDegreeInterval": 1000,
            "densifyFrom": 500, "densifyUntil": 15000, "densifyInterval": 100, "densifyGradThreshold": 0.0002,
            "percentDense": 0.01, "minOpacity": 0.005, "opacityResetInterval": 3000, "background": [0, 0, 0]}
    assert out["defaults"] == want
    # Both error paths throw JS errors carrying the library's message.
    assert len(out["errors"]) == 3
    assert "ctx is null" in out["errors"][1]
    assert out["errors"][2] == "No .ply header found."
    cm = out["colmap"]
    assert cm["names"] == ["a.jpg", "b.jpg"] and cm["fy"] == 500
    assert cm["position"] == [0, 0, -2]               # -R^T t with R = I, t = (0, 0, 2)
    # photo a: 90 degrees about y, so its forward axis (R's third row) is (-1, 0, 0)
    assert cm["forwardA"] == pytest.approx([-1, 0, 0], abs=1e-12)
    assert cm["positions"] == [0, 0, 0, 1, 2, -1, 2, 4, -2] and cm["colors"] == [0, 0, 0, 10, 20, 30, 20, 40, 60]
    # cameraView at the identity pose: x right, y up (= -down), z back (= -forward), eye at (0, 0, -2)
    assert cm["view"] == pytest.approx([1, 0, 0, 0, 0, -1, 0, 0, 0, 0, -1, 0, 0, 0, -2, 1], abs=1e-6)
    a = out["activated"]
    assert a[3] == 0.5 and a[4] == pytest.approx(2.0, rel=1e-6) and a[5] == 1.0
    assert a[12] == pytest.approx(0.5 + 0.28209479177387814, rel=1e-6)


@needs_node
@pytest.mark.gpu
def test_node_parity_on_gpu():
    out = run("gpu", 110)
    print(json.dumps(out))
    assert out["scan"]["mismatches"] == 0 and out["scan"]["total"] == out["scan"]["expected"]
    assert out["sort"] == {"high0": 0, "high16": 0}
    assert out["loss"]["error"] < 1e-5 and out["loss"]["gradientError"] < 1e-3
    r = out["render"]
    assert r["count"] == 10 and r["hookSaw"] >= 10 and r["hookPass"] and r["textureMatches"]
    assert r["imageError"] < 1e-4
    assert max(r["gradientErrors"].values()) < 1e-3 and not any(math.isnan(v) for v in r["gradientErrors"].values())
    t = out["tiles"]
    assert t["beforePass"] and t["pairs"] == t["rendererPairs"] > 0
    assert t["imageDiff"] == 0 and t["gradientDiff"] < 1e-5
    # every pre-drop-in argument list is refused with a TypeError
    assert out["refused"] == ["renderer.encode(target, camera, bg)", "renderer.encodeBackward(grads)",
                              "loss.encode(rendered, target)", "scan.encode()",
                              "new TileRasterizer(device, splats, count)"]
    assert out["mlp"]["logits"] == [0.5, -1.25, 2.0, 3.75] * 5 and out["mlp"]["weightsBack"] == 3.75
    assert out["batch"]["size"] == 2 and out["batch"]["iteration"] == 1 and math.isfinite(out["batch"]["loss"])
    p = out["ply"]
    assert p["count"] == 64 and p["degree"] == 3 and p["bytes"] == 1527 + 64 * 248  # header of "element vertex 64"
    assert p["paramsError"] < 1e-6 and p["shError"] == 0
    f = np.float32
    assert out["neighbors"] == [float(f(59) / f(3)), float(f(41) / f(3)), float(f(29) / f(3)), float(f(101) / f(3))]
    g = out["initial"]
    assert g[0:4] == [0, 0, 0, float(f(math.log(0.1 / 0.9)))] and g[8:12] == [1, 0, 0, 0]
    assert g[4] == g[5] == g[6] == float(f(0.5 * math.log(float(f(59) / f(3)))))
    assert g[12:15] == [float(f(0.5 / 0.28209479177387814)), float(f(-0.5 / 0.28209479177387814)),
                        float(f((128 / 255 - 0.5) / 0.28209479177387814))]
