"""Builds libdiorama_hip.so in-tree with hipcc for gfx950 (and the N-API addon when
Node's headers are present). Run as `python -m diorama_amd.build`."""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdiorama_hip.so")
SOURCES = ["scan.hip", "sort.hip", "render.hip", "depth_order.hip", "batch.hip", "loss.hip", "optim.hip", "mlp.hip", "ply.hip", "knn.hip", "metrics.hip", "api.cpp", "io.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950",
    # Packed FP32 instructions stay off: plain scalar/float2 code only.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", os.path.join(ROOT, "include"),
]


def _compile(src):
    obj = os.path.join(OBJ, src + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build_library():
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as pool:
        objs = list(pool.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_addon():
    """The N-API addon (raw C against node_api.h), linked to the library by rpath."""
    header = "/usr/include/node/node_api.h"
    if not os.path.exists(header) or shutil.which("gcc") is None:
        return None
    src = os.path.join(HERE, "node", "addon.c")
    if not os.path.exists(src):
        return None
    out = os.path.join(HERE, "node", "diorama.node")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-DNODE_GYP_MODULE_NAME=diorama",
           "-I/usr/include/node", "-I", os.path.join(ROOT, "include"),
           src, "-o", out + ".tmp", "-L", HERE, "-ldiorama_hip", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"addon build failed:\n{r.stderr}")
    os.replace(out + ".tmp", out)
    return out


def main():
    print(build_library())
    addon = build_addon()
    if addon:
        print(addon)


if __name__ == "__main__":
    sys.exit(main())
