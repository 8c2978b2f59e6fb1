"""Build libhrnn_hip.so (gfx950) and the host libraries in-tree.

    python -m hrnn_amd.build          # compile everything
    python -m hrnn_amd.build --check  # print what would be built

Outputs (git-ignored, shipped with the tree):
    hrnn_amd/lib/libhrnn_hip.so    HIP kernels + C-ABI (include/hrnn_dev.h)
    hrnn_amd/lib/libhrnn_host.so   hnswlib::HRNN<float> host API (include/hnswlib/hrnn.h)
                                   behind the C entry points of include/hrnn_host.h
    hrnn_amd/bin/hrnn_build, hrnn_search, hrnn_dynamic_test   CLI drivers
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
INC = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_SOURCES = ["api.hip", "csr.hip", "search.hip", "bruteforce.hip", "nndescent.hip",
               "mlp.hip", "beam.hip", "hnsw_build.hip", "insert.hip"]
HOST_SOURCES = ["synth.cpp"]

HIP_FLAGS = [
    # fused multiply-add across statements, except where a function turns it
    # off (the FP64 distance must round every operation like the oracle's)
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=fast-honor-pragmas",
    # packed FP32 instructions stay off (the compiler would otherwise form
    # v_pk_{add,mul,fma}_f32 from ordinary float code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + INC,
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-I" + INC]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + res.stdout + res.stderr)


def _compile_hip(src: str) -> str:
    out = os.path.join(OBJ, src + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", out])
    return out


def _compile_host(src: str) -> str:
    out = os.path.join(OBJ, src + ".o")
    _run(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", out])
    return out


def build_device_lib(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_hip, srcs))
        objs += list(ex.map(_compile_host, HOST_SOURCES))
    out = os.path.join(LIB, "libhrnn_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-fopenmp", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_host_lib() -> str | None:
    """libhrnn_host.so: the HRNN<float> host class behind a C ABI (for Python)."""
    src = os.path.join(CSRC, "host_capi.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(LIB, "libhrnn_host.so")
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-fPIC", "-std=c++17", "-fopenmp", "-march=x86-64-v3", "-shared",
          "-I" + INC, "-I" + os.path.join(ROCM, "include"), src, "-o", tmp, "-L" + LIB,
          "-lhrnn_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_drivers() -> list[str]:
    ddir = os.path.join(PKG, "drivers")
    if not os.path.isdir(ddir):
        return []
    os.makedirs(BIN, exist_ok=True)
    outs = []
    for name in ("hrnn_build", "hrnn_search", "hrnn_dynamic_test", "hrnn_gt"):
        src = os.path.join(ddir, name + ".cpp")
        if not os.path.exists(src):
            continue
        out = os.path.join(BIN, name)
        _run(["g++", "-O3", "-std=c++17", "-fopenmp", "-march=x86-64-v3", "-I" + INC,
              "-I" + ddir, src, "-o", out, "-L" + LIB, "-lhrnn_hip",
              "-Wl,-rpath,$ORIGIN/../lib"])
        outs.append(out)
    return outs


def build(verbose: bool = False) -> dict:
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    res = {"device": build_device_lib()}
    res["host"] = build_host_lib()
    res["drivers"] = build_drivers()
    if verbose:
        for k, v in res.items():
            print(f"built {k}: {v}")
    return res


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    if "--check" in sys.argv:
        print("hip:", HIP_SOURCES, "host:", HOST_SOURCES)
    else:
        build(verbose=True)
