"""Build libwnlqg_hip.so in-tree with hipcc for gfx950.

Every HIP source is compiled with ``-Xclang -target-feature -Xclang -packed-fp32-ops`` so that
no packed-FP32 VALU instructions are emitted.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwnlqg_hip.so")
SOURCES = ["wnlqg_kernels.hip", "wnlqg_eig.hip", "wnlqg_eigbig.hip", "wnlqg_geometry.hip", "wnlqg_shg.hip", "wnlqg_mlp.hip", "wnlqg_engine.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(os.path.dirname(HERE), "include"),
           "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libwnlqg_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
