"""In-tree build of the HIP library (libbbscite.so), the CLI (bbscite) and
the quality harness (bbscite_quality).

Everything is compiled for gfx950 with hipcc; no JIT cache is used, so the
built files travel with the source tree.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libbbscite.so")
CLI = os.path.join(HERE, "bbscite")
QUALITY = os.path.join(HERE, "bbscite_quality")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off keeps every float operation separately rounded, as the
# reference's host build evaluates them; packed FP32 is disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA results in VGPRs: the bulk scorer reads every count on the
    # vector unit, so accumulators in AGPRs would cost a copy each
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall",
]
LIB_SOURCES = ["bbs_kernels.hip", "bbscite_capi.cpp"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        l for l in proc.stdout.splitlines() if "not a recognized feature" not in l
    )
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_lib():
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


def _build_tool(src_name, out):
    src = os.path.join(HERE, "cli", src_name)
    tmp = out + ".tmp"
    _run([
        "g++", "-O2", "-std=c++17", "-Wall", "-o", tmp, src,
        f"-I{os.path.join(ROOT, 'include')}", f"-L{HERE}", "-lbbscite",
        f"-Wl,-rpath,$ORIGIN",
    ])
    os.replace(tmp, out)
    return out


def build_cli():
    _build_tool("bbscite_quality.cpp", QUALITY)
    return _build_tool("bbscite_main.cpp", CLI)


def build_all():
    build_lib()
    build_cli()


if __name__ == "__main__":
    build_all()
