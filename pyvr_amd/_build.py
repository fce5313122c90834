"""In-tree build of libhipvr.so with hipcc for gfx950.

Flags: -O3, C++17, no FP contraction (fused multiply-adds are written
explicitly where wanted), packed-FP32 instructions disabled.
"""

import os
import shutil
import subprocess

_PKG = os.path.dirname(os.path.abspath(__file__))
SOURCES = [os.path.join(_PKG, "csrc", "hipvr.hip")]
DEPENDS = SOURCES + [os.path.join(_PKG, "csrc", "hvr_device.h"), os.path.join(_PKG, "csrc", "hvr_rules.h"),
                     os.path.join(os.path.dirname(_PKG), "include", "hipvr.h")]
OUTPUT = os.path.join(_PKG, "libhipvr.so")
ARCH = os.environ.get("PYTORCH_ROCM_ARCH", "gfx950").split(";")[0] or "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def build(verbose: bool = False) -> str:
    """Compile libhipvr.so (always recompiles) and return its path."""
    tmp = OUTPUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + SOURCES
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose:
        print(" ".join(cmd))
    os.replace(tmp, OUTPUT)
    return OUTPUT


def is_stale() -> bool:
    if not os.path.exists(OUTPUT):
        return True
    t = os.path.getmtime(OUTPUT)
    return any(os.path.getmtime(p) > t for p in DEPENDS)
