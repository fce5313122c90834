"""In-tree build of libhipvr.so with hipcc for gfx950.

The library is compiled as separate translation units in parallel: the host
side and the kernels that do not depend on a texel format (``hipvr.hip``),
and one unit per texel format holding that format's march and batched-march
instances (``inst.hip`` compiled with ``-DHVR_INST_FMT=<format>``); the
objects are linked into one shared library.

Flags: -O3, C++17, no FP contraction (fused multiply-adds are written
explicitly where wanted), packed-FP32 instructions disabled.
"""

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

_PKG = os.path.dirname(os.path.abspath(__file__))
_CSRC = os.path.join(_PKG, "csrc")
HOST_SOURCE = os.path.join(_CSRC, "hipvr.hip")
INST_SOURCE = os.path.join(_CSRC, "inst.hip")
SOURCES = [HOST_SOURCE, INST_SOURCE]
DEPENDS = SOURCES + [os.path.join(_CSRC, n) for n in ("hvr_device.h", "hvr_rules.h", "hvr_launch.h", "hvr_copy.h")] + \
    [os.path.join(os.path.dirname(_PKG), "include", "hipvr.h")]
OUTPUT = os.path.join(_PKG, "libhipvr.so")
OBJ_DIR = os.path.join(_PKG, "_build")
ARCH = os.environ.get("PYTORCH_ROCM_ARCH", "gfx950").split(";")[0] or "gfx950"

# texel formats with fast-path kernel instances (HVR_FAST_FORMATS in hvr_launch.h)
FAST_FORMATS = ("TF_S_F32", "TF_S_F16", "TF_S_U8", "TF_C_F32_F32", "TF_C_F16_F32", "TF_C_U8_F32",
                "TF_C_F32_SN8", "TF_C_F16_SN8", "TF_C_U8_SN8", "TF_C_F16_F16")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def _units():
    """(source, extra flags, object path) of every translation unit."""
    units = [(HOST_SOURCE, [], os.path.join(OBJ_DIR, "hipvr.o"))]
    for f in FAST_FORMATS:
        units.append((INST_SOURCE, [f"-DHVR_INST_FMT={f}"], os.path.join(OBJ_DIR, f"inst_{f}.o")))
    return units


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + res.stdout + res.stderr)


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compile libhipvr.so (always recompiles every unit, in parallel) and
    return its path."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    cc = hipcc()
    cmds = [[cc] + HIPCC_FLAGS + extra + ["-c", src, "-o", obj] for src, extra, obj in _units()]
    jobs = jobs or max(1, min(len(cmds), os.cpu_count() or 1, int(os.environ.get("MAX_JOBS", "64"))))
    # the largest units first, so that the slowest compile starts earliest
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c) for c in cmds[1:] + cmds[:1]]:
            f.result()
    tmp = OUTPUT + ".tmp"
    link = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + [obj for _, _, obj in _units()]
    _run(link)
    if verbose:
        for c in cmds + [link]:
            print(" ".join(c))
    os.replace(tmp, OUTPUT)
    return OUTPUT


def is_stale() -> bool:
    if not os.path.exists(OUTPUT):
        return True
    t = os.path.getmtime(OUTPUT)
    return any(os.path.getmtime(p) > t for p in DEPENDS)
