"""Autotuner for batched rendering: picks the brick size (8^3 or 16^3 interior
texels), the ray tile of a wavefront (8x8 or 16x4 pixels) and the kernel
variant (waves per block 2 / 4 / 8, transparent-brick skipping on or off,
normals gated on non-zero alpha or not; ``batch.KERNEL_VARIANTS``) by timing
the persistent batch launch on the workload itself.  Every candidate gives
the same images; only the time differs.

Each candidate renders the requests (all, or an evenly spaced sample of
``max_requests``) in one batch; the figure of merit is the median device time
of ``runs`` launches after a warm-up launch (which also captures the graph).
Results can be kept in a JSON cache keyed by the workload's signature
(formats, volume shapes, image sizes, request count), so a server tunes once.
"""

import json
import os
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from .batch import KERNEL_VARIANTS, TILE_SHAPES, BatchRenderer

BRICK_LOG2 = (3, 4)


def signature(volumes: Sequence, requests: Sequence, scalar_format: str, normal_format: str) -> str:
    """Workload key: formats, the volumes' shapes (with/without normals) and
    the requests' (volume, width, height) multiset."""
    vols = [f"{'x'.join(map(str, _shape(v)))}{'n' if v.has_normals else ''}" for v in volumes]
    reqs = sorted(f"{q.volume}:{q.width}x{q.height}" for q in requests)
    return f"{scalar_format}+{normal_format}|{','.join(vols)}|{len(requests)}|{';'.join(reqs)}"


def _shape(volume):
    """Shape of a ``Volume`` (its data) or of a device-generated volume."""
    data = getattr(volume, "data", None)
    return tuple(data.shape) if data is not None else tuple(volume.shape)


def _sample(requests: Sequence, max_requests: int) -> List:
    if len(requests) <= max_requests:
        return list(requests)
    idx = np.linspace(0, len(requests) - 1, max_requests).round().astype(int)
    return [requests[i] for i in sorted(set(idx.tolist()))]


def tune(volumes: Sequence, requests: Sequence, scalar_format: str = "f16", normal_format: str = "sn8x3",
         runs: int = 5, max_requests: int = 64, device: int = 0, cache: Optional[str] = None,
         bricks: Iterable[int] = BRICK_LOG2, tiles: Iterable[str] = tuple(TILE_SHAPES),
         kernels: Iterable[str] = tuple(KERNEL_VARIANTS), margin: float = 0.05) -> Dict:
    """Time every (brick_log2, tile_shape, kernel) candidate; return
    ``{"brick_log2", "tile_shape", "kernel", "ms", "table": [(brick_log2,
    tile_shape, kernel, ms), ...], "signature", "cached"}``: the candidate
    with the smallest median time, unless it is within ``margin`` of the
    first candidate (the defaults: 8^3 bricks, 8x8 tiles, default kernel),
    which is then kept -- launch times vary by a few percent from run to
    run, and a pick that flips on noise helps nobody.  ``requests`` are
    objects with ``volume`` (index into ``volumes``), ``camera``, ``ctf``,
    ``otf``, ``width``, ``height``, ``config``, ``light`` and ``tau``."""
    key = signature(volumes, requests, scalar_format, normal_format)
    table = {}
    if cache and os.path.exists(cache):
        try:
            with open(cache) as f:
                table = json.load(f)
        except (OSError, ValueError):
            table = {}
        if key in table:
            out = dict(table[key])
            out["table"] = [tuple(t) for t in out["table"]]
            out["cached"] = True
            return out
    sample = _sample(requests, max_requests)
    timings: List[Tuple[int, str, str, float]] = []
    for lb in bricks:
        b = BatchRenderer(device=device, scalar_format=scalar_format, normal_format=normal_format,
                          brick_log2=lb, pipeline=1)
        try:
            for q in sample:
                b.add(volumes[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
            for tile in tiles:
                b.set_tile_shape(tile)
                for kernel in kernels:
                    b.set_kernel(kernel)
                    b.run()
                    ms = []
                    for _ in range(runs):
                        b.run()
                        ms.append(b.last_ms())
                    timings.append((lb, tile, kernel, float(np.median(ms))))
        finally:
            b.close()
    lb, tile, kernel, best = min(timings, key=lambda t: t[3])
    if best >= (1.0 - margin) * timings[0][3]:
        lb, tile, kernel, best = timings[0]
    out = {"brick_log2": lb, "tile_shape": tile, "kernel": kernel, "ms": best, "table": timings,
           "signature": key, "cached": False}
    if cache:
        table[key] = {k: v for k, v in out.items() if k != "cached"}
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1)
        os.replace(tmp, cache)
    return out
