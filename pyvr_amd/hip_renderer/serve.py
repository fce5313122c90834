"""Serving loop: continuous batching of render requests over volumes kept
resident in HBM, admitted by a device-memory budget.

The reference has no serving layer (one ``VolumeRenderer.render()`` per
frame, ``pyvr/moderngl_renderer/renderer.py:209-219``).  ``RenderServer``
turns a stream of requests into batches for ``BatchRenderer``:

* requests queue in arrival order; whenever a batch slot is free, the
  queued requests that fit (at most ``max_batch_requests`` and
  ``max_batch_pixels``; a request whose volume cannot be made resident yet
  waits while later ones go ahead) are launched together, so requests that
  arrive while others are on the GPU join the next batch, and up to
  ``pipeline`` batches run at once;
* a volume is uploaded (or generated) on first use and stays resident;
  when the budget would be exceeded, resident volumes that no queued-for-
  launch or in-flight batch uses are evicted, least recently used first;
  the budget also reserves what the batches in flight hold (output
  pools, per-request brick flags for the largest registered volume,
  staging), as those pools grow;
* with ``max_batch_samples`` a batch also stops at that many estimated
  samples (``estimated_samples``), so that a burst of large requests is
  split into batches of bounded duration;
* a request's latency runs from its arrival to its pixels being in host
  memory; ``poll()`` returns the requests that completed, with ``rgba``
  filled (the same bytes ``VolumeRenderer.render()`` gives for it) unless
  ``keep_results`` is off (load measurements that do not consume pixels);
* a request the batched path rejects (``BatchRenderer.add`` raises: a LUT
  of more than 2048 entries, a volume layout it has no kernel for) or
  whose batch fails to launch completes with ``error`` set and no pixels,
  and the server goes on with the others.
"""

import time
from collections import OrderedDict, deque
from dataclasses import dataclass
from typing import Deque, Dict, Hashable, List, Optional

import numpy as np

from . import _lib, trace
from .batch import BatchRenderer, DeviceBlobVolume

_TEXEL_BYTES = {("f32", "none"): 4, ("f16", "none"): 2, ("u8", "none"): 1,
                ("f32", "f32x3"): 16, ("f16", "f32x3"): 16, ("u8", "f32x3"): 16,
                ("f32", "sn8x3"): 8, ("f16", "sn8x3"): 8, ("u8", "sn8x3"): 4, ("f16", "f16x3"): 8}


def volume_bricks(shape, brick_log2: int = 3) -> int:
    """Bricks of a volume: per axis (s >> brick_log2) + 1 (the extended
    texture holds s + 2 texels, the edge copies included)."""
    nb = 1
    for s in shape:
        nb *= (int(s) >> brick_log2) + 1
    return nb


def volume_device_bytes(shape, scalar_format: str, normal_format: str, has_normals: bool,
                        brick_log2: int = 3) -> int:
    """Device bytes of a volume (apron bricks plus the per-brick range), as
    ``hvr_volume_create`` allocates them."""
    nf = normal_format if has_normals else "none"
    t = _TEXEL_BYTES[scalar_format, nf]
    e = (1 << brick_log2) + 1
    nb = volume_bricks(shape, brick_log2)
    return nb * e ** 3 * t + nb * 8


# Upper bounds of what one batch of hvr_batch_run_async stages per request:
# descriptor (below 512 bytes), LUT of at most 2048 RGBA float entries and
# its prefix counts, each block aligned to 256 bytes.
_DESC_BYTES = 512
_MAX_LUT = 2048


def _pool(need: int) -> int:
    """Bytes a batch pool holds for ``need``: pools start at 4 KiB and
    double until they fit (never shrink)."""
    n = 4096
    while n < need:
        n *= 2
    return n


def batch_pool_bytes(max_requests: int, max_pixels: int, max_bricks: int) -> int:
    """Device bytes one batch slot holds at most for batches of up to
    ``max_requests`` requests and ``max_pixels`` pixels over volumes of up
    to ``max_bricks`` bricks: the RGBA8 output pool (each request padded to
    64 pixels), the transparent-brick flags (one byte per brick per request,
    padded to 256), the staged block (header, descriptors, LUTs, prefix
    counts) and the tile counters; each pool rounded up as it grows."""
    out = 4 * (max_pixels + 64 * max_requests)
    flags = max_requests * (-(-max_bricks // 256) * 256)
    stage = 256 + max_requests * (_DESC_BYTES + 16 * _MAX_LUT + 4 * (_MAX_LUT + 1) + 3 * 256)
    return _pool(out) + _pool(flags) + _pool(stage) + 256


def estimated_samples(width: int, height: int, camera, config, bounds=None) -> float:
    """A cheap estimate of a request's reference samples, to bound a batch's
    work: pixels whose rays can hit the box (the bounding sphere's image,
    from the camera distance and the field of view) times the samples of a
    mean chord (4 V / S of the box) at the step size, capped by max_steps.
    Termination and empty space are ignored: it orders and bounds requests,
    it is not a count."""
    import math

    lo = np.asarray(bounds[0] if bounds is not None else (-1.0, -1.0, -1.0), np.float64)
    hi = np.asarray(bounds[1] if bounds is not None else (1.0, 1.0, 1.0), np.float64)
    ext = np.maximum(hi - lo, 1e-12)
    centre = 0.5 * (lo + hi)
    radius = 0.5 * float(np.linalg.norm(ext))
    fov = float(getattr(camera, "fov", math.pi / 4))
    dist, target = getattr(camera, "distance", None), getattr(camera, "target", None)
    if dist is not None and target is not None:
        # the camera sits at `distance` from its target (the placement
        # rotates an offset of that length), so at least
        # distance - |target - centre| from the centre: the nearer bound
        # gives the larger disc
        off = float(np.linalg.norm(np.asarray(target, np.float64) - centre))
        d = max(float(dist) - off, 0.0)
    else:
        d = 3.0 * radius
    if d <= radius:
        frac = 1.0
    else:
        r_ndc = math.tan(math.asin(radius / d)) / math.tan(fov / 2)     # disc radius, NDC units of y
        aspect = width / height
        frac = min(1.0, math.pi * r_ndc * r_ndc / (4.0 * aspect))
    chord = 4.0 * float(np.prod(ext)) / (2.0 * (ext[0] * ext[1] + ext[1] * ext[2] + ext[0] * ext[2]))
    step = float(getattr(config, "step_size", 0.01)) if config is not None else 0.01
    steps = float(getattr(config, "max_steps", 500)) if config is not None else 500.0
    return width * height * frac * min(steps, chord / step)


@dataclass
class ServeRequest:
    """One render request; the server fills the timing fields and ``rgba``."""
    volume: Hashable                  # key given to RenderServer.register
    camera: object
    ctf: object
    otf: object
    width: int = 512
    height: int = 512
    config: object = None
    light: object = None
    tau: float = 0.99
    arrival: Optional[float] = None   # perf_counter() time; stamped by submit if None
    launched: Optional[float] = None
    done: Optional[float] = None
    rgba: Optional[np.ndarray] = None  # width*height*4 uint8, bottom row first
    tag: object = None
    estimate: Optional[float] = None  # estimated samples (set by the server when it bounds batches by work)
    cancelled: bool = False           # set by RenderServer.cancel
    error: Optional[str] = None       # why the request was not rendered (then rgba is None)

    @property
    def latency(self) -> Optional[float]:
        return None if self.done is None else self.done - self.arrival


def path_requests(volume: Hashable, cameras, ctf, otf, width: int = 512, height: int = 512, config=None,
                  light=None, tau: float = 0.99, start: Optional[float] = None,
                  interval: Optional[float] = None, tag=None) -> List[ServeRequest]:
    """One request per camera of a sequence -- ``CameraPath.generate_frames``,
    ``CameraController.animate_to`` or any list of ``Camera`` -- on the
    registered ``volume``, in order.  With ``start`` and ``interval`` the
    arrivals are ``start + k * interval`` (a frame stream at 1 / interval
    frames per second); otherwise ``submit`` stamps them.  ``tag`` becomes
    ``(tag, k)``."""
    reqs = []
    for k, cam in enumerate(cameras):
        arrival = None if start is None or interval is None else start + k * interval
        reqs.append(ServeRequest(volume=volume, camera=cam, ctf=ctf, otf=otf, width=width, height=height,
                                 config=config, light=light, tau=tau, arrival=arrival,
                                 tag=None if tag is None else (tag, k)))
    return reqs


@dataclass
class _Resident:
    volume: object
    bytes: int
    pins: int = 0


class RenderServer:
    """Continuous-batching render server on one GPU (see the module doc)."""

    def __init__(self, device: int = 0, scalar_format: str = "f16", normal_format: str = "sn8x3",
                 brick_log2: int = 3, tile_shape: str = "8x8", memory_budget: Optional[int] = None,
                 max_batch_requests: int = 256, max_batch_pixels: int = 1 << 25, pipeline: int = 3,
                 keep_results: bool = True, backend=None, max_batch_samples: Optional[float] = None,
                 kernel: str = "default", readback: str = "direct"):
        self.scalar_format, self.normal_format, self.brick_log2 = scalar_format, normal_format, brick_log2
        self.backend = backend if backend is not None else BatchRenderer(
            device=device, scalar_format=scalar_format, normal_format=normal_format, brick_log2=brick_log2,
            tile_shape=tile_shape, pipeline=pipeline, kernel=kernel, readback=readback)
        if memory_budget is None:
            free, total = _lib.device_memory(device)
            memory_budget = int(0.9 * free)
        self.keep_results = keep_results
        self.max_batch_requests = max_batch_requests
        self.max_batch_pixels = max_batch_pixels
        self.max_batch_samples = max_batch_samples
        self.pipeline = pipeline
        # the pools of the batches in flight (outputs, brick flags, staging),
        # for the largest registered volume's brick count
        self.max_bricks = 0
        self.reserved = self._reserve(0)
        self.memory_budget = int(memory_budget)
        if self.memory_budget <= self.reserved:
            raise ValueError("memory budget does not cover the batches' pools")
        self._registry: Dict[Hashable, tuple] = {}
        self._resident: "OrderedDict[Hashable, _Resident]" = OrderedDict()
        self._queue: Deque[ServeRequest] = deque()
        self._flight: Dict[int, List[ServeRequest]] = {}
        self.batches = 0
        self.evictions = 0
        self.cancelled = 0
        self.failed = 0
        self._failed: List[ServeRequest] = []   # completed with an error, returned by the next poll

    # -- volumes -------------------------------------------------------------
    def register(self, key: Hashable, volume, device_normals: bool = False):
        """Make ``volume`` (``Volume`` or ``DeviceBlobVolume``) servable under
        ``key``; it is loaded on first use.  Raises ``ValueError`` if it can
        never fit the budget."""
        has_n = volume.has_normals or (device_normals and not isinstance(volume, DeviceBlobVolume))
        shape = volume.shape if isinstance(volume, DeviceBlobVolume) else volume.data.shape
        nbytes = volume_device_bytes(shape, self.scalar_format, self.normal_format, has_n, self.brick_log2)
        reserved = self._reserve(max(self.max_bricks, volume_bricks(shape, self.brick_log2)))
        if nbytes > self.memory_budget - reserved:
            raise ValueError(f"volume {key!r} needs {nbytes} bytes, more than the budget allows")
        bricks = volume_bricks(shape, self.brick_log2)
        if bricks > self.max_bricks:
            # the batches' pools at their largest now (what the budget
            # reserves), so that no launch grows one while others run
            reserve = getattr(self.backend, "reserve", None)
            if reserve is not None:
                reserve(self.max_batch_requests, self.max_batch_pixels, bricks)
            self.max_bricks = bricks
        self.reserved = reserved
        self._registry[key] = (volume, device_normals, nbytes)

    def _reserve(self, max_bricks: int) -> int:
        return self.pipeline * batch_pool_bytes(self.max_batch_requests, self.max_batch_pixels, max_bricks)

    def preload(self, keys=None):
        """Load registered volumes now (as many as the budget holds)."""
        for key in (self._registry if keys is None else keys):
            if not self._make_resident(key):
                break

    def resident_bytes(self) -> int:
        return sum(r.bytes for r in self._resident.values())

    @staticmethod
    def freeze_heap():
        """Collect once and move every object alive now into the garbage
        collector's permanent generation (``gc.freeze``).  The serving loop
        runs in the calling thread, so a full collection over a large
        long-lived heap (volumes, requests, the application's own objects)
        stalls launches and collections for its whole duration -- 17-35 ms
        measured with the serving benchmark's heap, which then sets the p99.
        Call it after setup, before taking traffic; objects created later
        are collected as usual.  Process-wide."""
        import gc

        gc.collect()
        gc.freeze()

    def resident_keys(self) -> List[Hashable]:
        return list(self._resident)

    def _make_resident(self, key) -> bool:
        r = self._resident.get(key)
        if r is not None:
            self._resident.move_to_end(key)
            return True
        volume, device_normals, nbytes = self._registry[key]
        room = self.memory_budget - self.reserved - self.resident_bytes()
        if nbytes > room:
            for k in list(self._resident):            # least recently used first
                if room >= nbytes:
                    break
                v = self._resident[k]
                if v.pins == 0:
                    self.backend.release(v.volume)
                    del self._resident[k]
                    room += v.bytes
                    self.evictions += 1
            if nbytes > room:
                return False
        self.backend.upload(volume, device_normals=device_normals)
        self._resident[key] = _Resident(volume, nbytes)
        return True

    # -- requests ------------------------------------------------------------
    def submit(self, req: ServeRequest) -> ServeRequest:
        if req.volume not in self._registry:
            raise KeyError(f"unknown volume {req.volume!r}")
        if req.arrival is None:
            req.arrival = time.perf_counter()
        self._queue.append(req)
        return req

    def pending(self) -> int:
        """Submitted requests that poll has not returned yet (queued, in
        flight, or failed and about to be returned)."""
        return len(self._queue) + len(self._failed) + sum(len(v) for v in self._flight.values())

    def cancel(self, req: ServeRequest) -> bool:
        """Withdraw a submitted request.  A queued request leaves the queue
        and is never launched; a request already in a batch on the GPU runs
        to the end of that batch (a launch is not interrupted), but its
        pixels are discarded when the batch is collected.  Cancelled
        requests are not returned by ``poll``.  Returns False when the
        request had already completed (or was never submitted)."""
        if req.cancelled:
            return True
        for i, r in enumerate(self._queue):
            if r is req:
                del self._queue[i]
                req.cancelled = True
                self.cancelled += 1
                return True
        for reqs in self._flight.values():
            if any(r is req for r in reqs):
                req.cancelled = True
                self.cancelled += 1
                return True
        return False

    def poll(self) -> List[ServeRequest]:
        """Collect finished batches, then launch queued requests into free
        slots; never waits for the GPU.  Returns the completed requests."""
        out, self._failed = self._failed, []
        for k in list(self._flight):
            if self.backend.done(k):
                out.extend(self._collect(k))
        while self._queue:
            k = self.backend.free_slot()
            if k is None:
                break
            if k in self._flight:
                # it finished after the pass above: collect its results
                # before the slot is refilled
                out.extend(self._collect(k))
            if not self._launch():
                break
        return out

    def _collect(self, k) -> List[ServeRequest]:
        with trace.span("hvr.serve.collect"):
            return self._collect_slot(k)

    def _collect_slot(self, k) -> List[ServeRequest]:
        now = time.perf_counter()
        reqs = self._flight.pop(k)
        for i, r in enumerate(reqs):
            self._resident[r.volume].pins -= 1
            if r.cancelled:
                continue
            r.done = now
            if self.keep_results:
                r.rgba = np.array(self.backend.view(i, k))
        return [r for r in reqs if not r.cancelled]

    def _estimate(self, r: ServeRequest) -> float:
        if r.estimate is None:
            vol = self._registry[r.volume][0]
            lo, hi = getattr(vol, "min_bounds", None), getattr(vol, "max_bounds", None)
            r.estimate = estimated_samples(r.width, r.height, r.camera, r.config,
                                           None if lo is None or hi is None else (lo, hi))
        return r.estimate

    def _launch(self) -> bool:
        with trace.span("hvr.serve.admit"):
            return self._admit_and_launch()

    def _fail(self, r: ServeRequest, err: BaseException):
        r.error = f"{type(err).__name__}: {err}"
        r.done = time.perf_counter()
        self.failed += 1
        self._failed.append(r)

    def _admit_and_launch(self) -> bool:
        chosen, pixels, work, waiting, popped = [], 0, 0.0, deque(), []
        try:
            while self._queue and len(chosen) < self.max_batch_requests:
                r = self._queue[0]
                px = r.width * r.height
                if chosen and pixels + px > self.max_batch_pixels:
                    break
                if self.max_batch_samples is not None:
                    est = self._estimate(r)
                    if chosen and work + est > self.max_batch_samples:
                        break
                    work += est
                popped.append(self._queue.popleft())
                # pin before loading others, so a request's volume is not
                # evicted by the next request of the same batch
                if self._make_resident(r.volume):
                    self._resident[r.volume].pins += 1
                    chosen.append(r)
                    pixels += px
                else:
                    waiting.append(r)
        except BaseException:
            # a volume could not be loaded: the queue and the pins as before
            for c in chosen:
                self._resident[c.volume].pins -= 1
            self._queue.extendleft(reversed(popped))
            raise
        self._queue.extendleft(reversed(waiting))
        if not chosen:
            return False
        added = 0
        try:
            for r in chosen:
                vol = self._registry[r.volume][0]
                self.backend.add(vol, r.camera, r.ctf, r.otf, r.width, r.height, r.config, r.light, r.tau)
                added += 1
            k = self.backend.submit()
        except Exception as e:
            # drop the partial batch and release the pins; a request the
            # batched path rejects completes with the error and the others
            # go back to the front of the queue, while a batch that fails to
            # launch completes every request in it with the error
            self.backend.clear()
            for r in chosen:
                self._resident[r.volume].pins -= 1
            if added < len(chosen):
                self._fail(chosen[added], e)
                self._queue.extendleft(reversed([r for i, r in enumerate(chosen) if i != added]))
            else:
                for r in chosen:
                    self._fail(r, e)
            return True
        now = time.perf_counter()
        for r in chosen:
            r.launched = now
        self._flight[k] = chosen
        self.batches += 1
        return True

    def drain(self, sleep: float = 0.0) -> List[ServeRequest]:
        """Poll until every submitted request completed."""
        out = []
        while self.pending():
            got = self.poll()
            out.extend(got)
            if not self._flight and self._queue:
                # nothing running and nothing launchable: cannot progress
                raise RuntimeError(f"{len(self._queue)} queued requests cannot be admitted")
            if not got and sleep:
                time.sleep(sleep)
        return out

    def close(self):
        if self.backend is not None and hasattr(self.backend, "close"):
            self.backend.close()
        self.backend = None
