"""RenderServer scheduling on the CPU (pyvr_amd/hip_renderer/serve.py) with a
stand-in backend: FIFO batches within the request and pixel limits,
continuous batching into free slots, least-recently-used eviction that
never touches volumes in use, requests waiting for memory while later ones
proceed, and the footprint rule of hvr_volume_create."""

from types import SimpleNamespace

import numpy as np
import pytest

from pyvr_amd.hip_renderer.serve import (RenderServer, ServeRequest, batch_pool_bytes, volume_bricks,
                                         volume_device_bytes)


class FakeBackend:
    def __init__(self, pipeline=2):
        self.slots = [[] for _ in range(pipeline)]
        self.busy = [False] * pipeline
        self.cur = 0
        self.reuse = False
        self.resident = set()
        self.uploads, self.releases, self.launches = [], [], []

    def upload(self, v, device_normals=False):
        assert id(v) not in self.resident
        self.resident.add(id(v))
        self.uploads.append(v.name)

    def release(self, v):
        self.resident.remove(id(v))
        self.releases.append(v.name)

    def free_slot(self):
        # as BatchRenderer.free_slot: asks done() (which may see it finish)
        return None if (self.reuse and not self.done(self.cur)) else self.cur

    def add(self, vol, camera, ctf, otf, w, h, config, light, tau):
        assert id(vol) in self.resident, "request on a volume that is not resident"
        if self.reuse:
            assert not self.busy[self.cur]
            self.slots[self.cur] = []
            self.reuse = False
        self.slots[self.cur].append((vol.name, w, h))

    def submit(self):
        k = self.cur
        self.busy[k] = True
        self.launches.append(list(self.slots[k]))
        self.cur = (k + 1) % len(self.slots)
        self.reuse = True
        return k

    def clear(self):
        # as BatchRenderer.clear: drop the batch being filled
        assert not self.busy[self.cur]
        self.slots[self.cur] = []
        self.reuse = False

    def done(self, k):
        return not self.busy[k]

    def finish(self, k):
        self.busy[k] = False

    def view(self, i, k):
        _, w, h = self.slots[k][i]
        return np.zeros(w * h * 4, np.uint8)


def vol(name, n):
    return SimpleNamespace(name=name, data=np.zeros((n, n, n), np.float32), has_normals=True)


def req(key, w=16, h=16):
    return ServeRequest(volume=key, camera=None, ctf=None, otf=None, width=w, height=h)


B16 = volume_device_bytes((16, 16, 16), "f16", "sn8x3", True)     # 3^3 bricks of 729 texels * 8 B + range


def test_footprint_rule():
    assert B16 == 27 * 729 * 8 + 27 * 8
    assert volume_device_bytes((96, 160, 224), "u8", "sn8x3", False, 4) == 7 * 11 * 15 * 17 ** 3 * 1 + 7 * 11 * 15 * 8


def make(budget_volumes, pipeline=2, **kw):
    be = FakeBackend(pipeline)
    kw.setdefault("max_batch_pixels", 1 << 12)
    kw.setdefault("max_batch_requests", 256)
    reserve = pipeline * batch_pool_bytes(kw["max_batch_requests"], kw["max_batch_pixels"], volume_bricks((16,) * 3))
    srv = RenderServer(backend=be, memory_budget=reserve + budget_volumes * B16 + 1, pipeline=pipeline, **kw)
    for k in "ABC":
        srv.register(k, vol(k, 16))
    return srv, be


def test_fifo_batches_and_continuous_batching():
    srv, be = make(3, max_batch_requests=3)
    for k in "ABCAB":
        srv.submit(req(k))
    assert srv.poll() == []
    # two slots: the first takes 3 requests, the second the remaining 2
    assert [[n for n, _, _ in l] for l in be.launches] == [["A", "B", "C"], ["A", "B"]]
    srv.submit(req("C"))
    assert srv.poll() == [] and len(be.launches) == 2        # no free slot: C waits
    be.finish(0)
    done = srv.poll()
    assert [r.volume for r in done] == ["A", "B", "C"] and all(r.rgba.size == 16 * 16 * 4 for r in done)
    assert [n for n, _, _ in be.launches[-1]] == ["C"]       # joined the freed slot
    be.finish(1)
    be.finish(0)
    assert len(srv.drain()) == 3 and srv.pending() == 0
    assert all(r.latency is not None and r.latency >= 0 for r in done)


def test_pixel_limit_splits_batches():
    srv, be = make(3, max_batch_pixels=1000)
    for k in "AB":
        srv.submit(req(k, 20, 20))                           # 400 px each
    srv.submit(req("C", 20, 20))
    srv.poll()
    assert [len(l) for l in be.launches] == [2, 1]


def test_lru_eviction_never_touches_volumes_in_use():
    srv, be = make(2, max_batch_requests=1)
    srv.submit(req("A"))
    srv.submit(req("B"))
    srv.poll()                                               # A and B in flight, both pinned
    srv.submit(req("C"))
    srv.poll()
    assert be.releases == [] and len(be.launches) == 2       # C waits: nothing evictable, no slot
    be.finish(0)
    srv.poll()                                               # A done and unpinned; C evicts A (LRU)
    assert be.releases == ["A"] and be.launches[-1][0][0] == "C"
    be.finish(1)
    be.finish(0)
    srv.drain()
    srv.submit(req("B"))                                     # resident: no upload
    srv.poll()
    assert be.uploads == ["A", "B", "C"]


def test_waiting_request_lets_later_ones_pass():
    srv, be = make(1, max_batch_requests=4, pipeline=2)
    srv.submit(req("A"))
    srv.poll()                                               # A resident and pinned
    srv.submit(req("B"))                                     # B cannot load while A is pinned
    srv.submit(req("A"))
    srv.poll()
    assert [n for n, _, _ in be.launches[1]] == ["A"]        # the later A request went ahead
    assert [r.volume for r in srv._queue] == ["B"]
    be.finish(0)
    be.finish(1)
    assert [r.volume for r in srv.poll()] == ["A", "A"]      # both done; B evicts A and launches
    assert be.releases == ["A"] and be.uploads == ["A", "B"]
    be.finish(0)
    assert [r.volume for r in srv.poll()] == ["B"] and srv.pending() == 0


class LateBackend(FakeBackend):
    """A slot finishes on its own after a number of completion queries, so a
    batch can finish between poll()'s collection pass and its launch pass."""

    def __init__(self, pipeline=2, after=2):
        super().__init__(pipeline)
        self.after = after
        self.queries = [0] * pipeline

    def submit(self):
        self.queries[self.cur] = 0
        return super().submit()

    def done(self, k):
        if self.busy[k]:
            self.queries[k] += 1
            if self.queries[k] > self.after:
                self.busy[k] = False
        return not self.busy[k]


@pytest.mark.parametrize("after", [0, 1, 2, 3])
def test_no_request_lost_when_a_batch_finishes_mid_poll(after):
    be = LateBackend(2, after)
    srv = RenderServer(backend=be, memory_budget=2 * batch_pool_bytes(2, 1 << 12, 27) + 3 * B16 + 1, pipeline=2,
                       max_batch_requests=2, max_batch_pixels=1 << 12)
    for k in "ABC":
        srv.register(k, vol(k, 16))
    reqs = [srv.submit(req(k)) for k in "ABCABCABCA"]
    done = []
    for _ in range(1000):
        done.extend(srv.poll())
        if not srv.pending():
            break
    assert sorted(id(r) for r in done) == sorted(id(r) for r in reqs)
    assert all(r.done is not None for r in reqs)
    assert all(v.pins == 0 for v in srv._resident.values())


def test_rejects_what_can_never_fit():
    srv, _ = make(1)
    with pytest.raises(ValueError):
        srv.register("big", vol("big", 64))
    with pytest.raises(KeyError):
        srv.submit(req("nope"))


def test_camera_path_requests_are_served_in_order():
    """A CameraPath turned into a request stream: one request per frame, in
    order, with the path's cameras and the given arrival spacing; the server
    batches them in that order."""
    from pyvr_amd.camera import Camera, CameraController, CameraPath
    from pyvr_amd.hip_renderer.serve import path_requests

    frames = CameraPath([Camera.front_view(), Camera.side_view(), Camera.top_view()]).generate_frames(7)
    reqs = path_requests("A", frames, None, None, width=16, height=16, start=10.0, interval=0.02, tag="orbit")
    assert [r.camera for r in reqs] == frames and [r.tag for r in reqs] == [("orbit", k) for k in range(7)]
    assert [r.arrival for r in reqs] == pytest.approx([10.0 + 0.02 * k for k in range(7)])
    srv, be = make(3, max_batch_requests=4)
    for r in reqs:
        srv.submit(r)
    srv.poll()
    assert [len(l) for l in be.launches] == [4, 3]
    be.finish(0)
    be.finish(1)
    done = srv.drain()
    assert [r.tag[1] for r in done] == list(range(7))
    moves = CameraController().animate_to(Camera.isometric_view(), n_frames=5)
    assert all(r.arrival is None for r in path_requests("B", moves, None, None))


def test_pool_reservation_covers_what_a_batch_allocates():
    """The budget's reservation per slot bounds what hvr_batch_run_async
    grows its pools to (hipvr.hip: output pool of 64-pixel-padded images,
    256-padded brick flags per request, staged block; 4 KiB start,
    doubling), for random batches within the limits."""
    rng = np.random.default_rng(5)

    def grow(cap, need):
        if need <= cap:
            return cap
        n = cap or 4096
        while n < need:
            n *= 2
        return n

    for _ in range(200):
        max_req, max_px, max_b = int(rng.integers(1, 40)), int(rng.integers(1, 1 << 22)), int(rng.integers(1, 1 << 21))
        out = flags = stage = 0
        for _ in range(5):                      # successive batches on one slot
            n = int(rng.integers(1, max_req + 1))
            px = rng.integers(1, max(2, max_px // n), n)
            bricks = rng.integers(1, max_b + 1, n)
            luts = rng.integers(1, 2049, n)
            out = grow(out, 4 * int(sum(-(-int(p) // 64) * 64 for p in px)))
            flags = grow(flags, int(sum(-(-int(b) // 256) * 256 for b in bricks)))
            desc_end = -(-(256 + n * 512) // 256) * 256
            lut_end = -(-(desc_end + int(sum(16 * luts))) // 256) * 256
            stage = grow(stage, -(-(lut_end + int(sum(4 * (luts + 1)))) // 256) * 256)
        assert out + flags + stage + 256 <= batch_pool_bytes(max_req, max_px, max_b)


def test_registering_a_larger_volume_raises_the_reservation():
    srv = RenderServer(backend=FakeBackend(2), memory_budget=1 << 30, pipeline=2, max_batch_pixels=1 << 12)
    srv.register("A", vol("A", 16))
    before = srv.reserved
    assert before == 2 * batch_pool_bytes(256, 1 << 12, volume_bricks((16,) * 3))
    srv.register("D", vol("D", 64))
    assert srv.reserved > before
    assert srv.reserved == 2 * batch_pool_bytes(256, 1 << 12, volume_bricks((64,) * 3))
    # a volume whose flags would push the reservation past the budget is refused
    tight = RenderServer(backend=FakeBackend(2), memory_budget=before + 2 * B16, pipeline=2,
                         max_batch_pixels=1 << 12)
    tight.register("A", vol("A", 16))
    with pytest.raises(ValueError):
        tight.register("D", vol("D", 64))
    assert tight.reserved == before


def test_batches_bounded_by_estimated_samples():
    """With max_batch_samples a burst of requests is split so that each
    batch's estimated work stays within the bound (a single request larger
    than the bound still launches alone)."""
    from pyvr_amd.camera import Camera
    from pyvr_amd.config import RenderConfig
    from pyvr_amd.hip_renderer.serve import estimated_samples

    cam = Camera.from_spherical(target=np.zeros(3, np.float32), azimuth=0.3, elevation=0.2, roll=0.0, distance=3.0)
    cfg = RenderConfig.balanced()
    one = estimated_samples(64, 64, cam, cfg)
    assert 0 < one <= 64 * 64 * cfg.max_steps
    # a larger image and a finer step cost more; a farther camera less
    assert estimated_samples(128, 128, cam, cfg) == pytest.approx(4 * one)
    assert estimated_samples(64, 64, cam, RenderConfig.high_quality()) > one
    far = Camera.from_spherical(target=np.zeros(3, np.float32), azimuth=0.3, elevation=0.2, roll=0.0, distance=12.0)
    assert estimated_samples(64, 64, far, cfg) < one
    srv, be = make(3, max_batch_requests=16, max_batch_pixels=1 << 20, max_batch_samples=2.5 * one)
    for k in "ABCABC":
        r = req(k, 64, 64)
        r.camera, r.config = cam, cfg
        srv.submit(r)
    srv.poll()
    assert [len(l) for l in be.launches] == [2, 2]
    be.finish(0)
    srv.poll()
    assert [len(l) for l in be.launches] == [2, 2, 2]


def test_cancel_queued_and_in_flight_requests():
    """A queued request that is cancelled is never launched; one cancelled
    while its batch is on the GPU is dropped at collection (its volume is
    unpinned like the others); neither comes back from poll."""
    srv, be = make(3, max_batch_requests=2, pipeline=1)
    a, b, c, d = (srv.submit(req(k)) for k in "ABCA")
    srv.poll()                                   # A, B launched; C, A queued
    assert [n for n, _, _ in be.launches[0]] == ["A", "B"]
    assert srv.cancel(b) and srv.cancel(c)       # in flight, queued
    assert srv.cancel(c)                         # idempotent
    assert srv.pending() == 3 and srv.cancelled == 2
    be.finish(0)
    done = srv.poll()                            # collects A (B dropped); launches the remaining A
    assert done == [a] and a.rgba is not None and b.rgba is None and b.done is None
    assert [n for n, _, _ in be.launches[1]] == ["A"]
    assert srv._resident["B"].pins == 0
    be.finish(0)
    assert srv.poll() == [d] and srv.pending() == 0
    assert not srv.cancel(a)                     # already completed
    assert not srv.cancel(req("A"))              # never submitted


def test_freeze_heap_moves_live_objects_out_of_collection():
    """RenderServer.freeze_heap: what is alive at setup is no longer scanned
    by the collector (gc.freeze), so a full collection during serving costs
    only the objects created since."""
    import gc

    from pyvr_amd.hip_renderer.serve import RenderServer

    keep = [[i] for i in range(1000)]
    try:
        RenderServer.freeze_heap()
        assert gc.get_freeze_count() >= len(keep)
    finally:
        gc.unfreeze()
    assert gc.get_freeze_count() == 0


def test_pools_reserved_when_a_larger_volume_registers():
    """The server grows the backend's batch pools to its limits (for the
    largest registered volume) at registration, so no launch grows a pool
    while others are in flight; smaller volumes reserve nothing more."""
    class ReservingBackend(FakeBackend):
        def __init__(self, pipeline=2):
            super().__init__(pipeline)
            self.reserved = []

        def reserve(self, requests, pixels, bricks):
            self.reserved.append((requests, pixels, bricks))

    be = ReservingBackend(2)
    srv = RenderServer(backend=be, memory_budget=1 << 30, pipeline=2, max_batch_requests=32,
                       max_batch_pixels=1 << 12)
    srv.register("A", vol("A", 32))
    srv.register("B", vol("B", 16))
    srv.register("C", vol("C", 64))
    assert be.reserved == [(32, 1 << 12, volume_bricks((32,) * 3)), (32, 1 << 12, volume_bricks((64,) * 3))]


def test_estimate_uses_the_camera_distance_bound():
    """estimated_samples places the camera at its distance from the target,
    less the target's offset from the box centre (the nearest it can be):
    a target off the centre can only raise the estimate."""
    from pyvr_amd.camera import Camera
    from pyvr_amd.config import RenderConfig
    from pyvr_amd.hip_renderer.serve import estimated_samples

    cfg = RenderConfig.balanced()
    centred = Camera.from_spherical(target=np.zeros(3, np.float32), azimuth=0.3, elevation=0.2, roll=0.0,
                                    distance=6.0)
    off = Camera.from_spherical(target=np.array([0.0, 1.0, 0.0], np.float32), azimuth=0.3, elevation=0.2,
                                roll=0.0, distance=6.0)
    inside = Camera.from_spherical(target=np.zeros(3, np.float32), azimuth=0.3, elevation=0.2, roll=0.0,
                                   distance=1.0)
    a, b = estimated_samples(128, 128, centred, cfg), estimated_samples(128, 128, off, cfg)
    assert 0 < a < b
    # inside the bounding sphere every pixel may hit the box
    assert estimated_samples(128, 128, inside, cfg) == pytest.approx(128 * 128 * min(cfg.max_steps,
                                                                                    4 * 8 / 24 / cfg.step_size))


class RejectingBackend(FakeBackend):
    """Rejects requests 13 pixels wide (as BatchRenderer.add rejects a LUT
    the batch path cannot hold), fails the launches listed in
    `fail_launches` (by launch attempt) and uploads of volumes in
    `fail_uploads`."""

    def __init__(self, pipeline=2, fail_launches=(), fail_uploads=()):
        super().__init__(pipeline)
        self.fail_launches, self.fail_uploads, self.attempts = set(fail_launches), set(fail_uploads), 0

    def add(self, vol, camera, ctf, otf, w, h, config, light, tau):
        if self.reuse:
            assert not self.busy[self.cur]
            self.slots[self.cur] = []
            self.reuse = False
        if w == 13:
            raise RuntimeError("hvr_batch_add failed (-1): batched LUTs hold 0 to 2048 entries")
        super().add(vol, camera, ctf, otf, w, h, config, light, tau)

    def submit(self):
        self.attempts += 1
        if self.attempts in self.fail_launches:
            raise RuntimeError("hvr_batch_run_async failed (-4): capturing the batch graph failed")
        return super().submit()

    def upload(self, v, device_normals=False):
        if v.name in self.fail_uploads:
            raise RuntimeError("hvr_volume_create failed (-2): out of memory")
        super().upload(v, device_normals)


def make_rejecting(**kw):
    be = RejectingBackend(**kw)
    reserve = 2 * batch_pool_bytes(256, 1 << 12, volume_bricks((16,) * 3))
    srv = RenderServer(backend=be, memory_budget=reserve + 3 * B16 + 1, pipeline=2, max_batch_pixels=1 << 12)
    for k in "ABC":
        srv.register(k, vol(k, 16))
    return srv, be


def finish_all(be):
    for k in range(len(be.busy)):
        be.finish(k)


def test_rejected_request_completes_with_error_and_the_rest_render():
    srv, be = make_rejecting()
    reqs = [srv.submit(req("A")), srv.submit(req("B", 13, 9)), srv.submit(req("C"))]
    out = srv.poll()
    # the rejected request comes back at once with its error; the others
    # were put back in order and launched as one batch
    assert out == [] and srv.failed == 1
    assert [[n for n, _, _ in l] for l in be.launches] == [["A", "C"]]
    finish_all(be)
    got = srv.drain()
    bad = reqs[1]
    assert bad.error is not None and "2048" in bad.error and bad.rgba is None and bad.done is not None
    # every request comes back exactly once, the rejected one included
    assert sorted(map(id, got)) == sorted(map(id, reqs))
    ok = [r for r in reqs if r is not bad]
    assert all(r.error is None and r.rgba is not None and r.rgba.size == 16 * 16 * 4 for r in ok)
    assert srv.pending() == 0 and all(v.pins == 0 for v in srv._resident.values())


def test_rejected_request_is_returned_by_poll():
    srv, be = make_rejecting()
    bad = srv.submit(req("A", 13, 5))
    assert srv.poll() == []                 # rejected during this poll's launch
    assert srv.poll() == [bad] and bad.error and srv.pending() == 0
    assert all(v.pins == 0 for v in srv._resident.values()) and be.launches == []


def test_failed_launch_completes_its_requests_with_the_error():
    srv, be = make_rejecting(fail_launches={1})
    first = [srv.submit(req(k)) for k in "AB"]
    srv.poll()                              # the first launch fails
    assert all("capturing" in r.error and r.rgba is None for r in first)
    assert srv.poll() == first and srv.failed == 2
    again = [srv.submit(req(k)) for k in "AB"]
    srv.poll()                              # the next launch succeeds
    finish_all(be)
    done = srv.drain()
    assert done == again and all(r.error is None and r.rgba is not None for r in again)
    assert all(v.pins == 0 for v in srv._resident.values())


def test_failed_upload_restores_the_queue_and_the_pins():
    srv, be = make_rejecting(fail_uploads={"B"})
    reqs = [srv.submit(req(k)) for k in "ABC"]
    with pytest.raises(RuntimeError, match="out of memory"):
        srv.poll()
    assert list(srv._queue) == reqs and all(v.pins == 0 for v in srv._resident.values())
    assert be.launches == []
    be.fail_uploads.clear()
    srv.poll()
    finish_all(be)
    assert srv.drain() == reqs and all(r.error is None for r in reqs)


def test_drain_returns_a_request_that_fails_last():
    srv, be = make_rejecting()
    bad = srv.submit(req("A", 13, 5))
    assert srv.drain() == [bad] and bad.error and srv.pending() == 0
