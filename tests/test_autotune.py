"""Autotuner logic on the CPU (pyvr_amd/hip_renderer/autotune.py), with a
stand-in for BatchRenderer whose launch times are synthetic."""

from types import SimpleNamespace

import numpy as np

from pyvr_amd.hip_renderer import autotune

TIMES = {(3, "8x8"): 5.0, (3, "16x4"): 4.0, (4, "8x8"): 6.0, (4, "16x4"): 3.5}
KERNEL_FACTOR = {"default": 1.0, "waves2": 1.1, "waves8": 0.97, "nogate": 1.05, "noskip": 1.2}


class FakeBatch:
    made = []

    def __init__(self, device=0, scalar_format="f16", normal_format="sn8x3", brick_log2=3, pipeline=1):
        self.lb, self.tile, self.kernel, self.n = brick_log2, "8x8", "default", 0
        FakeBatch.made.append(self)

    def add(self, *a, **k):
        self.n += 1

    def set_tile_shape(self, t):
        self.tile = t

    def set_kernel(self, k):
        self.kernel = k

    def run(self):
        pass

    def last_ms(self):
        return TIMES[self.lb, self.tile] * KERNEL_FACTOR[self.kernel]

    def close(self):
        pass


def workload(n):
    vols = [SimpleNamespace(data=np.zeros((4, 5, 6)), has_normals=True),
            SimpleNamespace(data=np.zeros((8, 8, 8)), has_normals=False)]
    reqs = [SimpleNamespace(volume=i % 2, camera=None, ctf=None, otf=None, width=32 + i, height=16,
                            config=None, light=None, tau=0.99) for i in range(n)]
    return vols, reqs


def test_picks_fastest_and_caches(monkeypatch, tmp_path):
    monkeypatch.setattr(autotune, "BatchRenderer", FakeBatch)
    FakeBatch.made = []
    vols, reqs = workload(10)
    cache = str(tmp_path / "tune.json")
    out = autotune.tune(vols, reqs, cache=cache)
    assert (out["brick_log2"], out["tile_shape"], out["kernel"]) == (4, "16x4", "waves8")
    assert out["ms"] == min(t[3] for t in out["table"])
    # a best candidate within the margin of the first (default) one is not taken
    near = autotune.tune(vols, reqs, bricks=(3,), tiles=("8x8",))
    assert (near["brick_log2"], near["tile_shape"], near["kernel"]) == (3, "8x8", "default")
    assert near["ms"] == 5.0
    assert not out["cached"] and len(out["table"]) == 4 * 5
    assert {b.lb for b in FakeBatch.made} == {3, 4} and all(b.n == 10 for b in FakeBatch.made)
    FakeBatch.made = []
    again = autotune.tune(vols, reqs, cache=cache)
    assert again["cached"] and (again["brick_log2"], again["tile_shape"], again["kernel"]) == (4, "16x4", "waves8")
    assert again["table"] == out["table"]
    assert FakeBatch.made == []                       # nothing timed on a cache hit
    # another workload misses the cache
    vols2, reqs2 = workload(11)
    assert not autotune.tune(vols2, reqs2, cache=cache)["cached"]


def test_sampling_and_signature(monkeypatch):
    monkeypatch.setattr(autotune, "BatchRenderer", FakeBatch)
    FakeBatch.made = []
    vols, reqs = workload(100)
    autotune.tune(vols, reqs, max_requests=16)
    assert all(b.n == 16 for b in FakeBatch.made)
    s = autotune.signature(vols, reqs, "f16", "sn8x3")
    assert s.startswith("f16+sn8x3|4x5x6n,8x8x8|100|")
    assert autotune.signature(vols, list(reversed(reqs)), "f16", "sn8x3") == s   # order-free
