"""The autotuner on the GPU, on the batched 64-request mix (FP16 +
SNORM8x3 volumes of 64^3 .. 256^3 and 96x160x224 at half size, images of
128^2 .. 512^2): it times every candidate (brick size x ray tile x kernel
variant), its pick is the fastest of them (or the default when nothing beats
it by more than 5 %) and stays at least as fast as the default configuration
when timed again, and the picked configuration's images equal the default
configuration's byte for byte."""

import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu


def _render_and_time(vols, reqs, lb, tile, kernel, runs=5):
    from pyvr_amd.hip_renderer import BatchRenderer

    b = BatchRenderer(scalar_format="f16", normal_format="sn8x3", brick_log2=lb, tile_shape=tile, kernel=kernel,
                      pipeline=1)
    try:
        idx = [b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
               for q in reqs]
        b.run()
        imgs = [b.result(i) for i in idx]
        ms = []
        for _ in range(runs):
            b.run()
            ms.append(b.last_ms())
    finally:
        b.close()
    return imgs, float(np.median(ms))


def test_autotuner_pick_is_fastest_and_renders_the_default_images(hip, tmp_path):
    from pyvr_amd.hip_renderer.autotune import tune

    vols = scenes.c3_volumes(scale=2)
    reqs = scenes.c3_requests(64, image_scale=2)
    out = tune(vols, reqs, "f16", "sn8x3", runs=3, cache=str(tmp_path / "tune.json"))
    table = out["table"]
    print(f"[autotune] {len(table)} candidates: " +
          ", ".join(f"lb{a}/{t}/{k} {ms:.3f}" for a, t, k, ms in table) +
          f" -> lb{out['brick_log2']}/{out['tile_shape']}/{out['kernel']}", flush=True)
    assert len(table) == 2 * 2 * 5
    best = min(t[3] for t in table)
    assert out["ms"] == best or (out["ms"] == table[0][3] and best >= 0.95 * table[0][3])
    pick = (out["brick_log2"], out["tile_shape"], out["kernel"])
    got, ms_pick = _render_and_time(vols, reqs, *pick)
    want, ms_default = _render_and_time(vols, reqs, 3, "8x8", "default")
    print(f"[autotune] retimed: pick {ms_pick:.3f} ms, default {ms_default:.3f} ms", flush=True)
    assert got == want
    assert ms_pick <= 1.05 * ms_default
    # a second call is answered from the cache
    again = tune(vols, reqs, "f16", "sn8x3", cache=str(tmp_path / "tune.json"))
    assert again["cached"] and (again["brick_log2"], again["tile_shape"], again["kernel"]) == pick
