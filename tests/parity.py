"""Helpers for GPU-vs-oracle parity tests."""

import numpy as np

import oracle
from pyvr_amd.hip_renderer import VolumeRenderer


def render_hip(scene, camera=None, scalar_format="f32", normal_format="f32x3",
               brick_log2=None, counts=False, renderer=None):
    """Render ``scene`` through the public VolumeRenderer API; returns the
    (H, W, 4) uint8 image (bottom row first) and optionally the counts."""
    r = renderer or VolumeRenderer(scene.width, scene.height, config=scene.config, light=scene.light,
                                   scalar_format=scalar_format, normal_format=normal_format)
    try:
        if brick_log2 is not None:
            from pyvr_amd.hip_renderer import _lib
            _lib.check(_lib.lib().hvr_set_brick_log2(r.gl_manager.ctx, brick_log2), "set_brick")
        r.load_volume(scene.volume())
        r.set_camera(camera or scene.camera)
        r.set_transfer_functions(scene.ctf, scene.otf, scene.lut_size)
        r.set_opacity_threshold(scene.tau)
        img = np.frombuffer(r.render(), np.uint8).reshape(scene.height, scene.width, 4).copy()
        cnt = r.sample_counts() if counts else None
    finally:
        if renderer is None:
            r.close()
    return (img, cnt) if counts else img


def oracle_render(scene, camera=None, mode="f64", scalar=None, normals="auto"):
    u = scene.uniforms(camera)
    sc = scene.data if scalar is None else scalar
    if isinstance(normals, str) and normals == "auto":
        nm = scene.normals if scene.normals is not None else sc
    else:
        nm = normals
    return oracle.render(u, sc, nm, scene.lut(), mode=mode)


def compare(img, ref_rgba8):
    """Per-pixel max channel error stats."""
    d = np.abs(img.astype(np.int16) - ref_rgba8.astype(np.int16)).max(axis=-1)
    return {
        "max": int(d.max()),
        "frac_le1": float((d <= 1).mean()),
        "frac_exact": float((d == 0).mean()),
        "n_gt1": int((d > 1).sum()),
        "mean": float(np.abs(img.astype(np.float64) - ref_rgba8).mean()),
    }


def assert_parity(img, ref_rgba8, frac=0.999, label=""):
    st = compare(img, ref_rgba8)
    print(f"[parity] {label} {st}", flush=True)
    assert st["frac_le1"] >= frac, f"{label}: {st}"
    return st
