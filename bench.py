#!/usr/bin/env python3
"""Benchmark of the HIP volume renderer on one MI355X.

Metric: ray samples/s (plus fps and p50/p99 latency) at 512^3 / 1024^2,
batched and served forms against the plain one-request-per-launch kernel.

Default run (what one ``python bench.py`` prints, one JSON line):

* workload: the 512^3 Gaussian-blob volume (32 blobs, seed 2, generated on
  the device), FP16 scalars + FP16x3 normals, 24 cameras on a sphere of
  radius 3 at 1024x1024, high_quality preset, termination at alpha 0.95;
  the volume is resident in HBM before any timed region;
* a *step* renders the 24 cameras once;
* batched leg (``value``): the 24 cameras as 24 requests of one persistent
  graph-captured launch per step (BatchRenderer; brick size and ray tile
  from the autotuner), two steps in flight, every step ending with its 24
  images in pinned host memory (the march writes them there directly);
* plain leg (``plain_kernel``): the same frames one launch (ray setup +
  march) per frame -- the single-frame kernels VolumeRenderer.render()
  runs -- back to back into resident device buffers; ``frame_ms_p50/p99``
  are end-to-end per-frame calls (camera update, draw, RGBA8 copy to the
  host);
* serving leg (``serving``): 16 resident device-generated volumes of 256^3
  to 1024^3, Poisson arrivals at 500 and 1000 req/s, images of 512^2 to
  2048^2 over the fast / balanced / high_quality presets; p50/p99 latency
  (arrival to pixels in host memory) and samples/s against the plain
  one-request-per-launch renderer on the same requests;
* samples: reference-algorithm samples (pixel, i) -- i < max_steps, alpha
  below the termination threshold, sample inside the box -- counted by the
  renderer's counting pass before the timed region (the tests pin that count
  to the oracle's);
* roofline of the dominant kernel (the batched launch): ``traffic`` is the
  HBM bytes per launch that rocprofv3 measured for these kernel sources
  (FETCH_SIZE and WRITE_SIZE passes, scaled by a copy-kernel calibration;
  profiles/traffic.json), ``achieved`` that over the launch's HIP-event
  duration, ``frac`` against 8 TB/s.  ``alg`` is the algorithmic figure (8
  texel reads per reference sample at the stored format, plus the images),
  which counts L1/L2 re-reads and exceeds HBM peak; ``valu`` is the VALU
  issue of the kernel: VALU instructions per sample from the SQ counter pass
  against 1024 SIMDs issuing one wave64 instruction per 4 cycles at 2.4 GHz
  (``frac``; gfx950 issues two VALU instructions in some quad-cycles, so
  this single-issue figure is no ceiling and can exceed 1), rocprof's
  VALUBusy (``busy_counter``, from SQ_ACTIVE_INST_VALU), the share of VALU
  quad-cycles with two instructions issued, the SIMD cycles issuing VALU
  work (``simd_issue_frac``) and the wave-cycles spent waiting;
* cpu_baseline: the oracle's FP32 port of the reference shader (its loop
  runs every i < max_steps, as the shader does) on all host cores, on the
  volume's stored values read back from the device, whole frames of the
  workload's cameras for about 12 s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --workload c0|c1|c3|c4 (single workloads)
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
VALU_PEAK = 1024 * 2.4e9 / 4 * 64   # lane instructions/s: 1024 SIMDs, wave64 per 4 cycles, 2.4 GHz
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
ORDER_MODES = {"off": 0, "on": 1, "auto": 2}   # hvr_set_dispatch_order (HVR_ORDER_OFF / _ON / _AUTO)
READBACK_MODES = {"copy": 0, "direct": 1}      # hvr_set_readback (HVR_READBACK_COPY / _DIRECT)
SCALAR_BYTES = {"f32": 4, "f16": 2, "u8": 1}
NORMAL_BYTES = {"f32x3": 12, "sn8x3": 3, "f16x3": 6, "none": 0}
METRIC = "ray samples/s + fps at 512^3/1024^2; p50/p99 frame latency vs plain kernel"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# workloads

class Workload:
    """Everything a frame of the workload needs besides the camera."""

    def __init__(self, name, volume, ctf, otf, config, light, tau, width, height, cams):
        self.name, self.volume, self.ctf, self.otf = name, volume, ctf, otf
        self.config, self.light, self.tau = config, light, tau
        self.width, self.height, self.cams = width, height, cams

    @property
    def shape(self):
        data = getattr(self.volume, "data", None)
        return tuple(data.shape) if data is not None else tuple(self.volume.shape)

    @property
    def has_normals(self):
        return bool(self.volume.has_normals)

    def requests(self):
        import scenes

        return [scenes.Request(0, cam, self.ctf, self.otf, self.width, self.height, self.config, self.light,
                               self.tau) for cam in self.cams]


def make_workload(args) -> Workload:
    import scenes
    from pyvr_amd.config import RenderConfig
    from pyvr_amd.hip_renderer.batch import DeviceBlobVolume
    from pyvr_amd.transferfunctions import ColorTransferFunction

    if args.workload == "c2":
        n, img = args.size or 512, args.image or 1024
        vol = DeviceBlobVolume(shape=(n, n, n), n_blobs=32, seed=2, noise=1e-3, with_normals=True)
        return Workload(f"c2dev_blobs32_{n}_{img}", vol, ColorTransferFunction.from_colormap("viridis"),
                        scenes.seeded_otf(6, seed=2), RenderConfig.high_quality(), scenes.standard_light(), 0.95,
                        img, img, scenes.c2_cameras(args.cameras))
    if args.workload == "c1":
        sc = scenes.c1(size=args.size or 256, image=args.image or 1024)
    elif args.workload == "c0":
        sc = scenes.c0(size=args.size or 128, image=args.image or 512)
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    return Workload(sc.name, sc.volume(), sc.ctf, sc.otf, sc.config, sc.light, sc.tau, sc.width, sc.height,
                    [sc.camera] * args.cameras)


# ---------------------------------------------------------------------------
# roofline

def kernel_source_hash() -> str:
    """Hash of the device sources: measured traffic is only reported for the
    kernels it was measured on."""
    import hashlib

    h = hashlib.sha256()
    for rel in ("pyvr_amd/csrc/hvr_device.h", "pyvr_amd/csrc/hvr_rules.h", "pyvr_amd/csrc/hvr_launch.h",
                "pyvr_amd/csrc/inst.hip", "pyvr_amd/csrc/hipvr.hip"):
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_profile(key: str) -> dict:
    """The counter figures rocprofv3 measured for this workload and these
    kernel sources (profiles/traffic.json), or {} when none match."""
    try:
        with open(TRAFFIC_FILE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return {}
    e = table.get(key)
    if not e or e.get("kernel_sources") != kernel_source_hash():
        return {}
    return e


def roofline_of(kernel, samples, bpv, out_bytes, ms, key, shaded_fraction=None, floor_bytes=None):
    """Roofline object of one launch of ``kernel``: measured HBM bytes per
    launch over its duration against HBM peak (``frac``), the algorithmic
    bytes beside it (``alg``), and the VALU issue roof (``valu``) from the
    SQ counter pass."""
    t = ms * 1e-3
    alg_bytes = samples * bpv + out_bytes
    prof = measured_profile(key)
    traffic = None
    if "read_bytes" in prof:
        traffic = prof["read_bytes"] + prof["write_bytes"]
    achieved = traffic / t if traffic is not None else None
    out = {
        "bound": "hbm", "achieved": None if achieved is None else achieved / 1e9, "peak": HBM_PEAK / 1e9,
        "unit": "GB/s", "frac": None if achieved is None else achieved / HBM_PEAK, "traffic": traffic,
        "kernel": kernel, "avg_kernel_ms": ms, "profile_key": key, "profile_source": prof.get("source"),
        "alg": {"bytes_per_sample": bpv, "bytes_per_launch": alg_bytes, "achieved": alg_bytes / t / 1e9,
                "frac": alg_bytes / t / HBM_PEAK},
        "limiter": None,
    }
    if shaded_fraction is not None:
        # the same count over the samples that load texels (those in
        # transparent bricks are stepped over), from the counting pass
        shaded = samples * shaded_fraction * bpv + out_bytes
        out["alg"].update({"shaded_fraction": shaded_fraction, "shaded_bytes_per_launch": shaded,
                           "shaded_achieved": shaded / t / 1e9, "shaded_frac": shaded / t / HBM_PEAK})
    if floor_bytes is not None:
        # compulsory traffic: the bricks the launch's samples read, once
        # each per frame, and the images
        out["floor"] = {"bytes_per_launch": floor_bytes, "achieved": floor_bytes / t / 1e9,
                        "frac": floor_bytes / t / HBM_PEAK,
                        "traffic_over_floor": None if traffic is None else traffic / floor_bytes}
    if "valu_per_sample" in prof:
        va = prof["valu_per_sample"] * samples / t
        busy, dual = prof.get("valu_busy_counter"), prof.get("valu_dual_issue_frac")
        simd = busy * (1.0 - dual) if busy is not None and dual is not None else None
        out["valu"] = {"instr_per_sample": prof["valu_per_sample"], "achieved": va / 1e12, "peak": VALU_PEAK / 1e12,
                       "unit": "T lane-instructions/s", "frac": va / VALU_PEAK, "busy_counter": busy,
                       "dual_issue_frac": dual, "simd_issue_frac": simd, "wave_wait_frac": prof.get("wave_wait_frac")}
        vf = simd if simd is not None else va / VALU_PEAK
        out["limiter"] = "valu" if vf > (achieved or 0) / HBM_PEAK else "hbm"
    for k in ("l2_hit", "resident_waves_frac"):
        if k in prof:
            out[k] = prof[k]
    return out


# ---------------------------------------------------------------------------
# CPU baseline

def cpu_baseline(wl: Workload, data, normals, seconds):
    """The oracle's FP32 port of the shader (reference loop to max_steps) on
    all host cores: whole frames of the workload's cameras, in order, until
    ``seconds`` have elapsed."""
    import oracle
    import scenes

    oracle.build()
    cores = len(os.sched_getaffinity(0))
    from pyvr_amd.hip_renderer.batch import rgba_lut

    lut = rgba_lut(wl.ctf, wl.otf)
    samples, frames, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        q = wl.requests()[frames % len(wl.cams)]
        u = scenes.request_uniforms(q, wl.volume)
        out = oracle.render(u, data, normals if normals is not None else data, lut, mode="f32", full_loop=True,
                            nthreads=cores, want=("counts",))
        samples += int(out["counts"].sum())
        frames += 1
    dt = time.perf_counter() - t0
    log(f"cpu baseline: {frames} frames, {samples} samples in {dt:.1f}s on {cores} threads")
    return {
        "value": samples / dt, "unit": "samples/s", "cores": cores, "kind": "port",
        "sample": (f"{frames} whole {wl.width}x{wl.height} frames (cameras 0..{frames - 1} of the workload, "
                   f"cycling) on the volume's stored values read back from the device, FP32 shader arithmetic, "
                   f"reference loop to max_steps, {samples} samples in {dt:.1f}s; Mesa llvmpipe is not "
                   "available on the host (no EGL/OSMesa/X server), so this C port of the shader stands in "
                   "for it"),
    }


# ---------------------------------------------------------------------------
# legs

def volume_renderer(wl: Workload, stored, sf, nf):
    """A VolumeRenderer of the workload's size holding the stored values of
    the workload's volume (read back from the device: FP16 values are
    re-stored exactly; SNORM8 normals come back as integers and are scaled to
    [-1, 1] so that they re-quantise to the same integers), with the
    workload's transfer functions and termination threshold."""
    from pyvr_amd.hip_renderer import VolumeRenderer
    from pyvr_amd.volume import Volume

    scalar, normals = stored
    if normals is not None and nf == "sn8x3":
        normals = normals / np.float32(127.0)
    vr = VolumeRenderer(wl.width, wl.height, config=wl.config, light=wl.light, scalar_format=sf, normal_format=nf)
    vr.load_volume(Volume(data=scalar, normals=normals, min_bounds=np.asarray(wl.volume.min_bounds),
                          max_bounds=np.asarray(wl.volume.max_bounds)))
    vr.set_transfer_functions(wl.ctf, wl.otf)
    vr.set_opacity_threshold(wl.tau)
    return vr


def plain_leg(wl: Workload, args, sf, nf):
    """One launch per frame (ray setup + march) into resident device
    buffers, back to back on one stream; then end-to-end per-frame calls."""
    from pyvr_amd.hip_renderer import _lib
    from pyvr_amd.hip_renderer.plain import PlainRenderer

    L = _lib.lib()
    lb = args.brick_log2 if args.brick_log2 >= 0 else 3
    pr = PlainRenderer(scalar_format=sf, normal_format=nf, brick_log2=lb)
    ctx = pr.ctx
    _lib.check(L.hvr_set_empty_skip(ctx, {"off": _lib.SKIP_OFF, "on": _lib.SKIP_ON, "auto": _lib.SKIP_AUTO}[args.skip]),
               "set_empty_skip")
    _lib.check(L.hvr_set_dispatch_order(ctx, ORDER_MODES[args.order]), "set_dispatch_order")
    t0 = time.perf_counter()
    pr.upload(wl.volume)
    vol_bytes = pr.resident_bytes()
    log(f"volume resident: {vol_bytes / 2**20:.0f} MiB in {time.perf_counter() - t0:.1f}s")
    params, counts, shaded, touched = [], [], [], []
    for cam in wl.cams:
        pr.prepare(wl.volume, cam, wl.ctf, wl.otf, wl.width, wl.height, wl.config, wl.light, wl.tau)
        params.append(pr.params)
        counts.append(int(pr.sample_counts().sum()))
        shaded.append(int(pr.shaded_sample_counts().sum()))
        touched.append(pr.touched_bytes())
    s_step = sum(counts)
    shaded_fraction = sum(shaded) / max(1, s_step)
    npix = wl.width * wl.height
    # compulsory floor: each frame reads the bricks its samples load from
    # once, and writes its image
    floor_frames = [t + 4 * npix for t in touched]
    log(f"samples per step: {s_step} ({len(wl.cams)} frames)")
    bufs = []
    for _ in wl.cams:
        p = ctypes.c_void_p()
        _lib.check(L.hvr_device_alloc(ctx, npix * 4, ctypes.byref(p)), "alloc")
        bufs.append(p)

    def step():
        for p, b in zip(params, bufs):
            _lib.check(L.hvr_set_params(ctx, ctypes.byref(p)), "set_params")
            _lib.check(L.hvr_render_async_to(ctx, b), "render")

    if args.copy_calibration_mib > 0:
        nb = args.copy_calibration_mib << 20
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(L.hvr_device_alloc(ctx, nb, ctypes.byref(a)), "alloc")
        _lib.check(L.hvr_device_alloc(ctx, nb, ctypes.byref(b)), "alloc")
        _lib.check(L.hvr_device_copy(ctx, b, a, nb), "copy")
        L.hvr_device_free(ctx, a)
        L.hvr_device_free(ctx, b)
        log(f"calibration copy: {args.copy_calibration_mib} MiB device-to-device")

    for _ in range(args.warmup):
        step()
    _lib.check(L.hvr_synchronize(ctx), "sync")
    _lib.check(L.hvr_timing_begin(ctx, args.steps * len(wl.cams)), "timing")
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step()
    _lib.check(L.hvr_synchronize(ctx), "sync")
    wall = time.perf_counter() - t_start
    march_ms, setup_ms, nfr = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    _lib.check(L.hvr_timing_end(ctx, ctypes.byref(march_ms), ctypes.byref(setup_ms), ctypes.byref(nfr)),
               "timing_end")
    value = s_step * args.steps / wall
    ms_step = wall * 1e3 / args.steps
    log(f"plain: {args.steps} steps, {ms_step:.2f} ms/step, {value / 1e9:.2f} G samples/s; march kernel avg "
        f"{march_ms.value / nfr.value:.3f} ms, setup {setup_ms.value / nfr.value:.3f} ms")
    clear_b, total_b, used = ctypes.c_ulonglong(), ctypes.c_ulonglong(), ctypes.c_int()
    _lib.check(L.hvr_skip_info(ctx, ctypes.byref(clear_b), ctypes.byref(total_b), ctypes.byref(used)), "skip_info")
    skip = {"mode": args.skip, "used": bool(used.value), "clear_bricks": clear_b.value, "bricks": total_b.value}
    ordered = ctypes.c_int()
    _lib.check(L.hvr_dispatch_order_info(ctx, ctypes.byref(ordered)), "dispatch_order_info")
    order = {"mode": args.order, "used": bool(ordered.value), "readback": args.readback}
    for b in bufs:
        L.hvr_device_free(ctx, b)

    # the volume's stored values, read back (for the per-frame API and the
    # CPU baseline)
    stored = pr.download(wl.volume)
    pr.close()
    # end-to-end frames through the reference API: VolumeRenderer.set_camera
    # + render() (camera update, draw, RGBA8 copy into a new bytes object),
    # on the same stored texels
    lat = []
    vr = volume_renderer(wl, stored, sf, nf)
    _lib.check(L.hvr_set_dispatch_order(vr.gl_manager.ctx, ORDER_MODES[args.order]), "set_dispatch_order")
    _lib.check(L.hvr_set_readback(vr.gl_manager.ctx, READBACK_MODES[args.readback]), "set_readback")
    try:
        for i in range(args.latency_frames + 2):
            cam = wl.cams[i % len(wl.cams)]
            t = time.perf_counter()
            vr.set_camera(cam)
            vr.render()
            if i >= 2:
                lat.append((time.perf_counter() - t) * 1e3)
    finally:
        vr.close()
    lat = np.array(lat)
    bpv = 8 * (SCALAR_BYTES[sf] + (NORMAL_BYTES[nf] if wl.has_normals else 0))
    key = (f"{wl.name}/{len(wl.cams)}cams/{sf}+{nf if wl.has_normals else 'none'}/plain/lb{lb}"
           f"/skip{int(used.value)}/order{int(ordered.value)}")
    roofline = roofline_of("hvr::k_march", s_step / len(wl.cams), bpv, 4 * npix, march_ms.value / nfr.value, key,
                           shaded_fraction, sum(floor_frames) / len(floor_frames))
    return {"value": value, "ms_per_step": ms_step, "samples": s_step, "march_kernel_ms": march_ms.value / nfr.value,
            "fps": 1e3 / float(lat.mean()), "frame_ms_p50": float(np.percentile(lat, 50)),
            "frame_ms_p99": float(np.percentile(lat, 99)), "frames_per_s_device": len(wl.cams) * args.steps / wall,
            "roofline": roofline, "shaded_fraction": shaded_fraction, "floor_bytes_per_step": sum(floor_frames),
            "brick_skip": skip, "dispatch_order": order, "volume_bytes": vol_bytes, "bpv": bpv}, stored


def batched_leg(wl: Workload, args, sf, nf, s_step):
    """The workload's cameras as one batch of requests (BatchRenderer: one
    persistent graph-captured launch per run, brick size and ray tile from
    the autotuner), two runs in flight on their own streams; every run ends
    with all its images in pinned host memory."""
    from pyvr_amd.hip_renderer import BatchRenderer

    reqs = wl.requests()
    vols = [wl.volume]
    tuning = None
    if args.tile_shape == "auto":
        from pyvr_amd.hip_renderer.autotune import tune

        t0 = time.perf_counter()
        bricks = (args.brick_log2,) if args.brick_log2 >= 0 else (3, 4)
        tuning = tune(vols, reqs, sf, nf, bricks=bricks)
        lb, tile, kernel = tuning["brick_log2"], tuning["tile_shape"], tuning["kernel"]
        log(f"autotuner ({time.perf_counter() - t0:.1f}s, {len(tuning['table'])} candidates): " +
            ", ".join(f"lb{a}/{t}/{k} {ms:.2f}" for a, t, k, ms in tuning["table"]) + f" ms -> lb{lb}/{tile}/{kernel}")
    else:
        lb, tile, kernel = (args.brick_log2 if args.brick_log2 >= 0 else 3), args.tile_shape, args.kernel
    b = BatchRenderer(scalar_format=sf, normal_format=nf, brick_log2=lb, tile_shape=tile, pipeline=2, kernel=kernel,
                      readback=args.batch_readback)
    try:
        for q in reqs:
            b.add(vols[0], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
        b.run(count=True)
        samples = sum(int(b.counts(i).sum()) for i in range(len(reqs)))
        if samples != s_step:
            log(f"warning: batched sample count {samples} differs from the single-frame count {s_step}")
        for _ in range(args.warmup):
            b.run()
        lat, dev_ms = [], []
        for _ in range(args.steps):
            t = time.perf_counter()
            b.run()
            lat.append((time.perf_counter() - t) * 1e3)
            dev_ms.append(b.last_ms())
        s0 = b.submit()
        for q in reqs:
            b.add(vols[0], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
        s1 = b.submit()
        b.wait(s0)
        b.wait(s1)
        for k in range(args.warmup):
            b.rerun((s0, s1)[k % 2])
        b.wait(s0)
        b.wait(s1)
        t_start = time.perf_counter()
        for k in range(args.steps):
            b.rerun((s0, s1)[k % 2])
        b.wait(s0)
        b.wait(s1)
        wall = time.perf_counter() - t_start
    finally:
        b.close()
    value = samples * args.steps / wall
    dev = float(np.mean(dev_ms))
    log(f"batched: {wall * 1e3 / args.steps:.2f} ms/step two in flight, {value / 1e9:.2f} G samples/s; "
        f"one at a time {np.mean(lat):.2f} ms/step (graph on the device {dev:.2f} ms)")
    return {"value": value, "ms_per_step": wall * 1e3 / args.steps, "samples": samples,
            "frames_per_s": len(reqs) * args.steps / wall, "device_ms_per_run": dev,
            "batch_ms_p50": float(np.percentile(lat, 50)), "batch_ms_p99": float(np.percentile(lat, 99)),
            "brick_log2": lb, "tile_shape": tile,
            "kernel": kernel, "readback": args.batch_readback,
            "autotune": None if tuning is None else [{"brick_log2": a, "tile_shape": t, "kernel": k, "device_ms": ms}
                                                     for a, t, k, ms in tuning["table"]]}


def serving_leg(args, rates, duration, nvol, plain_requests):
    """Serving: ``nvol`` resident FP16 blob volumes of 256^3 to 1024^3
    (generated on the device), requests with Zipf(1.1) volume popularity,
    square images of 512^2 to 2048^2 and the fast / balanced / high_quality
    presets, arriving as a Poisson process at each rate.  Latency runs from
    a request's (scheduled) arrival to its pixels being in pinned host
    memory; throughput is the requests' samples over the time from the first
    arrival to the last completion.  Samples come from a counting pass over
    the same requests before the timed region.  Plain: the same requests one
    per launch through the single-frame kernels, synchronously (each image
    read back before the next request), on the same volumes."""
    import scenes
    from pyvr_amd.hip_renderer.plain import PlainRenderer
    from pyvr_amd.hip_renderer.serve import RenderServer, ServeRequest

    sf, nf = args.serve_scalar_format, args.serve_normal_format
    lb = args.brick_log2 if args.brick_log2 >= 0 else 3
    tile = args.tile_shape if args.tile_shape != "auto" else "8x8"
    n_max = max(50, int(max(rates) * duration), args.serve_saturation)
    vols = scenes.c4_volumes(nvol)
    reqs = scenes.c4_requests(n_max, nvol=nvol)
    srv = RenderServer(scalar_format=sf, normal_format=nf, brick_log2=lb, tile_shape=tile,
                       max_batch_requests=args.serve_max_requests, max_batch_pixels=1 << 26,
                       pipeline=args.serve_pipeline,
                       keep_results=False, max_batch_samples=args.serve_max_samples or None,
                       readback=args.batch_readback)
    t0 = time.perf_counter()
    for k, v in enumerate(vols):
        srv.register(k, v)
    srv.preload()
    log(f"serving: {len(srv.resident_keys())} volumes resident, {srv.resident_bytes() / 2**30:.1f} GiB, "
        f"built in {time.perf_counter() - t0:.1f}s")
    b = srv.backend
    samples = np.zeros(len(reqs), np.int64)
    t0 = time.perf_counter()
    for c0 in range(0, len(reqs), 64):
        chunk = reqs[c0:c0 + 64]
        for q in chunk:
            b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
        b.run(count=True)
        for i in range(len(chunk)):
            samples[c0 + i] = int(b.counts(i).sum())
        b.clear()
    log(f"serving: counted {len(reqs)} requests in {time.perf_counter() - t0:.1f}s, "
        f"{samples.mean() / 1e6:.0f} M samples per request on average")

    def run_rate(rate, n):
        """Poisson arrivals at ``rate`` req/s, or every request at once
        (``rate`` None: the server's capacity)."""
        if rate is None:
            arr = np.zeros(n)
        else:
            rng = np.random.default_rng(int(rate) + 7)
            arr = np.cumsum(rng.exponential(1.0 / rate, n))
        pending = [ServeRequest(volume=q.volume, camera=q.camera, ctf=q.ctf, otf=q.otf, width=q.width,
                                height=q.height, config=q.config, light=q.light, tau=q.tau, tag=i)
                   for i, q in enumerate(reqs[:n])]
        done = []
        srv.batches = 0
        start = time.perf_counter()
        i = 0
        while len(done) < n:
            now = time.perf_counter()
            while i < n and start + arr[i] <= now:
                pending[i].arrival = start + arr[i]
                srv.submit(pending[i])
                i += 1
            done.extend(srv.poll())
        lat = np.array([r.latency for r in done]) * 1e3
        end = max(r.done for r in done)
        tot = int(samples[:n].sum())
        slow = sorted(done, key=lambda r: -r.latency)[:5]
        return {"rate": "saturation" if rate is None else rate, "requests": n,
                "slowest": [{"arrival_s": r.arrival - start, "queue_ms": (r.launched - r.arrival) * 1e3,
                             "run_ms": (r.done - r.launched) * 1e3, "size": [r.width, r.height],
                             "step": float(r.config.step_size)} for r in slow],
                "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                "samples_per_s": tot / (end - start), "achieved_req_per_s": n / (end - start),
                "batches": srv.batches}

    run_rate(max(rates), min(64, n_max))      # warm-up: graphs for the shapes in use
    srv.freeze_heap()                         # no full garbage collection inside the timed runs
    table = []
    for rate in rates:
        n = max(50, int(rate * duration))
        row = run_rate(rate, n)
        table.append(row)
        log(f"serving {rate:.0f}/s: {n} requests, p50 {row['p50_ms']:.2f} ms, p99 {row['p99_ms']:.2f} ms, "
            f"{row['samples_per_s'] / 1e9:.1f} G samples/s, {row['achieved_req_per_s']:.0f} req/s, "
            f"{row['batches']} batches")
    sat = None
    if args.serve_saturation > 0:
        # capacity: the first requests submitted at once, the queue drained
        # as fast as the server batches them
        sat = run_rate(None, args.serve_saturation)
        log(f"serving at saturation: {sat['requests']} requests at once, {sat['samples_per_s'] / 1e9:.1f} G samples/s, "
            f"{sat['achieved_req_per_s']:.0f} req/s, {sat['batches']} batches")
    srv.close()

    # plain: one request per launch, synchronously, same requests and
    # volumes (all resident before timing), each image copied to pinned host
    # memory as a served one is
    pr = PlainRenderer(scalar_format=sf, normal_format=nf, brick_log2=lb)
    n_plain = min(len(reqs), plain_requests)
    lat = []
    try:
        for v in vols:
            pr.upload(v)
        for q in reqs[:min(8, n_plain)]:      # warm-up (kernels loaded, buffers at size)
            pr.prepare(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
            pr.read_pinned()
        for q in reqs[:n_plain]:
            t = time.perf_counter()
            pr.prepare(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
            pr.read_pinned()
            lat.append(time.perf_counter() - t)
    finally:
        pr.close()
    lat = np.array(lat)
    plain = {"requests": n_plain, "p50_ms": float(np.percentile(lat, 50) * 1e3),
             "p99_ms": float(np.percentile(lat, 99) * 1e3),
             "samples_per_s": float(samples[:n_plain].sum() / lat.sum()),
             "path": "PlainRenderer: single-frame kernels, one request per launch, image to pinned host memory"}
    log(f"serving plain: p50 {plain['p50_ms']:.2f} ms, p99 {plain['p99_ms']:.2f} ms, "
        f"{plain['samples_per_s'] / 1e9:.1f} G samples/s")
    best = sat if sat is not None else max(table, key=lambda r: r["samples_per_s"])
    return {"volumes": nvol, "scalar_format": sf, "normal_format": nf, "brick_log2": lb, "tile_shape": tile,
            "readback": args.batch_readback,
            "max_batch_samples": args.serve_max_samples, "pipeline": args.serve_pipeline,
            "mean_samples_per_request": float(samples.mean()),
            "rates": table, "saturation": sat, "plain": plain,
            "served_vs_plain": best["samples_per_s"] / plain["samples_per_s"]}


# ---------------------------------------------------------------------------
# c3: one batch of 64 mixed requests

def bench_batch(args):
    """c3: one batch of ``--requests`` requests (FP16 volumes of 64^3,
    128^3, 256^3 and 96x160x224 with SNORM8x3 normals, images of 256^2 to
    1024^2, own cameras and transfer functions) rendered by one persistent
    graph-captured launch per run, against the same requests one per launch
    through the single-frame kernels (plain_kernel: launches back to back
    into resident buffers; plain_api: each image read back before the next
    request)."""
    import scenes
    from pyvr_amd.hip_renderer import BatchRenderer, _lib
    from pyvr_amd.hip_renderer.plain import PlainRenderer

    sf, nf = args.scalar_format, args.normal_format
    t0 = time.perf_counter()
    vols = scenes.c3_volumes()
    reqs = scenes.c3_requests(args.requests)
    log(f"c3: {len(vols)} volumes, {len(reqs)} requests generated in {time.perf_counter() - t0:.1f}s")
    L = _lib.lib()
    tuning = None
    if args.tile_shape == "auto":
        from pyvr_amd.hip_renderer.autotune import tune

        t0 = time.perf_counter()
        bricks = (args.brick_log2,) if args.brick_log2 >= 0 else (3, 4)
        tuning = tune(vols, reqs, sf, nf, bricks=bricks)
        lb, tile, kernel = tuning["brick_log2"], tuning["tile_shape"], tuning["kernel"]
        log(f"autotuner ({time.perf_counter() - t0:.1f}s, {len(tuning['table'])} candidates): " +
            ", ".join(f"lb{a}/{t}/{k} {ms:.2f}" for a, t, k, ms in tuning["table"]) + f" ms -> lb{lb}/{tile}/{kernel}")
    else:
        lb, tile, kernel = (args.brick_log2 if args.brick_log2 >= 0 else 3), args.tile_shape, args.kernel

    b = BatchRenderer(scalar_format=sf, normal_format=nf, brick_log2=lb, tile_shape=tile, pipeline=2, kernel=kernel,
                      readback=args.batch_readback)
    for q in reqs:
        b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
    b.run(count=True)
    samples = sum(int(b.counts(i).sum()) for i in range(len(reqs)))
    pixels = sum(q.width * q.height for q in reqs)
    log(f"batch: {samples} samples, {pixels} pixels per run; resident {b.resident_bytes() / 2**20:.0f} MiB")
    for _ in range(args.warmup):
        b.run()
    dev_ms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        b.run()
        dev_ms.append(b.last_ms())
    wall_sync = time.perf_counter() - t_start
    value_sync = samples * args.steps / wall_sync
    log(f"batched, one at a time: {wall_sync * 1e3 / args.steps:.2f} ms/run, {value_sync / 1e9:.2f} G samples/s "
        f"(device {np.mean(dev_ms):.2f} ms/run)")
    s0 = b.submit()
    for q in reqs:
        b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
    s1 = b.submit()
    b.wait(s0)
    b.wait(s1)
    for k in range(args.warmup):
        b.rerun((s0, s1)[k % 2])
    b.wait(s0)
    b.wait(s1)
    t_start = time.perf_counter()
    for k in range(args.steps):
        b.rerun((s0, s1)[k % 2])
    b.wait(s0)
    b.wait(s1)
    wall = time.perf_counter() - t_start
    value = samples * args.steps / wall
    log(f"batched, two in flight: {wall * 1e3 / args.steps:.2f} ms/run, {value / 1e9:.2f} G samples/s")
    b.close()

    pr = PlainRenderer(scalar_format=sf, normal_format=nf, brick_log2=lb)
    bufs = {}
    for q in reqs:
        key = (q.width, q.height)
        if key not in bufs:
            p = ctypes.c_void_p()
            _lib.check(L.hvr_device_alloc(pr.ctx, q.width * q.height * 4, ctypes.byref(p)), "alloc")
            bufs[key] = p

    def plain(read):
        for q in reqs:
            if read:
                pr.render(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
            else:
                pr.prepare(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
                pr.launch_to(bufs[q.width, q.height])
        _lib.check(L.hvr_synchronize(pr.ctx), "sync")

    res = {}
    for name, read in (("plain_kernel", False), ("plain_api", True)):
        for _ in range(max(1, args.warmup)):
            plain(read)
        t = time.perf_counter()
        for _ in range(args.steps):
            plain(read)
        dt = time.perf_counter() - t
        res[name] = samples * args.steps / dt
        log(f"{name}: {dt * 1e3 / args.steps:.2f} ms/run, {res[name] / 1e9:.2f} G samples/s")
    for p in bufs.values():
        L.hvr_device_free(pr.ctx, p)
    pr.close()
    line = {
        "metric": METRIC, "value": value, "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"c3: {len(reqs)} batched requests (FP16 blob volumes 64^3-256^3 and "
                               f"96x160x224, images 256^2-1024^2), one persistent graph launch per run",
                   "requests": len(reqs), "pixels_per_step": pixels, "samples_per_step": samples,
                   "scalar_format": sf, "normal_format": nf, "brick_log2": lb, "tile_shape": tile,
                   "kernel": kernel, "autotuned": tuning is not None, "readback": args.batch_readback},
        "autotune": None if tuning is None else [{"brick_log2": a, "tile_shape": t, "kernel": k, "device_ms": ms}
                                                 for a, t, k, ms in tuning["table"]],
        "batch_device_ms": float(np.mean(dev_ms)), "batched_one_at_a_time": value_sync,
        "plain_kernel": res["plain_kernel"], "plain_api": res["plain_api"],
        "batched_vs_plain_kernel": value / res["plain_kernel"],
        "batched_vs_plain_api": value / res["plain_api"],
    }
    print(json.dumps(line), flush=True)


def bench_serve(args):
    """c4 on its own: the serving leg over the full rate sweep."""
    rates = [float(r) for r in args.rates.split(",")]
    srv = serving_leg(args, rates, args.duration, args.volumes, 200)
    best = srv["saturation"] or max(srv["rates"], key=lambda r: r["samples_per_s"])
    line = {
        "metric": METRIC, "value": best["samples_per_s"], "unit": "samples/s", "n_gpus": 1,
        "steps": len(rates), "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"c4: serving loop, {args.volumes} resident FP16 blob volumes 256^3-1024^3, "
                               "Poisson arrivals, images 512^2-2048^2, fast/balanced/high_quality"},
        "serving": srv,
    }
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------

def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=["c0", "c1", "c2", "c3", "c4"])
    ap.add_argument("--rates", default="50,100,200,500,1000,2000", help="c4: Poisson arrival rates (req/s)")
    ap.add_argument("--duration", type=float, default=2.0, help="c4: seconds of arrivals per rate")
    ap.add_argument("--volumes", type=int, default=16, help="serving: resident volumes")
    ap.add_argument("--requests", type=int, default=64, help="c3: requests per batch")
    ap.add_argument("--tile-shape", default="auto", choices=["auto", "8x8", "16x4"],
                    help="ray tile of a wavefront (auto: the autotuner picks it and the brick size)")
    ap.add_argument("--kernel", default="default", help="batched kernel variant when --tile-shape is not auto "
                    "(default, waves2, waves8, nogate, noskip)")
    ap.add_argument("--size", type=int, default=0)
    ap.add_argument("--image", type=int, default=0)
    ap.add_argument("--cameras", type=int, default=24)
    ap.add_argument("--scalar-format", default=None, choices=["f32", "f16", "u8"],
                    help="volume storage (default: f32 for c0 and c1, whose configurations store float32 "
                         "volumes; f16 otherwise)")
    ap.add_argument("--normal-format", default=None, choices=["f32x3", "sn8x3", "f16x3"],
                    help="normal storage (default: f32x3 for c0 and c1, f16x3 otherwise)")
    ap.add_argument("--serve-scalar-format", default="f16", choices=["f32", "f16", "u8"])
    ap.add_argument("--serve-normal-format", default="sn8x3", choices=["f32x3", "sn8x3", "f16x3"])
    ap.add_argument("--serve-rates", default="500,1000", help="serving leg of the default run (req/s)")
    ap.add_argument("--serve-duration", type=float, default=1.0, help="serving leg: seconds of arrivals per rate")
    ap.add_argument("--serve-max-requests", type=int, default=64)
    ap.add_argument("--serve-pipeline", type=int, default=3, help="serving: batches in flight")
    ap.add_argument("--serve-saturation", type=int, default=400,
                    help="serving: requests submitted at once to measure capacity (0: skip)")
    ap.add_argument("--serve-max-samples", type=float, default=2e9,
                    help="serving: bound on a batch's estimated samples (0: requests and pixels only)")
    ap.add_argument("--no-serve", action="store_true", help="default run without the serving leg")
    ap.add_argument("--brick-log2", type=int, default=-1)
    ap.add_argument("--skip", choices=("auto", "on", "off"), default="auto",
                    help="transparent-brick skipping mode (result is identical in every mode)")
    ap.add_argument("--order", choices=tuple(ORDER_MODES), default="auto",
                    help="single-frame dispatch order: cost-ordered blocks (on), row order (off), or by frame size")
    ap.add_argument("--readback", choices=tuple(READBACK_MODES), default="direct",
                    help="per-frame latency leg: frames copied out of the device framebuffer, or written by the "
                         "march straight into pinned host memory")
    ap.add_argument("--batch-readback", choices=tuple(READBACK_MODES), default="direct",
                    help="batched and serving legs: images copied out of a device pool after each launch, or "
                         "written by the march straight into pinned host memory")
    ap.add_argument("--latency-frames", type=int, default=48)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-only", action="store_true",
                    help="c2: time only the one-frame-per-launch renderer (no batched or serving leg)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the timed region, copy this many MiB device-to-device once "
                         "(a known-size kernel for calibrating traffic counters)")
    args = ap.parse_args()
    fp32 = args.workload in ("c0", "c1")
    args.scalar_format = args.scalar_format or ("f32" if fp32 else "f16")
    args.normal_format = args.normal_format or ("f32x3" if fp32 else "f16x3")

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this renderer's "
              "multi-GPU split is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c3":
        return bench_batch(args)
    if args.workload == "c4":
        return bench_serve(args)

    sf, nf = args.scalar_format, args.normal_format
    wl = make_workload(args)
    log(f"workload {wl.name}: {len(wl.cams)} cameras, {wl.width}x{wl.height}, volume {wl.shape}")
    plain, stored = plain_leg(wl, args, sf, nf)
    s_step, bpv, npix = plain["samples"], plain["bpv"], wl.width * wl.height

    batched = None
    roofline = plain["roofline"]
    if args.workload == "c2" and not args.plain_only:
        batched = batched_leg(wl, args, sf, nf, s_step)
        bkey = (f"{wl.name}/{len(wl.cams)}cams/{sf}+{nf}/batch/lb{batched['brick_log2']}/{batched['tile_shape']}"
                f"/{batched['kernel']}/{batched['readback']}")
        # the skipping variants step over the same samples as the single-frame
        # kernel; the noskip variant loads texels for every sample
        sh = plain["shaded_fraction"] if batched["kernel"] != "noskip" else 1.0
        roofline = roofline_of("hvr::k_batch (graph launch: staging copy, counter fill, brick flags, k_batch)",
                               batched["samples"], bpv, 4 * npix * len(wl.cams), batched["device_ms_per_run"], bkey,
                               sh, plain["floor_bytes_per_step"])
    serving = None
    if args.workload == "c2" and not args.plain_only and not args.no_serve:
        serving = serving_leg(args, [float(r) for r in args.serve_rates.split(",")], args.serve_duration,
                              args.volumes, 100)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, stored[0], stored[1], args.cpu_seconds)

    plain_out = {k: plain[k] for k in ("value", "ms_per_step", "march_kernel_ms", "fps", "frame_ms_p50",
                                       "frame_ms_p99", "frames_per_s_device", "dispatch_order", "roofline")}
    if batched is not None:
        extra = {
            "fps": batched["frames_per_s"],
            "batch_ms_p50": batched["batch_ms_p50"], "batch_ms_p99": batched["batch_ms_p99"],
            "frame_ms_p50": plain["frame_ms_p50"], "frame_ms_p99": plain["frame_ms_p99"],
            "batched": {k: batched[k] for k in ("device_ms_per_run", "brick_log2", "tile_shape", "kernel",
                                                "autotune")},
            "plain_kernel": plain_out,
            "batched_vs_plain_kernel": batched["value"] / plain["value"],
        }
        value, ms_step = batched["value"], batched["ms_per_step"]
    else:
        extra = {"fps": plain["fps"], "frame_ms_p50": plain["frame_ms_p50"], "frame_ms_p99": plain["frame_ms_p99"],
                 "frames_per_s_device": plain["frames_per_s_device"]}
        value, ms_step = plain["value"], plain["ms_per_step"]
    if serving is not None:
        extra["serving"] = serving
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{wl.name}: {len(wl.cams)} cameras, {wl.width}x{wl.height}, "
                         f"{'x'.join(map(str, wl.shape))} volume"
                         + ("; the cameras batched into one persistent graph launch per step"
                            if batched is not None else "; one launch per frame")),
            "volume": list(wl.shape), "image": [wl.width, wl.height],
            "cameras": len(wl.cams), "preset": {"step_size": wl.config.step_size, "max_steps": wl.config.max_steps},
            "et_threshold": wl.tau, "scalar_format": sf, "normal_format": nf if wl.has_normals else "none",
            "samples_per_step": s_step, "volume_bytes": plain["volume_bytes"], "brick_skip": plain["brick_skip"],
        },
        **extra,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
