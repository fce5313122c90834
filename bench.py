#!/usr/bin/env python3
"""Benchmark of the HIP volume renderer on one MI355X.

Metric: ray samples/s (plus fps and p50/p99 latency) at 512^3 / 1024^2,
batched form against the plain one-request-per-launch kernel.  A *step*
renders the workload's camera set once (default: 24 cameras of the 512^3
blob volume at 1024x1024, high_quality preset, FP16 scalar + SNORM8x3
normals), with the volume already resident in HBM.

* plain leg: each frame one launch (ray setup + march) into its own resident
  device buffer, back to back on one stream; its figures are the
  ``plain_kernel`` object of the line (and the whole line for c0 / c1, or
  with --plain-only);
* batched leg (c2): the 24 cameras as 24 requests of one persistent
  graph-captured launch per step (BatchRenderer; brick size and ray tile
  from the autotuner), two steps in flight, every step ending with its 24
  images in pinned host memory; ``value`` is this leg's;
* samples: reference-algorithm samples (pixel, i) -- i < max_steps, alpha
  below the termination threshold, sample inside the box -- counted by the
  renderer's counting pass before the timed region (the tests pin that count
  to the oracle's);
* value = samples per step * steps / wall time of the timed steps;
* roofline: the dominant kernel (k_batch for the batched leg, k_march for
  the plain one) against HBM peak; algorithmic bytes per launch = samples *
  8 * (scalar + normal bytes per voxel, unpadded) + 4 * W * H per image,
  divided by the launch's average duration from HIP events recorded on its
  stream (for the batched leg: around the graph, which also holds the
  staging copy, a counter fill and the brick-flag kernel).  Most texel reads
  hit L1/L2, so that rate exceeds HBM peak; ``traffic`` and
  ``hbm_measured_*`` give the HBM bytes rocprofv3 measured for these
  sources, and ``profile`` the kernel's VALU-busy fraction (the limit);
* fps / latency: batched frames per second; ``batch_ms_p50/p99``: one step
  of the batch from launch to images in host memory; ``frame_ms_p50/p99``:
  end-to-end VolumeRenderer.render() (set_camera + draw + copy of the RGBA8
  image to the host) per frame;
* cpu_baseline: the oracle's FP32 port of the reference shader (its loop
  runs every i < max_steps, as the shader does) on all host cores, whole
  frames of the workload's cameras until ~15 s have elapsed.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
SCALAR_BYTES = {"f32": 4, "f16": 2, "u8": 1}
NORMAL_BYTES = {"f32x3": 12, "sn8x3": 3, "f16x3": 6, "none": 0}
METRIC = "ray samples/s + fps at 512^3/1024^2; p50/p99 frame latency vs plain kernel"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(args):
    import scenes

    if args.workload == "c2":
        sc = scenes.c2(size=args.size or 512, image=args.image or 1024)
        cams = scenes.c2_cameras(args.cameras)
    elif args.workload == "c1":
        sc = scenes.c1(size=args.size or 256, image=args.image or 1024)
        cams = [sc.camera] * args.cameras
    elif args.workload == "c0":
        sc = scenes.c0(size=args.size or 128, image=args.image or 512)
        cams = [sc.camera] * args.cameras
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    return sc, cams


def kernel_source_hash() -> str:
    """Hash of the device sources: measured traffic is only reported for the
    kernels it was measured on."""
    import hashlib

    h = hashlib.sha256()
    for rel in ("pyvr_amd/csrc/hvr_device.h", "pyvr_amd/csrc/hipvr.hip"):
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(key: str):
    """HBM bytes per launch from the rocprofv3 FETCH_SIZE/WRITE_SIZE passes
    recorded in profiles/traffic.json (scaled by the copy-kernel
    calibration), with the profile's other figures of that kernel (VALU
    busy, resident waves, L2 hit rate), or (None, None, {}) when no
    measurement matches this workload and these kernel sources."""
    try:
        with open(TRAFFIC_FILE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None, None, {}
    e = table.get(key)
    if not e or e.get("kernel_sources") != kernel_source_hash():
        return None, None, {}
    extra = {k: e[k] for k in ("valu_busy", "resident_waves_frac", "l2_hit") if k in e}
    return e["read_bytes"] + e["write_bytes"], e.get("source"), extra


def roofline_of(kernel, alg_bytes, ms, bpv, key):
    """The roofline object: algorithmic bytes (8 texel reads per sample at
    the stored format, plus the images) over the launch time against HBM
    peak, and, where a profile of these sources exists, the HBM bytes it
    measured and the kernel's issue figures.  Texels are re-read from L1/L2
    (neighbouring samples share taps), so the algorithmic rate can exceed
    HBM peak; the measured rate shows what HBM actually carried."""
    achieved = alg_bytes / (ms * 1e-3)
    traffic, src, prof = measured_traffic(key)
    out = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
           "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_source": src, "traffic_key": key,
           "alg_bytes_per_launch": alg_bytes, "bytes_per_sample": bpv, "kernel": kernel, "avg_kernel_ms": ms}
    if traffic is not None:
        hbm = traffic / (ms * 1e-3)
        out["hbm_measured_GBps"] = hbm / 1e9
        out["hbm_measured_frac"] = hbm / HBM_PEAK
        out["profile"] = prof
    return out


def cpu_baseline(sc, cams, args, sf, nf):
    """Oracle FP32 port of the shader (reference loop to max_steps) on all
    host cores: whole frames of the workload's cameras, in order, until
    ~cpu_seconds have elapsed."""
    import oracle
    import scenes

    oracle.build()
    cores = len(os.sched_getaffinity(0))
    data = scenes.dequant_scalar(sc.data, sf)
    nrm = scenes.dequant_normals(sc.normals, nf) if sc.normals is not None else data
    lut = sc.lut()
    samples, frames, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 < args.cpu_seconds:
        u = sc.uniforms(cams[frames % len(cams)])
        out = oracle.render(u, data, nrm, lut, mode="f32", full_loop=True, nthreads=cores,
                            want=("counts",))
        samples += int(out["counts"].sum())
        frames += 1
    dt = time.perf_counter() - t0
    log(f"cpu baseline: {frames} frames, {samples} samples in {dt:.1f}s on {cores} threads")
    return {
        "value": samples / dt, "unit": "samples/s", "cores": cores, "kind": "port",
        "sample": (f"{frames} whole {sc.width}x{sc.height} frames (cameras 0..{frames - 1} of the "
                   f"workload, cycling), FP32 shader arithmetic, reference loop to max_steps, "
                   f"{samples} samples in {dt:.1f}s; Mesa llvmpipe is not available on the host "
                   "(no EGL/OSMesa/X server), so this C port of the shader stands in for it"),
    }


def bench_batch(args):
    """c3: one batch of ``--requests`` requests (FP16 volumes of 64^3,
    128^3, 256^3 and 96x160x224 with SNORM8x3 normals, images of 256^2 to
    1024^2, own cameras and transfer functions) rendered by one persistent
    graph-captured launch per run, against the same requests rendered one
    per launch.  A step is one run of the whole batch, images copied to
    host memory.  Plain figures, same requests and resident volumes:
    plain_kernel -- back-to-back single-request launches on one stream
    (LUT upload, parameters, ray setup and march per request; one wait at
    the end); plain_api -- what a renderer user gets per frame (the same
    plus the synchronous copy of each image to the host)."""
    import scenes
    from pyvr_amd.hip_renderer import BatchRenderer, _lib
    from pyvr_amd.hip_renderer.batch import request_params, rgba_lut
    from pyvr_amd.hip_renderer.manager import _DeviceVolume

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
        lb, tile = tuning["brick_log2"], tuning["tile_shape"]
        log(f"autotuner ({time.perf_counter() - t0:.1f}s): " +
            ", ".join(f"lb{a}/{t} {ms:.2f} ms" for a, t, ms in tuning["table"]) + f" -> lb{lb}/{tile}")
    else:
        lb, tile = (args.brick_log2 if args.brick_log2 >= 0 else 3), args.tile_shape

    b = BatchRenderer(scalar_format=sf, normal_format=nf, brick_log2=lb, tile_shape=tile, pipeline=2)
    for q in reqs:
        b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
    b.run(count=True)
    samples = sum(int(b.counts(i).sum()) for i in range(len(reqs)))
    pixels = sum(q.width * q.height for q in reqs)
    log(f"batch: {samples} samples, {pixels} pixels per run; resident {b.resident_bytes() / 2**20:.0f} MiB")
    # one batch at a time: launch, wait for the images in host memory
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
    # two batches in flight (own streams): a run's copy back and tail
    # overlap the next run's work; every run still ends with its images in
    # host memory
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

    # plain: one context per image size, every volume resident in each
    ctxs, dvols = {}, {}
    for q in reqs:
        key = (q.width, q.height)
        if key in ctxs:
            continue
        c = ctypes.c_void_p()
        _lib.check(L.hvr_create(0, q.width, q.height, ctypes.byref(c)), "hvr_create")
        _lib.check(L.hvr_set_brick_log2(c, lb), "set_brick")
        ctxs[key] = c
        for k, v in enumerate(vols):
            dv = _DeviceVolume(L, c, v.data.shape, sf, nf if v.has_normals else "none")
            dv.upload_scalar(v.data)
            if v.has_normals:
                dv.upload_normals(v.normals)
            dvols[key, k] = dv
    bufs = {}
    for key, c in ctxs.items():
        p = ctypes.c_void_p()
        _lib.check(L.hvr_device_alloc(c, key[0] * key[1] * 4, ctypes.byref(p)), "alloc")
        bufs[key] = p
    plan = []
    for q in reqs:
        key = (q.width, q.height)
        lut = rgba_lut(q.ctf, q.otf)
        prm = request_params(vols[q.volume], q.camera, q.config, q.light, q.width, q.height, q.tau)
        plan.append((ctxs[key], dvols[key, q.volume].handle, lut, prm, bufs[key],
                     np.empty(q.width * q.height * 4, np.uint8)))

    def plain(read):
        for c, vh, lut, prm, buf, host in plan:
            nrole = _lib.HVR_ROLE_NORMAL
            _lib.check(L.hvr_bind(c, vh, _lib.HVR_ROLE_SCALAR, vh, nrole), "bind")
            _lib.check(L.hvr_set_lut(c, lut.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), lut.shape[0]), "lut")
            _lib.check(L.hvr_set_params(c, ctypes.byref(prm)), "params")
            if read:
                _lib.check(L.hvr_render(c, host.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))), "render")
            else:
                _lib.check(L.hvr_render_async_to(c, buf), "render")
        for c in ctxs.values():
            _lib.check(L.hvr_synchronize(c), "sync")

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
    for key, c in ctxs.items():
        L.hvr_device_free(c, bufs[key])
    for dv in dvols.values():
        dv.destroy()
    for c in ctxs.values():
        L.hvr_destroy(c)
    b.close()

    line = {
        "metric": METRIC, "value": value, "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"c3: {len(reqs)} batched requests (FP16 blob volumes 64^3-256^3 and "
                               f"96x160x224, images 256^2-1024^2), one persistent graph launch per run",
                   "requests": len(reqs), "pixels_per_step": pixels, "samples_per_step": samples,
                   "scalar_format": sf, "normal_format": nf, "brick_log2": lb, "tile_shape": tile,
                   "autotuned": tuning is not None},
        "autotune": None if tuning is None else [{"brick_log2": a, "tile_shape": t, "device_ms": ms}
                                                 for a, t, ms in tuning["table"]],
        "batch_device_ms": float(np.mean(dev_ms)), "batched_one_at_a_time": value_sync,
        "plain_kernel": res["plain_kernel"], "plain_api": res["plain_api"],
        "batched_vs_plain_kernel": value / res["plain_kernel"],
        "batched_vs_plain_api": value / res["plain_api"],
    }
    print(json.dumps(line), flush=True)


def batched_leg(sc, cams, args, sf, nf, s_step):
    """The workload's cameras as one batch of requests (BatchRenderer: one
    persistent graph-captured launch per run, brick size and ray tile from
    the autotuner), two runs in flight on their own streams; every run ends
    with all its images in pinned host memory.  Returns the figures for the
    JSON line."""
    import scenes
    from pyvr_amd.hip_renderer import BatchRenderer

    vols = [sc.volume()]
    reqs = [scenes.Request(0, cam, sc.ctf, sc.otf, sc.width, sc.height, sc.config, sc.light, sc.tau)
            for cam in cams]
    tuning = None
    if args.tile_shape == "auto":
        from pyvr_amd.hip_renderer.autotune import tune

        t0 = time.perf_counter()
        bricks = (args.brick_log2,) if args.brick_log2 >= 0 else (3, 4)
        tuning = tune(vols, reqs, sf, nf, bricks=bricks)
        lb, tile = tuning["brick_log2"], tuning["tile_shape"]
        log(f"autotuner ({time.perf_counter() - t0:.1f}s): " +
            ", ".join(f"lb{a}/{t} {ms:.2f} ms" for a, t, ms in tuning["table"]) + f" -> lb{lb}/{tile}")
    else:
        lb, tile = (args.brick_log2 if args.brick_log2 >= 0 else 3), args.tile_shape
    b = BatchRenderer(scalar_format=sf, normal_format=nf, brick_log2=lb, tile_shape=tile, pipeline=2)
    try:
        for q in reqs:
            b.add(vols[0], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
        b.run(count=True)
        samples = sum(int(b.counts(i).sum()) for i in range(len(reqs)))
        if samples != s_step:
            log(f"warning: batched sample count {samples} differs from the single-frame count {s_step}")
        # one run at a time: latency of a whole batch (launch to images in host memory)
        for _ in range(args.warmup):
            b.run()
        lat, dev_ms = [], []
        for _ in range(args.steps):
            t = time.perf_counter()
            b.run()
            lat.append((time.perf_counter() - t) * 1e3)
            dev_ms.append(b.last_ms())
        # two runs in flight: the timed region of the value
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
            "autotune": None if tuning is None else [{"brick_log2": a, "tile_shape": t, "device_ms": ms}
                                                     for a, t, ms in tuning["table"]]}


def bench_serve(args):
    """c4: a serving loop over 16 resident FP16+SNORM8x3 blob volumes of
    256^3 to 1024^3 (generated on the device), requests with Zipf(1.1)
    volume popularity, square images of 512^2 to 2048^2 and the fast /
    balanced / high_quality presets, arriving as a Poisson process at each
    rate.  Latency runs from a request's (scheduled) arrival to its pixels
    being in pinned host memory; throughput is the requests' samples over
    the time from the first arrival to the last completion.  Samples come
    from a counting pass over the same requests before the timed region."""
    import scenes
    from pyvr_amd.hip_renderer.serve import RenderServer, ServeRequest

    sf, nf = args.scalar_format, args.normal_format
    lb = args.brick_log2 if args.brick_log2 >= 0 else 3
    tile = args.tile_shape if args.tile_shape != "auto" else "8x8"
    rates = [float(r) for r in args.rates.split(",")]
    n_max = max(50, int(max(rates) * args.duration))
    vols = scenes.c4_volumes(args.volumes)
    reqs = scenes.c4_requests(n_max, nvol=args.volumes)
    srv = RenderServer(scalar_format=sf, normal_format=nf, brick_log2=lb, tile_shape=tile,
                       max_batch_requests=64, max_batch_pixels=1 << 26, pipeline=3, keep_results=False)
    t0 = time.perf_counter()
    for k, v in enumerate(vols):
        srv.register(k, v)
    srv.preload()
    log(f"c4: {len(srv.resident_keys())} volumes resident, {srv.resident_bytes() / 2**30:.1f} GiB, "
        f"built in {time.perf_counter() - t0:.1f}s")
    # counting pass (same kernels with sample counters), in chunks
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
    log(f"counted {len(reqs)} requests in {time.perf_counter() - t0:.1f}s: "
        f"{samples.mean() / 1e6:.0f} M samples per request on average")

    def run_rate(rate, n):
        rng = np.random.default_rng(int(rate) + 7)
        arr = np.cumsum(rng.exponential(1.0 / rate, n))
        pending = [ServeRequest(volume=q.volume, camera=q.camera, ctf=q.ctf, otf=q.otf, width=q.width,
                                height=q.height, config=q.config, light=q.light, tau=q.tau, tag=i)
                   for i, q in enumerate(reqs[:n])]
        done = []
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
        return {"rate": rate, "requests": n, "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                "samples_per_s": tot / (end - start), "achieved_req_per_s": n / (end - start),
                "batches": srv.batches}

    # plain: one request per launch, synchronously (no queueing), on the
    # same resident volumes and kernels
    n_plain = min(len(reqs), 200)
    plain_lat = []
    for q in reqs[:n_plain]:
        b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
        t = time.perf_counter()
        b.run()
        plain_lat.append(time.perf_counter() - t)
        b.clear()
    plain_lat = np.array(plain_lat)
    plain = {"requests": n_plain, "p50_ms": float(np.percentile(plain_lat, 50) * 1e3),
             "p99_ms": float(np.percentile(plain_lat, 99) * 1e3),
             "samples_per_s": float(samples[:n_plain].sum() / plain_lat.sum())}
    log(f"plain, one request per launch: p50 {plain['p50_ms']:.2f} ms, p99 {plain['p99_ms']:.2f} ms, "
        f"{plain['samples_per_s'] / 1e9:.1f} G samples/s")

    # warm-up (graph capture for the shapes in use)
    run_rate(max(rates), min(64, n_max))
    table = []
    for rate in rates:
        srv.batches = 0
        n = max(50, int(rate * args.duration))
        row = run_rate(rate, n)
        table.append(row)
        log(f"rate {rate:.0f}/s: {n} requests, p50 {row['p50_ms']:.2f} ms, p99 {row['p99_ms']:.2f} ms, "
            f"{row['samples_per_s'] / 1e9:.1f} G samples/s, {row['achieved_req_per_s']:.0f} req/s, "
            f"{row['batches']} batches")
    srv.close()
    best = max(table, key=lambda r: r["samples_per_s"])
    line = {
        "metric": METRIC, "value": best["samples_per_s"], "unit": "samples/s", "n_gpus": 1,
        "steps": len(rates), "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c4: serving loop, 16 resident FP16+SNORM8x3 blob volumes 256^3-1024^3, "
                               "Poisson arrivals, images 512^2-2048^2, fast/balanced/high_quality",
                   "volumes": args.volumes, "requests_max": n_max, "scalar_format": sf, "normal_format": nf,
                   "brick_log2": lb, "tile_shape": tile, "mean_samples_per_request": float(samples.mean())},
        "serving": table,
        "plain": plain,
        "served_vs_plain": best["samples_per_s"] / plain["samples_per_s"],
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=["c0", "c1", "c2", "c3", "c4"])
    ap.add_argument("--rates", default="50,100,200,500,1000,2000", help="c4: Poisson arrival rates (req/s)")
    ap.add_argument("--duration", type=float, default=2.0, help="c4: seconds of arrivals per rate")
    ap.add_argument("--volumes", type=int, default=16, help="c4: resident volumes")
    ap.add_argument("--requests", type=int, default=64, help="c3: requests per batch")
    ap.add_argument("--tile-shape", default="auto", choices=["auto", "8x8", "16x4"],
                    help="c3: ray tile of a wavefront (auto: the autotuner picks it and the brick size)")
    ap.add_argument("--size", type=int, default=0)
    ap.add_argument("--image", type=int, default=0)
    ap.add_argument("--cameras", type=int, default=24)
    ap.add_argument("--scalar-format", default="f16", choices=["f32", "f16", "u8"])
    ap.add_argument("--normal-format", default="sn8x3", choices=["f32x3", "sn8x3", "f16x3"])
    ap.add_argument("--brick-log2", type=int, default=-1)
    ap.add_argument("--skip", choices=("auto", "on", "off"), default="auto",
                    help="transparent-brick skipping mode (result is identical in every mode)")
    ap.add_argument("--latency-frames", type=int, default=48)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-only", action="store_true",
                    help="c2: time only the one-frame-per-launch renderer (no batched leg)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the timed region, copy this many MiB device-to-device once "
                         "(a known-size kernel for calibrating traffic counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this renderer's "
              "multi-GPU split is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c3":
        return bench_batch(args)
    if args.workload == "c4":
        return bench_serve(args)

    from pyvr_amd.hip_renderer import VolumeRenderer, _lib

    t0 = time.perf_counter()
    sc, cams = make_workload(args)
    sf, nf = args.scalar_format, args.normal_format
    log(f"workload {sc.name}: generated in {time.perf_counter() - t0:.1f}s")

    r = VolumeRenderer(sc.width, sc.height, config=sc.config, light=sc.light,
                       scalar_format=sf, normal_format=nf)
    L = _lib.lib()
    ctx = r.gl_manager.ctx
    if args.brick_log2 >= 0:
        _lib.check(L.hvr_set_brick_log2(ctx, args.brick_log2), "set_brick")
    _lib.check(L.hvr_set_empty_skip(ctx, {"off": _lib.SKIP_OFF, "on": _lib.SKIP_ON,
                                           "auto": _lib.SKIP_AUTO}[args.skip]), "set_empty_skip")
    t0 = time.perf_counter()
    r.load_volume(sc.volume())
    vol_bytes = sum(v.nbytes() for v in r.gl_manager._volumes)
    log(f"volume resident: {vol_bytes / 2**20:.0f} MiB in {time.perf_counter() - t0:.1f}s")
    r.set_transfer_functions(sc.ctf, sc.otf, sc.lut_size)
    r.set_opacity_threshold(sc.tau)

    # per-camera parameter blocks, exactly as set_camera + render would set them
    params, counts = [], []
    for cam in cams:
        r.set_camera(cam)
        r.gl_manager.clear_framebuffer(0, 0, 0, 0)
        r.gl_manager.setup_blending()
        r.gl_manager._sync_bindings()
        params.append(r.gl_manager._params())
        counts.append(int(r.sample_counts().sum()))
    s_step = sum(counts)
    npix = sc.width * sc.height
    log(f"samples per step: {s_step} ({len(cams)} frames)")

    bufs = []
    for _ in cams:
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
    _lib.check(L.hvr_timing_begin(ctx, args.steps * len(cams)), "timing")
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
    log(f"timed {args.steps} steps: {ms_step:.2f} ms/step, {value / 1e9:.2f} G samples/s; "
        f"march kernel avg {march_ms.value / nfr.value:.3f} ms, setup {setup_ms.value / nfr.value:.3f} ms")

    clear_b, total_b, used = ctypes.c_ulonglong(), ctypes.c_ulonglong(), ctypes.c_int()
    _lib.check(L.hvr_skip_info(ctx, ctypes.byref(clear_b), ctypes.byref(total_b), ctypes.byref(used)), "skip_info")
    skip = {"mode": args.skip, "used": bool(used.value), "clear_bricks": clear_b.value, "bricks": total_b.value}
    log(f"transparent bricks {clear_b.value}/{total_b.value}; stepped over: {bool(used.value)}")

    bpv = 8 * (SCALAR_BYTES[sf] + (NORMAL_BYTES[nf] if sc.normals is not None else 0))
    alg_bytes = s_step * args.steps * bpv + 4 * npix * nfr.value
    tkey = (f"{sc.name}/{len(cams)}cams/{sf}+{nf if sc.normals is not None else 'none'}/lb{args.brick_log2}"
            f"/skip{int(used.value)}")
    roofline = roofline_of("hvr::k_march", alg_bytes / nfr.value, march_ms.value / nfr.value, bpv, tkey)

    # end-to-end frames through the public API (set_camera + render + D2H)
    lat = []
    for i in range(args.latency_frames + 2):
        cam = cams[i % len(cams)]
        t = time.perf_counter()
        r.set_camera(cam)
        r.render()
        if i >= 2:
            lat.append((time.perf_counter() - t) * 1e3)
    lat = np.array(lat)

    for b in bufs:
        L.hvr_device_free(ctx, b)

    plain = {"value": value, "ms_per_step": ms_step, "march_kernel_ms": march_ms.value / nfr.value,
             "fps": 1e3 / float(lat.mean()), "frame_ms_p50": float(np.percentile(lat, 50)),
             "frame_ms_p99": float(np.percentile(lat, 99)),
             "frames_per_s_device": len(cams) * args.steps / wall, "roofline": roofline}
    r.close()

    batched = None
    if args.workload == "c2" and not args.plain_only:
        batched = batched_leg(sc, cams, args, sf, nf, s_step)
        alg_b = batched["samples"] * bpv + 4 * npix * len(cams)
        bkey = f"{sc.name}/{len(cams)}cams/{sf}+{nf}/batch/lb{batched['brick_log2']}/{batched['tile_shape']}"
        roofline = roofline_of("hvr::k_batch (graph launch: staging copy, counter fill, brick flags, k_batch)",
                               alg_b, batched["device_ms_per_run"], bpv, bkey)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sc, cams, args, sf, nf)

    if batched is not None:
        extra = {
            "fps": batched["frames_per_s"],
            "batch_ms_p50": batched["batch_ms_p50"], "batch_ms_p99": batched["batch_ms_p99"],
            "frame_ms_p50": plain["frame_ms_p50"], "frame_ms_p99": plain["frame_ms_p99"],
            "batched": {k: batched[k] for k in ("device_ms_per_run", "brick_log2", "tile_shape", "autotune")},
            "plain_kernel": plain,
            "batched_vs_plain_kernel": batched["value"] / value,
        }
        value, ms_step = batched["value"], batched["ms_per_step"]
    else:
        extra = {"fps": plain["fps"], "frame_ms_p50": plain["frame_ms_p50"],
                 "frame_ms_p99": plain["frame_ms_p99"], "frames_per_s_device": plain["frames_per_s_device"]}

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
            "workload": (f"{sc.name}: {len(cams)} cameras, {sc.width}x{sc.height}, "
                         f"{sc.data.shape[0]}^3 blob volume" if args.workload != "c0" else sc.name)
                        + ("; the cameras batched into one persistent graph launch per step"
                           if batched is not None else "; one launch per frame"),
            "volume": list(sc.data.shape), "image": [sc.width, sc.height],
            "cameras": len(cams), "preset": {"step_size": sc.config.step_size,
                                             "max_steps": sc.config.max_steps},
            "et_threshold": sc.tau, "scalar_format": sf,
            "normal_format": nf if sc.normals is not None else "none",
            "samples_per_step": s_step, "volume_bytes": vol_bytes, "brick_skip": skip,
        },
        **extra,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
