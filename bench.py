"""Benchmark: the per-pixel ray-tracing path on one MI355X.

Workload: the 1080p mesh + PBR scene at bounce depth 4 (three 81,920-triangle displaced icospheres on a textured
floor, Cook-Torrance GGX, 1024^2 mip-mapped checker+noise albedo, cube-map IBL, one glass and one mirror
instance).  A step is one frame over all 2,073,600 primary rays, with the scene resident in HBM and the canvas
(pixels + depths) written to HBM, in the fastest of eighteen frame forms (normally one launch of the per-sample
render kernel over the meshes' 8-wide BVHs, its tiles dispatched costliest first by an order computed on a second
stream during the previous frame).

Reported:
  value        whole-frame Mrays/s, rays = every SceneTracer::trace call (primary + bounce + shadow), counted
               by an instrumented pass of the same kernel (deterministic; equal to the CPU oracle's count)
  ms_per_step  time per frame between two HIP events around --steps back-to-back frames (synchronised before
               and after)
  roofline     HBM roofline of the render kernel: algorithmic bytes per frame / mean kernel time (HIP events on
               the launch stream); bytes per unit, counted on the reference's sibling-pair walk (the same for
               every form): 64 per sibling-pair visit, 64 per instance visit, 64 per triangle test, 64 per mesh
               closest-hit attribute fetch, 12 per texel-quad tap, 20 per sample out; traffic_frac = measured
               HBM bytes (committed PMC profile of the form) / kernel time / peak
  cpu_baseline the reference's own CPU renderer (compiled from its sources, oracle/_ref/slimref) on this host:
               one full frame with the threads this process may use (at most 16) and with one thread, CPU model
  neural_cache the plain CPU-runnable scene (seed 0) at 1920x1080 with bounce depth 8, rendered over a seeded
               orbit of the camera (300 frames) twice per frame: the plain full-depth trace (the cache's reference) and the
               neural-cache form (cache after the second bounce; 1 in 8 pixels, 2^18 at 1080p, trace full depth
               for training; confidence gate over the last ~16 frames), each timed with HIP events around its
               own launches; PSNR and MAE of every cached frame against its plain frame.  A second pass puts the
               cache at the primary hit, where it replaces the most work.
  batched_views 48 cameras (resolutions drawn from 128^2 .. 1920x1080) over the mesh + PBR scene; every frame
               the 3 instances get new random transforms.  Graph form: the frame (transform copy, device scene
               BVH rebuild, every view's tiles in one launch, launch shape from the autotuner) replayed as one
               HIP graph; baseline: the host rebuild + upload and one launch per view.  Per-frame GPU latency
               p50/p99 (HIP events) and aggregate Mrays/s (rays counted on the first frames' transforms).
  texel_storage (skip with --no-texel-storage) the same frame with the configuration's texture formats: RGBA8
               albedo texels, RGBA16F cube faces of an HDR sky (include/slimrt.h SLIMRT_TEXELS_*)
  large        (skip with --no-large) the combined 4K workload: 16 instances of a 262,144-triangle displaced sphere,
               32 implicit primitives, 4 emissive quads, 8 + 8 textures at 2048^2, 512^2 cube IBL, depth 8; one
               4K frame in every form, 8 uneven views (4K down to 640x360) in one launch vs one launch per view,
               and the neural cache at the primary hit against the plain trace.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER = {"pairs": 64, "instances": 64, "triangles": 64, "mesh_hits": 64, "taps": 12}
OUT_BYTES_PER_SAMPLE = 20


FORM_NAMES = ("megakernel", "megakernel_image_order", "wavefront", "wavefront_sorted", "wavefront_per_bounce", "wavefront_per_bounce_sorted",
              "wavefront_split", "wavefront_sorted_split")
FORM_NAMES = FORM_NAMES + tuple(f + "_bvh4" for f in FORM_NAMES) + ("megakernel_bvh8", "megakernel_image_order_bvh8")


def measured_traffic(width: int, height: int, form: str):
    """HBM bytes per frame of this form (the sum over the frame's launches) on this workload, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, each scaled by the factor the same counter reads on an
    elementwise kernel of known size); None when no summary covers the form."""
    import re
    path = os.path.join(ROOT, "profiles", f"{form}_mesh_pbr_{width}x{height}_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9]+)", open(path).read())
    return int(m.group(1)) if m else None


def algorithmic_bytes(cnt: dict, wide: bool = False) -> int:
    """Bytes the frame's work reads and writes on the reference's units, or (wide) on the 4-wide form's own records:
    128-byte node visits counted by 'pairs' and 48-byte triangle records."""
    per = dict(BYTES_PER, pairs=128, triangles=48) if wide else BYTES_PER
    return sum(cnt[k] * b for k, b in per.items()) + cnt["pixels"] * OUT_BYTES_PER_SAMPLE


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_reference(spec, threads: int, reps: int) -> tuple:
    """The compiled reference renderer (oracle/_ref/slimref) on one full frame: best-of-`reps` ms and process wall."""
    import oracle
    from slimtracincpp_amd.specfile import write_spec, read_container
    with tempfile.TemporaryDirectory() as td:
        sp = write_spec(spec, os.path.join(td, "scene.spec"))
        out = os.path.join(td, "out.bin")
        t0 = time.time()
        subprocess.run([oracle.REF_PATH, sp, out, "--threads", str(threads), "--reps", str(reps)], check=True,
                       capture_output=True, timeout=900)
        return float(read_container(out)["timing"][0]), time.time() - t0


def cpu_baseline(spec, rays_per_frame: int) -> dict:
    """The reference's own CPU renderer on this host: one full frame of the same scene, (i) with one thread, as the
    reference ships it (renderOnCPU is a serial loop over pixels), and (ii) with the host threads this process may
    use, capped at 16 (the host share of one GPU on these machines), rows split evenly.  value = mode (ii)."""
    import oracle
    if not os.path.exists(oracle.REF_PATH):
        return {"value": None, "unit": "Mrays/s", "cores": 0, "kind": "reference",
                "sample": "oracle/_ref/slimref not built"}
    nproc = os.cpu_count() or 1
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    threads = max(1, min(16, usable))
    ms_n, wall_n = time_reference(spec, threads, 2)
    ms_1, wall_1 = time_reference(spec, 1, 1)
    return {"value": rays_per_frame / (ms_n * 1e-3) / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "reference",
            "ms_per_frame": ms_n, "cpu_model": cpu_model(), "nproc": nproc, "usable_cpus": usable,
            "single_thread": {"value": rays_per_frame / (ms_1 * 1e-3) / 1e6, "unit": "Mrays/s", "cores": 1,
                              "ms_per_frame": ms_1},
            "sample": f"one full {spec.width}x{spec.height} frame of the same scene with the reference's renderPixel "
                      f"and closed-form primary rays: {threads} threads (rows split, best of 2; process wall "
                      f"{wall_n:.1f}s incl. its mesh BVH build) and 1 thread (one frame; wall {wall_1:.1f}s)"}


def orbit_camera(frame: int, frames: int, radius: float = 8.246, height: float = 2.0, target=(0.0, 0.0, 0.0)):
    """Orbit around `target` starting on its -z side (for the plain scene: its own camera position (0, 2, -8))."""
    import math
    import numpy as np
    from slimtracincpp_amd import scenes
    a = -math.pi / 2 + 2 * math.pi * frame / frames
    pos = np.array([target[0] + radius * math.cos(a), height, target[2] + radius * math.sin(a)], np.float32)
    return scenes.look_at_euler(pos, target), pos, scenes.FOV60_FOCAL


def neural_cache_leg(frames: int, width: int, height: int, stream, cache_depth: int, spec=None) -> dict:
    """spec None: the plain scene at width x height, depth 8, orbit radius 8.25; else that scene at depth 8 with
    an orbit of radius 6.6 at height 2.4 around (0, 1, 0.4)."""
    import numpy as np
    import torch
    from slimtracincpp_amd import scenes, runtime
    from slimtracincpp_amd.nrc import CachedFrames
    if spec is None:
        spec = scenes.implicit_scene(0, width, height, max_depth=8)
        orbit = dict(radius=8.246, height=2.0, target=(0.0, 0.0, 0.0))
    else:
        orbit = dict(radius=6.6, height=2.4, target=(0.0, 1.0, 0.4))
    r = runtime.Renderer(spec, "gpu")
    r.set_stream(stream.cuda_stream)
    n = width * height
    plain = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    cached = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    depths = torch.empty((n,), dtype=torch.float32, device="cuda")
    cf = CachedFrames(r, seed=1, cache_depth=cache_depth, train_one_in=8, train_steps=4, train_batch=16384,
                      learning_rate=5e-3, gate_threshold=0.1, gate_min_frames=4, gate_cell_size=0.25,
                      gate_cells_log2=20)
    ms_plain, ms_nrc, psnr, mae, queried, fallback = [], [], [], [], [], []
    plain_flags = None
    for f in range(frames):
        p = r.params(mode=scenes.MODE_BEAUTY, output=runtime.OUTPUT_RADIANCE, max_depth=8, size=(width, height),
                     camera=orbit_camera(f, frames, **orbit))
        if plain_flags is None:   # the plain trace in its fastest form on this workload
            plain_flags, _ = r.autotune_form(p, plain.data_ptr(), depths.data_ptr(), reps=2)
        p.flags = plain_flags
        r.render_device(p, plain.data_ptr(), depths.data_ptr())
        ms_plain.append(r.last_kernel_ms())
        p.flags = runtime.BVH8    # the cached frame traces through the 8-wide mesh BVHs as well
        cf.render_device(p, cached.data_ptr(), depths.data_ptr())
        ms_nrc.append(r.last_kernel_ms())
        d = (cached[:, :3] - plain[:, :3]).double()
        mse = float(torch.mean(d * d))
        psnr.append(10 * np.log10(1.0 / max(mse, 1e-20)))
        mae.append(float(torch.mean(d.abs())))
        st = cf.stats()
        queried.append(st["queries"] / n)
        fallback.append(st["fallbacks"] / n)
    r.synchronize()
    cf.cache.close()
    r.close()
    tail = slice(frames // 2, None)   # after the cache has trained on half the orbit
    mp, mn = float(np.median(ms_plain)), float(np.median(ms_nrc))
    return {"cache_depth": cache_depth, "frames": frames, "plain_form_flags": plain_flags,
            "ms_per_frame_plain": mp, "ms_per_frame_cache": mn, "speedup": mp / mn,
            "psnr_db_min": float(np.min(psnr[tail])), "psnr_db_mean": float(np.mean(psnr[tail])),
            "psnr_db_first": float(psnr[0]), "mae_mean": float(np.mean(mae[tail])),
            "queried_fraction": float(np.mean(queried[tail])), "fallback_fraction": float(np.mean(fallback[tail])),
            "psnr_db_per_frame": [round(float(v), 2) for v in psnr]}


def cache_bound(spec, stream, reps: int = 10) -> dict:
    """The shortest frame any cache of the path's radiance can give on this scene: it must still find each sample's
    primary hit (a frame that traces the camera rays and writes hit records only, SLIMRT_OUTPUT_HITS, in the
    fastest per-sample form) and evaluate the network once per sample with a hit (slimrt_nrc_infer).  Against the
    plain full-depth frame in its fastest form, that bounds the speed-up of the cache frame from above."""
    import numpy as np
    import torch
    from slimtracincpp_amd import runtime
    from slimtracincpp_amd.nrc import NeuralCache, scene_box
    r = runtime.Renderer(spec, "gpu")
    r.set_stream(stream.cuda_stream)
    n = spec.width * spec.height
    px = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    dp = torch.empty((n,), dtype=torch.float32, device="cuda")

    def timed(p):
        for _ in range(3):
            r.render_device(p, px.data_ptr(), dp.data_ptr())
        ts = []
        for _ in range(reps):
            r.render_device(p, px.data_ptr(), dp.data_ptr())
            ts.append(r.last_kernel_ms())
        r.synchronize()
        return float(np.median(ts))

    p = r.params(output=runtime.OUTPUT_RADIANCE)
    flags, _ = r.autotune_form(p, px.data_ptr(), dp.data_ptr(), reps=2)
    p.flags = flags
    plain_ms = timed(p)
    h = r.params(output=runtime.OUTPUT_HITS)
    trace = {}
    for name, f in (("megakernel", 0), ("megakernel_bvh4", runtime.BVH4), ("megakernel_bvh8", runtime.BVH8)):
        h.flags = f
        trace[name] = timed(h)
    hits = int(np.sum(px.view(torch.int32)[:, 0].cpu().numpy() >= 0))
    center, inv_half = scene_box(r.scene)
    cache = NeuralCache(center, inv_half, 1)
    cache.set_stream(stream.cuda_stream)
    raw = torch.rand((max(hits, 1), 14), dtype=torch.float32, device="cuda")
    out = torch.empty((max(hits, 1), 3), dtype=torch.float32, device="cuda")
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    cache.infer(raw, out)
    ev0.record(stream)
    for _ in range(reps):
        cache.infer(raw, out)
    ev1.record(stream)
    torch.cuda.synchronize()
    infer_ms = ev0.elapsed_time(ev1) / reps
    cache.close()
    r.close()
    trace_ms = min(trace.values())
    return {"plain_ms": plain_ms, "plain_form_flags": flags, "primary_trace_ms": trace_ms, "primary_trace_forms_ms": trace,
            "samples_with_hit": hits, "inference_ms": infer_ms, "bound_ms": trace_ms + infer_ms,
            "max_speedup": plain_ms / (trace_ms + infer_ms)}


def batched_views_leg(frames: int, stream, spec, n_views: int = 48) -> dict:
    import numpy as np
    import torch
    from slimtracincpp_amd import scenes, runtime
    r = runtime.Renderer(spec, "gpu")
    r.set_stream(stream.cuda_stream)
    cams = scenes.batched_views(n_views, 0)
    params, bufs = [], []
    for w, h, eul, pos, focal in cams:
        p = r.params(camera=(eul, pos, focal), size=(w, h))
        params.append(p)
        bufs.append((torch.empty((w * h, 4), dtype=torch.float32, device="cuda"),
                     torch.empty(w * h, dtype=torch.float32, device="cuda")))
    pixels = sum(p.width * p.height for p in params)
    views = runtime.Views(r, params, [b[0].data_ptr() for b in bufs], [b[1].data_ptr() for b in bufs])
    cfg, tuned_ms = views.autotune(reps=2)
    # rays per frame: instrumented passes of every view on the first frames' transforms
    rng = scenes.PCG32(7)
    transforms = [scenes.instance_transforms(spec, rng) for _ in range(frames)]
    rays = []
    for geo_f in transforms[:2]:
        r.update_scene_device(geo_f, runtime.SCENE_REBUILD)
        rays.append(sum(r.count_work(p)["rays"] for p in params))
    rays_frame = float(np.mean(rays))
    graph_ms = []
    for i, geo_f in enumerate(transforms):
        views.frame(geo_f, runtime.SCENE_REBUILD, graph=True)
        graph_ms.append(views.last_ms())
    r.synchronize()
    # baseline: host rebuild + upload, then one launch per view
    base_ms, base_wall = [], []
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for geo_f in transforms:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record(stream)
        r.update_geometries(geo_f)
        for p, (px, dp) in zip(params, bufs):
            r.render_device(p, px.data_ptr(), dp.data_ptr())
        ev1.record(stream)
        torch.cuda.synchronize()
        base_wall.append((time.perf_counter() - t0) * 1e3)
        base_ms.append(ev0.elapsed_time(ev1))
    r.synchronize()
    views.close()
    g = np.array(graph_ms[1:])   # the first replay includes the capture
    b = np.array(base_ms[1:])
    return {"views": n_views, "frames": frames, "pixels_per_frame": pixels, "rays_per_frame": rays_frame,
            "launch_config": cfg, "autotune_ms": tuned_ms,
            "graph_ms_p50": float(np.percentile(g, 50)), "graph_ms_p99": float(np.percentile(g, 99)),
            "graph_mrays_per_s": rays_frame / (float(np.mean(g)) * 1e-3) / 1e6,
            "per_view_ms_p50": float(np.percentile(b, 50)), "per_view_ms_p99": float(np.percentile(b, 99)),
            "per_view_wall_ms_p50": float(np.percentile(base_wall[1:], 50)),
            "per_view_mrays_per_s": rays_frame / (float(np.mean(b)) * 1e-3) / 1e6}


def texel_storage_leg(width: int, height: int, stream, steps: int, warmup: int) -> dict:
    """The headline workload with the configuration's texture formats: the 1024^2 albedo (and its mips) as RGBA8
    texels, the 6 x 256^2 sky and 6 x 32^2 irradiance faces as RGBA16F texels of an HDR sky (the same scene
    otherwise); the 4- and 8-wide megakernels in both tile orders, the faster reported, beside the same forms on
    the texel-quad scene rendered the same way (a context of its own next to the headline's)."""
    import torch
    from slimtracincpp_amd import scenes, runtime
    spec = scenes.mesh_pbr_scene(0, width, height, texel_storage=True)
    quads = scenes.mesh_pbr_scene(0, width, height)
    nbytes = lambda sp: int(sum(m.nbytes for t in sp.textures for m in t.mips))
    n = width * height
    px = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    dp = torch.empty(n, dtype=torch.float32, device="cuda")
    form_flags = (("megakernel_bvh4", runtime.BVH4),
                  ("megakernel_image_order_bvh4", runtime.BVH4 | runtime.TILE_ORDER_FIXED),
                  ("megakernel_bvh8", runtime.BVH8),
                  ("megakernel_image_order_bvh8", runtime.BVH8 | runtime.TILE_ORDER_FIXED))

    def time_forms(sp):
        r = runtime.Renderer(sp, "gpu")
        r.set_stream(stream.cuda_stream)
        p = r.params()
        cnt = r.count_work(p)
        forms = {}
        for name, flags in form_flags:
            p.flags = flags
            for _ in range(warmup):
                r.render_device(p, px.data_ptr(), dp.data_ptr())
            r.synchronize()
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            for _ in range(steps):
                r.render_device(p, px.data_ptr(), dp.data_ptr())
            ev1.record(stream)
            torch.cuda.synchronize()
            forms[name] = ev0.elapsed_time(ev1) / steps
        r.close()
        return forms, cnt["rays"]

    forms, rays = time_forms(spec)
    quad_forms, _ = time_forms(quads)
    best = min(forms, key=forms.get)
    return {"workload": "mesh_pbr as configured: 1024^2 RGBA8 albedo + mips, cube IBL 6x256^2 + 6x32^2 RGBA16F "
                        "faces (HDR sky), otherwise the headline scene",
            "texture_bytes": nbytes(spec), "texture_bytes_quads": nbytes(quads), "rays_per_frame": rays,
            "forms_ms": forms, "best_form": best, "ms_per_frame": forms[best],
            "mrays_per_s": rays / (forms[best] * 1e-3) / 1e6,
            "quads_forms_ms": quad_forms, "quads_ms_per_frame": min(quad_forms.values())}


def large_leg(stream, frames: int = 5) -> dict:
    import numpy as np
    import torch
    from slimtracincpp_amd import scenes, runtime
    from slimtracincpp_amd.nrc import CachedFrames
    t0 = time.time()
    spec = scenes.large_scene(0)
    print(f"large: scene generated in {time.time() - t0:.0f}s", file=sys.stderr, flush=True)
    t0 = time.time()
    r = runtime.Renderer(spec, "gpu")
    r.set_stream(stream.cuda_stream)
    print(f"large: host build + upload in {time.time() - t0:.0f}s", file=sys.stderr, flush=True)
    p = r.params()
    n = p.width * p.height
    px = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    dp = torch.empty(n, dtype=torch.float32, device="cuda")
    cnt = r.count_work(p)
    forms = {}
    for name, flags in (("megakernel", 0), ("wavefront", runtime.WAVEFRONT),
                        ("wavefront_sorted", runtime.WAVEFRONT | runtime.WAVEFRONT_SORT),
                        ("wavefront_per_bounce", runtime.WAVEFRONT | runtime.WAVEFRONT_PER_BOUNCE),
                        ("megakernel_bvh4", runtime.BVH4), ("megakernel_bvh8", runtime.BVH8),
                        ("wavefront_bvh4", runtime.WAVEFRONT | runtime.BVH4),
                        ("wavefront_sorted_bvh4", runtime.WAVEFRONT | runtime.WAVEFRONT_SORT | runtime.BVH4)):
        p.flags = flags
        ts = []
        for i in range(frames + 1):
            r.render_device(p, px.data_ptr(), dp.data_ptr())
            ts.append(r.last_kernel_ms())
        forms[name] = float(np.median(ts[1:]))
    best = min(forms, key=forms.get)
    print(f"large: frame forms {forms}", file=sys.stderr, flush=True)
    # 8 uneven views in one launch vs one launch per view
    params, bufs = [], []
    for w, h, eul, pos, focal in scenes.large_views(0):
        vp = r.params(camera=(eul, pos, focal), size=(w, h))
        params.append(vp)
        bufs.append((torch.empty((w * h, 4), dtype=torch.float32, device="cuda"),
                     torch.empty(w * h, dtype=torch.float32, device="cuda")))
    views = runtime.Views(r, params, [b[0].data_ptr() for b in bufs], [b[1].data_ptr() for b in bufs])
    cfg, _ = views.autotune(reps=1)
    vg = []
    for i in range(frames + 1):
        views.frame(graph=True)
        vg.append(views.last_ms())
    views.close()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    vs = []
    for i in range(frames + 1):
        torch.cuda.synchronize()
        ev0.record(stream)
        for vp, (bp, bd) in zip(params, bufs):
            r.render_device(vp, bp.data_ptr(), bd.data_ptr())
        ev1.record(stream)
        torch.cuda.synchronize()
        vs.append(ev0.elapsed_time(ev1))
    view_rays = sum(r.count_work(vp)["rays"] for vp in params)
    print("large: views done", file=sys.stderr, flush=True)
    # the neural cache at the primary hit, static camera, trained over the frames (both frames over 8-wide nodes)
    p.flags = runtime.BVH8
    p.output = runtime.OUTPUT_RADIANCE
    plain = torch.empty_like(px)
    r.render_device(p, plain.data_ptr(), dp.data_ptr())
    plain_ms = r.last_kernel_ms()
    cf = CachedFrames(r, seed=1, cache_depth=0, train_one_in=8, train_steps=4, train_batch=16384,
                      gate_threshold=0.1, gate_min_frames=4)
    psnr, nrc_ms = [], []
    for f in range(6 * frames):
        cf.render_device(p, px.data_ptr(), dp.data_ptr())
        nrc_ms.append(r.last_kernel_ms())
        d = (px[:, :3] - plain[:, :3]).double()
        psnr.append(10 * np.log10(1.0 / max(float(torch.mean(d * d)), 1e-20)))
    st = cf.stats()
    cf.cache.close()
    r.synchronize()
    rays = cnt["rays"]
    return {"workload": "large combined scene, 3840x2160, depth 8", "rays_per_frame": rays, "counters": cnt,
            "forms_ms": forms, "best_form": best, "ms_per_frame": forms[best],
            "mrays_per_s": rays / (forms[best] * 1e-3) / 1e6,
            "views": {"count": len(params), "launch_config": cfg, "rays_per_frame": view_rays,
                      "graph_ms_p50": float(np.median(vg[1:])), "per_view_ms_p50": float(np.median(vs[1:])),
                      "graph_mrays_per_s": view_rays / (float(np.median(vg[1:])) * 1e-3) / 1e6},
            "neural_cache_at_primary_hit": {"plain_ms": plain_ms, "cache_ms_median": float(np.median(nrc_ms[2:])),
                                            "psnr_db_last": psnr[-1], "psnr_db_max": max(psnr),
                                            "queried_fraction": st["queries"] / n,
                                            "fallback_fraction": st["fallbacks"] / n}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-neural-cache", action="store_true")
    ap.add_argument("--nrc-frames", type=int, default=300)
    ap.add_argument("--no-views", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the combined 4K workload (about 1 minute)")
    ap.add_argument("--large", action="store_true", help=argparse.SUPPRESS)   # the 4K workload runs by default
    ap.add_argument("--view-frames", type=int, default=20)
    ap.add_argument("--no-texel-storage", action="store_true", help="skip the RGBA8 / RGBA16F texture-format leg")
    ap.add_argument("--form", choices=FORM_NAMES, default=None,
                    help="time only this frame form (for profiler runs); by default every form, the fastest reported")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a plain device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import torch
    from slimtracincpp_amd import scenes, runtime

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    spec = scenes.mesh_pbr_scene(0, args.width, args.height)
    r = runtime.Renderer(spec, "gpu")
    p = r.params()
    n = args.width * args.height
    pixels = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    depths = torch.empty((n,), dtype=torch.float32, device="cuda")
    stream = torch.cuda.Stream()             # a real (non-NULL) stream shared by torch's events and the kernels
    torch.cuda.set_stream(stream)
    r.set_stream(stream.cuda_stream)

    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):   # an elementwise kernel: reads and writes exactly MIB MiB (a blit copy is not counted)
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    cnt = r.count_work(p)                       # sibling-pair walk (equal to the oracle's counts)
    p.flags = runtime.BVH4
    cnt4 = r.count_work(p)                      # 4-wide mesh walk: 'pairs' counts node visits
    p.flags = runtime.BVH8
    cnt8 = r.count_work(p)                      # 8-wide mesh walk (128-byte quantised nodes)
    p.flags = 0

    # Eighteen forms of the same frame (tests/test_wavefront.py, tests/test_bvh4.py, tests/test_bvh8.py,
    # tests/test_tile_order.py): the per-sample megakernel (tiles costliest first by the previous frame's durations,
    # or in image order) and the bounces as a wavefront over a compacted ray queue (run to the end in one launch, or
    # re-compacted after every bounce; sorted by direction octant or not; first hits traced by the shading kernel or
    # by a separate traversal kernel), each with the meshes' sibling-pair BVHs or their 4-wide collapse; and the
    # megakernel over the 8-wide collapse in both tile orders.  Each is timed over K back-to-back frames between
    # synchronisations; the fastest is the reported path.
    base_forms = {"megakernel": 0, "megakernel_image_order": runtime.TILE_ORDER_FIXED, "wavefront": runtime.WAVEFRONT,
                  "wavefront_sorted": runtime.WAVEFRONT | runtime.WAVEFRONT_SORT,
                  "wavefront_per_bounce": runtime.WAVEFRONT | runtime.WAVEFRONT_PER_BOUNCE,
                  "wavefront_per_bounce_sorted": runtime.WAVEFRONT | runtime.WAVEFRONT_PER_BOUNCE | runtime.WAVEFRONT_SORT,
                  "wavefront_split": runtime.WAVEFRONT | runtime.WAVEFRONT_SPLIT,
                  "wavefront_sorted_split": runtime.WAVEFRONT | runtime.WAVEFRONT_SORT | runtime.WAVEFRONT_SPLIT}
    forms = dict(base_forms)
    forms.update({k + "_bvh4": v | runtime.BVH4 for k, v in base_forms.items()})
    forms.update({"megakernel_bvh8": runtime.BVH8, "megakernel_image_order_bvh8": runtime.BVH8 | runtime.TILE_ORDER_FIXED})
    assert tuple(forms) == FORM_NAMES
    if args.form:
        forms = {args.form: forms[args.form]}
    timings = {}
    for name, flags in forms.items():
        p.flags = flags
        for _ in range(args.warmup):
            r.render_device(p, pixels.data_ptr(), depths.data_ptr())
        r.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(args.steps):
            r.render_device(p, pixels.data_ptr(), depths.data_ptr())
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / args.steps
        r.synchronize()
        # launch durations for the roofline: HIP events around each frame's launches on the launch stream
        kms = []
        for _ in range(min(args.steps, 20)):
            r.render_device(p, pixels.data_ptr(), depths.data_ptr())
            kms.append(r.last_kernel_ms())
        r.synchronize()
        timings[name] = {"ms_per_step": ev0.elapsed_time(ev1) / args.steps, "wall_ms_per_step": wall,
                         "kernel_ms": sum(kms) / len(kms)}
    form = min(timings, key=lambda k: timings[k]["ms_per_step"])
    ms_step, wall_ms, k_ms = (timings[form][k] for k in ("ms_per_step", "wall_ms_per_step", "kernel_ms"))
    p.flags = forms[form]

    wide = form.endswith("_bvh4") or form.endswith("_bvh8")
    cnt_form = cnt8 if form.endswith("_bvh8") else cnt4
    rays = cnt["rays"]
    mrays = rays / (ms_step * 1e-3) / 1e6
    # Algorithmic bytes on the reference's own traversal (sibling-pair walk, counts equal to the oracle's): the
    # same work for every frame form.  The 4-wide walk's own node bytes (128 per node visit) are reported beside.
    bytes_frame = algorithmic_bytes(cnt)
    bytes_form = algorithmic_bytes(cnt_form, True) if wide else bytes_frame
    achieved = bytes_frame / (k_ms * 1e-3) / 1e9
    traffic = measured_traffic(args.width, args.height, form)
    result = {
        "metric": "Mrays/s + ms/frame at 1080p, plain vs neural-cache at PSNR>=40 dB vs plain",
        "value": mrays,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, seed 0)",
        "config": {"workload": f"mesh_pbr {args.width}x{args.height} 1 spp bounce depth 4: 3 x 81,920-triangle "
                               "displaced icospheres (leaf 4) on a floor quad, GGX, 1024^2 albedo (9 mips), "
                               "cube IBL 6x256^2 + 6x32^2 irradiance, glass IOR 1.5 + mirror",
                   "rays_per_frame": rays, "primary_rays_per_frame": n,
                   "primary_mrays_per_s": n / (ms_step * 1e-3) / 1e6,
                   "wall_ms_per_step": wall_ms, "mode": "plain (full-depth trace)", "form": form,
                   "forms_ms_per_step": {k: v["ms_per_step"] for k, v in timings.items()}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_gbs": traffic / (k_ms * 1e-3) / 1e9 if traffic else None,
                     "traffic_frac": traffic / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
                     "basis": "algorithmic bytes of the reference's sibling-pair walk (64 B per pair visit); "
                              "traffic = measured HBM bytes per frame (PMC FETCH_SIZE + WRITE_SIZE, calibrated)",
                     "achieved_form_bytes": bytes_form / (k_ms * 1e-3) / 1e9, "form_bytes": bytes_form,
                     "kernel": ({"_bvh4": "4-wide mesh BVHs; ", "_bvh8": "8-wide mesh BVHs (quantised child boxes); "}
                                .get(form[-5:], "")) + {
                                "megakernel": "render_kernel<false> (tiles dispatched costliest first; their order, from "
                                              "the tile durations of three frames before, is counted and placed "
                                              "by the first workgroups of the two frames before it while their "
                                              "other workgroups render)",
                                "megakernel_image_order": "render_kernel<false> (tiles in image order)",
                                "wavefront": "the frame's launches: wf_primary_kernel + one wf_bounce_kernel<true> "
                                             "over the compacted bounce queue",
                                "wavefront_sorted": "the frame's launches: wf_primary_kernel + octant sort (count, "
                                                    "scan, scatter) + wf_bounce_kernel<true>",
                                "wavefront_per_bounce": "the frame's launches: wf_primary_kernel + "
                                                        "wf_bounce_kernel<false> per bounce",
                                "wavefront_per_bounce_sorted": "the frame's launches: wf_primary_kernel + per bounce "
                                                               "octant sort + wf_bounce_kernel<false>",
                                "wavefront_split": "the frame's launches: wf_trace_kernel (first hits) + "
                                                   "wf_primary_kernel (their shading) + one wf_bounce_kernel<true>",
                                "wavefront_sorted_split": "the frame's launches: wf_trace_kernel + wf_primary_kernel + "
                                                          "octant sort + wf_bounce_kernel<true>"}[form.replace("_bvh4", "").replace("_bvh8", "")],
                     "kernel_ms": k_ms, "algorithmic_bytes": bytes_frame,
                     "counters": cnt, "counters_form": cnt_form if wide else cnt},
        "neural_cache": None,
    }
    if form in ("megakernel", "megakernel_bvh4", "megakernel_bvh8"):
        # a frame's tile durations (slimrt_tile_costs): how far the frame is from its ideal length
        import numpy as np
        for _ in range(4):   # this form's frames (the order of a frame comes from the durations three frames before)
            r.render_device(p, pixels.data_ptr(), depths.data_ptr())
        costs, order = r.tile_costs()
        if costs.size:
            c = costs.astype(np.int64)
            # the ideal frame: the summed wave durations over the resident waves (a split tile records its last
            # part's duration times its parts, which stands for its part waves)
            ideal_us = float(c.sum()) / r.tile_resident / 100.0
            wave = c.copy()   # the duration of each tile's longest wave
            ordered = not np.array_equal(order, np.arange(order.size))
            if ordered:
                split = min(c.size // 32, 2048)
                quarter = min(c.size // 1024, 128, split) if form.endswith("_bvh8") else 0
                wave[order[:quarter]] //= 4
                wave[order[quarter:split]] //= 2
            result["tile_costs"] = {
                "tiles": int(c.size), "resident_waves": r.tile_resident, "median_us": float(np.median(wave)) / 100.0,
                "p99_us": float(np.percentile(wave, 99)) / 100.0, "longest_us": float(wave.max()) / 100.0,
                "ideal_frame_us": ideal_us, "order": "longest first" if ordered else "image"}
    if not args.no_texel_storage:
        result["texel_storage"] = texel_storage_leg(args.width, args.height, stream, args.steps, args.warmup)
    if not args.no_views:
        result["batched_views"] = batched_views_leg(args.view_frames, stream, spec)
    if not args.no_large:
        result["large"] = large_leg(stream)
    if not args.no_neural_cache:
        legs = [neural_cache_leg(args.nrc_frames, args.width, args.height, stream, d) for d in (2, 0)]
        mesh_legs = [neural_cache_leg(args.nrc_frames, args.width, args.height, stream, d, spec) for d in (2, 0)]
        result["neural_cache"] = {
            "workload": f"implicit scene seed 0 at {args.width}x{args.height}, bounce depth 8 (plain judge), "
                        f"{args.nrc_frames}-frame orbit at radius 8.25 and height 2 around the origin",
            "training": "1 in 8 waves of 64 pixels per frame trace full paths (2^18 rays at 1080p); "
                        "4 Adam steps x 16,384 samples, lr 5e-3",
            "gate": "hashed 0.25-unit cells, EMA 1/16 of target mean and mean square, >= 4 frames, "
                    "std <= 0.1 x mean",
            "timing": "median HIP-event time of each form's launches per frame (cache form includes inference, "
                      "training and the gate update); PSNR/MAE over the second half of the orbit",
            "after_second_bounce": legs[0], "at_primary_hit": legs[1],
            "mesh_pbr_depth8": {"workload": "the mesh + PBR scene at bounce depth 8, orbit radius 6.6 at height 2.4 "
                                            "around (0, 1, 0.4)",
                                "after_second_bounce": mesh_legs[0], "at_primary_hit": mesh_legs[1]},
            "bound": {"what": "the shortest frame any radiance cache can give: the camera rays' closest hits "
                              "(hit-record frame, fastest per-sample form) plus one network inference per sample "
                              "with a hit; max_speedup = plain frame / that",
                      "mesh_pbr_depth4": cache_bound(spec, stream),
                      "mesh_pbr_depth8": cache_bound(scenes.mesh_pbr_scene(0, args.width, args.height, max_depth=8),
                                                     stream),
                      "implicit_depth8": cache_bound(scenes.implicit_scene(0, args.width, args.height, max_depth=8),
                                                     stream)}}
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(spec, rays)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
