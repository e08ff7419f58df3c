#!/usr/bin/env python3
"""Closest-hit throughput and path-tracing benchmarks on one MI355X.

Headline workload (one step): the large scene -- an fBm heightfield on a 2237 x 2237 vertex grid
(8 octaves, gain 0.5, seed 3; 9,999,392 triangles) under an 8-wide SAH BVH -- traced with 2,097,152
camera rays (2048 x 1024 jittered pinhole framing the surface: 99% of them hit) and 2,097,152 cosine
bounce rays spawned from those rays' hits: two vt_trace_closest launches over device-resident SoA
ray buffers.  `value` = rays traced per second over the K timed steps (Mrays/s).

Also reported for the headline scene: the roofline of the traversal kernel (algorithmic node +
triangle bytes per launch / average launch time measured with HIP events on the launch stream,
against 8 TB/s HBM, with the measured HBM traffic of the committed rocprofv3 profile), the BVH build
time, the closest-hit rate on the reference's own incoherent ray recipe (4,194,304 rays, origins
uniform in the scene's box, directions uniform on the sphere, best of 3 passes after a warm-up,
scripts/measure_acceleration.swift:176-187, 237-269), and the CPU baseline: the oracle's restatement
of the reference's 4-wide traversal, built -O3 -march=native on this host, over every ray of the
step, median of 5 runs after a warm-up.  The traversal-stress scene (1024 icospheres, 1,310,720
triangles) follows as a secondary block with the same figures; then the integrator, learned-cache,
ReSTIR DI and batched-rendering workloads.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "metal-vibe-tracer_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "closest-hit Mrays/s + equal-time relMSE vs FP64 oracle, plain vs NRC, 1 GPU"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_PROFILES = {"fbm-heightfield-2237^2": "profiles/trace_closest_heightfield_10M_traffic.json",
                    "icosphere-field-1024x1280": "profiles/trace_closest_icosphere_field_1M3_traffic.json"}


def measured_traffic(name):
    """HBM bytes per traversal launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes of
    this workload (scaled by the copy-kernel calibration recorded beside them), or None."""
    path = TRAFFIC_PROFILES.get(name)
    if not path or not os.path.exists(os.path.join(ROOT, path)):
        return None, None
    with open(os.path.join(ROOT, path)) as f:
        t = json.load(f)
    return int(t["traffic_bytes_per_launch"]), path


def workload_stress():
    import vt
    mesh = vt.icosphere_field(spheres=1024, subdivisions=3, mu=0.0, sigma=0.5, half_extent=50.0, seed=7)
    prim = vt.camera_rays((0.0, 30.0, -170.0), (0.0, 0.0, 0.0), 45.0, 2048, 1024, 7)
    return "icosphere-field-1024x1280", mesh, prim, 7


def workload_heightfield():
    import vt
    mesh = vt.heightfield(grid=2237, octaves=8, gain=0.5, half_extent=100.0, height=12.0, seed=3)
    prim = vt.camera_rays((0.0, 70.0, -70.0), (0.0, 0.0, 0.0), 40.0, 2048, 1024, 3)
    return "fbm-heightfield-2237^2", mesh, prim, 3


def shadow_rays(mesh, bounce, seed):
    """(8, n) float32 rays from the bounce rays' origins to uniform points of a square light of half
    the scene box's largest extent, centred half an extent above the box (+y); tmax stops 1e-4 of
    the distance short of the light."""
    import numpy as np
    lo, hi = mesh.vertices[:, :3].min(0).astype(np.float64), mesh.vertices[:, :3].max(0).astype(np.float64)
    ext = float((hi - lo).max())
    n = bounce.shape[1]
    rng = np.random.default_rng(seed + 101)
    c = 0.5 * (lo + hi)
    light = np.stack([c[0] + (rng.random(n) - 0.5) * 0.5 * ext, np.full(n, hi[1] + 0.5 * ext),
                      c[2] + (rng.random(n) - 0.5) * 0.5 * ext])
    o = bounce[0:3].astype(np.float64)
    d = light - o
    dist = np.sqrt((d * d).sum(0))
    out = np.empty((8, n), np.float32)
    out[0:3] = bounce[0:3]
    out[3:6] = d / dist
    out[6] = bounce[6]
    out[7] = dist * (1.0 - 1e-4)
    return out


def timed_render(fn, *args, **kw):
    """(seconds, result) of one oracle render call."""
    c0 = time.perf_counter()
    out = fn(*args, **kw)
    return time.perf_counter() - c0, out


def render_setup(fn, *args, **kw):
    """Seconds of a render call's set-up (material records, emitter list, light tree): the faster of two
    calls over a single pixel sample, after one that absorbs the first call's one-time costs."""
    fn(*args, **kw)
    return min(timed_render(fn, *args, **kw)[0] for _ in range(2))


def measure_bandwidth(dev):
    """Device-to-device copy rates (read + write bytes / time): a 2 x 1 GiB copy (HBM) and a
    2 x 32 MiB copy repeated 20 times (a working set inside the 256 MB Infinity Cache), best of 3."""
    import torch
    out = {}
    for name, mib, reps in (("hbm_copy_gbs", 1024, 1), ("mall_copy_gbs", 32, 20)):
        src = torch.ones(mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        best = float("inf")
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                dst.copy_(src)
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) / reps)
        out[name] = round(2 * (mib << 20) / (best * 1e-3) / 1e9, 1)
        del src, dst
    torch.cuda.empty_cache()
    return out


def measure_traversal(args, ctx, name, mesh, prim, seed):
    """One traversal workload: K timed steps of camera + bounce closest-hit launches, per-launch
    kernel times, roofline, incoherent-ray rate, BVH build time and the CPU baseline."""
    import numpy as np
    import torch

    import vt
    dev = torch.device("cuda:0")
    log(f"workload {name}: {len(mesh.faces)} triangles, {prim.shape[1]} primary rays")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    scene = ctx.set_mesh(mesh)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    info = scene.info
    log(f"bvh: {info}, built and uploaded in {build_s:.2f} s")
    stream = torch.cuda.current_stream()
    prim_d = torch.from_numpy(prim).to(dev)
    h = scene.trace_closest(prim_d, stream=stream)
    torch.cuda.synchronize()
    hp = {k: v.cpu().numpy() for k, v in h.items()}
    bounce = vt.bounce_rays(mesh, prim, hp["face"], hp["b1"], hp["b2"], seed)
    bounce_d = torch.from_numpy(bounce).to(dev)
    n_p, n_b = prim.shape[1], bounce.shape[1]
    out_p = {k: torch.empty(n_p, dtype=torch.int32 if k == "face" else torch.float32, device=dev)
             for k in ("t", "b1", "b2", "face")}
    out_b = {k: torch.empty(n_b, dtype=torch.int32 if k == "face" else torch.float32, device=dev)
             for k in ("t", "b1", "b2", "face")}

    def step(ev=None):
        if ev:
            ev[0].record(stream)
        scene.trace_closest(prim_d, out=out_p, stream=stream)
        if ev:
            ev[1].record(stream)
        scene.trace_closest(bounce_d, out=out_b, stream=stream)
        if ev:
            ev[2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # Timed region: exactly K steps, bracketed by device synchronisation.
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = (n_p + n_b) * args.steps / elapsed / 1e6

    # Per-launch kernel durations (HIP events on the launch stream), a separate pass.
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    t_prim = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    t_bnc = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))

    # Algorithmic bytes: node records fetched x node size + triangle records tested x their size.
    np_, tp_ = scene.trace_stats(prim_d, stream=stream)
    nb_, tb_ = scene.trace_stats(bounce_d, stream=stream)
    sn, st = info["node_bytes"], info["tri_bytes"]
    bytes_p = np_ * sn + tp_ * st
    bytes_b = nb_ * sn + tb_ * st
    achieved = (bytes_p + bytes_b) / ((t_prim + t_bnc) * 1e-3) / 1e9
    hits_p = int((out_p["face"] >= 0).sum().item())
    hits_b = int((out_b["face"] >= 0).sum().item())
    traffic, traffic_src = measured_traffic(name)

    # The reference's incoherent probe: origins uniform in the scene's box, uniform directions,
    # 2^22 rays, best of 3 passes after a warm-up pass.
    lo, hi = mesh.vertices[:, :3].min(0), mesh.vertices[:, :3].max(0)
    inc = vt.random_rays(tuple(float(v) for v in lo), tuple(float(v) for v in hi), 1 << 22, seed)
    inc_d = torch.from_numpy(inc).to(dev)
    out_i = {k: torch.empty(inc.shape[1], dtype=torch.int32 if k == "face" else torch.float32, device=dev)
             for k in ("t", "b1", "b2", "face")}
    best = float("inf")
    for k in range(4):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        scene.trace_closest(inc_d, out=out_i, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        if k:
            best = min(best, e0.elapsed_time(e1))
    hits_i = int((out_i["face"] >= 0).sum().item())
    incoherent = {"rays": int(inc.shape[1]), "hits": hits_i, "ms": round(best, 4),
                  "mrays_per_s": round(inc.shape[1] / best / 1e3, 1),
                  "recipe": "origins uniform in the scene box, directions uniform on the sphere, best of 3 after a warm-up"}

    # Shadow set: from each bounce origin (a camera hit point) to a uniform point of a square area
    # light above the scene, ending just short of it; any-hit traversal.
    shadow = shadow_rays(mesh, bounce, seed)
    shadow_d = torch.from_numpy(shadow).to(dev)
    occ = scene.trace_any(shadow_d, stream=stream)
    ms_any = []
    for k in range(6):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        scene.trace_any(shadow_d, out=occ, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        if k:
            ms_any.append(e0.elapsed_time(e1))
    t_any = float(np.mean(ms_any))
    any_hit = {"rays": int(shadow.shape[1]), "occluded": int(occ.sum().item()), "ms": round(t_any, 4),
               "mrays_per_s": round(shadow.shape[1] / t_any / 1e3, 1),
               "recipe": "bounce origins to uniform points of a square light above the scene box (half the "
                         "box's largest extent wide, half an extent above it), mean of 5 after a warm-up"}

    cpu = None
    cross_check = None
    if not args.no_cpu:
        import oracle
        sample = np.concatenate([prim, bounce], axis=1)
        c0 = time.perf_counter()
        osc = oracle.Scene(mesh.vertices, mesh.faces)
        cpu_build = time.perf_counter() - c0
        osc.closest(sample, mode=0)                      # warm-up
        times = []
        for _ in range(5):
            c0 = time.perf_counter()
            osc.closest(sample, mode=0)
            times.append(time.perf_counter() - c0)
        cpu = {"value": round(sample.shape[1] / float(np.median(times)) / 1e6, 3), "unit": "Mrays/s",
               "cores": oracle.max_threads(), "kind": "port",
               "sample": f"every primary and bounce ray of the step ({sample.shape[1]} rays), reference 4-wide "
                         f"binned-SAH BVH (leaf 2) and traversal restated in C, gcc -O3 -march=native, OpenMP "
                         f"over rays, median of 5 after a warm-up",
               "bvh_build_seconds": round(cpu_build, 2)}
        osc.any(shadow, mode=0)
        times = []
        for _ in range(5):
            c0 = time.perf_counter()
            osc.any(shadow, mode=0)
            times.append(time.perf_counter() - c0)
        cpu["any_hit"] = {"value": round(shadow.shape[1] / float(np.median(times)) / 1e6, 3), "unit": "Mrays/s",
                          "sample": f"the shadow set ({shadow.shape[1]} rays), median of 5 after a warm-up"}
        # The same step priced with the reference 4-wide hierarchy's own work (its traversal's node
        # visits x 128 B + triangle tests x 112 B) over this build's kernel time.
        c4 = [0, 0]
        osc.closest(sample, mode=0, counters=c4)
        bytes4 = c4[0] * 128 + c4[1] * 112
        ach4 = bytes4 / ((t_prim + t_bnc) * 1e-3) / 1e9
        cross_check = {"achieved": round(ach4, 1), "frac": round(ach4 / HBM_PEAK_GBS, 4), "unit": "GB/s",
                       "node_visits": int(c4[0]), "triangle_tests": int(c4[1]), "bytes_per_step": int(bytes4),
                       "sizes": "128 B per 4-wide node, 112 B per triangle test"}
        del osc
    scene.close()
    return {
        "workload": name, "value": round(value, 2), "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "config": {"workload": name, "triangles": int(len(mesh.faces)), "primary_rays": n_p, "bounce_rays": n_b,
                   "primary_hits": hits_p, "bounce_hits": hits_b,
                   "bvh": f"8-wide binned SAH, leaf<={info['leaf_size']}, {sn} B nodes, {st} B triangles, "
                          f"depth {info['depth']}, {info['node_count']} nodes"},
        "bvh_build_seconds": round(build_s, 3),
        "kernel_ms": {"primary": round(t_prim, 4), "bounce": round(t_bnc, 4)},
        "mrays_per_s": {"primary": round(n_p / t_prim / 1e3, 1), "bounce": round(n_b / t_bnc / 1e3, 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": (bytes_p + bytes_b) // 2,
                     "work": {"node_fetches": np_ + nb_, "triangle_tests": tp_ + tb_, "bytes_per_step": bytes_p + bytes_b}},
        "incoherent": incoherent,
        "any_hit": any_hit,
        "roofline_reference_4wide": cross_check,
        "cpu_baseline": cpu,
    }


CORNELL_MATERIALS = [dict(type=0, albedo=(0.73, 0.73, 0.73)), dict(type=0, albedo=(0.65, 0.05, 0.05)),
                     dict(type=0, albedo=(0.12, 0.45, 0.15)), dict(type=3, albedo=(0, 0, 0), emission=(17, 17, 17))]


def relmse(x, ref):
    """Mean over pixels and channels of (x - ref)^2 / (ref^2 + 0.01)."""
    import numpy as np
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def bench_integrator(args, ctx):
    """Config 1: Cornell box (36 triangles), 512 x 512, 64 spp, depth 8, MIS, 4 wavelengths, PCG seed 1.

    GPU: time of the 64-sample render (wavefronts of 16 samples).  Error: relMSE of the 64-spp image
    against the FP64 oracle at `--reference-spp` on a fixed 16 x 16 stratified pixel subset.  CPU
    baseline: the reference integrator restated in C (oracle, float) on 4096 pixels x 64 spp."""
    import numpy as np
    import torch

    import vt
    w = h = 512
    spp = 64
    mesh = vt.cornell()
    scene = ctx.set_mesh(mesh)
    scene.set_materials(vt.cornell_materials(17.0))
    r = vt.Renderer(ctx)
    eye, target = vt.orbit_eye(), (0.0, -0.1, 0.0)
    r.render(scene, w, h, eye, target, 42.0, max_depth=8, samples=16, sample_index=1)  # warm-up
    torch.cuda.synchronize()
    r.reset()
    t0 = time.perf_counter()
    r.render(scene, w, h, eye, target, 42.0, max_depth=8, samples=spp, sample_index=1)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    img = r.readback()[..., :3].astype(np.float64)
    ys, xs = np.meshgrid((np.arange(16) * 32 + 16), (np.arange(16) * 32 + 16), indexing="ij")
    pix = (ys * w + xs).ravel().astype(np.uint32)
    out = {"workload": "cornell-36tri-512x512-64spp-depth8", "gpu_seconds": round(gpu_s, 4),
           "gpu_msamples_per_s": round(w * h * spp / gpu_s / 1e6, 2), "mean_radiance": float(img.mean())}
    if not args.no_cpu:
        import oracle
        sc = oracle.Scene(mesh.vertices, mesh.faces)
        frame = dict(width=w, height=h, eye=eye, target=target, fov=42.0, max_depth=8)
        c0 = time.perf_counter()
        ref = oracle.render(sc, CORNELL_MATERIALS, frame, 1, args.reference_spp, pixels=pix, precision="f64")
        ref_s = time.perf_counter() - c0
        sub = img.reshape(-1, 3)[pix]
        out["relmse_vs_fp64"] = relmse(sub, ref)
        out["fp64_reference"] = {"spp": args.reference_spp, "pixels": int(len(pix)), "seconds": round(ref_s, 2)}
        cpix = np.arange(0, w * h, 64, dtype=np.uint32)
        c0 = time.perf_counter()
        oracle.render(sc, CORNELL_MATERIALS, frame, 1, spp, pixels=cpix, precision="f32")
        cs = time.perf_counter() - c0
        out["cpu_baseline"] = {"value": round(len(cpix) * spp / cs / 1e6, 3), "unit": "Msamples/s",
                               "cores": oracle.max_threads(), "kind": "port",
                               "sample": f"every 64th pixel ({len(cpix)} pixels) x {spp} spp, reference integrator in C"}
    r.close()
    scene.close()
    return out


def bench_learned_cache(args, ctx):
    """Config 5: the emitter terrain with the OpenPBR glossy / metal / transmissive mix (~10M
    triangles, 4096 emitters in the light tree), 1920 x 1080, depth 8, MIS, 1 sample per frame.

    For each budget of B plain frames (`--cache-budgets`, default 16, 64 and 256): B frames of the
    plain tracer, each timed (synchronised) and accumulated; then, from a cold start, a fresh cache's
    frames -- online training included -- accumulated until the plain run's time is spent.  The
    headline ratio is the cold start's at the first budget.  A warm variant (the cache trained over
    `--cache-warmup` frames first, their time reported but not charged: an interactive session that
    has been running) is reported at the first budget.  Error: relMSE on the 64 x 36 stratified pixel
    subset against the committed FP64 oracle image at 65536 spp (tests/golden/terrain_openpbr_ref.npz,
    made by tests/golden/make_terrain_reference.py)."""
    import numpy as np
    import torch

    import vt
    from vt import workloads
    path = os.path.join(ROOT, "tests", "golden", "terrain_openpbr_ref.npz")
    with np.load(path) as z:
        pix, ref, depth = z["pixels"], z["radiance"], int(z["depth"])
        w, h = int(z["width"]), int(z["height"])
    mesh, mats = workloads.emitter_terrain("openpbr")
    scene = ctx.set_mesh(mesh)
    scene.set_materials(workloads.as_materials(mats))
    view = (workloads.TERRAIN_EYE, workloads.TERRAIN_TARGET, workloads.TERRAIN_FOV)
    lr = args.cache_lr

    def frame(r, k, cache=0):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r.render(scene, w, h, *view, max_depth=depth, samples=1, sample_index=k, cache=cache, learning_rate=lr)
        torch.cuda.synchronize()
        return time.perf_counter() - t

    def error(r):
        return relmse(r.readback()[..., :3].astype(np.float64).reshape(-1, 3)[pix], ref)

    def plain(frames):
        r = vt.Renderer(ctx)
        for k in range(3):
            frame(r, 1000 + k)
        r.reset()
        t = sum(frame(r, 1 + k) for k in range(frames))
        e = error(r)
        r.close()
        return e, t

    def learned(budget, warm):
        cache = vt.NRC(ctx, seed=1)
        r = vt.Renderer(ctx)
        r.set_cache(cache)
        t_warm = sum(frame(r, 5000 + k, cache=1) for k in range(warm))
        if warm:
            r.reset()
        spent, n = 0.0, 0
        while spent < budget:
            spent += frame(r, 1 + n, cache=1)
            n += 1
        e = error(r)
        trained.append(cache.params)
        r.close()
        cache.close()
        return e, n, spent, t_warm

    trained = []

    budgets = []
    for frames in args.cache_budgets:
        e_plain, t_plain = plain(frames)
        e_cold, n_cold, t_cold, _ = learned(t_plain, 0)
        budgets.append({"plain": {"spp": frames, "seconds": round(t_plain, 4), "relmse": e_plain,
                                  "ms_per_frame": round(t_plain / frames * 1e3, 3)},
                        "cold_start": {"spp": n_cold, "seconds": round(t_cold, 4), "relmse": e_cold,
                                       "ms_per_frame": round(t_cold / n_cold * 1e3, 3)},
                        "relmse_ratio": e_plain / e_cold})
        log(f"learned cache, {frames} plain frames ({t_plain * 1e3:.1f} ms): plain relMSE {e_plain:.4e}; "
            f"cold-start cache {n_cold} frames relMSE {e_cold:.4e}; ratio {e_plain / e_cold:.2f}")
    first = budgets[0]
    e_warm, n_warm, t_nrc, t_warm = learned(first["plain"]["seconds"], args.cache_warmup)
    scene.close()
    cpu = None
    if not args.no_cpu:
        # Reference integrator (MIS) on the relMSE pixel subset, and the cache-terminated estimator
        # with the warm cache's trained weights at a quarter of the resolution, both in C (oracle).
        import oracle
        osc = oracle.Scene(mesh.vertices, mesh.faces)
        full = dict(width=w, height=h, eye=view[0], target=view[1], fov=view[2], max_depth=depth)
        spp_cpu = 64
        setup = render_setup(oracle.render, osc, mats, full, 1, 1, pixels=pix[:1])
        dt, img = timed_render(oracle.render, osc, mats, full, 1, spp_cpu, pixels=pix)
        dt -= setup
        small = dict(full, width=w // 4, height=h // 4, indirect_reuse=1)
        setup_c = render_setup(oracle.render_nrc, osc, mats, dict(full, width=1, height=1), 1, 1, trained[-1])
        dt_c, img_c = timed_render(oracle.render_nrc, osc, mats, small, 1, 2, trained[-1])
        dt_c -= setup_c
        cpu = {"value": round(len(pix) * spp_cpu / dt / 1e6, 4), "unit": "Msamples/s", "cores": oracle.max_threads(),
               "kind": "port", "mean_radiance": float(img.mean()),
               "sample": f"the {len(pix)}-pixel relMSE subset x {spp_cpu} spp, reference MIS integrator in C "
                         f"(per-call set-up, {setup:.2f} s, measured on one pixel sample and left out)",
               "cache": {"value": round(small["width"] * small["height"] * 2 / dt_c / 1e6, 4), "unit": "Msamples/s",
                         "mean_radiance": float(img_c.mean()),
                         "sample": f"2 frames at {small['width']}x{small['height']}, the cache-terminated estimator "
                                   f"with the warm cache's weights (fixed), in C"}}
        del osc
        log(f"learned cache CPU baseline: {cpu}")
    # The cache's inference kernel against the matrix cores' dense bf16 peak: 2M queries (a 1080p
    # frame's worth), timed with HIP events on the launch stream.
    cache = vt.NRC(ctx, seed=1)
    stream = torch.cuda.current_stream()
    x = torch.rand((1 << 21, 11), device="cuda")
    y = cache.infer(x, stream=stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(10):
        cache.infer(x, out=y, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    infer_ms = e0.elapsed_time(e1) / 10
    cache.close()
    flop = 2.0 * x.shape[0] * (64 * 265 + 3 * 64 * 64 + 4 * 64)
    infer_tflops = flop / (infer_ms * 1e-3) / 1e12
    return {"workload": f"emitter-terrain-openpbr-{len(mesh.faces)}tri-{w}x{h}-depth{depth}",
            "reference": {"kind": "FP64 oracle", "spp": int(np.load(path)["spp"]), "pixels": int(len(pix))},
            "learning_rate": lr, "train_period": 16,
            "relmse_ratio_cold_start": first["relmse_ratio"],
            "gpu_msamples_per_s": {"plain": round(w * h * first["plain"]["spp"] / first["plain"]["seconds"] / 1e6, 2),
                                   "cold_start": round(w * h * first["cold_start"]["spp"] / first["cold_start"]["seconds"] / 1e6, 2)},
            "cpu_baseline": cpu,
            "budgets": budgets,
            "warm": {"plain_spp": first["plain"]["spp"], "spp": n_warm, "seconds": round(t_nrc, 4), "relmse": e_warm,
                     "warmup_frames": args.cache_warmup, "warmup_seconds": round(t_warm, 3),
                     "relmse_ratio": first["plain"]["relmse"] / e_warm},
            "inference_roofline": {"bound": "mfma", "achieved": round(infer_tflops, 1), "peak": 2500.0,
                                   "unit": "TFLOP/s", "frac": round(infer_tflops / 2500.0, 4),
                                   "queries": int(x.shape[0]), "ms": round(infer_ms, 4),
                                   "flops_per_query": 2 * (64 * 265 + 3 * 64 * 64 + 4 * 64)}}


def bench_restir(args, ctx):
    """Config 3: the emitter terrain with Lambert and GGX surfaces (~10M triangles, 4096 emitters in
    the light tree), 1920 x 1080, depth 8, ReSTIR DI (32 candidates, 1 temporal + 1 spatial reuse)
    against MIS, `--restir-frames` frames of 1 spp each, both timed per frame (synchronised), each with
    the per-pixel PCG streams and with the Z++ sampler; and ReSTIR DI + GI (the renderer's default
    mode-0 estimator: GI at the first indirect diffuse vertex).  Error: relMSE on the 64 x 36
    stratified pixel subset against the committed FP64 oracle image (MIS at 65536 spp,
    tests/golden/terrain_restir_ref.npz); ReSTIR DI converges to the same image, ReSTIR GI to one
    slightly brighter (the bias of its uniform spatial taps)."""
    import numpy as np
    import torch

    import vt
    from vt import workloads
    path = os.path.join(ROOT, "tests", "golden", "terrain_restir_ref.npz")
    with np.load(path) as z:
        pix, ref, depth = z["pixels"], z["radiance"], int(z["depth"])
        w, h, ref_spp = int(z["width"]), int(z["height"]), int(z["spp"])
    mesh, mats = workloads.emitter_terrain("restir")
    scene = ctx.set_mesh(mesh)
    scene.set_materials(workloads.as_materials(mats))
    view = (workloads.TERRAIN_EYE, workloads.TERRAIN_TARGET, workloads.TERRAIN_FOV)
    out = {"workload": f"emitter-terrain-restir-{len(mesh.faces)}tri-{w}x{h}-depth{depth}",
           "reference": {"kind": "FP64 oracle (MIS)", "spp": ref_spp, "pixels": int(len(pix))}}
    # *_z: the same with the Z++ sampler (sampler=1) in place of the per-pixel PCG streams.
    for name, mode, sampler, reuse in (("mis", 1, 0, 1), ("restir_di", 0, 0, 1), ("mis_z", 1, 1, 1),
                                       ("restir_di_z", 0, 1, 1), ("restir_gi", 0, 0, 0)):
        r = vt.Renderer(ctx)
        for k in range(2):
            r.render(scene, w, h, *view, max_depth=depth, samples=1, sample_index=900 + k, sampling_mode=mode,
                     candidates=32, sampler=sampler, indirect_reuse=reuse)
        r.reset()
        spent = 0.0
        for k in range(args.restir_frames):
            torch.cuda.synchronize()
            t = time.perf_counter()
            r.render(scene, w, h, *view, max_depth=depth, samples=1, sample_index=1 + k, sampling_mode=mode,
                     candidates=32, sampler=sampler, indirect_reuse=reuse)
            torch.cuda.synchronize()
            spent += time.perf_counter() - t
        full_img = r.readback()[..., :3].astype(np.float64)
        img = full_img.reshape(-1, 3)[pix]
        out[name] = {"frames": args.restir_frames, "ms_per_frame": round(spent / args.restir_frames * 1e3, 3),
                     "msamples_per_s": round(w * h * args.restir_frames / spent / 1e6, 2),
                     "relmse": relmse(img, ref), "mean_radiance": float(full_img.mean())}
        r.close()
    scene.close()
    if not args.no_cpu:
        # The reference's two ReSTIR DI passes and its MIS integrator restated in C (oracle): ReSTIR
        # at a quarter of the resolution (reuse needs whole frames), MIS on the relMSE pixel subset.
        import oracle
        osc = oracle.Scene(mesh.vertices, mesh.faces)
        full = dict(width=w, height=h, eye=view[0], target=view[1], fov=view[2], max_depth=depth)
        small = dict(full, width=w // 4, height=h // 4)
        frames = 2
        setup_r = render_setup(oracle.render_restir, osc, mats, dict(small, width=1, height=1), 1, 1, candidates=32)
        dt_r, img_r = timed_render(oracle.render_restir, osc, mats, small, 1, frames, candidates=32)
        dt_r -= setup_r
        spp_cpu = 64
        setup_m = render_setup(oracle.render, osc, mats, full, 1, 1, pixels=pix[:1])
        dt_m, img_m = timed_render(oracle.render, osc, mats, full, 1, spp_cpu, pixels=pix)
        dt_m -= setup_m
        out["cpu_baseline"] = {
            "value": round(small["width"] * small["height"] * frames / dt_r / 1e6, 4), "unit": "Msamples/s",
            "cores": oracle.max_threads(), "kind": "port", "mean_radiance": float(img_r.mean()),
            "sample": f"{frames} ReSTIR DI frames (32 candidates, temporal + spatial reuse) at "
                      f"{small['width']}x{small['height']}, same scene and camera, in C (per-call set-up, "
                      f"{setup_r:.2f} s, measured on a 1x1 frame and left out)",
            "mis": {"value": round(len(pix) * spp_cpu / dt_m / 1e6, 4), "unit": "Msamples/s",
                    "mean_radiance": float(img_m.mean()),
                    "sample": f"the {len(pix)}-pixel relMSE subset x {spp_cpu} spp, MIS integrator in C"}}
        del osc
        log(f"restir CPU baseline: {out['cpu_baseline']}")
    out["relmse_ratio_mis_over_restir"] = out["mis"]["relmse"] / out["restir_di"]["relmse"]
    out["efficiency_ratio"] = (out["mis"]["relmse"] * out["mis"]["ms_per_frame"]) / \
        (out["restir_di"]["relmse"] * out["restir_di"]["ms_per_frame"])
    out["relmse_ratio_pcg_over_z"] = {"mis": out["mis"]["relmse"] / out["mis_z"]["relmse"],
                                      "restir_di": out["restir_di"]["relmse"] / out["restir_di_z"]["relmse"]}
    out["efficiency_ratio_gi"] = (out["mis"]["relmse"] * out["mis"]["ms_per_frame"]) / \
        (out["restir_gi"]["relmse"] * out["restir_gi"]["ms_per_frame"])
    log(f"restir: MIS {out['mis']}, ReSTIR DI {out['restir_di']}, Z sampler: MIS {out['mis_z']}, "
        f"ReSTIR DI {out['restir_di_z']}, ReSTIR DI + GI {out['restir_gi']}")
    return out


def bench_batch(args, ctx):
    """Config 4: 64 icosphere-field scenes (triangle counts log-uniform over 10k-2M, seed 11) at image
    sizes drawn from {256^2, 512^2, 1024x576, 1920x1080}, 16 spp, depth 6, MIS, one persistent
    wavefront per sample captured in a hipGraph.  The autotuner (capacity, persistent grid, traversal
    work-group, and with --batch-tune-leaf the BVH leaf size) runs first; its time is reported apart.
    `value` of this part: camera paths (pixel samples) per second over the 16 timed samples."""
    import numpy as np
    import torch

    import vt
    from vt import workloads
    t0 = time.perf_counter()
    work = workloads.batch_workload(count=args.batch_scenes)
    scenes, jobs = [], []
    for mesh, mats, w, h, eye, target, fov in work:
        s = ctx.set_mesh(mesh)
        s.set_materials(workloads.as_materials(mats))
        scenes.append(s)
        jobs.append((s, w, h, eye, target, fov))
    build_s = time.perf_counter() - t0
    tris = sum(len(m.faces) for m, *_ in work)
    pixels = sum(w * h for _, _, w, h, *_ in work)
    b = vt.Batch(ctx, jobs, max_depth=6)
    t0 = time.perf_counter()
    best, best_ms = b.autotune(leaf_size=args.batch_tune_leaf)
    tune_s = time.perf_counter() - t0
    spp = 16
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.render(spp)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    images = [b.readback(k)[..., :3] for k in range(len(jobs))]
    mean = float(np.mean([im.mean() for im in images]))
    b.close()
    for s in scenes:
        s.close()
    log(f"batch: {len(jobs)} scenes, {tris} triangles, {pixels} pixels a sample: {spp} spp in {dt * 1e3:.1f} ms "
        f"({pixels * spp / dt / 1e6:.1f} Mpaths/s), tuning {best}")
    cpu = None
    if not args.no_cpu:
        # Per-scene sequential path tracing with the reference integrator in C (oracle): every 509th
        # pixel of each image at 2 spp, scene after scene (hierarchy builds untimed); the total for
        # the whole batch (every pixel, 16 spp) is extrapolated from the sampled paths' rate.
        import oracle
        spent, paths, acc, acc_gpu = 0.0, 0, 0.0, 0.0
        for k, (mesh, mats, w, h, eye, target, fov) in enumerate(work):
            osc = oracle.Scene(mesh.vertices, mesh.faces)
            cpix = np.arange(0, w * h, 509, dtype=np.uint32)
            acc_gpu += float(images[k].reshape(-1, 3)[cpix].astype(np.float64).sum())
            frame = dict(width=w, height=h, eye=eye, target=target, fov=fov, max_depth=6)
            if k == 0:
                oracle.render(osc, mats, frame, 1, 1, pixels=cpix[:1])   # the first call's one-time costs
            cpu_dt, img = timed_render(oracle.render, osc, mats, frame, 1, 2, pixels=cpix)
            setup, _ = timed_render(oracle.render, osc, mats, frame, 1, 1, pixels=cpix[:1])
            spent += cpu_dt - setup
            paths += len(cpix) * 2
            acc += float(img.sum()) * 2
            del osc
        rate = paths / spent
        cpu = {"value": round(rate / 1e6, 4), "unit": "Mpaths/s", "cores": oracle.max_threads(), "kind": "port",
               "sequential_seconds_extrapolated": round(pixels * spp / rate, 1),
               "mean_radiance": acc / (3 * paths), "gpu_mean_radiance_same_pixels": acc_gpu / (3 * paths / 2),
               "sample": f"every 509th pixel of each of the {len(work)} images at 2 spp ({paths} paths), scenes one "
                         f"after another, reference MIS integrator in C (each call's set-up, measured on one pixel "
                         f"sample, left out); the full batch's time extrapolated"}
        log(f"batch CPU baseline: {cpu}")
    return {"workload": f"batch-{len(jobs)}-icosphere-scenes-{tris}tri-{pixels}px-16spp-depth6",
            "scenes": len(jobs), "triangles": int(tris), "pixels_per_sample": int(pixels), "spp": spp,
            "seconds": round(dt, 4), "mpaths_per_s": round(pixels * spp / dt / 1e6, 1),
            "ms_per_sample": round(dt / spp * 1e3, 3), "tuning": best, "tuned_ms_per_sample": round(best_ms, 3),
            "autotune_seconds": round(tune_s, 2), "scene_build_seconds": round(build_s, 2), "mean_radiance": mean,
            "cpu_baseline": cpu}


def instanced_field(count=1024, assets=4, seed=7, half_extent=50.0):
    """Two-level version of the icosphere field: `assets` unit icospheres (1280 triangles each, 3
    subdivisions), `count` instances with a random rotation, a log-normal radius (sigma 0.5) and a
    position uniform in the box.  Returns (meshes, worlds (count, 3, 4), asset of each instance)."""
    import numpy as np

    import vt
    rng = np.random.default_rng(seed)
    meshes = []
    for a in range(assets):
        m = vt.icosphere_field(spheres=1, subdivisions=3, mu=0.0, sigma=0.0, half_extent=0.0, seed=seed + a)
        v = m.vertices.copy()
        c = v[:, :3].mean(0)
        v[:, :3] -= c
        v[:, :3] /= np.sqrt((v[:, :3] ** 2).sum(1)).max()
        meshes.append(vt.Mesh(v.astype(np.float32), m.faces))
    worlds = np.zeros((count, 3, 4), np.float32)
    for k in range(count):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        a, b, c, d = q
        R = np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                      [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                      [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])
        worlds[k, :, :3] = R * np.exp(rng.normal(0.0, 0.5))
        worlds[k, :, 3] = rng.uniform(-half_extent, half_extent, 3)
    return meshes, worlds, np.arange(count) % assets


def flatten_instances(meshes, worlds, which):
    """The instanced scene as one mesh: world positions in double rounded to float, instance after
    instance (flattened face = the instance's first face + the asset's face)."""
    import numpy as np

    import vt
    vs, fs, base = [], [], 0
    for w, a in zip(worlds, which):
        v, f = meshes[a].vertices, meshes[a].faces
        vw = np.zeros_like(v)
        vw[:, :3] = (v[:, :3].astype(np.float64) @ w[:, :3].T.astype(np.float64) + w[:, 3]).astype(np.float32)
        f2 = f.copy()
        f2[:, :3] += base
        vs.append(vw)
        fs.append(f2)
        base += len(v)
    return vt.Mesh(np.concatenate(vs), np.concatenate(fs))


def bench_instances(args, ctx):
    """Two-level scenes: 1024 instances of 4 unit icospheres (the icosphere field's density, 1.31 M
    rendered triangles from 5120 stored), 2048 x 1024 camera rays and their bounce rays through
    vt_tlas, beside the same rays through the flattened mesh's one-level hierarchy.  Each kernel time
    is the mean of 10 launches (HIP events on the launch stream).  CPU baseline: the flattened mesh
    through the reference 4-wide traversal (oracle, C)."""
    import numpy as np
    import torch

    import vt
    dev = torch.device("cuda:0")
    meshes, worlds, which = instanced_field()
    scenes = [ctx.set_mesh(m) for m in meshes]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tl = vt.Tlas(ctx, [(scenes[a], w) for w, a in zip(worlds, which)])
    torch.cuda.synchronize()
    tlas_s = time.perf_counter() - t0
    flat = flatten_instances(meshes, worlds, which)
    fscene = ctx.set_mesh(flat)
    prim = vt.camera_rays((0.0, 30.0, -170.0), (0.0, 0.0, 0.0), 45.0, 2048, 1024, 7)
    stream = torch.cuda.current_stream()
    pd = torch.from_numpy(prim).to(dev)
    h = fscene.trace_closest(pd)
    torch.cuda.synchronize()
    hp = {k: v.cpu().numpy() for k, v in h.items()}
    bounce = vt.bounce_rays(flat, prim, hp["face"], hp["b1"], hp["b2"], 7)
    bd = torch.from_numpy(bounce).to(dev)

    def timed(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(10):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 10

    out = {}
    for label, rd, n in (("camera", pd, prim.shape[1]), ("bounce", bd, bounce.shape[1])):
        ho = tl.trace_closest(rd, stream=stream)
        fo = fscene.trace_closest(rd, stream=stream)
        ms2 = timed(lambda: tl.trace_closest(rd, out=ho, stream=stream))
        ms1 = timed(lambda: fscene.trace_closest(rd, out=fo, stream=stream))
        torch.cuda.synchronize()
        same_t = float((ho["t"] == fo["t"]).float().mean().item())
        out[label] = {"rays": int(n), "two_level_ms": round(ms2, 4), "two_level_mrays_per_s": round(n / ms2 / 1e3, 1),
                      "flat_ms": round(ms1, 4), "flat_mrays_per_s": round(n / ms1 / 1e3, 1),
                      "hits": int((ho["face"] >= 0).sum().item()), "t_equal_to_flat": round(same_t, 5)}
    info = tl.info
    res = {"workload": f"instanced-icosphere-field-{len(worlds)}x{len(meshes)}", "instances": len(worlds),
           "assets": len(meshes), "rendered_triangles": int(info["rendered_triangles"]),
           "stored_triangles": int(sum(len(m.faces) for m in meshes)), "tlas_nodes": int(info["node_count"]),
           "tlas_build_seconds": round(tlas_s, 4), **out}
    if not args.no_cpu:
        # The same world geometry flattened into one mesh (the reference's graph-less path) through the
        # reference 4-wide traversal in C: the CPU's rate on this scene.
        import oracle
        osc = oracle.Scene(flat.vertices, flat.faces)
        sample = np.concatenate([prim, bounce], axis=1)
        osc.closest(sample, mode=0)
        times = []
        for _ in range(5):
            c0 = time.perf_counter()
            osc.closest(sample, mode=0)
            times.append(time.perf_counter() - c0)
        res["cpu_baseline"] = {"value": round(sample.shape[1] / float(np.median(times)) / 1e6, 3), "unit": "Mrays/s",
                               "cores": oracle.max_threads(), "kind": "port",
                               "sample": f"every camera and bounce ray ({sample.shape[1]} rays) through the flattened "
                                         f"world mesh by the reference 4-wide BVH and traversal in C, median of 5"}
        del osc
    log(f"instances: {res}")
    tl.close()
    fscene.close()
    for sc in scenes:
        sc.close()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["heightfield", "stress"], default="heightfield",
                    help="the headline traversal workload (the other one follows as a secondary block)")
    ap.add_argument("--no-secondary", action="store_true", help="skip the secondary traversal workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-integrator", action="store_true", help="skip the path-tracer (Cornell) measurement")
    ap.add_argument("--reference-spp", type=int, default=65536, help="FP64 oracle samples per pixel for relMSE")
    ap.add_argument("--no-cache", action="store_true", help="skip the learned-cache equal-time measurement")
    ap.add_argument("--cache-budgets", type=int, nargs="+", default=[16, 64, 256],
                    help="plain-tracer frames (1 spp each) setting each equal-time budget")
    ap.add_argument("--cache-warmup", type=int, default=64, help="online training frames before the equal-time run")
    ap.add_argument("--cache-lr", type=float, default=1e-3, help="Adam learning rate of the online training")
    ap.add_argument("--no-restir", action="store_true", help="skip the ReSTIR DI large-scene measurement")
    ap.add_argument("--restir-frames", type=int, default=16, help="frames (1 spp each) of the ReSTIR measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-scene measurement")
    ap.add_argument("--no-instances", action="store_true", help="skip the two-level (instanced) measurement")
    ap.add_argument("--batch-scenes", type=int, default=64, help="scenes in the batched measurement")
    ap.add_argument("--batch-tune-leaf", action="store_true", help="let the batch autotuner sweep the BVH leaf size")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device-to-device 3 times with a plain copy kernel (the "
                         "reference launch for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import vt

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    if not args.no_cpu:
        import oracle
        oracle.load(oracle.native_library())     # the CPU baseline, built for this host
    ctx = vt.Context(0)
    bandwidth = measure_bandwidth(dev)
    log(f"measured copy bandwidth: {bandwidth}")
    first, second = (workload_heightfield, workload_stress) if args.workload == "heightfield" else \
        (workload_stress, workload_heightfield)
    head = measure_traversal(args, ctx, *first())
    secondary = None if args.no_secondary else measure_traversal(args, ctx, *second())
    # Fractions of the measured copy rates: the HBM one for every workload, and the Infinity Cache
    # one for the workload whose hierarchy fits in it (the 1.3M-triangle field, 113 MB).
    for block in (head, secondary):
        if block:
            ach = block["roofline"]["achieved"]
            block["roofline"]["frac_of_measured"] = {"hbm_copy": round(ach / bandwidth["hbm_copy_gbs"], 4),
                                                     "mall_copy": round(ach / bandwidth["mall_copy_gbs"], 4)}

    integrator = None if args.no_integrator else bench_integrator(args, ctx)
    learned_cache = None if args.no_cache else bench_learned_cache(args, ctx)
    restir = None if args.no_restir else bench_restir(args, ctx)
    batch = None if args.no_batch else bench_batch(args, ctx)
    instances = None if args.no_instances else bench_instances(args, ctx)

    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": head["config"],
        "bvh_build_seconds": head["bvh_build_seconds"],
        "kernel_ms": head["kernel_ms"],
        "mrays_per_s": head["mrays_per_s"],
        "roofline": head["roofline"],
        "incoherent": head["incoherent"],
        "cpu_baseline": head["cpu_baseline"],
        "any_hit": head["any_hit"],
        "roofline_reference_4wide": head["roofline_reference_4wide"],
        "bandwidth_measured": bandwidth,
        "secondary": secondary,
        "integrator": integrator,
        "learned_cache": learned_cache,
        "restir": restir,
        "batch": batch,
        "instances": instances,
    }
    print(json.dumps(line), flush=True)
    if args.copy_calibration:
        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 x {args.copy_calibration} MiB device-to-device")
    ctx.close()


if __name__ == "__main__":
    main()
