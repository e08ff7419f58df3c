"""GPU tests of the cache-terminated estimator (vt_render_frame with radiance_cache): with the cache
frozen at fixed weights it matches the oracle's restatement (oracle/integrator.c, vo_render_nrc) --
same training-path subset, same termination vertex, same network inputs -- on the Lambert Cornell box
and on the learned-cache workload's OpenPBR terrain, and trained online it beats the plain tracer in
relative MSE against the FP64 oracle at equal render time, on glossy Cornell and on the full 10M
terrain at 1080p."""
import os
import time

import numpy as np
import pytest

import oracle
from oracle import nrc as onrc
import scenes
import vt
from vt import workloads

pytestmark = pytest.mark.gpu
TARGET = (0.0, -0.1, 0.0)


@pytest.mark.parametrize("sampler", [0, 1])
def test_frozen_cache_matches_oracle(gpu_ctx, sampler):
    """sampler=1: the Z++ sampler's events as on the plain path; the cache adds no draws of its own
    (its training subset is a hash of pixel and sample)."""
    mesh = vt.cornell()
    rng = np.random.default_rng(3)
    p = rng.uniform(-0.12, 0.12, onrc.N_PARAMS)
    Ws = onrc.split(p)
    Ws[0][:, 265:] = 0
    Ws[4][4:] = 0
    params = np.concatenate([w.reshape(-1) for w in Ws]).astype(np.float32)
    cache = vt.NRC(gpu_ctx)
    cache.params = params
    scene = gpu_ctx.set_mesh(mesh)
    scene.set_materials(vt.cornell_materials(17.0))
    r = vt.Renderer(gpu_ctx)
    r.set_cache(cache)
    w, spp = 96, 16
    r.render(scene, w, w, vt.orbit_eye(), TARGET, 42.0, max_depth=8, samples=spp, sample_index=1, cache=2,
             sampler=sampler)
    got = r.readback()[..., :3].astype(np.float64)
    assert np.array_equal(cache.params, params)             # frozen
    sc = oracle.Scene(mesh.vertices, mesh.faces)
    frame = dict(width=w, height=w, eye=vt.orbit_eye(), target=TARGET, fov=42.0, max_depth=8, sampler=sampler)
    ref = oracle.render_nrc(sc, scenes.CORNELL_MATS, frame, 1, spp, params)
    rel = abs(got.mean() / ref.mean() - 1)
    close = np.isclose(got, ref, rtol=1e-2, atol=1e-4).all(axis=2).mean()
    print(f"frozen cache sampler {sampler}: gpu {got.mean():.6f} oracle {ref.mean():.6f} rel {rel:.2e} close {close:.4f}", flush=True)
    assert rel < 3e-3 and close > 0.95
    r.close()
    scene.close()
    cache.close()


def _frozen_params(seed):
    rng = np.random.default_rng(seed)
    Ws = onrc.split(rng.uniform(-0.12, 0.12, onrc.N_PARAMS))
    Ws[0][:, 265:] = 0
    Ws[4][4:] = 0
    return np.concatenate([w.reshape(-1) for w in Ws]).astype(np.float32)


def test_frozen_cache_matches_oracle_on_openpbr_terrain(gpu_ctx):
    # The learned-cache workload's scene at a reduced grid (129^2): glossy, metal and transmissive
    # OpenPBR slots under 4096 emitters in the light tree, the cache frozen at fixed weights.
    mesh, mats = workloads.emitter_terrain("openpbr", grid=129)
    params = _frozen_params(5)
    cache = vt.NRC(gpu_ctx)
    cache.params = params
    scene = gpu_ctx.set_mesh(mesh)
    scene.set_materials(workloads.as_materials(mats))
    r = vt.Renderer(gpu_ctx)
    r.set_cache(cache)
    w, h, spp, depth = 160, 90, 16, 6
    view = (workloads.TERRAIN_EYE, workloads.TERRAIN_TARGET, workloads.TERRAIN_FOV)
    r.render(scene, w, h, *view, max_depth=depth, samples=spp, sample_index=1, cache=2)
    got = r.readback()[..., :3].astype(np.float64)
    assert np.array_equal(cache.params, params)
    frame = dict(width=w, height=h, eye=view[0], target=view[1], fov=view[2], max_depth=depth)
    ref = oracle.render_nrc(oracle.Scene(mesh.vertices, mesh.faces), mats, frame, 1, spp, params)
    rel = abs(got.mean() / ref.mean() - 1)
    close = np.isclose(got, ref, rtol=1e-2, atol=1e-4).all(axis=2).mean()
    print(f"frozen cache, OpenPBR terrain: gpu {got.mean():.6f} oracle {ref.mean():.6f} rel {rel:.2e} "
          f"close {close:.4f}", flush=True)
    assert rel < 5e-3 and close > 0.9
    r.close()
    scene.close()
    cache.close()


def test_trained_cache_beats_plain_on_10m_terrain(gpu_ctx):
    """The learned-cache workload itself: the OpenPBR terrain (10M triangles, 4096 emitters) at
    1920 x 1080, depth 8.  From a cold start (training included in its time), frames of the cache
    are accumulated until the plain tracer's time for 16 frames is spent; relMSE against the
    committed FP64 oracle image (65536 spp on 64 x 36 pixels) must be at least 1.5 x lower."""
    import torch
    path = os.path.join(os.path.dirname(__file__), "golden", "terrain_openpbr_ref.npz")
    with np.load(path) as z:
        pix, ref, depth, w, h = z["pixels"], z["radiance"], int(z["depth"]), int(z["width"]), int(z["height"])
    mesh, mats = workloads.emitter_terrain("openpbr")
    scene = gpu_ctx.set_mesh(mesh)
    scene.set_materials(workloads.as_materials(mats))
    view = (workloads.TERRAIN_EYE, workloads.TERRAIN_TARGET, workloads.TERRAIN_FOV)

    def timed(r, k, cache):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.render(scene, w, h, *view, max_depth=depth, samples=1, sample_index=k, cache=cache)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def err(r):
        return _relmse(r.readback()[..., :3].astype(np.float64).reshape(-1, 3)[pix], ref)

    plain = vt.Renderer(gpu_ctx)
    for k in range(3):
        timed(plain, 900 + k, 0)
    plain.reset()
    t_plain = sum(timed(plain, 1 + k, 0) for k in range(16))
    e_plain = err(plain)
    cache = vt.NRC(gpu_ctx, seed=1)
    learned = vt.Renderer(gpu_ctx)
    learned.set_cache(cache)
    spent, n = 0.0, 0
    while spent < t_plain:
        spent += timed(learned, 1 + n, 1)
        n += 1
    e_nrc = err(learned)
    print(f"10M terrain, equal time {t_plain * 1e3:.1f} ms: plain 16 spp relMSE {e_plain:.3e}; cold-start cache "
          f"{n} spp relMSE {e_nrc:.3e}; ratio {e_plain / e_nrc:.2f}", flush=True)
    assert e_plain / e_nrc > 1.5
    plain.close()
    learned.close()
    cache.close()
    scene.close()


def _relmse(img, ref):
    return float(((img - ref) ** 2 / (ref ** 2 + 0.01)).mean())


def test_trained_cache_beats_plain_tracer_at_equal_time(gpu_ctx):
    """The Cornell box with glossy metal, half-metallic and plastic surfaces at 512 x 512: the cache,
    trained online over 32 frames, then frames accumulated until the plain tracer's time for 16
    frames is spent.  relMSE against the FP64 oracle (8192 spp, 16 x 16 pixels) must be at least
    1.5 x lower than the plain tracer's."""
    import torch
    mesh, mats = scenes.materials_cornell("opaque")
    w = 512
    sc = oracle.Scene(mesh.vertices, mesh.faces)
    frame = dict(width=w, height=w, eye=vt.orbit_eye(), target=TARGET, fov=42.0, max_depth=8)
    pix = (np.arange(16)[:, None] * (w // 16) + w // 32) * w + (np.arange(16)[None, :] * (w // 16) + w // 32)
    pix = pix.reshape(-1)
    ref = oracle.render(sc, mats, frame, 100000, 8192, pixels=pix, precision="f64")
    scene = gpu_ctx.set_mesh(mesh)
    scene.set_materials(scenes.vt_materials(mats))

    def timed(r, k, cache):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.render(scene, w, w, vt.orbit_eye(), TARGET, 42.0, max_depth=8, samples=1, sample_index=k, cache=cache)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def image(r):
        return r.readback()[..., :3].astype(np.float64).reshape(-1, 3)[pix]

    plain = vt.Renderer(gpu_ctx)
    for k in range(3):
        timed(plain, 900 + k, 0)
    plain.reset()
    t_plain = sum(timed(plain, 1 + k, 0) for k in range(16))
    e_plain = _relmse(image(plain), ref)
    cache = vt.NRC(gpu_ctx, seed=1)
    learned = vt.Renderer(gpu_ctx)
    learned.set_cache(cache)
    for k in range(32):
        timed(learned, 5000 + k, 1)
    learned.reset()
    spent, n = 0.0, 0
    while spent < t_plain:
        spent += timed(learned, 1 + n, 1)
        n += 1
    e_nrc = _relmse(image(learned), ref)
    print(f"equal time {t_plain * 1e3:.1f} ms: plain 16 spp relMSE {e_plain:.3e}; cache {n} spp relMSE {e_nrc:.3e}; "
          f"ratio {e_plain / e_nrc:.2f}", flush=True)
    assert e_plain / e_nrc > 1.5
    plain.close()
    learned.close()
    cache.close()
    scene.close()
