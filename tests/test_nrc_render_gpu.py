"""GPU tests of the cache-terminated estimator (vt_render_frame with radiance_cache): with the cache
frozen at fixed weights it matches the oracle's restatement (oracle/integrator.c, vo_render_nrc) --
same training-path subset, same termination vertex, same network inputs -- and trained online it
beats the plain tracer in relative MSE against the FP64 oracle at equal render time."""
import time

import numpy as np
import pytest

import oracle
from oracle import nrc as onrc
import scenes
import vt

pytestmark = pytest.mark.gpu
TARGET = (0.0, -0.1, 0.0)


def test_frozen_cache_matches_oracle(gpu_ctx):
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
    r.render(scene, w, w, vt.orbit_eye(), TARGET, 42.0, max_depth=8, samples=spp, sample_index=1, cache=2)
    got = r.readback()[..., :3].astype(np.float64)
    assert np.array_equal(cache.params, params)             # frozen
    sc = oracle.Scene(mesh.vertices, mesh.faces)
    frame = dict(width=w, height=w, eye=vt.orbit_eye(), target=TARGET, fov=42.0, max_depth=8)
    ref = oracle.render_nrc(sc, scenes.CORNELL_MATS, frame, 1, spp, params)
    rel = abs(got.mean() / ref.mean() - 1)
    close = np.isclose(got, ref, rtol=1e-2, atol=1e-4).all(axis=2).mean()
    print(f"frozen cache: gpu {got.mean():.6f} oracle {ref.mean():.6f} rel {rel:.2e} close {close:.4f}", flush=True)
    assert rel < 3e-3 and close > 0.95
    r.close()
    scene.close()
    cache.close()


def _relmse(img, ref):
    return float(((img - ref) ** 2 / (ref ** 2 + 0.01)).mean())


def test_trained_cache_beats_plain_tracer_at_equal_time(gpu_ctx):
    import torch
    mesh = vt.cornell()
    w = 128
    sc = oracle.Scene(mesh.vertices, mesh.faces)
    frame = dict(width=w, height=w, eye=vt.orbit_eye(), target=TARGET, fov=42.0, max_depth=8)
    pix = (np.arange(16)[:, None] * (w // 16) + w // 32) * w + (np.arange(16)[None, :] * (w // 16) + w // 32)
    ref = oracle.render(sc, scenes.CORNELL_MATS, frame, 100000, 8192, pixels=pix.reshape(-1), precision="f64")
    scene = gpu_ctx.set_mesh(mesh)
    scene.set_materials(vt.cornell_materials(17.0))

    def run(use_cache, budget=None, spp=None):
        r = vt.Renderer(gpu_ctx)
        cache = None
        if use_cache:
            cache = vt.NRC(gpu_ctx, seed=1)
            r.set_cache(cache)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n = 0
        while (spp is not None and n < spp) or (budget is not None and time.perf_counter() - t0 < budget):
            r.render(scene, w, w, vt.orbit_eye(), TARGET, 42.0, max_depth=8, samples=1, sample_index=1 + n,
                     cache=1 if use_cache else 0)
            n += 1
            torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        img = r.readback()[..., :3].astype(np.float64).reshape(-1, 3)[pix.reshape(-1)]
        r.close()
        if cache:
            cache.close()
        return img, n, dt

    run(False, spp=2)                                         # warm up both paths
    run(True, spp=2)
    plain, n_plain, t_plain = run(False, spp=64)
    learned, n_nrc, t_nrc = run(True, budget=t_plain)
    e_plain, e_nrc = _relmse(plain, ref), _relmse(learned, ref)
    print(f"equal time {t_plain:.3f}s: plain {n_plain} spp relMSE {e_plain:.3e}; cache {n_nrc} spp "
          f"({t_nrc:.3f}s) relMSE {e_nrc:.3e}; ratio {e_plain / e_nrc:.2f}", flush=True)
    assert e_nrc < e_plain
    scene.close()
