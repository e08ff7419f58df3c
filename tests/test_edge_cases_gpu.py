"""GPU edge cases of the renderer, the learned cache and batched rendering: a scene without geometry,
one-pixel and one-vertex images, a wavefront with no cache training record, a batch of one job and a
batch job whose scene is empty."""
import numpy as np
import pytest

import vt

pytestmark = pytest.mark.gpu
TARGET = (0.0, -0.1, 0.0)


def _empty(ctx):
    s = ctx.set_mesh(vt.Mesh(np.zeros((0, 8), np.float32), np.zeros((0, 4), np.uint32)))
    s.set_materials([vt.material(0)])
    return s


def test_empty_scene_renders_black(gpu_ctx):
    s = _empty(gpu_ctx)
    r = vt.Renderer(gpu_ctx)
    for mode in (0, 1, 2, 3):
        r.reset()
        r.render(s, 16, 8, vt.orbit_eye(), TARGET, 42.0, max_depth=4, samples=2, sample_index=1, sampling_mode=mode)
        img = r.readback()
        assert np.all(img[..., :3] == 0) and np.all(np.isfinite(img))
    r.close()
    s.close()


def test_one_pixel_and_direct_light_only(gpu_ctx):
    s = gpu_ctx.set_mesh(vt.cornell())
    s.set_materials(vt.cornell_materials(17.0))
    r = vt.Renderer(gpu_ctx)
    r.render(s, 1, 1, vt.orbit_eye(), TARGET, 42.0, max_depth=1, samples=64, sample_index=1)
    one = r.readback()
    assert one.shape == (1, 1, 4) and np.all(np.isfinite(one)) and one[0, 0, :3].sum() > 0
    # max_depth 1 is direct light only: never brighter on average than the full path.
    r2 = vt.Renderer(gpu_ctx)
    r.render(s, 32, 32, vt.orbit_eye(), TARGET, 42.0, max_depth=1, samples=16, sample_index=1)
    r2.render(s, 32, 32, vt.orbit_eye(), TARGET, 42.0, max_depth=8, samples=16, sample_index=1)
    assert r.readback()[..., :3].mean() < r2.readback()[..., :3].mean()
    r.close()
    r2.close()
    s.close()


def test_cache_wavefront_without_training_records(gpu_ctx):
    # 4 x 4 pixels with one training path in 4096: most wavefronts make no record; the Adam steps are
    # skipped on the device and the parameters must not change.
    s = gpu_ctx.set_mesh(vt.cornell())
    s.set_materials(vt.cornell_materials(17.0))
    cache = vt.NRC(gpu_ctx, seed=4)
    before = cache.params
    r = vt.Renderer(gpu_ctx)
    r.set_cache(cache)
    made = []
    for k in range(6):
        r.render(s, 4, 4, vt.orbit_eye(), TARGET, 42.0, max_depth=8, samples=1, sample_index=1 + k, cache=1,
                 train_period=4096)
        made.append(r.cache_records()[2])
    img = r.readback()
    assert np.all(np.isfinite(img))
    if sum(made) == 0:
        assert np.array_equal(cache.params, before)
    print(f"records per wavefront {made}", flush=True)
    r.close()
    cache.close()
    s.close()


def test_batch_of_one_job_and_an_empty_scene(gpu_ctx):
    s = gpu_ctx.set_mesh(vt.cornell())
    s.set_materials(vt.cornell_materials(17.0))
    e = _empty(gpu_ctx)
    b = vt.Batch(gpu_ctx, [(s, 24, 16, vt.orbit_eye(), TARGET, 42.0)], max_depth=4)
    b.render(3)
    r = vt.Renderer(gpu_ctx)
    r.render(s, 24, 16, vt.orbit_eye(), TARGET, 42.0, max_depth=4, samples=3, sample_index=1)
    ref = r.readback()
    # Same estimator as the single-image renderer, up to floating-point contraction differences.
    assert np.isclose(b.readback(0), ref, rtol=1e-4, atol=1e-6).all(axis=2).mean() > 0.999
    b.close()
    b = vt.Batch(gpu_ctx, [(e, 8, 8, vt.orbit_eye(), TARGET, 42.0), (s, 24, 16, vt.orbit_eye(), TARGET, 42.0)],
                 max_depth=4)
    b.render(3)
    assert np.all(b.readback(0)[..., :3] == 0)
    assert np.isclose(b.readback(1), ref, rtol=1e-4, atol=1e-6).all(axis=2).mean() > 0.999
    b.close()
    r.close()
    e.close()
    s.close()


def test_completion_callbacks(gpu_ctx):
    # onFrameUpdate / onError stand-ins: each render call queues on_frame with the accumulated sample
    # count after its work; a refused call reports its message through on_error before returning.
    import threading
    import torch
    scene = gpu_ctx.set_mesh(vt.cornell())
    scene.set_materials(vt.cornell_materials(17.0))
    r = vt.Renderer(gpu_ctx)
    frames, errors, threads = [], [], set()

    def on_frame(n):
        frames.append(n)
        threads.add(threading.get_ident())

    r.set_callbacks(on_frame, errors.append)
    r.render(scene, 48, 48, vt.orbit_eye(), (0.0, -0.1, 0.0), 42.0, max_depth=4, samples=3, sample_index=1)
    r.render(scene, 48, 48, vt.orbit_eye(), (0.0, -0.1, 0.0), 42.0, max_depth=4, samples=2, sample_index=4)
    torch.cuda.synchronize()
    for _ in range(100):
        if len(frames) == 2:
            break
        threading.Event().wait(0.01)
    assert frames == [3, 5] and r.frames == 5
    with pytest.raises(vt.VTError):
        r.render(scene, 48, 48, vt.orbit_eye(), (0.0, -0.1, 0.0), 42.0, max_depth=0, samples=1)
    assert len(errors) == 1 and "max_depth" in errors[0]
    r.set_callbacks(None, None)
    r.close()
    scene.close()
