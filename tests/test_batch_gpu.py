"""GPU tests of batched rendering (vt_batch_*, metal-vibe-tracer_amd/csrc/batch.cpp): every job of a
batch of uneven scenes and image sizes is the single-image renderer's estimator sample for sample
(same pixel streams, same traversal), whatever the tuning -- chunk capacity, persistent grid,
traversal work-group size, BVH leaf size -- and the autotuner ends on a valid tuning with the
accumulation reset.  The single-image renderer is itself checked against the oracle elsewhere."""
import numpy as np
import pytest

import scenes
import vt
from vt import workloads

pytestmark = pytest.mark.gpu
DEPTH = 6


def _scenes(ctx):
    out = []
    mesh, mats, half = workloads.icosphere_scene(12_000, 5)
    out.append((ctx.set_mesh(mesh), mats, half))
    mesh, mats, half = workloads.icosphere_scene(60_000, 6)
    out.append((ctx.set_mesh(mesh), mats, half))
    mesh, mats = scenes.materials_cornell("opaque")
    out.append((ctx.set_mesh(mesh), mats, None))
    for s, m, _ in out:
        s.set_materials(scenes.vt_materials(m))
    return out


def _jobs(sc):
    jobs = []
    for (scene, _, half), (w, h) in zip(sc + sc[:1], [(64, 48), (96, 96), (80, 72), (72, 40)]):
        if half is None:
            eye, target, fov = vt.orbit_eye(), (0.0, -0.1, 0.0), 42.0
        else:
            eye, target, fov = (0.0, 0.6 * half, -3.4 * half), (0.0, 0.0, 0.0), 45.0
        jobs.append((scene, w, h, eye, target, fov))
    # The fourth job views the first scene from elsewhere.
    s0, w, h, eye, target, fov = jobs[3]
    jobs[3] = (s0, w, h, (eye[0] + 5.0, eye[1], eye[2]), target, fov)
    return jobs


def _single(ctx, job, spp):
    scene, w, h, eye, target, fov = job
    r = vt.Renderer(ctx)
    r.render(scene, w, h, eye, target, fov, max_depth=DEPTH, samples=spp, sample_index=1)
    img = r.readback()
    r.close()
    return img


def test_batch_matches_single_renders(gpu_ctx):
    sc = _scenes(gpu_ctx)
    jobs = _jobs(sc)
    b = vt.Batch(gpu_ctx, jobs, max_depth=DEPTH)
    spp = 4
    b.render(spp)
    assert b.frames == spp
    for k, job in enumerate(jobs):
        got = b.readback(k)
        ref = _single(gpu_ctx, job, spp)
        close = np.isclose(got, ref, rtol=1e-4, atol=1e-6).all(axis=2).mean()
        print(f"job {k} {job[1]}x{job[2]}: batch mean {got[..., :3].mean():.6f} single {ref[..., :3].mean():.6f} "
              f"max |diff| {np.abs(got - ref).max():.2e} close {close:.4f}", flush=True)
        assert got[..., :3].mean() > 0
        assert abs(got[..., :3].mean() / ref[..., :3].mean() - 1) < 1e-4 and close > 0.999
    b.close()
    for s, _, _ in sc:
        s.close()


def test_tuning_does_not_change_the_image(gpu_ctx):
    sc = _scenes(gpu_ctx)
    jobs = _jobs(sc)
    b = vt.Batch(gpu_ctx, jobs, max_depth=DEPTH)
    b.render(2)
    base = [b.readback(k) for k in range(len(jobs))]
    for tuning in (dict(capacity=4096, grid_waves=1, trace_block=64, leaf_size=0),
                   dict(capacity=0, grid_waves=8, trace_block=256, leaf_size=1),
                   dict(capacity=0, grid_waves=4, trace_block=512, leaf_size=0)):
        b.tuning = tuning
        b.reset()
        b.render(2)
        got = [b.readback(k) for k in range(len(jobs))]
        same = all(np.array_equal(g, r) for g, r in zip(got, base))
        print(f"tuning {b.tuning}: identical {same}", flush=True)
        assert same
    b.close()
    for s, _, _ in sc:
        s.close()


def test_autotune_and_errors(gpu_ctx):
    sc = _scenes(gpu_ctx)
    jobs = _jobs(sc)
    b = vt.Batch(gpu_ctx, jobs, max_depth=DEPTH)
    best, ms = b.autotune(leaf_size=True)
    print(f"autotune: {best}, {ms:.3f} ms per sample", flush=True)
    assert ms > 0 and best["trace_block"] in (64, 128, 256, 512) and 1 <= best["leaf_size"] <= 3
    assert b.tuning == best and b.frames == 0
    b.render(2)
    ref = _single(gpu_ctx, jobs[1], 2)
    assert np.isclose(b.readback(1), ref, rtol=1e-4, atol=1e-6).all(axis=2).mean() > 0.999
    b.close()
    # Delta materials and ReSTIR are refused.
    mesh, mats = scenes.materials_cornell("delta")
    delta = gpu_ctx.set_mesh(mesh)
    delta.set_materials(scenes.vt_materials(mats))
    with pytest.raises(vt.VTError):
        vt.Batch(gpu_ctx, [(delta, 32, 32, vt.orbit_eye(), (0.0, -0.1, 0.0), 42.0)])
    with pytest.raises(vt.VTError):
        vt.Batch(gpu_ctx, jobs[:1], sampling_mode=0)
    delta.close()
    for s, _, _ in sc:
        s.close()
