"""GPU tests of batched rendering (vt_batch_*, metal-vibe-tracer_amd/csrc/batch.cpp): every job of a
batch of uneven scenes and image sizes is the oracle's restatement of the reference integrator
(oracle/integrator.c) pixel for pixel up to float order, and the single-image renderer's estimator
sample for sample (same pixel streams, same traversal), whatever the tuning -- chunk capacity,
persistent grid, traversal work-group size, BVH leaf size; the autotuner ends on a valid tuning with
the accumulation reset; and the full 64-scene workload renders finite images whose jobs equal their
single-image renders."""
import numpy as np
import pytest

import oracle
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


def test_batch_jobs_match_oracle(gpu_ctx):
    # Each job against the oracle's MIS integrator with the same camera and sample indices 1..spp:
    # the same sample streams, so pixels agree up to float-order differences.
    sc = _scenes(gpu_ctx)
    mats_of = {id(s): (m, s) for s, m, _ in sc}
    jobs = _jobs(sc)
    b = vt.Batch(gpu_ctx, jobs, max_depth=DEPTH)
    spp = 4
    b.render(spp)
    meshes = [workloads.icosphere_scene(12_000, 5)[0], workloads.icosphere_scene(60_000, 6)[0],
              scenes.materials_cornell("opaque")[0]]
    mesh_of = {id(s): m for (s, _, _), m in zip(sc, meshes)}
    for k, (scene, w, h, eye, target, fov) in enumerate(jobs):
        mats, _ = mats_of[id(scene)]
        mesh = mesh_of[id(scene)]
        frame = dict(width=w, height=h, eye=eye, target=target, fov=fov, max_depth=DEPTH)
        ref = oracle.render(oracle.Scene(mesh.vertices, mesh.faces), mats, frame, 1, spp).reshape(h, w, 3)
        got = b.readback(k)[..., :3].astype(np.float64)
        rel = abs(got.mean() / ref.mean() - 1)
        close = np.isclose(got, ref, rtol=1e-3, atol=1e-4).all(axis=2).mean()
        print(f"job {k} {w}x{h}: batch {got.mean():.6f} oracle {ref.mean():.6f} rel {rel:.2e} close {close:.4f}",
              flush=True)
        assert rel < 5e-3 and close > 0.95
    b.close()
    for s, _, _ in sc:
        s.close()


def test_full_batch_workload_properties(gpu_ctx):
    # The benchmark's 64 icosphere scenes (10k-2M triangles, four image sizes) in one batch, 1 spp:
    # every image finite and lit, and the four smallest jobs equal to their single-image renders.
    work = workloads.batch_workload(count=64)
    scs, jobs = [], []
    for mesh, mats, w, h, eye, target, fov in work:
        s = gpu_ctx.set_mesh(mesh)
        s.set_materials(workloads.as_materials(mats))
        scs.append(s)
        jobs.append((s, w, h, eye, target, fov))
    b = vt.Batch(gpu_ctx, jobs, max_depth=DEPTH)
    b.render(1)
    means = []
    for k in range(len(jobs)):
        img = b.readback(k)
        assert np.isfinite(img).all(), k
        means.append(float(img[..., :3].mean()))
    print(f"64 jobs: mean radiance {np.mean(means):.5f}, min {min(means):.5f}", flush=True)
    assert min(means) > 0
    for k in sorted(range(len(jobs)), key=lambda j: jobs[j][1] * jobs[j][2])[:4]:
        ref = _single(gpu_ctx, jobs[k], 1)
        assert np.isclose(b.readback(k), ref, rtol=1e-4, atol=1e-6).all(axis=2).mean() > 0.999, k
    b.close()
    for s in scs:
        s.close()


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


def test_scene_rebuilt_under_a_batch(gpu_ctx):
    # A batched scene rebuilt with another leaf size (new node and triangle buffers, another depth)
    # between renders: the batch recaptures and renders what a freshly created batch renders.
    sc = _scenes(gpu_ctx)
    jobs = _jobs(sc)
    b = vt.Batch(gpu_ctx, jobs, max_depth=DEPTH)
    b.render(1)
    leaf = sc[1][0].info["leaf_size"]
    sc[1][0].rebuild(leaf_size=1 if leaf != 1 else 3)
    assert sc[1][0].info["leaf_size"] != leaf
    b.reset()
    b.render(2)
    got = [b.readback(k) for k in range(len(jobs))]
    b.close()
    fresh = vt.Batch(gpu_ctx, jobs, max_depth=DEPTH)
    fresh.render(2)
    for k in range(len(jobs)):
        assert np.array_equal(got[k], fresh.readback(k)), k
    fresh.close()
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


def test_batch_with_material_layers_matches_single_renders(gpu_ctx):
    # Coats, anisotropic specular lobes and fuzz go through the batch kernels' layered shading too: each
    # job equals its single-image render (the anisotropic scene's tangents come from its vertices' uv).
    made = []
    for variant in ("coated", "anisotropic", "fuzzy"):
        mesh, mats = scenes.materials_cornell(variant)
        s = gpu_ctx.set_mesh(mesh)
        s.set_materials(scenes.vt_materials(mats))
        made.append(s)
    jobs = [(s, 64, 64, vt.orbit_eye(), (0.0, -0.1, 0.0), 42.0) for s in made]
    b = vt.Batch(gpu_ctx, jobs, max_depth=DEPTH)
    b.render(4)
    for k, job in enumerate(jobs):
        got, ref = b.readback(k), _single(gpu_ctx, job, 4)
        close = np.isclose(got, ref, rtol=1e-4, atol=1e-6).all(axis=2).mean()
        print(f"layers job {k}: batch mean {got[..., :3].mean():.6f} single {ref[..., :3].mean():.6f} close {close:.4f}",
              flush=True)
        assert got[..., :3].mean() > 0 and close > 0.999
    b.close()
    for s in made:
        s.close()
