"""GPU test of the allocation-free entry point vt_trace_stats_noalloc (include/vt.h): its counters and
results against the allocating vt_trace_stats and vt_trace_closest / vt_trace_any, through the C ABI,
on two calls in a row, the second with fewer rays (the static scratch keeps the first call's
partial counters, which the second must clear)."""
import numpy as np
import pytest

import scenes
import vt

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.mark.parametrize("any_hit", [False, True])
def test_noalloc_stats_match_the_allocating_entry_point(gpu_ctx, any_hit):
    import torch
    v, f = scenes.scattered_triangles(3000, seed=21)
    scene = gpu_ctx.set_mesh(vt.Mesh(v, f))
    big = scenes.random_rays(20000, 31)
    for n in (20000, 1500):
        rays = _dev(big[:, :n])
        want = scene.trace_stats(rays, any_hit=any_hit)
        counters = torch.full((2,), -7, dtype=torch.int64, device="cuda:0")
        got = scene.trace_stats_noalloc(rays, counters, any_hit=any_hit)
        torch.cuda.synchronize()
        print(f"n={n} any={any_hit}: counters {counters.tolist()} vs {list(want)}", flush=True)
        assert tuple(counters.tolist()) == want and want[0] >= n
        if any_hit:
            ref = scene.trace_any(rays)
            torch.cuda.synchronize()
            assert torch.equal(got, ref)
        else:
            ref = scene.trace_closest(rays)
            torch.cuda.synchronize()
            for k in ("t", "b1", "b2", "face"):
                assert torch.equal(got[k], ref[k]), k
    scene.close()
