"""GPU parity of the Z++ sampler (vt_z_draws, metal-vibe-tracer_amd/csrc/zsampler.h) against the
oracle (oracle/zsampler.c): every draw bit-identical, for 1D / 2D / 3D draws, frames 0 and later,
the four temporal models, events of several vertices and streams, and candidates."""
import numpy as np
import pytest

import oracle
import vt

pytestmark = pytest.mark.gpu

CASES = [
    dict(frame_index=1, sample_index=1, temporal_model=0, path_vertex=0, stream=0, draws=4, kind=1),
    dict(frame_index=1, sample_index=1, temporal_model=1, path_vertex=1, stream=1, draws=3, kind=2),
    dict(frame_index=5, sample_index=9, temporal_model=2, path_vertex=2, stream=4, draws=2, kind=3),
    dict(frame_index=37, sample_index=40, temporal_model=3, path_vertex=7, stream=12, draws=5, kind=2),
    dict(frame_index=16385, sample_index=16390, temporal_model=0, path_vertex=3, stream=9, draws=2, kind=3),
    dict(frame_index=3, sample_index=3, temporal_model=0, path_vertex=1, stream=4, candidate=5, candidate_bits=5,
         draws=2, kind=3),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_z_draws_match_the_oracle(gpu_ctx, case):
    import torch
    q = CASES[case]
    w, h = 61, 37
    got = vt.z_draws(w, h, **q)
    torch.cuda.synchronize()
    ref, _ = oracle.z_draws(w, h, **q)
    got = got.cpu().numpy()
    print(f"case {case}: {q}, mean {got.mean():.4f}", flush=True)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_z_draws_errors(gpu_ctx):
    with pytest.raises(vt.VTError):
        vt.z_draws(4, 4, 1, 1, kind=4)
    with pytest.raises(vt.VTError):
        vt.z_draws(4, 4, 1, 1, candidate=4, candidate_bits=2)
