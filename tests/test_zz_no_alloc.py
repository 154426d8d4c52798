"""rt_trace_closest_noalloc (the allocation-free closest hit) against
rt_trace_closest on the same device buffers: hits identical bit for bit on
two calls in a row, the second with fewer rays, so the shared scratch array's
counters are cleared between calls."""
import numpy as np
import pytest

from master_thesis_amd import scenes
from master_thesis_amd._lib import RAY_DTYPE, RTError, RT_E_STATE

pytestmark = pytest.mark.gpu


def _device_rays(torch, rays):
    return torch.from_numpy(np.ascontiguousarray(rays).view(np.float32).reshape(-1).copy()).cuda()


def test_noalloc_trace_matches_trace_closest(tracer):
    import torch
    sc = scenes.triangle_soup(1 << 18, seed=5)
    tracer.upload(sc)
    tracer.build(4)
    first = np.concatenate(scenes.microbench_rays(512, 512, n_incoherent=1 << 18, seed=6))
    second = scenes.incoherent_rays(70001, seed=7)   # fewer rays, not a multiple of the block
    status = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    for rays in (first, second):
        print(f"noalloc trace: {rays.shape[0]} rays", flush=True)
        d_rays = _device_rays(torch, rays)
        n = rays.shape[0]
        want = tracer.trace(d_rays)
        got = torch.full((n * 4,), 7, dtype=torch.int32, device="cuda")
        status.fill_(-1)
        tracer.trace_noalloc(d_rays, got, status)
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        assert torch.equal(got, want)
        hit_ids = want.view(n, 4)[:, 0]
        assert 0 < int((hit_ids != -1).sum()) < n


def test_noalloc_trace_arguments(tracer):
    import torch
    from master_thesis_amd.tracer import Tracer
    with Tracer(0) as t:
        rays = _device_rays(torch, scenes.incoherent_rays(16, seed=1))
        hits = torch.empty(64, dtype=torch.int32, device="cuda")
        with pytest.raises(RTError) as e:
            t.trace_noalloc(rays, hits)
        assert e.value.code == RT_E_STATE
    tracer.upload(scenes.cornell_box())
    tracer.build(4)
    # no rays: nothing to do, and a NULL status is allowed
    assert tracer.lib.rt_trace_closest_noalloc(tracer.h, None, 0, None, None) == 0
    tracer.trace_noalloc(rays, hits)
    torch.cuda.synchronize()
    assert torch.equal(hits, tracer.trace(rays))
