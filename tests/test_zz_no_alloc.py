"""The allocation-free trace entry point (arc_trace_pools_noalloc) against arc_trace_device on the
same pools and rays, through the C ABI: identical compact records on two calls in a row, the
second with a smaller input (and another ray class), plus the argument checks that refuse a call
before any launch."""
from __future__ import annotations

import numpy as np
import pytest

import arcadia_amd as A
from arcadia_amd import native as N
from tests import worlds as W

pytestmark = pytest.mark.gpu


def test_noalloc_trace_equals_context_trace(ctx):
    import torch

    pools = W.caves_world(1)
    grid, roots, nodes, bricks = pools
    ctx.upload(*pools)
    dev = [torch.from_numpy(np.ascontiguousarray(a).view(np.uint8)).cuda() for a in (roots, nodes, bricks)]
    cam = A.camera_look((13.7, 30.2, 11.9), 0.6, -0.55, 65.0, 480, 270)
    prays, _ = A.primary_rays(cam, 480, 270, 250.0)
    p0 = A.TraceParams.exact(0)
    d_p = torch.from_numpy(prays.view(np.uint8)).cuda()
    d_ph = torch.zeros(len(prays) * 16, dtype=torch.uint8, device="cuda")
    ctx.trace_device(p0, d_p.data_ptr(), len(prays), d_ph.data_ptr())
    torch.cuda.synchronize()
    brays, _ = A.bounce_rays(prays, d_ph.cpu().numpy().view(A.HIT_DTYPE), seed=3, max_dist_m=26.0)
    brays = brays[: len(prays) // 3]
    cases = [(p0, prays), (A.TraceParams.exact(2, ignore_water=True), brays)]
    assert len(cases[1][1]) < len(cases[0][1]) and len(brays) > 1000
    for k, (params, rays) in enumerate(cases):
        print(f"[noalloc] call {k}: {len(rays)} rays", flush=True)
        d_r = torch.from_numpy(np.ascontiguousarray(rays).view(np.uint8)).cuda()
        want = torch.zeros(len(rays) * 16, dtype=torch.uint8, device="cuda")
        got = torch.full((len(rays) * 16,), 0xAB, dtype=torch.uint8, device="cuda")
        ctx.trace_device(params, d_r.data_ptr(), len(rays), want.data_ptr())
        rc = A.trace_pools_noalloc(grid, dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(), params,
                                   d_r.data_ptr(), len(rays), got.data_ptr())
        assert rc == N.ARC_OK
        torch.cuda.synchronize()
        w, g = want.cpu().numpy().view(A.HIT_DTYPE), got.cpu().numpy().view(A.HIT_DTYPE)
        bad = W.diff_records(g, w)
        assert len(bad) == 0, f"call {k}: {len(bad)} of {len(rays)} records differ, first {bad[:5]}"
        assert (w["flags"] & 1).sum() > 100
    # refused before any launch
    args = (dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(), p0, d_p.data_ptr())
    assert A.trace_pools_noalloc(grid, *args, 1 << 31, d_ph.data_ptr()) == N.ARC_ERR_CAPACITY
    assert A.trace_pools_noalloc((0, 2, 2), *args, 10, d_ph.data_ptr()) == N.ARC_ERR_INVALID
    assert A.trace_pools_noalloc((257, 2, 2), *args, 10, d_ph.data_ptr()) == N.ARC_ERR_INVALID
    assert A.trace_pools_noalloc(grid, *args, 10, 0) == N.ARC_ERR_INVALID
    assert A.trace_pools_noalloc(grid, *args, 0, 0) == N.ARC_OK  # nothing to trace
