"""ccd_ctcd_narrowphase_noalloc (no counterpart in the reference; for callers that cannot allocate)
through the C ABI: the same flags, time-of-impact bits and firing sub-tests as ccd_ctcd_narrowphase,
on two calls in a row (the second smaller, so the shared device scratch holds the first call's
contents), with no growth of the context's device memory; calls beyond the stated maximum and
invalid keys are refused."""
import ctypes

import numpy as np
import pytest

from collisiondetection_amd import scenes

pytestmark = pytest.mark.gpu


def _run(fn, ctx, qs, qe, V, vk, ek, eta):
    import torch

    dev = ctx.torch_device
    nvf, nee = vk.numel() // 4, ek.numel() // 4
    out = [torch.empty(n, dtype=dt, device=dev) for n, dt in ((nvf, torch.uint8), (nvf, torch.float64),
                                                                (nvf, torch.uint8), (nee, torch.uint8),
                                                                (nee, torch.float64), (nee, torch.uint8))]
    ctx.bind_stream()
    rc = fn(ctx.ptr, qs.data_ptr(), qe.data_ptr(), V, vk.data_ptr(), None, nvf, ek.data_ptr(), None, nee,
            float(eta), *[o.data_ptr() for o in out])
    torch.cuda.synchronize(dev)
    return rc, [o.cpu().numpy() for o in out]


def test_noalloc_narrowphase_equals_allocating_one():
    import torch

    from collisiondetection_amd import _native as N
    from collisiondetection_amd import api

    L = N.lib()
    ctx = api.Context(0)          # used only through the allocation-free entry point
    ref_ctx = api.Context(0)
    base_bytes = ctx.device_bytes()
    for n_grid in (110, 40):      # the second call is smaller than the first
        s = scenes.cloth(n_grid, drop=0.05, noise=0.01)
        print("case: no-alloc narrowphase, cloth", n_grid, flush=True)
        vf, ee = api.kdop_candidates(s["faces"], s["qstart"], s["qend"], s["eta"], ctx=ref_ctx, as_numpy=False)
        nvf, nee = vf.shape[0], ee.shape[0]
        assert 4096 <= nvf <= N.NOALLOC_MAX_STENCILS and 4096 <= nee <= N.NOALLOC_MAX_STENCILS, (nvf, nee)
        vk, ek = vf.reshape(-1).contiguous(), ee.reshape(-1).contiguous()
        qs = api.to_device(s["qstart"], torch.float64, ref_ctx).reshape(-1)
        qe = api.to_device(s["qend"], torch.float64, ref_ctx).reshape(-1)
        V = len(s["qstart"])
        rc0, ref = _run(L.ccd_ctcd_narrowphase, ref_ctx, qs, qe, V, vk, ek, s["eta"])
        rc1, got = _run(L.ccd_ctcd_narrowphase_noalloc, ctx, qs, qe, V, vk, ek, s["eta"])
        assert rc0 == N.CCD_OK and rc1 == N.CCD_OK, (rc0, rc1)
        for a, b, name in zip(ref, got, ("vf_hit", "vf_t", "vf_sub", "ee_hit", "ee_t", "ee_sub")):
            if a.dtype == np.float64:
                a, b = a.view(np.int64), b.view(np.int64)
            assert np.array_equal(a, b), name
        print(n_grid, "stencils", nvf, nee, "hits", int(ref[0].sum()), int(ref[3].sum()), flush=True)
        assert ref[0].sum() + ref[3].sum() > 0
        assert ctx.device_bytes() == base_bytes   # nothing grew on the no-alloc context
    # beyond the maximum: refused before anything is read
    big = torch.zeros(4 * (N.NOALLOC_MAX_STENCILS + 1), dtype=torch.int32, device=ctx.torch_device)
    none = torch.zeros(4, dtype=torch.int32, device=ctx.torch_device)
    rc, _ = _run(L.ccd_ctcd_narrowphase_noalloc, ctx, qs, qe, V, big, none[:0], s["eta"])
    assert rc == N.CCD_E_CAPACITY
    assert b"at most" in L.ccd_last_error(ctx.ptr)
    # invalid keys: refused by the device-side validation
    bad = vk[:4096 * 4].clone()
    bad[5] = V
    rc, _ = _run(L.ccd_ctcd_narrowphase_noalloc, ctx, qs, qe, V, bad, none[:0], s["eta"])
    assert rc == N.CCD_E_INVALID
    assert ctx.device_bytes() == base_bytes
