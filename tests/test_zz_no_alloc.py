"""The allocation-free ray cast (ncx_trimesh_cast_rays_noalloc) against ncx_trimesh_cast_rays.

Two calls in a row through the C ABI, the second with fewer rays and another mesh pose: the
shared scratch keeps the first call's work counters and spill stacks, so the second call shows
that the entry point clears what must start at zero.  Outputs must be bit-identical.
"""
import numpy as np
import pytest

import ncollide_amd as ncx
from ncollide_amd import _lib, workloads

pytestmark = pytest.mark.gpu


def _cast_noalloc(ctx, mesh, pose, soa, n, max_toi):
    import torch
    dev = ctx.device
    toi = torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
    nrm = torch.full((n, 3), float("nan"), dtype=torch.float32, device=dev)
    feat = torch.full((n,), 7, dtype=torch.int32, device=dev)
    p = mesh._pose(pose)
    _lib.check(ctx.lib.ncx_trimesh_cast_rays_noalloc(mesh.h, _lib.ptr(p), _lib.ptr6(soa), n, float(max_toi),
                                                     _lib.ptr(toi), _lib.ptr(nrm), _lib.ptr(feat)))
    _lib.check(ctx.lib.ncx_trimesh_check(mesh.h))
    return toi.cpu().numpy(), nrm.cpu().numpy(), feat.cpu().numpy().view(np.uint32)


def test_noalloc_cast_matches_cast_rays(ctx):
    pts, tri = workloads.heightfield_mesh(seed=2)
    mesh = ncx.TriMesh(pts, tri, ctx=ctx)
    poses = [None, np.array([0.5, -0.25, 0.1, 0, 0, 0, np.sin(0.05), np.cos(0.05)], np.float32)]
    for call, (n, seed, pose) in enumerate([(200_000, 11, poses[0]), (30_000, 12, poses[1])]):
        o, d = workloads.heightfield_rays(seed=seed, n=n)
        soa = mesh._rays(o, d)
        ref = mesh.toi_and_normal_with_ray(None, None, 100.0, pose=pose, soa=soa)
        rt, rn, rf = ref[0].cpu().numpy(), ref[1].cpu().numpy(), ref[2].cpu().numpy().view(np.uint32)
        t, nn, f = _cast_noalloc(ctx, mesh, pose, soa, n, 100.0)
        print(f"call {call}: {n} rays, hits {np.mean(f != 0xFFFFFFFF):.3f}", flush=True)
        assert np.array_equal(f, rf), f"call {call}: feature ids differ"
        assert np.array_equal(t.view(np.uint32), rt.view(np.uint32)), f"call {call}: toi differs"
        assert np.array_equal(nn.view(np.uint32), rn.view(np.uint32)), f"call {call}: normals differ"
