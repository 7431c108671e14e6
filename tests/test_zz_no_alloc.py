"""The allocation-free entry point (qpso_estimate_pose_noalloc) against
qpso_estimate_pose, through the C ABI: two calls in a row, the second with a
smaller input (fewer particles and iterations, a smaller rect), each bit-exact
against the allocating call; the device's free memory does not move across a
call; a solve larger than the scratch is refused and leaves it usable."""
import ctypes as C

import numpy as np
import pytest

import qpso
import workloads as W

pytestmark = pytest.mark.gpu


def same(a, b):
    return (np.array_equal(np.array(a.pose[:], np.float32).view(np.uint32),
                           np.array(b.pose[:], np.float32).view(np.uint32))
            and np.float32(a.fit).view(np.uint32) == np.float32(b.fit).view(np.uint32)
            and a.raster_px == b.raster_px and a.visit_px == b.visit_px
            and a.n_degenerate == b.n_degenerate and a.status == b.status)


def noalloc(ctx, mid, cl, p):
    r = qpso.Result()
    rc = qpso.lib().qpso_estimate_pose_noalloc(ctx.handle, mid, C.byref(cl), C.byref(p), C.byref(r))
    return rc, r


def free_device_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_noalloc_matches_estimate_pose():
    scene1, p1 = W.config1(0, n_particles=256, n_iters=20)  # 220x220 rect, 10,336 faces
    scene0, p0 = W.config0(0)  # smaller: 32 particles, 30 iterations, 1,280 faces
    print("noalloc: config1 (256 x 20) then config0 (32 x 30)", flush=True)
    with qpso.Context(0) as ctx:
        m1 = ctx.upload_mesh(scene1.mesh_m())
        m0 = ctx.upload_mesh(scene0.mesh_m())
        c1, c0 = scene1.cluster(), scene0.cluster()
        want1 = ctx.estimate_pose(m1, c1, p1)
        want0 = ctx.estimate_pose(m0, c0, p0)
        rc, got1 = noalloc(ctx, m1, c1, p1)
        assert rc == qpso.OK, qpso.lib().qpso_last_error(ctx.handle)
        assert same(got1, want1)
        before = free_device_bytes()
        rc, got0 = noalloc(ctx, m0, c0, p0)
        assert rc == qpso.OK and same(got0, want0)
        assert free_device_bytes() == before  # no device allocation in the call
        # larger than the scratch: refused with a clear error, scratch untouched
        big = qpso.make_params(2048, 200, w=0.7, c1=1.5, c2=1.5, cam=W.CAM, seed=0)
        rc, _ = noalloc(ctx, m1, c1, big)
        assert rc == qpso.E_NOMEM
        assert b"no-allocation scratch" in qpso.lib().qpso_last_error(ctx.handle)
        rc, again = noalloc(ctx, m1, c1, p1)
        assert rc == qpso.OK and same(again, want1)
        ctx.release_mesh(m1)
        ctx.release_mesh(m0)
