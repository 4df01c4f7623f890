"""ggem_local_velocity_points_noalloc (include/copss_ggem_noalloc.h): the
field-point local velocity with its staging in the fixed per-GPU device
array, against ggem_local_velocity_points on the same context: equal bit
for bit (the same kernels on the same inputs), on two calls in a row, the
second with fewer points and the elements located on the device; the size
limit and the argument errors of the plain entry point."""
import ctypes as C

import numpy as np
import pytest

import copss_ggem as G
import workloads as W

pytestmark = pytest.mark.gpu


def noalloc(ctx, pts, elem=None):
    L = G.lib()
    fn = L.ggem_local_velocity_points_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    el = None if elem is None else np.ascontiguousarray(elem, dtype=np.int32)
    u = np.zeros_like(pts)
    rc = fn(ctx._h, pts.shape[0], pts.ctypes.data, None if el is None else el.ctypes.data, u.ctypes.data)
    if rc != 0:
        raise G.GGEMError(rc, L.ggem_last_error(ctx._h).decode())
    return u


@pytest.mark.parametrize("prec", [G.FP64, G.FP32C])
def test_noalloc_equals_plain_entry_point(prec):
    cfg, x, f = W.make_inputs("beads20k")
    ctx = G.Context(dict(cfg, precision=prec))
    try:
        ctx.reinit(x, f)
        rng = np.random.default_rng(31)
        pts1 = rng.random((60000, 3)) * 200.0
        ref1 = ctx.local_velocity_points(pts1)
        got1 = noalloc(ctx, pts1)
        assert np.array_equal(got1, ref1)
        pts2 = rng.random((1234, 3)) * 200.0
        ids2 = np.floor(pts2 / (200.0 / 64)).astype(np.int64)
        el2 = (ids2[:, 0] + 64 * (ids2[:, 1] + 64 * ids2[:, 2])).astype(np.int32)
        ref2 = ctx.local_velocity_points(pts2, el2)
        got2 = noalloc(ctx, pts2, el2)
        assert np.array_equal(got2, ref2)
        got2b = noalloc(ctx, pts2)  # located on the device, after a larger call
        assert np.array_equal(got2b, ctx.local_velocity_points(pts2))
        assert np.abs(got1).max() > 0
    finally:
        ctx.close()


def test_noalloc_limits_and_errors():
    cfg, x, f = W.make_inputs("beads100")
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x, f)
        with pytest.raises(G.GGEMError) as ei:
            noalloc(ctx, np.full((65537, 3), 5.0))
        assert ei.value.code == -8  # GGEM_ERR_CAPACITY
        with pytest.raises(G.GGEMError) as ei:
            noalloc(ctx, np.array([[1.0, 1.0, 1.0]]), np.array([8000], np.int32))
        assert ei.value.code == -3  # invalid element id
        with pytest.raises(G.GGEMError) as ei:
            noalloc(ctx, np.array([[1.0, 1.0, 1.0], [41.0, 1.0, 1.0]]))
        assert ei.value.code == -2  # outside the mesh
        u = noalloc(ctx, np.full((65536, 3), 5.0))  # the largest call
        assert np.array_equal(u[0], ctx.local_velocity_points(np.array([[5.0, 5.0, 5.0]]))[0])
    finally:
        ctx.close()
