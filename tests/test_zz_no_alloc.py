"""hx_web_pack_device, the packing entry point that allocates nothing (its
tile counts in a static device array), through the C ABI on device buffers:
two calls in a row, the second on a smaller grid, each bit-identical to
hx_web_pack (the entry point that allocates) on the same inputs."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Dev:
    """Device buffers through the library's test hooks."""

    def __init__(self, ctx):
        self.ctx, self.L, self.bufs = ctx, ctx.L, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hx_test_device_alloc(self.ctx.h, max(a.nbytes, 16), C.byref(p)) == 0
        self.bufs.append(p)
        assert self.L.hx_test_memcpy(self.ctx.h, p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.L.hx_test_memcpy(self.ctx.h, out.ctypes.data_as(C.c_void_p), p, out.nbytes, 0) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hx_test_device_free(self.ctx.h, p)
        self.bufs = []


def _case(ctx, dims, seed, thr):
    ctx.set_grid(*dims, 0.7, 1.3, 0.9, nu=0.3)
    ne = ctx.ne
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0, 1, ne).astype(np.float32)
    rho[rng.choice(ne, ne // 50, replace=False)] = np.float32(thr)  # equality: kept
    rho[3] = np.nan  # never kept
    l1 = rng.uniform(0, 5, ne).astype(np.float32)
    pr = rng.standard_normal((ne, 3)).astype(np.float32)
    return rho, l1, pr


def test_web_pack_device_matches_web_pack_twice(hx):
    ctx = hx.Context(0)
    dev = _Dev(ctx)
    try:
        for dims, seed, thr in (((64, 40, 36), 1, 0.3), ((19, 11, 7), 2, 0.55)):  # second call: a smaller grid
            rho, l1, pr = _case(ctx, dims, seed, thr)
            ne = ctx.ne
            cent, dens, sl1 = ctx.web_pack(rho, l1, pr, thr)
            d_rho, d_l1, d_pr = dev.put(rho), dev.put(l1), dev.put(pr)
            nanf = np.full(3 * ne, np.nan, np.float32)
            d_cent, d_dens, d_sl1 = dev.put(nanf), dev.put(nanf[:ne]), dev.put(nanf[:ne])
            d_cnt = dev.put(np.full(1, 0xFFFFFFFF, np.uint32))
            rc = ctx.L.hx_web_pack_device(ctx.h, d_rho, d_l1, d_pr, C.c_float(np.float32(thr)), d_cent, d_dens, d_sl1,
                                          d_cnt)
            ctx.check(rc, "hx_web_pack_device")
            n = int(dev.get(d_cnt, np.zeros(1, np.uint32))[0])
            assert n == dens.size == int((rho >= np.float32(thr)).sum())
            c2 = dev.get(d_cent, nanf).reshape(ne, 3)
            assert np.array_equal(c2[:n], cent) and np.all(np.isnan(c2[n:]))
            assert np.array_equal(dev.get(d_dens, nanf[:ne])[:n], dens)
            assert np.array_equal(dev.get(d_sl1, nanf[:ne])[:n], sl1)
            dev.free()
    finally:
        dev.free()
        ctx.close()
