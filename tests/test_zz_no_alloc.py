"""vvhd_convective_noalloc (the allocation-free convective entry point)
against vvhd_convective through the C ABI.

Both run the same kernels; the allocation-free one takes its scratch (per-leaf
Taylor terms, per-vortex near sums) from the library's fixed device array, so
the velocities must be identical bit for bit.  Two calls in a row, the second
on a smaller input (the array keeps the first call's contents), with Kahan and
plain accumulation; then an input whose scratch exceeds the array must be
refused with VVHD_EINVAL before any launch.
"""
import numpy as np
import pytest

import vvflow_amd as V

pytestmark = pytest.mark.gpu


def _velocities(ctx, objs, noalloc, kahan, fc=8):
    ctx.set_params(re=600.0, dt=0.005)
    ctx.upload_bodies(np.zeros(0, V.SEG_DTYPE), np.zeros(0, V.BODY_DTYPE))
    ctx.upload(objs)
    ctx.tree_build(fc)
    ctx.epsilon(True)
    if noalloc:
        ctx.convective_noalloc((1.0, 0.0), kahan=kahan)
    else:
        ctx.convective((1.0, 0.0), kahan=kahan)
    return ctx.download()


@pytest.mark.parametrize("kahan", [True, False])
def test_noalloc_equals_convective_twice(kahan):
    ctx = V.Context(0)
    try:
        for n, seed in ((300_000, 0x5EED0101), (40_000, 0x5EED0102)):
            print(f"noalloc vs convective: n={n} seed={seed:#x} kahan={kahan}", flush=True)
            objs = V.gen_wake(n, seed)
            want = _velocities(ctx, objs, False, kahan)
            got = _velocities(ctx, objs, True, kahan)
            assert np.isfinite(got).all()
            assert np.array_equal(got, want), f"n={n}: allocation-free velocities differ"
            # the induced part is not trivially the free stream
            assert np.abs(got[:, 3] - 1.0).max() > 1e-3
    finally:
        ctx.close()


def test_noalloc_refuses_input_beyond_scratch():
    ctx = V.Context(0)
    try:
        n = V.NOALLOC_SCRATCH_BYTES // 16 + 1024   # the near sums alone exceed the array
        print(f"noalloc size limit: n={n}", flush=True)
        objs = V.gen_wake(n, 0x5EED0103)
        ctx.set_params(re=600.0, dt=0.005)
        ctx.upload_bodies(np.zeros(0, V.SEG_DTYPE), np.zeros(0, V.BODY_DTYPE))
        ctx.upload(objs)
        ctx.tree_build(8)
        ctx.epsilon(True)
        with pytest.raises(V.VVHDError) as e:
            ctx.convective_noalloc((1.0, 0.0))
        assert e.value.code == V.VVHD_EINVAL
        assert "scratch" in str(e.value)
        # the context stays usable: the allocating entry point still runs
        ctx.convective((1.0, 0.0))
        assert np.isfinite(ctx.download()[:, 3:5]).all()
    finally:
        ctx.close()
