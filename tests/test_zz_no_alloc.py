"""The allocation-free kinetic preconditioner (pisces_kinetic_precond_noalloc)
against pisces_kinetic_precond on the same inputs: the same kernels, so the
results must agree bit for bit, on consecutive calls that share the library's
scratch array (the second call smaller), and oversized requests are refused on
the host before anything is launched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [((40, 36, 33), (40.0, 38.0, 36.0)), ((20, 18, 16), (30.0, 28.0, 26.0))]


@pytest.mark.parametrize("precision", ["fp64", "bf16x2"])
def test_noalloc_equals_allocating_entry(pisces, precision):
    prec = pisces.PRECISIONS[precision]
    rng = np.random.default_rng(7)
    # two calls on one grid (3 then 2 vectors), then a smaller grid
    plan = [(CASES[0], 3, -0.05, -1.0), (CASES[0], 2, 0.3, 2.0), (CASES[1], 1, -0.02, -1.0)]
    ctxs = {}
    try:
        for (n, L), nvec, shift, scale in plan:
            print("noalloc", precision, n, nvec, flush=True)
            if n not in ctxs:
                ctxs[n] = pisces.Context(n, L)
            ctx = ctxs[n]
            r = rng.uniform(-1, 1, size=(nvec, ctx.ngp))
            ref = ctx.kinetic_precond(r, shift=shift, scale=scale, floor=1e-3, precision=prec)
            d_r = pisces.DeviceArray.from_host(r)
            d_t = pisces.DeviceArray((nvec, ctx.ngp))
            rc = ctx.kinetic_precond_noalloc(d_r, d_t, nvec, shift, scale, 1e-3, prec)
            assert rc == 0, pisces.lib().pisces_last_error()
            got = d_t.to_host()
            d_r.free()
            d_t.free()
            np.testing.assert_array_equal(got, ref)
    finally:
        for c in ctxs.values():
            c.close()


def test_noalloc_rejects_oversized_request(pisces):
    n, L = (64, 64, 64), (40.0, 40.0, 40.0)
    cap = pisces.noalloc_scratch_bytes()
    with pisces.Context(n, L) as ctx:
        nvec = cap // (2 * ctx.ngp)  # the BF16 intermediate alone fills the array
        print("noalloc oversized", nvec, flush=True)
        d_r = pisces.DeviceArray((nvec, ctx.ngp))
        d_t = pisces.DeviceArray((nvec, ctx.ngp))
        rc = ctx.kinetic_precond_noalloc(d_r, d_t, nvec, -0.05, -1.0, 1e-3, pisces.BF16X2)
        assert rc == -1
        assert b"scratch bytes" in pisces.lib().pisces_last_error()
        # the FP64 form needs only the descriptors: the same request fits
        r = np.random.default_rng(3).uniform(-1, 1, size=(2, ctx.ngp))
        d_r2 = pisces.DeviceArray.from_host(r)
        rc = ctx.kinetic_precond_noalloc(d_r2, d_t, 2, -0.05, -1.0, 1e-3, pisces.FP64)
        assert rc == 0, pisces.lib().pisces_last_error()
        got = d_t.to_host()[:2]
        ref = ctx.kinetic_precond(r, shift=-0.05, scale=-1.0, floor=1e-3, precision=pisces.FP64)
        np.testing.assert_array_equal(got, ref)
        for d in (d_r, d_t, d_r2):
            d.free()
