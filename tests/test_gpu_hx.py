"""H*X on the GPU (pisces_apply_h / pisces_apply_h_device) against the oracle's
restatement of DVR::MatrixTimesVector (src/mtx.cpp:22-108)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [((12, 11, 10), (22.0, 21.0, 20.0)), ((32, 32, 32), (36.0, 36.0, 36.0)),
          ((64, 64, 64), (40.0, 40.0, 40.0)), ((48, 40, 33), (44.0, 40.0, 36.0))]


def _case(oracle_mod, n, L, nvec, seed):
    g = oracle_mod.Grid(n, L)
    rng = np.random.default_rng(seed)
    V = rng.uniform(-1.0, 0.5, g.ngp)
    X = rng.uniform(-1.0, 1.0, (nvec, g.ngp))
    return g, V, X, g.apply_h(V, X)


@pytest.mark.parametrize("variant", [1, 0])
@pytest.mark.parametrize("n,L", SHAPES)
def test_apply_h_fp64(pisces, oracle_mod, n, L, variant):
    print("hx fp64", n, "variant", variant, flush=True)
    g, V, X, Yo = _case(oracle_mod, n, L, 3, 1)
    with pisces.Context(n, L) as ctx:
        ctx.set_hx_variant(variant)
        ctx.set_potential(V)
        Y = ctx.apply_h(X)
    # FP64 with a different summation order: a few ulps of the row sums
    assert np.abs(Y - Yo).max() <= 1e-13 * np.abs(Yo).max() * 10


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-6), ("bf16x2", 2e-4)])
@pytest.mark.parametrize("n,L", SHAPES)
def test_apply_h_low_precision(pisces, oracle_mod, n, L, precision, tol):
    print("hx", precision, n, flush=True)
    g, V, X, Yo = _case(oracle_mod, n, L, 4, 2)
    with pisces.Context(n, L) as ctx:
        ctx.set_potential(V)
        Y = ctx.apply_h_ex(X, pisces.PRECISIONS[precision])
    # reference for the rounded input: the FP32 input itself is exact in FP64
    Yr = g.apply_h(V, X.astype(np.float32).astype(np.float64))
    err = np.abs(Y - Yr).max() / np.abs(Yr).max()
    assert err <= tol, err


def test_apply_h_ex_fp64_is_apply_h(pisces, oracle_mod):
    n, L = (32, 32, 32), (36.0, 36.0, 36.0)
    g, V, X, Yo = _case(oracle_mod, n, L, 2, 3)
    with pisces.Context(n, L) as ctx:
        ctx.set_potential(V)
        np.testing.assert_array_equal(ctx.apply_h_ex(X, pisces.FP64), ctx.apply_h(X))


def test_apply_h_linearity_and_symmetry(pisces, oracle_mod):
    """Size-independent properties at the benchmark size: H(ax+by) = aHx + bHy
    and <u, Hv> = <Hu, v>."""
    n, L = (64, 64, 64), (40.0, 40.0, 40.0)
    rng = np.random.default_rng(9)
    ng = 64 ** 3
    V = rng.uniform(-1, 0.5, ng)
    u, v = rng.uniform(-1, 1, (2, ng))
    with pisces.Context(n, L) as ctx:
        ctx.set_potential(V)
        Hu, Hv, Hw = ctx.apply_h(np.stack([u, v, 0.3 * u - 1.7 * v]))
    np.testing.assert_allclose(Hw, 0.3 * Hu - 1.7 * Hv, rtol=0, atol=1e-12 * np.abs(Hw).max())
    assert abs(u @ Hv - Hu @ v) <= 1e-12 * abs(u @ Hv)


@pytest.mark.parametrize("n,L,nvec", [((64, 64, 64), (40.0, 40.0, 40.0), 8),
                                      ((48, 40, 33), (44.0, 40.0, 36.0), 5),
                                      ((32, 32, 32), (36.0, 36.0, 36.0), 13),
                                      ((12, 11, 10), (22.0, 21.0, 20.0), 1),
                                      ((96, 96, 96), (48.0, 48.0, 48.0), 3)])
def test_apply_h_bf16x2_one_launch_equals_two(pisces, oracle_mod, n, L, nvec):
    """The one-launch split-BF16 H*X (both passes ordered per XCD, pass B waiting
    on per-vector completion counts) does the same arithmetic per plane as the
    two launches: equal bit for bit, on every one of several repetitions."""
    print("hx fused", n, nvec, flush=True)
    g, V, X, Yo = _case(oracle_mod, n, L, nvec, 5)
    with pisces.Context(n, L) as ctx:
        ctx.set_potential(V)
        ctx.set_hx_variant(2)
        Y2 = ctx.apply_h_ex(X, pisces.BF16X2)
        ctx.set_hx_variant(3)
        for _ in range(4):
            Y1 = ctx.apply_h_ex(X, pisces.BF16X2)
            np.testing.assert_array_equal(Y1, Y2)
    assert np.abs(Y1 - Yo).max() <= 2e-4 * np.abs(Yo).max()
    assert pisces.hx_spin_timeouts() == 0
