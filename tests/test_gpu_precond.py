"""Kinetic preconditioner t = scale (Tx (+) Ty (+) Tz - shift)^{-1} r on the
matrix cores (FP64 and split-BF16 forms), against a dense numpy evaluation in
the eigenbases of the oracle's kinetic matrices and against the defining
identity (T - shift) t = scale r (the oracle's H*x with V = 0)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [((12, 11, 10), (22.0, 21.0, 20.0)), ((20, 18, 16), (30.0, 28.0, 26.0)),
          ((40, 36, 33), (40.0, 38.0, 36.0)), ((70, 9, 12), (60.0, 20.0, 22.0))]


def _reference(g, r, shift, scale, floor):
    lam, Q = zip(*(np.linalg.eigh(T) for T in g.T))
    nx, ny, nz = g.n
    R = r.reshape(nz, ny, nx)
    Rh = np.einsum("kji,ia,jb,kc->cba", R, Q[0], Q[1], Q[2], optimize=True)
    den = lam[2][:, None, None] + lam[1][None, :, None] + lam[0][None, None, :] - shift
    den = np.where(np.abs(den) < floor, np.where(den < 0, -floor, floor), den)
    Th = scale * Rh / den
    return np.einsum("cba,ia,jb,kc->kji", Th, Q[0], Q[1], Q[2], optimize=True).reshape(-1)


@pytest.mark.parametrize("n,L", SHAPES)
# bf16x2: the two intermediates between the plane passes are stored in BF16
# (rounding 2^-9); carried through the transforms in numpy that alone gives
# 1.8e-3 to 3.0e-3 relative on these shapes
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("bf16x2", 5e-3)])
def test_kinetic_precond(pisces, oracle_mod, n, L, precision, tol):
    print("precond", n, precision, flush=True)
    g = oracle_mod.Grid(n, L)
    rng = np.random.default_rng(sum(n))
    r = rng.uniform(-1, 1, size=(2, g.ngp))
    with pisces.Context(n, L) as ctx:
        t_lo = ctx.kinetic_precond(r, shift=-0.05, scale=-1.0, floor=1e-3,
                                   precision=pisces.PRECISIONS[precision])
        t_in = ctx.kinetic_precond(r, shift=0.3, scale=2.0, floor=1e-3,
                                   precision=pisces.PRECISIONS[precision])
    for v in range(2):
        ref = _reference(g, r[v], -0.05, -1.0, 1e-3)
        print("relative error", np.linalg.norm(t_lo[v] - ref) / np.linalg.norm(ref), flush=True)
        assert np.linalg.norm(t_lo[v] - ref) <= tol * np.linalg.norm(ref)
        ref = _reference(g, r[v], 0.3, 2.0, 1e-3)
        assert np.linalg.norm(t_in[v] - ref) <= tol * np.linalg.norm(ref)
    if precision == "fp64":
        # (T - shift) t = scale r, shift below the spectrum (no flooring)
        lhs = g.apply_h(np.zeros(g.ngp), t_lo) + 0.05 * t_lo
        np.testing.assert_allclose(lhs, -r, rtol=0, atol=1e-10 * np.abs(r).max())


@pytest.mark.parametrize("n,L,nvec", [((64, 64, 64), (40.0, 40.0, 40.0), 9),
                                      ((40, 36, 33), (40.0, 38.0, 36.0), 4),
                                      ((20, 18, 16), (30.0, 28.0, 26.0), 1)])
def test_precond_bf16x2_one_launch_equals_three(pisces, n, L, nvec):
    """The split-BF16 preconditioner with its three passes in one XCD-ordered
    launch (pass 1 waiting on pass 0's planes of the vector, pass 2 on pass
    1's) does the same arithmetic per plane as three launches: equal bit for
    bit, on every one of several repetitions."""
    print("precond fused", n, nvec, flush=True)
    r = np.random.default_rng(sum(n) + nvec).uniform(-1, 1, size=(nvec, int(np.prod(n))))
    with pisces.Context(n, L) as ctx:
        ctx.set_hx_variant(2)
        t3 = ctx.kinetic_precond(r, shift=-0.05, scale=-1.0, floor=1e-3, precision=pisces.BF16X2)
        ctx.set_hx_variant(3)
        for _ in range(4):
            t1 = ctx.kinetic_precond(r, shift=-0.05, scale=-1.0, floor=1e-3,
                                     precision=pisces.BF16X2)
            np.testing.assert_array_equal(t1, t3)
    assert pisces.hx_spin_timeouts() == 0
