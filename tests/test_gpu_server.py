"""Continuous-batching server: results are the same alone and in the mix,
admission respects the HBM cap, warm starts along an MD-like chain."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [((32, 32, 32), 30.0, 6, 3.5, 4100, 2), ((48, 48, 48), 36.0, 12, 4.5, 4101, 4),
         ((40, 36, 32), 34.0, 8, 4.0, 4102, 3), ((32, 32, 32), 30.0, 5, 3.5, 4103, 1)]


def _xyz(nw, rad, seed):
    import workload
    return workload.water_cluster(nw, rad, seed)[0]


@pytest.mark.parametrize("precision", ["fp64", "bf16x2"])
def test_same_alone_and_in_the_mix(pisces, precision):
    prec = pisces.PRECISIONS[precision]
    alone = []
    for n, L, nw, rad, seed, nr in SIZES:
        sv = pisces.Server(hbm_cap_bytes=64 << 30, max_slots=1, precision=prec)
        sv.submit(n, (L, L, L), _xyz(nw, rad, seed), nr)
        res = sv.run_until_done()
        sv.close()
        assert len(res) == 1 and res[0]["status"] == nr, res
        alone.append(res[0])
    sv = pisces.Server(hbm_cap_bytes=64 << 30, max_slots=8, precision=prec)
    ids = [sv.submit(n, (L, L, L), _xyz(nw, rad, seed), nr) for n, L, nw, rad, seed, nr in SIZES]
    mix = {r["id"]: r for r in sv.run_until_done()}
    sv.close()
    for i, rid in enumerate(ids):
        r = mix[rid]
        assert r["status"] == SIZES[i][5]
        # same eigenpairs within the stated tolerance (1e-6 Eh per eigenvalue)
        np.testing.assert_allclose(r["evals"], alone[i]["evals"], rtol=0, atol=1e-6)
        assert np.all(r["resid"] < 1e-7)


def test_matches_single_instance_solver(pisces):
    n, L, nw, rad, seed, nr = SIZES[1]
    xyz = _xyz(nw, rad, seed)
    with pisces.Context(n, (L, L, L)) as ctx:
        ctx.build_cluster_potential(xyz)
        r0 = ctx.solve_block(nr, pisces.start_vectors(ctx, nr), block=8, maxsub=32, ptol=8,
                             precision=pisces.FP64)
    sv = pisces.Server(hbm_cap_bytes=64 << 30, max_slots=4, precision=pisces.FP32)
    sv.submit(n, (L, L, L), xyz, nr, want_vectors=True)
    (r,) = sv.run_until_done()
    np.testing.assert_allclose(r["evals"], r0["evals"], rtol=0, atol=1e-6)


def test_admission_respects_hbm_cap(pisces):
    n, L, nw, rad, seed, nr = SIZES[0]
    sv = pisces.Server(hbm_cap_bytes=1 << 40, max_slots=16, precision=pisces.BF16X2)
    fp = sv.footprint(n, (L, L, L), _xyz(nw, rad, seed), nr)
    sv.close()
    cap = int(2.5 * fp)  # room for two instances at a time
    sv = pisces.Server(hbm_cap_bytes=cap, max_slots=16, precision=pisces.BF16X2)
    for i in range(5):
        sv.submit(n, (L, L, L), _xyz(nw, rad, seed + i), nr)
    peak, results = 0, []
    for _ in range(10000):
        live = sv.step()
        peak = max(peak, live)
        assert sv.used_bytes() <= cap
        results += sv.poll()
        if live == 0 and sv.pending() == 0:
            break
    sv.close()
    assert len(results) == 5 and all(r["status"] == nr for r in results)
    assert peak == 2


def test_warm_start_chain(pisces):
    import workload
    rng = np.random.default_rng(5)
    n, L, nw, rad, seed, nr = SIZES[0]
    xyz = _xyz(nw, rad, seed)
    chain = [xyz]
    for _ in range(3):
        chain.append(workload.perturb(chain[-1], rng))
    sv = pisces.Server(hbm_cap_bytes=64 << 30, max_slots=4, precision=pisces.BF16X2)
    cold, warm = [], []
    for i, x in enumerate(chain):
        sv.submit(n, (L, L, L), x, nr, chain=7, warm=(i > 0))
        warm += sv.run_until_done()
    for x in chain:
        sv.submit(n, (L, L, L), x, nr)
        cold += sv.run_until_done()
    sv.close()
    for c, w in zip(cold, warm):
        np.testing.assert_allclose(w["evals"], c["evals"], rtol=0, atol=1e-6)
    # warm starts from the previous snapshot need fewer iterations
    assert sum(w["n_iter"] for w in warm[1:]) < sum(c["n_iter"] for c in cold[1:])


def test_warm_chains_cross_lanes(pisces, monkeypatch):
    """Two chains submitted together in alternating order on a two-lane server:
    the lanes take the requests in turn, so each chain's next snapshot starts
    on the other lane from vectors the first lane wrote.  Warm results match
    cold solves."""
    import workload
    monkeypatch.setenv("PISCES_SERVER_LANES", "2")
    rng = np.random.default_rng(11)
    n, L, nw, rad, seed, nr = SIZES[0]
    chains = [[_xyz(nw, rad, seed)], [_xyz(nw, rad, seed + 50)]]
    for ch in chains:
        for _ in range(3):
            ch.append(workload.perturb(ch[-1], rng))
    sv = pisces.Server(hbm_cap_bytes=64 << 30, max_slots=4, precision=pisces.BF16X2)
    warm = {0: [], 1: []}
    for i in range(4):
        order = (0, 1) if i % 2 == 0 else (1, 0)
        ids = {sv.submit(n, (L, L, L), chains[c][i], nr, chain=c, warm=(i > 0)): c for c in order}
        for r in sv.run_until_done():
            warm[ids[r["id"]]].append(r)
    cold = {0: [], 1: []}
    for c in (0, 1):
        for x in chains[c]:
            sv.submit(n, (L, L, L), x, nr)
            cold[c] += sv.run_until_done()
    sv.close()
    for c in (0, 1):
        assert len(warm[c]) == 4
        for w, r in zip(warm[c], cold[c]):
            assert w["status"] == nr
            np.testing.assert_allclose(w["evals"], r["evals"], rtol=0, atol=1e-6)
        assert sum(w["n_iter"] for w in warm[c][1:]) < sum(r["n_iter"] for r in cold[c][1:])


def test_block_classes_and_lanes_in_one_mix(pisces):
    """Requests of 2, 6 and 4 states together: the 6-state one runs on a block-8
    engine, the others on block-4 engines, spread over the lanes; every
    result equals the same request solved alone."""
    cases = [((32, 32, 32), 30.0, 6, 3.5, 4200, 2), ((40, 40, 40), 34.0, 10, 4.0, 4201, 6),
             ((32, 32, 32), 30.0, 8, 3.5, 4202, 4), ((36, 36, 36), 32.0, 7, 3.5, 4203, 6)]
    alone = []
    for n, L, nw, rad, seed, nr in cases:
        sv = pisces.Server(hbm_cap_bytes=64 << 30, max_slots=1, precision=pisces.BF16X2)
        sv.submit(n, (L, L, L), _xyz(nw, rad, seed), nr)
        (r,) = sv.run_until_done()
        sv.close()
        assert r["status"] == nr, r
        alone.append(r)
    sv = pisces.Server(hbm_cap_bytes=64 << 30, max_slots=8, precision=pisces.BF16X2)
    ids = [sv.submit(n, (L, L, L), _xyz(nw, rad, seed), nr) for n, L, nw, rad, seed, nr in cases]
    mix = {r["id"]: r for r in sv.run_until_done()}
    sv.close()
    for i, rid in enumerate(ids):
        r = mix[rid]
        assert r["status"] == cases[i][5] and len(r["evals"]) == cases[i][5]
        np.testing.assert_allclose(r["evals"], alone[i]["evals"], rtol=0, atol=1e-6)
        assert np.all(r["resid"] < 1e-7)
