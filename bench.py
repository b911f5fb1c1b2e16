"""Benchmark of the PISCES excess-electron eigensolve on one MI355X.

Workload: (H2O)24 anions, 24 rigid waters placed from a seed in a 6 A sphere
(one seed per snapshot), DPP electron-water potential with polarization, a
sine-DVR 64^3 grid on a 40 bohr cube, lowest 4 states (block 4, subspace 36
for the mixed-precision form and 28 for the FP64 forms, each its measured best;
residual tolerance 1e-7).  A step is one batch of B such snapshots submitted
together to the continuous-batching server and solved to completion; the
potential build (rigid-water projection, induced dipoles, grid potential) is
part of every solve.

Reported next to each other on the same snapshots:
  * the served mixed-precision solver (H*X in FP32 or split BF16 on the matrix
    cores, FP64 Ritz/orthogonalisation/residual refresh): `value`, solves/s;
  * the plain FP64 GPU form (the same engine, FP64 throughout, one solve at a
    time) and the FP64 form batched through the server;
  * the eigenvalue error of every state of every snapshot against the FP64
    GPU result (which matches the CPU oracle within 1e-8 Eh in the tests);
  * HIP-event timing of every batched launch family over one more step of
    the batch on a single engine lane (the headline's three lanes overlap their
    launches), the roofline of the dominant one;
  * the CPU restatement of the reference path (its Davidson, dense-T H*x, the
    DPP potential) timed on the host cores on a bounded sample.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pisces-1_amd"))

N_GRID = 64
L_BOX = 40.0
N_WATER = 24
RADIUS = 6.0
SEED = 2000
NROOTS = 4
BLOCK = 4        # block engine: block and subspace of the served solver's 4-state class
MAXSUB = 36       # served mixed-precision form (36: 564 solves/s; 28: 546, 44: 546)
MAXSUB_FP64 = 28  # FP64 comparators (28: 267 batched solves/s; 36: 253)
MAXSUB_MIX = 28   # the 32^3-96^3 mixes (512-snapshot batch: 996-1000 solves/s; 36: 960-980)
FP64_MATRIX_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix
REF_MAXSUB = 32  # the reference algorithm's subspace (cpu_baseline's matvec count)
PTOL = 7
DE_TOL = 1e-6
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
METRIC = ("solves/s + p50/p99 latency, 64³ grid 4 states, ΔE≤1e-6 Eh: "
          "mixed vs FP64 GPU")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def hx_algorithmic(n, nvec, elem):
    """One H*X application: bytes = X read + Y written once per vector, V once;
    flops = (6n + 4) per point per vector (three dense 1-D contractions of
    length n, the potential multiply and the sums)."""
    ngp = n ** 3
    return (2 * nvec + 1) * ngp * elem, (6 * n + 4) * ngp * nvec


def run_single(P, geoms, precision, maxsub=MAXSUB_FP64):
    """One solve at a time on one context (single-instance block engine)."""
    n, L = (N_GRID,) * 3, (L_BOX,) * 3
    ctx = P.Context(n, L)
    out = []
    lat = []
    for i, xyz in enumerate(geoms):
        t0 = time.perf_counter()
        ctx.build_cluster_potential(xyz)
        r = ctx.solve_block(NROOTS, P.start_vectors(ctx, NROOTS, seed=i), block=BLOCK,
                            maxsub=maxsub, ptol=PTOL, precision=precision)
        lat.append(time.perf_counter() - t0)
        out.append(r)
    ctx.close()
    return out, lat


def run_server(P, geoms, precision, steps, warmup, prof=False, lanes=None, maxsub=None):
    """`lanes`: the server's engine lanes (None: its default, three); `maxsub`:
    the subspace (None: MAXSUB for the mixed forms, MAXSUB_FP64 for FP64)."""
    if maxsub is None:
        maxsub = MAXSUB_FP64 if precision == P.FP64 else MAXSUB
    if lanes is None:
        os.environ.pop("PISCES_SERVER_LANES", None)
    else:
        os.environ["PISCES_SERVER_LANES"] = str(lanes)
    sv = P.Server(hbm_cap_bytes=200 << 30, max_slots=len(geoms), precision=precision,
                  maxsub=maxsub, ptol=PTOL)
    n, L = (N_GRID,) * 3, (L_BOX,) * 3

    def one_step():
        ids = [sv.submit(n, L, xyz, NROOTS) for xyz in geoms]
        res = {r["id"]: r for r in sv.run_until_done()}
        return [res[i] for i in ids]

    for _ in range(warmup):
        one_step()
    if prof:
        P.prof_enable(True)
        P.prof_reset()
    results, t_steps = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        results.append(one_step())
        t_steps.append(time.perf_counter() - s0)
    dt = time.perf_counter() - t0
    fam = None
    if prof:
        fam = P.prof_read()
        P.prof_enable(False)
    sv.close()
    return results, dt, t_steps, fam


def roofline_pass(P, geoms, precision):
    """One step of the served batch on one engine lane with HIP-event timing of
    every launch family.  The headline runs three lanes whose launches overlap,
    so there an event pair around a launch also spans the other lane's work;
    on one lane each family's time is its own."""
    res, dt, _, fam = run_server(P, geoms, precision, 1, 0, prof=True, lanes=1)
    return fam, dt


def families(fam):
    out = {}
    for k, (ms, nl, by) in fam.items():
        if nl:
            out[k] = {"ms_total": round(ms, 3), "launches": nl,
                      "ms_per_launch": round(ms / nl, 5),
                      "GBs_algorithmic": round(by / (ms * 1e-3) / 1e9, 1) if by else None}
    return out


# matrix-core peaks of MI355X (dense): FP64 78.6, FP32 157.3, BF16 2500 TFLOP/s
PEAK_TFS = {"fp64": 78.6, "fp32": 157.3, "bf16x2": 2500.0}


def hx_roofline(hx, fam):
    """The fused H*X against both roofline terms: the standalone 64^3 x 8
    application per precision (bytes (2 + 1/k) N s, flops (6n + 4) N k, the
    split-BF16 form executing three BF16 products per contraction) as
    max(bytes / 8 TB/s, flops / peak) / measured time, and the served batch's
    split-BF16 launches (HBM-bound: bytes only)."""
    out = {}
    for name, e in (("fp64", 8), ("fp32", 4), ("bf16x2", 4)):
        by, fl = hx_algorithmic(N_GRID, 8, e)
        fl_exec = fl * (3 if name == "bf16x2" else 1)
        t = hx[name]["ms"] * 1e-3
        t_bytes, t_flops = by / (HBM_PEAK_GBS * 1e9), fl_exec / (PEAK_TFS[name] * 1e12)
        out[name + "_64cube_x8"] = {
            "bytes": by, "flops": fl, "flops_executed": fl_exec, "ms": hx[name]["ms"],
            "t_bytes_us": round(t_bytes * 1e6, 3), "t_flops_us": round(t_flops * 1e6, 3),
            "bound": "hbm" if t_bytes >= t_flops else "mfma",
            "frac": round(max(t_bytes, t_flops) / t, 4),
            "hbm_frac": round(t_bytes / t, 4)}
    if fam and "hx_lo" in fam and fam["hx_lo"][1]:
        ms, nl, by = fam["hx_lo"]
        gbs = by / (ms * 1e-3) / 1e9
        out["served_bf16x2"] = {"launches": nl, "ms_per_launch": round(ms / nl, 5),
                                "bytes_per_launch": round(by / nl), "GBs_algorithmic": round(gbs, 1),
                                "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def potential_roofline(fam, batch):
    """The grid build of the potential (FP64 VALU-bound: rsqrt and exp per site
    term) in the served batch: executed FP64 flops per launch from the
    per-point count that the SQ_INSTS_VALU_*_F64 counters give for this
    workload (profiles/potential_pmc.json), over the launch time, against the
    FP64 vector peak."""
    if not fam or "potential" not in fam or not fam["potential"][1]:
        return None
    ms, nl, _ = fam["potential"]
    path = os.path.join(ROOT, "profiles", "potential_pmc.json")
    try:
        with open(path) as f:
            fpp = json.load(f)["flops_per_point"]
    except (OSError, KeyError, ValueError):
        return {"ms_per_launch": round(ms / nl, 4), "TFLOPs": None}
    fl = fpp * N_GRID ** 3 * batch / nl
    tf = fl / (ms / nl * 1e-3) / 1e12
    return {"ms_per_launch": round(ms / nl, 4), "flops_per_point": fpp,
            "flops_per_launch": fl, "TFLOPs": round(tf, 2), "peak_TFLOPs": FP64_PEAK_TFS,
            "frac": round(tf / FP64_PEAK_TFS, 4), "bound": "fp64 valu"}


def roofline_of(fam, dt):
    """Roofline of the launch family with the most time in the one-lane pass."""
    fo = families(fam)
    dom = max(fo, key=lambda k: fo[k]["ms_total"])
    ms_d, nl_d, by_d = fam[dom]
    achieved = by_d / (ms_d * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
            "ms_per_launch": round(ms_d / nl_d, 5),
            "bytes_per_launch": round(by_d / nl_d),
            "share_of_timed": round(ms_d * 1e-3 / dt, 3),
            "measured_on": "one step of the same batch on one engine lane (no overlap)"}
    roof["traffic"], roof["traffic_over_algorithmic"] = measured_traffic(dom, by_d / nl_d)
    return roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--precision", default="bf16x2", choices=["fp32", "bf16x2"])
    ap.add_argument("--fp64-single", type=int, default=8,
                    help="snapshots solved one at a time in FP64 (the plain form)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--served-only", action="store_true",
                    help="time only the served mixed-precision steps (profiling runs)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the one-lane roofline pass after a warm-up step (profiling runs)")
    ap.add_argument("--gradient-only", action="store_true",
                    help="only the electron side of the analytic gradient: (H2O)24 on 64^3 with a "
                         "PolType 3 potential, time per pisces_gradient_terms call")
    ap.add_argument("--pol3-only", action="store_true",
                    help="only the PolType 3 potential builds (profiling runs)")
    ap.add_argument("--config", default="c2", choices=["c2", "c3", "c4", "c5"],
                    help="c2: the headline workload; c4: 512 uneven snapshots in one batch; "
                         "c5: open-loop Poisson-served MD chains")
    ap.add_argument("--duration", type=float, default=8.0, help="c5: seconds per offered rate")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the Gram kernel configuration on the served workload and "
                         "write pisces-1_amd/tuning.txt")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only a plain device-to-device copy of BYTES (the calibration of the "
                         "FETCH_SIZE / WRITE_SIZE byte counters: one launch moves BYTES in and "
                         "BYTES out), then read-only streams of BYTES with the default and the "
                         "nontemporal load policy")
    ap.add_argument("--pmc-summary", nargs=5, default=None,
                    metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV",
                             "ROOF_JSON"),
                    help="only print the HBM traffic summary per launch family (the "
                         "profiles/traffic.json format) from rocprofv3 --pmc FETCH_SIZE / "
                         "WRITE_SIZE counter CSVs of a --roofline-only run and of a "
                         "--copy-probe run, with the algorithmic bytes of the --roofline-only "
                         "JSON line; no GPU")
    a = ap.parse_args()
    if a.pmc_summary:
        print(json.dumps(pmc_summary(*a.pmc_summary), indent=1))
        return
    if a.gpus != 1:
        print("bench.py: one GPU only; the library has no multi-GPU path (--gpus %d)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)

    if a.tune:
        print(json.dumps(tune_gram(a)), flush=True)
        return

    import pisces_amd as P
    import workload

    if a.copy_probe:
        print(json.dumps(copy_probe(P, a.copy_probe, max(1, a.steps))), flush=True)
        return
    prec = P.PRECISIONS[a.precision]
    if a.pol3_only:
        print(json.dumps({"pol3_potential": pol3_measure(P, workload)}), flush=True)
        return
    if a.gradient_only:
        print(json.dumps({"gradient": gradient_measure(P, workload, max(1, a.steps))}), flush=True)
        return
    if a.config == "c3":
        print(json.dumps(bench_c3(P, workload, prec)), flush=True)
        return
    if a.config == "c4":
        print(json.dumps(bench_c4(P, workload, prec, a)), flush=True)
        return
    if a.config == "c5":
        print(json.dumps(bench_c5(P, workload, prec, a)), flush=True)
        return
    B = a.batch
    geoms = [workload.water_cluster(N_WATER, RADIUS, SEED + i)[0] for i in range(B)]
    log("batch of %d snapshots generated" % B)

    nsingle = max(1, min(a.fp64_single, B))
    if a.roofline_only:
        if a.warmup:
            run_server(P, geoms, prec, a.warmup, 0, lanes=1)
        fam, dt1 = roofline_pass(P, geoms, prec)
        print(json.dumps({"roofline": roofline_of(fam, dt1), "launch_families": families(fam),
                          "solves_per_s": round(B / dt1, 3)}), flush=True)
        return
    if not a.served_only:
        # ---- plain FP64 GPU form, one solve at a time
        run_single(P, geoms[:1], P.FP64)  # warm-up
        r64, lat64 = run_single(P, geoms[:nsingle], P.FP64)
        assert all(r["nconv"] == NROOTS for r in r64), [r["nconv"] for r in r64]
        fp64_single_sps = nsingle / sum(lat64)
        log("fp64 single: %.2f solves/s" % fp64_single_sps)

        # ---- FP64 through the server (same batching as the mixed form)
        # three timed batches after a warm-up one: a single batch varied by
        # 10-15% between runs
        res64b, dt64b, _, _ = run_server(P, geoms, P.FP64, 3, 1)
        fp64_batched_sps = 3 * B / dt64b
        ev64 = np.array([r["evals"] for r in res64b[-1]])
        assert all(r["status"] == NROOTS for step in res64b for r in step), \
            [(r["status"], r["n_iter"]) for r in res64b[-1]]
        log("fp64 batched: %.2f solves/s" % fp64_batched_sps)

    # ---- mixed precision through the server: the headline, timed
    res, dt, t_steps, _ = run_server(P, geoms, prec, a.steps, a.warmup)
    sps = B * a.steps / dt
    lat = np.array([r["latency_s"] for step in res for r in step])
    n_ok = sum(r["status"] == NROOTS for step in res for r in step)
    evm = np.array([r["evals"] for r in res[-1]])
    if a.served_only:
        de = de_single = np.zeros(1)
        fp64_single_sps = fp64_batched_sps = float("nan")
        lat64 = [float("nan")]
    else:
        de = np.abs(evm - ev64)
        de_single = np.abs(evm[:nsingle] - np.array([r["evals"] for r in r64]))
    log("%s batched: %.2f solves/s, max |dE| %.2e" % (a.precision, sps, de.max()))

    # ---- launch families and the roofline of the dominant one, on one lane
    fam, dt1 = roofline_pass(P, geoms, prec)
    fam_out = families(fam)
    roof = roofline_of(fam, dt1)

    # ---- standalone H*X at 64^3 x 8 vectors (entries uniform in [-1, 1])
    hx = {}
    with P.Context((N_GRID,) * 3, (L_BOX,) * 3) as ctx:
        ctx.build_cluster_potential(geoms[0])
        for name, pr, e in (("fp64", P.FP64, 8), ("fp32", P.FP32, 4), ("bf16x2", P.BF16X2, 4)):
            ms = ctx.time_hx(8, pr, warmup=10, iters=100)
            by, fl = hx_algorithmic(N_GRID, 8, e)
            hx[name] = {"ms": round(ms, 5), "GBs_algorithmic": round(by / (ms * 1e-3) / 1e9, 1),
                        "TFLOPs": round(fl / (ms * 1e-3) / 1e12, 2)}
    hx_roof = hx_roofline(hx, fam)

    # ---- potential with self-consistent polarizable sites (PolType 3)
    pol3 = None if a.served_only else pol3_measure(P, workload)

    cpu = None
    if not a.no_cpu and not a.served_only:
        cpu = cpu_baseline(P, geoms[0], a.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": round(sps, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1000.0 * dt / a.steps, 3),
        "p50_ms": round(1000 * float(np.percentile(lat, 50)), 3),
        "p99_ms": round(1000 * float(np.percentile(lat, 99)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32+f64", "bf16x2": "bf16x2+f64"}[a.precision],
        "data": "synthetic (seeded water clusters)",
        "config": {"workload": "(H2O)24 anion snapshots, sine-DVR 64^3 on 40 bohr, 4 states, "
                               "block %d, subspace %d, residual tol 1e-%d, batch %d per step, "
                               "H*X %s" % (BLOCK, MAXSUB, PTOL, B, a.precision),
                   "batch": B, "parallelism": "single GPU",
                   "note": "the C2 configuration states block 8 and subspace 32; the served "
                           "4-state class runs block 4 and each form its measured best "
                           "subspace: 36 mixed, 28 FP64 (fewer bytes per step)"},
        "fp64_gpu": {"single_solves_per_s": round(fp64_single_sps, 3),
                     "single_p50_ms": round(1000 * float(np.percentile(lat64, 50)), 3),
                     "batched_solves_per_s": round(fp64_batched_sps, 3)},
        "speedup_vs_fp64_single": round(sps / fp64_single_sps, 2),
        "speedup_vs_fp64_batched": round(sps / fp64_batched_sps, 2),
        "max_abs_dE_vs_fp64": float("%.3e" % max(de.max(), de_single.max())),
        # per state: the largest |E_mixed - E_fp64| over the batch's snapshots
        "max_abs_dE_vs_fp64_per_state": [float("%.3e" % v) for v in
                                         np.max(np.atleast_2d(de), axis=0)],
        "dE_ok": bool(max(de.max(), de_single.max()) <= DE_TOL and n_ok == B * a.steps),
        "evals_snapshot0": [round(float(x), 10) for x in evm[0]],
        "iters_mean": round(float(np.mean([r["n_iter"] for r in res[-1]])), 1),
        "roofline": roof,
        "launch_families": fam_out,
        "lanes": 3,
        "autotuned": P.tuning(),
        "hx_standalone_64cube_x8": hx,
        "hx_roofline": hx_roof,
        "potential_build": potential_roofline(fam, B),
        "pol3_potential": pol3,
        "cpu_baseline": cpu,
    }
    if a.served_only:
        for key in ("fp64_gpu", "speedup_vs_fp64_single", "speedup_vs_fp64_batched",
                    "max_abs_dE_vs_fp64", "max_abs_dE_vs_fp64_per_state", "dE_ok"):
            out[key] = None
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------- batched / served configs
SEED_C3 = 3000
SEED_C4 = 4000


def bench_c3(P, workload, prec, nroots=8, block=8, maxsub=40):
    """(H2O)45 + Na+ on 96^3 (48 bohr), 8 states, in the mixed form: a cold
    solve from the particle-in-a-box start against the dual-grid start of
    ClusterAnion (src/ClusterAnion.cpp:207-299) -- the same cluster solved on
    the coarse grid (every third point, 32^3), its states carried to the fine
    grid by ExtendWfThree's weights -- with the fine solve's matvecs, steps and
    time for both, the coarse grid's cost, and the largest eigenvalue
    difference between the two fine results."""
    xyz, ion = workload.water_cluster(45, 8.0, SEED_C3, ion_min=2.3)
    ions = (np.asarray(ion) * workload.ANGS2BOHR, [1.0])
    n, L = 96, 48.0
    (nc, _, _), (Lc, _, _) = P.dual_coarse((n,) * 3, (L,) * 3)

    def solve(ctx, start):
        t0 = time.perf_counter()
        r = ctx.solve_block(nroots, start, block=block, maxsub=maxsub, max_iter=2000,
                            ptol=PTOL, precision=prec)
        return r, time.perf_counter() - t0

    with P.Context((nc,) * 3, (Lc,) * 3) as cc:
        t0 = time.perf_counter()
        cc.build_cluster_potential(xyz, ions=ions, ion_r=2.9)
        t_build = time.perf_counter() - t0
        solve(cc, P.start_vectors(cc, nroots, seed=3))  # warm-up of the coarse shapes
        rc_, tc = solve(cc, P.start_vectors(cc, nroots, seed=3))
        t_coarse = t_build + tc
    assert rc_["nconv"] == nroots, rc_["nconv"]
    with P.Context((n,) * 3, (L,) * 3) as ctx:
        ctx.build_cluster_potential(xyz, ions=ions, ion_r=2.9)
        solve(ctx, P.start_vectors(ctx, nroots, seed=3))  # warm-up
        cold, t_cold = solve(ctx, P.start_vectors(ctx, nroots, seed=3))
        t0 = time.perf_counter()
        start = np.stack([ctx.dual_extend((nc,) * 3, v, 0) for v in rc_["vecs"]])
        t_ext = time.perf_counter() - t0
        warm, t_warm = solve(ctx, start)
    assert cold["nconv"] == nroots and warm["nconv"] == nroots, (cold["nconv"], warm["nconv"])
    de = float(np.abs(warm["evals"] - cold["evals"]).max())
    mv = lambda r: r["n_matvec"] + r["n_matvec_lo"]
    log("c3: cold %d steps %d matvecs %.3f s; dual-grid start %d steps %d matvecs %.3f s "
        "(+ coarse %.3f s, transfer %.3f s); max |dE| %.2e"
        % (cold["n_iter"], mv(cold), t_cold, warm["n_iter"], mv(warm), t_warm, t_coarse, t_ext, de))
    return {"config": {"workload": "(H2O)45 + Na+ anion, sine-DVR 96^3 on 48 bohr, 8 states, "
                                   "block %d, subspace %d, residual tol 1e-%d" % (block, maxsub, PTOL),
                       "coarse_grid": "%d^3 on %.3f bohr" % (nc, Lc)},
            "dtype": "bf16x2+f64" if prec == P.BF16X2 else "f32+f64",
            "cold": {"steps": cold["n_iter"], "matvecs_fp64": cold["n_matvec"],
                     "matvecs_low": cold["n_matvec_lo"], "seconds": round(t_cold, 4)},
            "dual_grid": {"steps": warm["n_iter"], "matvecs_fp64": warm["n_matvec"],
                          "matvecs_low": warm["n_matvec_lo"], "seconds": round(t_warm, 4),
                          "transfer_s": round(t_ext, 4),
                          "coarse_build_and_solve_s": round(t_coarse, 4),
                          "coarse_steps": rc_["n_iter"]},
            "note": "seconds of the fine solve alone (potential built beforehand); the dual-grid "
                    "total adds the coarse build and solve and the host-buffer transfer",
            "evals_cold": [round(float(e), 10) for e in cold["evals"]],
            "max_abs_dE_dual_vs_cold": de, "dE_ok": de <= 1e-6}
SEED_C5 = 5000


def bench_c4(P, workload, prec, a, count=512):
    """512 independent snapshots of uneven size submitted together; solves/s of
    the whole batch and per-request latency."""
    rng = np.random.default_rng(SEED_C4)
    snaps = [workload.mixed_snapshot(i, rng, SEED_C4) for i in range(count)]
    log("c4: %d snapshots, grids %s" % (count, {n: sum(s[1] == n for s in snaps) for n, _ in workload.MIX_GRIDS}))
    sv = P.Server(hbm_cap_bytes=200 << 30, max_slots=count, precision=prec, maxsub=MAXSUB_MIX,
                  ptol=PTOL)
    for xyz, n, L, nr in snaps[:32]:  # warm-up: shapes, pools, kernels
        sv.submit((n,) * 3, (L,) * 3, xyz, nr)
    sv.run_until_done()
    t0 = time.perf_counter()
    ids = [sv.submit((n,) * 3, (L,) * 3, xyz, nr) for xyz, n, L, nr in snaps]
    res = sv.run_until_done()
    dt = time.perf_counter() - t0
    sv.close()
    ok = [r for r in res if r["status"] == r["evals"].size]
    lat = np.array([r["latency_s"] for r in res])
    # eigenvalue error against the plain FP64 form on a seeded subset of 32
    byid = {r["id"]: r for r in res}
    sub = np.random.default_rng(SEED_C4 + 1).choice(count, 32, replace=False)
    ev64 = fp64_reference(P, [snaps[i] for i in sub])
    de = max(float(np.abs(byid[ids[i]]["evals"] - e).max()) for i, e in zip(sub, ev64))
    log("c4: %.1f solves/s, max |dE| vs FP64 on 32 snapshots %.2e" % (count / dt, de))
    return {"metric": "solves/s, 512 uneven snapshots in one batch (4-32 waters, 32^3/48^3/64^3, "
                      "1-4 states)",
            "max_abs_dE_vs_fp64_subset32": float("%.3e" % de),
            "dE_ok": bool(de <= DE_TOL and len(ok) == count),
            "value": round(count / dt, 3), "unit": "solves/s", "n_gpus": 1,
            "higher_is_better": True, "seconds": round(dt, 3),
            "p50_ms": round(1000 * float(np.percentile(lat, 50)), 2),
            "p99_ms": round(1000 * float(np.percentile(lat, 99)), 2),
            "converged": len(ok), "requests": len(res),
            "iters_mean": round(float(np.mean([r["n_iter"] for r in res])), 1),
            "dtype": {"fp32": "f32+f64", "bf16x2": "bf16x2+f64"}[a.precision],
            "data": "synthetic (seeded water clusters)",
            "config": {"workload": "C4: 512 snapshots, water count uniform in [4, 32], grid "
                                   "32^3/32, 48^3/38 or 64^3/44 bohr by water count (box >= "
                                   "radius + 20), 1-4 states, block 4, subspace 28, tol 1e-7",
                       "grids": {str(n): sum(s[1] == n for s in snaps) for n, _ in workload.MIX_GRIDS}}}


def fp64_reference(P, items):
    """Eigenvalues of the plain FP64 form (cold starts, the same server code)
    for (xyz, n, L, nroots) items."""
    sv = P.Server(hbm_cap_bytes=200 << 30, max_slots=len(items), precision=P.FP64,
                  maxsub=MAXSUB_FP64, ptol=PTOL)
    ids = [sv.submit((n,) * 3, (L,) * 3, xyz, nr) for xyz, n, L, nr in items]
    res = {r["id"]: r for r in sv.run_until_done()}
    sv.close()
    assert all(res[i]["status"] == res[i]["evals"].size for i in ids)
    return [res[i]["evals"] for i in ids]


def bench_c5(P, workload, prec, a, nchains=64):
    """Open-loop Poisson arrivals at four offered rates.  Requests belong to
    MD-like chains (each snapshot a rigid Gaussian move of the chain's previous
    one, 0.05 A / 2 deg) that warm-start from the chain's last converged
    subspace; sizes as in the batched mix plus 20% at 96^3; admission capped at
    200 GB of device memory."""
    rng = np.random.default_rng(SEED_C5)
    chains = []
    for c in range(nchains):
        xyz, n, L, nr = workload.mixed_snapshot(10000 + c, rng, SEED_C4)
        if rng.uniform() < 0.2:
            n, L = workload.MIX_BIG
        chains.append({"xyz": xyz, "n": n, "L": L, "nr": nr})
    sv = P.Server(hbm_cap_bytes=200 << 30, max_slots=512, precision=prec, maxsub=MAXSUB_MIX,
                  ptol=PTOL)
    # warm-up and capacity estimate: one request per chain, all at once
    t0 = time.perf_counter()
    for c, ch in enumerate(chains):
        sv.submit((ch["n"],) * 3, (ch["L"],) * 3, ch["xyz"], ch["nr"], chain=c, warm=False)
    sv.run_until_done()
    cap = nchains / (time.perf_counter() - t0)
    log("c5: closed-batch capacity %.1f solves/s" % cap)
    out_rates = []
    submitted, done = {}, {}
    for frac in (0.25, 0.5, 0.75, 0.95):
        rate = frac * cap
        t_start = time.perf_counter()
        t_end = t_start + a.duration
        next_t = t_start + rng.exponential(1.0 / rate)
        k = 0
        res = []
        while True:
            now = time.perf_counter()
            while next_t <= now and next_t < t_end:
                c = k % nchains
                ch = chains[c]
                ch["xyz"] = workload.perturb(ch["xyz"], rng)
                rid = sv.submit((ch["n"],) * 3, (ch["L"],) * 3, ch["xyz"], ch["nr"], chain=c,
                                warm=True)
                submitted[rid] = (ch["xyz"], ch["n"], ch["L"], ch["nr"])
                k += 1
                next_t += rng.exponential(1.0 / rate)
            live = sv.step()
            res += sv.poll()
            if now >= t_end and live == 0 and sv.pending() == 0:
                break
            if live == 0 and sv.pending() == 0:
                time.sleep(min(0.0005, max(0.0, next_t - now)))
        elapsed = time.perf_counter() - t_start
        done.update({r["id"]: r for r in res})
        lat = np.array([r["latency_s"] for r in res]) if res else np.zeros(1)
        out_rates.append({"offered_per_s": round(rate, 2), "submitted": k, "completed": len(res),
                          "solves_per_s": round(len(res) / elapsed, 2),
                          "p50_ms": round(1000 * float(np.percentile(lat, 50)), 2),
                          "p99_ms": round(1000 * float(np.percentile(lat, 99)), 2),
                          "iters_mean": round(float(np.mean([r["n_iter"] for r in res])), 1)
                          if res else None})
        log("c5: offered %.1f/s -> %.1f/s, p50 %.1f ms, p99 %.1f ms"
            % (rate, out_rates[-1]["solves_per_s"], out_rates[-1]["p50_ms"],
               out_rates[-1]["p99_ms"]))
    sv.close()
    # warm-started results against cold plain-FP64 solves of the same snapshots:
    # a seeded sample of 24 completed requests, the 96^3 ones among them
    big = [i for i in done if submitted[i][1] == workload.MIX_BIG[0]]
    small = [i for i in done if submitted[i][1] != workload.MIX_BIG[0]]
    pick = np.random.default_rng(SEED_C5 + 1)
    sample = list(pick.choice(big, min(8, len(big)), replace=False)) + \
        list(pick.choice(small, min(16, len(small)), replace=False))
    ev64 = fp64_reference(P, [submitted[i] for i in sample])
    de = max(float(np.abs(done[i]["evals"] - e).max()) for i, e in zip(sample, ev64))
    nconv = sum(r["status"] == r["evals"].size for r in done.values())
    log("c5: max |dE| of %d warm results vs cold FP64 %.2e" % (len(sample), de))
    best = max(out_rates, key=lambda r: r["solves_per_s"])
    return {"metric": "served solves/s and p50/p99 latency under open-loop Poisson load "
                      "(warm-started MD chains, 20% at 96^3)",
            "value": best["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
            "higher_is_better": True, "closed_batch_capacity_per_s": round(cap, 2),
            "rates": out_rates,
            "max_abs_dE_vs_fp64_cold_sample": float("%.3e" % de), "dE_sample": len(sample),
            "converged": nconv, "completed": len(done),
            "dE_ok": bool(de <= DE_TOL and nconv == len(done)),
            "dtype": {"fp32": "f32+f64", "bf16x2": "bf16x2+f64"}[a.precision],
            "data": "synthetic (seeded water clusters, rigid-move chains)",
            "config": {"workload": "C5: %d chains, sizes as C4 plus 20%% at 96^3/56 bohr, "
                                   "rigid moves 0.05 A / 2 deg, warm starts, admission cap "
                                   "200 GB, %.0f s per offered rate" % (nchains, a.duration),
                       "chains_96": sum(ch["n"] == 96 for ch in chains)}}


GRAM_CANDIDATES = ["0:8192", "1:4096", "2:2048", "2:4096", "2:8192"]


def tune_gram(a):
    """Autotuner of the Gram kernel (variant x rows per workgroup): one served
    run of the headline workload per candidate, each in a fresh process with the
    candidate forced through PISCES_GRAM; the candidate with the highest Gram
    throughput (algorithmic bytes over the HIP-event time of the Gram launches
    of the timed region) is written to pisces-1_amd/tuning.txt, which the
    library reads at start-up.  Whole served runs are the measure because the
    kernels' speed depends on the cache state the surrounding solver leaves."""
    import subprocess
    res = {}
    for cand in GRAM_CANDIDATES:
        env = dict(os.environ, PISCES_GRAM=cand)
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--served-only",
                              "--steps", "2", "--warmup", "1", "--precision", a.precision],
                             env=env, capture_output=True, text=True, check=True)
        d = json.loads(out.stdout.strip().splitlines()[-1])
        res[cand] = {"gram_GBs": d["launch_families"]["gram"]["GBs_algorithmic"],
                     "solves_per_s": d["value"]}
        log("tune: gram %s -> %.1f GB/s, %.2f solves/s" % (cand, res[cand]["gram_GBs"],
                                                            res[cand]["solves_per_s"]))
    best = max(res, key=lambda c: res[c]["gram_GBs"])
    v, r = best.split(":")
    path = os.path.join(ROOT, "pisces-1_amd", "tuning.txt")
    with open(path, "w") as f:
        f.write("# Gram kernel configuration: variant rows-per-workgroup\n")
        f.write("# written by `python bench.py --tune` from served runs of the (H2O)24 64^3 "
                "workload\n")
        for c in GRAM_CANDIDATES:
            f.write("# %s: %.1f GB/s\n" % (c, res[c]["gram_GBs"]))
        f.write("gram %s %s\n" % (v, r))
    return {"tuned": "gram", "best": best, "candidates": res, "written": path}


def copy_probe(P, nbytes, reps):
    """A plain streaming copy of nbytes, reps timed launches after one warm-up
    (pisces_copy_probe): under rocprofv3 --pmc it calibrates the byte counters,
    alone it gives the achievable copy bandwidth; then read-only streams of the
    same size with the default and the nontemporal load policy."""
    import ctypes
    L = P.lib()
    L.pisces_copy_probe.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double)]
    ms = ctypes.c_double()
    P.check(L.pisces_copy_probe(0, nbytes, reps, ctypes.byref(ms)), "pisces_copy_probe")
    out = {"bytes": nbytes, "launches": reps + 1, "ms_per_launch": ms.value,
           "GBs_read_plus_write": 2 * nbytes / (ms.value * 1e-3) / 1e9}
    # read-only streams over the same size (pisces_stream_probe), default and
    # nontemporal load policy
    L.pisces_stream_probe.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double)]
    for nt, key in ((0, "GBs_read_default"), (1, "GBs_read_nontemporal")):
        P.check(L.pisces_stream_probe(0, nbytes, reps, nt, ctypes.byref(ms)), "pisces_stream_probe")
        out[key] = nbytes / (ms.value * 1e-3) / 1e9
    return out


# kernel name fragments -> launch family of the solver's HIP-event brackets
# (first match wins)
KERNEL_FAMILIES = (("hx_f64_kernel", "hx_fp64"), ("hx_f_kernel<double", "hx_fp64"),
                   ("hx_bf2", "hx_lo"), ("hx_f_kernel<float", "hx_lo"), ("gram", "gram"),
                   ("comb_kernel", "combine"), ("eig_kernel", "eig"), ("ritz", "ritz_resid"),
                   ("resid", "ritz_resid"), ("fdm_", "precond"), ("precond_kernel", "precond"),
                   ("cholinv", "factors"), ("orth", "factors"), ("coef_kernel", "factors"),
                   ("potential_grid", "potential"), ("water_scf", "dipole_scf"))


def pmc_summary(fetch_csv, write_csv, copyf_csv, copyw_csv, roof_json):
    """HBM bytes per launch of each launch family: FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch) summed over the family's kernels, each scaled so that
    the plain copy kernel of --copy-probe reads and writes its byte count, over
    the launch count and algorithmic bytes of the --roofline-only line."""
    import csv

    def per_kernel(path):
        out = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                k = r["Kernel_Name"]
                out.setdefault(k, []).append(float(r["Counter_Value"]))
        return out

    def copy_kib(path):
        v = [x for k, xs in per_kernel(path).items() if "copy_probe_kernel" in k for x in xs]
        return sum(v) / len(v)

    with open(roof_json) as f:
        line = json.loads(f.read().strip().splitlines()[-1])
    fam = line["launch_families"]
    # the copy probe moves 1 GiB in and 1 GiB out per launch
    probe = 1 << 30
    cf, cw = copy_kib(copyf_csv), copy_kib(copyw_csv)
    fs, ws = probe / 1024.0 / cf, probe / 1024.0 / cw
    tot = {}
    for which, path, scale in (("fetch", fetch_csv, fs), ("write", write_csv, ws)):
        for k, xs in per_kernel(path).items():
            name = next((f for frag, f in KERNEL_FAMILIES if frag in k), None)
            if name is None:
                continue
            t = tot.setdefault(name, {"fetch": 0.0, "write": 0.0})
            t[which] += sum(xs) * 1024.0 * scale
    out = {"note": "HBM bytes per launch of each solver launch family in a one-lane roofline "
                   "pass (bench.py --roofline-only --warmup 0), from FETCH_SIZE and WRITE_SIZE "
                   "(KiB per dispatch, summed over the family's kernels), each scaled by what "
                   "the counter reads for a plain copy of 1 GiB; ratio = measured / algorithmic "
                   "bytes of the same pass; $R: the repository root, $OUT: the output folder",
           "fetch_scale": round(fs, 4), "write_scale": round(ws, 4),
           "copy_fetch_kib": cf, "copy_write_kib": cw,
           "commands": {
               "fetch": "cd /tmp && rocprofv3 --pmc FETCH_SIZE --output-format csv -d "
                        "$OUT/pmc_fetch -o f -- python3 $R/bench.py --roofline-only "
                        "--warmup 0",
               "write": "cd /tmp && rocprofv3 --pmc WRITE_SIZE --output-format csv -d "
                        "$OUT/pmc_write -o w -- python3 $R/bench.py --roofline-only "
                        "--warmup 0",
               "copy": "cd /tmp && rocprofv3 --pmc FETCH_SIZE --output-format csv -d "
                       "$OUT/pmc_copyf -o c -- python3 $R/bench.py --copy-probe "
                       "1073741824 --steps 2 (and WRITE_SIZE into pmc_copyw)",
               "algorithmic": "cd /tmp && rocprofv3 --kernel-trace --stats --output-format csv "
                              "-d $OUT/prof_roof -o bench -- python3 $R/bench.py "
                              "--roofline-only --warmup 1 > $OUT/roof.json",
               "summary": "python3 bench.py --pmc-summary $OUT/pmc_fetch/f_counter_collection.csv "
                          "$OUT/pmc_write/w_counter_collection.csv "
                          "$OUT/pmc_copyf/c_counter_collection.csv "
                          "$OUT/pmc_copyw/c_counter_collection.csv $OUT/roof.json"},
           "families": {}}
    for name, t in tot.items():
        f = fam.get(name)
        if not f or not f["launches"]:
            continue
        nl = f["launches"]
        e = {"launches": nl,
             "hbm_bytes_per_launch": round((t["fetch"] + t["write"]) / nl),
             "fetch_bytes_per_launch": round(t["fetch"] / nl),
             "write_bytes_per_launch": round(t["write"] / nl)}
        if f.get("GBs_algorithmic"):
            alg = f["GBs_algorithmic"] * 1e9 * f["ms_per_launch"] * 1e-3
            e["algorithmic_bytes_per_launch"] = round(alg)
            e["ratio"] = round((t["fetch"] + t["write"]) / nl / alg, 4)
        out["families"][name] = e
    return out


def measured_traffic(family, algorithmic_bytes_per_launch):
    """HBM bytes per launch of a launch family from the hardware byte counters
    (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE of the same served run,
    calibrated on a plain copy kernel), as the measured-to-algorithmic ratio of
    that run applied to this run's algorithmic bytes per launch."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        ratio = t["families"][family]["ratio"]
    except (OSError, KeyError, ValueError):
        return None, None
    return round(ratio * algorithmic_bytes_per_launch), round(ratio, 3)


def gradient_measure(P, workload, reps):
    """Electron side of the analytic gradient (DVR::ComputeGradient's grid loop)
    for the ground state of (H2O)24 on 64^3 with a PolType 3 potential (216
    polarizable dimensions, padded to 224): wall time per pisces_gradient_terms
    call (the wavefunction crosses PCIe once per call) after a warm-up, and the
    matrix work per grid point -- mu = (E + Epc) InvA, 2 n3p^2 flops, and the
    upper 16 x 16 blocks of sum psi^2 (E + Epc)(E + Epc)^T, ncb (ncb + 1) / 2 x
    512 flops -- against the FP64 peak."""
    flags = (3, 2, 0, 3, 0, 0)
    xyz, _ = workload.water_cluster(N_WATER, RADIUS, SEED)
    n, L = N_GRID, 40.0
    with P.Context((n,) * 3, (L,) * 3) as ctx:
        sites, _, _ = ctx.build_cluster_potential(xyz, flags=flags)
        r = ctx.solve(1, P.start_vectors(ctx, 1), maxsub=20, maxmacro=200, ptol=8, corrflag=1)
        psi = r["vecs"][0]
        ns, npol = len(sites), 3 * N_WATER
        ctx.gradient_terms(psi, ns, npol)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.gradient_terms(psi, ns, npol)
        ms = (time.perf_counter() - t0) / reps * 1e3
    n3p = (3 * npol + 15) // 16 * 16
    ncb = n3p // 16
    fpp = 2.0 * n3p * n3p + ncb * (ncb + 1) / 2 * 512
    fl = fpp * n ** 3
    return {"workload": "(H2O)24 anion ground state, 64^3 on 40 bohr, PolType 3 (216 dims)",
            "ms_per_call_wall": round(ms, 4), "flops_per_point": fpp, "flops_per_call": fl,
            "TFLOPs_wall": round(fl / (ms * 1e-3) / 1e12, 2), "peak_TFLOPs": FP64_PEAK_TFS,
            "note": "wall time includes the 2 MB wavefunction upload and the host's final sums; "
                    "the kernels' own time is in the rocprofv3 kernel trace"}


def pol3_measure(P, workload):
    """Grid build of the potential with self-consistent polarizable sites
    (PolType 3) for (H2O)24 on 64^3 (216 polarizable dimensions) and (H2O)45 on
    96^3 (405): the FP64 matrix-core kernel's time from HIP events around its
    launch, the flops it executes (the upper block triangle of the symmetric
    InvA, padded to 16) and the fraction of the FP64 matrix peak."""
    out = {}
    flags = (3, 2, 0, 3, 0, 0)
    for tag, nw, radius, seed, n, L in (("h2o24_64cube", 24, 6.0, 2000, 64, 40.0),
                                        ("h2o45_96cube", 45, 8.0, 3000, 96, 48.0)):
        xyz, _ = workload.water_cluster(nw, radius, seed)
        with P.Context((n,) * 3, (L,) * 3) as ctx:
            ctx.build_cluster_potential(xyz, flags=flags)  # warm-up
            P.prof_enable(True)
            P.prof_reset()
            t0 = time.perf_counter()
            for _ in range(5):
                ctx.build_cluster_potential(xyz, flags=flags)
            t_build = (time.perf_counter() - t0) / 5
            fam = P.prof_read()
            P.prof_enable(False)
        ms, nl, _ = fam["pol_scp"]
        t = ms / nl * 1e-3
        n3 = 9 * nw
        ncb = (n3 + 15) // 16
        ng = n ** 3
        # MFMA flops of the launch: per 16 points, (cb + 1) 16x16x16 tiles of
        # 2 * 16^3 flops for column block cb of the upper block triangle
        fl_exec = 256.0 * ncb * (ncb + 1) * ng
        fl_ref = 2.0 * n3 * n3 * ng                     # the reference's dsymv per point
        out[tag] = {"n_pol_dims": n3, "grid": n, "ms_kernel": round(1e3 * t, 4),
                    # the whole build: water model, InvA on the host, uploads, kernels
                    "ms_build_total": round(1e3 * t_build, 3),
                    "TFLOPs_executed": round(fl_exec / t / 1e12, 2),
                    "TFLOPs_dsymv_equivalent": round(fl_ref / t / 1e12, 2),
                    "peak_TFLOPs": FP64_MATRIX_PEAK_TFLOPS,
                    "frac_executed": round(fl_exec / t / 1e12 / FP64_MATRIX_PEAK_TFLOPS, 4)}
    return out


def cpu_baseline(P, xyz, seconds):
    """The oracle (C restatement of the reference path: DPP potential with
    induced dipoles, dense-T H*x, the reference Davidson) on the host cores, on
    a bounded sample: one potential build and a Davidson truncated to a few
    macro iterations.  The solve time is extrapolated to the matvec count the
    reference algorithm takes on this snapshot, measured with the GPU
    restatement of that algorithm (pisces_solve, FP64), which reproduces the
    reference's matvec count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    oracle.set_threads(cores)
    n, L = (N_GRID,) * 3, (L_BOX,) * 3
    # matvec count of the reference algorithm (one vector per root, subspace 32,
    # at most 200 macro iterations) on this snapshot
    with P.Context(n, L) as ctx:
        ctx.build_cluster_potential(xyz)
        start = P.start_vectors(ctx, NROOTS, seed=1)
        rref = ctx.solve(NROOTS, start, REF_MAXSUB, 200, PTOL, 1, P.FP64)
    t0 = time.perf_counter()
    cl = oracle.Cluster(xyz)
    g = oracle.Grid(n, L)
    V = cl.potential_grid(g)
    t_pot = time.perf_counter() - t0
    macro = 1
    while True:
        t1 = time.perf_counter()
        rr = g.solve(V, NROOTS, REF_MAXSUB, macro, PTOL, 1, start)
        t_dav = time.perf_counter() - t1
        if t_dav > seconds / 3 or rr["nconv"] == NROOTS or macro >= 64:
            break
        macro *= 2
    per_mv = t_dav / max(1, rr["n_mtx"])
    t_solve = t_pot + per_mv * rref["n_matvec"]
    # one core: the same per-matvec time on a single thread (two macro iterations)
    oracle.set_threads(1)
    t1 = time.perf_counter()
    r1 = g.solve(V, NROOTS, REF_MAXSUB, 2, PTOL, 1, start)
    per_mv1 = (time.perf_counter() - t1) / max(1, r1["n_mtx"])
    t1p = time.perf_counter()
    cl.potential_grid(g)
    t_pot1 = time.perf_counter() - t1p
    oracle.set_threads(cores)
    # a completed solve at the CPU-runnable configuration: (H2O)6, 32^3 on 36
    # bohr, one state, residual 1e-7, the reference algorithm to the end
    xyz1, _ = __import__("workload").water_cluster(6, 3.5, 1000)
    t2 = time.perf_counter()
    cl1 = oracle.Cluster(xyz1)
    g1 = oracle.Grid((32,) * 3, (36.0,) * 3)
    V1 = cl1.potential_grid(g1)
    r1c = g1.solve(V1, 1, 20, 200, PTOL, 1, g1.piab().reshape(1, -1))
    t_c1 = time.perf_counter() - t2
    return {"value": round(1.0 / t_solve, 6), "unit": "solves/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": "oracle potential build (%.2f s) + its Davidson limited to %d macro "
                      "iterations (%d matvecs, %.1f s); extrapolated to the reference "
                      "algorithm's %d matvecs on this snapshot (it converges %d of %d states "
                      "within 200 macro iterations)"
                      % (t_pot, macro, rr["n_mtx"], t_dav, rref["n_matvec"], rref["nconv"],
                         NROOTS),
            "one_core": {"solves_per_s": round(1.0 / (t_pot1 + per_mv1 * rref["n_matvec"]), 7),
                         "s_per_matvec": round(per_mv1, 4), "potential_s": round(t_pot1, 3),
                         "sample": "the same extrapolation from 2 macro iterations on 1 thread"},
            "c1_completed": {"solves_per_s": round(1.0 / t_c1, 4), "seconds": round(t_c1, 3),
                             "cores": cores, "nconv": r1c["nconv"], "matvecs": r1c["n_mtx"],
                             "eval": round(float(r1c["evals"][0]), 10),
                             "workload": "(H2O)6, 32^3 on 36 bohr, 1 state, tol 1e-7, potential "
                                         "build + the reference Davidson to convergence"}}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
