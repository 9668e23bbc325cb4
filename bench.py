"""Benchmark of the force hot path on one MI355X.

Workload (the SPH adiabatic box): N = 2^20 gas particles uniform in a
periodic unit cube, equal mass, u = 1, gamma = 5/3, Gaussian velocities
(sigma 0.1), M4 kernel, nSmooth = 64, Monaghan viscosity alpha = 1,
beta = 2.  One step = gg_build_local + gg_smooth(DenDVDX) + gg_gas_pressure
+ gg_resmooth(SphPressureTermsSym), inputs resident in HBM, the previous
step's smoothing lengths discarded so every kNN search starts cold.

value = SPH pair evaluations per second (density pass + force pass, each
pass counted as sum_p |N(p)|), FP32 pair arithmetic.  The same step in plain
FP64 is timed beside it.  A gravity leg (Plummer sphere N = 2^20, theta 0.7,
quadrupole, bucket 16; the fused walk + interaction kernel, tree build
excluded) reports interactions/s in both precisions.  The CPU restatement of the same path is
timed on a bounded sample on the host cores.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grav interactions/s + SPH pair-evals/s at rms force err<=1e-4, vs plain FP64"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3          # MI355X FP32 vector peak (packed FMA)
FP32_NONPACKED_TFLOPS = 78.6      # without packed FP32, which this build does not emit
# Algorithmic bytes per unit of work (the data the operator must touch once):
# kNN per query: its own position and fBall2, and per neighbour the j-side
# position a distance needs (24 B, as the density pass counts its j record)
# plus the list entry written (4 B); the strict output-only model (32 B +
# 4 K per query) is reported beside it
KNN_BYTES_PER_QUERY = 24 + 8     # + 28 K
KNN_BYTES_PER_NEIGHBOUR = 24 + 4
# SPH passes per pair: the j-side record as an FP32 SoA record plus flags,
# and the materialised neighbour index: density r 12, m 4, vPred 12, uPred 4,
# flags 4 (+4); force r 12, m 4, vPred 12, rho 4, P/rho^2 4, c 4, Balsara 4,
# dt 4, flags 4 (+4)
DENS_BYTES_PER_PAIR = 36 + 4
HYDRO_BYTES_PER_PAIR = 52 + 4
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def sph_setup(gg, n, prec, seed=2):
    from gasoline_amd import ics
    d = ics.sph_box(n, seed=seed)
    pk = gg.Pkd(n, prec=prec)
    pk.put("r", d["r"])
    pk.put("fMass", d["m"])
    pk.put("vPred", d["v"])
    pk.put("uPred", d["u"])
    pk.build_local(period=[1.0, 1.0, 1.0])
    pk.reorder()                         # the store in the structure's (Morton) order
    return pk


def sph_step(pk, smf, zero_fb2, K=64):
    pk.put("fBall2", zero_fb2)           # device-to-device: cold kNN search
    pk.build_local(period=[1.0, 1.0, 1.0])
    pk.smooth(smf, K, True)
    pk.gas_pressure(smf)
    pk.resmooth(smf, K, True)


def time_sph(gg, torch, n, prec, steps, warmup):
    smf = gg.Smf.default()
    pk = sph_setup(gg, n, prec)
    zero = torch.zeros(n, dtype=torch.float64, device="cuda")
    for _ in range(warmup):
        sph_step(pk, smf, zero)
    torch.cuda.synchronize()
    phases = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sph_step(pk, smf, zero)
        t = pk.timings()
        phases.append(t)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    pk.close()
    return dt, ph


def time_grav(gg, torch, n, prec, steps, warmup):
    from gasoline_amd import ics
    r, v, m = ics.plummer(n, seed=1)
    soft = np.full(n, 0.01)
    pk = gg.Pkd(n, prec=prec)
    pk.put("r", r)
    pk.put("fMass", m)
    pk.put("fSoft", soft)
    pk.build_binary(16, 1, 0.7, 2)
    st = None
    for _ in range(warmup):
        st = pk.grav_all(2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    walk, inter = [], []
    for _ in range(steps):
        st = pk.grav_all(2)
        t = pk.timings()
        walk.append(t["walk"])
        inter.append(t["interact"])
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(3):
        pk.build_binary(16, 1, 0.7, 2)
    tb = (time.perf_counter() - t0) / 3
    pk.close()
    nint = st["dPartSum"] + st["dCellSum"] + st["dSoftSum"]
    return dt, nint, st["dFlop"], float(np.mean(walk)), float(np.mean(inter)), tb


def log(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def periodic_box(n, seed=21):
    """Half uniform, half in three Gaussian clumps (sigma 0.03), wrapped into
    the periodic unit box [-1/2, 1/2)^3; equal masses 1/n."""
    from gasoline_amd import ics
    rng = np.random.default_rng(seed)
    u = rng.uniform(-0.5, 0.5, (n // 2, 3))
    c = rng.uniform(-0.5, 0.5, (3, 3))
    k = rng.integers(0, 3, n - n // 2)
    cl = c[k] + rng.normal(0, 0.03, (n - n // 2, 3))
    return ics._f32((np.concatenate([u, cl]) + 0.5) % 1.0 - 0.5), np.full(n, 1.0 / n)


def time_periodic(gg, torch, n, prec, steps, warmup):
    """Periodic gravity with the reference's defaults for it: nReps 1,
    hexadecapole walk and Ewald (iOrder = iEwOrder = 4), dEwCut 2.6, dEwhCut
    2.8, theta 0.7; the same step without Ewald gives the Ewald pass's time."""
    r, m = periodic_box(n)
    pk = gg.Pkd(n, prec=prec)
    pk.put("r", r)
    pk.put("fMass", m)
    pk.put("fSoft", np.full(n, 0.2 / n ** (1 / 3)))
    pk.set_period([1.0, 1.0, 1.0])
    pk.build_binary(16, 1, 0.7, 4)
    out = {}
    for ew in (True, False):
        for _ in range(warmup):
            st = pk.grav_all(4, nReps=1, periodic=True, bEwald=ew)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = pk.grav_all(4, nReps=1, periodic=True, bEwald=ew)
        torch.cuda.synchronize()
        out[ew] = ((time.perf_counter() - t0) / steps, st)
    pk.close()
    dt, st = out[True]
    nint = st["dPartSum"] + st["dCellSum"] + st["dSoftSum"]
    return {"ms": dt * 1e3, "ms_without_ewald": out[False][0] * 1e3,
            "interactions": nint, "interactions_per_s": nint / dt, "flop": st["dFlop"]}


def time_ensemble(gg, torch, e, steps, prec):
    """Batched ensemble: nsys King models (W0 in [3, 9], N log-uniform in
    [1024, 65536]), eps = 0.005 King radii, theta 0.6, quadrupole; one tree
    per system in one forest, bucket size by the autotuner, the gravity step
    replayed from a HIP graph.  Returns times per step (eager and graph)."""
    n = len(e["m"])
    nsys = len(e["off"]) - 1
    pk = gg.Pkd(n, prec=prec)
    pk.put("r", e["r"])
    pk.put("fMass", e["m"])
    pk.put("fSoft", np.full(n, 0.005))
    best, tune = pk.tune_bucket((8, 12, 16, 24, 32), sysOff=e["off"], dCrit=0.6)
    st = pk.grav_all(2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = pk.grav_all(2)
    torch.cuda.synchronize()
    eager = (time.perf_counter() - t0) / steps
    pk.grav_graph(2)
    pk.grav_replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = pk.grav_replay()
    torch.cuda.synchronize()
    graph = (time.perf_counter() - t0) / steps
    pk.close()
    nint = st["dPartSum"] + st["dCellSum"] + st["dSoftSum"]
    return dict(n=n, nsys=nsys, nBucket=best, tune_ms=tune, ms_eager=eager * 1e3,
                ms_graph=graph * 1e3, interactions=nint, interactions_per_s=nint / graph)


def force_errors(gg, d, soft, bits, theta, nsample=10000, seed=1, nBucket=16):
    """rms and max relative gravity-force error of the FP32 and FP64 tree
    forces against the FP64 direct sum (gg_direct_grav) on nsample random
    targets, and of FP32 against FP64 on the same tree."""
    from gasoline_amd import evolve
    n = len(d["m"])
    idx = np.sort(np.random.default_rng(seed).choice(n, min(nsample, n), replace=False))
    acc = {}
    for name, prec in (("fp32", gg.PREC_FP32), ("fp64", gg.PREC_FP64)):
        with gg.Pkd(n, prec=prec) as pk:
            evolve.start(pk, d, soft, bits)
            pk.build_binary(nBucket, 1, theta, 2)
            pk.init_accel()
            pk.grav_all(2)
            acc[name] = pk.get("a")[idx]
            if name == "fp64":
                acc["direct"], _ = pk.direct_grav(idx)

    def err(x, ref):
        e = np.linalg.norm(x - ref, axis=1) / np.linalg.norm(ref, axis=1)
        return {"rms": float(np.sqrt(np.mean(e * e))), "max": float(e.max())}
    return {"targets": int(len(idx)), "fp32_vs_direct": err(acc["fp32"], acc["direct"]),
            "fp64_vs_direct": err(acc["fp64"], acc["direct"]),
            "fp32_vs_fp64": err(acc["fp32"], acc["fp64"])}


def time_kdk(gg, torch, d, soft, bits, prec, mode, K, dt, theta, steps, warmup, visc=1):
    """Kick-drift-kick steps with gravity + SPH forces (evolve.KDK): the tree
    build, gravity walk and interactions, SPH tree, kNN + DenDVDX, EOS and
    SphPressureTermsSym each step.  mode: "mlp" (h-proposal network seeds
    the search, trained online), "prev" (previous fBall2 seeds it) or "cold"
    (no seed: the plain search).  The store is kept in the SPH structure's
    order (gg_reorder after each SPH build)."""
    from gasoline_amd import evolve
    n = len(d["m"])
    pk = gg.Pkd(n, prec=prec)
    evolve.start(pk, d, soft, bits)
    hm = None
    if mode == "mlp":
        pk.hmlp_init(1)
        hm = dict(train_iters=50, batch=8192, lr=3e-3)
    pk.set_seed_slack({"cold": 0.0, "mlp": 1.08, "prev": 1.15}[mode])
    run = evolve.KDK(pk, gg.Smf.default(iViscosityLimiter=visc), nSmooth=K, theta=theta,
                     hmlp=hm, reorder=True)
    run.forces(dt)
    for _ in range(warmup):
        run.step(dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.step(dt)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    st, sm = run.grav_stats, run.stats[-1]
    pk.close()
    return {"ms_per_step": el * 1e3, "mode": mode,
            "grav_interactions": st["dPartSum"] + st["dCellSum"] + st["dSoftSum"],
            "knn_candidates": int(sm["candidates"]), "knn_extra_passes": int(sm["extra_passes"])}


def cpu_baseline_sph(n=1 << 20):
    """The oracle's SPH step (kNN + DenDVDX + EOS + SphPressureTermsSym) on a
    periodic box of n particles, pair-evals/s on the host cores."""
    from gasoline_amd import ics
    from oracle import oracle as O
    d = ics.sph_box(n, seed=3)
    smf = O.default_smf()
    t0 = time.perf_counter()
    nbr, fb2, _ = O.knn(d["r"], 64, period=[1, 1, 1])
    lists = O.ball_gather(d["r"], fb2, period=[1, 1, 1])
    den = O.dendvdx(d["m"], d["v"], d["u"], fb2, lists, smf)
    pr2, c, dt = O.gas_pressure(d["u"], den["rho"], fb2, smf)
    O.sph_pressure_sym(d["m"], d["v"], d["u"], den["rho"], pr2, c, den["balsara"],
                       den["diff"], dt, fb2, lists, smf)
    el = time.perf_counter() - t0
    pairs = 2 * len(lists[1])
    return pairs / el, el


def cpu_baseline_grav(n=1 << 20):
    from gasoline_amd import ics
    from oracle import oracle as O
    r, v, m = ics.plummer(n, seed=1)
    soft = np.full(n, 0.01)
    T = O.Tree(r, m, soft, nBucket=16, dTheta=0.7, iOrder=2)
    t0 = time.perf_counter()
    _, _, _, st = T.grav_all()
    el = time.perf_counter() - t0
    return (st["dPartSum"] + st["dCellSum"] + st["dSoftSum"]) / el, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--ngrav", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nsys", type=int, default=256, help="ensemble systems (0: skip)")
    ap.add_argument("--nevrard", type=int, default=1 << 24,
                    help="Evrard collapse particles of the full-step leg (0: skip)")
    ap.add_argument("--disk", type=int, default=1, help="disk-galaxy full-step leg (0: skip)")
    ap.add_argument("--periodic", type=int, default=1 << 20,
                    help="particles of the periodic-gravity (Ewald) leg (0: skip)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run five device-to-device copies of MIB MiB (the known-size "
                         "kernel that scales rocprof FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": a.numel() * 4, "copies": 5}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the force path runs on one GPU's share "
              "and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import torch
    import gasoline_amd as gg

    n = args.n
    K = 64
    pairs = 2 * K * n                      # density pass + force pass
    t32, ph32 = time_sph(gg, torch, n, gg.PREC_FP32, args.steps, args.warmup)
    t64, ph64 = time_sph(gg, torch, n, gg.PREC_FP64, max(2, args.steps // 2), 1)
    log("SPH box done")
    g32 = time_grav(gg, torch, args.ngrav, gg.PREC_FP32, args.steps, args.warmup)
    g64 = time_grav(gg, torch, args.ngrav, gg.PREC_FP64, max(2, args.steps // 2), 1)
    log("gravity done")

    per = None
    if args.periodic > 0:
        p32 = time_periodic(gg, torch, args.periodic, gg.PREC_FP32, args.steps, args.warmup)
        p64 = time_periodic(gg, torch, args.periodic, gg.PREC_FP64, max(2, args.steps // 2), 1)
        per = {"workload": "periodic unit box N=%d (half uniform, half in three clumps of "
                           "sigma 0.03), eps 0.2 n^-1/3, theta 0.7, hexadecapole walk, nReps 1, "
                           "Ewald iEwOrder 4, dEwCut 2.6, dEwhCut 2.8; gg_grav_all (tree build "
                           "excluded)" % args.periodic,
               "fp32": p32, "fp64": p64, "speedup_vs_fp64": p64["ms"] / p32["ms"]}
        log("periodic gravity done")
    ens32 = ens64 = None
    if args.nsys > 0:
        from gasoline_amd import ics
        e = ics.king_ensemble(args.nsys, seed=7)
        ens32 = time_ensemble(gg, torch, e, args.steps, gg.PREC_FP32)
        ens64 = time_ensemble(gg, torch, e, max(2, args.steps // 2), gg.PREC_FP64)
        log("ensemble done")
    evr = None
    if args.nevrard > 0:
        from gasoline_amd import ics
        d = ics.evrard(args.nevrard, seed=5)
        soft = np.full(args.nevrard, 0.004)
        ks = max(2, args.steps // 2)
        fast = time_kdk(gg, torch, d, soft, None, gg.PREC_FP32, "mlp", 64, 1e-3, 0.7, ks, 3)
        log("Evrard FP32 done")
        plain = time_kdk(gg, torch, d, soft, None, gg.PREC_FP64, "cold", 64, 1e-3, 0.7, ks, 1)
        log("Evrard FP64 done")
        evr = {"workload": "Evrard collapse N=%d (rho ~ 1/r, M=R=G=1, u=0.05, gamma 5/3), "
                           "eps 0.004, theta 0.7, quadrupole, bucket 16, nSmooth 64, dt 0.001 "
                           "(about the central Courant limit); "
                           "KDK steps (gravity tree build + walk/interactions, SPH tree, kNN, "
                           "DenDVDX, EOS, SphPressureTermsSym)" % args.nevrard,
               "fp32_mlp": fast, "fp64_plain": plain,
               "speedup_vs_fp64_plain": plain["ms_per_step"] / fast["ms_per_step"],
               "force_error": force_errors(gg, d, soft, None, 0.7)}
        del d
    disk = None
    if args.disk:
        from gasoline_amd import ics
        d = ics.disk_galaxy(seed=11)
        gas_bits = gg.TYPE_ACTIVE | gg.TYPE_TREEACTIVE | gg.TYPE_SMOOTHACTIVE | gg.TYPE_GAS
        bits = np.where(d["gas"], gas_bits, 3).astype(np.uint32)
        ks = max(2, args.steps // 2)
        f32 = time_kdk(gg, torch, d, d["soft"], bits, gg.PREC_FP32, "prev", 32, 5e-4, 0.7, ks, 1)
        f64 = time_kdk(gg, torch, d, d["soft"], bits, gg.PREC_FP64, "prev", 32, 5e-4, 0.7, ks, 1)
        disk = {"workload": "isolated disk galaxy: Hernquist halo 4194304 DM (1e12 Msun, a 30 "
                            "kpc, cut at 10a) + exponential gas disk 1048576 (R_d 3 kpc, sech^2 "
                            "z0 0.3 kpc, min Toomre Q 1.5, T 1e4 K, mu 1.22), eps 0.1/0.05 kpc, "
                            "theta 0.7, quadrupole, nSmooth 32, dt 5e-4 (%s)" % d["units"],
                "fp32": f32, "fp64": f64, "speedup_vs_fp64": f64["ms_per_step"] / f32["ms_per_step"],
                "force_error": force_errors(gg, d, d["soft"], bits, 0.7)}
        del d
        log("disk galaxy done")
    # roofline of each SPH kernel: algorithmic bytes per launch over its
    # HIP-event duration on the handle's stream; the dominant one is reported
    kern = {
        "k_knn_bucket": (ph32["knn"], n * (KNN_BYTES_PER_QUERY + KNN_BYTES_PER_NEIGHBOUR * K)),
        "k_dendvdx<float>": (ph32["density"], n * K * DENS_BYTES_PER_PAIR),
        "k_hydro2<float>": (ph32["hydro"], n * K * HYDRO_BYTES_PER_PAIR),
    }
    traffic = {}
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f)

    def roof(name):
        ms, by = kern[name]
        a = by / (ms * 1e-3) / 1e9
        t = traffic.get(name)
        return {"bound": "hbm", "kernel": name, "achieved": a, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": a / HBM_PEAK_GBS, "ms": ms, "algorithmic_bytes": by,
                "traffic": (t["fetch_bytes"] + t["write_bytes"]) if t else None}

    dom = max(kern, key=lambda k: kern[k][0])
    knn_strict_gbs = n * (KNN_BYTES_PER_QUERY + 4 * K) / (ph32["knn"] * 1e-3) / 1e9
    g_inter_s = g32[1] / (g32[0])
    line = {
        "metric": METRIC,
        "value": pairs / t32,
        "unit": "SPH pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "SPH adiabatic box: N=%d periodic unit cube, nSmooth=64, M4, "
                        "alpha=1 beta=2, one DenDVDX + one SphPressureTermsSym pass "
                        "(tree build, kNN, EOS and reverse lists inside the step)" % n,
            "n": n, "nSmooth": K, "pairs_per_step": pairs,
            "gravity_workload": "Plummer N=%d, eps=0.01, theta=0.7, iOrder=2, nBucket=16; "
                                "fused walk+interact (tree build excluded)" % args.ngrav,
        },
        "fp64_value": pairs / t64,
        "speedup_vs_fp64": t64 / t32,
        "phases_ms_fp32": ph32,
        "phases_ms_fp64": ph64,
        "grav_interactions_per_s": g_inter_s,
        "grav_interactions_per_s_fp64": g64[1] / g64[0],
        "grav_speedup_vs_fp64": g64[0] / g32[0],
        "grav_ms_walk": g32[3], "grav_ms_interact_fp32": g32[4], "grav_ms_interact_fp64": g64[4],
        "grav_ms_tree_build": g32[5] * 1e3,
        "roofline": roof(dom),
        "rooflines_other": [roof(k) for k in kern if k != dom],
        "knn_output_only_gbs": knn_strict_gbs,
        "roofline_gravity": {
            "bound": "valu", "kernel": "k_walk_interact<float,2> (walk included)",
            "achieved": g32[2] / (g32[4] * 1e-3) / 1e12, "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": g32[2] / (g32[4] * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
            "frac_nonpacked": g32[2] / (g32[4] * 1e-3) / 1e12 / FP32_NONPACKED_TFLOPS,
            "traffic": (traffic["k_walk_interact<float,2>"]["fetch_bytes"]
                        + traffic["k_walk_interact<float,2>"]["write_bytes"])
            if "k_walk_interact<float,2>" in traffic else None,
        },
    }
    if ens32:
        line["ensemble"] = {
            "workload": "%d King models (W0 in [3,9], N log-uniform in [1024,65536], %d "
                        "particles), eps 0.005 r0, theta 0.6, quadrupole; one forest, one "
                        "launch; bucket size autotuned; HIP-graph replay" % (ens32["nsys"],
                                                                           ens32["n"]),
            "fp32": ens32, "fp64": ens64,
            "speedup_vs_fp64": ens64["ms_graph"] / ens32["ms_graph"],
        }
    if per:
        line["periodic_gravity"] = per
    if evr:
        line["evrard"] = evr
    if disk:
        line["disk_galaxy"] = disk
    if not args.no_cpu:
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        v_sph, el_sph = cpu_baseline_sph()
        log("CPU baseline (SPH) done")
        v_gr, el_gr = cpu_baseline_grav()
        log("CPU baseline (gravity) done")
        line["cpu_baseline"] = {
            "value": v_sph, "unit": "SPH pair-evals/s", "cores": cores, "kind": "port",
            "sample": "oracle SPH step on the benchmark's box, 1048576 particles (%.1f s): "
                      "kNN and DenDVDX on %d OpenMP threads, the symmetric force pass "
                      "serial as in the reference" % (el_sph, cores),
            "grav_interactions_per_s": v_gr,
            "grav_sample": "oracle gravity (1 thread, walk + interactions) on Plummer "
                           "N=1048576, theta 0.7, quadrupole, bucket 16 (%.1f s)" % el_gr,
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
