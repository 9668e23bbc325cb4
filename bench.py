"""Benchmark of the force hot path on one MI355X.

Headline (the SPH adiabatic box): N = 2^20 gas particles uniform in a
periodic unit cube, equal mass, u = 1, gamma = 5/3, Gaussian velocities
(sigma 0.1), M4 kernel, nSmooth = 64, Monaghan viscosity alpha = 1,
beta = 2.  One step = gg_build_local + gg_smooth(DenDVDX) + gg_gas_pressure
+ gg_resmooth(SphPressureTermsSym), inputs resident in HBM, the previous
step's smoothing lengths discarded so every kNN search starts cold.

value = SPH pair evaluations per second (density pass + force pass, each
pass counted as sum_p |N(p)|) in FP64 pair arithmetic, the reference's
precision; the FP32 form is timed beside it with its per-particle difference
from the FP64 form.  Gravity legs (interactions/s, FP32 against the plain
FP64 form) run at an operating point whose rms force error against the FP64
direct sum is at most 1e-4 (hexadecapole cells, theta 0.55 or 0.5) and at the
configurations' own theta / quadrupole for comparison, each with its rms and
max force error against the FP64 direct sum on sampled targets: a Plummer
sphere, the batched King ensemble, the Evrard collapse (the full gravity +
SPH step, 100 steps), the disk galaxy and the periodic box (against the
Ewald-summed periodic direct sum, at the reference's nReps 1 and at two
replica shells).  The CPU
restatement of the same path is timed on a bounded sample on the host cores.

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


# Gravity operating points: per workload the cheapest point of a measured
# theta / iOrder sweep (FP32 time of gg_grav_all) whose rms force error
# against the FP64 direct sum is at most 1e-4 with a quarter to spare
# (hexadecapole cells at theta 0.55; 0.5 for the King ensemble, whose
# error at 0.55 is 1.2e-4); the configurations' own quadrupole thetas are
# timed beside it.
GATE = dict(theta=0.55, iOrder=4)
GATE_ENSEMBLE = dict(theta=0.5, iOrder=4)
# The periodic box: (theta, nReps), hexadecapole walk and Ewald.
PERIODIC_GATE = (0.35, 2)


def log(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def rel_errors(x, ref):
    """rms and max of |x - ref| / |ref| per particle (rows)."""
    x = np.asarray(x, float).reshape(len(x), -1)
    ref = np.asarray(ref, float).reshape(len(ref), -1)
    e = np.linalg.norm(x - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
    return {"rms": float(np.sqrt(np.mean(e * e))), "max": float(e.max())}


# ------------------------------------------------------------ SPH box
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


def sph_step(pk, smf, zero_fb2, zero_a, K=64, nBucket=32):
    pk.put("fBall2", zero_fb2)           # device-to-device: cold kNN search
    pk.put("a", zero_a)                  # the force pass adds onto a
    pk.build_local(period=[1.0, 1.0, 1.0], nBucket=nBucket)
    pk.smooth(smf, K, True)
    pk.gas_pressure(smf)
    pk.resmooth(smf, K, True)


def time_sph(gg, torch, n, prec, steps, warmup, tune=None, nBucket=32, lanes=(4, 8, 16)):
    """The SPH step, its phases and outputs; with tune = candidate bucket
    sizes, gg_tune_sph chooses the structure's bucket size first (from a
    cold start, as every timed step searches), then gg_tune_sph_lanes the
    lanes per particle of DenDVDX and the force pass."""
    smf = gg.Smf.default()
    pk = sph_setup(gg, n, prec)
    zero = torch.zeros(n, dtype=torch.float64, device="cuda")
    zero3 = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    nb, tuned = nBucket, {"nBucket": nBucket}
    if tune:
        pk.put("fBall2", zero)
        nb, cand_ms = pk.tune_sph(smf, 64, [1.0, 1.0, 1.0], cand=tune)
        tuned = {"nBucket": nb, "candidate_ms": cand_ms}
    if lanes:
        pk.put("fBall2", zero)
        pk.put("a", zero3)
        pk.build_local(period=[1.0, 1.0, 1.0], nBucket=nb)
        lp, lane_ms = pk.tune_sph_lanes(smf, 64, True, cand=lanes)
        tuned.update(lanes=lp, lanes_candidate_ms=lane_ms)
    for _ in range(warmup):
        sph_step(pk, smf, zero, zero3, nBucket=nb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sph_step(pk, smf, zero, zero3, nBucket=nb)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    # the phase breakdown from steps of their own: reading the timers waits
    # for the stream, which the timed steps above do not do
    phases = []
    for _ in range(3):
        sph_step(pk, smf, zero, zero3, nBucket=nb)
        phases.append(pk.timings())
    ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    out = {f: pk.get(f) for f in ("a", "fDensity", "uDotPdV", "uDotAV")}
    pk.close()
    return dt, ph, out, tuned


# ------------------------------------------------------------ gravity
def direct_forces(gg, r, m, soft, idx, sysOff=None):
    """FP64 direct-sum accelerations of the sampled particles idx (each
    system on its own when sysOff is given)."""
    if sysOff is None:
        with gg.Pkd(len(m), prec=gg.PREC_FP64) as pk:
            pk.put("r", r)
            pk.put("fMass", m)
            pk.put("fSoft", soft)
            return pk.direct_grav(idx)[0]
    ad = np.zeros((len(idx), 3))
    for s in range(len(sysOff) - 1):
        a0, a1 = int(sysOff[s]), int(sysOff[s + 1])
        sel = (idx >= a0) & (idx < a1)
        if not sel.any():
            continue
        with gg.Pkd(a1 - a0, prec=gg.PREC_FP64) as pk:
            pk.put("r", r[a0:a1])
            pk.put("fMass", m[a0:a1])
            pk.put("fSoft", soft[a0:a1])
            ad[sel] = pk.direct_grav(idx[sel] - a0)[0]
    return ad


def sample(n, k=10000, seed=1):
    return np.sort(np.random.default_rng(seed).choice(n, min(k, n), replace=False))


def time_grav(gg, torch, r, m, soft, prec, theta, order, steps, warmup, idx=None,
              sysOff=None, bits=None):
    """gg_grav_all on one tree (or forest): time per call, walk/interaction
    kernel times, counters, the tree build time and the accelerations of the
    sampled particles idx."""
    n = len(m)
    pk = gg.Pkd(n, prec=prec)
    pk.put("r", r)
    pk.put("fMass", m)
    pk.put("fSoft", soft)
    if bits is not None:
        pk.put("iActive", bits)

    def build():
        if sysOff is None:
            pk.build_binary(16, 1, theta, order)
        else:
            pk.build_forest(sysOff, 16, 1, theta, order)
    build()
    pk.init_accel()
    st = pk.grav_all(order)
    acc = pk.get("a")[idx] if idx is not None else None
    for _ in range(warmup):
        pk.grav_all(order)
    torch.cuda.synchronize()
    inter = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = pk.grav_all(order)
        inter.append(pk.timings()["interact"])
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(3):
        build()
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / 3
    pk.close()
    nint = st["dPartSum"] + st["dCellSum"] + st["dSoftSum"]
    return {"ms": dt * 1e3, "ms_interact": float(np.mean(inter)), "ms_tree_build": tb * 1e3,
            "interactions": nint, "interactions_per_s": nint / dt, "flop": st["dFlop"],
            "tflops_interact": st["dFlop"] / (np.mean(inter) * 1e-3) / 1e12}, acc


def grav_leg(gg, torch, r, m, soft, points, steps, warmup, sysOff=None, nsample=10000):
    """Each operating point in FP32 and FP64, with force errors against the
    direct sum on sampled targets."""
    idx = sample(len(m), nsample)
    ad = direct_forces(gg, r, m, soft, idx, sysOff)
    out = {}
    for name, (theta, order) in points.items():
        f32, a32 = time_grav(gg, torch, r, m, soft, gg.PREC_FP32, theta, order, steps, warmup,
                             idx, sysOff)
        f64, a64 = time_grav(gg, torch, r, m, soft, gg.PREC_FP64, theta, order,
                             max(2, steps // 2), 1, idx, sysOff)
        out[name] = {"theta": theta, "iOrder": order, "fp32": f32, "fp64": f64,
                     "speedup_vs_fp64": f64["ms"] / f32["ms"],
                     "force_error": {"targets": len(idx),
                                     "fp32_vs_direct": rel_errors(a32, ad),
                                     "fp64_vs_direct": rel_errors(a64, ad),
                                     "fp32_vs_fp64": rel_errors(a32, a64)}}
    return out


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


def time_periodic(gg, torch, r, m, soft, prec, theta, nReps, steps, warmup, idx):
    """Periodic gravity in the unit box: hexadecapole walk and Ewald (iOrder =
    iEwOrder = 4), dEwCut 2.6, dEwhCut 2.8, nReps replica shells; the same
    call without Ewald gives the Ewald pass's time.  Returns the timings and
    the accelerations of the sampled particles idx."""
    n = len(m)
    pk = gg.Pkd(n, prec=prec)
    pk.put("r", r)
    pk.put("fMass", m)
    pk.put("fSoft", soft)
    pk.set_period([1.0, 1.0, 1.0])
    pk.build_binary(16, 1, theta, 4)
    out = {}
    for ew in (True, False):
        pk.init_accel()
        st = pk.grav_all(4, nReps=nReps, periodic=True, bEwald=ew)
        if ew:
            acc = pk.get("a")[idx]
        for _ in range(warmup):
            st = pk.grav_all(4, nReps=nReps, periodic=True, bEwald=ew)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = pk.grav_all(4, nReps=nReps, periodic=True, bEwald=ew)
        torch.cuda.synchronize()
        out[ew] = ((time.perf_counter() - t0) / steps, st)
    pk.close()
    dt, st = out[True]
    nint = st["dPartSum"] + st["dCellSum"] + st["dSoftSum"]
    return {"ms": dt * 1e3, "ms_without_ewald": out[False][0] * 1e3,
            "interactions": nint, "interactions_per_s": nint / dt, "flop": st["dFlop"]}, acc


def periodic_leg(gg, torch, n, steps, warmup, nsample=2000):
    """The periodic box at the reference's defaults (theta 0.7, nReps 1) and
    at the point whose rms error against the periodic direct sum is under
    1e-4 (two replica shells, theta 0.35: with one shell the hexadecapole
    expansion of the root in the Ewald correction leaves ~5e-4 at any
    theta), FP32 and FP64, each with its force error against the FP64
    Ewald-summed direct sum (gg_direct_ewald) on sampled targets."""
    r, m = periodic_box(n)
    soft = np.full(n, 0.2 / n ** (1 / 3))
    idx = sample(n, nsample)
    t0 = time.perf_counter()
    with gg.Pkd(n, prec=gg.PREC_FP64) as pk:
        pk.put("r", r)
        pk.put("fMass", m)
        pk.put("fSoft", soft)
        pk.set_period([1.0, 1.0, 1.0])
        ad = pk.direct_ewald(idx)[0]
    t_direct = time.perf_counter() - t0
    out = {"direct_ewald_s": t_direct}
    for name, (theta, nReps) in {"reference_config": (0.7, 1), "gate": PERIODIC_GATE}.items():
        p32, a32 = time_periodic(gg, torch, r, m, soft, gg.PREC_FP32, theta, nReps, steps,
                                 warmup, idx)
        p64, a64 = time_periodic(gg, torch, r, m, soft, gg.PREC_FP64, theta, nReps,
                                 max(2, steps // 2), 1, idx)
        out[name] = {"theta": theta, "nReps": nReps, "iOrder": 4, "fp32": p32, "fp64": p64,
                     "speedup_vs_fp64": p64["ms"] / p32["ms"],
                     "force_error": {"targets": len(idx),
                                     "fp32_vs_direct_ewald": rel_errors(a32, ad),
                                     "fp64_vs_direct_ewald": rel_errors(a64, ad),
                                     "fp32_vs_fp64": rel_errors(a32, a64)}}
    return out


def time_ensemble(gg, torch, e, steps, prec, theta, order):
    """Batched ensemble: nsys King models (W0 in [3, 9], N log-uniform in
    [1024, 65536]), eps = 0.005 King radii; one tree per system in one
    forest, bucket size by the autotuner, the gravity step replayed from a
    HIP graph.  Returns times per step (eager and graph)."""
    n = len(e["m"])
    nsys = len(e["off"]) - 1
    pk = gg.Pkd(n, prec=prec)
    pk.put("r", e["r"])
    pk.put("fMass", e["m"])
    pk.put("fSoft", np.full(n, 0.005))
    best, tune = pk.tune_bucket((8, 12, 16, 24, 32), sysOff=e["off"], dCrit=theta, iOrder=order)
    st = pk.grav_all(order)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = pk.grav_all(order)
    torch.cuda.synchronize()
    eager = (time.perf_counter() - t0) / steps
    pk.grav_graph(order)
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


def ensemble_errors(gg, e, theta, order, nsample=4000):
    """Force errors of the forest's FP32 and FP64 forces against each
    system's own direct sum on sampled particles."""
    n = len(e["m"])
    soft = np.full(n, 0.005)
    idx = sample(n, nsample)
    ad = direct_forces(gg, e["r"], e["m"], soft, idx, e["off"])
    acc = {}
    for name, prec in (("fp32", gg.PREC_FP32), ("fp64", gg.PREC_FP64)):
        with gg.Pkd(n, prec=prec) as pk:
            pk.put("r", e["r"])
            pk.put("fMass", e["m"])
            pk.put("fSoft", soft)
            pk.build_forest(e["off"], 16, 1, theta, order)
            pk.init_accel()
            pk.grav_all(order)
            acc[name] = pk.get("a")[idx]
    return {"targets": len(idx), "fp32_vs_direct": rel_errors(acc["fp32"], ad),
            "fp64_vs_direct": rel_errors(acc["fp64"], ad),
            "fp32_vs_fp64": rel_errors(acc["fp32"], acc["fp64"])}


def force_errors(gg, d, soft, bits, theta, order, nsample=10000, seed=1, nBucket=16):
    """rms and max relative gravity-force error of the FP32 and FP64 tree
    forces against the FP64 direct sum (gg_direct_grav) on nsample random
    targets, and of FP32 against FP64 on the same tree."""
    from gasoline_amd import evolve
    n = len(d["m"])
    idx = sample(n, nsample, seed)
    acc = {}
    for name, prec in (("fp32", gg.PREC_FP32), ("fp64", gg.PREC_FP64)):
        with gg.Pkd(n, prec=prec) as pk:
            evolve.start(pk, d, soft, bits)
            pk.build_binary(nBucket, 1, theta, order)
            pk.init_accel()
            pk.grav_all(order)
            acc[name] = pk.get("a")[idx]
            if name == "fp64":
                acc["direct"], _ = pk.direct_grav(idx)
    return {"targets": int(len(idx)), "fp32_vs_direct": rel_errors(acc["fp32"], acc["direct"]),
            "fp64_vs_direct": rel_errors(acc["fp64"], acc["direct"]),
            "fp32_vs_fp64": rel_errors(acc["fp32"], acc["fp64"])}


def time_kdk(gg, torch, d, soft, bits, prec, mode, K, dt, theta, order, steps, warmup, visc=1):
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
                     iOrder=order, hmlp=hm, reorder=True)
    run.forces(dt)
    for _ in range(warmup):
        run.step(dt)
    torch.cuda.synchronize()
    phases = []
    t0 = time.perf_counter()
    for k in range(steps):
        run.step(dt)
        if k % 10 == 9:
            phases.append(pk.timings())
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    st, sm = run.grav_stats, run.stats[-1]
    pk.close()
    return {"ms_per_step": el * 1e3, "mode": mode, "steps": steps,
            "grav_interactions": st["dPartSum"] + st["dCellSum"] + st["dSoftSum"],
            "knn_candidates": int(sm["candidates"]), "knn_extra_passes": int(sm["extra_passes"]),
            "phases_ms_last_call": {k: float(np.mean([p[k] for p in phases]))
                                    for k in phases[0]} if phases else None}


# ------------------------------------------------------------ CPU baseline
def cpu_model():
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def cpu_baseline_sph(n=1 << 20):
    """The oracle's SPH step (kNN + DenDVDX + EOS + SphPressureTermsSym) on a
    periodic box of n particles, pair-evals/s on the host cores."""
    from gasoline_amd import ics
    from oracle import oracle as O
    d = ics.sph_box(n, seed=3)
    smf = O.default_smf()
    t = {}
    t0 = time.perf_counter()
    nbr, fb2, _ = O.knn(d["r"], 64, period=[1, 1, 1])
    lists = O.ball_gather(d["r"], fb2, period=[1, 1, 1])
    t["knn"] = time.perf_counter() - t0
    t1 = time.perf_counter()
    den = O.dendvdx(d["m"], d["v"], d["u"], fb2, lists, smf)
    pr2, c, dt = O.gas_pressure(d["u"], den["rho"], fb2, smf)
    t["dendvdx"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    O.sph_pressure_sym(d["m"], d["v"], d["u"], den["rho"], pr2, c, den["balsara"],
                       den["diff"], dt, fb2, lists, smf)
    t["hydro"] = time.perf_counter() - t1
    el = time.perf_counter() - t0
    pairs = 2 * len(lists[1])
    return pairs / el, el, t


def cpu_baseline_grav(n=1 << 20):
    from gasoline_amd import ics
    from oracle import oracle as O
    r, v, m = ics.plummer(n, seed=1)
    soft = np.full(n, 0.01)
    T = O.Tree(r, m, soft, nBucket=16, dTheta=GATE["theta"], iOrder=GATE["iOrder"])
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
    ap.add_argument("--evrard-steps", type=int, default=100,
                    help="timed steps of the Evrard leg (the configuration runs 100)")
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
    from gasoline_amd import ics

    n = args.n
    K = 64
    pairs = 2 * K * n                      # density pass + force pass
    t64, ph64, o64, tune64 = time_sph(gg, torch, n, gg.PREC_FP64, args.steps, args.warmup,
                                      tune=(32, 24, 16))
    t32, ph32, o32, tune32 = time_sph(gg, torch, n, gg.PREC_FP32, args.steps, args.warmup,
                                      nBucket=tune64["nBucket"])
    tune64["fp32_lanes"] = tune32["lanes"]
    tune64["fp32_lanes_candidate_ms"] = tune32["lanes_candidate_ms"]
    sph_prec = {f: rel_errors(o32[f], o64[f]) for f in o32}
    log("SPH box done")

    r, v, m = ics.plummer(args.ngrav, seed=1)
    grav = grav_leg(gg, torch, r, m, np.full(args.ngrav, 0.01),
                    {"gate": (GATE["theta"], GATE["iOrder"]), "reference_config": (0.7, 2)},
                    args.steps, args.warmup)
    del r, v, m
    log("gravity done")

    per = None
    if args.periodic > 0:
        per = {"workload": "periodic unit box N=%d (half uniform, half in three clumps of "
                           "sigma 0.03), eps 0.2 n^-1/3, hexadecapole walk, Ewald iEwOrder 4, "
                           "dEwCut 2.6, dEwhCut 2.8; gg_grav_all (tree build excluded)"
                           % args.periodic}
        per.update(periodic_leg(gg, torch, args.periodic, args.steps, args.warmup))
        log("periodic gravity done")
    ens = None
    if args.nsys > 0:
        e = ics.king_ensemble(args.nsys, seed=7)
        ens = {"workload": "%d King models (W0 in [3,9], N log-uniform in [1024,65536], %d "
                           "particles), eps 0.005 r0; one forest, one launch; bucket size "
                           "autotuned; HIP-graph replay" % (args.nsys, len(e["m"]))}
        for name, (theta, order) in (("gate", (GATE_ENSEMBLE["theta"],
                                               GATE_ENSEMBLE["iOrder"])),
                                     ("reference_config", (0.6, 2))):
            e32 = time_ensemble(gg, torch, e, args.steps, gg.PREC_FP32, theta, order)
            e64 = time_ensemble(gg, torch, e, max(2, args.steps // 2), gg.PREC_FP64, theta, order)
            ens[name] = {"theta": theta, "iOrder": order, "fp32": e32, "fp64": e64,
                         "speedup_vs_fp64": e64["ms_graph"] / e32["ms_graph"],
                         "force_error": ensemble_errors(gg, e, theta, order)}
        del e
        log("ensemble done")
    evr = None
    if args.nevrard > 0:
        d = ics.evrard(args.nevrard, seed=5)
        bits = None                    # every particle active gas (evolve.start)
        soft = np.full(args.nevrard, 0.004)
        ks = args.evrard_steps
        th, od = GATE["theta"], GATE["iOrder"]
        fast = time_kdk(gg, torch, d, soft, bits, gg.PREC_FP32, "mlp", 64, 1e-3, th, od, ks, 3)
        log("Evrard FP32 done")
        plain = time_kdk(gg, torch, d, soft, bits, gg.PREC_FP64, "cold", 64, 1e-3, th, od, ks, 1)
        log("Evrard FP64 done")
        evr = {"workload": "Evrard collapse N=%d (rho ~ 1/r, M=R=G=1, u=0.05, gamma 5/3), "
                           "eps 0.004, theta %.2f, iOrder %d, bucket 16, nSmooth 64, dt 0.001 "
                           "(about the central Courant limit); %d timed KDK steps (gravity tree "
                           "build + walk/interactions, SPH tree, kNN, DenDVDX, EOS, "
                           "SphPressureTermsSym)" % (args.nevrard, th, od, ks),
               "fp32_mlp": fast, "fp64_plain": plain,
               "speedup_vs_fp64_plain": plain["ms_per_step"] / fast["ms_per_step"],
               "force_error": force_errors(gg, d, soft, bits, th, od)}
        del d
    disk = None
    if args.disk:
        d = ics.disk_galaxy(seed=11)
        gas_bits = gg.TYPE_ACTIVE | gg.TYPE_TREEACTIVE | gg.TYPE_SMOOTHACTIVE | gg.TYPE_GAS
        bits = np.where(d["gas"], gas_bits, 3).astype(np.uint32)
        ks = max(2, args.steps // 2)
        th, od = GATE["theta"], GATE["iOrder"]
        f32 = time_kdk(gg, torch, d, d["soft"], bits, gg.PREC_FP32, "prev", 32, 5e-4, th, od,
                       ks, 1)
        f64 = time_kdk(gg, torch, d, d["soft"], bits, gg.PREC_FP64, "prev", 32, 5e-4, th, od,
                       ks, 1)
        disk = {"workload": "isolated disk galaxy: Hernquist halo 4194304 DM (1e12 Msun, a 30 "
                            "kpc, cut at 10a) + exponential gas disk 1048576 (R_d 3 kpc, sech^2 "
                            "z0 0.3 kpc, min Toomre Q 1.5, T 1e4 K, mu 1.22), eps 0.1/0.05 kpc, "
                            "theta %.2f, iOrder %d, nSmooth 32, dt 5e-4 (%s)" % (th, od,
                                                                                d["units"]),
                "fp32": f32, "fp64": f64, "speedup_vs_fp64": f64["ms_per_step"] / f32["ms_per_step"],
                "force_error": force_errors(gg, d, d["soft"], bits, th, od)}
        del d
        log("disk galaxy done")
    # roofline of each SPH kernel: algorithmic bytes per launch over its
    # HIP-event duration on the handle's stream (the headline FP64 step; the
    # kNN is the same kernel in both precisions); the dominant one is reported
    L64, L32 = tune64.get("lanes", 4), tune64.get("fp32_lanes", 4)   # the tuners' lanes
    kern = {
        "k_knn_bucket": (ph64["knn"], n * (KNN_BYTES_PER_QUERY + KNN_BYTES_PER_NEIGHBOUR * K)),
        "k_dendvdx<double,%d>" % L64: (ph64["density"], n * K * DENS_BYTES_PER_PAIR),
        "k_hydro2<double,%d>" % L64: (ph64["hydro"], n * K * HYDRO_BYTES_PER_PAIR),
    }
    kern32 = {
        "k_dendvdx<float,%d>" % L32: (ph32["density"], n * K * DENS_BYTES_PER_PAIR),
        "k_hydro2<float,%d>" % L32: (ph32["hydro"], n * K * HYDRO_BYTES_PER_PAIR),
    }
    traffic = {}
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f)

    def roof(name, table):
        ms, by = table[name]
        a = by / (ms * 1e-3) / 1e9
        t = traffic.get(name)
        return {"bound": "hbm", "kernel": name, "achieved": a, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": a / HBM_PEAK_GBS, "ms": ms, "algorithmic_bytes": by,
                "traffic": (t["fetch_bytes"] + t["write_bytes"]) if t else None}

    def gtraffic(name):
        t = traffic.get(name)
        return (t["fetch_bytes"] + t["write_bytes"]) if t else None

    dom = max(kern, key=lambda k: kern[k][0])
    knn_strict_gbs = n * (KNN_BYTES_PER_QUERY + 4 * K) / (ph64["knn"] * 1e-3) / 1e9
    gate = grav["gate"]
    line = {
        "metric": METRIC,
        "value": pairs / t64,
        "unit": "SPH pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "SPH adiabatic box: N=%d periodic unit cube, nSmooth=64, M4, "
                        "alpha=1 beta=2, one DenDVDX + one SphPressureTermsSym pass "
                        "(tree build, kNN, EOS and reverse lists inside the step), FP64 pair "
                        "arithmetic" % n,
            "n": n, "nSmooth": K, "pairs_per_step": pairs,
            "gravity_workload": "Plummer N=%d, eps=0.01, bucket 16; theta %.2f iOrder %d (rms "
                                "force error <= 1e-4) and the reference configuration theta 0.7 "
                                "iOrder 2; fused walk+interact (tree build excluded)"
                                % (args.ngrav, GATE["theta"], GATE["iOrder"]),
        },
        "fp32_value": pairs / t32,
        "fp32_ms_per_step": t32 * 1e3,
        "speedup_fp32_vs_fp64": t64 / t32,
        "sph_tuner": tune64,
        "fp32_vs_fp64": sph_prec,
        "phases_ms_fp64": ph64,
        "phases_ms_fp32": ph32,
        "grav_interactions_per_s": gate["fp64"]["interactions_per_s"],
        "grav_interactions_per_s_fp32": gate["fp32"]["interactions_per_s"],
        "grav_speedup_fp32_vs_fp64": gate["speedup_vs_fp64"],
        "grav_force_error": gate["force_error"],
        "gravity": grav,
        "roofline": roof(dom, kern),
        "rooflines_other": [roof(k, kern) for k in kern if k != dom]
                           + [roof(k, kern32) for k in kern32],
        "knn_output_only_gbs": knn_strict_gbs,
        "roofline_gravity": {
            "bound": "valu", "kernel": "k_walk_interact<float,%d> (walk included)" % GATE["iOrder"],
            "achieved": gate["fp32"]["tflops_interact"], "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": gate["fp32"]["tflops_interact"] / FP32_PEAK_TFLOPS,
            "frac_nonpacked": gate["fp32"]["tflops_interact"] / FP32_NONPACKED_TFLOPS,
            "traffic": gtraffic("k_walk_interact<float,%d>" % GATE["iOrder"]),
        },
        "roofline_gravity_fp64": {
            "bound": "valu", "kernel": "k_walk_interact<double,%d> (walk included)"
                                       % GATE["iOrder"],
            "achieved": gate["fp64"]["tflops_interact"], "peak": 78.6, "unit": "TFLOP/s",
            "frac": gate["fp64"]["tflops_interact"] / 78.6,
            "traffic": gtraffic("k_walk_interact<double,%d>" % GATE["iOrder"]),
        },
    }
    if ens:
        line["ensemble"] = ens
    if per:
        line["periodic_gravity"] = per
    if evr:
        line["evrard"] = evr
    if disk:
        line["disk_galaxy"] = disk
    if not args.no_cpu:
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        v_sph, el_sph, t_sph = cpu_baseline_sph()
        log("CPU baseline (SPH) done")
        v_gr, el_gr = cpu_baseline_grav()
        log("CPU baseline (gravity) done")
        line["cpu_baseline"] = {
            "value": v_sph, "unit": "SPH pair-evals/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": "oracle SPH step on the benchmark's box, 1048576 particles (%.1f s: kNN "
                      "%.1f s, DenDVDX + EOS %.1f s, force pass %.1f s), all phases on %d "
                      "OpenMP threads" % (el_sph, t_sph["knn"], t_sph["dendvdx"],
                                          t_sph["hydro"], cores),
            "grav_interactions_per_s": v_gr,
            "grav_sample": "oracle gravity (walk + interactions, %d OpenMP threads over "
                           "buckets) on Plummer N=1048576, theta %.2f, iOrder %d, bucket 16 "
                           "(%.1f s)" % (cores, GATE["theta"], GATE["iOrder"], el_gr),
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
