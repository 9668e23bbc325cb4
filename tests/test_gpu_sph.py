"""GPU parity of the SPH path (gg_build_local, gg_smooth, gg_gas_pressure,
gg_resmooth) against the CPU oracle, through the C ABI.

Bars: neighbour sets and fBall2 bit-exact, with ties at the K-th distance
broken by the smaller caller index in both (a list's order is not compared:
the library writes it in ball order, the oracle by (d2, caller index), the
reference in its priority-queue order); FP64 operator outputs within 1e-12
relative; FP32 within 1e-5 relative to the field's scale (1e-4 for the
second-derivative quantities dvds and for single values near zero), and per
particle: FP64 within 1e-10 of each particle's own value, FP32 within 1e-5
rms and 1e-3 max (floored at 1% of the field's rms)."""
import numpy as np
import pytest

from gasoline_amd import evolve, ics
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DENS = {"rho": "fDensity", "divv": "divv", "dvds": "dvds", "fDivv_t": "fDivv_t",
        "balsara": "BalsaraSwitch", "diff": "diff"}


def put_gas(pk, d):
    pk.put("r", d["r"])
    pk.put("fMass", d["m"])
    pk.put("vPred", d["v"])
    pk.put("uPred", d["u"])


def scaled(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def per_particle(got, ref):
    """rms and max of |got - ref| / max(|ref|, 0.01 rms|ref|), per particle
    (vectors by their norm)."""
    g = np.asarray(got, float).reshape(len(got), -1)
    r = np.asarray(ref, float).reshape(len(ref), -1)
    mag = np.linalg.norm(r, axis=1)
    floor = 0.01 * np.sqrt(np.mean(mag * mag))
    e = np.linalg.norm(g - r, axis=1) / np.maximum(mag, max(floor, 1e-300))
    return float(np.sqrt(np.mean(e * e))), float(e.max())


def oracle_step(d, period, K, smf, fb2=None):
    nbr, ofb2, ties = O.knn(d["r"], K, period=period)
    if fb2 is None:
        fb2 = ofb2
    lists = O.ball_gather(d["r"], fb2, period=period)
    den = O.dendvdx(d["m"], d["v"], d["u"], ofb2, O.ball_gather(d["r"], ofb2, period=period), smf)
    pr2, c, dt = O.gas_pressure(d["u"], den["rho"], fb2, smf)
    hyd = O.sph_pressure_sym(d["m"], d["v"], d["u"], den["rho"], pr2, c, den["balsara"],
                             den["diff"], dt, fb2, lists, smf)
    return dict(nbr=nbr, fb2=ofb2, ties=ties, den=den, pr2=pr2, c=c, dt=dt, hyd=hyd)


@pytest.fixture(scope="module")
def box64k():
    d = ics.sph_box(1 << 16, seed=2)
    return d, oracle_step(d, [1, 1, 1], 64, O.default_smf())


@pytest.fixture(scope="module")
def evrard32k():
    d = ics.evrard(1 << 15, seed=6)
    d["v"] = ics._f32(np.random.default_rng(1).normal(0, 0.05, (len(d["m"]), 3)))
    return d, oracle_step(d, None, 32, O.default_smf(iViscosityLimiter=1))


@pytest.mark.parametrize("case", ["box64k", "evrard32k"])
def test_knn_bitexact(gpu, case, request):
    d, ref = request.getfixturevalue(case)
    periodic = case == "box64k"
    K = 64 if periodic else 32
    assert ref["ties"] == 0
    with gpu.Pkd(len(d["m"])) as pk:
        put_gas(pk, d)
        pk.build_local(period=[1, 1, 1] if periodic else None)
        ns = pk.smooth(gpu.Smf.default(), K, periodic)
        nbr, fb2 = pk.smooth_lists(K), pk.get("fBall2")
    assert ns == len(d["m"])
    assert np.array_equal(fb2, ref["fb2"])
    assert np.array_equal(np.sort(nbr, axis=1), np.sort(ref["nbr"], axis=1))


def test_knn_lattice_ties(gpu):
    # periodic 16^3 lattice, K = 32: 27 particles lie within sqrt(3) spacings
    # and 6 at two spacings, so every list takes 5 of 6 equal distances and
    # the caller index decides which
    g = np.arange(16) / 16.0
    r = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    r = r[np.random.default_rng(4).permutation(len(r))]
    nbr_o, fb2_o, ties_o = O.knn(r, 32, period=[1, 1, 1])
    assert ties_o == len(r)
    with gpu.Pkd(len(r)) as pk:
        pk.put("r", r)
        pk.put("fMass", np.full(len(r), 1.0 / len(r)))
        pk.build_local(period=[1, 1, 1])
        pk.smooth(gpu.Smf.default(), 32, True)
        nbr, fb2 = pk.smooth_lists(32), pk.get("fBall2")
        st = pk.smooth_stats()
    assert st["ties"] == len(r)
    assert np.array_equal(fb2, fb2_o)
    assert np.array_equal(np.sort(nbr, axis=1), np.sort(nbr_o, axis=1))


def test_gas_subset_of_a_mixed_store(gpu):
    # a disk galaxy's store holds halo and gas particles: the SPH path works on
    # the TYPE_GAS ones only, exactly as on a store of the gas alone
    d = ics.disk_galaxy(1 << 14, 1 << 13, seed=5)
    g = d["gas"]
    smf = gpu.Smf.default(iViscosityLimiter=1)
    res = []
    gas_bits = gpu.TYPE_ACTIVE | gpu.TYPE_TREEACTIVE | gpu.TYPE_SMOOTHACTIVE | gpu.TYPE_GAS
    for sel, bits in ((np.ones(len(g), bool), np.where(g, gas_bits, 3)),
                      (g, np.full(g.sum(), gas_bits))):
        with gpu.Pkd(int(sel.sum())) as pk:
            for f, k in (("r", "r"), ("v", "v"), ("vPred", "v"), ("fMass", "m"), ("u", "u"),
                         ("uPred", "u")):
                pk.put(f, d[k][sel])
            pk.put("iActive", bits.astype(np.uint32))
            pk.build_local()
            pk.smooth(smf, 32, False)
            pk.gas_pressure(smf)
            pk.resmooth(smf, 32, False)
            gi = np.nonzero(g[sel])[0]
            res.append({f: pk.get(f)[gi] for f in ("fBall2", "fDensity", "divv", "a", "uDotPdV",
                                                    "uDotAV")})
            st = pk.smooth_stats()
            res[-1]["work"] = np.array([st["queries"], st["candidates"]])
            assert st["queries"] == g.sum() and st["candidates"] > 32 * g.sum()
            if sel.all():
                halo = ~g
                assert np.all(pk.get("fDensity")[halo] == 0) and np.all(pk.get("a")[halo] == 0)
                # reordering puts the gas first, in the structure's order, and
                # the halo after it in its previous order
                perm = pk.reorder()
                ng = int(g.sum())
                assert np.all(g[perm[:ng]]) and np.array_equal(perm[ng:], np.nonzero(halo)[0])
                assert np.array_equal(pk.get("fDensity")[:ng], res[-1]["fDensity"][
                    np.searchsorted(gi, perm[:ng])])
    for f in res[0]:
        assert np.array_equal(res[0][f], res[1][f]), f


def test_knn_isolated_particles(gpu):
    # particles far outside a uniform ball: their K nearest lie in a thin
    # shell on the ball's near side (the 64th and 256th squared distances
    # differ by about 1.4%), which the search must still bracket
    rng = np.random.default_rng(8)
    n = 20000
    v = rng.normal(size=(n, 3))
    r = (rng.uniform(size=n) ** (1 / 3))[:, None] * v / np.linalg.norm(v, axis=1)[:, None]
    far = np.array([[9.0, 0, 0], [0, -12.0, 0.5], [3.0, 4.0, -7.0]])
    r = ics._f32(np.concatenate([r, far]))
    nbr_o, fb2_o, _ = O.knn(r, 64)
    with gpu.Pkd(len(r)) as pk:
        pk.put("r", r)
        pk.put("fMass", np.full(len(r), 1.0 / len(r)))
        for slack in (0.0, 1.15):
            pk.set_seed_slack(slack)
            pk.build_local()
            pk.smooth(gpu.Smf.default(), 64, False)
            assert np.array_equal(pk.get("fBall2"), fb2_o)
            assert np.array_equal(np.sort(pk.smooth_lists(64), 1), np.sort(nbr_o, 1))


def test_reorder_moves_particles_and_keeps_results(gpu):
    # gg_reorder permutes the store into the SPH structure's order: every
    # field and the h-proposal history move with their particles, the SPH
    # results are the same particle by particle, and the gravity tree must be
    # built again
    d = ics.evrard(20000, seed=9)
    d["v"] = -0.3 * d["r"]
    n = len(d["m"])
    smf = gpu.Smf.default(iViscosityLimiter=1)
    out = []
    for reorder in (False, True):
        with gpu.Pkd(n) as pk:
            evolve.start(pk, d, 0.01)
            pk.build_local()
            pk.smooth(smf, 32, False)
            pk.hmlp_init(2)
            pk.hmlp_record(0.05)
            pk.build_binary(16, 1, 0.7, 2)
            perm = np.arange(n)
            if reorder:
                pk.build_local()
                perm = pk.reorder()
                assert np.array_equal(np.sort(perm), np.arange(n)) and not np.all(perm == np.arange(n))
                with pytest.raises(gpu.GGError):
                    pk.grav_all(2)
                assert np.array_equal(pk.get("r"), d["r"][perm])
            pk.hmlp_predict()
            res = {"hint": pk.get("fBall2")}
            pk.build_local()
            pk.smooth(smf, 32, False)
            pk.gas_pressure(smf)
            pk.resmooth(smf, 32, False)
            for f in ("fBall2", "fDensity", "divv", "curlv", "a", "uDotPdV", "uDotAV", "fMass"):
                res[f] = pk.get(f)
            pk.build_binary(16, 1, 0.7, 2)
            pk.init_accel()
            pk.grav_all(2)
            out.append((perm, res))
    perm = out[1][0]
    for f in out[0][1]:
        assert np.array_equal(out[1][1][f], out[0][1][f][perm]), f


def test_knn_cold_switch(gpu):
    # slack 0 ignores a stored fBall2: the search, its work counters and its
    # result equal those of a store with no hint at all
    d = ics.evrard(20000, seed=6)
    out = []
    for fb2, slack in ((np.zeros(20000), 1.15), (np.full(20000, 1e-4), 0.0)):
        with gpu.Pkd(20000) as pk:
            pk.put("r", d["r"])
            pk.put("fMass", d["m"])
            pk.put("fBall2", fb2)
            pk.set_seed_slack(slack)
            pk.build_local()
            pk.smooth(gpu.Smf.default(), 32, False)
            st = pk.smooth_stats()
            out.append((pk.get("fBall2"), np.sort(pk.smooth_lists(32), 1), st["candidates"],
                        st["extra_passes"]))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][2:] == out[1][2:]


@pytest.mark.parametrize("prec", ["FP64", "FP32"])
@pytest.mark.parametrize("case", ["box64k", "evrard32k"])
def test_density_and_hydro_vs_oracle(gpu, case, prec, request):
    d, ref = request.getfixturevalue(case)
    periodic = case == "box64k"
    K = 64 if periodic else 32
    smf = gpu.Smf.default(iViscosityLimiter=0 if periodic else 1)
    p = gpu.PREC_FP64 if prec == "FP64" else gpu.PREC_FP32
    tol = 1e-12 if prec == "FP64" else 1e-5
    with gpu.Pkd(len(d["m"]), prec=p) as pk:
        put_gas(pk, d)
        pk.build_local(period=[1, 1, 1] if periodic else None)
        pk.smooth(smf, K, periodic)
        got = {k: pk.get(v) for k, v in DENS.items()}
        curl = pk.get("curlv")
        pk.gas_pressure(smf)
        pr2, c, dt = pk.get("PoverRho2"), pk.get("c"), pk.get("dt")
        pk.resmooth(smf, K, periodic)
        h = {k: pk.get(k) for k in ("a", "uDotPdV", "uDotAV", "uDotDiff", "mumax", "dtNew")}
    den = ref["den"]
    for k in DENS:
        t = 1e-4 if (k == "dvds" and prec == "FP32") else tol
        assert scaled(got[k], den[k]) < t, k
    assert scaled(curl, den["curlv"]) < tol
    assert scaled(pr2, ref["pr2"]) < tol and scaled(c, ref["c"]) < tol
    assert scaled(dt, ref["dt"]) < tol
    hy = ref["hyd"]
    assert scaled(h["a"], hy["acc"]) < tol
    for k in ("uDotPdV", "uDotAV", "mumax", "dtNew"):
        assert scaled(h[k], hy[k]) < tol, k
    assert (h["uDotDiff"] == 0).all()
    # per particle as well (relative to each particle's own value, floored
    # at 1% of the field's rms so that values crossing zero do not divide
    # by nothing)
    for name, g, r in (("rho", got["rho"], den["rho"]), ("a", h["a"], hy["acc"]),
                       ("uDotPdV", h["uDotPdV"], hy["uDotPdV"]),
                       ("uDotAV", h["uDotAV"], hy["uDotAV"])):
        rms, mx = per_particle(g, r)
        if prec == "FP64":
            assert mx < 1e-10, (name, rms, mx)
        else:
            assert rms < 1e-5 and mx < 1e-3, (name, rms, mx)


@pytest.mark.parametrize("prec", ["FP64", "FP32"])
@pytest.mark.parametrize("case", ["box64k", "evrard32k"])
def test_density_operators(gpu, case, prec, request):
    # Density equals DenDVDX's density bit for bit and marks no neighbour of
    # an active particle; DensitySym against the oracle's serial scatter
    d, ref = request.getfixturevalue(case)
    periodic = case == "box64k"
    per = [1, 1, 1] if periodic else None
    K = 64 if periodic else 32
    smf = gpu.Smf.default()
    p = gpu.PREC_FP64 if prec == "FP64" else gpu.PREC_FP32
    tol = 1e-12 if prec == "FP64" else 1e-5
    n = len(d["m"])
    assert ref["ties"] == 0
    sym_ref = O.density_sym(d["m"], ref["fb2"], O.ball_gather(d["r"], ref["fb2"], period=per))
    bits = np.full(n, 7, np.uint32)
    with gpu.Pkd(n, prec=p) as pk:
        put_gas(pk, d)
        pk.put("iActive", bits)
        pk.build_local(period=per)
        pk.smooth(smf, K, periodic)
        rho_dv = pk.get("fDensity")
        pk.put("iActive", bits)               # drops the structure: build it again
        pk.put("fDensity", np.zeros(n))
        pk.put("fBall2", np.zeros(n))         # the same cold search: the same list order
        pk.build_local(period=per)
        assert pk.smooth(smf, K, periodic, iSmoothType=gpu.SMX_DENSITY) == n
        rho = pk.get("fDensity")
        flags = pk.get("iActive")
        fb2 = pk.get("fBall2")
        pk.put("fDensity", np.full(n, 5.0))      # overwritten: every particle is smoothed
        pk.smooth(smf, K, periodic, iSmoothType=gpu.SMX_DENSITY, symmetric=True)
        sym = pk.get("fDensity")
    assert np.array_equal(rho, rho_dv)
    assert np.array_equal(fb2, ref["fb2"])
    assert (flags & 32 == 0).all() and (flags & 8).all()
    assert scaled(sym, sym_ref) < tol
    # each pair counted once per end: the mass-weighted sums agree
    assert np.dot(d["m"], sym) == pytest.approx(np.dot(d["m"], rho), rel=10 * tol)


def test_density_sym_partial_smooth_active(gpu):
    # particles that are not smooth-active keep their density plus what
    # their smooth-active neighbours add; their fBall2 is left alone
    d = ics.sph_box(1 << 14, seed=33)
    n = len(d["m"])
    rng = np.random.default_rng(6)
    sact = rng.uniform(size=n) < 0.05
    m = d["m"] * rng.uniform(0.5, 1.5, n)
    rho0 = rng.uniform(1, 2, n)
    fb0 = rng.uniform(1e-3, 2e-3, n)
    per = [1, 1, 1]
    nbr, fb2, ties = O.knn(d["r"], 64, period=per)
    assert ties == 0
    ref = O.density_sym(m, fb2, O.ball_gather(d["r"], fb2, period=per),
                        smooth_active=sact.astype(np.uint8), rho=rho0)
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        pk.put("r", d["r"])
        pk.put("fMass", m)
        pk.put("iActive", np.where(sact, 7, 3).astype(np.uint32))
        pk.put("fDensity", rho0)
        pk.put("fBall2", fb0)
        pk.build_local(period=per)
        ns = pk.smooth(gpu.Smf.default(), 64, True, iSmoothType=gpu.SMX_DENSITY, symmetric=True)
        got, gfb2, flags = pk.get("fDensity"), pk.get("fBall2"), pk.get("iActive")
    assert ns == sact.sum()
    assert scaled(got, ref) < 1e-12
    assert np.array_equal(gfb2[sact], fb2[sact]) and np.array_equal(gfb2[~sact], fb0[~sact])
    assert np.array_equal((flags & 8) != 0, sact)


def test_resmooth_ball_gather_path(gpu, box64k):
    # enlarged radii force the explicit ball gather (lists longer than nSmooth)
    d, ref = box64k
    smf = gpu.Smf.default()
    fb2 = ref["fb2"] * 1.21
    oref = oracle_step(d, [1, 1, 1], 64, O.default_smf(), fb2=fb2)
    with gpu.Pkd(len(d["m"]), prec=gpu.PREC_FP64) as pk:
        put_gas(pk, d)
        pk.build_local(period=[1, 1, 1])
        pk.smooth(smf, 64, True)
        pk.put("fBall2", fb2)
        pk.gas_pressure(smf)
        pk.resmooth(smf, 64, True)
        acc, pdv = pk.get("a"), pk.get("uDotPdV")
    assert scaled(acc, oref["hyd"]["acc"]) < 1e-12
    assert scaled(pdv, oref["hyd"]["uDotPdV"]) < 1e-12


def test_resmooth_dendvdx(gpu, box64k):
    # at the stored radii the kNN lists are the ball lists: the re-smooth
    # repeats the smooth bit for bit; at enlarged radii it gathers the balls
    d, ref = box64k
    smf = gpu.Smf.default()
    fb2 = ref["fb2"] * 1.21
    oden = O.dendvdx(d["m"], d["v"], d["u"], fb2, O.ball_gather(d["r"], fb2, period=[1, 1, 1]),
                     O.default_smf())
    with gpu.Pkd(len(d["m"]), prec=gpu.PREC_FP64) as pk:
        put_gas(pk, d)
        pk.build_local(period=[1, 1, 1])
        pk.smooth(smf, 64, True, symmetric=True)
        first = {k: pk.get(v) for k, v in DENS.items()}
        pk.resmooth(smf, 64, True, iSmoothType=gpu.SMX_DENDVDX)
        again = {k: pk.get(v) for k, v in DENS.items()}
        pk.put("fBall2", fb2)
        pk.resmooth(smf, 64, True, iSmoothType=gpu.SMX_DENDVDX)
        wide = {k: pk.get(v) for k, v in DENS.items()}
        curl = pk.get("curlv")
    for k in DENS:
        assert np.array_equal(again[k], first[k]), k
        assert scaled(wide[k], oden[k]) < 1e-12, k
    assert scaled(curl, oden["curlv"]) < 1e-12


def test_partial_activity(gpu):
    d = ics.sph_box(1 << 14, seed=31)
    n = len(d["m"])
    rng = np.random.default_rng(5)
    active = rng.uniform(size=n) < 0.4
    bits = np.where(active, 7, 6).astype(np.uint32)     # all smooth-active
    smf = O.default_smf()
    per = [1, 1, 1]
    nbr, fb2, _ = O.knn(d["r"], 64, period=per)
    lists = O.ball_gather(d["r"], fb2, period=per)
    den = O.dendvdx(d["m"], d["v"], d["u"], fb2, lists, smf, active=active.astype(np.uint8))
    pr2, c, dt = O.gas_pressure(d["u"], den["rho"], fb2, smf)
    hyd = O.sph_pressure_sym(d["m"], d["v"], d["u"], den["rho"], pr2, c, den["balsara"],
                             den["diff"], dt, fb2, lists, smf, active=active.astype(np.uint8))
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        put_gas(pk, d)
        pk.put("iActive", bits)
        pk.build_local(period=per)
        pk.smooth(gpu.Smf.default(), 64, True)
        flags = pk.get("iActive")
        pk.gas_pressure(gpu.Smf.default())
        pk.resmooth(gpu.Smf.default(), 64, True)
        acc, mumax, dtnew = pk.get("a"), pk.get("mumax"), pk.get("dtNew")
        flags2 = pk.get("iActive")
    assert np.array_equal((flags & 32) != 0, den["nbrOfActive"] != 0)
    assert np.array_equal(flags2 & 32, flags & 32)   # SphPressureTerms leaves the mark
    assert (flags & 8).all()                     # TYPE_SMOOTHDONE set
    assert (acc[~active] == 0).all()
    assert scaled(acc[active], hyd["acc"][active]) < 1e-12
    assert scaled(mumax, hyd["mumax"]) < 1e-12
    assert scaled(dtnew, hyd["dtNew"]) < 1e-12


def test_momentum_conservation_fp32(gpu, box64k):
    d, _ = box64k
    with gpu.Pkd(len(d["m"]), prec=gpu.PREC_FP32) as pk:
        put_gas(pk, d)
        pk.build_local(period=[1, 1, 1])
        smf = gpu.Smf.default()
        pk.smooth(smf, 64, True)
        pk.gas_pressure(smf)
        pk.resmooth(smf, 64, True)
        acc = pk.get("a")
    p = (d["m"][:, None] * acc).sum(0)
    assert np.abs(p).max() < 1e-6 * (d["m"][:, None] * np.abs(acc)).sum()


def test_sph_bucket_tuner(gpu, box64k):
    # gg_tune_sph times each candidate bucket size, keeps the fastest and
    # leaves the structure and the DenDVDX outputs as a plain build at that
    # size gives them; the neighbour sets do not depend on the bucket size
    d, ref = box64k
    smf = gpu.Smf.default()
    n = len(d["m"])
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        put_gas(pk, d)
        best, ms = pk.tune_sph(smf, 64, [1.0] * 3, cand=(32, 16, 8))
        assert set(ms) == {32, 16, 8} and all(v > 0 for v in ms.values())
        assert ms[best] == min(ms.values())
        tuned = {f: pk.get(f) for f in ("fBall2", "fDensity", "divv")}
        lists = pk.smooth_lists(64)
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        put_gas(pk, d)
        pk.build_local(period=[1.0] * 3, nBucket=best)
        pk.smooth(smf, 64, True)
        for f, v in tuned.items():
            assert np.array_equal(pk.get(f), v), f
        assert np.array_equal(np.sort(lists, axis=1), np.sort(pk.smooth_lists(64), axis=1))
    assert np.array_equal(tuned["fBall2"], ref["fb2"])
    assert np.array_equal(np.sort(lists, axis=1), np.sort(ref["nbr"], axis=1))
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        put_gas(pk, d)
        with pytest.raises(gpu.GGError, match="outside 1..32"):
            pk.tune_sph(smf, 64, [1.0] * 3, cand=(32, 40))


def _sph_step(pk, smf, lanes=None):
    if lanes is not None:
        pk.set_sph_lanes(lanes)
    pk.smooth(smf, 64, True)
    pk.gas_pressure(smf)
    pk.resmooth(smf, 64, True)


SPH_OUT = ("fBall2", "fDensity", "divv", "a", "uDotPdV", "uDotAV", "mumax", "dtNew", "dt")


def test_sph_lanes_tuner(gpu, box64k):
    # gg_tune_sph_lanes times DenDVDX + the force pass at 4, 8 and 16 lanes
    # per particle, each run from the caller's state (a is accumulated by the
    # force pass, dt lowered by the EOS), and leaves every output as one
    # plain step at the chosen setting
    d, _ = box64k
    smf = gpu.Smf.default()
    n = len(d["m"])
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        put_gas(pk, d)
        pk.put("a", np.full((n, 3), 0.25))
        pk.put("dt", np.full(n, 1e-3))
        pk.build_local(period=[1.0] * 3)
        best, ms = pk.tune_sph_lanes(smf, 64, True, cand=(4, 8, 16))
        assert set(ms) == {4, 8, 16} and all(v > 0 for v in ms.values())
        assert ms[best] == min(ms.values())
        tuned = {f: pk.get(f) for f in SPH_OUT}
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        put_gas(pk, d)
        pk.put("a", np.full((n, 3), 0.25))
        pk.put("dt", np.full(n, 1e-3))
        pk.build_local(period=[1.0] * 3)
        _sph_step(pk, smf, best)
        for f in SPH_OUT:
            assert np.array_equal(pk.get(f), tuned[f]), f
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        put_gas(pk, d)
        with pytest.raises(gpu.GGError, match="before gg_build_local"):
            pk.tune_sph_lanes(smf, 64, True)
        pk.build_local(period=[1.0] * 3)
        with pytest.raises(gpu.GGError, match="not 4, 8 or 16"):
            pk.tune_sph_lanes(smf, 64, True, cand=(8, 6))
        with pytest.raises(gpu.GGError, match="not 4, 8 or 16"):
            pk.set_sph_lanes(32)


@pytest.mark.parametrize("prec", ["FP64", "FP32"])
def test_sph_lanes_agree(gpu, box64k, prec):
    # the lanes per particle change only the order of each particle's sums:
    # the same neighbour sets, and the outputs at 4 and 16 lanes equal those
    # at 8 to the rounding of the precision's sums (per particle, against the
    # field's own scale)
    d, _ = box64k
    smf = gpu.Smf.default()
    n = len(d["m"])
    p = gpu.PREC_FP64 if prec == "FP64" else gpu.PREC_FP32
    out = {}
    for lanes in (8, 4, 16):
        with gpu.Pkd(n, prec=p) as pk:
            put_gas(pk, d)
            pk.build_local(period=[1.0] * 3)
            _sph_step(pk, smf, lanes)
            out[lanes] = {f: pk.get(f) for f in SPH_OUT}
            out[lanes]["lists"] = np.sort(pk.smooth_lists(64), axis=1)
    tol = 1e-12 if prec == "FP64" else 1e-5
    for lanes in (4, 16):
        assert np.array_equal(out[lanes]["lists"], out[8]["lists"])
        assert np.array_equal(out[lanes]["fBall2"], out[8]["fBall2"])
        for f in ("fDensity", "divv", "a", "uDotPdV", "uDotAV"):
            got, ref = out[lanes][f], out[8][f]
            scale = np.abs(ref).max()
            assert np.abs(got - ref).max() <= tol * scale, (lanes, f)
