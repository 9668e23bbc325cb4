"""GPU parity of the gravity path (gg_build_binary, gg_grav_all) against the
CPU oracle, through the C ABI.

Bars: tree bit-exact; interaction lists bit-exact as sets per bucket and
list kind (the library's walk is breadth-first, so members come in another
order than the reference's preorder walk); counters exact; FP64
accelerations/potentials within 1e-8 relative (the oracle uses gasoline's
v_sqrt1, itself accurate to 2e-10); FP32 within 1e-5 rms / 1e-4 max relative
to the FP64 oracle."""
import numpy as np
import pytest

from gasoline_amd import ics
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def setup_pkd(gg, r, m, soft, prec, iactive=None):
    pk = gg.Pkd(len(m), prec=prec)
    pk.put("r", r)
    pk.put("fMass", m)
    pk.put("fSoft", soft)
    if iactive is not None:
        pk.put("iActive", iactive)
    return pk


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


@pytest.fixture(scope="module")
def plummer32k():
    r, v, m = ics.plummer(32768, seed=1)
    soft = np.full(len(m), 0.01)
    T = O.Tree(r, m, soft, nBucket=16, dTheta=0.7, iOrder=2)
    return r, m, soft, T, T.grav_all()


def test_tree_bitexact(gpu, plummer32k):
    r, m, soft, T, _ = plummer32k
    with setup_pkd(gpu, r, m, soft, gpu.PREC_FP64) as pk:
        pk.build_binary(16, 1, 0.7, 2)
        root, perm, ints, dbls = pk.tree_info()
    oi, od = T.nodes()
    assert root == T.root
    assert np.array_equal(perm, T.perm())
    assert np.array_equal(ints[:, :5], oi[:, :5])
    # mass, softening, c.o.m., fOpen2, Bmax: bit for bit
    assert np.array_equal(dbls[:, 0], od[:, 7])
    assert np.array_equal(dbls[:, 1], od[:, 8])
    assert np.array_equal(dbls[:, 2:5], od[:, 9:12])
    assert np.array_equal(dbls[:, 5], od[:, 12])
    assert np.array_equal(dbls[:, 6], od[:, 44])


def test_tree_bitexact_lattice_ties(gpu):
    # a 21^3 lattice: midpoint splits fall on lattice planes, so particles
    # lie exactly on the split and the partition's tie rule decides
    g = np.arange(21) * 0.05
    r = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    m = np.full(len(r), 1.0 / len(r))
    soft = np.full(len(r), 0.01)
    act = np.where(np.arange(len(r)) % 3 == 0, 3, 2).astype(np.uint32)
    T = O.Tree(r, m, soft, iactive=act, nBucket=8, dTheta=0.7, iOrder=4)
    with setup_pkd(gpu, r, m, soft, gpu.PREC_FP64, iactive=act) as pk:
        pk.build_binary(8, 1, 0.7, 4)
        root, perm, ints, dbls = pk.tree_info()
    oi, od = T.nodes()
    assert np.array_equal(perm, T.perm())
    assert np.array_equal(ints[:, :5], oi[:, :5])
    assert np.array_equal(dbls[:, :6], od[:, [7, 8, 9, 10, 11, 12]])
    assert np.array_equal(dbls[:, 6], od[:, 44])


def _tree_vs_oracle(gpu, r, m, soft, nb, order=2, act=None):
    T = O.Tree(r, m, soft, iactive=act, nBucket=nb, dTheta=0.7, iOrder=order)
    with setup_pkd(gpu, r, m, soft, gpu.PREC_FP64, iactive=act) as pk:
        pk.build_binary(nb, 1, 0.7, order)
        root, perm, ints, dbls = pk.tree_info()
        st = pk.grav_all(order)
        a = pk.get("a")
    oi, od = T.nodes()
    assert np.array_equal(perm, T.perm())
    assert np.array_equal(ints[:, :5], oi[:, :5])
    assert np.array_equal(dbls[:, :6], od[:, [7, 8, 9, 10, 11, 12]])
    assert np.array_equal(dbls[:, 6], od[:, 44])
    ao, _, _, sto = T.grav_all()
    assert st["dPartSum"] == sto["dPartSum"] and st["dCellSum"] == sto["dCellSum"]
    e = rel_err(a, ao)        # the FP64 bar of test_forces_vs_oracle
    assert np.sqrt(np.mean(e * e)) < 1e-8 and e.max() < 1e-7


@pytest.mark.parametrize("case", ["bucket1", "clump", "flat", "tiny"])
def test_tree_bitexact_shapes(gpu, case):
    # trees the device build finishes mostly in its per-workgroup subtrees:
    # one particle per bucket, a dense clump many levels deep inside a sphere,
    # a flat disk (no extent in z), and stores smaller than a bucket
    rng = np.random.default_rng(17)
    if case == "bucket1":
        r, v, m = ics.plummer(3000, seed=3)
        nb = 1
    elif case == "clump":
        r, v, m = ics.plummer(6000, seed=4)
        r[:2000] = ics._f32(rng.normal(0, 1e-4, (2000, 3)) + 0.3)
        nb = 8
    elif case == "flat":
        r, v, m = ics.plummer(8000, seed=5)
        r[:, 2] = 0.0
        nb = 16
    else:
        r, v, m = ics.plummer(7, seed=6)
        nb = 16
    soft = np.full(len(m), 0.01)
    _tree_vs_oracle(gpu, r, m, soft, nb)


def test_interaction_lists_bitexact(gpu, plummer32k):
    r, m, soft, T, (ao, po, dto, sto) = plummer32k
    with setup_pkd(gpu, r, m, soft, gpu.PREC_FP32) as pk:
        pk.build_binary(16, 1, 0.7, 2)
        st = pk.grav_all(2)
        bk, off, cnt, ent = pk.grav_lists()
    assert np.array_equal(bk, T.buckets())
    for b, ib in enumerate(bk):
        part, scell, ncell = T.walk(ib)
        for k, ref in enumerate((part, scell, ncell)):
            got = ent[off[b, k]:off[b, k] + cnt[b, k]]
            assert np.array_equal(np.sort(got), np.sort(ref)), (ib, k)
    for key in ("nActive", "dPartSum", "dCellSum", "dSoftSum", "dFlop"):
        assert st[key] == sto[key], key


def mix64(x):
    """splitmix64's finaliser, as gg_grav_walk_checksums sums it."""
    z = np.asarray(x, np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


@pytest.mark.parametrize("case", ["plummer32k", "softened", "hexadecapole"])
def test_fused_walk_lists_checksums(gpu, plummer32k, case):
    # the walk inside the fused kernel that gg_grav_all runs evaluates, per
    # bucket and kind, exactly the oracle's members: their count, sum and
    # sum of mixes (order-independent) agree bucket by bucket
    if case == "plummer32k":
        r, m, soft, T, _ = plummer32k
        nb, theta, order, act = 16, 0.7, 2, None
    else:
        r, v, m = ics.plummer(16384, seed=41)
        soft = np.full(len(m), 0.15 if case == "softened" else 0.01)
        nb, theta, order = (16, 0.7, 2) if case == "softened" else (8, 0.5, 4)
        act = (np.random.default_rng(4).uniform(size=len(m)) < 0.5).astype(np.uint32) | 2
        T = O.Tree(r, m, soft, iactive=act, nBucket=nb, dTheta=theta, iOrder=order)
    for prec in (gpu.PREC_FP32, gpu.PREC_FP64):
        with setup_pkd(gpu, r, m, soft, prec, iactive=act) as pk:
            pk.build_binary(nb, 1, theta, order)
            bk = pk.grav_lists()[0]
            chk = pk.walk_checksums(order)
            assert pk.grav_all(order) is not None
        # the active buckets (with at least one active particle), in tree order
        ints, _ = T.nodes()
        perm, on = T.perm(), (np.full(len(m), 3) if act is None else act) & 1
        assert np.array_equal(bk, [b for b in T.buckets()
                                   if on[perm[ints[b, 1]:ints[b, 2] + 1]].any()])
        nsoft = 0
        for b, ib in enumerate(bk):
            for k, ref in enumerate(T.walk(ib)):
                ref = ref.astype(np.uint64)
                assert chk[b, k, 0] == len(ref), (ib, k)
                assert chk[b, k, 1] == ref.sum(dtype=np.uint64), (ib, k)
                assert chk[b, k, 2] == mix64(ref).sum(dtype=np.uint64), (ib, k)
                nsoft += len(ref) if k == 1 else 0
        if case == "softened":
            assert nsoft > 0


@pytest.mark.parametrize("prec,rms,mx", [("FP64", 1e-8, 1e-7), ("FP32", 1e-5, 1e-4)])
def test_forces_vs_oracle(gpu, plummer32k, prec, rms, mx):
    r, m, soft, T, (ao, po, dto, sto) = plummer32k
    p = gpu.PREC_FP64 if prec == "FP64" else gpu.PREC_FP32
    with setup_pkd(gpu, r, m, soft, p) as pk:
        pk.build_binary(16, 1, 0.7, 2)
        pk.grav_all(2)
        a, pot, dtg = pk.get("a"), pk.get("fPot"), pk.get("dtGrav")
    e = rel_err(a, ao)
    assert np.sqrt(np.mean(e ** 2)) < rms and e.max() < mx
    assert np.max(np.abs(pot / po - 1)) < mx
    assert np.max(np.abs(dtg / dto - 1)) < mx


@pytest.mark.parametrize("order", [1, 3, 4])
def test_multipole_orders_fp64(gpu, order):
    r, v, m = ics.plummer(8192, seed=7)
    soft = np.full(len(m), 0.01)
    T = O.Tree(r, m, soft, nBucket=8, dTheta=0.6, iOrder=order)
    ao, po, _, sto = T.grav_all()
    with setup_pkd(gpu, r, m, soft, gpu.PREC_FP64) as pk:
        pk.build_binary(8, 1, 0.6, order)
        st = pk.grav_all(order)
        a, pot = pk.get("a"), pk.get("fPot")
    assert st["dCellSum"] == sto["dCellSum"] and st["dPartSum"] == sto["dPartSum"]
    assert rel_err(a, ao).max() < 1e-7
    assert np.max(np.abs(pot / po - 1)) < 1e-7


def test_small_opening_angle_uses_the_listed_walk(gpu):
    # theta = 0.05 opens nearly every cell: breadth-first frontiers outgrow
    # the fused kernel's buffer and the listed walk (and, deeper, the
    # one-thread-per-bucket walk) takes over; results must not change
    r, v, m = ics.plummer(8192, seed=12)
    soft = np.full(len(m), 0.01)
    T = O.Tree(r, m, soft, nBucket=8, dTheta=0.05, iOrder=2)
    ao, po, _, sto = T.grav_all()
    with setup_pkd(gpu, r, m, soft, gpu.PREC_FP64) as pk:
        pk.build_binary(8, 1, 0.05, 2)
        st = pk.grav_all(2)
        a, pot = pk.get("a"), pk.get("fPot")
        bk, off, cnt, ent = pk.grav_lists()
    for key in ("dPartSum", "dCellSum", "dSoftSum"):
        assert st[key] == sto[key], key
    assert rel_err(a, ao).max() < 1e-7 and np.max(np.abs(pot / po - 1)) < 1e-7
    for b in range(0, len(bk), 37):
        part, scell, ncell = T.walk(bk[b])
        for k, ref in enumerate((part, scell, ncell)):
            got = ent[off[b, k]:off[b, k] + cnt[b, k]]
            assert np.array_equal(np.sort(got), np.sort(ref))


def test_softened_cells_and_partial_activity(gpu):
    # large softening puts cells on the softened list; a third of the
    # particles active: inactive ones must be left untouched
    r, v, m = ics.plummer(16384, seed=8)
    soft = np.full(len(m), 0.15)
    rng = np.random.default_rng(3)
    act = (rng.uniform(size=len(m)) < 0.33).astype(np.uint32) | 2
    T = O.Tree(r, m, soft, iactive=act, nBucket=16, dTheta=0.7, iOrder=2)
    ao, po, dto, sto = T.grav_all()
    assert sto["dSoftSum"] > 0
    for prec, tol in ((gpu.PREC_FP64, 1e-7), (gpu.PREC_FP32, 1e-4)):
        with setup_pkd(gpu, r, m, soft, prec, iactive=act) as pk:
            pk.build_binary(16, 1, 0.7, 2)
            st = pk.grav_all(2)
            a, pot = pk.get("a"), pk.get("fPot")
        assert st["dSoftSum"] == sto["dSoftSum"] and st["nActive"] == sto["nActive"]
        on = (act & 1) == 1
        assert (a[~on] == 0).all() and (pot[~on] == 0).all()
        assert rel_err(a[on], ao[on]).max() < tol


def test_accumulates_onto_stored_fields(gpu):
    r, v, m = ics.plummer(4096, seed=9)
    soft = np.full(len(m), 0.02)
    with setup_pkd(gpu, r, m, soft, gpu.PREC_FP64) as pk:
        pk.build_binary(16, 1, 0.7, 2)
        pk.grav_all(2)
        a1 = pk.get("a")
        pk.grav_all(2)
        a2 = pk.get("a")
    assert np.allclose(a2, 2 * a1, rtol=1e-14, atol=0)


def test_small_and_degenerate_inputs(gpu):
    # fewer particles than a bucket: intra-bucket pairs only
    r = np.array([[0.0, 0, 0], [1, 0, 0], [0, 2, 0], [0, 0, 3], [1, 1, 1]], float)
    m = np.array([1.0, 2, 3, 4, 5])
    soft = np.full(5, 1e-3)
    ad, pd = O.direct_sum(r, m, soft)
    with setup_pkd(gpu, r, m, soft, gpu.PREC_FP64) as pk:
        pk.build_binary(16, 1, 0.7, 2)
        st = pk.grav_all(2)
        a = pk.get("a")
    assert st["dPartSum"] == 10 and st["dCellSum"] == 0
    assert np.allclose(a, ad, rtol=1e-13)
    # 100 coincident particles: one zero-extent bucket of 100 (pkd.c:2489),
    # every pair at r = 0 inside the softening
    r = np.zeros((100, 3))
    ad, pd = O.direct_sum(r, np.ones(100), np.full(100, 0.1))
    for prec in (gpu.PREC_FP64, gpu.PREC_FP32):
        with setup_pkd(gpu, r, np.ones(100), np.full(100, 0.1), prec) as pk:
            pk.build_binary(16, 1, 0.7, 2)
            st = pk.grav_all(2)
            a, pot = pk.get("a"), pk.get("fPot")
        assert st["dPartSum"] == 100 * 99 / 2
        assert (a == 0).all() and np.allclose(pot, pd, rtol=1e-6)


def test_direct_sum_vs_oracle(gpu):
    # the FP64 direct sum used to state force errors at large N: the oracle's
    # O(N^2) sum on sampled targets, coincident pairs included
    r, v, m = ics.plummer(6000, seed=12)
    r[17] = r[5]
    soft = np.where(np.arange(len(m)) % 2 == 0, 0.01, 0.03)
    idx = np.random.default_rng(2).choice(len(m), 700, replace=False)
    idx[0] = 5
    ao, po = O.direct_sum(r, m, soft, idx)
    with setup_pkd(gpu, r, m, soft, gpu.PREC_FP64) as pk:
        a, pot = pk.direct_grav(idx)
    assert rel_err(a, ao).max() < 1e-12
    assert np.max(np.abs(pot / po - 1)) < 1e-12


def test_c_host_driver(gpu):
    # the C host sequences (gasoline_amd/host) drive a gravity and an SPH
    # step through the C ABI and check conservation
    import os
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "gasoline_amd", "host", "ggforce")
    assert os.path.exists(exe), "ggforce not built"
    out = subprocess.run([exe, "32768", "32768"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "Gravity Calculated" in out.stdout and "ggforce ok" in out.stdout


# ------------------------------------------------ periodic gravity, Ewald
@pytest.fixture(scope="module")
def periodic_box():
    # a clustered periodic unit box: half uniform, half in three clumps,
    # inside [-1/2, 1/2) as the reference's drift keeps it
    rng = np.random.default_rng(21)
    n = 8000
    u = rng.uniform(-0.5, 0.5, (n // 2, 3))
    c = rng.uniform(-0.5, 0.5, (3, 3))
    k = rng.integers(0, 3, n - n // 2)
    cl = c[k] + rng.normal(0, 0.03, (n - n // 2, 3))
    r = ics._f32((np.concatenate([u, cl]) + 0.5) % 1.0 - 0.5)
    m = np.full(n, 1.0 / n)
    soft = np.full(n, 0.01)
    return r, m, soft


@pytest.mark.parametrize("order,nReps,theta", [(2, 1, 0.7), (4, 1, 0.7), (4, 2, 0.35)])
def test_periodic_ewald_vs_oracle(gpu, periodic_box, order, nReps, theta):
    r, m, soft = periodic_box
    T = O.Tree(r, m, soft, nBucket=16, dTheta=theta, iOrder=order)
    ao, po, do, so = T.grav_all(nReps=nReps, period=[1.0] * 3, bEwald=True, iEwOrder=order)
    for prec, rms, mx in ((gpu.PREC_FP64, 1e-8, 1e-7), (gpu.PREC_FP32, 1e-5, 1e-4)):
        with setup_pkd(gpu, r, m, soft, prec) as pk:
            pk.set_period([1.0] * 3)
            pk.build_binary(16, 1, theta, order)
            st = pk.grav_all(order, nReps=nReps, periodic=True, bEwald=True, iEwOrder=order)
            a, pot = pk.get("a"), pk.get("fPot")
        for k in ("nActive", "dPartSum", "dCellSum", "dSoftSum", "dFlop"):
            assert st[k] == so[k], (prec, k)
        e = rel_err(a, ao)
        assert np.sqrt(np.mean(e * e)) < rms and e.max() < mx, prec
        ep = np.abs(pot - po) / np.abs(po).max()
        assert ep.max() < mx, prec


def test_periodic_without_ewald_and_errors(gpu, periodic_box):
    r, m, soft = periodic_box
    T = O.Tree(r, m, soft, nBucket=16, dTheta=0.7, iOrder=2)
    ao, _, _, so = T.grav_all(nReps=1, period=[1.0] * 3, bEwald=False)
    with setup_pkd(gpu, r, m, soft, gpu.PREC_FP64) as pk:
        pk.build_binary(16, 1, 0.7, 2)
        with pytest.raises(gpu.GGError, match="gg_set_period"):
            pk.grav_all(2, nReps=1, periodic=True)
        pk.set_period([1.0, 1.0, 2.0])
        with pytest.raises(gpu.GGError, match="cubic"):
            pk.grav_all(2, nReps=1, periodic=True, bEwald=True)
        pk.set_period([1.0] * 3)
        st = pk.grav_all(2, nReps=1, periodic=True, bEwald=False)   # replicas only
        a = pk.get("a")
    assert st["dPartSum"] == so["dPartSum"] and st["dCellSum"] == so["dCellSum"]
    e = rel_err(a, ao)
    assert np.sqrt(np.mean(e * e)) < 1e-8 and e.max() < 1e-7


# ------------------------------------------------ periodic direct sum
def test_direct_ewald_vs_oracle(gpu, periodic_box):
    # the periodic FP64 direct sum (wave vectors through structure factors,
    # split alpha = 6 / L) against the oracle's pair-by-pair sum at another
    # split (alpha = 3 / L): two softenings and a coincident pair included
    r, m, soft = periodic_box
    r = np.array(r, dtype=np.float64)
    r[17] = r[5]
    soft = np.where(np.arange(len(m)) % 2 == 0, 0.01, 0.02)
    idx = np.random.default_rng(3).choice(len(m), 150, replace=False)
    idx[0] = 5
    ao, po = O.direct_ewald(r, m, soft, 1.0, idx=idx, alphaL=3.0)
    with setup_pkd(gpu, r, m, soft, gpu.PREC_FP64) as pk:
        with pytest.raises(gpu.GGError, match="cubic"):
            pk.direct_ewald(idx)
        pk.set_period([1.0] * 3)
        a, pot = pk.direct_ewald(idx)
    assert rel_err(a, ao).max() < 1e-11
    assert np.abs(pot - po).max() < 1e-12 * np.abs(po).max()


def test_direct_ewald_rock_salt(gpu):
    # signed masses on the rock-salt lattice: every ion at q M / a with the
    # Madelung constant M, and no force
    L = 1.3
    pts = np.array([[x, y, z] for x in range(2) for y in range(2) for z in range(2)], float)
    q = np.where(pts.sum(1) % 2 == 0, 1.0, -1.0)
    r = pts * L / 2 - L / 4
    with setup_pkd(gpu, r, q, np.full(8, 1e-9), gpu.PREC_FP64) as pk:
        pk.set_period([L] * 3)
        a, pot = pk.direct_ewald(np.arange(8))
    assert np.allclose(pot * (L / 2), q * 1.747564594633182, rtol=1e-12, atol=0)
    assert np.abs(a).max() < 1e-12


def test_periodic_tree_error_vs_direct_ewald(gpu, periodic_box):
    # the tree + Ewald forces against the periodic direct sum: at the
    # reference's nReps 1 the hexadecapole expansion of the root in the Ewald
    # correction leaves ~7e-4 rms whatever theta; two replica shells and
    # theta 0.35 bring the rms under 1e-4
    r, m, soft = periodic_box
    idx = np.arange(0, len(m), 7)
    with setup_pkd(gpu, r, m, soft, gpu.PREC_FP64) as pk:
        pk.set_period([1.0] * 3)
        ad, _ = pk.direct_ewald(idx)
        pk.build_binary(16, 1, 0.35, 4)
        pk.init_accel()
        pk.grav_all(4, nReps=2, periodic=True, bEwald=True, iEwOrder=4)
        a2 = pk.get("a")[idx]
        pk.build_binary(16, 1, 0.2, 4)
        pk.init_accel()                  # gg_grav_all adds onto a
        pk.grav_all(4, nReps=1, periodic=True, bEwald=True, iEwOrder=4)
        a1 = pk.get("a")[idx]
    e2, e1 = rel_err(a2, ad), rel_err(a1, ad)
    assert np.sqrt(np.mean(e2 * e2)) < 1e-4
    assert 2e-4 < np.sqrt(np.mean(e1 * e1)) < 2e-3
