"""The batched ensemble through the C ABI: independent systems in one store,
one tree per system (gg_build_forest), one launch for all (gg_grav_all), the
HIP-graph replay of that launch (gg_grav_graph / gg_grav_replay) and the
bucket-size autotuner (gg_tune_bucket).

Bars: the forest's accelerations, potentials and dtGrav equal, bit for bit,
those of each system evaluated alone, and its counters are their sums; every
graph replay equals the eager evaluation bit for bit; the autotuner returns
one of its candidates and leaves the caller's fields as they were."""
import numpy as np
import pytest

from gasoline_amd import ics
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ens():
    return ics.king_ensemble(7, seed=4, nmin=300, nmax=3000)


def load(pk, d, eps):
    pk.put("r", d["r"])
    pk.put("fMass", d["m"])
    pk.put("fSoft", np.full(len(d["m"]), eps))


@pytest.mark.parametrize("prec", ["FP32", "FP64"])
def test_forest_equals_separate_runs(gpu, ens, prec):
    p = gpu.PREC_FP32 if prec == "FP32" else gpu.PREC_FP64
    off = ens["off"]
    n = len(ens["m"])
    with gpu.Pkd(n, prec=p) as pk:
        load(pk, ens, 0.005)
        pk.build_forest(off, nBucket=16, dCrit=0.6)
        st = pk.grav_all(2)
        a, pot, dtg = pk.get("a"), pk.get("fPot"), pk.get("dtGrav")
    tot = dict.fromkeys(st, 0.0)
    for s in range(len(off) - 1):
        sl = slice(off[s], off[s + 1])
        d = {k: ens[k][sl] for k in ("r", "m")}
        with gpu.Pkd(off[s + 1] - off[s], prec=p) as pk:
            load(pk, d, 0.005)
            pk.build_binary(16, 1, 0.6, 2)
            ss = pk.grav_all(2)
            assert np.array_equal(pk.get("a"), a[sl])
            assert np.array_equal(pk.get("fPot"), pot[sl])
            assert np.array_equal(pk.get("dtGrav"), dtg[sl])
        for k in ss:
            tot[k] += ss[k]
    for k in ("nActive", "dPartSum", "dCellSum", "dSoftSum", "dFlop"):
        assert st[k] == tot[k], k


def test_forest_with_empty_systems(gpu, ens):
    # empty systems have no tree; the others are evaluated as alone
    off = ens["off"]
    n = len(ens["m"])
    off2 = np.r_[0, 0, off[1:3], off[3], off[3:], n].astype(np.int32)
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        load(pk, ens, 0.005)
        pk.build_forest(off, nBucket=16, dCrit=0.6)
        pk.grav_all(2)
        a = pk.get("a")
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        load(pk, ens, 0.005)
        pk.build_forest(off2, nBucket=16, dCrit=0.6)
        pk.grav_all(2)
        assert np.array_equal(pk.get("a"), a)


def test_forest_of_tiny_and_uneven_systems(gpu):
    # systems of 1, 2, 17, 600 and 5000 particles (the small ones become
    # per-workgroup subtrees at the root level): each equals its own run
    sizes = [1, 2, 17, 600, 5000]
    parts = [ics.king(k, 5.0, seed=30 + k) if k > 1 else
             dict(r=np.zeros((1, 3)), v=np.zeros((1, 3)), m=np.ones(1)) for k in sizes]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    cat = {k: np.concatenate([p[k] for p in parts]) for k in ("r", "m")}
    n = off[-1]
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        load(pk, cat, 0.01)
        pk.build_forest(off, nBucket=16, dCrit=0.6)
        pk.grav_all(2)
        a = pk.get("a")
    for s in range(len(sizes)):
        sl = slice(off[s], off[s + 1])
        with gpu.Pkd(sizes[s], prec=gpu.PREC_FP64) as pk:
            load(pk, {k: cat[k][sl] for k in ("r", "m")}, 0.01)
            pk.build_binary(16, 1, 0.6, 2)
            pk.grav_all(2)
            assert np.array_equal(pk.get("a"), a[sl]), s


def test_forest_system_vs_oracle(gpu, ens):
    off = ens["off"]
    s = 2
    sl = slice(off[s], off[s + 1])
    soft = np.full(off[s + 1] - off[s], 0.005)
    T = O.Tree(ens["r"][sl], ens["m"][sl], soft, nBucket=16, dTheta=0.6, iOrder=2)
    ao, po, _, sto = T.grav_all()
    with gpu.Pkd(len(ens["m"]), prec=gpu.PREC_FP32) as pk:
        load(pk, ens, 0.005)
        pk.build_forest(off, nBucket=16, dCrit=0.6)
        pk.grav_all(2)
        a = pk.get("a")[sl]
    e = np.linalg.norm(a - ao, axis=1) / np.linalg.norm(ao, axis=1)
    assert np.sqrt(np.mean(e ** 2)) < 1e-5 and e.max() < 1e-4


def test_graph_replay_equals_eager(gpu, ens):
    n = len(ens["m"])
    with gpu.Pkd(n, prec=gpu.PREC_FP32) as pk, gpu.Pkd(4096) as other:
        load(pk, ens, 0.005)
        pk.build_forest(ens["off"], nBucket=16, dCrit=0.6)
        pk.init_accel()
        st0 = pk.grav_all(2)
        a0, p0, d0 = pk.get("a"), pk.get("fPot"), pk.get("dtGrav")
        pk.grav_graph(2)
        r, v, m = ics.plummer(4096, seed=5)
        load(other, dict(r=r, m=m), 0.01)
        for rep in range(4):
            pk.init_accel()
            st = pk.grav_replay()
            assert st == st0
            assert np.array_equal(pk.get("a"), a0), rep
            assert np.array_equal(pk.get("fPot"), p0), rep
            assert np.array_equal(pk.get("dtGrav"), d0), rep
            other.build_binary(16, 1, 0.7, 2)     # eager work between replays
            other.grav_all(2)
        # a replay adds, as gg_grav_all does
        st = pk.grav_replay()
        assert np.allclose(pk.get("a"), 2 * a0, rtol=1e-6, atol=0)


def test_autotuner(gpu, ens):
    n = len(ens["m"])
    with gpu.Pkd(n, prec=gpu.PREC_FP32) as pk:
        load(pk, ens, 0.005)
        pk.put("a", np.full((n, 3), 7.0))
        best, ms = pk.tune_bucket((8, 16, 32), sysOff=ens["off"], dCrit=0.6)
        assert best in (8, 16, 32) and all(t > 0 for t in ms.values())
        assert np.all(pk.get("a") == 7.0)
        pk.init_accel()
        pk.grav_all(2)       # the handle is built with the chosen bucket size
        a = pk.get("a")
    with gpu.Pkd(n, prec=gpu.PREC_FP32) as pk:
        load(pk, ens, 0.005)
        pk.build_forest(ens["off"], nBucket=best, dCrit=0.6)
        pk.grav_all(2)
        assert np.array_equal(pk.get("a"), a)


@pytest.mark.parametrize("prec", ["FP32", "FP64"])
def test_persistent_schedule_equals_per_bucket(gpu, ens, prec):
    # the persistent grid (workgroups taking buckets from a queue, heaviest
    # first once a first evaluation has measured them) gives the per-bucket
    # launch's results bit for bit: each bucket's sums do not depend on when
    # it runs; also replayed from a graph captured with the queue; and the
    # persistent grid whose XCDs take the tree in shares and then help
    # one another
    p = gpu.PREC_FP32 if prec == "FP32" else gpu.PREC_FP64
    n = len(ens["m"])
    out = []
    for mode in (0, 1, 2):
        with gpu.Pkd(n, prec=p) as pk:
            load(pk, ens, 0.005)
            pk.set_grav_schedule(mode)
            pk.build_forest(ens["off"], nBucket=16, dCrit=0.6)
            res = []
            for call in range(3):            # identity queue, then heaviest first
                pk.init_accel()
                st = pk.grav_all(2)
                res.append((pk.get("a"), pk.get("fPot"), pk.get("dtGrav"), st))
            pk.grav_graph(2)
            for _ in range(2):
                pk.init_accel()
                st = pk.grav_replay()
                res.append((pk.get("a"), pk.get("fPot"), pk.get("dtGrav"), st))
            out.append(res)
    ref = out[0][0]
    for res in out:
        for a, pot, dtg, st in res:
            assert np.array_equal(a, ref[0]) and np.array_equal(pot, ref[1])
            assert np.array_equal(dtg, ref[2])
            for k in ("nActive", "dPartSum", "dCellSum", "dSoftSum", "dFlop"):
                assert st[k] == ref[3][k], k
