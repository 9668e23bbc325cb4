"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_longlong)
_up = C.POINTER(C.c_ubyte)
_uip = C.POINTER(C.c_uint)


class Smf(C.Structure):
    """orc_smf: the subset of gasoline's SMF (smoothfcn.h:22-130) the
    DenDVDX / SphPressureTermsSym / pkdGasPressure path reads."""
    _fields_ = [("H", C.c_double), ("a", C.c_double), ("alpha", C.c_double),
                ("beta", C.c_double), ("gamma", C.c_double),
                ("dtFacCourant", C.c_double), ("dtFacDiffusion", C.c_double),
                ("dMetalDiffusionCoeff", C.c_double),
                ("dThermalDiffusionCoeff", C.c_double),
                ("dResolveJeans", C.c_double), ("bCannonical", C.c_int),
                ("iViscosityLimiter", C.c_int), ("bConstantDiffusion", C.c_int),
                ("pad", C.c_int)]


def default_smf(**kw):
    """Parameters of the SPH box workload: H=0, a=1, alpha=1, beta=2,
    gamma=5/3, dEtaCourant=0.4 (dtFacCourant = 0.4*2/1.6, pkd.c:5893),
    plain Monaghan viscosity (iViscosityLimiter=0), no diffusion."""
    s = Smf(H=0.0, a=1.0, alpha=1.0, beta=2.0, gamma=5.0 / 3.0,
            dtFacCourant=0.4 * 2 / 1.6, dtFacDiffusion=2 * 0.1,
            dMetalDiffusionCoeff=0.0, dThermalDiffusionCoeff=0.0,
            dResolveJeans=0.0, bCannonical=1, iViscosityLimiter=0,
            bConstantDiffusion=0, pad=0)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_vsqrt1.argtypes = [C.c_int, _dp, _dp]
        L.orc_build_binary.restype = C.c_void_p
        L.orc_build_binary.argtypes = [C.c_int, _dp, _dp, _dp, _uip, C.c_int,
                                       C.c_int, C.c_double, C.c_int]
        L.orc_tree_free.argtypes = [C.c_void_p]
        L.orc_tree_nnodes.argtypes = [C.c_void_p]
        L.orc_tree_root.argtypes = [C.c_void_p]
        L.orc_tree_perm.argtypes = [C.c_void_p, _ip]
        L.orc_tree_nodes.argtypes = [C.c_void_p, _ip, _dp]
        L.orc_walk_bucket.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int,
                                      _ip, _ip, _ip, _ip]
        L.orc_grav_all.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_grav_all_periodic.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_int, C.c_int,
                                            C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_grav_all_weighted.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_int, C.c_int,
                                            C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.orc_direct_sum.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _ip, _dp, _dp]
        L.orc_direct_ewald.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int,
                                       _ip, _dp, _dp]
        L.orc_ewald_self.restype = C.c_double
        L.orc_ewald_self.argtypes = [C.c_double, C.c_double]
        L.orc_spline.argtypes = [C.c_double, C.c_double, _dp, _dp]
        L.orc_splineq.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_knn.restype = C.c_int
        L.orc_knn.argtypes = [C.c_int, _dp, C.c_int, C.c_int, _dp, _ip, _dp]
        L.orc_ball_gather.restype = C.c_longlong
        L.orc_ball_gather.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _lp, _ip, _dp, _dp]
        L.orc_dendvdx.argtypes = [C.c_int, _dp, _dp, _dp, _up, _up, _dp, _lp, _ip,
                                  _dp, _dp, C.POINTER(Smf), _dp, _dp, _dp, _dp,
                                  _dp, _dp, _dp, _up]
        L.orc_density_sym.argtypes = [C.c_int, _dp, _up, _dp, _lp, _ip, _dp, _dp]
        L.orc_gas_pressure.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.POINTER(Smf),
                                       _dp, _dp, _dp]
        L.orc_sph_pressure_sym.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                           _dp, _dp, _dp, _up, _dp, _lp, _ip, _dp,
                                           _dp, C.POINTER(Smf), _dp, _dp, _dp, _dp,
                                           _dp, _dp]
        L.orc_sph_pressure_sym_serial.argtypes = L.orc_sph_pressure_sym.argtypes
        _u = _uip
        L.orc_init_accel.argtypes = [C.c_int, _u, _dp, _dp, _dp]
        L.orc_kick.argtypes = [C.c_int, _u] + [C.c_double] * 6 + [C.c_int] + [_dp] * 8
        L.orc_kick_vpred.restype = C.c_int
        L.orc_kick_vpred.argtypes = [C.c_int, _u, C.c_double, C.c_double, C.c_double, C.c_int,
                                     _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_drift.restype = C.c_int
        L.orc_drift.argtypes = [C.c_int, C.c_double, _dp, C.c_int, _dp, _dp, _dp]
        L.orc_init_dt.argtypes = [C.c_int, _u, C.c_double, _dp]
        L.orc_grav_step.restype = C.c_int
        L.orc_grav_step.argtypes = [C.c_int, _u, C.c_double, _dp, _dp]
        L.orc_accel_step.argtypes = [C.c_int, _u, C.c_double, C.c_double, C.c_double, C.c_int,
                                     C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_density_step.argtypes = [C.c_int, _u, C.c_double, C.c_double, _dp, _dp]
        L.orc_sph_step.argtypes = [C.c_int, _u] + [C.c_double] * 7 + [_dp] * 11
        L.orc_one_dt_to_rung.restype = C.c_int
        L.orc_one_dt_to_rung.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_dt_to_rung.restype = C.c_int
        L.orc_dt_to_rung.argtypes = [C.c_int, _u, C.c_int, C.c_double, C.c_int, C.c_int, _dp,
                                     _dp, _ip, _dp, _ip, _ip]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own v_sqrt1.c built into oracle/_ref (None if absent)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libvsqrt1_ref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.v_sqrt1.argtypes = [C.c_int, _dp, _dp]
        _REF = L
    return _REF


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def P(a, t=_dp):
    return a.ctypes.data_as(t)


def vsqrt1(r2):
    r2 = _d(r2)
    out = np.empty_like(r2)
    lib().orc_vsqrt1(r2.size, P(r2), P(out))
    return out


def ref_vsqrt1(r2):
    L = ref_lib()
    if L is None:
        return None
    r2 = _d(r2)
    out = np.empty_like(r2)
    L.v_sqrt1(r2.size, P(r2), P(out))
    return out


class Tree:
    """Gravity tree of pkdBuildBinary (pkd.c:2627) over input-ordered arrays."""

    def __init__(self, r, m, soft, iactive=None, nBucket=16, dTheta=0.7, iOrder=2,
                 iOpenType=1):
        self.r = _d(r).reshape(-1, 3)
        self.n = self.r.shape[0]
        self.m = _d(m)
        self.soft = _d(soft)
        self.iactive = (np.full(self.n, 3, np.uint32) if iactive is None
                        else np.ascontiguousarray(iactive, dtype=np.uint32))
        self.iOrder = iOrder
        self.h = lib().orc_build_binary(self.n, P(self.r), P(self.m), P(self.soft),
                                        P(self.iactive, _uip), nBucket, iOpenType,
                                        float(dTheta), iOrder)
        if not self.h:
            raise MemoryError("orc_build_binary failed")

    def __del__(self):
        h = getattr(self, "h", None)
        self.h = None
        if h and _LIB is not None:
            try:
                _LIB.orc_tree_free(h)
            except Exception:
                pass

    @property
    def nnodes(self):
        return lib().orc_tree_nnodes(self.h)

    @property
    def root(self):
        return lib().orc_tree_root(self.h)

    def perm(self):
        p = np.empty(self.n, np.int32)
        lib().orc_tree_perm(self.h, P(p, _ip))
        return p

    def nodes(self):
        nn = self.nnodes
        ints = np.empty((nn, 6), np.int32)
        dbls = np.empty((nn, 47), np.float64)
        lib().orc_tree_nodes(self.h, P(ints, _ip), P(dbls))
        return ints, dbls

    def buckets(self):
        ints, _ = self.nodes()
        return np.nonzero(ints[:, 0] == -1)[0]

    def walk(self, ib, cap=4096):
        while True:
            part = np.empty(cap, np.int32)
            soft = np.empty(cap, np.int32)
            newt = np.empty(cap, np.int32)
            cnt = np.zeros(3, np.int32)
            lib().orc_walk_bucket(self.h, int(ib), self.iOrder, cap, P(part, _ip),
                                  P(soft, _ip), P(newt, _ip), P(cnt, _ip))
            if cnt.max() <= cap:
                return part[:cnt[0]].copy(), soft[:cnt[1]].copy(), newt[:cnt[2]].copy()
            cap = int(cnt.max()) * 2

    def grav_all(self, nReps=0, period=None, bEwald=False, iEwOrder=4, fEwCut=2.6,
                 fEwhCut=2.8, fweight=None):
        """a, fPot, dtGrav (input order) and the counters; with fweight (an
        array of n doubles) its active entries receive their bucket's flops."""
        a = np.zeros((self.n, 3))
        pot = np.zeros(self.n)
        dtg = np.zeros(self.n)
        st = np.zeros(5)
        L = None if period is None else _d(period)
        fw = None
        if fweight is not None:
            assert fweight.dtype == np.float64 and fweight.shape == (self.n,)
            fw = P(fweight)
        lib().orc_grav_all_weighted(self.h, self.iOrder, int(nReps) if L is not None else 0,
                                    None if L is None else P(L), int(bEwald), int(iEwOrder),
                                    fEwCut, fEwhCut, P(a), P(pot), P(dtg), P(st), fw)
        return a, pot, dtg, dict(nActive=st[0], dPartSum=st[1], dCellSum=st[2],
                                 dSoftSum=st[3], dFlop=st[4])


def direct_sum(r, m, soft, idx=None):
    r = _d(r).reshape(-1, 3)
    m = _d(m)
    soft = _d(soft)
    n = r.shape[0]
    if idx is None:
        nt, ip = n, None
    else:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        nt, ip = idx.size, P(idx, _ip)
    a = np.zeros((nt, 3))
    pot = np.zeros(nt)
    lib().orc_direct_sum(n, P(r), P(m), P(soft), nt, ip, P(a), P(pot))
    return a, pot


def direct_ewald(r, m, soft, L, idx=None, alphaL=6.0):
    """Periodic direct sum (orc_direct_ewald) in the cube of side L."""
    r = _d(r).reshape(-1, 3)
    m = _d(m)
    soft = _d(soft)
    n = r.shape[0]
    if idx is None:
        nt, ip = n, None
    else:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        nt, ip = idx.size, P(idx, _ip)
    a = np.zeros((nt, 3))
    pot = np.zeros(nt)
    lib().orc_direct_ewald(n, P(r), P(m), P(soft), float(L), float(alphaL), nt, ip, P(a), P(pot))
    return a, pot


def ewald_self(L, alphaL=6.0):
    return lib().orc_ewald_self(float(L), float(alphaL))


def spline(r2, twoh):
    a = C.c_double()
    b = C.c_double()
    lib().orc_spline(r2, twoh, C.byref(a), C.byref(b))
    return a.value, b.value


def splineq(invr, r2, twoh):
    v = [C.c_double() for _ in range(4)]
    lib().orc_splineq(invr, r2, twoh, *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def knn(r, nSmooth, period=None):
    r = _d(r).reshape(-1, 3)
    n = r.shape[0]
    L = _d(period if period is not None else [0, 0, 0])
    nbr = np.empty((n, nSmooth), np.int32)
    fb2 = np.empty(n)
    ties = lib().orc_knn(n, P(r), nSmooth, int(period is not None), P(L), P(nbr, _ip), P(fb2))
    return nbr, fb2, ties


def ball_gather(r, fBall2, period=None):
    """CSR ball gather: (off[n+1], idx, dxyz[.,3], d2)."""
    r = _d(r).reshape(-1, 3)
    n = r.shape[0]
    fBall2 = _d(fBall2)
    L = _d(period if period is not None else [0, 0, 0])
    per = int(period is not None)
    off = np.zeros(n + 1, np.int64)
    tot = lib().orc_ball_gather(n, P(r), P(fBall2), per, P(L), P(off, _lp), None, None, None)
    idx = np.empty(max(tot, 1), np.int32)
    dxyz = np.empty((max(tot, 1), 3))
    d2 = np.empty(max(tot, 1))
    lib().orc_ball_gather(n, P(r), P(fBall2), per, P(L), P(off, _lp), P(idx, _ip), P(dxyz), P(d2))
    return off, idx[:tot], dxyz[:tot], d2[:tot]


def dendvdx(m, vpred, upred, fBall2, lists, smf, active=None, smooth_active=None):
    off, idx, dxyz, d2 = lists
    n = len(m)
    m = _d(m)
    vpred = _d(vpred).reshape(-1, 3)
    upred = _d(upred)
    fBall2 = _d(fBall2)
    act = np.ones(n, np.uint8) if active is None else np.ascontiguousarray(active, np.uint8)
    sact = np.ones(n, np.uint8) if smooth_active is None else np.ascontiguousarray(smooth_active, np.uint8)
    out = {k: np.zeros(n) for k in ("rho", "divv", "dvds", "fDivv_t", "balsara", "diff")}
    curlv = np.zeros((n, 3))
    nbract = np.zeros(n, np.uint8)
    dxyz = _d(dxyz)
    d2 = _d(d2)
    idx = np.ascontiguousarray(idx, np.int32)
    off = np.ascontiguousarray(off, np.int64)
    lib().orc_dendvdx(n, P(m), P(vpred), P(upred), P(act, _up), P(sact, _up), P(fBall2),
                      P(off, _lp), P(idx, _ip), P(dxyz), P(d2), C.byref(smf),
                      P(out["rho"]), P(out["divv"]), P(curlv), P(out["dvds"]),
                      P(out["fDivv_t"]), P(out["balsara"]), P(out["diff"]), P(nbract, _up))
    out["curlv"] = curlv
    out["nbrOfActive"] = nbract
    return out


def density_sym(m, fBall2, lists, smooth_active=None, rho=None):
    """DensitySym over CSR lists (off, idx, dxyz, d2); rho (if given) is the
    value the particles that are not smooth-active start from."""
    off, idx, _, d2 = lists
    n = len(m)
    sact = np.ones(n, np.uint8) if smooth_active is None else np.ascontiguousarray(smooth_active, np.uint8)
    out = np.zeros(n) if rho is None else _d(rho).copy()
    lib().orc_density_sym(n, P(_d(m)), P(sact, _up), P(_d(fBall2)),
                          P(np.ascontiguousarray(off, np.int64), _lp),
                          P(np.ascontiguousarray(idx, np.int32), _ip), P(_d(d2)), P(out))
    return out


def gas_pressure(upred, rho, fBall2, smf, dt=None, udot=None):
    n = len(upred)
    upred = _d(upred)
    rho = _d(rho)
    fBall2 = _d(fBall2)
    dt = np.full(n, 1e30) if dt is None else _d(dt).copy()
    ud = None if udot is None else _d(udot)
    pr2 = np.zeros(n)
    c = np.zeros(n)
    lib().orc_gas_pressure(n, P(upred), P(rho), P(fBall2), None if ud is None else P(ud),
                           C.byref(smf), P(pr2), P(c), P(dt))
    return pr2, c, dt


def sph_pressure_sym(m, vpred, upred, rho, PoverRho2, c, balsara, diff, dt, fBall2,
                     lists, smf, active=None, acc=None, serial=False):
    """SphPressureTermsSym over CSR lists: the threaded gather form, or with
    serial=True the serial scatter form."""
    off, idx, dxyz, d2 = lists
    n = len(m)
    act = np.ones(n, np.uint8) if active is None else np.ascontiguousarray(active, np.uint8)
    acc = np.zeros((n, 3)) if acc is None else _d(acc).reshape(-1, 3).copy()
    outs = {k: np.zeros(n) for k in ("uDotPdV", "uDotAV", "uDotDiff", "mumax")}
    dtNew = np.full(n, np.float64(np.float32(3.402823466e38)))
    arrs = [_d(x) for x in (m, vpred, upred, rho, PoverRho2, c, balsara, diff, dt, fBall2)]
    arrs[1] = arrs[1].reshape(-1, 3)
    idx = np.ascontiguousarray(idx, np.int32)
    off = np.ascontiguousarray(off, np.int64)
    dxyz = _d(dxyz)
    d2 = _d(d2)
    fn = lib().orc_sph_pressure_sym_serial if serial else lib().orc_sph_pressure_sym
    fn(n, P(arrs[0]), P(arrs[1]), P(arrs[2]), P(arrs[3]), P(arrs[4]), P(arrs[5]), P(arrs[6]),
       P(arrs[7]), P(arrs[8]), P(act, _up), P(arrs[9]), P(off, _lp), P(idx, _ip), P(dxyz),
       P(d2), C.byref(smf), P(acc), P(outs["uDotPdV"]), P(outs["uDotAV"]),
       P(outs["uDotDiff"]), P(outs["mumax"]), P(dtNew))
    outs["acc"] = acc
    outs["dtNew"] = dtNew
    return outs


# ---------------------------------------------------- updates either side
# Each takes a dict of SoA fields (names as the library's) and returns the
# updated fields as new arrays; the inputs are not modified.
def _bits(b):
    return np.ascontiguousarray(b, np.uint32)


def _cp(f, k):
    return _d(f[k]).copy()


def init_accel(f):
    a, pot, dtg = _cp(f, "a"), _cp(f, "fPot"), _cp(f, "dtGrav")
    b = _bits(f["iActive"])
    lib().orc_init_accel(len(b), P(b, _uip), P(a), P(pot), P(dtg))
    return dict(a=a, fPot=pot, dtGrav=dtg)


def kick(f, dvFacOne, dvFacTwo, dvPredFacOne, dvPredFacTwo, duDelta, duPredDelta,
         bIsothermal=False):
    v, vp, u, up = _cp(f, "v"), _cp(f, "vPred"), _cp(f, "u"), _cp(f, "uPred")
    b = _bits(f["iActive"])
    lib().orc_kick(len(b), P(b, _uip), dvFacOne, dvFacTwo, dvPredFacOne, dvPredFacTwo, duDelta,
                   duPredDelta, int(bIsothermal), P(_d(f["a"])), P(_d(f["uDotPdV"])),
                   P(_d(f["uDotAV"])), P(_d(f["uDotDiff"])), P(v), P(vp), P(u), P(up))
    return dict(v=v, vPred=vp, u=u, uPred=up)


def kick_vpred(f, dvFacOne, dvFacTwo, duDelta, bIsothermal=False):
    vp, up = _cp(f, "vPred"), _cp(f, "uPred")
    b = _bits(f["iActive"])
    bad = lib().orc_kick_vpred(len(b), P(b, _uip), dvFacOne, dvFacTwo, duDelta, int(bIsothermal),
                               P(_d(f["a"])), P(_d(f["uDotPdV"])), P(_d(f["uDotAV"])),
                               P(_d(f["uDotDiff"])), P(vp), P(up))
    return dict(vPred=vp, uPred=up, bad=bad)


def drift(f, dDelta, fCenter=None, fPeriod=None):
    r = _cp(f, "r")
    per = fPeriod is not None
    c = _d(fCenter if per else [0, 0, 0])
    L = _d(fPeriod if per else [0, 0, 0])
    bad = lib().orc_drift(len(r), dDelta, P(c), int(per), P(L), P(_d(f["v"])), P(r))
    return dict(r=r, bad=bad)


def init_dt(f, dDelta):
    dt = _cp(f, "dt")
    b = _bits(f["iActive"])
    lib().orc_init_dt(len(b), P(b, _uip), dDelta, P(dt))
    return dict(dt=dt)


def grav_step(f, dEta):
    dt = _cp(f, "dt")
    b = _bits(f["iActive"])
    bad = lib().orc_grav_step(len(b), P(b, _uip), dEta, P(_d(f["dtGrav"])), P(dt))
    return dict(dt=dt, bad=bad)


def accel_step(f, dEta, dVelFac, dAccFac, bEpsAcc, bSqrtPhi):
    dt = _cp(f, "dt")
    b = _bits(f["iActive"])
    lib().orc_accel_step(len(b), P(b, _uip), dEta, dVelFac, dAccFac, int(bEpsAcc), int(bSqrtPhi),
                         P(_d(f["v"])), P(_d(f["a"])), P(_d(f["fSoft"])), P(_d(f["fBall2"])),
                         P(_d(f["fPot"])), P(dt))
    return dict(dt=dt)


def density_step(f, dEta, dRhoFac):
    dt = _cp(f, "dt")
    b = _bits(f["iActive"])
    lib().orc_density_step(len(b), P(b, _uip), dEta, dRhoFac, P(_d(f["fDensity"])), P(dt))
    return dict(dt=dt)


def sph_step(f, dCosmoFac, dEtaCourant, dEtauDot, duMinDt, dDiffCoeff, dEtaDiffusion,
             dResolveJeans):
    dt, dtn = _cp(f, "dt"), _cp(f, "dtNew")
    b = _bits(f["iActive"])
    m = C.c_double()
    lib().orc_sph_step(len(b), P(b, _uip), dCosmoFac, dEtaCourant, dEtauDot, duMinDt, dDiffCoeff,
                       dEtaDiffusion, dResolveJeans, P(_d(f["fBall2"])), P(_d(f["c"])),
                       P(_d(f["u"])), P(_d(f["uDotPdV"])), P(_d(f["uDotAV"])),
                       P(_d(f["uDotDiff"])), P(_d(f["fDensity"])), P(_d(f["diff"])), P(dtn),
                       P(dt), C.byref(m))
    return dict(dt=dt, dtNew=dtn, dtMinGas=m.value)


def one_dt_to_rung(iRung, dDelta, dt):
    return lib().orc_one_dt_to_rung(iRung, dDelta, dt)


def dt_to_rung(f, iRung, dDelta, iMaxRung, bAll):
    rung = np.ascontiguousarray(f["iRung"], np.int32).copy()
    dto = _cp(f, "dtOld")
    b = _bits(f["iActive"])
    nmax, ideal = C.c_int(), C.c_int()
    top = lib().orc_dt_to_rung(len(b), P(b, _uip), iRung, dDelta, iMaxRung, int(bAll),
                               P(_d(f["fSoft"])), P(_d(f["dt"])), P(rung, _ip), P(dto),
                               C.byref(nmax), C.byref(ideal))
    return dict(iRung=rung, dtOld=dto, iMaxRung=top, nMaxRung=nmax.value,
                iMaxRungIdeal=ideal.value)
