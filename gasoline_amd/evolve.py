"""Single-rung kick-drift-kick stepping over the device-resident store: the
host sequence of msrTopStepKDK (master.c:8242) for one rung, calling the
force path (gravity tree + SPH) and the device kick/drift through the C ABI.

    step: open kick (v, u by dt/2; vPred, uPred predicted to the step's end)
          -> drift (r by dt) -> forces at the new positions -> close kick.

Forces: gg_init_accel, then gravity (gg_build_binary + gg_grav_all) if
enabled, then SPH (gg_build_local, optionally gg_reorder, which permutes the
store into the SPH structure's order as pkdBuildLocal does, gg_smooth(DenDVDX),
gg_gas_pressure,
gg_resmooth(SphPressureTermsSym), which adds the hydro acceleration and the
uDot terms).  The h-proposal network, when given, seeds the density pass's
search radius from the recorded history and is trained after it.
"""
import numpy as np

from . import pkd as P


class KDK:
    def __init__(self, pk, smf, nSmooth=64, gravity=True, theta=0.7, iOrder=2, nBucket=16,
                 period=None, hmlp=None, reorder=False, track_order=False):
        self.pk, self.smf, self.K = pk, smf, nSmooth
        self.gravity, self.theta, self.iOrder, self.nBucket = gravity, theta, iOrder, nBucket
        self.period = period
        self.hmlp = hmlp            # dict(train_iters, batch, lr) or None
        self.dt_prev = None
        self.stats = []             # kNN counters of each density pass
        self.grav_stats = None      # counters of the last gravity evaluation
        self.reorder = reorder      # keep the store in the SPH structure's order
        self.track_order = track_order
        self.iord = None            # with track_order: original index of each stored particle

    def seed_search(self):
        """Before a density pass: the network's proposal, if trained."""
        if self.hmlp and self.dt_prev is not None and self.hmlp.get("predict", True):
            self.pk.hmlp_predict()

    def forces(self, dt):
        pk = self.pk
        pk.init_accel()
        if self.gravity:
            pk.build_binary(self.nBucket, 1, self.theta, self.iOrder)
            self.grav_stats = pk.grav_all(self.iOrder)
        pk.build_local(period=self.period)
        if self.reorder:
            perm = pk.reorder(want_perm=self.track_order)
            if self.track_order:
                self.iord = perm if self.iord is None else self.iord[perm]
        self.seed_search()
        pk.smooth(self.smf, self.K, self.period is not None)
        self.stats.append(pk.smooth_stats())
        if self.hmlp:
            if self.dt_prev is not None:
                h = self.hmlp
                self.last_loss = pk.hmlp_train(h.get("train_iters", 50), h.get("batch", 8192),
                                               h.get("lr", 3e-3), h.get("seed", 1))
            pk.hmlp_record(dt)
        pk.gas_pressure(self.smf)
        pk.resmooth(self.smf, self.K, self.period is not None)

    def step(self, dt):
        pk = self.pk
        pk.kick(1.0, 0.5 * dt, 1.0, dt, 0.5 * dt, dt)
        pk.drift(dt, period=self.period)
        self.forces(dt)
        self.dt_prev = dt
        pk.kick(1.0, 0.5 * dt, 1.0, 0.5 * dt, 0.5 * dt, 0.5 * dt)


ALL_GAS = P.TYPE_ACTIVE | P.TYPE_TREEACTIVE | P.TYPE_SMOOTHACTIVE | P.TYPE_GAS


def start(pk, d, soft, bits=None):
    """Put initial conditions: r, v, fMass, fSoft, u (vPred = v, uPred = u)
    and the type bits (default: every particle active gas, the all-gas
    workloads; the kick predicts vPred and advances u of gas only, as
    pkdKick does, pkd.c:3805-3812)."""
    pk.put("r", d["r"])
    pk.put("v", d["v"])
    pk.put("vPred", d["v"])
    pk.put("fMass", d["m"])
    pk.put("fSoft", np.broadcast_to(np.asarray(soft, float), d["m"].shape).copy())
    pk.put("u", d["u"])
    pk.put("uPred", d["u"])
    if bits is None:
        bits = np.full(d["m"].shape, ALL_GAS, np.uint32)
    pk.put("iActive", bits)
    return pk


class RungKDK:
    """msrTopStepKDK (master.c:8242-8400) for collisionless gravity: block
    time steps on rungs 0..iMaxRung-1, particle rungs from the timestep
    criteria (gg_init_dt, gg_grav_step, gg_accel_step, gg_dt_to_rung) at each
    adjust, an open and a close kick of half the rung's step for the
    particles on that rung, recursion into the next rung while any particle
    is deeper, and one drift of every particle plus a force evaluation for
    the rungs from iKickRung down at the bottom.  Kicks are those of
    msrKickKDKOpen/Close without comoving factors (dvFacOne 1, dvFacTwo the
    half step).  bAdjust = False keeps the rungs the store holds."""

    def __init__(self, pk, iMaxRung=8, dEta=0.2, theta=0.7, iOrder=2, nBucket=16,
                 bGravStep=True, bAccelStep=False, bAdjust=True):
        self.pk, self.iMaxRung, self.dEta = pk, iMaxRung, dEta
        self.theta, self.iOrder, self.nBucket = theta, iOrder, nBucket
        self.bGravStep, self.bAccelStep, self.bAdjust = bGravStep, bAccelStep, bAdjust
        self.iCurrMaxRung = 0
        self.force_calls = []        # (iKickRung, particles active) of each force evaluation

    def forces(self, iKickRung):
        """Accelerations of the particles on rungs >= iKickRung (the rest
        keep theirs); all particles are sources."""
        pk = self.pk
        n = pk.active_rung(iKickRung, True)
        self.force_calls.append((iKickRung, n))
        if n:
            pk.init_accel()
            pk.build_binary(self.nBucket, 1, self.theta, self.iOrder)
            pk.grav_all(self.iOrder)

    def adjust(self, iRung, dDelta):
        pk = self.pk
        pk.active_rung(iRung, True)
        pk.init_dt(dDelta)
        if self.bGravStep:
            pk.grav_step(self.dEta)
        if self.bAccelStep:
            pk.accel_step(self.dEta)
        top, _, _ = pk.dt_to_rung(iRung, dDelta, self.iMaxRung, bAll=True)
        self.iCurrMaxRung = top

    def start(self):
        """Forces of every particle before the first step (msrInitAccel +
        msrGravity of the start-up); the rungs come from the first adjust."""
        self.forces(0)
        if not self.bAdjust:
            r = self.pk.get("iRung")
            self.iCurrMaxRung = int(r.max()) if len(r) else 0

    def top_step(self, dDelta, iRung=0, iKickRung=0, iAdjust=True):
        pk = self.pk
        if self.bAdjust and iAdjust and iRung < self.iMaxRung - 1:
            self.adjust(iRung, dDelta)
        pk.active_rung(iRung, False)
        h = 0.5 * dDelta
        pk.kick(1.0, h, 1.0, h, h, h)                  # open kick of this rung
        if self.iCurrMaxRung > iRung:
            self.top_step(h, iRung + 1, iRung + 1, False)
            pk.active_rung(iRung, False)
            self.top_step(h, iRung + 1, iKickRung, True)
        else:
            pk.drift(dDelta)
            self.forces(iKickRung)
        pk.active_rung(iRung, False)
        pk.kick(1.0, h, 1.0, h, h, h)                  # close kick of this rung
