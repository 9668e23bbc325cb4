"""Block time steps (evolve.RungKDK, the msrTopStepKDK recursion) over the
device entry points: gg_active_mask_rung, the timestep criteria,
gg_dt_to_rung, gg_kick, gg_drift and the gravity step.

Bars (no oracle restates the master's recursion; these are what the
scheme fixes): with rungs held fixed, N top steps forward and N back with
the velocities reversed return every particle to its start (the KDK
recursion is time-symmetric) to 1e-9 of the system's size; each top step
evaluates each particle's force 2^rung times; with the rungs adjusted from
gravity's timestep criterion the deep rungs are used, the total energy is
conserved to 1e-3 over three top steps, and its error falls more than
fourfold when dEta halves."""
import numpy as np
import pytest

from gasoline_amd import evolve, ics

pytestmark = pytest.mark.gpu


def plummer_pkd(gpu, n, seed, rungs=None):
    r, v, m = ics.plummer(n, seed=seed)
    pk = gpu.Pkd(n, prec=gpu.PREC_FP64)
    pk.put("r", r)
    pk.put("v", v)
    pk.put("fMass", m)
    pk.put("fSoft", np.full(n, 0.02))
    if rungs is not None:
        pk.put("iRung", rungs)
    return pk, r, v, m


def fixed_rungs(r):
    rad = np.linalg.norm(r, axis=1)
    return np.clip(np.floor(np.log2(2.0 / np.maximum(rad, 1e-3))), 0, 3).astype(np.int32)


def test_fixed_rungs_are_time_reversible_and_counted(gpu):
    n = 4000
    r0 = ics.plummer(n, seed=3)[0]
    rungs = fixed_rungs(r0)
    assert rungs.max() == 3 and rungs.min() == 0
    pk, r, v, m = plummer_pkd(gpu, n, 3, rungs)
    with pk:
        drv = evolve.RungKDK(pk, iMaxRung=4, bAdjust=False)
        drv.start()
        drv.force_calls.clear()
        drv.top_step(0.02)
        # one top step: 8 force evaluations at the bottom, each particle's 2^rung times
        assert len(drv.force_calls) == 8
        assert sum(c for _, c in drv.force_calls) == int((2 ** rungs).sum())
        for _ in range(3):
            drv.top_step(0.02)
        pk.put("v", -pk.get("v"))
        for _ in range(4):
            drv.top_step(0.02)
        back = pk.get("r")
        assert np.array_equal(pk.get("iRung"), rungs)
    assert np.abs(back - r).max() < 1e-9


def energy(pk, m):
    v, pot = pk.get("v"), pk.get("fPot")
    return 0.5 * np.sum(m * np.sum(v * v, axis=1)) + 0.5 * np.sum(m * pot)


def _adaptive_energy_error(gpu, eta):
    n = 4000
    pk, r, v, m = plummer_pkd(gpu, n, 4)
    with pk:
        # accurate forces (theta 0.4, hexadecapole) so that the integration,
        # not the tree's force error, sets the energy error
        drv = evolve.RungKDK(pk, iMaxRung=6, dEta=eta, theta=0.4, iOrder=4)
        drv.start()
        e0 = energy(pk, m)
        for _ in range(3):
            drv.top_step(0.1)
        e1 = energy(pk, m)
        rungs = pk.get("iRung")
    return abs(e1 - e0) / abs(e0), drv.iCurrMaxRung, rungs


def test_adjusted_rungs_conserve_energy(gpu):
    # the error falls fast as dEta shrinks (deeper rungs); with
    # dEta = 0.05 it is below 1e-3 over three top steps of 0.1
    e1, top1, r1 = _adaptive_energy_error(gpu, 0.05)
    e2, top2, r2 = _adaptive_energy_error(gpu, 0.025)
    assert top1 >= 3 and r1.min() == 0 and top2 >= top1
    assert e1 < 1e-3 and e2 < e1 / 4
