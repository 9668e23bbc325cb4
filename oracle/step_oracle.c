/*
 * step_oracle.c -- CPU restatement of the per-particle updates on either
 * side of the force path: pkdInitAccel, pkdKick, pkdKickVpred, pkdDrift,
 * pkdInitDt, pkdGravStep, pkdAccelStep, pkdDensityStep, pkdSphStep and
 * pkdDtToRung of a -DGASOLINE -DNOCOOLING build (pkd.c).  Arrays are SoA in
 * the caller's order; vectors are xyz interleaved.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with -ffp-contract=off
 * so that every product and sum rounds as on x86-64 without FMA.
 */
#include <float.h>
#include <math.h>

#include "oracle.h"

#define ACT 1u      /* TYPE_ACTIVE, pkd.h:288 */
#define GAS 256u    /* TYPE_GAS, pkd.h:296 */

/* pkdInitAccel (pkd.c:5443-5456) */
void orc_init_accel(int n, const unsigned *bits, double *a, double *pot, double *dtg)
{
    int i, j;
    for (i = 0; i < n; ++i) {
        if (!(bits[i] & ACT)) continue;
        for (j = 0; j < 3; ++j) a[3 * i + j] = 0;
        pot[i] = 0;
        dtg[i] = 0;
    }
}

/* pkdKick (pkd.c:3780-3967), NOCOOLING GASOLINE build: active gas advances
 * vPred, v and, unless isothermal, uPred and u by UDOT_HYDRO (pkd.h:65);
 * other active particles only v (vPred is written for them only in builds
 * without GASOLINE, pkd.c:3953-3957). */
void orc_kick(int n, const unsigned *bits, double dvFacOne, double dvFacTwo,
              double dvPredFacOne, double dvPredFacTwo, double duDelta, double duPredDelta,
              int bIsothermal, const double *a, const double *uDotPdV, const double *uDotAV,
              const double *uDotDiff, double *v, double *vPred, double *u, double *uPred)
{
    int i, j;
    for (i = 0; i < n; ++i) {
        if (!(bits[i] & ACT)) continue;
        for (j = 0; j < 3; ++j) {
            if (bits[i] & GAS)
                vPred[3 * i + j] = v[3 * i + j] * dvPredFacOne + a[3 * i + j] * dvPredFacTwo;
            v[3 * i + j] = v[3 * i + j] * dvFacOne + a[3 * i + j] * dvFacTwo;
        }
        if ((bits[i] & GAS) && !bIsothermal) {
            double ud = uDotPdV[i] + uDotAV[i] + uDotDiff[i];
            uPred[i] = u[i] + ud * duPredDelta;
            u[i] = u[i] + ud * duDelta;
        }
    }
}

/* pkdKickVpred (pkd.c:6489-6552), NOCOOLING: gas particles only.  Returns
 * the number of negative uPred (the reference asserts uPred >= 0). */
int orc_kick_vpred(int n, const unsigned *bits, double dvFacOne, double dvFacTwo,
                   double duDelta, int bIsothermal, const double *a, const double *uDotPdV,
                   const double *uDotAV, const double *uDotDiff, double *vPred, double *uPred)
{
    int i, j, bad = 0;
    for (i = 0; i < n; ++i) {
        if (!(bits[i] & GAS)) continue;
        for (j = 0; j < 3; ++j)
            vPred[3 * i + j] = vPred[3 * i + j] * dvFacOne + a[3 * i + j] * dvFacTwo;
        if (!bIsothermal) {
            uPred[i] = uPred[i] + (uDotPdV[i] + uDotAV[i] + uDotDiff[i]) * duDelta;
            if (!(uPred[i] >= 0.0)) ++bad;
        }
    }
    return bad;
}

/* pkdDrift (pkd.c:3686-3778): r += dDelta v, wrapped once into
 * [c - L/2, c + L/2) when periodic.  Returns the number of coordinates
 * still outside (the reference asserts none). */
int orc_drift(int n, double dDelta, const double *fCenter, int bPeriodic, const double *fPeriod,
              const double *v, double *r)
{
    int i, j, bad = 0;
    for (i = 0; i < n; ++i)
        for (j = 0; j < 3; ++j) {
            double *x = &r[3 * i + j];
            *x += dDelta * v[3 * i + j];
            if (bPeriodic) {
                if (*x >= fCenter[j] + 0.5 * fPeriod[j]) *x -= fPeriod[j];
                if (*x < fCenter[j] - 0.5 * fPeriod[j]) *x += fPeriod[j];
                if (!(*x >= fCenter[j] - 0.5 * fPeriod[j] && *x < fCenter[j] + 0.5 * fPeriod[j]))
                    ++bad;
            }
        }
    return bad;
}

/* pkdInitDt (pkd.c:4817-4826) */
void orc_init_dt(int n, const unsigned *bits, double dDelta, double *dt)
{
    int i;
    for (i = 0; i < n; ++i)
        if (bits[i] & ACT) dt[i] = dDelta;
}

/* pkdGravStep (pkd.c:4627-4640).  Returns the number of active particles
 * with dtGrav <= 0 (the reference asserts none). */
int orc_grav_step(int n, const unsigned *bits, double dEta, const double *dtGrav, double *dt)
{
    int i, bad = 0;
    for (i = 0; i < n; ++i) {
        double d;
        if (!(bits[i] & ACT)) continue;
        if (!(dtGrav[i] > 0.0)) {
            ++bad;
            continue;
        }
        d = dEta / sqrt(dtGrav[i]);
        if (d < dt[i]) dt[i] = d;
    }
    return bad;
}

/* pkdAccelStep (pkd.c:4643-4686) */
void orc_accel_step(int n, const unsigned *bits, double dEta, double dVelFac, double dAccFac,
                    int bEpsAcc, int bSqrtPhi, const double *v, const double *a,
                    const double *fSoft, const double *fBall2, const double *fPot, double *dt)
{
    int i, j;
    for (i = 0; i < n; ++i) {
        double vel = 0, acc = 0, dT;
        if (!(bits[i] & ACT)) continue;
        for (j = 0; j < 3; ++j) {
            vel += v[3 * i + j] * v[3 * i + j];
            acc += a[3 * i + j] * a[3 * i + j];
        }
        vel = sqrt(vel) * dVelFac;
        (void)vel;
        acc = sqrt(acc) * dAccFac;
        dT = DBL_MAX;
        if (bEpsAcc && acc > 0) {
            if (bits[i] & GAS) dT = dEta * sqrt(sqrt(0.25 * fBall2[i]) / acc);
            else dT = dEta * sqrt(fSoft[i] / acc);
        }
        if (bSqrtPhi && acc > 0) {
            double dtemp = dEta * 3.5 * sqrt(dAccFac * fabs(fPot[i])) / acc;
            if (dtemp < dT) dT = dtemp;
        }
        if (dT < dt[i]) dt[i] = dT;
    }
}

/* pkdDensityStep (pkd.c:4689-4701) */
void orc_density_step(int n, const unsigned *bits, double dEta, double dRhoFac,
                      const double *rho, double *dt)
{
    int i;
    for (i = 0; i < n; ++i) {
        double dT;
        if (!(bits[i] & ACT)) continue;
        dT = dEta / sqrt(rho[i] * dRhoFac);
        if (dT < dt[i]) dt[i] = dT;
    }
}

/* pkdSphStep (pkd.c:5901-5979), NOCOOLING, without THERMALCOND/TWOPHASE/
 * UNONCOOL: the Courant estimate of the force pass (dtNew, then reset to
 * FLT_MAX), the extrapolated Courant step for heating gas, the adiabatic-
 * cooling limit and the diffusion limit; *dtMinGas = the smallest dt of the
 * gas particles, active or not. */
void orc_sph_step(int n, const unsigned *bits, double dCosmoFac, double dEtaCourant,
                  double dEtauDot, double duMinDt, double dDiffCoeff, double dEtaDiffusion,
                  double dResolveJeans, const double *fBall2, const double *c,
                  const double *u, const double *uDotPdV, const double *uDotAV,
                  const double *uDotDiff, const double *rho, const double *diff, double *dtNew,
                  double *dt, double *dtMinGas)
{
    int i;
    *dtMinGas = DBL_MAX;
    for (i = 0; i < n; ++i) {
        if (!(bits[i] & GAS)) continue;
        if (bits[i] & ACT) {
            double ph = sqrt(0.25 * fBall2[i]);
            double dT = dtNew[i], uTotDot;
            dtNew[i] = FLT_MAX;
            uTotDot = uDotPdV[i] + uDotAV[i] + uDotDiff[i];
            if (uTotDot > 0) {
                /* pkdDtFacCourant, pkd.c:5893-5895; GAMMA_NONCOOL pkd.h:228 */
                double dtExtrap = dEtaCourant * dCosmoFac * 2 / 1.6
                    * sqrt(fBall2[i] * 0.25 / (4 * (c[i] * c[i] + (5. / 3.) * uTotDot * dt[i])));
                if (dtExtrap < dT) dT = dtExtrap;
            }
            if (dEtauDot > 0.0 && uDotPdV[i] < 0.0) {
                /* pkdPoverRhoFloorJeansParticle pkd.c:5668-5688; PONRHOFLOOR 0 and
                 * GAMMA_JEANS 2 (pkd.h:66,227) */
                double floorJ = 0.25 * fBall2[i] * dResolveJeans * rho[i];
                double uEff = 0 + floorJ / (2.0 - 1) + u[i];
                double dTu;
                if (uEff < duMinDt) uEff = duMinDt;
                dTu = dEtauDot * uEff / fabs(uDotPdV[i]);
                if (dTu < dT) dT = dTu;
            }
            if (diff[i] > 0 && dDiffCoeff > 0) {
                double dTD = dEtaDiffusion * ph * ph / (dDiffCoeff * diff[i]);
                if (dTD < dT) dT = dTD;
            }
            if (dT < dt[i]) dt[i] = dT;
        }
        if (dt[i] < *dtMinGas) *dtMinGas = dt[i];
    }
}

/* pkdOneParticleDtToRung (pkd.c:4704-4726) */
int orc_one_dt_to_rung(int iRung, double dDelta, double dt)
{
    double dSteps = dDelta / dt;
    int iSteps = (int)floor(dSteps), r = iRung;
    if (fmod(dDelta, dt) == 0.0) iSteps--;   /* an integer boundary goes to the lower rung */
    if (iSteps < 0) iSteps = 0;
    while (iSteps) {
        ++r;
        iSteps >>= 1;
    }
    return r;
}

/* pkdDtToRung (pkd.c:4729-4809) without the diagnostic print: particles on
 * rung >= iRung get their rung from dt (bAll) or iRung / iRung+1; returns
 * the largest rung, with *nMaxRung particles on it and *iMaxRungIdeal the
 * rung the largest dt would want.  Returns -1 (and nothing else is
 * defined) if a particle breaks the reference's assertions. */
int orc_dt_to_rung(int n, const unsigned *bits, int iRung, double dDelta, int iMaxRung, int bAll,
                   const double *fSoft, const double *dt, int *rung, double *dtOld,
                   int *nMaxRung, int *iMaxRungIdeal)
{
    int i, iMaxRungOut = 0, nMax = 0, ideal = 0;
    for (i = 0; i < n; ++i) {
        if (rung[i] >= iRung) {
            if (!(bits[i] & ACT)) return -1;
            if (bAll) {
                int t;
                if (!(fSoft[i] > 0.0) || !(dt[i] > 0.0) || !(dDelta / dt[i] < 2.1e9)) return -1;
                t = orc_one_dt_to_rung(iRung, dDelta, dt[i]);
                if (t >= ideal) ideal = t + 1;
                if (t >= iMaxRung) t = iMaxRung - 1;
                rung[i] = t;
                dtOld[i] = dt[i];
            } else {
                rung[i] = dDelta <= dt[i] ? iRung : iRung + 1;
            }
        }
        if (rung[i] > iMaxRungOut) {
            iMaxRungOut = rung[i];
            nMax = 1;
        } else if (rung[i] == iMaxRungOut) {
            ++nMax;
        }
    }
    *nMaxRung = nMax;
    *iMaxRungIdeal = ideal;
    return iMaxRungOut;
}
