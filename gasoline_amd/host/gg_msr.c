/*
 * gg_msr.c -- C host sequences over the gasoline_amd C ABI (see gg_msr.h).
 */
#include <stdio.h>
#include <sys/time.h>

#include "gg_msr.h"

static double wallclock(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

#define GG_RET(call)                                                       \
    do {                                                                   \
        int rc_ = (call);                                                  \
        if (rc_ != GG_OK) {                                                \
            fprintf(stderr, "%s failed (%d): %s\n", #call, rc_,            \
                    gg_last_error());                                      \
            return rc_;                                                    \
        }                                                                  \
    } while (0)

int ggMsrGravity(GGPKD pkd, int nBucket, double dTheta, int iOrder, int bVerbose,
                 ggGravityOut *out)
{
    double t0;
    /* msrBuildTree(msr, 0, dMass, 0): all particles, OPEN_JOSH with dTheta */
    GG_RET(gg_build_binary(pkd, nBucket, 1, dTheta, iOrder, 0, 1));
    t0 = wallclock();
    GG_RET(gg_grav_all(pkd, 0, 0, iOrder, 0, iOrder, 0.0, 0.0, 0, 0.0, 0, 0.0, NULL,
                       &out->nActive, &out->dPartSum, &out->dCellSum, &out->dSoftSum,
                       &out->dFlop));
    out->dWallclock = wallclock() - t0;
    if (bVerbose) {
        double dsec = out->dWallclock;
        if (dsec > 0.0)
            printf("Gravity Calculated, Wallclock: %f secs, MFlops:%.1f, Flop:%.3g\n", dsec,
                   out->dFlop / dsec * 1e-6, out->dFlop);
        if (out->nActive > 0)
            printf("dPartAvg:%f dCellAvg:%f dSoftAvg:%f\n", out->dPartSum / out->nActive,
                   out->dCellSum / out->nActive, out->dSoftSum / out->nActive);
    }
    return GG_OK;
}

int ggMsrSph(GGPKD pkd, const gg_smf *smf, int nSmooth, int bPeriodic,
             const double *fPeriod, int bVerbose)
{
    int nSmoothed = 0;
    double t0 = wallclock(), t1, t2, t3;
    GG_RET(gg_build_local(pkd, 32, bPeriodic, fPeriod));
    GG_RET(gg_smooth(pkd, smf, nSmooth, bPeriodic, 0, GG_SMX_DENDVDX, &nSmoothed));
    t1 = wallclock();
    GG_RET(gg_gas_pressure(pkd, smf));
    t2 = wallclock();
    GG_RET(gg_resmooth(pkd, smf, nSmooth, bPeriodic, 1, GG_SMX_SPHPRESSURETERMS));
    t3 = wallclock();
    if (bVerbose)
        printf("SPH: Smooth Active Particles: %d  smooth %f s  pressure %f s  resmooth %f s\n",
               nSmoothed, t1 - t0, t2 - t1, t3 - t2);
    return GG_OK;
}

/* ------------------------------------------------ block time steps */
#define TYPE_ALL (GG_TYPE_GAS | GG_TYPE_DARK | GG_TYPE_STAR)

/* msrDrift (master.c:6306-6420), non-comoving: pkdDrift of every particle,
 * then pkdKickVpred of the gas (vPred += a dDelta, uPred += uDot dDelta). */
static int msr_drift(GGPKD pkd, const ggStepParam *p, double dDelta)
{
    GG_RET(gg_drift(pkd, dDelta, p->fCenter, p->bPeriodic, p->fPeriod));
    if (p->bDoGas) GG_RET(gg_kick_vpred(pkd, 1.0, dDelta, dDelta, 0));
    return GG_OK;
}

int ggMsrSphRung(GGPKD pkd, const ggStepParam *p, int iKickRung)
{
    int nSmoothed = 0;
    /* tree over all gas, the gas on the kicked rungs active */
    GG_RET(gg_active_type_rung(pkd, GG_TYPE_GAS, GG_TYPE_ACTIVE, iKickRung, 1, NULL));
    GG_RET(gg_active_type(pkd, GG_TYPE_GAS, GG_TYPE_TREEACTIVE | GG_TYPE_SMOOTHACTIVE, NULL));
    GG_RET(gg_build_local(pkd, p->nBucketSph, p->bPeriodic, p->fPeriod));
    GG_RET(gg_active_type_rung(pkd, GG_TYPE_GAS, GG_TYPE_ACTIVE, iKickRung, 1, NULL));
    /* not FastGas: smooth every gas particle, marking the neighbours of the
     * active ones (master.c:9307-9320) */
    GG_RET(gg_reset_type(pkd, GG_TYPE_GAS, GG_TYPE_SMOOTHDONE | GG_TYPE_NbrOfACTIVE, NULL));
    GG_RET(gg_active_type(pkd, GG_TYPE_ACTIVE, GG_TYPE_SMOOTHACTIVE, NULL));
    GG_RET(gg_active_type(pkd, GG_TYPE_GAS, GG_TYPE_SMOOTHACTIVE, NULL));
    GG_RET(gg_smooth(pkd, &p->smf, p->nSmooth, p->bPeriodic, 1, GG_SMX_DENDVDX, &nSmoothed));
    GG_RET(gg_active_type(pkd, GG_TYPE_NbrOfACTIVE, GG_TYPE_SMOOTHACTIVE, NULL));
    GG_RET(gg_gas_pressure(pkd, &p->smf));
    GG_RET(gg_resmooth(pkd, &p->smf, p->nSmooth, p->bPeriodic, 1, GG_SMX_SPHPRESSURETERMS));
    return GG_OK;
}

int ggMsrForces(GGPKD pkd, const ggStepParam *p, ggStepState *s, int iKickRung)
{
    int nActive = 0, cur = 0;
    GG_RET(gg_active_mask_rung(pkd, GG_TYPE_ACTIVE, iKickRung, 1, &nActive));
    if (s->nForce < s->nLogCap) {
        if (s->logKickRung) s->logKickRung[s->nForce] = iKickRung;
        if (s->logActive) s->logActive[s->nForce] = nActive;
    }
    s->nForce++;
    if (!nActive) return GG_OK;
    GG_RET(gg_init_accel(pkd));
    if (p->bDoGravity) {
        int n;
        double ps, cs, ss, fl;
        GG_RET(gg_active_mask_rung(pkd, GG_TYPE_ACTIVE, iKickRung, 1, NULL));
        GG_RET(gg_active_type(pkd, TYPE_ALL, GG_TYPE_TREEACTIVE, NULL));
        GG_RET(gg_build_binary(pkd, p->nBucket, 1, p->dTheta, p->iOrder, 0, 1));
        GG_RET(gg_grav_all(pkd, 0, 0, p->iOrder, 0, p->iOrder, 0.0, 0.0, 0, 0.0, 0, 0.0, NULL,
                           &n, &ps, &cs, &ss, &fl));
    }
    if (p->bDoGas) {
        GG_RET(gg_sph_curr_rung(pkd, iKickRung, 1, &cur));
        if (cur) GG_RET(ggMsrSphRung(pkd, p, iKickRung));
    }
    return GG_OK;
}

int ggMsrTopStepKDK(GGPKD pkd, const ggStepParam *p, ggStepState *s, double dStep, double dTime,
                    double dDelta, int iRung, int iKickRung, int iAdjust)
{
    const double h = 0.5 * dDelta;
    iAdjust = iAdjust && p->bAdjust;
    if (iAdjust && p->bDoGas) {      /* msrEmergencyAdjust (master.c:6742) */
        int nUn, ideal, nMax, top;
        GG_RET(gg_active_mask_rung(pkd, GG_TYPE_ACTIVE, iRung, 1, NULL));
        GG_RET(gg_emergency_adjust(pkd, iRung, p->iMaxRung, dDelta, 0.5 * dDelta, &nUn, &ideal,
                                   &nMax, &top));
        s->nEmergency += nUn;
        if (nUn && top > s->iCurrMaxRung) s->iCurrMaxRung = top;
    }
    if (iAdjust && iRung < p->iMaxRung - 1) {
        int nMax, ideal, top;
        GG_RET(gg_active_mask_rung(pkd, GG_TYPE_ACTIVE, iRung, 1, NULL));
        GG_RET(gg_init_dt(pkd, p->dDelta));
        if (p->bGravStep) GG_RET(gg_grav_step(pkd, p->dEta));
        if (p->bAccelStep)
            GG_RET(gg_accel_step(pkd, p->dEta, 1.0, 1.0, p->bDoGravity, p->bEpsAccStep,
                                 p->bSqrtPhiStep, 0.0));
        if (p->bDoGas && p->bSphStep) {
            const double diff = p->smf.dMetalDiffusionCoeff > p->smf.dThermalDiffusionCoeff
                                    ? p->smf.dMetalDiffusionCoeff
                                    : p->smf.dThermalDiffusionCoeff;
            /* msrSphStep (master.c:9181-9199); msrSmoothFcnParam set
             * dtFacDiffusion = 2 dEtaDiffusion (master.c:5615) */
            GG_RET(gg_sph_step(pkd, 1.0, p->dEtaCourant, p->dEtauDot, p->duMinDt, diff,
                               p->smf.dtFacDiffusion / 2.0, p->smf.dResolveJeans,
                               p->bViscosityLimitdt, &s->dtMinGas));
        }
        GG_RET(gg_dt_to_rung(pkd, iRung, dDelta, p->iMaxRung, 1, 1, &nMax, &ideal, &top));
        s->iCurrMaxRung = top;
    }
    /* open kick: vPred and uPred keep the start-of-step values */
    GG_RET(gg_active_mask_rung(pkd, GG_TYPE_ACTIVE, iRung, 0, NULL));
    if (p->bDoGas) GG_RET(gg_active_type(pkd, TYPE_ALL, GG_TYPE_TREEACTIVE | GG_TYPE_SMOOTHACTIVE,
                                         NULL));
    GG_RET(gg_kick(pkd, 1.0, h, 1.0, 0.0, h, 0.0, 0));
    if (s->iCurrMaxRung > iRung) {
        GG_RET(ggMsrTopStepKDK(pkd, p, s, dStep, dTime, h, iRung + 1, iRung + 1, 0));
        dStep += 1.0 / (2 << iRung);
        dTime += h;
        GG_RET(gg_active_mask_rung(pkd, GG_TYPE_ACTIVE, iRung, 0, NULL));
        GG_RET(ggMsrTopStepKDK(pkd, p, s, dStep, dTime, h, iRung + 1, iKickRung, 1));
    } else {
        /* drift everyone in two halves (the GASOLINE branch) and evaluate
         * the forces of rungs >= iKickRung */
        GG_RET(msr_drift(pkd, p, h));
        dTime += h;
        GG_RET(gg_active_mask_rung(pkd, GG_TYPE_ACTIVE, iRung, 0, NULL));
        GG_RET(msr_drift(pkd, p, h));
        dTime += h;
        GG_RET(ggMsrForces(pkd, p, s, iKickRung));
        dTime -= h;
    }
    /* closing kick: vPred and uPred to the end of the step */
    GG_RET(gg_active_mask_rung(pkd, GG_TYPE_ACTIVE, iRung, 0, NULL));
    GG_RET(gg_kick(pkd, 1.0, h, 1.0, h, h, h, 0));
    (void)dStep;
    (void)dTime;
    return GG_OK;
}
