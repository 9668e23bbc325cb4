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
