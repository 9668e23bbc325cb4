/*
 * gg_msr.h -- C host sequences over the gasoline_amd C ABI.
 *
 * These are the per-process parts of gasoline's master-level force steps,
 * written against include/gasoline_amd.h the way gasoline's PST leaves
 * would call it: msrBuildTree + msrGravity (master.c:4249, 5869) and the
 * SPH block of msrSph (master.c:9258-9362) for a step in which every gas
 * particle is active.
 */
#ifndef GG_MSR_H
#define GG_MSR_H

#include "../../include/gasoline_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ggGravityOut {     /* the counters of outGravity (pst.h:540-571) */
    int nActive;
    double dPartSum, dCellSum, dSoftSum, dFlop;
    double dWallclock;            /* seconds, tree build excluded */
} ggGravityOut;

/* Build the gravity tree and compute gravity on all active particles;
 * prints gasoline's "Gravity Calculated" lines (master.c:6149-6189) when
 * bVerbose.  Returns 0 or a GG_E* code. */
int ggMsrGravity(GGPKD pkd, int nBucket, double dTheta, int iOrder, int bVerbose,
                 ggGravityOut *out);

/* SPH force step on the gas: pkdBuildLocal, smooth with DenDVDX, gas
 * pressure, resmooth with SphPressureTermsSym (master.c:9266-9351, the
 * not-FastGas branch).  Returns 0 or a GG_E* code. */
int ggMsrSph(GGPKD pkd, const gg_smf *smf, int nSmooth, int bPeriodic,
             const double *fPeriod, int bVerbose);

#ifdef __cplusplus
}
#endif
#endif
