/*
 * gasoline_amd.h -- C ABI of the MI355X force hot path for gasoline.
 *
 * The library (gasoline_amd/libgasoline_amd.so) replaces, for one process,
 * the leaf calls that gasoline's PST layer makes into its per-processor code:
 *
 *   pstBuildTree leaf  (pst.c:2955-2973) -> pkdBuildBinary / pkdBuildLocal
 *   pstGravity leaf    (pst.c:3305-3336) -> pkdGravAll
 *   pstSmooth leaf     (pst.c:3012-3039) -> smInitialize, smSmooth, smFinish
 *   pstReSmooth leaf   (pst.c:3169-3196) -> smInitialize, smReSmooth, smFinish
 *   pstGetGasPressure                    -> pkdGasPressure
 *
 * A GGPKD handle plays the role of gasoline's PKD: it owns a device-resident
 * particle store (structure of arrays, FP64), the trees and the interaction
 * and neighbour lists.  Fields are exchanged in the caller's particle order
 * (the order of the arrays given to gg_put); the trees keep their own
 * locality orders internally, so unlike pkdBuildBinary the caller's order is
 * never permuted.
 *
 * Conventions: plain pointers and sizes; every entry point returns 0 on
 * success and a negative GG_E* code on error, with a message available from
 * gg_last_error() (gasoline itself asserts; errors here are returned).
 * All calls on one GGPKD are issued on that handle's HIP stream, in order.
 * gg_build_local, gg_smooth, gg_gas_pressure and gg_resmooth may return
 * before their last kernels have finished (each still waits for what its
 * host side must read: counts, error words); every call that hands results
 * to the host (gg_get, gg_grav_all's counters, gg_timings, ...) waits for
 * the stream first.  A caller that reads the store's device memory on a
 * stream of its own synchronises the handle's stream (or the device)
 * before it does.  The other calls are synchronous with respect to the
 * host unless stated otherwise.
 */
#ifndef GASOLINE_AMD_H
#define GASOLINE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gg_pkd_s *GGPKD;

/* error codes */
#define GG_OK          0
#define GG_EINVAL     -1   /* bad argument */
#define GG_ENOMEM     -2   /* device or host allocation failed */
#define GG_EHIP       -3   /* HIP runtime error */
#define GG_ESTATE     -4   /* call out of order (e.g. no tree built) */
#define GG_EUNSUP     -5   /* feature outside this library's scope */

/* Particle fields (one FP64 value per particle unless noted).  Names follow
 * gasoline's PARTICLE (pkd.h:90-225). */
enum gg_field {
    GG_R = 0,        /* r[3], x y z interleaved (3 doubles) */
    GG_V,            /* v[3] */
    GG_VPRED,        /* vPred[3] */
    GG_A,            /* a[3] (accumulated by gravity and hydro) */
    GG_MASS,         /* fMass */
    GG_SOFT,         /* fSoft */
    GG_POT,          /* fPot */
    GG_DTGRAV,       /* dtGrav: max (m_i+m_j) b, a 1/dt^2 estimate */
    GG_U,            /* u */
    GG_UPRED,        /* uPred */
    GG_FBALL2,       /* fBall2 = (2h)^2 */
    GG_DENSITY,      /* fDensity */
    GG_DIVV,         /* divv */
    GG_CURLV,        /* curlv[3] */
    GG_DVDS,         /* dvds */
    GG_FDIVV_T,      /* fDivv_t */
    GG_BALSARA,      /* BalsaraSwitch */
    GG_DIFF,         /* diff */
    GG_POVERRHO2,    /* PoverRho2 = (P/rho)/rho */
    GG_C,            /* c, sound speed */
    GG_DT,           /* dt, Courant time step (in/out of gas pressure) */
    GG_UDOTPDV,      /* uDotPdV */
    GG_UDOTAV,       /* uDotAV */
    GG_UDOTDIFF,     /* uDotDiff */
    GG_MUMAX,        /* mumax */
    GG_DTNEW,        /* dtNew */
    GG_IACTIVE,      /* iActive type bits (uint32 per particle, pkd.h:288-329) */
    GG_IRUNG,        /* iRung (int32 per particle) */
    GG_DTOLD,        /* dtOld */
    GG_FWEIGHT,      /* fWeight: flops of the particle's bucket in the last gravity
                        pass (pkdBucketWeight, pkd.c:2851-2861), for load balance */
    GG_NFIELDS
};

/* iActive bits used by this path (values as pkd.h:288-308) */
#define GG_TYPE_ACTIVE        (1u << 0)
#define GG_TYPE_TREEACTIVE    (1u << 1)
#define GG_TYPE_SMOOTHACTIVE  (1u << 2)
#define GG_TYPE_SMOOTHDONE    (1u << 3)
#define GG_TYPE_NbrOfACTIVE   (1u << 5)
#define GG_TYPE_GAS           (1u << 8)
#define GG_TYPE_DARK          (1u << 9)
#define GG_TYPE_STAR          (1u << 10)

/* Arithmetic of the interaction kernels. */
#define GG_PREC_FP64  0   /* plain FP64 throughout */
#define GG_PREC_FP32  1   /* FP32 on cell-relative coordinates, compensated sums */

/* ---------------------------------------------------------------- context */

/* Create a handle on HIP device `device` for up to nStore particles.
 * Replaces pkdInitialize (pkd.c:150-296) for the fields this path uses. */
int gg_initialize(GGPKD *ppkd, int device, int nStore);

/* Release every device and host buffer of the handle (pkdFinish). */
int gg_finish(GGPKD pkd);

/* Message of the last error on this thread. */
const char *gg_last_error(void);

/* The periodic box (pkdInitialize's fPeriod, pkd.c:150): gg_grav_all with
 * bPeriodic uses it.  Every side must be positive. */
int gg_set_period(GGPKD pkd, const double *fPeriod);

/* Number of particles currently in the store (pkdLocal). */
int gg_set_local(GGPKD pkd, int nLocal);
int gg_local(GGPKD pkd);

/* Copy a field in (gg_put) or out (gg_get), in the caller's particle order.
 * src/dst are host pointers when bDevice == 0, device pointers otherwise.
 * gg_put from a host source returns with the copy done.  From a device
 * source, it queues the copy on the handle's stream, ahead of the handle's
 * later work: the source must hold its values when gg_put is called (work
 * on other streams that writes it has finished), and keep them, and stay
 * allocated, until a call that reads back to the host (gg_get,
 * gg_timings ...) has returned.
 * Stands in for the particle I/O the reference does through pStore
 * (pkdReadTipsy pkd.c:297, pkdOutVector outtype.c:806). */
int gg_put(GGPKD pkd, int field, const void *src, int bDevice);
int gg_get(GGPKD pkd, int field, void *dst, int bDevice);

/* One member of the caller's particle struct: the field it maps to and its
 * byte offset (offsetof(PARTICLE, r) ...).  FP64 members (FLOAT is double,
 * floattype.h:19), iActive and iRung as 32-bit ints. */
typedef struct gg_aos_field {
    int field;          /* GG_R ... GG_IACTIVE */
    int pad;
    size_t offset;      /* bytes from the start of a particle */
} gg_aos_field;

/* Copy fields in from (gg_put_aos) or out to (gg_get_aos) an array of
 * structs of nLocal particles, `stride` bytes apart (sizeof(PARTICLE)):
 * gasoline's pStore, whose PARTICLE layout depends on compile flags
 * (pkd.h:90-225), used as it is.  One pass over the particles per chunk,
 * staged through pinned host memory, the copies of one chunk overlapping
 * the packing of the next.  gg_get_aos writes only the listed members.
 * The pack/unpack loops a caller of pkdGravAll/smSmooth would otherwise
 * write around each call. */
int gg_put_aos(GGPKD pkd, const void *base, size_t stride, const gg_aos_field *f, int nf);
int gg_get_aos(GGPKD pkd, void *base, size_t stride, const gg_aos_field *f, int nf);

/* ------------------------------------------------------- Tipsy files */
/* The header of a Tipsy file (struct dump, tipsydefs.h:38-45): bStandard
 * selects XDR (big-endian, 32 bytes with a pad word) over the host's struct.
 * Needs no device. */
int gg_tipsy_header(const char *path, int bStandard, double *dTime, int *nBodies, int *nGas,
                    int *nDark, int *nStar);
/* pkdReadTipsy (pkd.c:297-640) of a whole file into the store (nLocal =
 * nbodies <= nStore): gas, dark and star records in that order set r,
 * v = dvFac vel (and vPred for gas), fMass, fSoft = eps, fPot = phi; gas
 * also fDensity = rho and u = uPred = dTuFac temp.  iActive = the type bit
 * only (TYPE_GAS, TYPE_DARK, TYPE_STAR), iRung 0, dt = dtNew = FLT_MAX,
 * fBall2 0, as pkdReadTipsy leaves them (the caller then sets activity as
 * msrActiveType does).  Metals and star formation times are not carried.
 * *pdTime receives the header's time. */
int gg_read_tipsy(GGPKD pkd, const char *path, int bStandard, double dvFac, double dTuFac,
                  double *pdTime);
/* pkdWriteTipsy (pkd.c:1628-1800) with the header msrWriteTipsy writes:
 * gas, dark and star records (a particle's type by its bits, dark first as
 * the reference tests them), each type in store order; vel = dvFac v, gas
 * temp = duTFac u (the non-cooling gas models), eps = fSoft; metals and
 * formation times are written as 0. */
int gg_write_tipsy(GGPKD pkd, const char *path, int bStandard, double dTime, double dvFac,
                   double duTFac);

/* An output array as msrOutArray / msrOutVector write it in ASCII
 * (iBinaryOutput 0; master.c:5420-5545, pkdOutVector outtype.c:946-960):
 * the particle count on the first line, then one value per line with
 * "%.14g", a vector field component by component (all x, all y, all z).
 * The caller picks the field its output type shows (fDensity for .den, fPot
 * for .pot, a for .acc ...); values in store order. */
int gg_out_array(GGPKD pkd, const char *path, int field);

/* Select FP32 or FP64 interaction arithmetic (default GG_PREC_FP32). */
int gg_set_precision(GGPKD pkd, int prec);

/* ------------------------------------------------------------ gravity */

/* Build the gravity tree: spatial bisection at the midpoint of the squeezed
 * bounding box's longest axis, buckets of <= nBucket particles, FP64 c.o.m.,
 * mass-weighted softening, multipoles to iOrder and OPEN_JOSH opening radii.
 * Replaces pkdBuildBinary (pkd.h:794, pkd.c:2627-2723).  iOpenType must be 1
 * (OPEN_JOSH, opentype.h:5) or 0 (minimal Bmax); bGravity must be 1. */
int gg_build_binary(GGPKD pkd, int nBucket, int iOpenType, double dCrit,
                    int iOrder, int bTreeActiveOnly, int bGravity);

/* Gravity on all active particles: per-bucket interaction lists from the
 * threaded-tree walk (pkdBucketWalk walk.c:306-439), then particle-particle,
 * softened-cell and Newtonian-cell interactions and the intra-bucket pairs
 * (pkdBucketInteract grav.c:23-249).  a, fPot and dtGrav of active particles
 * are ACCUMULATED onto their stored values (gasoline zeroes them first in
 * msrInitAccel).  Replaces pkdGravAll (pkd.h:797-801, pkd.c:2869-3060) with
 * the same signature.  bPeriodic: every bucket also walks the replicas
 * -nReps..nReps of the box set by gg_set_period (walk.c:325-438; the own
 * bucket enters a replica's PP list) and, with bEwald, the Ewald correction
 * of the root's multipoles to order iEwOrder is added (pkdEwaldInit
 * ewald.c:180, pkdBucketEwald ewald.c:15; real-space images within fEwCut
 * boxes, k-space |h| <= fEwhCut; cubic box).  A periodic tree must be a
 * single system.  bComove and bDoSun must be 0.  Counters follow
 * pkd.c:2946-2949 and the flop models of grav.c:246-247 and ewald.c:175.
 * fWeight of every active particle is set to its bucket's flops (the
 * interaction and Ewald counts of that bucket), as pkdBucketWeight does
 * (pkd.c:2851-2861, 2998); inactive particles keep theirs. */
int gg_grav_all(GGPKD pkd, int nReps, int bPeriodic, int iOrder, int bEwald,
                int iEwOrder, double fEwCut, double fEwhCut, int bComove,
                double dRhoFac, int bDoSun, double dSunSoft, double *aSun,
                int *nActive, double *pdPartSum, double *pdCellSum,
                double *pdSoftSum, double *pdFlop);

/* Copy the interaction lists of the last gg_grav_all to the host, for
 * parity checks.  Lists are per active bucket b (buckets in tree order):
 * bucket_node[b], then for kind k in {PP, soft cell, Newtonian cell}
 * off[3*b+k] .. off[3*b+k] + cnt[3*b+k] index into ent[].  PP entries are
 * particle slots in the tree's storage order, cell entries are node ids.
 * Call with ent == NULL to get the sizes (*nBuckets, *nEnt). */
int gg_grav_lists(GGPKD pkd, int *nBuckets, int64_t *nEnt, int *bucket_node,
                  int64_t *off, int *cnt, int *ent);

/* Checksums of the lists that the fused walk + interaction kernel (the walk
 * gg_grav_all runs) evaluates, for parity checks of that kernel's own walk:
 * for each active bucket b (the order of gg_grav_lists) and list kind k
 * (0 PP, 1 softened cell, 2 Newtonian cell), sums[9b+3k] = the number of
 * members, sums[9b+3k+1] = their sum (particle slots or node ids, as
 * gg_grav_lists numbers them) and sums[9b+3k+2] = the sum of mix(member),
 * the sums modulo 2^64, where mix is
 * splitmix64's finaliser (z += 0x9E3779B97F4A7C15; z = (z ^ z>>30) *
 * 0xBF58476D1CE4E5B9; z = (z ^ z>>27) * 0x94D049BB133111EB; z ^ z>>31).
 * The isolated walk (nReps 0); a, fPot and dtGrav are not touched.
 * GG_EUNSUP if this tree is walked by the listed walk.  No counterpart in
 * the reference. */
int gg_grav_walk_checksums(GGPKD pkd, int iOrder, uint64_t *sums);

/* Tree shape of the last gg_build_binary, for parity checks: perm[k] is
 * the caller index of storage slot k; per node ints[6*i..] = iDim, pLower,
 * pUpper, iLower, iUpper (threaded next), 0 and dbls[8*i..] = fMass, fSoft,
 * r[3], fOpen2, Bmax, fSplit. */
int gg_tree_info(GGPKD pkd, int *nNodes, int *iRoot, int *perm, int *ints,
                 double *dbls);

/* FP64 direct-sum accelerations and potentials of the particles target[k]
 * (caller indices) from every other particle, with the particle-particle
 * spline softening twoh = h_i + h_j (grav.c:81-108) and G = 1: the
 * reference force against which the tree's force errors are stated.  No
 * counterpart in the reference (it has no direct-sum path); a[3*k..],
 * pot[k] are host arrays. */
int gg_direct_grav(GGPKD pkd, int nTarget, const int *target, double *a, double *pot);

/* The periodic counterpart of gg_direct_grav in the cubic box set by
 * gg_set_period (side L): FP64 accelerations and potentials of target[k]
 * from every other particle and all images of every particle, Ewald-summed
 * (split alpha = 6 / L, both sums complete to 1e-16) with the uniform
 * background that removes the k = 0 mode, and the spline softening of the
 * images nearer than twoh = h_i + h_j.  This is the field whose expansion
 * about the root's moments pkdBucketEwald (ewald.c:15-178) evaluates, so it
 * is the reference force of the periodic tree + Ewald forces.  pot includes
 * the target's own images, m_i xi (xi L = 2.8372974795, the simple cubic
 * lattice's Madelung constant).  No counterpart in the reference; host
 * arrays; GG_EINVAL without a cubic period. */
int gg_direct_ewald(GGPKD pkd, int nTarget, const int *target, double *a, double *pot);

/* The largest nTarget of gg_direct_grav_noalloc. */
#define GG_DIRECT_NOALLOC_MAX 4096

/* gg_direct_grav without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  target
 * (nTarget caller indices), a (3 per target) and pot (1 per target) are
 * DEVICE pointers; the call only enqueues its kernels on the handle's
 * stream and returns (order later work after it, or synchronize, before
 * reading a and pot).  The same kernels as gg_direct_grav, so the results
 * are the same bit for bit.  The partial sums live in a fixed device
 * array of 8 MiB (64 source shares x 4 sums x GG_DIRECT_NOALLOC_MAX
 * targets) that each GPU the library runs on holds until the process ends;
 * calls that overlap on one GPU (other streams, threads or handles) race
 * on it.  A target outside 0..nLocal-1 gives NaN.  GG_EINVAL if nTarget
 * exceeds GG_DIRECT_NOALLOC_MAX. */
int gg_direct_grav_noalloc(GGPKD pkd, int nTarget, const int *target, double *a, double *pot);

/* ----------------------------------------------------------------- SPH */

/* Smoothing parameters: the subset of gasoline's SMF (smoothfcn.h:22-130)
 * the density and pressure-force operators read. */
typedef struct gg_smf {
    double H, a;
    double alpha, beta, gamma;
    double dtFacCourant, dtFacDiffusion;
    double dMetalDiffusionCoeff, dThermalDiffusionCoeff;
    double dResolveJeans;
    int bCannonical, iViscosityLimiter, bConstantDiffusion, pad;
} gg_smf;

/* Smooth operator ids, numbered as enum smx_smoothtype (smoothfcn.h:147-170)
 * of a -DGASOLINE build,
 * for the three this library implements. */
#define GG_SMX_DENSITY           1
#define GG_SMX_SPHPRESSURETERMS  21
#define GG_SMX_DENDVDX           22

/* Build the SPH search structure over the gas particles (TYPE_GAS, or all
 * particles if none is flagged).  Replaces pkdBuildLocal (pkd.h:795,
 * pkd.c:2732-2848); the neighbour sets found below do not depend on the
 * tree's shape.  fPeriod[3] is the box when bPeriodic. */
int gg_build_local(GGPKD pkd, int nBucket, int bPeriodic, const double *fPeriod);

/* kNN smooth: for every smooth-active particle, the nSmooth nearest
 * neighbours by FP64 distance (minimum periodic image), fBall2 = largest
 * squared distance, then the operator iSmoothType (GG_SMX_DENSITY or
 * GG_SMX_DENDVDX) over that list.  Replaces smInitialize + smSmooth + smFinish
 * (smooth.h:616-617,638; smooth.c:1439-1713) for those operators.
 * bSymmetric selects the operator as smInitialize does (smooth.c:101-106,
 * 272-279): SMX_DENSITY runs Density (0, smoothfcn.c:239-256) or DensitySym
 * (1, smoothfcn.c:258-274: each pair adds half of the list owner's kernel to
 * both ends; particles of the structure that are not smooth-active keep
 * their fDensity plus what their neighbours add).  SMX_DENDVDX runs DenDVDX
 * either way.  Flags are only set: TYPE_SMOOTHDONE on the smoothed
 * particles, TYPE_NbrOfACTIVE by DenDVDX (smoothfcn.c:3424,3487); the caller
 * resets them as msrResetType does (master.c:9279).
 * *pnSmoothed receives smf->nSmoothed. */
int gg_smooth(GGPKD pkd, const gg_smf *smf, int nSmooth, int bPeriodic,
              int bSymmetric, int iSmoothType, int *pnSmoothed);

/* SPH bucket-size autotuner (no counterpart in the reference; the tiling
 * the reference fixes with nBucket): for each candidate nBucket (1..32)
 * build the structure (gg_build_local) and run the DenDVDX pass (gg_smooth,
 * iSmoothType GG_SMX_DENDVDX, not symmetric) twice from the caller's fBall2,
 * time the device work of the second run, keep the fastest, and finish with
 * a build and a pass at it: the structure and every output are those of
 * gg_build_local(best) + gg_smooth.  ms[k] (may be NULL) gets candidate k's
 * milliseconds, *pBest the chosen nBucket. */
int gg_tune_sph(GGPKD pkd, const gg_smf *smf, int nSmooth, int bPeriodic, const double *fPeriod,
                const int *cand, int nCand, double *ms, int *pBest);

/* Lanes per particle of the SPH pair kernels, DenDVDX and the symmetric force
 * pass (no counterpart in the reference, which runs one particle per thread):
 * 4 (the default), 8 or 16 lanes share a particle's neighbour list and sum
 * their parts in a fixed order.  The results agree to FP64 rounding between
 * settings and are bit-identical between runs at one setting.  GG_EINVAL for
 * other values. */
int gg_set_sph_lanes(GGPKD pkd, int lanes);

/* Autotuner of the lanes per particle on the structure already built (no
 * counterpart in the reference).  For each candidate (4, 8 or 16) it runs
 * twice, from the caller's fBall2, a, uDot terms, mumax, dtNew and dt each
 * time:
 *   gg_smooth(GG_SMX_DENDVDX, not symmetric) + gg_gas_pressure
 *   + gg_resmooth(GG_SMX_SPHPRESSURETERMS).
 * It times the DenDVDX and force-pass kernels of the second run and keeps
 * the fastest.  It finishes with one such sequence at the winner, so every
 * output is that of the sequence run once at gg_set_sph_lanes(best).
 * ms[k] (may be NULL) gets candidate k's milliseconds; *pBest gets the
 * choice. */
int gg_tune_sph_lanes(GGPKD pkd, const gg_smf *smf, int nSmooth, int bPeriodic, const int *cand,
                      int nCand, double *ms, int *pBest);

/* Ball-gather re-smooth at each particle's stored fBall2 (d2 <= fBall2) and
 * the operator iSmoothType: GG_SMX_SPHPRESSURETERMS (bSymmetric=1,
 * SphPressureTermsSym smoothfcn.c:2276-2385), evaluated in gather form over
 * the neighbour and reverse-neighbour lists, so no scatter or atomics are
 * used; or GG_SMX_DENDVDX (DenDVDX over the ball lists, as msrInitSph calls
 * it, master.c:9044).  Replaces smInitialize + smReSmooth + smFinish
 * (smooth.h:636; smooth.c:1716-1749) for those operators. */
int gg_resmooth(GGPKD pkd, const gg_smf *smf, int nSmooth, int bPeriodic,
                int bSymmetric, int iSmoothType);

/* Adiabatic equation of state: PoverRho2, c and the Courant dt update for
 * gas particles.  Replaces pkdGasPressure (pkd.h:912, pkd.c:5750-5785)
 * without cooling. */
int gg_gas_pressure(GGPKD pkd, const gg_smf *smf);

/* The kNN search tries first a ball of slack * fBall2 around a particle
 * whose stored fBall2 is positive (a hint: the previous step's value, or a
 * proposal of the h-proposal network), widening it if it holds fewer than
 * nSmooth particles; the result does not depend on it.  Default 1.15; an
 * accurate hint does with less.  slack = 0 ignores the stored fBall2: every
 * search starts cold, from the density of the particle's tree cell.  No
 * counterpart in the reference, whose search starts from the bucket's own
 * particles (smooth.c:1204-1305). */
int gg_set_seed_slack(GGPKD pkd, double slack);

/* Permute the particle store into the order of the SPH structure built by
 * gg_build_local: its particles in Morton order first, then the others in
 * their previous order; every stored field and the h-proposal history move
 * with their particles.  pkdBuildLocal reorders pStore in place the same way
 * while it builds (pkd.c:2732-2848); here it is a separate call so that a
 * caller keeping its own indices can skip it.  Afterwards the gathers and
 * scatters between the structure and the store read consecutive memory.
 * perm (host, nLocal ints, may be null) receives for each new index k the
 * previous index of the particle now there.  The gravity tree is dropped
 * (build it again before gg_grav_all). */
int gg_reorder(GGPKD pkd, int *perm);

/* Neighbour lists of the last gg_smooth, for parity checks: nbr[i*nSmooth+k]
 * in caller order.  The order within a list is the search's own (the
 * reference's is its priority-queue order); the set is exact. */
int gg_smooth_lists(GGPKD pkd, int nSmooth, int *nbr);

/* Counters of the last gg_smooth / gg_resmooth: s[0] kNN queries that the
 * bucket-group search handed to the single-query search, s[1] particles whose
 * nSmooth-th and (nSmooth+1)-th distances tie (their neighbour set is not
 * unique; gasoline's priority queue keeps the one met first), s[2] entries
 * of the force pass's extras lists, s[3] smooth-active particles; search
 * diagnostics of the bucket-group kNN: s[4] groups too large to search
 * (more than 192 buckets or 4096 particles), s[5] queries with fewer than nSmooth candidates in
 * the group radius, s[6] queries whose radius search did not settle, s[7]
 * largest number of buckets in a group; search work of the last gg_smooth:
 * s[8] trial balls beyond the first (widened or narrowed), s[9] candidate
 * particles whose distance was computed.  n <= 10. */
int gg_smooth_stats(GGPKD pkd, int64_t *s, int n);

/* --------------------------------------------- ensembles, graphs, tuning */

/* Many independent systems in one store: system s is caller indices
 * [sysOff[s], sysOff[s+1]), sysOff[0] = 0, sysOff[nSys] = the particle count.
 * One tree per system, as gg_build_binary would build it alone (no system
 * sees another), walked and evaluated by gg_grav_all in one launch.  No
 * counterpart in the reference, which runs such systems as separate jobs. */
int gg_build_forest(GGPKD pkd, int nSys, const int *sysOff, int nBucket, int iOpenType,
                    double dCrit, int iOrder);
/* Capture the device work of one isolated gg_grav_all (nReps 0, no Ewald)
 * for the current tree in a HIP graph (kernel launches only), and replay it:
 * gg_grav_replay adds to a, fPot and dtGrav and sets fWeight exactly as
 * gg_grav_all does and returns the same counters.  A rebuild invalidates
 * the graph.  GG_EUNSUP if the tree needs the listed walk (gg_grav_all
 * handles it) or if the handle has a periodic box (gg_set_period): the
 * periodic step is gg_grav_all's. */
int gg_grav_graph(GGPKD pkd, int iOrder);
int gg_grav_replay(GGPKD pkd, int *nActive, double *pdPartSum, double *pdCellSum,
                   double *pdSoftSum, double *pdFlop);
/* How the fused walk + interaction kernel of gg_grav_all (and of a graph
 * captured afterwards) is scheduled: 0 (default) one workgroup per active
 * bucket, dealt by the hardware in tree order; 1 a persistent grid that
 * fills the chip once, whose workgroups take buckets from a queue: after a
 * first evaluation of a tree the queue holds its buckets heaviest first (by
 * gasoline's flop model on that evaluation's list sizes, grouped by binary
 * exponent), so the longest buckets start first across systems of uneven
 * N; 2 a persistent grid in which each of the eight XCDs takes the buckets
 * of its own share of the tree (in tree order, dealt in chunks of 128) and,
 * once that is done, helps with the other shares.  Results do not depend on
 * the schedule. */
int gg_set_grav_schedule(GGPKD pkd, int mode);

/* Bucket-size autotuner: builds the tree (sysOff = NULL) or the forest with
 * each of the nCand candidate bucket sizes, times the device work of
 * gg_grav_all, and leaves the handle built with the fastest (*pBest);
 * ms[k] receives candidate k's time.  a, fPot, dtGrav and fWeight are
 * preserved. */
int gg_tune_bucket(GGPKD pkd, int nSys, const int *sysOff, int iOpenType, double dCrit,
                   int iOrder, const int *cand, int nCand, double *ms, int *pBest);

/* ------------------------------------------- kick, drift, accel reset */

/* pkdInitAccel (pkd.c:5443): a, fPot, dtGrav = 0 for TYPE_ACTIVE particles. */
int gg_init_accel(GGPKD pkd);
/* pkdKick (pkd.c:3780, NOCOOLING build): for TYPE_ACTIVE particles
 * vPred = v dvPredFacOne + a dvPredFacTwo, v = v dvFacOne + a dvFacTwo; for
 * TYPE_GAS ones unless bIsothermal, uPred = u + uDot duPredDelta and
 * u = u + uDot duDelta with uDot = uDotPdV + uDotAV + uDotDiff (pkd.h:65). */
int gg_kick(GGPKD pkd, double dvFacOne, double dvFacTwo, double dvPredFacOne,
            double dvPredFacTwo, double duDelta, double duPredDelta, int bIsothermal);
/* pkdDrift (pkd.c:3686): r += dDelta v, wrapped into the box
 * [fCenter - fPeriod/2, fCenter + fPeriod/2) when bPeriodic; GG_EINVAL if a
 * coordinate is still outside (the reference asserts).  The trees are stale
 * afterwards: build again before the next force call. */
int gg_drift(GGPKD pkd, double dDelta, const double *fCenter, int bPeriodic,
             const double *fPeriod);
/* pkdKickVpred (pkd.c:6489, NOCOOLING build): for TYPE_GAS particles
 * vPred = vPred dvFacOne + a dvFacTwo and, unless bIsothermal,
 * uPred += uDot duDelta; GG_EINVAL if a uPred turns negative (the
 * reference asserts). */
int gg_kick_vpred(GGPKD pkd, double dvFacOne, double dvFacTwo, double duDelta, int bIsothermal);

/* ------------------------------------------- timesteps and rungs */

/* pkdActiveMaskRung (pkd.c:5228; pkdActiveRung pkd.c:4569 is iSetMask =
 * TYPE_ACTIVE): the bits of iSetMask set on the particles of rung iRung
 * (and above when bGreater) and cleared on the others; *pnActive (may be
 * null) receives their number.  A gravity tree has to be built again
 * afterwards if a TYPE_ACTIVE bit changed. */
int gg_active_mask_rung(GGPKD pkd, unsigned iSetMask, int iRung, int bGreater, int *pnActive);
/* Type bits, as the reference's pkd functions set them (TYPETest: any bit
 * of iTestMask; TYPEFilter: (bits & iFilterMask) == iTestMask); *pn (may
 * be null) receives the number of particles that matched.  A gravity tree
 * is kept unless some TYPE_ACTIVE bit changed, the SPH structure unless
 * some TYPE_GAS bit did.
 *   gg_active_type       pkdActiveType (pkd.c:5203): iSetMask set where the
 *                        bits test iTestMask, cleared elsewhere;
 *   gg_active_type_rung  pkdActiveTypeRung (pkd.c:5255): the same, restricted
 *                        to rung iRung (and above with bGreater);
 *   gg_active_exact_type pkdActiveExactType (pkd.c:5114): set where the bits
 *                        filter to iTestMask, cleared elsewhere;
 *   gg_set_type          pkdSetType (pkd.c:5138): set where the bits test;
 *   gg_reset_type        pkdResetType (pkd.c:5161): cleared where they test;
 *   gg_sph_curr_rung     pkdSphCurrRung (pkd.c:5876): *piCurrent = 1 if a
 *                        TYPE_GAS particle sits on the rung(s). */
int gg_active_type(GGPKD pkd, unsigned iTestMask, unsigned iSetMask, int *pn);
/* pkdEmergencyAdjust (pkd.c:3971, GASOLINE with NOCOOLING): every gas particle on
 * a rung below iRung whose dtNew fell under dDeltaThresh (< dDelta) takes
 * dt = dtNew, the rung of that dt (capped at iMaxRung - 1) and v = vPred.
 * *pnUn particles moved, the largest rung they took (*piMaxRungOut) with
 * *pnMaxRung of them on it, and the ideal largest rung. */
int gg_emergency_adjust(GGPKD pkd, int iRung, int iMaxRung, double dDelta, double dDeltaThresh,
                        int *pnUn, int *piMaxRungIdeal, int *pnMaxRung, int *piMaxRungOut);
int gg_active_type_rung(GGPKD pkd, unsigned iTestMask, unsigned iSetMask, int iRung,
                        int bGreater, int *pn);
int gg_active_exact_type(GGPKD pkd, unsigned iFilterMask, unsigned iTestMask, unsigned iSetMask,
                         int *pn);
int gg_set_type(GGPKD pkd, unsigned iTestMask, unsigned iSetMask, int *pn);
int gg_reset_type(GGPKD pkd, unsigned iTestMask, unsigned iSetMask, int *pn);
int gg_sph_curr_rung(GGPKD pkd, int iRung, int bGreater, int *piCurrent);

/* These update dt (and iRung, dtOld) of the store like their reference
 * functions, bit for bit: no contraction into FMA. */

/* pkdInitDt (pkd.c:4817): dt = dDelta for TYPE_ACTIVE particles. */
int gg_init_dt(GGPKD pkd, double dDelta);
/* pkdGravStep (pkd.c:4627): dt = min(dt, dEta/sqrt(dtGrav)) for active
 * particles; GG_EINVAL if one has dtGrav <= 0 (the reference asserts). */
int gg_grav_step(GGPKD pkd, double dEta);
/* pkdAccelStep (pkd.c:4643): the acceleration criterion (bEpsAcc:
 * dEta sqrt(fSoft/|a|), or dEta sqrt(h/|a|) for gas with h = sqrt(fBall2)/2)
 * and the potential criterion (bSqrtPhi: 3.5 dEta sqrt(dAccFac |fPot|)/|a|),
 * |a| scaled by dAccFac.  dVelFac, bDoGravity and dhMinOverSoft enter no
 * criterion, as in the reference. */
int gg_accel_step(GGPKD pkd, double dEta, double dVelFac, double dAccFac, int bDoGravity,
                  int bEpsAcc, int bSqrtPhi, double dhMinOverSoft);
/* pkdDensityStep (pkd.c:4689): dt = min(dt, dEta/sqrt(fDensity dRhoFac)). */
int gg_density_step(GGPKD pkd, double dEta, double dRhoFac);
/* pkdSphStep (pkd.c:5901, NOCOOLING): for active gas the Courant estimate
 * of the last force pass (dtNew, reset to FLT_MAX), the extrapolated
 * Courant step of heating gas, the adiabatic-cooling limit (dEtauDot) and
 * the diffusion limit; *pdtMinGas = the smallest dt of all gas particles
 * (DBL_MAX if none). */
int gg_sph_step(GGPKD pkd, double dCosmoFac, double dEtaCourant, double dEtauDot,
                double duMinDt, double dDiffCoeff, double dEtaDiffusion, double dResolveJeans,
                int bViscosityLimitdt, double *pdtMinGas);
/* pkdDtToRung (pkd.c:4729): particles on rung >= iRung take the rung of
 * their dt (bAll, capped at iMaxRung-1, dtOld = dt) or iRung / iRung+1;
 * *piMaxRungOut = the largest rung (the reference's return value) with
 * *pnMaxRung particles on it, *piMaxRungIdeal the rung the smallest dt
 * wants.  GG_EINVAL where the reference asserts (an inactive particle on an
 * updated rung, fSoft <= 0, dt <= 0, dDelta/dt >= 2.1e9); the rungs are
 * then undefined.  bDiagExceed selects diagnostic prints only. */
int gg_dt_to_rung(GGPKD pkd, int iRung, double dDelta, int iMaxRung, int bAll, int bDiagExceed,
                  int *pnMaxRung, int *piMaxRungIdeal, int *piMaxRungOut);

/* ------------------------------------------------- h-proposal MLP */

/* A 4-32-32-1 tanh network (FP32) that proposes each gas particle's next
 * smoothing length from its previous h, density, velocity divergence and
 * the step dt, and writes the proposal as the fBall2 hint gg_smooth starts
 * its search from.  The kNN search stays exact whatever the hint, so the
 * network changes how much work the search does, never its result.  No
 * counterpart in the reference: it stands beside smSmooth (smooth.c:1439),
 * whose search starts from the bucket's own particles.
 *
 * Use per step: gg_smooth(DENDVDX) -> gg_hmlp_train (previous history ->
 * the fBall2 just found) -> gg_hmlp_record(dt) -> ... forces, kick, drift
 * ... -> gg_hmlp_predict -> next gg_smooth.  The history is kept by store
 * index: a caller that reorders its particles records again before
 * predicting.
 *
 * Parameter vector (GG_HMLP_NP floats): W1[32][4], b1[32], W2[32][32],
 * b2[32], W3[32], b3, input mean[4], input scale[4].  Inputs
 * x = (log h, log rho, divv dt, log dt), standardised by mean and scale
 * (taken from the first training set); output y = log h_next - log h. */
#define GG_HMLP_NW 1249
#define GG_HMLP_NP 1257

/* Deterministic initialisation from seed (Glorot-uniform W1, W2; W3 = b3 = 0,
 * so the untrained network proposes h_next = h); clears Adam state and
 * history. */
int gg_hmlp_init(GGPKD pkd, unsigned seed);
/* Snapshot the features of the particles the last gg_smooth smoothed. */
int gg_hmlp_record(GGPKD pkd, double dt);
/* nIter Adam steps (beta 0.9/0.999) on minibatches of nBatch samples drawn
 * (seeded, reproducibly: sample k of the handle's Adam step s is drawn by a
 * hash of (seed, s, k)) from the particles with a history that the last
 * gg_smooth smoothed again; target t = log h_now - log h_recorded.  Returns
 * the first and last minibatch losses (mean (y-t)^2) and the sample count. */
int gg_hmlp_train(GGPKD pkd, int nIter, int nBatch, double lr, unsigned seed,
                  double *pLossFirst, double *pLossLast, int *pnSamples);
/* Write fBall2 = (2 h exp(y))^2 from the recorded features, for the
 * particles with a history; others keep their fBall2. */
int gg_hmlp_predict(GGPKD pkd);
/* Get (bSet = 0) or set the GG_HMLP_NP parameters. */
int gg_hmlp_params(GGPKD pkd, float *w, int n, int bSet);
/* Forward pass on n host samples x[4][n] (feature-major): y[n]. */
int gg_hmlp_eval(GGPKD pkd, const float *x, int n, float *y);
/* Gradient of mean (y - t)^2 over n host samples: g[GG_HMLP_NW], loss. */
int gg_hmlp_grad(GGPKD pkd, const float *x, const float *t, int n, float *g, double *loss);

/* ------------------------------------------------------------ timing */

/* Device time in milliseconds of the last call of each kind, by phase:
 * t[0] tree walk (list build), t[1] interaction kernel, t[2] SPH structure,
 * t[3] kNN kernel, t[4] density kernel, t[5] resmooth lists, t[6] reverse
 * lists, t[7] hydro kernel, t[8] other (gathers, scatters, counters), t[9]
 * kNN group lists, t[10] kNN retries (with the count read-back before them).
 * Measured with HIP events on the handle's stream; n <= 12. */
int gg_timings(GGPKD pkd, double *t, int n);

#ifdef __cplusplus
}
#endif
#endif
