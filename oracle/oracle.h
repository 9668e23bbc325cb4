/*
 * oracle.h -- CPU restatement of gasoline's force hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the gasoline_amd
 * package, its C-ABI library, bench.py's timed GPU region) may link, load or
 * call this code.  Only tests/, __graft_entry__.smoke() (as a checker) and
 * bench.py's cpu_baseline leg use it.
 *
 * Every routine is an FP64 restatement of the reference algorithm, cited by
 * file:line of N-BodyShop/gasoline.  Arithmetic is written in the same
 * operation order as the reference and the library is compiled with
 * -ffp-contract=off, so that FP64 decisions (tree splits, opening tests,
 * neighbour selection) round exactly as the reference does on x86-64.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   orc_vsqrt1             bit-exact vs the reference's own v_sqrt1.c compiled
 *                          into oracle/_ref (v_sqrt1.c:11-113).
 *   orc_build_binary       invariants: partition is a permutation, every bucket
 *                          holds <= nBucket particles or has zero extent, bounds
 *                          are exact, c.o.m./mass/Q equal direct sums, fOpen2
 *                          equals the OPEN_JOSH closed form.
 *   orc_walk_bucket        every particle of the system is covered exactly once
 *                          per bucket (PP + cells + own bucket), accepted cells
 *                          satisfy the opening test, opened ones fail it.
 *   orc_grav_all           FP64 O(N^2) direct sum with the same spline
 *                          softening (rms error of a theta=0.7 quadrupole tree),
 *                          closed forms of SPLINE at r=0 and r>=2h, single-
 *                          particle cells (Q=0) vs point masses.
 *   orc_direct_sum         Newton's law / Plummer analytic field.
 *   orc_direct_ewald       the Madelung constants of the simple cubic lattice
 *                          (own images and background) and of rock salt
 *                          (signed masses); the same sums for three splits
 *                          alpha; Newton plus the background's (4 pi/3) m r /
 *                          L^3 at small r; the spline inside twoh; and the
 *                          reference's own Ewald (orc_grav_all_periodic) for a
 *                          point-mass root, which expands this Green's function.
 *   orc_grav_all_periodic  a test mass near a point mass in the periodic box:
 *                          zero force at the symmetry points of the lattice,
 *                          odd in the separation, Newton plus the uniform
 *                          background's (4 pi/3) m r / L^3 at small r; with
 *                          the box large, the isolated result.
 *   orc_knn                brute-force full sort on small inputs.
 *   orc_dendvdx            uniform density, linear velocity field (div, curl
 *                          exact), M4 normalisation.
 *   orc_density_sym        sum of m*rho equal to the gather density's (each pair
 *                          counted once per end), gather equality on a lattice
 *                          whose lists are symmetric, unsmoothed particles keep
 *                          their value plus what their neighbours scatter.
 *   orc_kick, orc_drift    leapfrog of a uniform acceleration is exact
 *                          (x0 + v0 t + a t^2/2, v0 + a t); a drift by one
 *                          period returns every coordinate to itself.
 *   orc_*_step             closed forms at special values (dtGrav = 4,
 *                          rho = 1/dRhoFac ...), each limit alone.
 *   orc_dt_to_rung         dt = dDelta/2^k lands on rung k (an integer
 *                          boundary goes to the lower rung), a hair below
 *                          on rung k+1, floor(log2) on random dt.
 *   orc_sph_pressure_sym   momentum conservation, energy conservation, uniform
 *                          pressure => zero force, AV only for approaching pairs.
 */
#ifndef GG_ORACLE_H
#define GG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- v_sqrt1 */
void orc_vsqrt1(int n, const double *r2, double *out);

/* ---------------------------------------------------------------- gravity */
typedef struct orc_tree orc_tree;

/* Opening types, numbered as opentype.h:5-9. */
#define ORC_OPEN_JOSH   1
#define ORC_OPEN_ABSPAR 2

/* Build the gravity tree (pkdBuildBinary, pkd.c:2627-2723) over n particles.
 * r is n*3 (xyz interleaved).  iActive[i] holds the type bits (pkd.h:288;
 * bit 0 TYPE_ACTIVE, bit 1 TYPE_TREEACTIVE); NULL means all set.
 * Returns NULL on allocation failure. */
orc_tree *orc_build_binary(int n, const double *r, const double *m,
                           const double *soft, const unsigned *iActive,
                           int nBucket, int iOpenType, double dCrit, int iOrder);
void orc_tree_free(orc_tree *t);
int  orc_tree_nnodes(const orc_tree *t);
int  orc_tree_root(const orc_tree *t);
/* perm[k] = input index of the particle stored at slot k after the build. */
void orc_tree_perm(const orc_tree *t, int *perm);
/* Per node: ints[6*i..] = iDim,pLower,pUpper,iLower,iUpper(thread next),0.
 * dbls[ORC_NODE_NDBL*i..] = fSplit, bnd min[3], bnd max[3], fMass, fSoft,
 * r[3], fOpen2, Q[6] (xx yy zz xy xz yz), O[10], H[15], Bmax, B2..B6. */
#define ORC_NODE_NDBL 47
void orc_tree_nodes(const orc_tree *t, int *ints, double *dbls);

/* Interaction list of one bucket (pkdBucketWalk, walk.c:306, nReps=0).
 * part[] receives the storage slots of the PP particles in walk order, soft[]
 * and newt[] the node ids of softened and Newtonian cells.  cap bounds each
 * array; counts[3] = nPart, nCellSoft, nCellNewt (may exceed cap -> retry).
 * The bucket's bounds used for the opening test are those of its active
 * particles, as pkdGravAll (pkd.c:2916-2937) substitutes them. */
void orc_walk_bucket(orc_tree *t, int iBucket, int iOrder, int cap,
                     int *part, int *soft, int *newt, int counts[3]);

/* pkdGravAll (pkd.c:2869-3060) for one process, nReps=0, no Ewald, no
 * comoving term.  Outputs are in INPUT order (length n / n*3) and are the
 * accumulation from zero (msrInitAccel).  stats[5] = nActive, dPartSum,
 * dCellSum, dSoftSum, dFlop. */
void orc_grav_all(orc_tree *t, int iOrder, double *a, double *pot,
                  double *dtgrav, double *stats);

/* The same with the replicas -nReps..nReps of the periodic box L[3]
 * (pkdBucketWalk walk.c:306-439; L = NULL: isolated) and, when bEwald, the
 * Ewald correction of the root's multipoles to order iEwOrder (pkdCalcRoot
 * pkd.c:4395, pkdDistribRoot :4472, pkdEwaldInit ewald.c:180,
 * pkdBucketEwald ewald.c:15; cubic box L[0]). */
void orc_grav_all_periodic(orc_tree *t, int iOrder, int nReps, const double *L, int bEwald,
                           int iEwOrder, double fEwCut, double fEwhCut, double *a, double *pot,
                           double *dtgrav, double *stats);
/* The same, and fweight[i] (input order, may be NULL) of each active
 * particle = its bucket's interaction + Ewald flops (pkdBucketWeight,
 * pkd.c:2851-2861); inactive particles keep the value fweight holds. */
void orc_grav_all_weighted(orc_tree *t, int iOrder, int nReps, const double *L, int bEwald,
                           int iEwOrder, double fEwCut, double fEwhCut, double *a, double *pot,
                           double *dtgrav, double *stats, double *fweight);

/* FP64 O(N^2) direct sum with gasoline's SPLINE softening (grav.h:74-97),
 * twoh = h_i + h_j.  Targets are the nt indices in idx (or all if idx is
 * NULL, nt = n).  OpenMP over targets. */
void orc_direct_sum(int n, const double *r, const double *m, const double *soft,
                    int nt, const int *idx, double *a, double *pot);

/* The periodic counterpart of orc_direct_sum in a cubic box of side L:
 * the Ewald sum (split alpha = alphaL / L) of every other particle and all
 * images of every particle, with the uniform background, and the spline
 * softening of images nearer than twoh.  pot includes the particle's own
 * images, m_i * orc_ewald_self(L, alphaL).  OpenMP over targets. */
void orc_direct_ewald(int n, const double *r, const double *m, const double *soft, double L,
                      double alphaL, int nt, const int *idx, double *a, double *pot);
double orc_ewald_self(double L, double alphaL);

/* Spline kernels of grav.h, exported for closed-form tests. */
void orc_spline(double r2, double twoh, double *a, double *b);
void orc_splineq(double invr, double r2, double twoh, double *a, double *b,
                 double *c, double *d);

/* -------------------------------------------------------------------- SPH */
typedef struct orc_smf {
    double H, a;                /* Hubble, expansion factor (smoothfcn.h:22-23) */
    double alpha, beta, gamma;  /* viscosity and adiabatic index */
    double dtFacCourant;        /* pkdDtFacCourant(dEtaCourant, a), pkd.c:5893 */
    double dtFacDiffusion;      /* 2*dEtaDiffusion, master.c:5615 */
    double dMetalDiffusionCoeff, dThermalDiffusionCoeff;
    double dResolveJeans;       /* Jeans floor factor (0 = off) */
    int bCannonical, iViscosityLimiter, bConstantDiffusion, pad;
} orc_smf;

/* Exact nSmooth nearest neighbours of every particle by FP64 distance, the
 * set smSmooth (smooth.c:1439-1713) finds.  Periodic images use the minimum
 * image (INTERSECT, smooth.h:291-390) with period L[3] when bPeriodic.
 * nbr[i*nSmooth + k] sorted by (d2, index).  fBall2[i] = largest d2.
 * Returns the number of particles whose nSmooth-th and (nSmooth+1)-th
 * distances tie (those neighbour sets are not unique). */
int orc_knn(int n, const double *r, int nSmooth, int bPeriodic,
            const double *L, int *nbr, double *fBall2);

/* Ball gather at the stored radius (smBallGather, smooth.c:788-891): all q
 * with d2 <= fBall2[i].  Two-phase CSR: call with idx == NULL to get the
 * counts in off[i+1]-off[i] (off has n+1 entries, prefix-summed), then with
 * idx/dx buffers sized off[n].  Entries sorted by (d2, index). */
long long orc_ball_gather(int n, const double *r, const double *fBall2,
                          int bPeriodic, const double *L, long long *off,
                          int *idx, double *dxyz, double *d2);

/* DenDVDX (smoothfcn.c:3390-3627) over CSR neighbour lists, M4 kernel.
 * Outputs per particle: rho, divv, curlv[3], dvds, fDivv_t, Balsara, diff,
 * nbrOfActive flag (TYPE_NbrOfACTIVE bit).  Only particles with
 * smoothActive[i] are evaluated. */
void orc_dendvdx(int n, const double *m, const double *vpred,
                 const double *upred, const unsigned char *active,
                 const unsigned char *smoothActive, const double *fBall2,
                 const long long *off, const int *idx, const double *dxyz,
                 const double *d2, const orc_smf *smf, double *rho,
                 double *divv, double *curlv, double *dvds, double *fDivv_t,
                 double *balsara, double *diff, unsigned char *nbrOfActive);

/* DensitySym (smoothfcn.c:258-274) over CSR lists (the nSmooth nearest of
 * each smooth-active particle, d2 per entry): rho of smooth-active
 * particles starts at 0 (initDensity, smoothfcn.c:229-232), the others keep
 * the rho passed in; each pair adds the half-normalised kernel of the list
 * owner to both ends. */
void orc_density_sym(int n, const double *m, const unsigned char *smoothActive,
                     const double *fBall2, const long long *off, const int *idx,
                     const double *d2, double *rho);

/* pkdGasPressure (pkd.c:5750-5785), adiabatic, no cooling: PoverRho2, c and
 * the Courant update of dt (in/out). */
void orc_gas_pressure(int n, const double *upred, const double *rho,
                      const double *fBall2, const double *udot_hydro,
                      const orc_smf *smf, double *PoverRho2, double *c,
                      double *dt);

/* SphPressureTermsSym (smoothfcn.c:2276-2385 + SphPressureTerms.h), the
 * reference's symmetric scatter form, over CSR ball-gather lists, with
 * initSphPressureTermsParticle (smoothfcn.c:2204-2221) applied first to
 * active particles.  In/out: a (added to), uDotPdV, uDotAV, uDotDiff,
 * mumax, dtNew.  Non-GDFORCE (fDivv_Corrector = 1), PRES_PDV/PRES_ACC of
 * the Benz form (smoothfcn.c:20-23). */
void orc_sph_pressure_sym(int n, const double *m, const double *vpred,
                          const double *upred, const double *rho,
                          const double *PoverRho2, const double *c,
                          const double *balsara, const double *diff,
                          const double *dt, const unsigned char *active,
                          const double *fBall2, const long long *off,
                          const int *idx, const double *dxyz, const double *d2,
                          const orc_smf *smf, double *acc, double *uDotPdV,
                          double *uDotAV, double *uDotDiff, double *mumax,
                          double *dtNew);
/* The same in the serial scatter form (every pair updates both ends as it
 * is met); the threaded gather form above sums the same terms in another
 * order. */
void orc_sph_pressure_sym_serial(int n, const double *m, const double *vpred,
                          const double *upred, const double *rho,
                          const double *PoverRho2, const double *c,
                          const double *balsara, const double *diff,
                          const double *dt, const unsigned char *active,
                          const double *fBall2, const long long *off,
                          const int *idx, const double *dxyz, const double *d2,
                          const orc_smf *smf, double *acc, double *uDotPdV,
                          double *uDotAV, double *uDotDiff, double *mumax,
                          double *dtNew);

/* ------------------------------------------- updates either side (step_oracle.c) */
void orc_init_accel(int n, const unsigned *bits, double *a, double *pot, double *dtg);
void orc_kick(int n, const unsigned *bits, double dvFacOne, double dvFacTwo,
              double dvPredFacOne, double dvPredFacTwo, double duDelta, double duPredDelta,
              int bIsothermal, const double *a, const double *uDotPdV, const double *uDotAV,
              const double *uDotDiff, double *v, double *vPred, double *u, double *uPred);
int orc_kick_vpred(int n, const unsigned *bits, double dvFacOne, double dvFacTwo,
                   double duDelta, int bIsothermal, const double *a, const double *uDotPdV,
                   const double *uDotAV, const double *uDotDiff, double *vPred, double *uPred);
int orc_drift(int n, double dDelta, const double *fCenter, int bPeriodic, const double *fPeriod,
              const double *v, double *r);
void orc_init_dt(int n, const unsigned *bits, double dDelta, double *dt);
int orc_grav_step(int n, const unsigned *bits, double dEta, const double *dtGrav, double *dt);
void orc_accel_step(int n, const unsigned *bits, double dEta, double dVelFac, double dAccFac,
                    int bEpsAcc, int bSqrtPhi, const double *v, const double *a,
                    const double *fSoft, const double *fBall2, const double *fPot, double *dt);
void orc_density_step(int n, const unsigned *bits, double dEta, double dRhoFac,
                      const double *rho, double *dt);
void orc_sph_step(int n, const unsigned *bits, double dCosmoFac, double dEtaCourant,
                  double dEtauDot, double duMinDt, double dDiffCoeff, double dEtaDiffusion,
                  double dResolveJeans, const double *fBall2, const double *c,
                  const double *u, const double *uDotPdV, const double *uDotAV,
                  const double *uDotDiff, const double *rho, const double *diff, double *dtNew,
                  double *dt, double *dtMinGas);
int orc_one_dt_to_rung(int iRung, double dDelta, double dt);
int orc_dt_to_rung(int n, const unsigned *bits, int iRung, double dDelta, int iMaxRung, int bAll,
                   const double *fSoft, const double *dt, int *rung, double *dtOld,
                   int *nMaxRung, int *iMaxRungIdeal);

#ifdef __cplusplus
}
#endif
#endif
