/*
 * smooth_oracle.c -- FP64 CPU restatement of gasoline's SPH neighbour loop:
 * exact kNN (the set smSmooth finds), ball gather (smBallGather), DenDVDX,
 * pkdGasPressure and SphPressureTermsSym with the M4 kernel.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with -ffp-contract=off.
 *
 * Neighbour search uses a uniform grid and an explicit sort, not the
 * reference's priority queue and snake order: the neighbour SET is fixed by
 * the FP64 distances (strict '<' replacement, smooth.c:602-643, keeps the
 * nSmooth smallest), and only ties at the boundary -- which orc_knn counts --
 * could make the two differ.  Lists are ordered by (d2, index); the
 * reference's list order is its PQ slot order, which changes FP64 sums only
 * in the last bits.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "oracle.h"

#define ORC_FLT_MAX ((double)3.402823466E+38F)   /* floattype.h:9 */
#define ORC_ALPHAMIN 0.01                        /* pkd.h:41 */

/* ------------------------------------------------------------ grid */
typedef struct {
    int g[3], periodic;
    double x0[3], cs[3], L[3];
    int *start;    /* g0*g1*g2 + 1 */
    int *item;     /* particle ids sorted by cell */
} OGrid;

static int cell_of(const OGrid *G, const double *x, int d)
{
    int c = (int)floor((x[d] - G->x0[d]) / G->cs[d]);
    if (c < 0) c = 0;
    if (c >= G->g[d]) c = G->g[d] - 1;
    return c;
}

static void grid_build(OGrid *G, int n, const double *r, int per_cell,
                       int bPeriodic, const double *L)
{
    double lo[3], hi[3];
    int i, d, ncell, *cnt;
    long long gg;
    for (d = 0; d < 3; ++d) { lo[d] = DBL_MAX; hi[d] = -DBL_MAX; }
    for (i = 0; i < n; ++i)
        for (d = 0; d < 3; ++d) {
            if (r[3 * i + d] < lo[d]) lo[d] = r[3 * i + d];
            if (r[3 * i + d] > hi[d]) hi[d] = r[3 * i + d];
        }
    gg = (long long)cbrt((double)n / per_cell);
    if (gg < 1) gg = 1;
    if (gg > 256) gg = 256;
    G->periodic = bPeriodic;
    for (d = 0; d < 3; ++d) {
        double span = hi[d] - lo[d];
        G->g[d] = (int)gg;
        if (bPeriodic) {
            G->L[d] = L[d];
            G->x0[d] = lo[d];
            G->cs[d] = L[d] / gg;
        } else {
            G->L[d] = DBL_MAX;
            G->x0[d] = lo[d];
            G->cs[d] = span > 0 ? span / gg * (1.0 + 1e-12) : 1.0;
        }
    }
    ncell = G->g[0] * G->g[1] * G->g[2];
    G->start = (int *)calloc((size_t)ncell + 1, sizeof(int));
    G->item = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    cnt = (int *)calloc((size_t)ncell + 1, sizeof(int));
    for (i = 0; i < n; ++i) {
        int c = (cell_of(G, r + 3 * i, 2) * G->g[1] + cell_of(G, r + 3 * i, 1)) * G->g[0]
                + cell_of(G, r + 3 * i, 0);
        cnt[c + 1]++;
    }
    for (i = 0; i < ncell; ++i) cnt[i + 1] += cnt[i];
    memcpy(G->start, cnt, ((size_t)ncell + 1) * sizeof(int));
    for (i = 0; i < n; ++i) {
        int c = (cell_of(G, r + 3 * i, 2) * G->g[1] + cell_of(G, r + 3 * i, 1)) * G->g[0]
                + cell_of(G, r + 3 * i, 0);
        G->item[cnt[c]++] = i;
    }
    free(cnt);
}

static void grid_free(OGrid *G) { free(G->start); free(G->item); }

/* Separation p - q with the minimum periodic image chosen per axis; the
 * image is applied to the query point first and then subtracted, as the
 * reference does with sx = x +/- lx (INTERSECT, smooth.h:291-390). */
static inline void separation(const double *p, const double *q, int bPeriodic,
                              const double *L, double *dx)
{
    int d;
    for (d = 0; d < 3; ++d) {
        double s = p[d], t = p[d] - q[d];
        if (bPeriodic) {
            if (t > 0.5 * L[d]) s = p[d] - L[d];
            else if (t < -0.5 * L[d]) s = p[d] + L[d];
        }
        dx[d] = s - q[d];
    }
}

typedef struct { double d2; int j; double dx[3]; } OCand;

static int cand_cmp(const void *a, const void *b)
{
    const OCand *x = (const OCand *)a, *y = (const OCand *)b;
    if (x->d2 < y->d2) return -1;
    if (x->d2 > y->d2) return 1;
    return (x->j > y->j) - (x->j < y->j);
}

/* Collect every particle whose grid cell lies in the index box
 * [lo,hi] (per axis; wrapped when periodic).  Returns count. */
static int gather_box(const OGrid *G, const double *r, const double *x,
                      const int lo[3], const int hi[3], OCand **buf, int *cap)
{
    int n = 0, a, b, c, d, full[3], lo2[3], hi2[3];
    for (d = 0; d < 3; ++d) {
        full[d] = (hi[d] - lo[d] + 1 >= G->g[d]);
        lo2[d] = full[d] ? 0 : lo[d];
        hi2[d] = full[d] ? G->g[d] - 1 : hi[d];
        if (!G->periodic && !full[d]) {
            if (lo2[d] < 0) lo2[d] = 0;
            if (hi2[d] > G->g[d] - 1) hi2[d] = G->g[d] - 1;
        }
    }
    for (c = lo2[2]; c <= hi2[2]; ++c)
        for (b = lo2[1]; b <= hi2[1]; ++b)
            for (a = lo2[0]; a <= hi2[0]; ++a) {
                int ia = ((a % G->g[0]) + G->g[0]) % G->g[0];
                int ib = ((b % G->g[1]) + G->g[1]) % G->g[1];
                int ic = ((c % G->g[2]) + G->g[2]) % G->g[2];
                int cell = (ic * G->g[1] + ib) * G->g[0] + ia, k;
                for (k = G->start[cell]; k < G->start[cell + 1]; ++k) {
                    int j = G->item[k];
                    OCand *o;
                    if (n >= *cap) {
                        *cap = *cap * 2 + 64;
                        *buf = (OCand *)realloc(*buf, (size_t)*cap * sizeof(OCand));
                    }
                    o = &(*buf)[n++];
                    o->j = j;
                    separation(x, r + 3 * j, G->periodic, G->L, o->dx);
                    o->d2 = o->dx[0] * o->dx[0] + o->dx[1] * o->dx[1] + o->dx[2] * o->dx[2];
                }
            }
    return n;
}

int orc_knn(int n, const double *r, int nSmooth, int bPeriodic,
            const double *L, int *nbr, double *fBall2)
{
    OGrid G;
    int i, ties = 0;
    grid_build(&G, n, r, 16, bPeriodic, L);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : ties)
    for (i = 0; i < n; ++i) {
        OCand *buf = NULL;
        int cap = 0, k, d;
        const double *x = r + 3 * i;
        int ci[3];
        for (d = 0; d < 3; ++d) ci[d] = cell_of(&G, x, d);
        for (k = 1;; ++k) {
            int lo[3], hi[3], m, all = 1;
            double cover = DBL_MAX;
            for (d = 0; d < 3; ++d) {
                lo[d] = ci[d] - k;
                hi[d] = ci[d] + k;
                if (hi[d] - lo[d] + 1 < G.g[d]) {
                    all = 0;
                    if (G.cs[d] * k < cover) cover = G.cs[d] * k;
                }
            }
            /* non-periodic: cells beyond the grid are empty space, so
             * clamping at the edge does not shrink the covered radius */
            m = gather_box(&G, r, x, lo, hi, &buf, &cap);
            if (m < nSmooth && !all) continue;
            qsort(buf, (size_t)m, sizeof(OCand), cand_cmp);
            if (!all && sqrt(buf[nSmooth - 1].d2) > cover) continue;
            for (d = 0; d < nSmooth; ++d) nbr[(size_t)i * nSmooth + d] = buf[d].j;
            fBall2[i] = buf[nSmooth - 1].d2;
            if (m > nSmooth && buf[nSmooth].d2 == buf[nSmooth - 1].d2) ties++;
            break;
        }
        free(buf);
    }
    grid_free(&G);
    return ties;
}

long long orc_ball_gather(int n, const double *r, const double *fBall2,
                          int bPeriodic, const double *L, long long *off,
                          int *idx, double *dxyz, double *d2out)
{
    OGrid G;
    int i;
    long long total = 0;
    grid_build(&G, n, r, 16, bPeriodic, L);
    if (!idx) off[0] = 0;
#pragma omp parallel for schedule(dynamic, 256)
    for (i = 0; i < n; ++i) {
        OCand *buf = NULL;
        int cap = 0, m, k, d, lo[3], hi[3], cnt = 0;
        const double *x = r + 3 * i;
        double R = sqrt(fBall2[i]);
        for (d = 0; d < 3; ++d) {
            lo[d] = (int)floor((x[d] - R - G.x0[d]) / G.cs[d]);
            hi[d] = (int)floor((x[d] + R - G.x0[d]) / G.cs[d]);
        }
        m = gather_box(&G, r, x, lo, hi, &buf, &cap);
        qsort(buf, (size_t)m, sizeof(OCand), cand_cmp);
        for (k = 0; k < m; ++k) {
            if (buf[k].d2 <= fBall2[i]) {
                if (idx) {
                    long long o = off[i] + cnt;
                    idx[o] = buf[k].j;
                    dxyz[3 * o] = buf[k].dx[0];
                    dxyz[3 * o + 1] = buf[k].dx[1];
                    dxyz[3 * o + 2] = buf[k].dx[2];
                    d2out[o] = buf[k].d2;
                }
                cnt++;
            }
        }
        if (!idx) off[i + 1] = cnt;
        free(buf);
    }
    if (!idx) {
        for (i = 0; i < n; ++i) off[i + 1] += off[i];
    }
    total = off[n];
    grid_free(&G);
    return total;
}

/* ------------------------------------------------ M4 kernel (smoothfcn.c:199-214) */
static inline double m4_w(double ar2)
{
    double ak = 2.0 - sqrt(ar2);
    if (ar2 < 1.0) ak = (1.0 - 0.75 * ak * ar2);
    else ak = 0.25 * ak * ak * ak;
    return ak;
}

static inline double m4_dw(double ar2)
{
    double adk = sqrt(ar2);
    if (ar2 < 1.0) adk = -3 + 2.25 * adk;
    else adk = -0.75 * (2.0 - adk) * (2.0 - adk) / adk;
    return adk;
}

void orc_dendvdx(int n, const double *m, const double *vpred,
                 const double *upred, const unsigned char *active,
                 const unsigned char *smoothActive, const double *fBall2,
                 const long long *off, const int *idx, const double *dxyz,
                 const double *d2, const orc_smf *smf, double *rho,
                 double *divv, double *curlv, double *dvds, double *fDivv_t,
                 double *balsara, double *diff, unsigned char *nbrOfActive)
{
    int i;
    /* TYPE_NbrOfACTIVE marking is a scatter (smoothfcn.c:3424); done
     * serially first so the parallel loop below only reads. */
    for (i = 0; i < n; ++i) {
        long long k;
        if (!smoothActive[i]) continue;
        if (active[i])
            for (k = off[i]; k < off[i + 1]; ++k) nbrOfActive[idx[k]] = 1;
        for (k = off[i]; k < off[i + 1]; ++k)
            if (active[idx[k]]) { nbrOfActive[i] = 1; break; }
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (i = 0; i < n; ++i) {
        double ih2, ih, vFac, fNorm, fNorm1, fDensity = 0.0;
        double dvxdx = 0, dvxdy = 0, dvxdz = 0, dvydx = 0, dvydy = 0, dvydz = 0;
        double dvzdx = 0, dvzdy = 0, dvzdz = 0, grx = 0, gry = 0, grz = 0;
        double divvnorm = 0, trace, gnorm, dvdr, dv_s;
        const double *vp = vpred + 3 * i;
        long long k;
        if (!smoothActive[i]) continue;
        ih2 = 4.0 / fBall2[i];
        ih = sqrt(ih2);
        vFac = (smf->bCannonical ? 1. / (smf->a * smf->a) : 1.0);
        fNorm = M_1_PI * ih2 * ih;
        fNorm1 = fNorm * ih2;
        for (k = off[i]; k < off[i + 1]; ++k) {
            int q = idx[k];
            double r2 = d2[k] * ih2, rs, rs1, dx, dy, dz, dvx, dvy, dvz;
            rs = m4_w(r2);
            fDensity += rs * m[q];
            rs1 = m4_dw(r2);
            rs1 *= m[q];
            dx = dxyz[3 * k]; dy = dxyz[3 * k + 1]; dz = dxyz[3 * k + 2];
            dvx = (-vp[0] + vpred[3 * q]) * vFac;
            dvy = (-vp[1] + vpred[3 * q + 1]) * vFac;
            dvz = (-vp[2] + vpred[3 * q + 2]) * vFac;
            dvxdx += dvx * dx * rs1;
            dvxdy += dvx * dy * rs1;
            dvxdz += dvx * dz * rs1;
            dvydx += dvy * dx * rs1;
            dvydy += dvy * dy * rs1;
            dvydz += dvy * dz * rs1;
            dvzdx += dvz * dx * rs1;
            dvzdy += dvz * dy * rs1;
            dvzdz += dvz * dz * rs1;
            divvnorm += (dx * dx + dy * dy + dz * dz) * rs1;
            grx += (upred[q]) * dx * rs1;
            gry += (upred[q]) * dy * rs1;
            grz += (upred[q]) * dz * rs1;
        }
        fDensity *= fNorm;
        rho[i] = fDensity;
        trace = dvxdx + dvydy + dvzdz;
        fNorm1 = (divvnorm != 0 ? 3 / fabs(divvnorm) : 0);
        fDivv_t[i] = fNorm1 * trace;
        divv[i] = fNorm1 * trace + 3 * smf->H;
        curlv[3 * i] = fNorm1 * (dvzdy - dvydz);
        curlv[3 * i + 1] = fNorm1 * (dvxdz - dvzdx);
        curlv[3 * i + 2] = fNorm1 * (dvydx - dvxdy);
        {
            double Hcorr = (fNorm1 != 0 ? smf->H / fNorm1 : 0);
            gnorm = (grx * grx + gry * gry + grz * grz);
            if (gnorm > 0) gnorm = 1 / sqrt(gnorm);
            grx *= gnorm; gry *= gnorm; grz *= gnorm;
            dvdr = (((dvxdx + Hcorr) * grx + dvxdy * gry + dvxdz * grz) * grx
                    + (dvydx * grx + (dvydy + Hcorr) * gry + dvydz * grz) * gry
                    + (dvzdx * grx + dvzdy * gry + (dvzdz + Hcorr) * grz) * grz) * fNorm1;
            dvds[i] = dv_s = (divv[i] < 0 ? 1.5 * (dvdr - (1. / 3.) * divv[i]) : dvdr);
        }
        {
            const double *cv = curlv + 3 * i;
            double cmag = sqrt(cv[0] * cv[0] + cv[1] * cv[1] + cv[2] * cv[2]);
            switch (smf->iViscosityLimiter) {
            case 0: balsara[i] = 1; break;
            case 1:
                if (divv[i] != 0.0) balsara[i] = fabs(divv[i]) / (fabs(divv[i]) + cmag);
                else balsara[i] = 0;
                break;
            default:
                if (dv_s != 0) balsara[i] = fabs(dv_s) / (fabs(dv_s) + cmag);
                else balsara[i] = ORC_ALPHAMIN;
            }
        }
        {
            double ot = (1. / 3.) * trace;
            double sxx = dvxdx - ot, syy = dvydy - ot, szz = dvzdz - ot;
            double sxy = 0.5 * (dvxdy + dvydx), sxz = 0.5 * (dvxdz + dvzdx);
            double syz = 0.5 * (dvydz + dvzdy);
            double S2 = fNorm1 * fNorm1 * (sxx * sxx + syy * syy + szz * szz
                        + 2 * (sxy * sxy + sxz * sxz + syz * syz));
            diff[i] = smf->bConstantDiffusion ? 1 : 0.25 * fBall2[i] * sqrt(2 * S2);
        }
    }
}

/* ------------------------------------ DensitySym (smoothfcn.c:258-274) */
void orc_density_sym(int n, const double *m, const unsigned char *smoothActive,
                     const double *fBall2, const long long *off, const int *idx,
                     const double *d2, double *rho)
{
    int i;
    /* initDensity on the smooth-active particles (smooth.c:466-472,
     * smoothfcn.c:229-232); the others keep what they held. */
    for (i = 0; i < n; ++i)
        if (smoothActive[i]) rho[i] = 0.0;
    /* serial: every pair updates both ends */
    for (i = 0; i < n; ++i) {
        double ih2, fNorm;
        long long k;
        if (!smoothActive[i]) continue;
        ih2 = 4.0 / fBall2[i];
        fNorm = 0.5 * M_1_PI * sqrt(ih2) * ih2;
        for (k = off[i]; k < off[i + 1]; ++k) {
            int q = idx[k];
            double rs = m4_w(d2[k] * ih2) * fNorm;
            rho[i] += rs * m[q];
            rho[q] += rs * m[i];
        }
    }
}

void orc_gas_pressure(int n, const double *upred, const double *rho,
                      const double *fBall2, const double *udot_hydro,
                      const orc_smf *smf, double *PoverRho2, double *c,
                      double *dt)
{
    int i;
    double gammam1 = smf->gamma - 1;
    for (i = 0; i < n; ++i) {
        double PoverRhoGas = gammam1 * upred[i];
        double cGas = sqrt(smf->gamma * PoverRhoGas);
        double floorJ = 0.25 * fBall2[i] * smf->dResolveJeans * rho[i];
        double PoverRho = PoverRhoGas;                 /* + PoverRhoHot (0) */
        double PoverRhoJeans = (PoverRho < floorJ ? floorJ - PoverRho : 0);
        double udot, dti;
        PoverRho += PoverRhoJeans;
        PoverRho2[i] = PoverRho / rho[i];
        /* GAMMA_NONCOOL*PoverRhoHot vanishes without UNONCOOL; GAMMA_JEANS = 2 */
        c[i] = sqrt(cGas * cGas + 2.0 * PoverRhoJeans);
        udot = udot_hydro ? udot_hydro[i] : 0.0;
        if (udot > 0)
            dti = smf->dtFacCourant * sqrt(fBall2[i] * 0.25 / (4 * (c[i] * c[i] + (5. / 3.) * udot * dt[i])));
        else
            dti = smf->dtFacCourant * sqrt(fBall2[i] * 0.25) / (2 * (c[i]));
        if (dti < dt[i]) dt[i] = dti;
    }
}

/* The terms of one pair of SphPressureTermsSym (smoothfcn.c:2276-2385,
 * SphPressureTerms.h:97-152): entry k of p's list, q = idx[k], evaluated
 * with p's kernel.  Each end's increments come out separately: tp for p,
 * tq for q (acceleration 3, uDotPdV, uDotAV, uDotDiff), and absmu >= 0
 * (0 for a receding pair) and dtv for mumax and dtNew of both.  Returns 0
 * when neither end is active (the pair is skipped). */
typedef struct {
    const double *m, *vpred, *upred, *rho, *PoverRho2, *c, *balsara, *diff, *dt, *fBall2;
    const unsigned char *active;
    const int *idx;
    const double *dxyz, *d2;
    const orc_smf *smf;
} OHydro;

static int hydro_pair(const OHydro *H, int p, long long k, double tp[6], double tq[6],
                      double *absmu_out, double *dtv_out)
{
    const orc_smf *smf = H->smf;
    const int q = H->idx[k];
    const double *m = H->m, *vpred = H->vpred, *rho = H->rho, *c = H->c;
    double ph, ih2, fNorm, fNorm1, aFac, vFac, pPoverRho2, pPoverRho2f;
    double pc = c[p], pMass = m[p], pDensity = rho[p];
    double r2, rs1, rp, rq, dx, dy, dz, dvx, dvy, dvz, dvdotdr;
    double qPoverRho2, qPoverRho2f, Accp = 0, Accq = 0, dtv, visc, absmu = 0;
    int pa = H->active[p] != 0, qa = H->active[q] != 0, j;
    for (j = 0; j < 6; ++j) tp[j] = tq[j] = 0.0;
    if (!pa && !qa) return 0;
    pPoverRho2 = H->PoverRho2[p];
    {
        double pd2 = rho[p] * rho[p];
        pPoverRho2f = (pPoverRho2 * pd2 + 0) / pd2;
    }
    ph = sqrt(0.25 * H->fBall2[p]);
    ih2 = 4.0 / H->fBall2[p];
    fNorm = 0.5 * M_1_PI * ih2 / ph;
    fNorm1 = fNorm * ih2;
    aFac = smf->a;
    vFac = (smf->bCannonical ? 1. / (smf->a * smf->a) : 1.0);
    r2 = H->d2[k] * ih2;
    rs1 = m4_dw(r2);
    rs1 *= fNorm1;
    rs1 *= 1.0;                 /* fDivv_Corrector (non-GDFORCE) */
    rp = rs1 * pMass;
    rq = rs1 * m[q];
    dx = H->dxyz[3 * k]; dy = H->dxyz[3 * k + 1]; dz = H->dxyz[3 * k + 2];
    dvx = vpred[3 * p] - vpred[3 * q];
    dvy = vpred[3 * p + 1] - vpred[3 * q + 1];
    dvz = vpred[3 * p + 2] - vpred[3 * q + 2];
    dvdotdr = vFac * (dvx * dx + dvy * dy + dvz * dz) + H->d2[k] * smf->H;
    qPoverRho2 = H->PoverRho2[q];
    {
        double qd2 = rho[q] * rho[q];
        qPoverRho2f = (qPoverRho2 * qd2 + 0) / qd2;
    }
    /* body of SphPressureTerms.h:112-152 */
    if (pa) tp[3] = rq * pPoverRho2 * dvdotdr;
    if (qa) tq[3] = rp * qPoverRho2 * dvdotdr;
    if (pa) Accp = (pPoverRho2f + qPoverRho2f);
    if (qa) Accq = (qPoverRho2f + pPoverRho2f);
    if (dvdotdr >= 0.0) {
        dtv = smf->dtFacCourant * ph / (2 * (pc > c[q] ? pc : c[q]));
    } else {
        absmu = -dvdotdr * smf->a / sqrt(H->d2[k]);
        visc = (smf->alpha * (pc + c[q]) + smf->beta * 1.5 * absmu);
        dtv = smf->dtFacCourant * ph / (0.625 * (pc + c[q]) + 0.375 * visc);
        visc = (0.5 * (H->balsara[p] + H->balsara[q])) * visc * absmu / (pDensity + rho[q]);
        if (pa) tp[4] = rq * (0.5 * visc) * dvdotdr;
        if (qa) tq[4] = rp * (0.5 * visc) * dvdotdr;
        if (pa) Accp += visc;
        if (qa) Accq += visc;
    }
    if (pa) Accp *= rq * aFac;
    if (qa) Accq *= rp * aFac;
    if (pa) {
        tp[0] = -(Accp * dx);
        tp[1] = -(Accp * dy);
        tp[2] = -(Accp * dz);
    }
    if (qa) {
        tq[0] = Accq * dx;
        tq[1] = Accq * dy;
        tq[2] = Accq * dz;
    }
    {   /* thermal diffusion (SphPressureTerms.h:32-70) */
        double diffSum = H->diff[p] + H->diff[q];
        double diffBase = (diffSum == 0 ? 0 : diffSum);
        double diffTh = (2 * smf->dThermalDiffusionCoeff * diffBase / (pDensity + rho[q]));
        double dt_diff, diffu;
        if (diffTh > 0 && (dt_diff = smf->dtFacDiffusion * ph * ph / (diffTh * pDensity)) < dtv)
            dtv = dt_diff;
        diffu = (diffTh + 0 + 0) * (H->upred[p] - H->upred[q]);
        if (pa) tp[5] = diffu * rq * 1;
        if (qa) tq[5] = -(diffu * rp * 1);
    }
    *absmu_out = absmu;
    *dtv_out = dtv;
    return 1;
}

static void hydro_init(int n, const unsigned char *active, double *mumax, double *dtNew,
                       double *uDotPdV, double *uDotAV, double *uDotDiff)
{
    int p;
    /* initSphPressureTermsParticle (smoothfcn.c:2204-2221): a is NOT zeroed */
    for (p = 0; p < n; ++p) {
        if (!active[p]) continue;
        mumax[p] = 0.0;
        dtNew[p] = ORC_FLT_MAX;
        uDotDiff[p] = 0.0;
        uDotPdV[p] = 0.0;
        uDotAV[p] = 0.0;
    }
}

/* one end's share of a pair: its increments, then mumax and dtNew
 * (SETDTNEW_PQ, SphPressureTerms.h:97-100: the pair's dt and 4 dt of the
 * other end) */
static void hydro_apply(int x, int other, const double t[6], double absmu, double dtv,
                        const double *dt, double *acc, double *uDotPdV, double *uDotAV,
                        double *uDotDiff, double *mumax, double *dtNew)
{
    acc[3 * x] += t[0];
    acc[3 * x + 1] += t[1];
    acc[3 * x + 2] += t[2];
    uDotPdV[x] += t[3];
    uDotAV[x] += t[4];
    uDotDiff[x] += t[5];
    if (absmu > mumax[x]) mumax[x] = absmu;
    if (dtv < dtNew[x]) dtNew[x] = dtv;
    if (4 * dt[other] < dtNew[x]) dtNew[x] = 4 * dt[other];
}

/* The serial scatter form: the reference's order, every pair updating both
 * ends as it is met. */
void orc_sph_pressure_sym_serial(int n, const double *m, const double *vpred,
                                 const double *upred, const double *rho,
                                 const double *PoverRho2, const double *c,
                                 const double *balsara, const double *diff,
                                 const double *dt, const unsigned char *active,
                                 const double *fBall2, const long long *off,
                                 const int *idx, const double *dxyz, const double *d2,
                                 const orc_smf *smf, double *acc, double *uDotPdV,
                                 double *uDotAV, double *uDotDiff, double *mumax,
                                 double *dtNew)
{
    OHydro H = {m, vpred, upred, rho, PoverRho2, c, balsara, diff, dt, fBall2, active, idx,
                dxyz, d2, smf};
    int p;
    hydro_init(n, active, mumax, dtNew, uDotPdV, uDotAV, uDotDiff);
    for (p = 0; p < n; ++p) {
        long long k;
        for (k = off[p]; k < off[p + 1]; ++k) {
            double tp[6], tq[6], absmu, dtv;
            const int q = idx[k];
            if (!hydro_pair(&H, p, k, tp, tq, &absmu, &dtv)) continue;
            hydro_apply(p, q, tp, absmu, dtv, dt, acc, uDotPdV, uDotAV, uDotDiff, mumax, dtNew);
            hydro_apply(q, p, tq, absmu, dtv, dt, acc, uDotPdV, uDotAV, uDotDiff, mumax, dtNew);
        }
    }
}

/* The gather form on OpenMP threads: each particle x takes its share of
 * the pairs of its own list (x as p) and of the lists holding it (x as q,
 * from the reverse lists, in list order), the same terms as the serial
 * form summed in another order. */
void orc_sph_pressure_sym(int n, const double *m, const double *vpred,
                          const double *upred, const double *rho,
                          const double *PoverRho2, const double *c,
                          const double *balsara, const double *diff,
                          const double *dt, const unsigned char *active,
                          const double *fBall2, const long long *off,
                          const int *idx, const double *dxyz, const double *d2,
                          const orc_smf *smf, double *acc, double *uDotPdV,
                          double *uDotAV, double *uDotDiff, double *mumax,
                          double *dtNew)
{
    OHydro H = {m, vpred, upred, rho, PoverRho2, c, balsara, diff, dt, fBall2, active, idx,
                dxyz, d2, smf};
    const long long ne = off[n];
    long long *roff = (long long *)calloc((size_t)n + 1, sizeof(long long));
    long long *rk = (long long *)malloc((size_t)(ne > 0 ? ne : 1) * sizeof(long long));
    long long *rpos = (long long *)malloc(((size_t)n + 1) * sizeof(long long));
    int *rp = (int *)malloc((size_t)(ne > 0 ? ne : 1) * sizeof(int));
    int p;
    long long k;
    hydro_init(n, active, mumax, dtNew, uDotPdV, uDotAV, uDotDiff);
    /* reverse lists: for each q the entries k (and their owner p) that
     * hold it, in increasing k */
    for (k = 0; k < ne; ++k) roff[idx[k] + 1]++;
    for (p = 0; p < n; ++p) roff[p + 1] += roff[p];
    memcpy(rpos, roff, ((size_t)n + 1) * sizeof(long long));
    for (p = 0; p < n; ++p)
        for (k = off[p]; k < off[p + 1]; ++k) {
            const long long o = rpos[idx[k]]++;
            rk[o] = k;
            rp[o] = p;
        }
#pragma omp parallel for schedule(dynamic, 256)
    for (p = 0; p < n; ++p) {
        long long e;
        for (e = off[p]; e < off[p + 1]; ++e) {
            double tp[6], tq[6], absmu, dtv;
            if (!hydro_pair(&H, p, e, tp, tq, &absmu, &dtv)) continue;
            hydro_apply(p, idx[e], tp, absmu, dtv, dt, acc, uDotPdV, uDotAV, uDotDiff, mumax,
                        dtNew);
        }
        for (e = roff[p]; e < roff[p + 1]; ++e) {
            double tp[6], tq[6], absmu, dtv;
            const int o = rp[e];
            if (!hydro_pair(&H, o, rk[e], tp, tq, &absmu, &dtv)) continue;
            hydro_apply(p, o, tq, absmu, dtv, dt, acc, uDotPdV, uDotAV, uDotDiff, mumax, dtNew);
        }
    }
    free(roff);
    free(rk);
    free(rpos);
    free(rp);
}
