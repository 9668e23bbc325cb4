/*
 * grav_oracle.c -- FP64 CPU restatement of gasoline's gravity hot path:
 * tree build (pkd.c), bucket walk (walk.c), bucket interaction (grav.c,
 * grav.h, qeval.h) and the batched 1/sqrt (v_sqrt1.c).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "oracle.h"

/* ======================================================================
 * v_sqrt1 (v_sqrt1.c:11-113): exponent-indexed scale tables, a degree-5
 * polynomial for x^-1/2 on [1,2) obtained from a 200-point Chebyshev fit,
 * and one Newton step.  The operation order is the reference's, because the
 * restatement is checked bit-for-bit against the reference's own file.
 * ====================================================================== */
#define VS_NCHEB 200
static double vs_scale[2048];   /* 2^(-(e-1023)/2)  (t[] of v_sqrt1.c:14) */
static double vs_norm[2048];    /* 2^-(e-1023)      (u[] of v_sqrt1.c:14) */
static double vs_coef[6];
static int vs_ready = 0;

static double vs_target(double y) { return pow(1.5 + 0.5 * y, -0.5); }

static void vs_setup(void)
{
    double half = 1.0, node[VS_NCHEB + 1], fval[VS_NCHEB + 1], cheb[VS_NCHEB];
    double p0, p1, p2, p3, p4, p5;
    int i, j, k;

    vs_scale[1023] = 1.0;
    vs_norm[1023] = 1.0;
    for (i = 1; i < 1024; ++i) {
        half *= 0.5;
        vs_scale[1023 + i] = sqrt(half);
        if (half != 0.0) vs_scale[1023 - i] = 1.0 / vs_scale[1023 + i];
        vs_norm[1023 + i] = half;
        if (half != 0.0) vs_norm[1023 - i] = 1.0 / half;
    }
    for (k = 1; k <= VS_NCHEB; ++k) {
        node[k] = (M_PI * (k - 0.5)) / VS_NCHEB;
        fval[k] = vs_target(cos(node[k]));
    }
    for (j = 0; j < VS_NCHEB; ++j) {
        double acc = 0.0;
        for (k = 1; k <= VS_NCHEB; ++k) acc += fval[k] * cos(node[k] * j);
        cheb[j] = (2.0 / VS_NCHEB) * acc;
    }
    /* power-basis coefficients on [-1,1], then shifted to x in [1,2) */
    p0 = 0.5 * cheb[0] - cheb[2] + cheb[4];
    p1 = cheb[1] - 3 * cheb[3] + 5 * cheb[5];
    p2 = 2 * cheb[2] - 8 * cheb[4];
    p3 = 4 * cheb[3] - 20 * cheb[5];
    p4 = 8 * cheb[4];
    p5 = 16 * cheb[5];
    vs_coef[0] = p0 - 3 * p1 + 9 * p2 - 27 * p3 + 81 * p4 - 243 * p5;
    vs_coef[1] = 2 * p1 - 12 * p2 + 54 * p3 - 216 * p4 + 810 * p5;
    vs_coef[2] = 4 * p2 - 36 * p3 + 216 * p4 - 1080 * p5;
    vs_coef[3] = 8 * p3 - 96 * p4 + 720 * p5;
    vs_coef[4] = 16 * p4 - 240 * p5;
    vs_coef[5] = 32 * p5;
    vs_ready = 1;
}

void orc_vsqrt1(int n, const double *r2, double *out)
{
    int i;
    if (!vs_ready) {
#pragma omp critical(orc_vs_setup)
        if (!vs_ready) vs_setup();
    }
    for (i = 0; i < n; ++i) {
        unsigned long long bits;
        int e;
        double x, s;
        memcpy(&bits, &r2[i], sizeof bits);
        e = (int)(bits >> 52);      /* sign + biased exponent, r2 >= 0 */
        x = r2[i] * vs_norm[e];
        s = vs_coef[0] + x * (vs_coef[1] + x * (vs_coef[2] + x * (vs_coef[3]
            + x * (vs_coef[4] + x * vs_coef[5]))));
        s *= 1.5 - 0.5 * x * s * s;
        out[i] = s * vs_scale[e];
    }
}

/* ======================================================================
 * Spline softening (grav.h:17-97, Hernquist & Katz 1989 eqs. A1/A2 with
 * Stadel's higher derivatives).  dih = 2/twoh, u = r*dih.
 * ====================================================================== */
static inline void spline_m(double dir, double r2, double twoh, double *pa,
                            double *pb)
{   /* SPLINEM, grav.h:50-71 */
    if (r2 < twoh * twoh) {
        double dih = 2.0 / twoh, u = dih / dir;
        if (u < 1.0) {
            *pa = dih * (7.0 / 5.0 - 2.0 / 3.0 * u * u + 3.0 / 10.0 * u * u * u * u
                         - 1.0 / 10.0 * u * u * u * u * u);
            *pb = dih * dih * dih * (4.0 / 3.0 - 6.0 / 5.0 * u * u + 1.0 / 2.0 * u * u * u);
        } else {
            *pa = -1.0 / 15.0 * dir + dih * (8.0 / 5.0 - 4.0 / 3.0 * u * u + u * u * u
                         - 3.0 / 10.0 * u * u * u * u + 1.0 / 30.0 * u * u * u * u * u);
            *pb = -1.0 / 15.0 * dir * dir * dir + dih * dih * dih * (8.0 / 3.0 - 3.0 * u
                         + 6.0 / 5.0 * u * u - 1.0 / 6.0 * u * u * u);
        }
    } else {
        *pa = dir;
        *pb = *pa * *pa * *pa;
    }
}

void orc_splineq(double dir, double r2, double twoh, double *pa, double *pb,
                 double *pc, double *pd)
{   /* SPLINEQ, grav.h:17-47 */
    if (r2 < twoh * twoh) {
        double dih = 2.0 / twoh, u = dih / dir;
        if (u < 1.0) {
            *pa = dih * (7.0 / 5.0 - 2.0 / 3.0 * u * u + 3.0 / 10.0 * u * u * u * u
                         - 1.0 / 10.0 * u * u * u * u * u);
            *pb = dih * dih * dih * (4.0 / 3.0 - 6.0 / 5.0 * u * u + 1.0 / 2.0 * u * u * u);
            *pc = dih * dih * dih * dih * dih * (12.0 / 5.0 - 3.0 / 2.0 * u);
            *pd = 3.0 / 2.0 * dih * dih * dih * dih * dih * dih * dir;
        } else {
            *pa = -1.0 / 15.0 * dir + dih * (8.0 / 5.0 - 4.0 / 3.0 * u * u + u * u * u
                         - 3.0 / 10.0 * u * u * u * u + 1.0 / 30.0 * u * u * u * u * u);
            *pb = -1.0 / 15.0 * dir * dir * dir + dih * dih * dih * (8.0 / 3.0 - 3.0 * u
                         + 6.0 / 5.0 * u * u - 1.0 / 6.0 * u * u * u);
            *pc = -1.0 / 5.0 * dir * dir * dir * dir * dir + 3.0 * dih * dih * dih * dih * dir
                  + dih * dih * dih * dih * dih * (-12.0 / 5.0 + 1.0 / 2.0 * u);
            *pd = -dir * dir * dir * dir * dir * dir * dir
                  + 3.0 * dih * dih * dih * dih * dir * dir * dir
                  - 1.0 / 2.0 * dih * dih * dih * dih * dih * dih * dir;
        }
    } else {
        *pa = dir;
        *pb = *pa * *pa * *pa;
        *pc = 3.0 * *pb * *pa * *pa;
        *pd = 5.0 * *pc * *pa * *pa;
    }
}

void orc_spline(double r2, double twoh, double *pa, double *pb)
{   /* SPLINE, grav.h:74-97 (sqrt form, used for intra-bucket pairs) */
    double r = sqrt(r2);
    if (r < twoh) {
        double dih = 2.0 / twoh, u = r * dih;
        if (u < 1.0) {
            *pa = dih * (7.0 / 5.0 - 2.0 / 3.0 * u * u + 3.0 / 10.0 * u * u * u * u
                         - 1.0 / 10.0 * u * u * u * u * u);
            *pb = dih * dih * dih * (4.0 / 3.0 - 6.0 / 5.0 * u * u + 1.0 / 2.0 * u * u * u);
        } else {
            double dir = 1.0 / r;
            *pa = -1.0 / 15.0 * dir + dih * (8.0 / 5.0 - 4.0 / 3.0 * u * u + u * u * u
                         - 3.0 / 10.0 * u * u * u * u + 1.0 / 30.0 * u * u * u * u * u);
            *pb = -1.0 / 15.0 * dir * dir * dir + dih * dih * dih * (8.0 / 3.0 - 3.0 * u
                         + 6.0 / 5.0 * u * u - 1.0 / 6.0 * u * u * u);
        }
    } else {
        *pa = 1.0 / r;
        *pb = *pa * *pa * *pa;
    }
}

/* ======================================================================
 * Tree (pkd.h:454-469 KDN, built by pkd.c:2373-2723)
 * ====================================================================== */
typedef struct {
    double r[3], m, soft;
    unsigned char active;    /* TYPE_ACTIVE */
    unsigned bits;           /* iActive type bits (pkd.h:288-329) */
    int src;                 /* input index */
} OPart;

typedef struct {
    int iDim, pLower, pUpper, iLower, iUpper;
    double fSplit, bmin[3], bmax[3], fMass, fSoft, r[3], fOpen2;
    /* pkdCalcCellStruct order (pkd.h:441-451) */
    double Q[6];    /* xx yy zz xy xz yz */
    double O[10];   /* xxx xyy xxy yyy xxz yyz xyz xzz yzz zzz */
    double H[15];   /* xxxx xyyy xxxy yyyy xxxz yyyz xxyy xxyz xyyz xxzz xyzz xzzz yyzz yzzz zzzz */
    double B[6];    /* Bmax B2 B3 B4 B5 B6 */
} ONode;

struct orc_tree {
    int n, nNodes, iRoot, nBucket, iOpenType, iOrder;
    double dCrit;
    OPart *p;
    ONode *c;
    int iFree;
};

/* Hoare partition about fSplit on axis d (pkdUpperPart, pkd.c:1109-1157).
 * Returns the first index of the upper part. */
static int upper_part(OPart *p, int d, double fSplit, int i, int j)
{
    OPart tmp;
    if (i > j) return i;
    i--; j++;
    while (++i < j && p[i].r[d] < fSplit);
    while (i < --j && p[j].r[d] >= fSplit);
    if (i >= j) return i;
    tmp = p[i]; p[i] = p[j]; p[j] = tmp;
    for (;;) {
        while (p[++i].r[d] < fSplit);
        while (p[--j].r[d] >= fSplit);
        if (i > j) return i;
        tmp = p[i]; p[i] = p[j]; p[j] = tmp;
    }
}

/* Type-first Hoare partition (pkdActiveTypeOrder, pkd.c:1304-1355): the
 * particles whose iActive has any bit of mask move to the front. */
static void active_order(OPart *p, int n, unsigned mask)
{
    int i = 0, j = n - 1;
    OPart tmp;
    if (i > j) return;
    i--; j++;
    while (++i < j && (p[i].bits & mask));
    while (i < --j && !(p[j].bits & mask));
    if (i >= j) return;
    tmp = p[i]; p[i] = p[j]; p[j] = tmp;
    for (;;) {
        while (p[++i].bits & mask);
        while (!(p[--j].bits & mask));
        if (i > j) return;
        tmp = p[i]; p[i] = p[j]; p[j] = tmp;
    }
}

/* Count nodes and partition (NumBinaryNodes, pkd.c:2373-2434). */
static int count_nodes(OPart *p, int nBucket, int lo, int hi)
{
    double bmin[3], bmax[3];
    int i, j, d, m, good = 0;
    if (lo > hi) return 0;
    for (j = 0; j < 3; ++j) bmin[j] = bmax[j] = p[lo].r[j];
    for (i = lo + 1; i <= hi; ++i)
        for (j = 0; j < 3; ++j) {
            if (p[i].r[j] < bmin[j]) bmin[j] = p[i].r[j];
            else if (p[i].r[j] > bmax[j]) bmax[j] = p[i].r[j];
        }
    for (j = 0; j < 3; ++j) if (bmax[j] > bmin[j]) good = 1;
    if (hi - lo + 1 > nBucket && good) {
        d = 0;
        for (j = 1; j < 3; ++j)
            if (bmax[j] - bmin[j] > bmax[d] - bmin[d]) d = j;
        m = upper_part(p, d, 0.5 * (bmin[d] + bmax[d]), lo, hi);
        return count_nodes(p, nBucket, lo, m - 1) + count_nodes(p, nBucket, m, hi) + 1;
    }
    return 1;
}

/* Multipole moments and B numbers about rcm (pkdCalcCell, pkd.c:2018-2135).
 * All orders are accumulated (the reference builds to max(iOrder,iEwOrder)). */
static void calc_cell(const OPart *p, ONode *c)
{
    int pj, k;
    for (k = 0; k < 6; ++k) c->Q[k] = 0.0;
    for (k = 0; k < 10; ++k) c->O[k] = 0.0;
    for (k = 0; k < 15; ++k) c->H[k] = 0.0;
    for (k = 0; k < 6; ++k) c->B[k] = 0.0;
    for (pj = c->pLower; pj <= c->pUpper; ++pj) {
        double m = p[pj].m;
        double dx = p[pj].r[0] - c->r[0];
        double dy = p[pj].r[1] - c->r[1];
        double dz = p[pj].r[2] - c->r[2];
        double d2 = dx * dx + dy * dy + dz * dz;
        double d1 = sqrt(d2);
        double *H = c->H, *O = c->O, *Q = c->Q, *B = c->B;
        if (d1 > B[0]) B[0] = d1;
        B[1] += m * d2;
        B[2] += m * d2 * d1;
        B[3] += m * d2 * d2;
        B[4] += m * d2 * d2 * d1;
        B[5] += m * d2 * d2 * d2;
        H[0] += m * (dx * dx * dx * dx - 6.0 / 7.0 * d2 * (dx * dx - 0.1 * d2));
        H[1] += m * (dx * dy * dy * dy - 3.0 / 7.0 * d2 * dx * dy);
        H[2] += m * (dx * dx * dx * dy - 3.0 / 7.0 * d2 * dx * dy);
        H[3] += m * (dy * dy * dy * dy - 6.0 / 7.0 * d2 * (dy * dy - 0.1 * d2));
        H[4] += m * (dx * dx * dx * dz - 3.0 / 7.0 * d2 * dx * dz);
        H[5] += m * (dy * dy * dy * dz - 3.0 / 7.0 * d2 * dy * dz);
        H[6] += m * (dx * dx * dy * dy - 1.0 / 7.0 * d2 * (dx * dx + dy * dy - 0.2 * d2));
        H[7] += m * (dx * dx * dy * dz - 1.0 / 7.0 * d2 * dy * dz);
        H[8] += m * (dx * dy * dy * dz - 1.0 / 7.0 * d2 * dx * dz);
        H[9] += m * (dx * dx * dz * dz - 1.0 / 7.0 * d2 * (dx * dx + dz * dz - 0.2 * d2));
        H[10] += m * (dx * dy * dz * dz - 1.0 / 7.0 * d2 * dx * dy);
        H[11] += m * (dx * dz * dz * dz - 3.0 / 7.0 * d2 * dx * dz);
        H[12] += m * (dy * dy * dz * dz - 1.0 / 7.0 * d2 * (dy * dy + dz * dz - 0.2 * d2));
        H[13] += m * (dy * dz * dz * dz - 3.0 / 7.0 * d2 * dy * dz);
        H[14] += m * (dz * dz * dz * dz - 6.0 / 7.0 * d2 * (dz * dz - 0.1 * d2));
        O[0] += m * (dx * dx * dx - 0.6 * d2 * dx);
        O[1] += m * (dx * dy * dy - 0.2 * d2 * dx);
        O[2] += m * (dx * dx * dy - 0.2 * d2 * dy);
        O[3] += m * (dy * dy * dy - 0.6 * d2 * dy);
        O[4] += m * (dx * dx * dz - 0.2 * d2 * dz);
        O[5] += m * (dy * dy * dz - 0.2 * d2 * dz);
        O[6] += m * dx * dy * dz;
        O[7] += m * (dx * dz * dz - 0.2 * d2 * dx);
        O[8] += m * (dy * dz * dz - 0.2 * d2 * dy);
        O[9] += m * (dz * dz * dz - 0.6 * d2 * dz);
        Q[0] += m * dx * dx;
        Q[1] += m * dy * dy;
        Q[2] += m * dz * dz;
        Q[3] += m * dx * dy;
        Q[4] += m * dx * dz;
        Q[5] += m * dy * dz;
    }
}

/* Opening radius (pkdCalcOpen, pkd.c:2228-2264); OPEN_JOSH and the minimal
 * Bmax form.  The OPEN_ABSPAR root bracketing is not restated. */
static double calc_open(const ONode *c, int iOpenType, double dCrit)
{
    double dOpen;
    if (iOpenType == ORC_OPEN_JOSH) {
        dOpen = 2 / sqrt(3.0) * c->B[0] / dCrit;
        if (dOpen < c->B[0]) dOpen = c->B[0];
    } else {
        dOpen = c->B[0];
    }
    return dOpen;
}

/* BuildBinary (pkd.c:2437-2587): preorder cell allocation, squeezed bounds,
 * midpoint split of the longest axis, c.o.m. from the children. */
static int build(orc_tree *t, int lo, int hi)
{
    OPart *p = t->p;
    ONode *c;
    int ic, i, j, d, m, good = 0;
    double bmin[3], bmax[3], fm, dOpen;
    if (lo > hi) return -1;
    ic = t->iFree++;
    c = &t->c[ic];
    c->pLower = lo;
    c->pUpper = hi;
    for (j = 0; j < 3; ++j) bmin[j] = bmax[j] = p[lo].r[j];
    for (i = lo + 1; i <= hi; ++i)
        for (j = 0; j < 3; ++j) {
            if (p[i].r[j] < bmin[j]) bmin[j] = p[i].r[j];
            else if (p[i].r[j] > bmax[j]) bmax[j] = p[i].r[j];
        }
    for (j = 0; j < 3; ++j) if (bmax[j] > bmin[j]) good = 1;
    for (j = 0; j < 3; ++j) { c->bmin[j] = bmin[j]; c->bmax[j] = bmax[j]; }
    if (hi - lo + 1 > t->nBucket && good) {
        int il, iu;
        d = 0;
        for (j = 1; j < 3; ++j)
            if (bmax[j] - bmin[j] > bmax[d] - bmin[d]) d = j;
        c->iDim = d;
        c->fSplit = 0.5 * (bmin[d] + bmax[d]);
        m = upper_part(p, d, c->fSplit, lo, hi);
        il = build(t, lo, m - 1);
        iu = build(t, m, hi);
        c = &t->c[ic];
        c->iLower = il;
        c->iUpper = iu;
        c->fMass = 0.0;
        c->fSoft = 0.0;
        for (j = 0; j < 3; ++j) c->r[j] = 0.0;
        if (il != -1) {
            fm = t->c[il].fMass;
            c->fMass += fm;
            c->fSoft += fm * t->c[il].fSoft;
            for (j = 0; j < 3; ++j) c->r[j] += fm * t->c[il].r[j];
        }
        if (iu != -1) {
            fm = t->c[iu].fMass;
            c->fMass += fm;
            c->fSoft += fm * t->c[iu].fSoft;
            for (j = 0; j < 3; ++j) c->r[j] += fm * t->c[iu].r[j];
        }
    } else {
        c->iDim = -1;
        c->fSplit = 0.0;
        c->iLower = -1;
        c->iUpper = -1;
        c->fMass = 0.0;
        c->fSoft = 0.0;
        for (j = 0; j < 3; ++j) c->r[j] = 0.0;
        for (i = lo; i <= hi; ++i) {
            fm = p[i].m;
            c->fMass += fm;
            c->fSoft += fm * p[i].soft;
            for (j = 0; j < 3; ++j) c->r[j] += fm * p[i].r[j];
        }
    }
    if (c->fMass > 0) {
        c->fSoft /= c->fMass;
        for (j = 0; j < 3; ++j) c->r[j] /= c->fMass;
    }
    calc_cell(p, c);
    dOpen = calc_open(c, t->iOpenType, t->dCrit);
    c->fOpen2 = dOpen * dOpen;
    return ic;
}

/* Threading (pkdThreadTree, pkd.c:2590-2620): iLower = first child,
 * iUpper = next node in a depth-first walk, -1 at the end. */
static void thread_tree(ONode *c, int ic, int inext)
{
    if (ic == -1) return;
    if (c[ic].iDim != -1) {
        int l = c[ic].iLower, u = c[ic].iUpper;
        if (u == -1) {
            thread_tree(c, l, inext);
        } else if (l == -1) {
            thread_tree(c, u, inext);
            c[ic].iLower = u;
        } else {
            thread_tree(c, l, u);
            thread_tree(c, u, inext);
        }
        c[ic].iUpper = inext;
    } else {
        c[ic].iLower = -1;
        c[ic].iUpper = inext;
    }
}

orc_tree *orc_build_binary(int n, const double *r, const double *m,
                           const double *soft, const unsigned *iActive,
                           int nBucket, int iOpenType, double dCrit, int iOrder)
{
    orc_tree *t = (orc_tree *)calloc(1, sizeof *t);
    int i, nn;
    if (!t) return NULL;
    t->n = n;
    t->nBucket = nBucket;
    t->iOpenType = iOpenType;
    t->dCrit = dCrit;
    t->iOrder = iOrder;
    t->p = (OPart *)malloc((size_t)(n > 0 ? n : 1) * sizeof(OPart));
    if (!t->p) { free(t); return NULL; }
    for (i = 0; i < n; ++i) {
        t->p[i].r[0] = r[3 * i];
        t->p[i].r[1] = r[3 * i + 1];
        t->p[i].r[2] = r[3 * i + 2];
        t->p[i].m = m[i];
        t->p[i].soft = soft[i];
        t->p[i].bits = iActive ? iActive[i] : 3u;
        t->p[i].active = (t->p[i].bits & 1u) != 0;
        t->p[i].src = i;
    }
    /* msrBuildTree orders particles TYPE_ACTIVE|TYPE_TREEACTIVE first
     * (master.c:4262), pkdBuildBinary then TYPE_TREEACTIVE first
     * (pkd.c:2639); the gravity tree takes all particles. */
    active_order(t->p, n, 3u);
    active_order(t->p, n, 2u);
    nn = count_nodes(t->p, nBucket, 0, n - 1);
    if (nn == 0) nn = 1;
    t->nNodes = nn;
    t->c = (ONode *)calloc((size_t)nn + 1, sizeof(ONode));
    if (!t->c) { free(t->p); free(t); return NULL; }
    t->iFree = 0;
    if (n > 0) {
        t->iRoot = build(t, 0, n - 1);
    } else {
        ONode *c = &t->c[0];
        int j;
        t->iRoot = 0;
        c->pLower = 0; c->pUpper = -1; c->iDim = -1;
        c->iLower = c->iUpper = -1;
        for (j = 0; j < 3; ++j) { c->bmin[j] = DBL_MAX; c->bmax[j] = -DBL_MAX; }
    }
    thread_tree(t->c, t->iRoot, -1);
    return t;
}

void orc_tree_free(orc_tree *t)
{
    if (!t) return;
    free(t->p);
    free(t->c);
    free(t);
}

int orc_tree_nnodes(const orc_tree *t) { return t->nNodes; }
int orc_tree_root(const orc_tree *t) { return t->iRoot; }

void orc_tree_perm(const orc_tree *t, int *perm)
{
    int i;
    for (i = 0; i < t->n; ++i) perm[i] = t->p[i].src;
}

void orc_tree_nodes(const orc_tree *t, int *ints, double *dbls)
{
    int i, k;
    for (i = 0; i < t->nNodes; ++i) {
        const ONode *c = &t->c[i];
        double *o = dbls + (size_t)ORC_NODE_NDBL * i;
        ints[6 * i + 0] = c->iDim;
        ints[6 * i + 1] = c->pLower;
        ints[6 * i + 2] = c->pUpper;
        ints[6 * i + 3] = c->iLower;
        ints[6 * i + 4] = c->iUpper;
        ints[6 * i + 5] = 0;
        o[0] = c->fSplit;
        for (k = 0; k < 3; ++k) { o[1 + k] = c->bmin[k]; o[4 + k] = c->bmax[k]; }
        o[7] = c->fMass;
        o[8] = c->fSoft;
        for (k = 0; k < 3; ++k) o[9 + k] = c->r[k];
        o[12] = c->fOpen2;
        for (k = 0; k < 6; ++k) o[13 + k] = c->Q[k];
        for (k = 0; k < 10; ++k) o[19 + k] = c->O[k];
        for (k = 0; k < 15; ++k) o[29 + k] = c->H[k];
        /* Bmax, B2, B3 (B4..B6 are not exported) */
        o[44] = c->B[0]; o[45] = c->B[1]; o[46] = c->B[2];
    }
}

/* ======================================================================
 * Bucket walk (pkdLocalWalk, walk.c:58-178, with the top tree of one
 * process being the local root, walk.c:342-349).
 * ====================================================================== */
typedef struct { double bmin[3], bmax[3]; } OBnd;

static inline int intersect_np(const OBnd *b, double fBall2, double x,
                               double y, double z)
{   /* INTERSECTNP, walk.h:12-30 */
    double dx = b->bmin[0] - x, dx1 = x - b->bmax[0];
    double dy = b->bmin[1] - y, dy1 = y - b->bmax[1];
    double dz = b->bmin[2] - z, dz1 = z - b->bmax[2];
    double d2;
    if (dx > 0.0) d2 = dx * dx;
    else if (dx1 > 0.0) d2 = dx1 * dx1;
    else d2 = 0.0;
    if (dy > 0.0) d2 += dy * dy;
    else if (dy1 > 0.0) d2 += dy1 * dy1;
    if (dz > 0.0) d2 += dz * dz;
    else if (dz1 > 0.0) d2 += dz1 * dz1;
    return d2 <= fBall2;
}

/* Active bounds and max softening of a bucket (pkd.c:2916-2937,
 * walk.c:319-324).  Returns the number of active particles. */
static int bucket_setup(const orc_tree *t, int ib, OBnd *b, double *fSoftMax)
{
    const ONode *c = &t->c[ib];
    int i, j, n = 0;
    for (j = 0; j < 3; ++j) { b->bmin[j] = DBL_MAX; b->bmax[j] = -DBL_MAX; }
    for (i = c->pLower; i <= c->pUpper; ++i) {
        if (!t->p[i].active) continue;
        ++n;
        for (j = 0; j < 3; ++j) {
            if (t->p[i].r[j] < b->bmin[j]) b->bmin[j] = t->p[i].r[j];
            if (t->p[i].r[j] > b->bmax[j]) b->bmax[j] = t->p[i].r[j];
        }
    }
    *fSoftMax = 0.0;
    for (i = c->pLower; i <= c->pUpper; ++i)
        if (t->p[i].soft > *fSoftMax) *fSoftMax = t->p[i].soft;
    return n;
}

void orc_walk_bucket(orc_tree *t, int iBucket, int iOrder, int cap,
                     int *part, int *soft, int *newt, int counts[3])
{
    OBnd b;
    double fSoftMax;
    int nPart = 0, nSoft = 0, nNewt = 0, ic = t->iRoot;
    (void)iOrder;
    bucket_setup(t, iBucket, &b, &fSoftMax);
    while (ic != -1) {
        const ONode *c = &t->c[ic];
        double x = c->r[0], y = c->r[1], z = c->r[2];
        int open = intersect_np(&b, c->fOpen2, x, y, z);
        if (c->pUpper - c->pLower < 3) open = 1;
        if (open) {
            if (c->iLower != -1) { ic = c->iLower; continue; }
            if (ic != iBucket) {
                int pj;
                for (pj = c->pLower; pj <= c->pUpper; ++pj, ++nPart)
                    if (nPart < cap) part[nPart] = pj;
            }
            ic = c->iUpper;
        } else {
            double twoh2 = c->fSoft + fSoftMax;
            int isSoft;
            twoh2 *= twoh2;
            if (twoh2 < c->fOpen2) isSoft = 0;
            else isSoft = intersect_np(&b, twoh2, x, y, z);
            if (isSoft) { if (nSoft < cap) soft[nSoft] = ic; ++nSoft; }
            else { if (nNewt < cap) newt[nNewt] = ic; ++nNewt; }
            ic = c->iUpper;
        }
    }
    counts[0] = nPart;
    counts[1] = nSoft;
    counts[2] = nNewt;
}

/* ======================================================================
 * Bucket interaction (pkdBucketInteract, grav.c:23-249), iOrder <= 4.
 * ====================================================================== */
static const int qeval_flop[5] = {10, 10, 41, 120, 277};   /* qeval.h:66 */

/* Reduced multipole evaluation (QEVAL, qeval.h:21-64), falling through from
 * iOrder down to the monopole. */
static inline void qeval(int iOrder, double m, const double *tq, const double *O,
                         const double *H, const double *gam, double dx, double dy, double dz,
                         double *ax, double *ay, double *az, double *pot)
{
    double Qmirx, Qmiry, Qmirz, Qmir, Qta = 0.0;
    /* traceless quadrupole of SETILIST (walk.c:41-48): xx yy zz xy xz yz */
    double xx = tq[0], yy = tq[1], zz = tq[2], xy = tq[3], xz = tq[4], yz = tq[5];
    switch (iOrder) {
    case 4: {
        double xxxx = H[0], xyyy = H[1], xxxy = H[2], yyyy = H[3], xxxz = H[4],
               yyyz = H[5], xxyy = H[6], xxyz = H[7], xyyz = H[8], xxzz = H[9],
               xyzz = H[10], xzzz = H[11], yyzz = H[12], yzzz = H[13], zzzz = H[14];
        Qmirx = (1.0 / 6.0) * (xzzz * dz * dz * dz + 3 * xyzz * dy * dz * dz + 3 * xyyz * dy * dy * dz
                + xyyy * dy * dy * dy + 3 * xxzz * dx * dz * dz + 6 * xxyz * dx * dy * dz
                + 3 * xxyy * dx * dy * dy + 3 * xxxz * dx * dx * dz + 3 * xxxy * dx * dx * dy
                + xxxx * dx * dx * dx);
        Qmiry = (1.0 / 6.0) * (yzzz * dz * dz * dz + 3 * xyzz * dx * dz * dz + 3 * xxyz * dx * dx * dz
                + xxxy * dx * dx * dx + 3 * yyzz * dy * dz * dz + 6 * xyyz * dx * dy * dz
                + 3 * xxyy * dx * dx * dy + 3 * yyyz * dy * dy * dz + 3 * xyyy * dx * dy * dy
                + yyyy * dy * dy * dy);
        Qmirz = (1.0 / 6.0) * (yyyz * dy * dy * dy + 3 * xyyz * dx * dy * dy + 3 * xxyz * dx * dx * dy
                + xxxz * dx * dx * dx + 3 * yyzz * dy * dy * dz + 6 * xyzz * dx * dy * dz
                + 3 * xxzz * dx * dx * dz + 3 * yzzz * dy * dz * dz + 3 * xzzz * dx * dz * dz
                + zzzz * dz * dz * dz);
        Qmir = (1.0 / 4.0) * (Qmirx * dx + Qmiry * dy + Qmirz * dz);
        *pot -= gam[4] * Qmir;
        Qta += gam[5] * Qmir;
        *ax += gam[4] * Qmirx;
        *ay += gam[4] * Qmiry;
        *az += gam[4] * Qmirz;
    }
    /* fall through */
    case 3: {
        double xxx = O[0], xyy = O[1], xxy = O[2], yyy = O[3], xxz = O[4],
               yyz = O[5], xyz = O[6], xzz = O[7], yzz = O[8], zzz = O[9];
        Qmirx = (1.0 / 2.0) * (xzz * dz * dz + 2 * xyz * dy * dz + xyy * dy * dy
                + 2 * xxz * dx * dz + 2 * xxy * dx * dy + xxx * dx * dx);
        Qmiry = (1.0 / 2.0) * (yzz * dz * dz + 2 * xyz * dx * dz + xxy * dx * dx
                + 2 * yyz * dy * dz + 2 * xyy * dx * dy + yyy * dy * dy);
        Qmirz = (1.0 / 2.0) * (yyz * dy * dy + 2 * xyz * dx * dy + xxz * dx * dx
                + 2 * yzz * dy * dz + 2 * xzz * dx * dz + zzz * dz * dz);
        Qmir = (1.0 / 3.0) * (Qmirx * dx + Qmiry * dy + Qmirz * dz);
        *pot -= gam[3] * Qmir;
        Qta += gam[4] * Qmir;
        *ax += gam[3] * Qmirx;
        *ay += gam[3] * Qmiry;
        *az += gam[3] * Qmirz;
    }
    /* fall through */
    case 2:
        Qmirx = (1.0 / 1.0) * (xz * dz + xy * dy + xx * dx);
        Qmiry = (1.0 / 1.0) * (yz * dz + xy * dx + yy * dy);
        Qmirz = (1.0 / 1.0) * (yz * dy + xz * dx + zz * dz);
        Qmir = (1.0 / 2.0) * (Qmirx * dx + Qmiry * dy + Qmirz * dz);
        *pot -= gam[2] * Qmir;
        Qta += gam[3] * Qmir;
        *ax += gam[2] * Qmirx;
        *ay += gam[2] * Qmiry;
        *az += gam[2] * Qmirz;
        /* fall through */
    case 1:
    default:
        *pot -= gam[0] * m;
        Qta += gam[1] * m;
        *ax -= dx * Qta;
        *ay -= dy * Qta;
        *az -= dz * Qta;
    }
}

/* Traceless quadrupole as SETILIST stores it (walk.c:41-48). */
static inline void traceless(const ONode *c, double *tq)
{
    double tr = c->Q[0] + c->Q[1] + c->Q[2];
    tq[0] = c->Q[0] - tr / 3.0;
    tq[1] = c->Q[1] - tr / 3.0;
    tq[2] = c->Q[2] - tr / 3.0;
    tq[3] = c->Q[3];
    tq[4] = c->Q[4];
    tq[5] = c->Q[5];
}

typedef struct {
    int *part, *soft, *newt, cap;
    double *d2a, *sq;
    double *po, *so, *no;    /* replica offset of each entry (3 per entry) */
} OScratch;

static int scratch_fit(OScratch *s, int need)
{
    if (need <= s->cap) return 1;
    s->cap = need * 2;
    s->part = (int *)realloc(s->part, (size_t)s->cap * sizeof(int));
    s->soft = (int *)realloc(s->soft, (size_t)s->cap * sizeof(int));
    s->newt = (int *)realloc(s->newt, (size_t)s->cap * sizeof(int));
    s->d2a = (double *)realloc(s->d2a, (size_t)s->cap * sizeof(double));
    s->sq = (double *)realloc(s->sq, (size_t)s->cap * sizeof(double));
    s->po = (double *)realloc(s->po, (size_t)s->cap * 3 * sizeof(double));
    s->so = (double *)realloc(s->so, (size_t)s->cap * 3 * sizeof(double));
    s->no = (double *)realloc(s->no, (size_t)s->cap * 3 * sizeof(double));
    return s->part && s->soft && s->newt && s->d2a && s->sq && s->po && s->so && s->no;
}

/* a, pot, dtg are in storage order here (length t->n). */
static int bucket_interact(orc_tree *t, int ib, int iOrder, OScratch *s,
                           int nPart, int nSoft, int nNewt, double *a,
                           double *pot, double *dtg)
{
    const ONode *bk = &t->c[ib];
    OPart *p = t->p;
    int lo = bk->pLower, n = bk->pUpper - bk->pLower + 1, i, j, nActive = 0;
    for (i = lo; i < lo + n; ++i) {
        double ax = 0.0, ay = 0.0, az = 0.0, fPot = 0.0;
        double x, y, z, h, dx, dy, dz, twoh, aa, bb, cc, dd, idt2;
        if (!p[i].active) continue;
        ++nActive;
        x = p[i].r[0]; y = p[i].r[1]; z = p[i].r[2];
        h = p[i].soft;
        /* particle-particle (grav.c:81-108) */
        /* an entry of replica o sits at r + o (ILP x = r + rOffset, walk.c:106-110) */
        for (j = 0; j < nPart; ++j) {
            const OPart *q = &p[s->part[j]];
            const double *o = s->po + 3 * j;
            dx = x - (q->r[0] + o[0]); dy = y - (q->r[1] + o[1]); dz = z - (q->r[2] + o[2]);
            s->d2a[j] = dx * dx + dy * dy + dz * dz;
        }
        if (nPart > 0) orc_vsqrt1(nPart, s->d2a, s->sq);
        for (j = 0; j < nPart; ++j) {
            const OPart *q = &p[s->part[j]];
            const double *o = s->po + 3 * j;
            dx = x - (q->r[0] + o[0]); dy = y - (q->r[1] + o[1]); dz = z - (q->r[2] + o[2]);
            twoh = h + q->soft;
            spline_m(s->sq[j], s->d2a[j], twoh, &aa, &bb);
            idt2 = (p[i].m + q->m) * bb;
            if (idt2 > dtg[i]) dtg[i] = idt2;
            aa *= q->m;
            bb *= q->m;
            fPot -= aa;
            ax -= dx * bb;
            ay -= dy * bb;
            az -= dz * bb;
        }
        /* softened cells, monopole + raw quadrupole (grav.c:117-150) */
        for (j = 0; j < nSoft; ++j) {
            const ONode *c = &t->c[s->soft[j]];
            const double *o = s->so + 3 * j;
            dx = x - (c->r[0] + o[0]); dy = y - (c->r[1] + o[1]); dz = z - (c->r[2] + o[2]);
            s->d2a[j] = dx * dx + dy * dy + dz * dz;
        }
        if (nSoft > 0) orc_vsqrt1(nSoft, s->d2a, s->sq);
        for (j = 0; j < nSoft; ++j) {
            const ONode *c = &t->c[s->soft[j]];
            const double *o = s->so + 3 * j;
            double qirx, qiry, qirz, qir, tr, qir3;
            const double *Q = c->Q;   /* xx yy zz xy xz yz */
            dx = x - (c->r[0] + o[0]); dy = y - (c->r[1] + o[1]); dz = z - (c->r[2] + o[2]);
            twoh = h + c->fSoft;
            orc_splineq(s->sq[j], s->d2a[j], twoh, &aa, &bb, &cc, &dd);
            qirx = Q[0] * dx + Q[3] * dy + Q[4] * dz;
            qiry = Q[3] * dx + Q[1] * dy + Q[5] * dz;
            qirz = Q[4] * dx + Q[5] * dy + Q[2] * dz;
            qir = 0.5 * (qirx * dx + qiry * dy + qirz * dz);
            tr = 0.5 * (Q[0] + Q[1] + Q[2]);
            qir3 = bb * c->fMass + dd * qir - cc * tr;
            fPot -= aa * c->fMass + cc * qir - bb * tr;
            ax -= qir3 * dx - cc * qirx;
            ay -= qir3 * dy - cc * qiry;
            az -= qir3 * dz - cc * qirz;
            idt2 = (p[i].m + c->fMass) * bb;
            if (idt2 > dtg[i]) dtg[i] = idt2;
        }
        /* Newtonian cells, reduced multipoles (grav.c:164-191) */
        for (j = 0; j < nNewt; ++j) {
            const ONode *c = &t->c[s->newt[j]];
            const double *o = s->no + 3 * j;
            dx = x - (c->r[0] + o[0]); dy = y - (c->r[1] + o[1]); dz = z - (c->r[2] + o[2]);
            s->d2a[j] = dx * dx + dy * dy + dz * dz;
        }
        if (nNewt > 0) orc_vsqrt1(nNewt, s->d2a, s->sq);
        for (j = 0; j < nNewt; ++j) {
            const ONode *c = &t->c[s->newt[j]];
            const double *o = s->no + 3 * j;
            double gam[6], dir2, tq[6];
            dx = x - (c->r[0] + o[0]); dy = y - (c->r[1] + o[1]); dz = z - (c->r[2] + o[2]);
            gam[0] = s->sq[j];
            dir2 = gam[0] * gam[0];
            gam[1] = gam[0] * dir2;
            gam[2] = 3 * gam[1] * dir2;
            gam[3] = 5 * gam[2] * dir2;
            gam[4] = 7 * gam[3] * dir2;
            gam[5] = 9 * gam[4] * dir2;
            traceless(c, tq);
            qeval(iOrder, c->fMass, tq, c->O, c->H, gam, dx, dy, dz, &ax, &ay, &az, &fPot);
            idt2 = (p[i].m + c->fMass) * gam[1];
            if (idt2 > dtg[i]) dtg[i] = idt2;
        }
        pot[i] += fPot;
        a[3 * i] += ax;
        a[3 * i + 1] += ay;
        a[3 * i + 2] += az;
    }
    /* intra-bucket symmetric pairs (grav.c:211-242) */
    for (i = lo; i < lo + n - 1; ++i) {
        for (j = i + 1; j < lo + n; ++j) {
            double dx, dy, dz, d2, twoh, aa, bb, idt2;
            if (!p[i].active && !p[j].active) continue;
            dx = p[j].r[0] - p[i].r[0];
            dy = p[j].r[1] - p[i].r[1];
            dz = p[j].r[2] - p[i].r[2];
            d2 = dx * dx + dy * dy + dz * dz;
            twoh = p[i].soft + p[j].soft;
            orc_spline(d2, twoh, &aa, &bb);
            idt2 = (p[i].m + p[j].m) * bb;
            if (p[j].active) {
                pot[j] -= aa * p[i].m;
                a[3 * j] -= dx * bb * p[i].m;
                a[3 * j + 1] -= dy * bb * p[i].m;
                a[3 * j + 2] -= dz * bb * p[i].m;
                if (idt2 > dtg[j]) dtg[j] = idt2;
            }
            if (p[i].active) {
                pot[i] -= aa * p[j].m;
                a[3 * i] += dx * bb * p[j].m;
                a[3 * i + 1] += dy * bb * p[j].m;
                a[3 * i + 2] += dz * bb * p[j].m;
                if (idt2 > dtg[i]) dtg[i] = idt2;
            }
        }
    }
    /* flop model (grav.c:246-247) */
    return nActive * ((nPart + n) * 38 + nSoft * 82 + nNewt * (35 + qeval_flop[iOrder]));
}

/* One replica of pkdLocalWalk (walk.c:58-178) for bucket iBucket: cell
 * centres and particles seen at r + ro; the bucket's own particles enter the
 * PP list only in a replica (bRep).  Entries are appended from the counts
 * in n[3] (stored while below cap, counted always). */
static void walk_rep(const orc_tree *t, int iBucket, const OBnd *b, double fSoftMax,
                     const double *ro, int bRep, OScratch *s, int n[3])
{
    int ic = t->iRoot, k;
    while (ic != -1) {
        const ONode *c = &t->c[ic];
        double x = c->r[0] + ro[0], y = c->r[1] + ro[1], z = c->r[2] + ro[2];
        int open = intersect_np(b, c->fOpen2, x, y, z);
        if (c->pUpper - c->pLower < 3) open = 1;
        if (open) {
            if (c->iLower != -1) { ic = c->iLower; continue; }
            if (ic != iBucket || bRep) {
                int pj;
                for (pj = c->pLower; pj <= c->pUpper; ++pj, ++n[0])
                    if (n[0] < s->cap) {
                        s->part[n[0]] = pj;
                        for (k = 0; k < 3; ++k) s->po[3 * n[0] + k] = ro[k];
                    }
            }
            ic = c->iUpper;
        } else {
            double twoh2 = c->fSoft + fSoftMax;
            int isSoft;
            twoh2 *= twoh2;
            if (twoh2 < c->fOpen2) isSoft = 0;
            else isSoft = intersect_np(b, twoh2, x, y, z);
            if (isSoft) {
                if (n[1] < s->cap) {
                    s->soft[n[1]] = ic;
                    for (k = 0; k < 3; ++k) s->so[3 * n[1] + k] = ro[k];
                }
                ++n[1];
            } else {
                if (n[2] < s->cap) {
                    s->newt[n[2]] = ic;
                    for (k = 0; k < 3; ++k) s->no[3 * n[2] + k] = ro[k];
                }
                ++n[2];
            }
            ic = c->iUpper;
        }
    }
}

/* pkdBucketWalk (walk.c:306-439): the replicas -nReps..nReps of each
 * periodic dimension (L[d] > 0), x outermost. */
static void walk_all_reps(const orc_tree *t, int iBucket, const OBnd *b, double fSoftMax,
                          int nReps, const double *L, OScratch *s, int n[3])
{
    int ix, iy, iz;
    n[0] = n[1] = n[2] = 0;
    for (ix = -nReps; ix <= nReps; ++ix) {
        if (ix && !(L && L[0] > 0)) continue;
        for (iy = -nReps; iy <= nReps; ++iy) {
            if (iy && !(L && L[1] > 0)) continue;
            for (iz = -nReps; iz <= nReps; ++iz) {
                double ro[3];
                if (iz && !(L && L[2] > 0)) continue;
                ro[0] = L ? ix * L[0] : 0.0;
                ro[1] = L ? iy * L[1] : 0.0;
                ro[2] = L ? iz * L[2] : 0.0;
                walk_rep(t, iBucket, b, fSoftMax, ro, ix || iy || iz, s, n);
            }
        }
    }
}

/* ----------------------------------------------------------------- Ewald */
/* The root's moments as pkdCalcRoot + pkdDistribRoot leave them in
 * ilcnRoot (pkd.c:4395-4493): mass and c.o.m. of the root, its raw
 * quadrupole (trace not removed), and the complete octupole and
 * hexadecapole about the c.o.m. (d2 = 0 in pkd.c:4447-4469). */
typedef struct {
    double m, x, y, z;
    double Q[6];    /* xx yy zz xy xz yz */
    double O[10];   /* xxx xyy xxy yyy xxz yyz xyz xzz yzz zzz */
    double H[15];   /* xxxx xyyy xxxy yyyy xxxz yyyz xxyy xxyz xyyz xxzz xyzz xzzz yyzz yzzz zzzz */
} OMom;

static void calc_root(const orc_tree *t, OMom *mom)
{
    const ONode *r = &t->c[t->iRoot];
    int pj, k;
    memset(mom, 0, sizeof(*mom));
    for (pj = r->pLower; pj <= r->pUpper; ++pj) {
        double m = t->p[pj].m;
        double dx = t->p[pj].r[0] - r->r[0], dy = t->p[pj].r[1] - r->r[1],
               dz = t->p[pj].r[2] - r->r[2];
        double *H = mom->H, *O = mom->O;
        H[0] += m * (dx * dx * dx * dx);
        H[1] += m * (dx * dy * dy * dy);
        H[2] += m * (dx * dx * dx * dy);
        H[3] += m * (dy * dy * dy * dy);
        H[4] += m * (dx * dx * dx * dz);
        H[5] += m * (dy * dy * dy * dz);
        H[6] += m * (dx * dx * dy * dy);
        H[7] += m * (dx * dx * dy * dz);
        H[8] += m * (dx * dy * dy * dz);
        H[9] += m * (dx * dx * dz * dz);
        H[10] += m * (dx * dy * dz * dz);
        H[11] += m * (dx * dz * dz * dz);
        H[12] += m * (dy * dy * dz * dz);
        H[13] += m * (dy * dz * dz * dz);
        H[14] += m * (dz * dz * dz * dz);
        O[0] += m * (dx * dx * dx);
        O[1] += m * (dx * dy * dy);
        O[2] += m * (dx * dx * dy);
        O[3] += m * (dy * dy * dy);
        O[4] += m * (dx * dx * dz);
        O[5] += m * (dy * dy * dz);
        O[6] += m * dx * dy * dz;
        O[7] += m * (dx * dz * dz);
        O[8] += m * (dy * dz * dz);
        O[9] += m * (dz * dz * dz);
    }
    mom->m = r->fMass;
    mom->x = r->r[0];
    mom->y = r->r[1];
    mom->z = r->r[2];
    for (k = 0; k < 6; ++k) mom->Q[k] = r->Q[k];
}

/* MEVAL (meval.h:21-80): the complete multipole at (dx, dy, dz) from its
 * centre, falling through from iOrder to the monopole. */
static void meval(int iOrder, const OMom *mom, const double *gam, double dx, double dy,
                  double dz, double *ax, double *ay, double *az, double *fPot)
{
    const double *Q = mom->Q, *O = mom->O, *H = mom->H;
    const double xx = Q[0], yy = Q[1], zz = Q[2], xy = Q[3], xz = Q[4], yz = Q[5];
    double Qmirx, Qmiry, Qmirz, Qmir, Qta = 0.0;
    switch (iOrder) {
    case 4: {
        const double xxxx = H[0], xyyy = H[1], xxxy = H[2], yyyy = H[3], xxxz = H[4],
                     yyyz = H[5], xxyy = H[6], xxyz = H[7], xyyz = H[8], xxzz = H[9],
                     xyzz = H[10], xzzz = H[11], yyzz = H[12], yzzz = H[13], zzzz = H[14];
        double Qxx, Qyy, Qzz, Qxy, Qxz, Qyz, Qtr, Qhx, Qhy, Qhz, Qh;
        Qmirx = (1.0 / 6.0) * (xzzz * dz * dz * dz + 3 * xyzz * dy * dz * dz + 3 * xyyz * dy * dy * dz
                + xyyy * dy * dy * dy + 3 * xxzz * dx * dz * dz + 6 * xxyz * dx * dy * dz
                + 3 * xxyy * dx * dy * dy + 3 * xxxz * dx * dx * dz + 3 * xxxy * dx * dx * dy
                + xxxx * dx * dx * dx);
        Qmiry = (1.0 / 6.0) * (yzzz * dz * dz * dz + 3 * xyzz * dx * dz * dz + 3 * xxyz * dx * dx * dz
                + xxxy * dx * dx * dx + 3 * yyzz * dy * dz * dz + 6 * xyyz * dx * dy * dz
                + 3 * xxyy * dx * dx * dy + 3 * yyyz * dy * dy * dz + 3 * xyyy * dx * dy * dy
                + yyyy * dy * dy * dy);
        Qmirz = (1.0 / 6.0) * (yyyz * dy * dy * dy + 3 * xyyz * dx * dy * dy + 3 * xxyz * dx * dx * dy
                + xxxz * dx * dx * dx + 3 * yyzz * dy * dy * dz + 6 * xyzz * dx * dy * dz
                + 3 * xxzz * dx * dx * dz + 3 * yzzz * dy * dz * dz + 3 * xzzz * dx * dz * dz
                + zzzz * dz * dz * dz);
        Qmir = (1.0 / 4.0) * (Qmirx * dx + Qmiry * dy + Qmirz * dz);
        Qxx = xxxx + xxyy + xxzz;
        Qxy = xxxy + xyyy + xyzz;
        Qxz = xxxz + xyyz + xzzz;
        Qyy = xxyy + yyyy + yyzz;
        Qyz = xxyz + yyyz + yzzz;
        Qzz = xxzz + yyzz + zzzz;
        Qtr = (1.0 / 8.0) * (Qxx + Qyy + Qzz);
        Qhx = 0.5 * (Qxx * dx + Qxy * dy + Qxz * dz);
        Qhy = 0.5 * (Qxy * dx + Qyy * dy + Qyz * dz);
        Qhz = 0.5 * (Qxz * dx + Qyz * dy + Qzz * dz);
        Qh = 0.5 * (Qhx * dx + Qhy * dy + Qhz * dz);
        *fPot -= gam[4] * Qmir - gam[3] * Qh + gam[2] * Qtr;
        Qta += gam[5] * Qmir - gam[4] * Qh + gam[3] * Qtr;
        *ax += gam[4] * Qmirx - gam[3] * Qhx;
        *ay += gam[4] * Qmiry - gam[3] * Qhy;
        *az += gam[4] * Qmirz - gam[3] * Qhz;
    }
    /* fall through */
    case 3: {
        const double xxx = O[0], xyy = O[1], xxy = O[2], yyy = O[3], xxz = O[4], yyz = O[5],
                     xyz = O[6], xzz = O[7], yzz = O[8], zzz = O[9];
        double Qx, Qy, Qz, Qtr;
        Qmirx = (1.0 / 2.0) * (xzz * dz * dz + 2 * xyz * dy * dz + xyy * dy * dy
                + 2 * xxz * dx * dz + 2 * xxy * dx * dy + xxx * dx * dx);
        Qmiry = (1.0 / 2.0) * (yzz * dz * dz + 2 * xyz * dx * dz + xxy * dx * dx
                + 2 * yyz * dy * dz + 2 * xyy * dx * dy + yyy * dy * dy);
        Qmirz = (1.0 / 2.0) * (yyz * dy * dy + 2 * xyz * dx * dy + xxz * dx * dx
                + 2 * yzz * dy * dz + 2 * xzz * dx * dz + zzz * dz * dz);
        Qmir = (1.0 / 3.0) * (Qmirx * dx + Qmiry * dy + Qmirz * dz);
        Qx = 0.5 * (xxx + xyy + xzz);
        Qy = 0.5 * (xxy + yyy + yzz);
        Qz = 0.5 * (xxz + yyz + zzz);
        Qtr = Qx * dx + Qy * dy + Qz * dz;
        *fPot -= gam[3] * Qmir - gam[2] * Qtr;
        Qta += gam[4] * Qmir - gam[3] * Qtr;
        *ax += gam[3] * Qmirx - gam[2] * Qx;
        *ay += gam[3] * Qmiry - gam[2] * Qy;
        *az += gam[3] * Qmirz - gam[2] * Qz;
    }
    /* fall through */
    case 2: {
        double Qtr;
        Qmirx = (1.0 / 1.0) * (xz * dz + xy * dy + xx * dx);
        Qmiry = (1.0 / 1.0) * (yz * dz + xy * dx + yy * dy);
        Qmirz = (1.0 / 1.0) * (yz * dy + xz * dx + zz * dz);
        Qmir = (1.0 / 2.0) * (Qmirx * dx + Qmiry * dy + Qmirz * dz);
        Qtr = 0.5 * (xx + yy + zz);
        *fPot -= gam[2] * Qmir - gam[1] * Qtr;
        Qta += gam[3] * Qmir - gam[2] * Qtr;
        *ax += gam[2] * Qmirx;
        *ay += gam[2] * Qmiry;
        *az += gam[2] * Qmirz;
    }
    /* fall through */
    case 1:
    default:
        *fPot -= gam[0] * mom->m;
        Qta += gam[1] * mom->m;
        *ax -= dx * Qta;
        *ay -= dy * Qta;
        *az -= dz * Qta;
    }
}

typedef struct { double hx, hy, hz, hCfac, hSfac; } OEwt;

/* pkdEwaldInit (ewald.c:180-247): the h-loop table, |h| <= fhCut, h != 0. */
static int ewald_init(const OMom *mom, double L, double fhCut, int iOrder, OEwt **pewt)
{
    int hReps = (int)ceil(fhCut), hx, hy, hz, i = 0, cap = 64;
    double alpha = 2.0 / L, k4 = M_PI * M_PI / (alpha * alpha * L * L);
    OEwt *ewt = (OEwt *)malloc(cap * sizeof(OEwt));
    for (hx = -hReps; hx <= hReps; ++hx)
        for (hy = -hReps; hy <= hReps; ++hy)
            for (hz = -hReps; hz <= hReps; ++hz) {
                int h2 = hx * hx + hy * hy + hz * hz, k;
                double gam[6], ax, ay, az, mfacc, mfacs, tq[6];
                if (h2 == 0) continue;
                if (h2 > fhCut * fhCut) continue;
                if (i == cap) {
                    cap *= 2;
                    ewt = (OEwt *)realloc(ewt, cap * sizeof(OEwt));
                }
                /* QEVAL (qeval.h) over the root moments as ilcnRoot holds them */
                for (k = 0; k < 6; ++k) tq[k] = mom->Q[k];
                gam[0] = exp(-k4 * h2) / (M_PI * h2 * L);
                gam[1] = 2 * M_PI / L * gam[0];
                gam[2] = -2 * M_PI / L * gam[1];
                gam[3] = 2 * M_PI / L * gam[2];
                gam[4] = -2 * M_PI / L * gam[3];
                gam[5] = 2 * M_PI / L * gam[4];
                gam[1] = 0.0;
                gam[3] = 0.0;
                gam[5] = 0.0;
                ax = ay = az = 0.0;
                mfacc = 0.0;
                qeval(iOrder, mom->m, tq, mom->O, mom->H, gam, hx, hy, hz, &ax, &ay, &az, &mfacc);
                gam[0] = exp(-k4 * h2) / (M_PI * h2 * L);
                gam[1] = 2 * M_PI / L * gam[0];
                gam[2] = -2 * M_PI / L * gam[1];
                gam[3] = 2 * M_PI / L * gam[2];
                gam[4] = -2 * M_PI / L * gam[3];
                gam[5] = 2 * M_PI / L * gam[4];
                gam[0] = 0.0;
                gam[2] = 0.0;
                gam[4] = 0.0;
                ax = ay = az = 0.0;
                mfacs = 0.0;
                qeval(iOrder, mom->m, tq, mom->O, mom->H, gam, hx, hy, hz, &ax, &ay, &az, &mfacs);
                ewt[i].hx = 2 * M_PI / L * hx;
                ewt[i].hy = 2 * M_PI / L * hy;
                ewt[i].hz = 2 * M_PI / L * hz;
                ewt[i].hCfac = mfacc;
                ewt[i].hSfac = mfacs;
                ++i;
            }
    *pewt = ewt;
    return i;
}

/* pkdBucketEwald (ewald.c:15-178) for the active particles of bucket ib;
 * returns its flop count. */
static double bucket_ewald(orc_tree *t, int ib, int nReps, double L, double fEwCut, int iOrder,
                           const OMom *mom, const OEwt *ewt, int nEwh, double *a, double *pot)
{
    static const int nMultiFlop[5] = {10, 10, 48, 151, 343};   /* MEVAL_FLOP, meval.h:84 */
    const ONode *bk = &t->c[ib];
    int i, j, ix, iy, iz, nEwReps = (int)ceil(fEwCut), nLoop = 0, nActive = 0;
    double alpha = 2.0 / L, alpha2 = alpha * alpha, k1 = M_PI / (alpha2 * L * L * L);
    double ka = 2.0 * alpha / sqrt(M_PI), fEwCut2 = fEwCut * fEwCut * L * L;
    if (!iOrder) return 0;
    nEwReps = nEwReps > nReps ? nEwReps : nReps;
    for (j = bk->pLower; j <= bk->pUpper; ++j) {
        const OPart *p = &t->p[j];
        double fPot, ax = 0, ay = 0, az = 0, dx, dy, dz;
        if (!p->active) continue;
        fPot = mom->m * k1;
        dx = p->r[0] - mom->x;
        dy = p->r[1] - mom->y;
        dz = p->r[2] - mom->z;
        for (ix = -nEwReps; ix <= nEwReps; ++ix) {
            int bInHolex = (ix >= -nReps && ix <= nReps);
            double dxo = dx + ix * L;
            for (iy = -nEwReps; iy <= nEwReps; ++iy) {
                int bInHolexy = (bInHolex && iy >= -nReps && iy <= nReps);
                double dyo = dy + iy * L;
                for (iz = -nEwReps; iz <= nEwReps; ++iz) {
                    int bInHole = (bInHolexy && iz >= -nReps && iz <= nReps);
                    double dzo = dz + iz * L, r2, gam[6], alphan;
                    r2 = dxo * dxo + dyo * dyo + dzo * dzo;
                    if (r2 > fEwCut2 && !bInHole) continue;
                    if (r2 < 3.0e-3 * L * L) {   /* series about the origin */
                        alphan = ka;
                        r2 *= alpha2;
                        gam[0] = alphan * (r2 / 3 - 1);
                        alphan *= 2 * alpha2;
                        gam[1] = alphan * (r2 / 5 - 1.0 / 3.0);
                        alphan *= 2 * alpha2;
                        gam[2] = alphan * (r2 / 7 - 1.0 / 5.0);
                        alphan *= 2 * alpha2;
                        gam[3] = alphan * (r2 / 9 - 1.0 / 7.0);
                        alphan *= 2 * alpha2;
                        gam[4] = alphan * (r2 / 11 - 1.0 / 9.0);
                        alphan *= 2 * alpha2;
                        gam[5] = alphan * (r2 / 13 - 1.0 / 11.0);
                    } else {
                        double r = sqrt(r2), dir = 1 / r, dir2 = dir * dir, ea;
                        ea = exp(-r2 * alpha2);
                        ea *= ka * dir2;
                        if (bInHole) gam[0] = -erf(alpha * r);
                        else gam[0] = erfc(alpha * r);
                        gam[0] *= dir;
                        gam[1] = gam[0] * dir2 + ea;
                        alphan = 2 * alpha2;
                        gam[2] = 3 * gam[1] * dir2 + alphan * ea;
                        alphan *= 2 * alpha2;
                        gam[3] = 5 * gam[2] * dir2 + alphan * ea;
                        alphan *= 2 * alpha2;
                        gam[4] = 7 * gam[3] * dir2 + alphan * ea;
                        alphan *= 2 * alpha2;
                        gam[5] = 9 * gam[4] * dir2 + alphan * ea;
                    }
                    meval(iOrder, mom, gam, dxo, dyo, dzo, &ax, &ay, &az, &fPot);
                    ++nLoop;
                }
            }
        }
        for (i = 0; i < nEwh; ++i) {
            double hdotx = ewt[i].hx * dx + ewt[i].hy * dy + ewt[i].hz * dz;
            double c = cos(hdotx), sn = sin(hdotx);
            fPot += ewt[i].hCfac * c + ewt[i].hSfac * sn;
            ax += ewt[i].hx * (ewt[i].hCfac * sn - ewt[i].hSfac * c);
            ay += ewt[i].hy * (ewt[i].hCfac * sn - ewt[i].hSfac * c);
            az += ewt[i].hz * (ewt[i].hCfac * sn - ewt[i].hSfac * c);
        }
        pot[j] += fPot;
        a[3 * j] += ax;
        a[3 * j + 1] += ay;
        a[3 * j + 2] += az;
        ++nActive;
    }
    return (double)nLoop * (104 + nMultiFlop[iOrder]) + (double)nActive * nEwh * 58;
}

/* pkdGravAll (pkd.c:2869-3060): every bucket with active particles walks
 * the replicas -nReps..nReps (periodic dimensions L[d] > 0; L = NULL for an
 * isolated system) and, when bEwald, adds the Ewald correction of the
 * root's multipoles (pkdEwaldInit + pkdBucketEwald, cubic box L[0]).
 * fweight (may be NULL): the active particles of each bucket receive the
 * bucket's interaction + Ewald flops (pkdBucketWeight, pkd.c:2851-2861,
 * 2998); the others keep the value the array holds. */
void orc_grav_all_weighted(orc_tree *t, int iOrder, int nReps, const double *L, int bEwald,
                           int iEwOrder, double fEwCut, double fEwhCut, double *a, double *pot,
                           double *dtgrav, double *stats, double *fweight)
{
    int n = t->n, ic, i, nEwh = 0, nbk = 0, k;
    double *sa = (double *)calloc((size_t)(n > 0 ? n : 1) * 3, sizeof(double));
    double *sp = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double *sd = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    int *bk = (int *)malloc((size_t)(t->nNodes > 0 ? t->nNodes : 1) * sizeof(int));
    OMom mom;
    OEwt *ewt = NULL;
    double nActive = 0, dPart = 0, dCell = 0, dSoft = 0, dFlop = 0;
    const int ewald = L && bEwald;
    if (ewald) {
        calc_root(t, &mom);
        nEwh = ewald_init(&mom, L[0], fEwhCut, iEwOrder, &ewt);
    }
    {   /* the 1/sqrt tables, before any thread reads them */
        double one = 1.0, r;
        orc_vsqrt1(1, &one, &r);
    }
    /* the buckets in the threaded walk order; each one's walk and
     * interactions touch only its own particles' results, so the buckets
     * are shared out over OpenMP threads (every counter is a sum of
     * integers, exact in any order) */
    ic = t->iRoot;
    while (ic != -1) {
        const ONode *c = &t->c[ic];
        if (c->iLower != -1) { ic = c->iLower; continue; }
        bk[nbk++] = ic;
        ic = c->iUpper;
    }
#pragma omp parallel reduction(+ : nActive, dPart, dCell, dSoft, dFlop)
    {
        OScratch s = {0};
        scratch_fit(&s, 1024);
#pragma omp for schedule(dynamic, 16)
        for (k = 0; k < nbk; ++k) {
            const int ib = bk[k];
            const ONode *c = &t->c[ib];
            OBnd b;
            double fsm;
            int na, counts[3], j;
            na = bucket_setup(t, ib, &b, &fsm);
            if (na <= 0) continue;
            for (;;) {
                walk_all_reps(t, ib, &b, fsm, L ? nReps : 0, L, &s, counts);
                if (counts[0] <= s.cap && counts[1] <= s.cap && counts[2] <= s.cap) break;
                scratch_fit(&s, counts[0] > counts[1] ? (counts[0] > counts[2] ? counts[0] : counts[2])
                                                      : (counts[1] > counts[2] ? counts[1] : counts[2]));
            }
            nActive += na;
            /* counters of pkd.c:2946-2949 (integer pair count as written) */
            dPart += na * counts[0] + na * (2 * (c->pUpper - c->pLower) - na + 1) / 2;
            dCell += na * counts[2];
            dSoft += na * counts[1];
            {
                double fI = bucket_interact(t, ib, iOrder, &s, counts[0], counts[1], counts[2],
                                            sa, sp, sd), fE = 0;
                if (ewald)
                    fE = bucket_ewald(t, ib, nReps, L[0], fEwCut, iEwOrder, &mom, ewt, nEwh,
                                      sa, sp);
                dFlop += fI + fE;
                if (fweight)
                    for (j = c->pLower; j <= c->pUpper; ++j)
                        if (t->p[j].active) fweight[t->p[j].src] = fI + fE;
            }
        }
        free(s.part); free(s.soft); free(s.newt); free(s.d2a); free(s.sq);
        free(s.po); free(s.so); free(s.no);
    }
    free(bk);
    for (i = 0; i < n; ++i) {
        int src = t->p[i].src;
        a[3 * src] = sa[3 * i];
        a[3 * src + 1] = sa[3 * i + 1];
        a[3 * src + 2] = sa[3 * i + 2];
        pot[src] = sp[i];
        dtgrav[src] = sd[i];
    }
    stats[0] = nActive; stats[1] = dPart; stats[2] = dCell; stats[3] = dSoft; stats[4] = dFlop;
    free(sa); free(sp); free(sd); free(ewt);
}

void orc_grav_all_periodic(orc_tree *t, int iOrder, int nReps, const double *L, int bEwald,
                           int iEwOrder, double fEwCut, double fEwhCut, double *a, double *pot,
                           double *dtgrav, double *stats)
{
    orc_grav_all_weighted(t, iOrder, nReps, L, bEwald, iEwOrder, fEwCut, fEwhCut, a, pot,
                          dtgrav, stats, NULL);
}

void orc_grav_all(orc_tree *t, int iOrder, double *a, double *pot,
                  double *dtgrav, double *stats)
{
    orc_grav_all_weighted(t, iOrder, 0, NULL, 0, 0, 0.0, 0.0, a, pot, dtgrav, stats, NULL);
}

/* ======================================================================
 * Direct sum with the same softening as the intra-bucket pairs.
 * ====================================================================== */
void orc_direct_sum(int n, const double *r, const double *m, const double *soft,
                    int nt, const int *idx, double *a, double *pot)
{
    int k;
#pragma omp parallel for schedule(dynamic, 64)
    for (k = 0; k < nt; ++k) {
        int i = idx ? idx[k] : k, j;
        double ax = 0, ay = 0, az = 0, ph = 0;
        for (j = 0; j < n; ++j) {
            double dx, dy, dz, d2, aa, bb;
            if (j == i) continue;
            dx = r[3 * j] - r[3 * i];
            dy = r[3 * j + 1] - r[3 * i + 1];
            dz = r[3 * j + 2] - r[3 * i + 2];
            d2 = dx * dx + dy * dy + dz * dz;
            orc_spline(d2, soft[i] + soft[j], &aa, &bb);
            ph -= aa * m[j];
            ax += dx * bb * m[j];
            ay += dy * bb * m[j];
            az += dz * bb * m[j];
        }
        a[3 * k] = ax; a[3 * k + 1] = ay; a[3 * k + 2] = az;
        pot[k] = ph;
    }
}

/* ======================================================================
 * Periodic direct sum: the Ewald-summed field of every particle and all of
 * its images in the cubic box of side L, with the uniform background that
 * makes the k = 0 mode vanish (the Green's function pkdBucketEwald,
 * ewald.c:15-178, expands for the root's moments), and the spline
 * softening (grav.h:74-97) of the images nearer than twoh = h_i + h_j.
 * Pair potential, G = 1, d = x_i - x_j:
 *   psi(d) = - sum_n erfc(alpha |d + nL|) / |d + nL|
 *            - sum_{h != 0} exp(-pi^2 h^2 / (alpha L)^2) / (pi L h^2)
 *                           cos(2 pi h.d / L)
 *            + pi / (alpha^2 L^3)
 * (Ewald 1921; the two sums' split alpha cancels), plus, for each image
 * within twoh, the spline's potential minus -1/r.  The particle's own
 * images add m_i xi, xi = lim_{d->0} (psi(d) + 1/|d|), to its potential and
 * nothing to its force.  Every pair sums its images and wave vectors
 * itself (no structure factors), so this is a plain statement of the sum.
 * alphaL = alpha L; the real-space sum runs to erfc(alpha r) < 3e-17
 * (alpha r = 6), the wave-vector sum to exp(-pi^2 h^2/(alpha L)^2) < 1e-17.
 * ====================================================================== */
typedef struct { double kx, ky, kz, c; } OEwk;

static int ewald_k_table(double L, double alphaL, OEwk **pk)
{
    const double kc2 = 39.0 * alphaL * alphaL / (M_PI * M_PI);
    const int hmax = (int)ceil(sqrt(kc2));
    int hx, hy, hz, nk = 0, cap = 256;
    OEwk *k = (OEwk *)malloc(cap * sizeof(OEwk));
    for (hx = 0; hx <= hmax; ++hx)
        for (hy = -hmax; hy <= hmax; ++hy)
            for (hz = -hmax; hz <= hmax; ++hz) {
                int h2 = hx * hx + hy * hy + hz * hz;
                /* one of each pair +h, -h: the cos terms are even in h */
                if (hx == 0 && (hy < 0 || (hy == 0 && hz <= 0))) continue;
                if (h2 > kc2) continue;
                if (nk == cap) {
                    cap *= 2;
                    k = (OEwk *)realloc(k, cap * sizeof(OEwk));
                }
                k[nk].kx = 2 * M_PI * hx / L;
                k[nk].ky = 2 * M_PI * hy / L;
                k[nk].kz = 2 * M_PI * hz / L;
                k[nk].c = 2.0 * exp(-M_PI * M_PI * h2 / (alphaL * alphaL)) / (M_PI * L * h2);
                ++nk;
            }
    *pk = k;
    return nk;
}

/* xi: the potential at a particle of unit mass from its own images and the
 * background, lim_{d->0} (psi(d) + 1/|d|). */
double orc_ewald_self(double L, double alphaL)
{
    const double alpha = alphaL / L, rc = 6.0 / alpha;
    const int nr = (int)floor(rc / L + 0.5);
    int nx, ny, nz, i, nk;
    double xi = 2 * alpha / sqrt(M_PI) + M_PI / (alpha * alpha * L * L * L);
    OEwk *k;
    for (nx = -nr; nx <= nr; ++nx)
        for (ny = -nr; ny <= nr; ++ny)
            for (nz = -nr; nz <= nr; ++nz) {
                double r;
                if (!nx && !ny && !nz) continue;
                r = L * sqrt((double)(nx * nx + ny * ny + nz * nz));
                if (r < rc) xi -= erfc(alpha * r) / r;
            }
    nk = ewald_k_table(L, alphaL, &k);
    for (i = 0; i < nk; ++i) xi -= k[i].c;
    free(k);
    return xi;
}

void orc_direct_ewald(int n, const double *r, const double *m, const double *soft, double L,
                      double alphaL, int nt, const int *idx, double *a, double *pot)
{
    const double alpha = alphaL / L, rc = 6.0 / alpha, rc2 = rc * rc;
    const double ka = 2 * alpha / sqrt(M_PI), bg = M_PI / (alpha * alpha * L * L * L);
    const int nr = (int)floor(rc / L + 0.5);
    const double xi = orc_ewald_self(L, alphaL);
    OEwk *kt;
    const int nk = ewald_k_table(L, alphaL, &kt);
    int k;
#pragma omp parallel for schedule(dynamic, 4)
    for (k = 0; k < nt; ++k) {
        const int i = idx ? idx[k] : k;
        double ax = 0, ay = 0, az = 0, ph = m[i] * xi, mt = 0;
        int j, nx, ny, nz, q;
        for (j = 0; j < n; ++j) {
            double dx, dy, dz, twoh = soft[i] + soft[j];
            mt += m[j];
            if (j == i) continue;
            /* real space: the source's images about the nearest one */
            dx = r[3 * j] - r[3 * i];
            dy = r[3 * j + 1] - r[3 * i + 1];
            dz = r[3 * j + 2] - r[3 * i + 2];
            dx -= L * rint(dx / L);
            dy -= L * rint(dy / L);
            dz -= L * rint(dz / L);
            for (nx = -nr; nx <= nr; ++nx)
                for (ny = -nr; ny <= nr; ++ny)
                    for (nz = -nr; nz <= nr; ++nz) {
                        double sx = dx + nx * L, sy = dy + ny * L, sz = dz + nz * L;
                        double s2 = sx * sx + sy * sy + sz * sz, s, e, g;
                        if (s2 >= rc2) continue;
                        if (s2 == 0.0) {   /* coincident: spline(0) minus the erf part */
                            double aa, bb;
                            orc_spline(0.0, twoh, &aa, &bb);
                            ph -= m[j] * (aa - ka);
                            continue;
                        }
                        s = sqrt(s2);
                        e = erfc(alpha * s) / s;
                        g = (e + ka * exp(-alpha * alpha * s2)) / s2;
                        if (s2 < twoh * twoh) {   /* softened: spline minus Newton */
                            double aa, bb;
                            orc_spline(s2, twoh, &aa, &bb);
                            e += aa - 1 / s;
                            g += bb - 1 / (s * s2);
                        }
                        ph -= m[j] * e;
                        ax += m[j] * g * sx;
                        ay += m[j] * g * sy;
                        az += m[j] * g * sz;
                    }
            /* wave vectors, d = x_i - x_j */
            dx = r[3 * i] - r[3 * j];
            dy = r[3 * i + 1] - r[3 * j + 1];
            dz = r[3 * i + 2] - r[3 * j + 2];
            for (q = 0; q < nk; ++q) {
                const double kd = kt[q].kx * dx + kt[q].ky * dy + kt[q].kz * dz;
                const double cs = cos(kd), sn = sin(kd), f = m[j] * kt[q].c * sn;
                ph -= m[j] * kt[q].c * cs;
                ax -= f * kt[q].kx;
                ay -= f * kt[q].ky;
                az -= f * kt[q].kz;
            }
        }
        ph += mt * bg - m[i] * bg;   /* the others' background (its own is in xi) */
        a[3 * k] = ax;
        a[3 * k + 1] = ay;
        a[3 * k + 2] = az;
        pot[k] = ph;
    }
    free(kt);
}
