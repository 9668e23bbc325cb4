// Tree gravity on the GPU: per-bucket interaction-list walk and bucket
// interaction kernels (FP64 plain and FP32 cell-relative with compensated
// accumulation), behind gg_build_binary / gg_grav_all.
//
// Reference path: pkdGravAll (pkd.c:2869-3060) -> pkdBucketWalk
// (walk.c:306-439) -> pkdBucketInteract (grav.c:23-249).
#include "gg_internal.h"

#include <mutex>

#include <algorithm>
#include <cstring>

namespace gg {

// ------------------------------------------------------------------ walk
// INTERSECTNP (walk.h:12-30): squared distance from point to box <= fBall2.
__device__ __forceinline__ bool box_within(const double *bb, double fBall2,
                                           double x, double y, double z)
{
#pragma clang fp contract(off)
    double dx = bb[0] - x, dx1 = x - bb[3];
    double dy = bb[1] - y, dy1 = y - bb[4];
    double dz = bb[2] - z, dz1 = z - bb[5];
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

// One thread per active bucket walks the threaded tree (pkdLocalWalk,
// walk.c:58-178).  fill == 0: count the PP particles, softened cells and
// Newtonian cells.  fill == 1: write them at the scanned offsets.
__global__ void k_walk(int nb, const int *__restrict__ bucket,
                       const double *__restrict__ bbox,
                       const WalkNode *__restrict__ wn, const int *__restrict__ broot, int fill,
                       int *__restrict__ cnt, const int64_t *__restrict__ off,
                       int *__restrict__ ent)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int ib = bucket[b];
    double bb[6];
    for (int k = 0; k < 6; ++k) bb[k] = bbox[8 * (size_t)b + k];
    const double fSoftMax = bbox[8 * (size_t)b + 6];
    int nP = 0, nS = 0, nN = 0;
    int64_t oP = 0, oS = 0, oN = 0;
    if (fill) {
        oP = off[3 * (size_t)b];
        oS = off[3 * (size_t)b + 1];
        oN = off[3 * (size_t)b + 2];
    }
    int ic = broot[b];
    while (ic != -1) {
        const WalkNode c = wn[ic];
        bool open = box_within(bb, c.fOpen2, c.cx, c.cy, c.cz);
        if (c.pUpper - c.pLower < 3) open = true;   // < 4 particles: open
        if (open) {
            if (c.iLower != -1) {
                ic = c.iLower;
                continue;
            }
            if (ic != ib) {
                if (fill)
                    for (int pj = c.pLower; pj <= c.pUpper; ++pj) ent[oP + nP++] = pj;
                else
                    nP += c.pUpper - c.pLower + 1;
            }
            ic = c.iNext;
        } else {
            double twoh2 = c.fSoft + fSoftMax;
            twoh2 *= twoh2;
            bool soft = (twoh2 < c.fOpen2) ? false : box_within(bb, twoh2, c.cx, c.cy, c.cz);
            if (soft) {
                if (fill) ent[oS + nS] = ic;
                ++nS;
            } else {
                if (fill) ent[oN + nN] = ic;
                ++nN;
            }
            ic = c.iNext;
        }
    }
    if (!fill) {
        cnt[3 * (size_t)b] = nP;
        cnt[3 * (size_t)b + 1] = nS;
        cnt[3 * (size_t)b + 2] = nN;
    }
}

// The same walk, one wavefront per bucket: a breadth-first descent that
// takes the decisions of pkdLocalWalk (walk.c:58-178) for up to 64 cells at
// a time.  Every cell gets the decision the threaded walk gives it (open:
// the bucket's active box within fOpen2 of the c.o.m., or fewer than 4
// particles; an opened bucket's particles go on the PP list unless it is
// this bucket; an accepted cell goes on the softened or the Newtonian list),
// so each list holds the reference's members; within a list they come in
// breadth-first rather than preorder order.  Pass fill = 0 counts, fill = 1
// writes; *err is set if a frontier outgrows WK_FC (the caller then uses
// k_walk).
constexpr int WK_FC = 1024;

// pkdLocalWalk's decision for cell c seen from bucket ib (active box bb,
// largest softening fSoftMax): open (children, or the bucket's particles
// for the PP list unless c is the bucket itself), or accept as a softened
// or a Newtonian cell (walk.c:77-171).
struct WalkDec {
    bool kids, pp, soft, newt;
    int cl, cu, pl, np;
};

// A replica (PER, bRep) sees every cell at its centre + (ox, oy, oz), its
// own bucket included among the opened ones (walk.c:74-93).
template <bool PER = false>
__device__ __forceinline__ WalkDec walk_decide(const WalkNode &n, int c, int ib, const double *bb,
                                               double fSoftMax, double ox = 0.0, double oy = 0.0,
                                               double oz = 0.0, bool bRep = false)
{
#pragma clang fp contract(off)
    WalkDec d{false, false, false, false, 0, 0, 0, 0};
    double x = n.cx, y = n.cy, z = n.cz;
    if constexpr (PER) {
        x += ox;
        y += oy;
        z += oz;
    } else {
        bRep = false;
    }
    bool open = box_within(bb, n.fOpen2, x, y, z);
    if (n.pUpper - n.pLower < 3) open = true;   // < 4 particles: open
    if (open) {
        if (n.iLower != -1) {
            d.kids = true;
            d.cl = n.iLower;
            d.cu = n.iUpper;
        } else if (c != ib || bRep) {
            d.pp = true;
            d.pl = n.pLower;
            d.np = n.pUpper - n.pLower + 1;
        }
    } else {
        double twoh2 = n.fSoft + fSoftMax;
        twoh2 *= twoh2;
        d.soft = (twoh2 < n.fOpen2) ? false : box_within(bb, twoh2, x, y, z);
        d.newt = !d.soft;
    }
    return d;
}

__global__ __launch_bounds__(256) void k_walk_bfs(int nb, const int *__restrict__ bucket,
                                                  const double *__restrict__ bbox,
                                                  const WalkNode *__restrict__ wn,
                                                  const int *__restrict__ broot,
                                                  int fill, int *__restrict__ cnt,
                                                  const int64_t *__restrict__ off,
                                                  int *__restrict__ ent, int *__restrict__ err)
{
#pragma clang fp contract(off)
    __shared__ int fr[4][2][WK_FC];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int b = blockIdx.x * 4 + w;
    if (b >= nb) return;
    const int ib = bucket[b];
    double bb[6];
    for (int k = 0; k < 6; ++k) bb[k] = bbox[8 * (size_t)b + k];
    const double fSoftMax = bbox[8 * (size_t)b + 6];
    int64_t oP = 0, oS = 0, oN = 0;
    if (fill) {
        oP = off[3 * (size_t)b];
        oS = off[3 * (size_t)b + 1];
        oN = off[3 * (size_t)b + 2];
    }
    const unsigned long long below = (1ull << lane) - 1;
    int nP = 0, nS = 0, nN = 0, nf = 1, cur = 0;
    if (lane == 0) fr[w][0][0] = broot[b];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    while (nf > 0) {
        int nn = 0;
        for (int base = 0; base < nf; base += 64) {
            const int k = base + lane;
            int c = 0;
            WalkDec dc{false, false, false, false, 0, 0, 0, 0};
            if (k < nf) {
                c = fr[w][cur][k];
                dc = walk_decide(wn[c], c, ib, bb, fSoftMax);
            }
            const bool kids = dc.kids, pp = dc.pp, soft = dc.soft, newt = dc.newt;
            const int cl = dc.cl, cu = dc.cu, pl = dc.pl, np = dc.np;
            const unsigned long long mk = __ballot(kids);
            if (kids) {
                const int p = nn + 2 * __popcll(mk & below);
                if (p + 1 < WK_FC) {
                    fr[w][cur ^ 1][p] = cl;
                    fr[w][cur ^ 1][p + 1] = cu;
                }
            }
            nn += 2 * __popcll(mk);
            int incl = np;     // PP: exclusive prefix of the particle counts
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(incl, o);
                if (lane >= o) incl += u;
            }
            if (fill && pp)
                for (int j = 0; j < np; ++j) ent[oP + nP + incl - np + j] = pl + j;
            nP += __shfl(incl, 63);
            const unsigned long long ms = __ballot(soft), mn = __ballot(newt);
            if (fill && soft) ent[oS + nS + __popcll(ms & below)] = c;
            if (fill && newt) ent[oN + nN + __popcll(mn & below)] = c;
            nS += __popcll(ms);
            nN += __popcll(mn);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (nn > WK_FC) {
            if (lane == 0) atomicOr(err, 1);
            return;
        }
        nf = nn;
        cur ^= 1;
    }
    if (!fill && lane == 0) {
        cnt[3 * (size_t)b] = nP;
        cnt[3 * (size_t)b + 1] = nS;
        cnt[3 * (size_t)b + 2] = nN;
    }
}

// Zero n 32-bit words (a kernel, not a memset, so that it can sit in a
// captured graph).
__global__ void k_zero32(int n, uint32_t *__restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Interaction counters of pkd.c:2946-2949 and the flop model of
// grav.c:246-247 (QEVAL_FLOP, qeval.h:66), summed exactly in integers.
__global__ void k_grav_stats(int nb, const int *__restrict__ bucket,
                             const double *__restrict__ bbox,
                             const WalkNode *__restrict__ wn,
                             const int *__restrict__ cnt, int iOrder,
                             unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long red[5][256];
    const int qf[5] = {10, 10, 41, 120, 277};
    unsigned long long v[5] = {0, 0, 0, 0, 0};
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb;
         b += gridDim.x * blockDim.x) {
        const WalkNode c = wn[bucket[b]];
        const long long na = (long long)bbox[8 * (size_t)b + 7];
        const long long n = c.pUpper - c.pLower + 1;
        const long long nP = cnt[3 * (size_t)b], nS = cnt[3 * (size_t)b + 1],
                        nN = cnt[3 * (size_t)b + 2];
        v[0] += na;
        v[1] += na * nP + na * (2 * (long long)(c.pUpper - c.pLower) - na + 1) / 2;
        v[2] += na * nN;
        v[3] += na * nS;
        v[4] += na * ((nP + n) * 38 + nS * 82 + nN * (35 + qf[iOrder]));
    }
    for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 5; ++k) atomicAdd(&out[k], red[k][0]);
}

// pkdBucketWeight (pkd.c:2851-2861, called at pkd.c:2998): every active
// particle of a bucket gets the bucket's flops, pkdBucketInteract's count
// (grav.c:246-247) plus, with Ewald, pkdBucketEwald's (ewald.c:175-176: per
// real-space term 104 + MEVAL_FLOP, per active particle and k-space term 58).
// One thread per active bucket.
__global__ void k_grav_weight(int nb, const int *__restrict__ bucket,
                              const double *__restrict__ bbox,
                              const WalkNode *__restrict__ wn, const int *__restrict__ cnt,
                              int iOrder, const int *__restrict__ perm,
                              const uint32_t *__restrict__ pact,
                              const int *__restrict__ ewLoops, int ewTermFlop, int nEwhFlop,
                              double *__restrict__ fWeight)
{
    const int qf[5] = {10, 10, 41, 120, 277};
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const WalkNode c = wn[bucket[b]];
    const long long na = (long long)bbox[8 * (size_t)b + 7];
    const long long n = c.pUpper - c.pLower + 1;
    const long long nP = cnt[3 * (size_t)b], nS = cnt[3 * (size_t)b + 1],
                    nN = cnt[3 * (size_t)b + 2];
    long long flop = na * ((nP + n) * 38 + nS * 82 + nN * (35 + qf[iOrder]));
    if (ewLoops) {
        long long loops = 0;
        for (int s = c.pLower; s <= c.pUpper; ++s)
            if (pact[s] & 1u) loops += ewLoops[s];
        flop += loops * ewTermFlop + na * nEwhFlop;
    }
    for (int s = c.pLower; s <= c.pUpper; ++s)
        if (pact[s] & 1u) fWeight[perm[s]] = (double)flop;
}

// --------------------------------------------------------------- interact
template <typename T> __device__ __forceinline__ T rsq(T x);
template <> __device__ __forceinline__ float rsq<float>(float x)
{
    return __builtin_amdgcn_rsqf(x);
}
// FP64: the hardware estimate refined by two Newton steps, y <- y + y (1/2 -
// x y^2 / 2), each of which roughly doubles its correct bits: within 2 ulp
// of the correctly rounded 1/sqrt(x) for positive finite x (the reference's
// own v_sqrt1 is good to 2e-10), in a third of the instructions of the
// correctly rounded square root and division.  Subnormal x is scaled by
// 2^104 first (the estimate does not take subnormals); 0 and +inf keep the
// estimate, inf and 0, as 1/sqrt gives them.
__device__ __forceinline__ double rsq_f64(double x)
{
    const bool sub = x < 0x1p-1022;
    const double xs = sub ? x * 0x1p104 : x;
    const double y0 = __builtin_amdgcn_rsq(xs);
    const double h = 0.5 * xs;
    double y = y0;
    double t = fma(-h * y, y, 0.5);
    y = fma(y, t, y);
    t = fma(-h * y, y, 0.5);
    y = fma(y, t, y);
    y = (xs > 0.0 && xs < __builtin_inf()) ? y : y0;
    return sub ? y * 0x1p52 : y;
}
template <> __device__ __forceinline__ double rsq<double>(double x) { return rsq_f64(x); }

// Running sum of per-chunk partials: Kahan-compensated in FP32, plain in FP64.
template <typename T> struct Sum {
    T s = 0, c = 0;
    __device__ __forceinline__ void add(T v)
    {
        if constexpr (sizeof(T) == 4) {
            T y = v - c;
            T t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            s += v;
        }
    }
    __device__ __forceinline__ double value() const
    {
        return (double)s - (double)c;
    }
};

template <typename T> struct Part {   // per-chunk partial sums of one target
    T ax = 0, ay = 0, az = 0, pot = 0;
};

// Spline-softened monopole (SPLINEM, grav.h:50-71) with
// twoh = h_i + h_j; Newtonian beyond twoh.
template <typename T>
__device__ __forceinline__ void spline_ab(T r2, T dir, T twoh, T &a, T &b)
{
    if (r2 < twoh * twoh) {
        const T dih = T(2) / twoh;
        T u;
        if constexpr (sizeof(T) == 4) u = r2 > T(0) ? dih * (r2 * dir) : T(0);   // r = 0: coincident
        else u = dih / dir;
        if (u < T(1)) {
            a = dih * (T(7.0 / 5.0) - T(2.0 / 3.0) * u * u + T(3.0 / 10.0) * u * u * u * u
                       - T(1.0 / 10.0) * u * u * u * u * u);
            b = dih * dih * dih * (T(4.0 / 3.0) - T(6.0 / 5.0) * u * u + T(1.0 / 2.0) * u * u * u);
        } else {
            a = T(-1.0 / 15.0) * dir + dih * (T(8.0 / 5.0) - T(4.0 / 3.0) * u * u + u * u * u
                       - T(3.0 / 10.0) * u * u * u * u + T(1.0 / 30.0) * u * u * u * u * u);
            b = T(-1.0 / 15.0) * dir * dir * dir + dih * dih * dih * (T(8.0 / 3.0) - T(3) * u
                       + T(6.0 / 5.0) * u * u - T(1.0 / 6.0) * u * u * u);
        }
    } else {
        a = dir;
        b = a * a * a;
    }
}

// Spline-softened quadrupole (SPLINEQ, grav.h:17-47).
template <typename T>
__device__ __forceinline__ void spline_abcd(T r2, T dir, T twoh, T &a, T &b, T &c, T &d)
{
    if (r2 < twoh * twoh) {
        const T dih = T(2) / twoh;
        T u;
        if constexpr (sizeof(T) == 4) u = r2 > T(0) ? dih * (r2 * dir) : T(0);   // r = 0: coincident
        else u = dih / dir;
        const T dih2 = dih * dih, dih3 = dih2 * dih, dih5 = dih3 * dih2;
        if (u < T(1)) {
            a = dih * (T(7.0 / 5.0) - T(2.0 / 3.0) * u * u + T(3.0 / 10.0) * u * u * u * u
                       - T(1.0 / 10.0) * u * u * u * u * u);
            b = dih3 * (T(4.0 / 3.0) - T(6.0 / 5.0) * u * u + T(1.0 / 2.0) * u * u * u);
            c = dih5 * (T(12.0 / 5.0) - T(3.0 / 2.0) * u);
            d = T(3.0 / 2.0) * dih5 * dih * dir;
        } else {
            a = T(-1.0 / 15.0) * dir + dih * (T(8.0 / 5.0) - T(4.0 / 3.0) * u * u + u * u * u
                       - T(3.0 / 10.0) * u * u * u * u + T(1.0 / 30.0) * u * u * u * u * u);
            b = T(-1.0 / 15.0) * dir * dir * dir + dih3 * (T(8.0 / 3.0) - T(3) * u
                       + T(6.0 / 5.0) * u * u - T(1.0 / 6.0) * u * u * u);
            c = T(-1.0 / 5.0) * dir * dir * dir * dir * dir + T(3) * dih2 * dih2 * dir
                + dih5 * (T(-12.0 / 5.0) + T(1.0 / 2.0) * u);
            d = -dir * dir * dir * dir * dir * dir * dir + T(3) * dih2 * dih2 * dir * dir * dir
                - T(1.0 / 2.0) * dih5 * dih * dir;
        }
    } else {
        a = dir;
        b = a * a * a;
        c = T(3) * b * a * a;
        d = T(5) * c * a * a;
    }
}

// LDS staging of one chunk of 64 list entries, structure of arrays.
// PP: x y z m h.  Softened cell: x y z m h Q(6, raw).  Newtonian cell:
// x y z m Q(6, traceless) [O(10) H(15)].
// A chunk of up to 64 staged interactions, field-major.  Newtonian cells
// carry the traceless quadrupole (6), and for ORDER >= 3 only the
// independent components of the traceless octupole (7: xxx xyy xxy yyy xxz
// yyz xyz) and hexadecapole (9: xxxx xyyy xxxy yyyy xxxz yyyz xxyy xxyz
// xyyz); the others follow from the vanishing traces when evaluated, which
// keeps the chunk small enough for one more workgroup per CU.
template <typename T, int ORDER> struct Stage {
    static constexpr int NF = ORDER >= 4 ? 26 : (ORDER == 3 ? 17 : 11);
    T f[NF][64];
};

struct IArgs {
    int nb, iOrder;
    int nReps;                 // periodic replicas -nReps..nReps of each L[d] > 0
    double L[3];
    const int *bucket;
    const WalkNode *wn;
    const double *cell, *momOH;
    const double *px, *py, *pz, *pm, *ph;
    const uint32_t *pact;
    const int *cnt;
    const int64_t *off;
    const int *ent;
    double *acc;   // per slot: ax ay az pot dtgrav
};

// One 64-lane workgroup per active bucket.  Lane = (target t, slice s):
// TPW = bucket size targets, S = floor(64/TPW) slices that
// split each staged chunk; slice partials are reduced at the end.
// One target of an interaction workgroup, and its running sums.
template <typename T> struct Tgt {
    T xi, yi, zi, hi, mi;   // position relative to the bucket's c.o.m., softening, mass
    int s, S;               // slice of the staged chunk this lane evaluates
    bool act;
};
template <typename T> struct Acc {
    Sum<T> SX, SY, SZ, SP;
    T dtg = 0;
};

// Staging of one interaction per lane into chunk slot `lane`, relative to
// the bucket's centre of mass (cx, cy, cz).
template <typename T, int ORDER>
__device__ __forceinline__ void stage_pp(Stage<T, ORDER> &L, int lane, int j, const IArgs &A,
                                         double cx, double cy, double cz)
{
    L.f[0][lane] = (T)(A.px[j] - cx);
    L.f[1][lane] = (T)(A.py[j] - cy);
    L.f[2][lane] = (T)(A.pz[j] - cz);
    L.f[3][lane] = (T)A.pm[j];
    L.f[4][lane] = (T)A.ph[j];
}

template <typename T, int ORDER>
__device__ __forceinline__ void stage_soft(Stage<T, ORDER> &L, int lane, int ic, const IArgs &A,
                                           double cx, double cy, double cz)
{
    const double *c = A.cell + (size_t)CR_N * ic;
    L.f[0][lane] = (T)(c[CR_X] - cx);
    L.f[1][lane] = (T)(c[CR_X + 1] - cy);
    L.f[2][lane] = (T)(c[CR_X + 2] - cz);
    L.f[3][lane] = (T)c[CR_M];
    L.f[4][lane] = (T)c[CR_H];
    for (int k = 0; k < 6; ++k) L.f[5 + k][lane] = (T)c[CR_QR + k];
}

template <typename T, int ORDER>
__device__ __forceinline__ void stage_newt(Stage<T, ORDER> &L, int lane, int ic, const IArgs &A,
                                           double cx, double cy, double cz)
{
    const double *c = A.cell + (size_t)CR_N * ic;
    L.f[0][lane] = (T)(c[CR_X] - cx);
    L.f[1][lane] = (T)(c[CR_X + 1] - cy);
    L.f[2][lane] = (T)(c[CR_X + 2] - cz);
    L.f[3][lane] = (T)c[CR_M];
    for (int k = 0; k < 6; ++k) L.f[4 + k][lane] = (T)c[CR_QT + k];
    if constexpr (ORDER >= 3) {
        const double *oh = A.momOH + (size_t)OH_N * ic;
        for (int k = 0; k < 7; ++k) L.f[10 + k][lane] = (T)oh[k];
        if constexpr (ORDER >= 4)
            for (int k = 0; k < 9; ++k) L.f[17 + k][lane] = (T)oh[10 + k];
    }
}

// Particle-particle interactions with the m staged particles
// (grav.c:81-108; the intra-bucket pairs of grav.c:211-242 skip e == self).
template <typename T, int ORDER>
__device__ __forceinline__ void eval_pp(const Stage<T, ORDER> &L, int m, const Tgt<T> &g,
                                        Acc<T> &R, int self = -1)
{
    if (!g.act) return;
    const T xi = g.xi, yi = g.yi, zi = g.zi, hi = g.hi, mi = g.mi;
    const int s = g.s, S = g.S;
    T &dtg = R.dtg;
    Part<T> P;
    for (int e = s; e < m; e += S) {
        if (e == self) continue;
        const T dx = xi - L.f[0][e], dy = yi - L.f[1][e], dz = zi - L.f[2][e];
        const T mj = L.f[3][e];
        const T r2 = dx * dx + dy * dy + dz * dz;
        T a, bb;
        spline_ab<T>(r2, rsq<T>(r2), hi + L.f[4][e], a, bb);
        const T idt2 = (mi + mj) * bb;
        dtg = idt2 > dtg ? idt2 : dtg;
        a *= mj;
        bb *= mj;
        P.pot -= a;
        P.ax -= dx * bb;
        P.ay -= dy * bb;
        P.az -= dz * bb;
    }
    R.SX.add(P.ax); R.SY.add(P.ay); R.SZ.add(P.az); R.SP.add(P.pot);
}

// Softened cells: monopole + raw quadrupole (grav.c:117-150).
template <typename T, int ORDER>
__device__ __forceinline__ void eval_soft(const Stage<T, ORDER> &L, int m, const Tgt<T> &g,
                                          Acc<T> &R)
{
    if (!g.act) return;
    const T xi = g.xi, yi = g.yi, zi = g.zi, hi = g.hi, mi = g.mi;
    const int s = g.s, S = g.S;
    T &dtg = R.dtg;
    Part<T> P;
    for (int e = s; e < m; e += S) {
        const T dx = xi - L.f[0][e], dy = yi - L.f[1][e], dz = zi - L.f[2][e];
        const T mc = L.f[3][e];
        const T xx = L.f[5][e], yy = L.f[6][e], zz = L.f[7][e];
        const T xy = L.f[8][e], xz = L.f[9][e], yz = L.f[10][e];
        const T r2 = dx * dx + dy * dy + dz * dz;
        T a, bb, c, d;
        spline_abcd<T>(r2, rsq<T>(r2), hi + L.f[4][e], a, bb, c, d);
        const T qirx = xx * dx + xy * dy + xz * dz;
        const T qiry = xy * dx + yy * dy + yz * dz;
        const T qirz = xz * dx + yz * dy + zz * dz;
        const T qir = T(0.5) * (qirx * dx + qiry * dy + qirz * dz);
        const T tr = T(0.5) * (xx + yy + zz);
        const T qir3 = bb * mc + d * qir - c * tr;
        P.pot -= a * mc + c * qir - bb * tr;
        P.ax -= qir3 * dx - c * qirx;
        P.ay -= qir3 * dy - c * qiry;
        P.az -= qir3 * dz - c * qirz;
        const T idt2 = (mi + mc) * bb;
        dtg = idt2 > dtg ? idt2 : dtg;
    }

    R.SX.add(P.ax); R.SY.add(P.ay); R.SZ.add(P.az); R.SP.add(P.pot);
}

// Newtonian cells: reduced multipoles to ORDER (grav.c:164-191, QEVAL
// qeval.h:21-64).
template <typename T, int ORDER>
__device__ __forceinline__ void eval_newt(const Stage<T, ORDER> &L, int m, const Tgt<T> &g,
                                          Acc<T> &R)
{
    if (!g.act) return;
    const T xi = g.xi, yi = g.yi, zi = g.zi, mi = g.mi;
    const int s = g.s, S = g.S;
    T &dtg = R.dtg;
    Part<T> P;
    for (int e = s; e < m; e += S) {
        const T dx = xi - L.f[0][e], dy = yi - L.f[1][e], dz = zi - L.f[2][e];
        const T mc = L.f[3][e];
        const T r2 = dx * dx + dy * dy + dz * dz;
        T g[6];
        g[0] = rsq<T>(r2);
        const T dir2 = g[0] * g[0];
        g[1] = g[0] * dir2;
        g[2] = T(3) * g[1] * dir2;
        g[3] = T(5) * g[2] * dir2;
        g[4] = T(7) * g[3] * dir2;
        g[5] = T(9) * g[4] * dir2;
        T Qx, Qy, Qz, Qm, Qta = 0;
        // Orders 3 and 4 in the independent components of the traceless
        // moments: with the traces' vanishing folded into the monomials, the
        // sums QEVAL (qeval.h:21-64) forms over all components regroup into
        // seven cubic (octupole: five quadratic) combinations shared by the
        // three components of the field.
        const T x2 = dx * dx, y2 = dy * dy, z2 = dz * dz;
        if constexpr (ORDER >= 4) {
            const T xxxx = L.f[17][e], xyyy = L.f[18][e], xxxy = L.f[19][e],
                    yyyy = L.f[20][e], xxxz = L.f[21][e], yyyz = L.f[22][e],
                    xxyy = L.f[23][e], xxyz = L.f[24][e], xyyz = L.f[25][e];
            const T xz2 = dx * z2, yz2 = dy * z2, z3 = dz * z2;
            const T cA = dx * x2 - T(3) * xz2;               // x^3 - 3 x z^2
            const T cB = T(3) * (x2 * dy - yz2);             // 3 x^2 y - 3 y z^2
            const T cC = T(3) * x2 * dz - z3;                // 3 x^2 z - z^3
            const T cD = T(3) * (dx * y2 - xz2);             // 3 x y^2 - 3 x z^2
            const T cE = T(6) * dx * dy * dz;                // 6 x y z
            const T cF = dy * y2 - T(3) * yz2;               // y^3 - 3 y z^2
            const T cG = T(3) * y2 * dz - z3;                // 3 y^2 z - z^3
            Qx = T(1.0 / 6.0) * (xxxx * cA + xxxy * cB + xxxz * cC + xxyy * cD + xxyz * cE
                                 + xyyy * cF + xyyz * cG);
            Qy = T(1.0 / 6.0) * (xxxy * cA + xxyy * cB + xxyz * cC + xyyy * cD + xyyz * cE
                                 + yyyy * cF + yyyz * cG);
            Qz = T(1.0 / 6.0) * (xxxz * cA + xxyz * cB - xxxx * cC - xxyy * (cC + cG)
                                 + xyyz * cD - (xxxy + xyyy) * cE + yyyz * cF - yyyy * cG);
            Qm = T(1.0 / 4.0) * (Qx * dx + Qy * dy + Qz * dz);
            P.pot -= g[4] * Qm;
            Qta += g[5] * Qm;
            P.ax += g[4] * Qx;
            P.ay += g[4] * Qy;
            P.az += g[4] * Qz;
        }
        if constexpr (ORDER >= 3) {
            const T xxx = L.f[10][e], xyy = L.f[11][e], xxy = L.f[12][e],
                    yyy = L.f[13][e], xxz = L.f[14][e], yyz = L.f[15][e],
                    xyz = L.f[16][e];
            const T cP = x2 - z2, cR = y2 - z2;
            const T cxy = T(2) * dx * dy, cxz = T(2) * dx * dz, cyz = T(2) * dy * dz;
            Qx = T(0.5) * (xxx * cP + xyy * cR + xxy * cxy + xxz * cxz + xyz * cyz);
            Qy = T(0.5) * (xxy * cP + yyy * cR + xyy * cxy + xyz * cxz + yyz * cyz);
            Qz = T(0.5) * (xxz * cP + yyz * cR + xyz * cxy - (xxx + xyy) * cxz
                           - (xxy + yyy) * cyz);
            Qm = T(1.0 / 3.0) * (Qx * dx + Qy * dy + Qz * dz);
            P.pot -= g[3] * Qm;
            Qta += g[4] * Qm;
            P.ax += g[3] * Qx;
            P.ay += g[3] * Qy;
            P.az += g[3] * Qz;
        }
        if constexpr (ORDER >= 2) {
            const T xx = L.f[4][e], yy = L.f[5][e], zz = L.f[6][e];
            const T xy = L.f[7][e], xz = L.f[8][e], yz = L.f[9][e];
            Qx = xz * dz + xy * dy + xx * dx;
            Qy = yz * dz + xy * dx + yy * dy;
            Qz = yz * dy + xz * dx + zz * dz;
            Qm = T(0.5) * (Qx * dx + Qy * dy + Qz * dz);
            P.pot -= g[2] * Qm;
            Qta += g[3] * Qm;
            P.ax += g[2] * Qx;
            P.ay += g[2] * Qy;
            P.az += g[2] * Qz;
        }
        P.pot -= g[0] * mc;
        Qta += g[1] * mc;
        P.ax -= dx * Qta;
        P.ay -= dy * Qta;
        P.az -= dz * Qta;
        const T idt2 = (mi + mc) * g[1];
        dtg = idt2 > dtg ? idt2 : dtg;
    }

    R.SX.add(P.ax); R.SY.add(P.ay); R.SZ.add(P.az); R.SP.add(P.pot);
}

// Target set-up of a bucket's workgroup and its intra-bucket pairs.
// Lane = (target t, slice s): TPW = bucket size targets, S = floor(64/TPW)
// slices that split each staged chunk (at most TPW - 1 lanes idle).
template <typename T, int ORDER>
__device__ __forceinline__ void bucket_targets(const IArgs &A, const WalkNode &bn,
                                               Stage<T, ORDER> &L, Tgt<T> &g, Acc<T> &R,
                                               int &it, int &TPW)
{
    const int lane = threadIdx.x;
    const int pL = bn.pLower, nT = bn.pUpper - bn.pLower + 1;
    TPW = nT;
    g.S = 64 / TPW;
    const int t = lane % TPW;
    g.s = lane / TPW;
    const bool valid = g.s < g.S;
    it = pL + (valid ? t : 0);
    g.act = valid && (A.pact[it] & 1u);
    const double cx = bn.cx, cy = bn.cy, cz = bn.cz;     // bucket centre of mass
    g.xi = (T)(A.px[it] - cx);
    g.yi = (T)(A.py[it] - cy);
    g.zi = (T)(A.pz[it] - cz);
    g.hi = (T)A.ph[it];
    g.mi = (T)A.pm[it];
    if (lane < nT) stage_pp(L, lane, pL + lane, A, cx, cy, cz);
    __syncthreads();
    eval_pp(L, nT, g, R, t);       // every pair i != j of the bucket
    __syncthreads();
}

// Sum the S slices of each target and write the slot's results.
template <typename T>
__device__ __forceinline__ void bucket_store(const IArgs &A, const Tgt<T> &g, const Acc<T> &R,
                                             int it, int TPW)
{
    const double vx = R.SX.value(), vy = R.SY.value(), vz = R.SZ.value(), vp = R.SP.value();
    const double vd = (double)R.dtg;
    double ax = vx, ay = vy, az = vz, pot = vp, dg = vd;
    const int lane = threadIdx.x;
    for (int k = 1; k < g.S; ++k) {      // slice k of this lane's target, in order
        const int src = lane + k * TPW < 64 ? lane + k * TPW : lane;
        ax += __shfl(vx, src);
        ay += __shfl(vy, src);
        az += __shfl(vz, src);
        pot += __shfl(vp, src);
        const double od = __shfl(vd, src);
        dg = od > dg ? od : dg;
    }
    if (g.act && g.s == 0) {
        double *o = A.acc + 5 * (size_t)it;
        o[0] = ax;
        o[1] = ay;
        o[2] = az;
        o[3] = pot;
        o[4] = dg;
    }
}

// The three CSR lists of bucket b, in chunks of 64 entries staged in LDS as
// cell-relative T.
template <typename T, int ORDER>
__device__ __forceinline__ void interact_lists(const IArgs &A, int b, const WalkNode &bn,
                                               Stage<T, ORDER> &L, const Tgt<T> &g, Acc<T> &R)
{
    const int lane = threadIdx.x;
    const size_t b3 = 3 * (size_t)b;
    for (int kind = 0; kind < 3; ++kind) {
        const int n = A.cnt[b3 + kind];
        const int *E = A.ent + A.off[b3 + kind];
        for (int base = 0; base < n; base += 64) {
            const int m = min(64, n - base);
            if (lane < m) {
                const int v = E[base + lane];
                if (kind == 0) stage_pp(L, lane, v, A, bn.cx, bn.cy, bn.cz);
                else if (kind == 1) stage_soft(L, lane, v, A, bn.cx, bn.cy, bn.cz);
                else stage_newt(L, lane, v, A, bn.cx, bn.cy, bn.cz);
            }
            __syncthreads();
            if (kind == 0) eval_pp(L, m, g, R);
            else if (kind == 1) eval_soft(L, m, g, R);
            else eval_newt(L, m, g, R);
            __syncthreads();
        }
    }
}

// One 64-lane workgroup per active bucket over its three lists.  A bucket
// of more than 64 particles (a zero-extent cell of coincident particles,
// which pkd.c:2489 makes a bucket whatever its size) takes its targets 64 at
// a time, one per lane, each group over its own bucket in chunks of 64
// (every pair i != j) and then over the lists.
template <typename T, int ORDER>
__global__ __launch_bounds__(64) void k_interact(IArgs A)
{
    __shared__ Stage<T, ORDER> L;
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    const WalkNode bn = A.wn[A.bucket[b]];
    const int pL = bn.pLower, nT = bn.pUpper - bn.pLower + 1;
    if (nT <= 64) {
        Tgt<T> g;
        Acc<T> R;
        int it, TPW;
        bucket_targets(A, bn, L, g, R, it, TPW);
        interact_lists(A, b, bn, L, g, R);
        bucket_store(A, g, R, it, TPW);
        return;
    }
    for (int g0 = 0; g0 < nT; g0 += 64) {
        Tgt<T> g;
        Acc<T> R;
        g.S = 1;
        g.s = 0;
        const bool valid = g0 + lane < nT;
        const int it = pL + (valid ? g0 + lane : 0);
        g.act = valid && (A.pact[it] & 1u);
        g.xi = (T)(A.px[it] - bn.cx);
        g.yi = (T)(A.py[it] - bn.cy);
        g.zi = (T)(A.pz[it] - bn.cz);
        g.hi = (T)A.ph[it];
        g.mi = (T)A.pm[it];
        for (int c0 = 0; c0 < nT; c0 += 64) {
            const int m = min(64, nT - c0);
            if (lane < m) stage_pp(L, lane, pL + c0 + lane, A, bn.cx, bn.cy, bn.cz);
            __syncthreads();
            eval_pp(L, m, g, R, g0 + lane - c0);     // skips the target itself
            __syncthreads();
        }
        interact_lists(A, b, bn, L, g, R);
        bucket_store(A, g, R, it, 64);
    }
}

// The walk and the interactions of one bucket in one 64-lane workgroup,
// with no list in memory.  The walk is breadth-first with walk_decide's
// decisions for up to 64 cells at a time; accepted cells and opened buckets
// queue in LDS and are evaluated in chunks of 64 as soon as a chunk is full,
// a PP chunk taking whole opened buckets (at most 64 particles).  Per-kind
// counts go to cnt for the interaction counters.  A frontier that outgrows
// WI_FC sets *err and leaves the bucket unevaluated (the caller then runs
// the listed walk and k_interact).
constexpr int WI_FC = 512;
constexpr int WI_Q = 128;

template <typename T, int ORDER>
__device__ __forceinline__ void flush_cells(Stage<T, ORDER> &L, int *q, int &nq, int m, int kind,
                                            const IArgs &A, const WalkNode &bn, const Tgt<T> &g,
                                            Acc<T> &R)
{
    const int lane = threadIdx.x;
    if (lane < m) {
        if (kind == 1) stage_soft(L, lane, q[lane], A, bn.cx, bn.cy, bn.cz);
        else stage_newt(L, lane, q[lane], A, bn.cx, bn.cy, bn.cz);
    }
    __syncthreads();
    if (kind == 1) eval_soft(L, m, g, R);
    else eval_newt(L, m, g, R);
    const int rest = nq - m;           // < 64 remain: move them to the front
    const int v = lane < rest ? q[m + lane] : 0;
    __syncthreads();
    if (lane < rest) q[lane] = v;
    __syncthreads();
    nq = rest;
}

// One PP chunk: the longest run of pending buckets with at most 64 particles.
template <typename T, int ORDER>
__device__ __forceinline__ void flush_pp(Stage<T, ORDER> &L, int *qpl, int *qpn, int *pidx,
                                         int &nr, int &np, const IArgs &A, const WalkNode &bn,
                                         const Tgt<T> &g, Acc<T> &R)
{
    const int lane = threadIdx.x;
    const int cntr = lane < nr ? qpn[lane] : 0;
    int incl = cntr;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
    }
    const bool take = lane < nr && incl <= 64;
    const int ntake = __popcll(__ballot(take));
    const int m = __shfl(incl, ntake - 1);
    if (take)
        for (int j = 0; j < cntr; ++j) pidx[incl - cntr + j] = qpl[lane] + j;
    __syncthreads();
    if (lane < m) stage_pp(L, lane, pidx[lane], A, bn.cx, bn.cy, bn.cz);
    __syncthreads();
    eval_pp(L, m, g, R);
    const int rest = nr - ntake;       // <= WI_Q - 1 ranges remain
    int vl[2], vn[2];
    for (int h = 0; h < 2; ++h) {
        const int k = lane + 64 * h;
        vl[h] = k < rest ? qpl[ntake + k] : 0;
        vn[h] = k < rest ? qpn[ntake + k] : 0;
    }
    __syncthreads();
    for (int h = 0; h < 2; ++h) {
        const int k = lane + 64 * h;
        if (k < rest) {
            qpl[k] = vl[h];
            qpn[k] = vn[h];
        }
    }
    __syncthreads();
    nr = rest;
    np -= m;
}

// splitmix64's finaliser: the per-entry term of the walk checksums.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// CHK: also count and sum, per list kind k, the members the walk hands to
// the evaluation and their mix64 (order-independent checksums of the lists
// this kernel evaluates, for parity checks): chk[9b + 3k ..] = count, sum,
// sum of mixes.
// The LDS of one fused walk + interaction workgroup.
template <typename T, int ORDER> struct WIShared {
    Stage<T, ORDER> L;
    int fr[2][WI_FC];
    int qs[WI_Q], qn[WI_Q], qpl[WI_Q], qpn[WI_Q];
    int pidx[64];
};

// The walk and interactions of active bucket b.  Returns false (with *err
// set, the bucket unevaluated) if a frontier outgrew WI_FC.
template <typename T, int ORDER, bool PER, bool CHK>
__device__ __forceinline__ bool walk_interact_one(const IArgs &A, int b,
                                                  const double *__restrict__ bbox,
                                                  const int *__restrict__ broot,
                                                  int *__restrict__ cnt, int *__restrict__ err,
                                                  unsigned long long *__restrict__ chk,
                                                  WIShared<T, ORDER> &S)
{
    Stage<T, ORDER> &L = S.L;
    auto &fr = S.fr;
    int *qs = S.qs, *qn = S.qn, *qpl = S.qpl, *qpn = S.qpn, *pidx = S.pidx;
    const int lane = threadIdx.x;
    const int ib = A.bucket[b];
    const WalkNode bn = A.wn[ib];
    Tgt<T> g;
    Acc<T> R;
    int it, TPW;
    bucket_targets(A, bn, L, g, R, it, TPW);
    double bb[6];
    for (int k = 0; k < 6; ++k) bb[k] = bbox[8 * (size_t)b + k];
    const double fSoftMax = bbox[8 * (size_t)b + 6];
    const unsigned long long below = (1ull << lane) - 1;
    int nS = 0, nN = 0, nR = 0, nRp = 0;          // pending (wave-uniform)
    int totP = 0, totS = 0, totN = 0;
    unsigned long long cs[6] = {0, 0, 0, 0, 0, 0};  // CHK: per lane, sum and mix sum per kind
    const int rx = PER && A.L[0] > 0 ? A.nReps : 0, ry = PER && A.L[1] > 0 ? A.nReps : 0,
              rz = PER && A.L[2] > 0 ? A.nReps : 0;
    // replicas x outermost (walk.c:325-338); each walks the whole tree with
    // its cells and particles seen at r + o, and its queues are evaluated
    // before the next one starts
    for (int ix = -rx; ix <= rx; ++ix)
    for (int iy = -ry; iy <= ry; ++iy)
    for (int iz = -rz; iz <= rz; ++iz) {
    const double ox = PER ? ix * A.L[0] : 0.0, oy = PER ? iy * A.L[1] : 0.0,
                 oz = PER ? iz * A.L[2] : 0.0;
    const bool bRep = PER && (ix || iy || iz);
    WalkNode sbo = bn;                 // staging reference: the bucket's c.o.m. less o
    if constexpr (PER) {
        sbo.cx = bn.cx - ox;
        sbo.cy = bn.cy - oy;
        sbo.cz = bn.cz - oz;
    }
    const WalkNode &sb = PER ? sbo : bn;
    int nf = 1, cur = 0;
    if (lane == 0) fr[0][0] = broot[b];
    __syncthreads();
    while (nf > 0) {
        int nn = 0;
        for (int base = 0; base < nf; base += 64) {
            const int k = base + lane;
            int c = 0;
            WalkDec dc{false, false, false, false, 0, 0, 0, 0};
            if (k < nf) {
                c = fr[cur][k];
                dc = walk_decide<PER>(A.wn[c], c, ib, bb, fSoftMax, ox, oy, oz, bRep);
            }
            const unsigned long long mk = __ballot(dc.kids);
            if (dc.kids) {
                const int p = nn + 2 * __popcll(mk & below);
                if (p + 1 < WI_FC) {
                    fr[cur ^ 1][p] = dc.cl;
                    fr[cur ^ 1][p + 1] = dc.cu;
                }
            }
            nn += 2 * __popcll(mk);
            const unsigned long long ms = __ballot(dc.soft), mn = __ballot(dc.newt);
            const unsigned long long mp = __ballot(dc.pp);
            if (dc.soft) qs[nS + __popcll(ms & below)] = c;
            if (dc.newt) qn[nN + __popcll(mn & below)] = c;
            if constexpr (CHK) {
                if (dc.soft) { cs[2] += (unsigned)c; cs[3] += mix64((unsigned)c); }
                if (dc.newt) { cs[4] += (unsigned)c; cs[5] += mix64((unsigned)c); }
                if (dc.pp)
                    for (int j = 0; j < dc.np; ++j) {
                        cs[0] += (unsigned)(dc.pl + j);
                        cs[1] += mix64((unsigned)(dc.pl + j));
                    }
            }
            if (dc.pp) {
                const int r = nR + __popcll(mp & below);
                qpl[r] = dc.pl;
                qpn[r] = dc.np;
            }
            int add = dc.np;
            for (int o = 32; o > 0; o >>= 1) add += __shfl_xor(add, o);
            nS += __popcll(ms);
            nN += __popcll(mn);
            nR += __popcll(mp);
            nRp += add;
            totS += __popcll(ms);
            totN += __popcll(mn);
            totP += add;
            __syncthreads();
            while (nS >= 64) flush_cells(L, qs, nS, 64, 1, A, sb, g, R);
            while (nN >= 64) flush_cells(L, qn, nN, 64, 2, A, sb, g, R);
            while (nRp >= 64) flush_pp(L, qpl, qpn, pidx, nR, nRp, A, sb, g, R);
        }
        __syncthreads();
        if (nn > WI_FC) {
            if (lane == 0) atomicOr(err, 1);
            return false;
        }
        nf = nn;
        cur ^= 1;
    }
    if (nS > 0) flush_cells(L, qs, nS, nS, 1, A, sb, g, R);
    if (nN > 0) flush_cells(L, qn, nN, nN, 2, A, sb, g, R);
    while (nR > 0) flush_pp(L, qpl, qpn, pidx, nR, nRp, A, sb, g, R);
    }
    if (lane == 0) {
        cnt[3 * (size_t)b] = totP;
        cnt[3 * (size_t)b + 1] = totS;
        cnt[3 * (size_t)b + 2] = totN;
    }
    if constexpr (CHK) {
        const int tot[3] = {totP, totS, totN};
        for (int k = 0; k < 6; ++k) {
            unsigned long long v = cs[k];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (lane == 0) chk[9 * (size_t)b + 3 * (k >> 1) + 1 + (k & 1)] = v;
        }
        if (lane == 0)
            for (int k = 0; k < 3; ++k) chk[9 * (size_t)b + 3 * k] = (unsigned)tot[k];
    }
    bucket_store(A, g, R, it, TPW);
    return true;
}

// One workgroup per active bucket, dealt to the XCDs in chunks of 128.
template <typename T, int ORDER, bool PER, bool CHK = false>
__global__ __launch_bounds__(64) void k_walk_interact(IArgs A, const double *__restrict__ bbox,
                                                      const int *__restrict__ broot,
                                                      int *__restrict__ cnt,
                                                      int *__restrict__ err,
                                                      unsigned long long *__restrict__ chk = nullptr)
{
    __shared__ WIShared<T, ORDER> S;
    walk_interact_one<T, ORDER, PER, CHK>(A, xcd_chunk(blockIdx.x, gridDim.x, 128), bbox, broot,
                                          cnt, err, chk, S);
}

// Persistent form: a grid that fills the chip once; each workgroup takes
// the next bucket of queue (a permutation of the active buckets, heaviest
// first when the previous evaluation's costs are known) until none is left.
// *next must be 0 at launch.  Every workgroup leaves the loop once the
// queue is exhausted.
// The persistent kernels' loop around walk_interact_one costs the compiler
// a few registers, which at its own choice would drop a wave per SIMD below
// the per-bucket kernel's occupancy; they are held to that occupancy.
template <typename T, int ORDER> constexpr int persistent_waves()
{
    return sizeof(T) == 4 ? (ORDER >= 4 ? 3 : 4) : (ORDER >= 4 ? 2 : 3);
}

template <typename T, int ORDER>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(persistent_waves<T, ORDER>())))
void k_walk_interact_queue(IArgs A,
                                                            const double *__restrict__ bbox,
                                                            const int *__restrict__ broot,
                                                            int *__restrict__ cnt,
                                                            int *__restrict__ err,
                                                            const int *__restrict__ queue,
                                                            unsigned *__restrict__ next)
{
    __shared__ WIShared<T, ORDER> S;
    for (;;) {
        unsigned slot = 0;
        if (threadIdx.x == 0) slot = atomicAdd(next, 1u);
        slot = __shfl(slot, 0);
        if (slot >= (unsigned)A.nb) break;
        walk_interact_one<T, ORDER, false, false>(A, queue[slot], bbox, broot, cnt, err, nullptr,
                                                  S);
        __syncthreads();
    }
}

// Persistent form that keeps each XCD's L2 on its own part of the tree:
// the active buckets, in tree order, are dealt to the eight XCDs in chunks
// of QC as xcd_chunk deals them, and next[x] counts the buckets XCD x's
// workgroups have taken from its share.  A workgroup (on XCD blockIdx.x & 7,
// the order in which the hardware deals workgroups) takes from its own
// XCD's share and, once that is exhausted, from the others' in turn, so the
// tail of a clumpy tree is shared out.  next[0..7] must be 0 at launch.
// Every workgroup visits each of the eight shares once and stops at the
// first slot past a share's end, so every workgroup leaves.
constexpr int QC = 128;
template <typename T, int ORDER>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(persistent_waves<T, ORDER>())))
void k_walk_interact_xcdq(IArgs A,
                                                           const double *__restrict__ bbox,
                                                           const int *__restrict__ broot,
                                                           int *__restrict__ cnt,
                                                           int *__restrict__ err,
                                                           unsigned *__restrict__ next)
{
    __shared__ WIShared<T, ORDER> S;
    const int x0 = blockIdx.x & 7;
    const unsigned nb = (unsigned)A.nb;
    for (int v = 0; v < 8; ++v) {
        const int x = (x0 + v) & 7;
        for (;;) {
            unsigned slot = 0;
            if (threadIdx.x == 0) slot = atomicAdd(&next[x], 1u);
            slot = __shfl(slot, 0);
            // slot s of XCD x: chunk (s / QC) * 8 + x, position s % QC in it;
            // later slots of x lie further on, so the first one past nb ends x
            const unsigned long long b =
                ((unsigned long long)(slot / QC) * 8 + x) * QC + slot % QC;
            if (b >= nb) break;
            walk_interact_one<T, ORDER, false, false>(A, (int)b, bbox, broot, cnt, err, nullptr,
                                                      S);
            __syncthreads();
        }
    }
}

// The persistent walk's queue, heaviest buckets first: each active bucket's
// cost by gasoline's flop model (grav.c:246-247) from the list sizes of the
// last walk, binned by its binary exponent (bin 0 the heaviest), the bins
// counted, scanned and scattered.  Buckets within a bin come in any order;
// each bucket's results do not depend on when it runs.
constexpr int QBINS = 64;
__global__ void k_cost_bin(int nb, const int *__restrict__ bucket,
                           const double *__restrict__ bbox, const WalkNode *__restrict__ wn,
                           const int *__restrict__ cnt, int *__restrict__ bin,
                           unsigned *__restrict__ hist)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const WalkNode c = wn[bucket[b]];
    const double na = bbox[8 * (size_t)b + 7];
    const double n = c.pUpper - c.pLower + 1;
    const double f = na * ((cnt[3 * (size_t)b] + n) * 38.0 + cnt[3 * (size_t)b + 1] * 82.0
                           + cnt[3 * (size_t)b + 2] * 76.0);
    const int e = f >= 1.0 ? ilogb(f) : 0;
    const int k = QBINS - 1 - min(e, QBINS - 1);
    bin[b] = k;
    atomicAdd(&hist[k], 1u);
}

// hist -> exclusive offsets, in place (one wavefront).
__global__ void k_bin_scan(unsigned *__restrict__ hist)
{
    const int l = threadIdx.x;
    const unsigned h = hist[l];
    unsigned incl = h;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned u = __shfl_up(incl, o);
        if (l >= o) incl += u;
    }
    hist[l] = incl - h;
}

__global__ void k_bin_scatter(int nb, const int *__restrict__ bin, unsigned *__restrict__ pos,
                              int *__restrict__ queue)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    queue[atomicAdd(&pos[bin[b]], 1u)] = b;
}

__global__ void k_queue_iota(int n, int *__restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// Add the per-slot results to the caller-order fields of active particles.
__global__ void k_grav_scatter(int n, const int *__restrict__ perm,
                               const uint32_t *__restrict__ pact,
                               const double *__restrict__ acc,
                               double *__restrict__ a, double *__restrict__ pot,
                               double *__restrict__ dtg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !(pact[i] & 1u)) return;
    const int c = perm[i];
    const double *o = acc + 5 * (size_t)i;
    a[3 * (size_t)c] += o[0];
    a[3 * (size_t)c + 1] += o[1];
    a[3 * (size_t)c + 2] += o[2];
    pot[c] += o[3];
    if (o[4] > dtg[c]) dtg[c] = o[4];
}

// ------------------------------------------------------------- host side
template <typename T>
static int launch_interact(int order, const IArgs &A, hipStream_t st)
{
    dim3 grid(A.nb), block(64);
    switch (order) {
    case 1: hipLaunchKernelGGL((k_interact<T, 1>), grid, block, 0, st, A); break;
    case 2: hipLaunchKernelGGL((k_interact<T, 2>), grid, block, 0, st, A); break;
    case 3: hipLaunchKernelGGL((k_interact<T, 3>), grid, block, 0, st, A); break;
    case 4: hipLaunchKernelGGL((k_interact<T, 4>), grid, block, 0, st, A); break;
    default: set_error("iOrder %d not in 1..4", order); return GG_EINVAL;
    }
    GG_HIP(hipGetLastError());
    return GG_OK;
}

}  // namespace gg

using namespace gg;

extern "C" int gg_build_binary(GGPKD pkd, int nBucket, int iOpenType, double dCrit,
                               int iOrder, int bTreeActiveOnly, int bGravity)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GG_CHECK(nBucket >= 1 && nBucket <= 64, GG_EINVAL, "nBucket %d outside 1..64", nBucket);
    GG_CHECK(iOpenType == 0 || iOpenType == 1, GG_EUNSUP,
             "iOpenType %d: only OPEN_JOSH (1) and Bmax (0) are implemented", iOpenType);
    GG_CHECK(iOrder >= 1 && iOrder <= 4, GG_EINVAL, "iOrder %d outside 1..4", iOrder);
    GG_CHECK(bGravity == 1, GG_EUNSUP, "gg_build_binary builds gravity trees only");
    GG_CHECK(bTreeActiveOnly == 0, GG_EUNSUP, "bTreeActiveOnly is not supported");
    GG_CHECK(dCrit > 0, GG_EINVAL, "dCrit must be positive");
    GG_CHECK(pkd->f[GG_R].p && pkd->f[GG_MASS].p && pkd->f[GG_SOFT].p, GG_ESTATE,
             "positions, masses and softenings must be put before gg_build_binary");
    GG_TRY(ensure_field(pkd, GG_IACTIVE));
    const int off[2] = {0, pkd->nLocal};
    return gtree_build(pkd, 1, off, nBucket, iOpenType, dCrit, iOrder);
}

extern "C" int gg_build_forest(GGPKD pkd, int nSys, const int *sysOff, int nBucket,
                               int iOpenType, double dCrit, int iOrder)
{
    GG_CHECK(pkd && sysOff, GG_EINVAL, "null argument");
    GG_CHECK(nSys >= 1, GG_EINVAL, "nSys must be >= 1");
    GG_CHECK(nBucket >= 1 && nBucket <= 64, GG_EINVAL, "nBucket %d outside 1..64", nBucket);
    GG_CHECK(iOpenType == 0 || iOpenType == 1, GG_EUNSUP,
             "iOpenType %d: only OPEN_JOSH (1) and Bmax (0) are implemented", iOpenType);
    GG_CHECK(iOrder >= 1 && iOrder <= 4, GG_EINVAL, "iOrder %d outside 1..4", iOrder);
    GG_CHECK(dCrit > 0, GG_EINVAL, "dCrit must be positive");
    const int n = pkd->nLocal;
    GG_CHECK(sysOff[0] == 0 && sysOff[nSys] == n, GG_EINVAL,
             "system offsets must run from 0 to the %d particles", n);
    for (int s = 0; s < nSys; ++s)
        GG_CHECK(sysOff[s + 1] >= sysOff[s], GG_EINVAL, "system offsets decrease at %d", s);
    GG_CHECK(pkd->f[GG_R].p && pkd->f[GG_MASS].p && pkd->f[GG_SOFT].p, GG_ESTATE,
             "positions, masses and softenings must be put before gg_build_forest");
    GG_TRY(ensure_field(pkd, GG_IACTIVE));
    return gtree_build(pkd, nSys, sysOff, nBucket, iOpenType, dCrit, iOrder);
}

static IArgs make_iargs(GGPKD pkd, int iOrder)
{
    GravDev &G = pkd->grav;
    IArgs A;
    A.nb = G.nb;
    A.iOrder = iOrder;
    A.bucket = G.bucket.as<int>();
    A.wn = G.wnode.as<WalkNode>();
    A.cell = G.cell.as<double>();
    A.momOH = G.momOH.as<double>();
    A.px = G.px.as<double>();
    A.py = G.py.as<double>();
    A.pz = G.pz.as<double>();
    A.pm = G.pm.as<double>();
    A.ph = G.ph.as<double>();
    A.pact = G.pact.as<uint32_t>();
    A.cnt = G.cnt.as<int>();
    A.off = G.off.as<int64_t>();
    A.ent = G.ent.as<int>();
    A.acc = G.acc.as<double>();
    A.nReps = 0;
    A.L[0] = A.L[1] = A.L[2] = 0.0;
    return A;
}

// Workgroups of the persistent grid: as many as fit on the chip at once.
template <typename T, int O>
static int persistent_grid(GravDev &G)
{
    const int key = (int)sizeof(T) * 10 + O;
    if (G.qGrid > 0 && G.qGridKey == key) return GG_OK;
    G.qGridKey = key;
    int dev = 0, ncu = 0, per = 0;
    GG_HIP(hipGetDevice(&dev));
    GG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    GG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_walk_interact_queue<T, O>, 64, 0));
    G.qGrid = std::max(1, ncu * std::max(per, 1));
    return GG_OK;
}

// The persistent queue for the current tree: heaviest buckets first when a
// fused walk of this tree has left its list sizes in G.cnt, else the
// buckets in order.  Kernels only (it may sit in a captured graph).
static int queue_order(GravDev &G, hipStream_t st, bool fromCosts)
{
    const int nb = G.nb;
    GG_TRY(G.qorder.reserve(sizeof(int) * std::max(nb, 1)));
    if (!fromCosts) {
        hipLaunchKernelGGL(k_queue_iota, dim3((nb + 255) / 256), dim3(256), 0, st, nb,
                           G.qorder.as<int>());
        G.qTree = -1;
    } else {
        GG_TRY(G.qbin.reserve(sizeof(int) * std::max(nb, 1)));
        GG_TRY(G.qhist.reserve(sizeof(unsigned) * QBINS));
        hipLaunchKernelGGL(k_zero32, dim3(1), dim3(64), 0, st, QBINS, G.qhist.as<uint32_t>());
        hipLaunchKernelGGL(k_cost_bin, dim3((nb + 255) / 256), dim3(256), 0, st, nb,
                           G.bucket.as<int>(), G.bbox.as<double>(), G.wnode.as<WalkNode>(),
                           G.cnt.as<int>(), G.qbin.as<int>(), G.qhist.as<unsigned>());
        hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(64), 0, st, G.qhist.as<unsigned>());
        hipLaunchKernelGGL(k_bin_scatter, dim3((nb + 255) / 256), dim3(256), 0, st, nb,
                           G.qbin.as<int>(), G.qhist.as<unsigned>(), G.qorder.as<int>());
        G.qTree = G.treeId;
    }
    GG_HIP(hipGetLastError());
    return GG_OK;
}

template <typename T>
static int launch_fused(int order, const IArgs &A, GravDev &G, hipStream_t st)
{
    dim3 grid(A.nb), block(64);
    const double *bb = G.bbox.as<double>();
    int *cnt = G.cnt.as<int>(), *err = G.werr.as<int>();
    const int *br = G.broot.as<int>();
    if (G.sched == 2 && A.nReps == 0) {   // persistent, a share of the tree per XCD
        GG_TRY(G.qnext.reserve(sizeof(unsigned) * 8));
        hipLaunchKernelGGL(k_zero32, dim3(1), dim3(64), 0, st, 8, G.qnext.as<uint32_t>());
        unsigned *nx = G.qnext.as<unsigned>();
#define GG_XQ(O)                                                                                 \
    GG_TRY((persistent_grid<T, O>(G)));                                                           \
    hipLaunchKernelGGL((k_walk_interact_xcdq<T, O>), dim3(G.qGrid), block, 0, st, A, bb, br,    \
                       cnt, err, nx)
        switch (order) {
        case 1: GG_XQ(1); break;
        case 2: GG_XQ(2); break;
        case 3: GG_XQ(3); break;
        case 4: GG_XQ(4); break;
        default: set_error("iOrder %d not in 1..4", order); return GG_EINVAL;
        }
#undef GG_XQ
        GG_HIP(hipGetLastError());
        return GG_OK;
    }
    if (G.sched == 1 && A.nReps == 0) {   // persistent: workgroups take buckets from the queue
        if (G.qTree != G.treeId) GG_TRY(queue_order(G, st, false));
        GG_TRY(G.qnext.reserve(sizeof(unsigned)));
        hipLaunchKernelGGL(k_zero32, dim3(1), dim3(64), 0, st, 1, G.qnext.as<uint32_t>());
        const int *q = G.qorder.as<int>();
        unsigned *nx = G.qnext.as<unsigned>();
#define GG_QUEUE(O)                                                                              \
    GG_TRY((persistent_grid<T, O>(G)));                                                           \
    hipLaunchKernelGGL((k_walk_interact_queue<T, O>), dim3(G.qGrid), block, 0, st, A, bb, br,   \
                       cnt, err, q, nx)
        switch (order) {
        case 1: GG_QUEUE(1); break;
        case 2: GG_QUEUE(2); break;
        case 3: GG_QUEUE(3); break;
        case 4: GG_QUEUE(4); break;
        default: set_error("iOrder %d not in 1..4", order); return GG_EINVAL;
        }
#undef GG_QUEUE
        GG_HIP(hipGetLastError());
        return GG_OK;
    }
    // replicas only when there are some: the isolated walk keeps its code
#define GG_FUSED(O)                                                                              \
    if (A.nReps > 0)                                                                             \
        hipLaunchKernelGGL((k_walk_interact<T, O, true>), grid, block, 0, st, A, bb, br, cnt, err); \
    else                                                                                         \
        hipLaunchKernelGGL((k_walk_interact<T, O, false>), grid, block, 0, st, A, bb, br, cnt, err)
    switch (order) {
    case 1: GG_FUSED(1); break;
    case 2: GG_FUSED(2); break;
    case 3: GG_FUSED(3); break;
    case 4: GG_FUSED(4); break;
    default: set_error("iOrder %d not in 1..4", order); return GG_EINVAL;
    }
#undef GG_FUSED
    GG_HIP(hipGetLastError());
    return GG_OK;
}

// The listed walk (count, scan, fill) into G.cnt / G.off / G.ent: one
// wavefront per bucket, or one thread per bucket if a breadth-first
// frontier outgrows its buffer.
static int walk_lists(GGPKD pkd)
{
    GravDev &G = pkd->grav;
    hipStream_t st = pkd->stream;
    const int nb = G.nb;
    GG_TRY(G.werr.reserve(sizeof(int)));
    GG_HIP(hipMemsetAsync(G.werr.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_walk_bfs, dim3((nb + 3) / 4), dim3(256), 0, st, nb, G.bucket.as<int>(),
                       G.bbox.as<double>(), G.wnode.as<WalkNode>(), G.broot.as<int>(), 0, G.cnt.as<int>(),
                       (const int64_t *)nullptr, (int *)nullptr, G.werr.as<int>());
    GG_HIP(hipGetLastError());
    int werr = 0;
    GG_HIP(hipMemcpyAsync(&werr, G.werr.p, sizeof(int), hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    if (werr)
        hipLaunchKernelGGL(k_walk, dim3((nb + 127) / 128), dim3(128), 0, st, nb,
                           G.bucket.as<int>(), G.bbox.as<double>(), G.wnode.as<WalkNode>(), G.broot.as<int>(), 0, G.cnt.as<int>(), (const int64_t *)nullptr, (int *)nullptr);
    int64_t total = 0;
    GG_TRY(exclusive_scan_i64(st, G.cnt.as<int>(), G.off.as<int64_t>(), 3 * nb, G.scan, &total));
    G.nEnt = total;
    GG_TRY(G.ent.reserve(sizeof(int) * std::max<int64_t>(total, 1)));
    if (werr)
        hipLaunchKernelGGL(k_walk, dim3((nb + 127) / 128), dim3(128), 0, st, nb,
                           G.bucket.as<int>(), G.bbox.as<double>(), G.wnode.as<WalkNode>(), G.broot.as<int>(), 1, G.cnt.as<int>(), G.off.as<int64_t>(), G.ent.as<int>());
    else
        hipLaunchKernelGGL(k_walk_bfs, dim3((nb + 3) / 4), dim3(256), 0, st, nb,
                           G.bucket.as<int>(), G.bbox.as<double>(), G.wnode.as<WalkNode>(), G.broot.as<int>(), 1, G.cnt.as<int>(), G.off.as<int64_t>(), G.ent.as<int>(),
                           G.werr.as<int>());
    GG_HIP(hipGetLastError());
    G.serialWalk = werr;
    G.listsCurrent = 1;
    return GG_OK;
}

extern "C" int gg_grav_all(GGPKD pkd, int nReps, int bPeriodic, int iOrder, int bEwald,
                           int iEwOrder, double fEwCut, double fEwhCut, int bComove,
                           double dRhoFac, int bDoSun, double dSunSoft, double *aSun,
                           int *nActive, double *pdPartSum, double *pdCellSum,
                           double *pdSoftSum, double *pdFlop)
{
    (void)dRhoFac; (void)dSunSoft; (void)aSun;
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GravDev &G = pkd->grav;
    GG_CHECK(G.built, GG_ESTATE, "gg_grav_all before gg_build_binary");
    GG_CHECK(G.n == pkd->nLocal, GG_ESTATE, "particle count changed since the tree was built");
    const double *L = pkd->fPeriod;
    if (bPeriodic) {
        GG_CHECK(L[0] > 0 && L[1] > 0 && L[2] > 0, GG_ESTATE,
                 "periodic gravity needs the box (gg_set_period)");
        GG_CHECK(nReps >= 0 && nReps <= 8, GG_EINVAL, "nReps %d outside 0..8", nReps);
        GG_CHECK(G.nSys == 1, GG_EUNSUP, "periodic gravity needs a single tree, not a forest");
        GG_CHECK(!bEwald || (L[0] == L[1] && L[1] == L[2]), GG_EUNSUP,
                 "Ewald summation needs a cubic box (ewald.c uses fPeriod[0])");
        GG_CHECK(!bEwald || (iEwOrder >= 0 && iEwOrder <= 4), GG_EINVAL,
                 "iEwOrder %d outside 0..4", iEwOrder);
        GG_CHECK(!bEwald || (fEwCut > 0 && fEwCut <= 8 && fEwhCut > 0 && fEwhCut <= 8), GG_EINVAL,
                 "Ewald cuts outside (0, 8]");
    }
    const bool per = bPeriodic != 0, ewald = per && bEwald;
    GG_CHECK(!bComove && !bDoSun, GG_EUNSUP, "comoving and solar-system terms are not supported");
    GG_CHECK(iOrder >= 1 && iOrder <= G.iOrder, GG_EINVAL,
             "iOrder %d exceeds the tree's moments (%d)", iOrder, G.iOrder);
    hipStream_t st = pkd->stream;
    GG_TRY(ensure_field(pkd, GG_A));
    GG_TRY(ensure_field(pkd, GG_POT));
    GG_TRY(ensure_field(pkd, GG_DTGRAV));
    GG_TRY(ensure_field(pkd, GG_FWEIGHT));
    const int nb = G.nb;
    unsigned long long hs[5] = {0, 0, 0, 0, 0};
    if (nb > 0) {
        timer_mark(pkd, 0);
        // walk and interactions fused, one workgroup per bucket; the listed
        // walk and k_interact if a breadth-first frontier outgrew its buffer
        IArgs A = make_iargs(pkd, iOrder);
        if (per) {
            A.nReps = nReps;
            for (int d = 0; d < 3; ++d) A.L[d] = L[d];
        }
        GG_TRY(G.werr.reserve(sizeof(int)));
        G.listsCurrent = 0;
        // buckets of more than 64 coincident particles do not fit the fused
        // kernel's wavefront: the listed walk and k_interact take them
        const bool big = G.maxBucket > 64;
        GG_CHECK(!(big && per), GG_EUNSUP,
                 "periodic gravity with a bucket of %d coincident particles (> 64)",
                 G.maxBucket);
        int werr = 1;
        if (!big) {
            GG_HIP(hipMemsetAsync(G.werr.p, 0, sizeof(int), st));
            if (pkd->prec == GG_PREC_FP32) GG_TRY(launch_fused<float>(iOrder, A, G, st));
            else GG_TRY(launch_fused<double>(iOrder, A, G, st));
            GG_HIP(hipMemcpyAsync(&werr, G.werr.p, sizeof(int), hipMemcpyDeviceToHost, st));
            GG_HIP(hipStreamSynchronize(st));
        }
        G.fusedWalk = !werr;
        GG_CHECK(!(werr && per), GG_EUNSUP,
                 "a breadth-first frontier of the periodic walk outgrew its buffer");
        if (werr) {
            GG_TRY(walk_lists(pkd));
            const IArgs B = make_iargs(pkd, iOrder);
            if (pkd->prec == GG_PREC_FP32) GG_TRY(launch_interact<float>(iOrder, B, st));
            else GG_TRY(launch_interact<double>(iOrder, B, st));
        }
        double ewFlop = 0;
        if (ewald) GG_TRY(ewald_apply(pkd, nReps, iEwOrder, L[0], fEwCut, fEwhCut, &ewFlop));
        timer_mark(pkd, 1);
        GG_TRY(G.stats.reserve(sizeof(unsigned long long) * 5));
        GG_HIP(hipMemsetAsync(G.stats.p, 0, sizeof(unsigned long long) * 5, st));
        hipLaunchKernelGGL(k_grav_stats, dim3(std::min((nb + 255) / 256, 1024)), dim3(256), 0, st,
                           nb, G.bucket.as<int>(), G.bbox.as<double>(), G.wnode.as<WalkNode>(),
                           G.cnt.as<int>(), iOrder, G.stats.as<unsigned long long>());
        GG_HIP(hipGetLastError());
        if (G.sched == 1 && G.fusedWalk && !per)   // the next call's queue: heaviest first
            GG_TRY(queue_order(G, st, true));
        {
            static const int meval[5] = {10, 10, 48, 151, 343};   // MEVAL_FLOP, meval.h:84
            const bool ew = ewald && iEwOrder;
            hipLaunchKernelGGL(k_grav_weight, dim3((nb + 255) / 256), dim3(256), 0, st, nb,
                               G.bucket.as<int>(), G.bbox.as<double>(), G.wnode.as<WalkNode>(),
                               G.cnt.as<int>(), iOrder, G.perm.as<int>(), G.pact.as<uint32_t>(),
                               ew ? G.ewl.as<int>() : (const int *)nullptr,
                               ew ? 104 + meval[iEwOrder > 4 ? 4 : iEwOrder] : 0,
                               ew ? G.nEwh * 58 : 0, pkd->f[GG_FWEIGHT].as<double>());
            GG_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_grav_scatter, dim3((G.n + 255) / 256), dim3(256), 0, st, G.n,
                           G.perm.as<int>(), G.pact.as<uint32_t>(), G.acc.as<double>(),
                           pkd->f[GG_A].as<double>(), pkd->f[GG_POT].as<double>(),
                           pkd->f[GG_DTGRAV].as<double>());
        GG_HIP(hipGetLastError());
        timer_mark(pkd, 2);
        GG_HIP(hipMemcpyAsync(hs, G.stats.p, sizeof hs, hipMemcpyDeviceToHost, st));
        GG_HIP(hipStreamSynchronize(st));
        const int phase[2] = {1, 8};
        timer_collect_now(pkd, 0, 2, phase);
        pkd->tm.ms[0] = 0;     // the walk is inside the interaction kernel
        if (ewald && iEwOrder)   // pkdBucketEwald's count (ewald.c:175-176)
            ewFlop += (double)hs[0] * G.nEwh * 58;
        hs[4] += (unsigned long long)ewFlop;
    }
    if (nActive) *nActive = (int)hs[0];
    if (pdPartSum) *pdPartSum = (double)hs[1];
    if (pdCellSum) *pdCellSum = (double)hs[2];
    if (pdSoftSum) *pdSoftSum = (double)hs[3];
    if (pdFlop) *pdFlop = (double)hs[4];
    return GG_OK;
}

extern "C" int gg_grav_walk_checksums(GGPKD pkd, int iOrder, uint64_t *sums)
{
    GG_CHECK(pkd && sums, GG_EINVAL, "null argument");
    GravDev &G = pkd->grav;
    GG_CHECK(G.built, GG_ESTATE, "no gravity tree");
    GG_CHECK(G.n == pkd->nLocal, GG_ESTATE, "particle count changed since the tree was built");
    GG_CHECK(iOrder >= 1 && iOrder <= G.iOrder, GG_EINVAL,
             "iOrder %d exceeds the tree's moments (%d)", iOrder, G.iOrder);
    GG_CHECK(G.maxBucket <= 64, GG_EUNSUP, "a bucket of %d coincident particles (> 64) is "
             "walked by the listed walk only", G.maxBucket);
    const int nb = G.nb;
    if (nb == 0) return GG_OK;
    hipStream_t st = pkd->stream;
    DevBuf out;
    GG_TRY(out.reserve(sizeof(unsigned long long) * 9 * (size_t)nb));
    GG_TRY(G.werr.reserve(sizeof(int)));
    GG_HIP(hipMemsetAsync(G.werr.p, 0, sizeof(int), st));
    const IArgs A = make_iargs(pkd, iOrder);
    const dim3 grid(nb), block(64);
    const double *bb = G.bbox.as<double>();
    const int *br = G.broot.as<int>();
    int *cnt = G.cnt.as<int>(), *err = G.werr.as<int>();
    unsigned long long *o = out.as<unsigned long long>();
#define GG_CHKW(O)                                                                              \
    if (pkd->prec == GG_PREC_FP32)                                                              \
        hipLaunchKernelGGL((k_walk_interact<float, O, false, true>), grid, block, 0, st, A, bb, \
                           br, cnt, err, o);                                                    \
    else                                                                                        \
        hipLaunchKernelGGL((k_walk_interact<double, O, false, true>), grid, block, 0, st, A,    \
                           bb, br, cnt, err, o)
    switch (iOrder) {
    case 1: GG_CHKW(1); break;
    case 2: GG_CHKW(2); break;
    case 3: GG_CHKW(3); break;
    default: GG_CHKW(4); break;
    }
#undef GG_CHKW
    GG_HIP(hipGetLastError());
    int werr = 0;
    GG_HIP(hipMemcpyAsync(&werr, G.werr.p, sizeof(int), hipMemcpyDeviceToHost, st));
    GG_HIP(hipMemcpyAsync(sums, o, sizeof(unsigned long long) * 9 * (size_t)nb,
                          hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    out.release();
    G.listsCurrent = 0;
    GG_CHECK(!werr, GG_EUNSUP, "a frontier of the fused walk outgrew its buffer: this tree is "
             "walked by the listed walk");
    return GG_OK;
}

extern "C" int gg_grav_lists(GGPKD pkd, int *nBuckets, int64_t *nEnt, int *bucket_node,
                             int64_t *off, int *cnt, int *ent)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GravDev &G = pkd->grav;
    GG_CHECK(G.built, GG_ESTATE, "no gravity tree");
    if (G.nb && !G.listsCurrent) GG_TRY(walk_lists(pkd));
    if (nBuckets) *nBuckets = G.nb;
    if (nEnt) *nEnt = G.nEnt;
    if (!ent) return GG_OK;
    hipStream_t st = pkd->stream;
    if (G.nb) {
        GG_HIP(hipMemcpyAsync(bucket_node, G.bucket.p, sizeof(int) * G.nb, hipMemcpyDeviceToHost, st));
        GG_HIP(hipMemcpyAsync(off, G.off.p, sizeof(int64_t) * 3 * G.nb, hipMemcpyDeviceToHost, st));
        GG_HIP(hipMemcpyAsync(cnt, G.cnt.p, sizeof(int) * 3 * G.nb, hipMemcpyDeviceToHost, st));
    }
    if (G.nEnt) GG_HIP(hipMemcpyAsync(ent, G.ent.p, sizeof(int) * G.nEnt, hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    return GG_OK;
}

namespace gg {
// FP64 direct sum for sampled targets (the accuracy check of the tree
// forces): every other particle of the store with the spline softening of
// the particle-particle interaction, twoh = h_i + h_j (grav.c:81-108), no
// tree.  One thread per target and DSPLIT workgroups per 256 targets, each
// over its share of the sources, staged through LDS in tiles of 256, with
// Kahan-compensated sums; the shares are added in order afterwards.
constexpr int DSPLIT = 64;

__global__ __launch_bounds__(256) void k_direct(int n, int nt, const int *__restrict__ tgt,
                                                const double *__restrict__ r,
                                                const double *__restrict__ m,
                                                const double *__restrict__ h,
                                                double *__restrict__ part)
{
    __shared__ double sx[256], sy[256], sz[256], sm[256], sh[256];
    const int k = blockIdx.x * 256 + threadIdx.x;
    int i = k < nt ? tgt[k] : -1;
    if (i >= n) i = -1;                  // a target outside the store computes nothing
    const double xi = i >= 0 ? r[3 * (size_t)i] : 0.0, yi = i >= 0 ? r[3 * (size_t)i + 1] : 0.0;
    const double zi = i >= 0 ? r[3 * (size_t)i + 2] : 0.0, hi = i >= 0 ? h[i] : 0.0;
    double s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    const int per = (n + DSPLIT - 1) / DSPLIT;
    const int j0 = blockIdx.y * per, j1 = min(n, j0 + per);
    for (int t0 = j0; t0 < j1; t0 += 256) {
        const int j = t0 + threadIdx.x;
        if (j < j1) {
            sx[threadIdx.x] = r[3 * (size_t)j];
            sy[threadIdx.x] = r[3 * (size_t)j + 1];
            sz[threadIdx.x] = r[3 * (size_t)j + 2];
            sm[threadIdx.x] = m[j];
            sh[threadIdx.x] = h[j];
        }
        __syncthreads();
        const int nj = min(256, j1 - t0);
        if (i >= 0)
            for (int e = 0; e < nj; ++e) {
                if (t0 + e == i) continue;
                const double dx = sx[e] - xi, dy = sy[e] - yi, dz = sz[e] - zi;
                const double r2 = dx * dx + dy * dy + dz * dz;
                double a, b;
                spline_ab<double>(r2, 1.0 / sqrt(r2), hi + sh[e], a, b);
                const double v[4] = {dx * b * sm[e], dy * b * sm[e], dz * b * sm[e], -a * sm[e]};
                for (int q = 0; q < 4; ++q) {   // Kahan
                    const double y = v[q] - c[q], t = s[q] + y;
                    c[q] = (t - s[q]) - y;
                    s[q] = t;
                }
            }
        __syncthreads();
    }
    if (i >= 0)
        for (int q = 0; q < 4; ++q) part[((size_t)blockIdx.y * nt + k) * 4 + q] = s[q] - c[q];
}

__global__ void k_direct_sum(int nt, const double *__restrict__ part, double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nt) return;
    for (int q = 0; q < 4; ++q) {
        double s = 0.0;
        for (int p = 0; p < DSPLIT; ++p) s += part[((size_t)p * nt + k) * 4 + q];
        out[4 * (size_t)k + q] = s;
    }
}
// The shares' sums of gg_direct_grav_noalloc, added in order into the
// caller's device arrays; a target outside the store gives NaN.
__global__ void k_direct_out(int nt, int n, const int *__restrict__ tgt,
                             const double *__restrict__ part, double *__restrict__ a,
                             double *__restrict__ pot)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nt) return;
    const int i = tgt[k];
    double s[4];
    for (int q = 0; q < 4; ++q) {
        s[q] = 0.0;
        for (int p = 0; p < DSPLIT; ++p) s[q] += part[((size_t)p * nt + k) * 4 + q];
        if (i < 0 || i >= n) s[q] = __builtin_nan("");
    }
    for (int d = 0; d < 3; ++d) a[3 * (size_t)k + d] = s[d];
    pot[k] = s[3];
}

// Scratch of gg_direct_grav_noalloc: the DSPLIT shares' partial sums (4 per
// target) of at most GG_DIRECT_NOALLOC_MAX targets, 8 MiB, a fixed device
// array that every GPU the library runs on holds until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host the name
// is only a stand-in): it uses the device address hipGetSymbolAddress gives,
// looked up once per device by direct_scratch_setup.
constexpr int NOALLOC_MAX = GG_DIRECT_NOALLOC_MAX;
constexpr int NOALLOC_DEV = 64;
static __device__ __align__(256) double scratch[(size_t)DSPLIT * 4 * NOALLOC_MAX];
static std::mutex scratch_mu;
static double *scratch_ptr[NOALLOC_DEV];

int direct_scratch_setup(int device)
{
    std::lock_guard<std::mutex> lock(scratch_mu);
    GG_CHECK(device >= 0 && device < NOALLOC_DEV, GG_EINVAL, "device %d outside 0..%d", device,
             NOALLOC_DEV - 1);
    if (scratch_ptr[device]) return GG_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    GG_CHECK(e == hipSuccess && p, GG_EHIP, "hipGetSymbolAddress of the direct-sum scratch: %s",
             hipGetErrorString(e));
    scratch_ptr[device] = static_cast<double *>(p);
    return GG_OK;
}

static double *direct_scratch(int device)
{
    std::lock_guard<std::mutex> lock(scratch_mu);
    return device >= 0 && device < NOALLOC_DEV ? scratch_ptr[device] : nullptr;
}
}  // namespace gg

extern "C" int gg_direct_grav_noalloc(GGPKD pkd, int nTarget, const int *target, double *a,
                                      double *pot)
{
    GG_CHECK(pkd && (nTarget == 0 || (target && a && pot)), GG_EINVAL, "null argument");
    GG_CHECK(nTarget >= 0 && nTarget <= NOALLOC_MAX, GG_EINVAL,
             "nTarget %d outside 0..%d (GG_DIRECT_NOALLOC_MAX)", nTarget, NOALLOC_MAX);
    GG_CHECK(pkd->f[GG_R].p && pkd->f[GG_MASS].p && pkd->f[GG_SOFT].p, GG_ESTATE,
             "positions, masses and softenings must be put before gg_direct_grav_noalloc");
    double *part = direct_scratch(pkd->device);
    GG_CHECK(part, GG_ESTATE, "no direct-sum scratch on device %d", pkd->device);
    if (nTarget == 0) return GG_OK;
    hipStream_t st = pkd->stream;
    const int n = pkd->nLocal;
    hipLaunchKernelGGL(k_direct, dim3((nTarget + 255) / 256, DSPLIT), dim3(256), 0, st, n,
                       nTarget, target, pkd->f[GG_R].as<double>(), pkd->f[GG_MASS].as<double>(),
                       pkd->f[GG_SOFT].as<double>(), part);
    hipLaunchKernelGGL(k_direct_out, dim3((nTarget + 255) / 256), dim3(256), 0, st, nTarget, n,
                       target, part, a, pot);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

extern "C" int gg_direct_grav(GGPKD pkd, int nTarget, const int *target, double *a, double *pot)
{
    GG_CHECK(pkd && (nTarget == 0 || (target && a && pot)), GG_EINVAL, "null argument");
    GG_CHECK(nTarget >= 0, GG_EINVAL, "nTarget must be >= 0");
    GG_CHECK(pkd->f[GG_R].p && pkd->f[GG_MASS].p && pkd->f[GG_SOFT].p, GG_ESTATE,
             "positions, masses and softenings must be put before gg_direct_grav");
    const int n = pkd->nLocal;
    for (int k = 0; k < nTarget; ++k)
        GG_CHECK(target[k] >= 0 && target[k] < n, GG_EINVAL, "target %d outside 0..%d",
                 target[k], n - 1);
    if (nTarget == 0) return GG_OK;
    hipStream_t st = pkd->stream;
    DevBuf t, o;
    GG_TRY(t.reserve(sizeof(int) * nTarget));
    DevBuf pt;
    if (o.reserve(sizeof(double) * 4 * nTarget) != GG_OK ||
        pt.reserve(sizeof(double) * 4 * (size_t)DSPLIT * nTarget) != GG_OK) {
        t.release();
        o.release();
        return GG_EHIP;
    }
    int rc = GG_OK;
    std::vector<double> res(4 * (size_t)nTarget);
    if (hipMemcpyAsync(t.p, target, sizeof(int) * nTarget, hipMemcpyHostToDevice, st) != hipSuccess)
        rc = GG_EHIP;
    if (rc == GG_OK) {
        hipLaunchKernelGGL(k_direct, dim3((nTarget + 255) / 256, DSPLIT), dim3(256), 0, st, n,
                           nTarget, t.as<int>(), pkd->f[GG_R].as<double>(),
                           pkd->f[GG_MASS].as<double>(), pkd->f[GG_SOFT].as<double>(),
                           pt.as<double>());
        hipLaunchKernelGGL(k_direct_sum, dim3((nTarget + 255) / 256), dim3(256), 0, st, nTarget,
                           pt.as<double>(), o.as<double>());
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(res.data(), o.p, sizeof(double) * 4 * nTarget, hipMemcpyDeviceToHost,
                           st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = GG_EHIP;
    }
    t.release();
    o.release();
    pt.release();
    GG_CHECK(rc == GG_OK, GG_EHIP, "gg_direct_grav: HIP error");
    for (int k = 0; k < nTarget; ++k) {
        for (int d = 0; d < 3; ++d) a[3 * (size_t)k + d] = res[4 * (size_t)k + d];
        pot[k] = res[4 * (size_t)k + 3];
    }
    return GG_OK;
}

namespace gg {
// FP64 periodic direct sum for sampled targets, the reference force of the
// periodic tree + Ewald forces: the Ewald-summed field (Ewald 1921) of every
// other particle and of all images of every particle in the cube of side L,
// with the uniform background that removes the k = 0 mode (the Green's
// function pkdBucketEwald, ewald.c:15-178, expands for the root's moments),
// and the spline softening (twoh = h_i + h_j, grav.c:81-108) of images
// nearer than twoh.  Split alpha = 6 / L: the real-space sum is complete at
// r = L (erfc(6) = 2e-17), so a source's 27 images about its nearest one
// hold every term; the wave-vector sum runs to exp(-pi^2 h^2 / 36) < 1e-17
// through structure factors S(k) = sum_j m_j exp(i k.x_j), one pass over
// the sources per block of wave vectors.
constexpr double EW_ALPHAL = 6.0;
struct EwK {
    double kx, ky, kz, c;   // k = 2 pi h / L (one of +h, -h), c = 2 e^(-pi^2 h^2/(aL)^2) / (pi L h^2)
};
constexpr int SF_T = 256, SF_KV = 4;

__device__ __forceinline__ void kahan_add(double &s, double &c, double v)
{
    const double y = v - c, t = s + y;
    c = (t - s) - y;
    s = t;
}

// Structure factors of SF_KV wave vectors per workgroup: sf[2q] = sum_j m_j
// cos(k_q.x_j), sf[2q+1] = sum_j m_j sin(k_q.x_j), each thread over a fixed
// stride of sources, then a fixed-order tree over the workgroup.
__global__ __launch_bounds__(SF_T) void k_ewald_sf(int n, int nk, const EwK *__restrict__ kt,
                                                   const double *__restrict__ r,
                                                   const double *__restrict__ m,
                                                   double *__restrict__ sf)
{
    __shared__ double red[SF_T];
    const int q0 = blockIdx.x * SF_KV;
    double kx[SF_KV], ky[SF_KV], kz[SF_KV], s[2 * SF_KV], c[2 * SF_KV];
    for (int q = 0; q < SF_KV; ++q) {
        const bool ok = q0 + q < nk;
        kx[q] = ok ? kt[q0 + q].kx : 0.0;
        ky[q] = ok ? kt[q0 + q].ky : 0.0;
        kz[q] = ok ? kt[q0 + q].kz : 0.0;
        s[2 * q] = s[2 * q + 1] = c[2 * q] = c[2 * q + 1] = 0.0;
    }
    for (int j = threadIdx.x; j < n; j += SF_T) {
        const double x = r[3 * (size_t)j], y = r[3 * (size_t)j + 1], z = r[3 * (size_t)j + 2];
        const double mj = m[j];
        for (int q = 0; q < SF_KV; ++q) {
            double sn, cs;
            sincos(kx[q] * x + ky[q] * y + kz[q] * z, &sn, &cs);
            kahan_add(s[2 * q], c[2 * q], mj * cs);
            kahan_add(s[2 * q + 1], c[2 * q + 1], mj * sn);
        }
    }
    for (int v = 0; v < 2 * SF_KV; ++v) {
        red[threadIdx.x] = s[v] - c[v];
        __syncthreads();
        for (int w = SF_T / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0 && q0 + v / 2 < nk) sf[2 * (q0 + v / 2) + (v & 1)] = red[0];
        __syncthreads();
    }
}

// Real-space share: one thread per target, DSPLIT workgroups per 256
// targets, each over its share of the sources staged through LDS.
__global__ __launch_bounds__(256) void k_direct_ewald(int n, int nt, const int *__restrict__ tgt,
                                                      const double *__restrict__ r,
                                                      const double *__restrict__ m,
                                                      const double *__restrict__ h, double L,
                                                      double *__restrict__ part)
{
    __shared__ double sx[256], sy[256], sz[256], sm[256], sh[256];
    const int k = blockIdx.x * 256 + threadIdx.x;
    int i = k < nt ? tgt[k] : -1;
    if (i >= n) i = -1;
    const double xi = i >= 0 ? r[3 * (size_t)i] : 0.0, yi = i >= 0 ? r[3 * (size_t)i + 1] : 0.0;
    const double zi = i >= 0 ? r[3 * (size_t)i + 2] : 0.0, hi = i >= 0 ? h[i] : 0.0;
    const double alpha = EW_ALPHAL / L, rc2 = L * L, ka = 2 * alpha / sqrt(M_PI);
    double s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    const int per = (n + DSPLIT - 1) / DSPLIT;
    const int j0 = blockIdx.y * per, j1 = min(n, j0 + per);
    for (int t0 = j0; t0 < j1; t0 += 256) {
        const int j = t0 + threadIdx.x;
        if (j < j1) {
            sx[threadIdx.x] = r[3 * (size_t)j];
            sy[threadIdx.x] = r[3 * (size_t)j + 1];
            sz[threadIdx.x] = r[3 * (size_t)j + 2];
            sm[threadIdx.x] = m[j];
            sh[threadIdx.x] = h[j];
        }
        __syncthreads();
        const int nj = min(256, j1 - t0);
        if (i >= 0)
            for (int e = 0; e < nj; ++e) {
                if (t0 + e == i) continue;
                double dx = sx[e] - xi, dy = sy[e] - yi, dz = sz[e] - zi;
                dx -= L * rint(dx / L);
                dy -= L * rint(dy / L);
                dz -= L * rint(dz / L);
                const double twoh = hi + sh[e], mj = sm[e];
                for (int nx = -1; nx <= 1; ++nx)
                    for (int ny = -1; ny <= 1; ++ny)
                        for (int nz = -1; nz <= 1; ++nz) {
                            const double px = dx + nx * L, py = dy + ny * L, pz = dz + nz * L;
                            const double s2 = px * px + py * py + pz * pz;
                            if (s2 >= rc2) continue;
                            if (s2 == 0.0) {   // coincident: spline(0) minus the erf part
                                double aa, bb;
                                spline_ab<double>(0.0, __builtin_inf(), twoh, aa, bb);
                                kahan_add(s[3], c[3], -mj * (aa - ka));
                                continue;
                            }
                            const double sr = sqrt(s2), dir = 1.0 / sr;
                            double ee = erfc(alpha * sr) * dir;
                            double g = (ee + ka * exp(-alpha * alpha * s2)) / s2;
                            if (s2 < twoh * twoh) {   // softened: spline minus Newton
                                double aa, bb;
                                spline_ab<double>(s2, dir, twoh, aa, bb);
                                ee += aa - dir;
                                g += bb - dir / s2;
                            }
                            kahan_add(s[0], c[0], mj * g * px);
                            kahan_add(s[1], c[1], mj * g * py);
                            kahan_add(s[2], c[2], mj * g * pz);
                            kahan_add(s[3], c[3], -mj * ee);
                        }
            }
        __syncthreads();
    }
    if (i >= 0)
        for (int q = 0; q < 4; ++q) part[((size_t)blockIdx.y * nt + k) * 4 + q] = s[q] - c[q];
}

// Each target: its real-space shares in order, the wave-vector sum through
// the structure factors (its own term has no force), the background of
// every other particle and xi m_i for its own images.  sf[2 nk] holds the
// total mass (the k = 0 entry of the table).
__global__ void k_direct_ewald_out(int nt, int nk, const int *__restrict__ tgt,
                                   const double *__restrict__ r, const double *__restrict__ m,
                                   const EwK *__restrict__ kt, const double *__restrict__ sf,
                                   const double *__restrict__ part, double bg, double xi,
                                   double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nt) return;
    const int i = tgt[k];
    const double x = r[3 * (size_t)i], y = r[3 * (size_t)i + 1], z = r[3 * (size_t)i + 2];
    double s[4];
    for (int q = 0; q < 4; ++q) {
        s[q] = 0.0;
        for (int p = 0; p < DSPLIT; ++p) s[q] += part[((size_t)p * nt + k) * 4 + q];
    }
    double ws[4] = {0, 0, 0, 0}, wc[4] = {0, 0, 0, 0};
    for (int q = 0; q < nk; ++q) {
        double sn, cs;
        sincos(kt[q].kx * x + kt[q].ky * y + kt[q].kz * z, &sn, &cs);
        const double C = sf[2 * q], S = sf[2 * q + 1];
        // sum_j m_j sin(k.(x_i - x_j)) and cos(...)
        const double f = kt[q].c * (sn * C - cs * S);
        kahan_add(ws[0], wc[0], -f * kt[q].kx);
        kahan_add(ws[1], wc[1], -f * kt[q].ky);
        kahan_add(ws[2], wc[2], -f * kt[q].kz);
        kahan_add(ws[3], wc[3], -kt[q].c * (cs * C + sn * S));
    }
    const double mi = m[i];
    for (int q = 0; q < 3; ++q) out[4 * (size_t)k + q] = s[q] + (ws[q] - wc[q]);
    // the k-space self term -sum c m_i is part of xi: take it back out of the sum
    double cself = 0.0;
    for (int q = 0; q < nk; ++q) cself += kt[q].c;
    out[4 * (size_t)k + 3] = s[3] + (ws[3] - wc[3]) + cself * mi + (sf[2 * nk] - mi) * bg + xi * mi;
}
}  // namespace gg

extern "C" int gg_direct_ewald(GGPKD pkd, int nTarget, const int *target, double *a, double *pot)
{
    GG_CHECK(pkd && (nTarget == 0 || (target && a && pot)), GG_EINVAL, "null argument");
    GG_CHECK(nTarget >= 0, GG_EINVAL, "nTarget must be >= 0");
    GG_CHECK(pkd->f[GG_R].p && pkd->f[GG_MASS].p && pkd->f[GG_SOFT].p, GG_ESTATE,
             "positions, masses and softenings must be put before gg_direct_ewald");
    const double L = pkd->fPeriod[0];
    GG_CHECK(L > 0 && pkd->fPeriod[1] == L && pkd->fPeriod[2] == L, GG_EINVAL,
             "gg_direct_ewald needs a cubic box (gg_set_period)");
    const int n = pkd->nLocal;
    for (int k = 0; k < nTarget; ++k)
        GG_CHECK(target[k] >= 0 && target[k] < n, GG_EINVAL, "target %d outside 0..%d",
                 target[k], n - 1);
    if (nTarget == 0) return GG_OK;
    // the wave-vector table (one of each +h, -h) and a k = 0 entry last
    const double aL = EW_ALPHAL, alpha = aL / L, kc2 = 39.0 * aL * aL / (M_PI * M_PI);
    const int hmax = (int)ceil(sqrt(kc2));
    std::vector<EwK> kt;
    for (int hx = 0; hx <= hmax; ++hx)
        for (int hy = -hmax; hy <= hmax; ++hy)
            for (int hz = -hmax; hz <= hmax; ++hz) {
                const int h2 = hx * hx + hy * hy + hz * hz;
                if (hx == 0 && (hy < 0 || (hy == 0 && hz <= 0))) continue;
                if (h2 > kc2) continue;
                kt.push_back({2 * M_PI * hx / L, 2 * M_PI * hy / L, 2 * M_PI * hz / L,
                              2.0 * exp(-M_PI * M_PI * h2 / (aL * aL)) / (M_PI * L * h2)});
            }
    const int nk = (int)kt.size();
    kt.push_back({0.0, 0.0, 0.0, 0.0});
    // xi = lim_{d->0} (psi(d) + 1/|d|): own images, background, k-space
    double xi = 2 * alpha / sqrt(M_PI) + M_PI / (alpha * alpha * L * L * L);
    for (int nx = -1; nx <= 1; ++nx)
        for (int ny = -1; ny <= 1; ++ny)
            for (int nz = -1; nz <= 1; ++nz) {
                if (!nx && !ny && !nz) continue;
                const double rr = L * sqrt((double)(nx * nx + ny * ny + nz * nz));
                if (rr < L) xi -= erfc(alpha * rr) / rr;
            }
    for (int q = 0; q < nk; ++q) xi -= kt[q].c;
    const double bg = M_PI / (alpha * alpha * L * L * L);

    hipStream_t st = pkd->stream;
    DevBuf t, kd, sf, pt, o;
    int rc = GG_OK;
    if (t.reserve(sizeof(int) * nTarget) != GG_OK || kd.reserve(sizeof(EwK) * (nk + 1)) != GG_OK ||
        sf.reserve(sizeof(double) * 2 * (nk + 1)) != GG_OK ||
        pt.reserve(sizeof(double) * 4 * (size_t)DSPLIT * nTarget) != GG_OK ||
        o.reserve(sizeof(double) * 4 * nTarget) != GG_OK)
        rc = GG_EHIP;
    std::vector<double> res(4 * (size_t)nTarget);
    if (rc == GG_OK &&
        (hipMemcpyAsync(t.p, target, sizeof(int) * nTarget, hipMemcpyHostToDevice, st) != hipSuccess ||
         hipMemcpyAsync(kd.p, kt.data(), sizeof(EwK) * (nk + 1), hipMemcpyHostToDevice, st) !=
             hipSuccess))
        rc = GG_EHIP;
    if (rc == GG_OK) {
        const double *r = pkd->f[GG_R].as<double>(), *m = pkd->f[GG_MASS].as<double>();
        hipLaunchKernelGGL(k_ewald_sf, dim3((nk + 1 + SF_KV - 1) / SF_KV), dim3(SF_T), 0, st, n,
                           nk + 1, kd.as<EwK>(), r, m, sf.as<double>());
        hipLaunchKernelGGL(k_direct_ewald, dim3((nTarget + 255) / 256, DSPLIT), dim3(256), 0, st,
                           n, nTarget, t.as<int>(), r, m, pkd->f[GG_SOFT].as<double>(), L,
                           pt.as<double>());
        hipLaunchKernelGGL(k_direct_ewald_out, dim3((nTarget + 255) / 256), dim3(256), 0, st,
                           nTarget, nk, t.as<int>(), r, m, kd.as<EwK>(), sf.as<double>(),
                           pt.as<double>(), bg, xi, o.as<double>());
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(res.data(), o.p, sizeof(double) * 4 * nTarget, hipMemcpyDeviceToHost,
                           st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = GG_EHIP;
    }
    t.release();
    kd.release();
    sf.release();
    pt.release();
    o.release();
    GG_CHECK(rc == GG_OK, GG_EHIP, "gg_direct_ewald: HIP error");
    for (int k = 0; k < nTarget; ++k) {
        for (int d = 0; d < 3; ++d) a[3 * (size_t)k + d] = res[4 * (size_t)k + d];
        pot[k] = res[4 * (size_t)k + 3];
    }
    return GG_OK;
}

extern "C" int gg_tree_info(GGPKD pkd, int *nNodes, int *iRoot, int *perm, int *ints,
                            double *dbls)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GravDev &G = pkd->grav;
    GG_CHECK(G.built, GG_ESTATE, "no gravity tree");
    hipStream_t st = pkd->stream;
    if (nNodes) *nNodes = G.nNodes;
    if (iRoot) *iRoot = G.iRoot;
    if (perm && G.n)
        GG_HIP(hipMemcpyAsync(perm, G.perm.p, sizeof(int) * G.n, hipMemcpyDeviceToHost, st));
    if (ints)
        GG_HIP(hipMemcpyAsync(ints, G.tint.p, sizeof(int) * 6 * G.nNodes, hipMemcpyDeviceToHost,
                              st));
    if (dbls)
        GG_HIP(hipMemcpyAsync(dbls, G.tdbl.p, sizeof(double) * 8 * G.nNodes,
                              hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    return GG_OK;
}

// ------------------------------------------------------- graph and tuner
// The device work of one gg_grav_all (fused walk + interactions, counters,
// scatter into a / fPot / dtGrav) as kernel launches only.
static int grav_enqueue(GGPKD pkd, int iOrder)
{
    GravDev &G = pkd->grav;
    hipStream_t st = pkd->stream;
    const int nb = G.nb;
    hipLaunchKernelGGL(k_zero32, dim3(1), dim3(64), 0, st, 1, G.werr.as<uint32_t>());
    hipLaunchKernelGGL(k_zero32, dim3(1), dim3(64), 0, st, 10, G.stats.as<uint32_t>());
    const IArgs A = make_iargs(pkd, iOrder);
    if (pkd->prec == GG_PREC_FP32) GG_TRY(launch_fused<float>(iOrder, A, G, st));
    else GG_TRY(launch_fused<double>(iOrder, A, G, st));
    hipLaunchKernelGGL(k_grav_stats, dim3(std::min((nb + 255) / 256, 1024)), dim3(256), 0, st, nb,
                       G.bucket.as<int>(), G.bbox.as<double>(), G.wnode.as<WalkNode>(),
                       G.cnt.as<int>(), iOrder, G.stats.as<unsigned long long>());
    hipLaunchKernelGGL(k_grav_weight, dim3((nb + 255) / 256), dim3(256), 0, st, nb,
                       G.bucket.as<int>(), G.bbox.as<double>(), G.wnode.as<WalkNode>(),
                       G.cnt.as<int>(), iOrder, G.perm.as<int>(), G.pact.as<uint32_t>(),
                       (const int *)nullptr, 0, 0, pkd->f[GG_FWEIGHT].as<double>());
    hipLaunchKernelGGL(k_grav_scatter, dim3((G.n + 255) / 256), dim3(256), 0, st, G.n,
                       G.perm.as<int>(), G.pact.as<uint32_t>(), G.acc.as<double>(),
                       pkd->f[GG_A].as<double>(), pkd->f[GG_POT].as<double>(),
                       pkd->f[GG_DTGRAV].as<double>());
    GG_HIP(hipGetLastError());
    return GG_OK;
}

static void graph_release(GravDev &G)
{
    if (G.gexec) (void)hipGraphExecDestroy(G.gexec);
    if (G.graph) (void)hipGraphDestroy(G.graph);
    G.gexec = nullptr;
    G.graph = nullptr;
    G.graphOrder = 0;
}

extern "C" int gg_grav_graph(GGPKD pkd, int iOrder)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GravDev &G = pkd->grav;
    GG_CHECK(G.built, GG_ESTATE, "gg_grav_graph before gg_build_binary / gg_build_forest");
    GG_CHECK(G.n == pkd->nLocal, GG_ESTATE, "particle count changed since the tree was built");
    GG_CHECK(iOrder >= 1 && iOrder <= G.iOrder, GG_EINVAL,
             "iOrder %d exceeds the tree's moments (%d)", iOrder, G.iOrder);
    GG_CHECK(G.nb > 0, GG_EINVAL, "no active bucket");
    GG_CHECK(G.maxBucket <= 64, GG_EUNSUP,
             "a bucket of %d coincident particles needs the listed walk: use gg_grav_all",
             G.maxBucket);
    GG_CHECK(!(pkd->fPeriod[0] > 0 || pkd->fPeriod[1] > 0 || pkd->fPeriod[2] > 0), GG_EUNSUP,
             "the captured step is the isolated one: a handle with a periodic box uses "
             "gg_grav_all");
    const int f[] = {GG_A, GG_POT, GG_DTGRAV, GG_FWEIGHT};
    for (int k : f) GG_TRY(ensure_field(pkd, k));
    GG_TRY(G.werr.reserve(sizeof(int) * 11));
    GG_TRY(G.stats.reserve(sizeof(unsigned long long) * 5));
    graph_release(G);
    hipStream_t st = pkd->stream;
    {   // the fused walk must hold every frontier of this tree: try it once
        hipLaunchKernelGGL(k_zero32, dim3(1), dim3(64), 0, st, 1, G.werr.as<uint32_t>());
        const IArgs A = make_iargs(pkd, iOrder);
        if (pkd->prec == GG_PREC_FP32) GG_TRY(launch_fused<float>(iOrder, A, G, st));
        else GG_TRY(launch_fused<double>(iOrder, A, G, st));
        int werr = 0;
        GG_HIP(hipMemcpyAsync(&werr, G.werr.p, sizeof(int), hipMemcpyDeviceToHost, st));
        GG_HIP(hipStreamSynchronize(st));
        GG_CHECK(!werr, GG_EUNSUP,
                 "this tree needs the listed walk (a frontier outgrows the fused kernel): "
                 "use gg_grav_all");
    }
    GG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = grav_enqueue(pkd, iOrder);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    GG_CHECK(rc == GG_OK, rc, "capture of the gravity step failed: %s", gg_last_error());
    GG_HIP(e);
    G.graph = g;
    GG_HIP(hipGraphInstantiate(&G.gexec, g, nullptr, nullptr, 0));
    G.graphOrder = iOrder;
    G.graphTree = G.treeId;
    return GG_OK;
}

extern "C" int gg_grav_replay(GGPKD pkd, int *nActive, double *pdPartSum, double *pdCellSum,
                              double *pdSoftSum, double *pdFlop)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GravDev &G = pkd->grav;
    GG_CHECK(G.gexec && G.graphTree == G.treeId && G.built, GG_ESTATE,
             "no graph for the current tree: gg_grav_graph after building");
    hipStream_t st = pkd->stream;
    GG_HIP(hipGraphLaunch(G.gexec, st));
    int werr = 0;
    unsigned long long hs[5];
    GG_HIP(hipMemcpyAsync(&werr, G.werr.p, sizeof(int), hipMemcpyDeviceToHost, st));
    GG_HIP(hipMemcpyAsync(hs, G.stats.p, sizeof hs, hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    G.listsCurrent = 0;
    // gg_grav_graph checked that no frontier of this tree overflows
    GG_CHECK(!werr, GG_EHIP, "a walk frontier overflowed inside the graph");
    if (nActive) *nActive = (int)hs[0];
    if (pdPartSum) *pdPartSum = (double)hs[1];
    if (pdCellSum) *pdCellSum = (double)hs[2];
    if (pdSoftSum) *pdSoftSum = (double)hs[3];
    if (pdFlop) *pdFlop = (double)hs[4];
    return GG_OK;
}

extern "C" int gg_set_grav_schedule(GGPKD pkd, int mode)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GG_CHECK(mode >= 0 && mode <= 2, GG_EINVAL, "schedule %d: 0, 1 or 2", mode);
    pkd->grav.sched = mode;
    pkd->grav.qTree = -1;
    return GG_OK;
}

// Bucket-size autotuner: build with each candidate nBucket, time the device
// work of gg_grav_all (second of two runs), keep the fastest.  The caller's
// a, fPot and dtGrav are left as they were.
extern "C" int gg_tune_bucket(GGPKD pkd, int nSys, const int *sysOff, int iOpenType,
                              double dCrit, int iOrder, const int *cand, int nCand, double *ms,
                              int *pBest)
{
    GG_CHECK(pkd && cand && nCand >= 1, GG_EINVAL, "null argument or no candidate");
    const int n = pkd->nLocal;
    const int f[] = {GG_A, GG_POT, GG_DTGRAV, GG_FWEIGHT};
    for (int k : f) GG_TRY(ensure_field(pkd, k));
    hipStream_t st = pkd->stream;
    DevBuf save;
    GG_TRY(save.reserve(sizeof(double) * 6 * (size_t)std::max(n, 1)));
    double *sv = save.as<double>();
    int rc = GG_OK;
    auto copy_fields = [&](bool out) -> int {
        const size_t w3 = sizeof(double) * 3 * (size_t)n, w1 = sizeof(double) * (size_t)n;
        void *fa = pkd->f[GG_A].p, *fp = pkd->f[GG_POT].p, *fd = pkd->f[GG_DTGRAV].p;
        void *fw = pkd->f[GG_FWEIGHT].p;
        char *b = reinterpret_cast<char *>(sv);
        if (out) {
            GG_HIP(hipMemcpyAsync(b, fa, w3, hipMemcpyDeviceToDevice, st));
            GG_HIP(hipMemcpyAsync(b + w3, fp, w1, hipMemcpyDeviceToDevice, st));
            GG_HIP(hipMemcpyAsync(b + w3 + w1, fd, w1, hipMemcpyDeviceToDevice, st));
            GG_HIP(hipMemcpyAsync(b + w3 + 2 * w1, fw, w1, hipMemcpyDeviceToDevice, st));
        } else {
            GG_HIP(hipMemcpyAsync(fa, b, w3, hipMemcpyDeviceToDevice, st));
            GG_HIP(hipMemcpyAsync(fp, b + w3, w1, hipMemcpyDeviceToDevice, st));
            GG_HIP(hipMemcpyAsync(fd, b + w3 + w1, w1, hipMemcpyDeviceToDevice, st));
            GG_HIP(hipMemcpyAsync(fw, b + w3 + 2 * w1, w1, hipMemcpyDeviceToDevice, st));
        }
        GG_HIP(hipStreamSynchronize(st));
        return GG_OK;
    };
    auto build = [&](int nb) -> int {
        return sysOff ? gg_build_forest(pkd, nSys, sysOff, nb, iOpenType, dCrit, iOrder)
                      : gg_build_binary(pkd, nb, iOpenType, dCrit, iOrder, 0, 1);
    };
    if (n > 0) rc = copy_fields(true);
    int best = -1;
    double bestMs = 0;
    for (int k = 0; k < nCand && rc == GG_OK; ++k) {
        rc = build(cand[k]);
        for (int rep = 0; rep < 2 && rc == GG_OK; ++rep)
            rc = gg_grav_all(pkd, 0, 0, iOrder, 0, 0, 0, 0, 0, 0, 0, 0, nullptr, nullptr,
                             nullptr, nullptr, nullptr, nullptr);
        if (rc != GG_OK) break;
        timer_resolve(pkd);
        const double t = pkd->tm.ms[1] + pkd->tm.ms[8];
        if (ms) ms[k] = t;
        if (best < 0 || t < bestMs) {
            best = k;
            bestMs = t;
        }
    }
    if (rc == GG_OK) rc = build(cand[best]);
    if (n > 0) {
        const int rc2 = copy_fields(false);
        if (rc == GG_OK) rc = rc2;
    }
    save.release();
    if (rc != GG_OK) return rc;
    if (pBest) *pBest = cand[best];
    return GG_OK;
}
