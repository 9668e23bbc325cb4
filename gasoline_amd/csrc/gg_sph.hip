// SPH neighbour loop on the GPU: search structure, exact kNN, DenDVDX,
// equation of state, ball gather and the symmetric pressure/viscosity force
// evaluated in gather form.
//
// Reference path: msrSph (master.c:9258-9362) -> pkdBuildLocal
// (pkd.c:2732-2848) -> smSmooth + DenDVDX (smooth.c:1439-1713,
// smoothfcn.c:3390-3627) -> pkdGasPressure (pkd.c:5750-5785) -> smReSmooth
// + SphPressureTermsSym (smooth.c:1716-1749, smoothfcn.c:2276-2385,
// SphPressureTerms.h).
//
// Search structure: gas particles sorted by 63-bit Morton key, cut into
// leaves of 32 consecutive particles, with a complete binary hierarchy of
// FP64 bounding boxes over the leaves (heap layout, root = 1).  The
// neighbour SET of a particle is fixed by FP64 distances alone, so it does
// not depend on this structure.
#include "gg_internal.h"

#include <utility>

#include <cstring>
#include <algorithm>
#include <cmath>
#include <cfloat>
#include <climits>

#include <rocprim/device/device_radix_sort.hpp>

namespace gg {

constexpr int FCAP = 1024;         // BFS frontier capacity (nodes)
constexpr int CAP = 256;           // candidate buffer of a single-query search (4 per lane)
constexpr int ECAP = 48;           // force-pass extras kept per particle by the density pass
static_assert(ECAP % 4 == 0 && ECAP >= 16, "extras lists are sorted 16 bytes at a time");
constexpr int MAXTRY = 48;         // radius adjustments per query
constexpr double GG_FLT_MAX = (double)3.402823466E+38F;   // floattype.h:9

// ------------------------------------------------------------ geometry
// Separation p - q, minimum periodic image chosen per axis, image applied to
// the query point first (sx = x +/- lx, INTERSECT smooth.h:291-390).
__device__ __forceinline__ double img_sep(double p, double q, int per, double L)
{
#pragma clang fp contract(off)
    double s = p;
    if (per) {
        const double t = p - q;
        if (t > 0.5 * L) s = p - L;
        else if (t < -0.5 * L) s = p + L;
    }
    return s - q;
}

__device__ __forceinline__ double sq_dist(double dx, double dy, double dz)
{
#pragma clang fp contract(off)
    return dx * dx + dy * dy + dz * dz;
}

// FP32 bounds rounded outwards: f_down(d) <= d <= f_up(d).
__device__ __forceinline__ float f_down(double d)
{
    float f = (float)d;
    if ((double)f > d) {
        int b = __float_as_int(f);
        b = f > 0 ? b - 1 : (f == 0 ? (int)0x80000001u : b + 1);
        f = __int_as_float(b);
    }
    return f;
}

__device__ __forceinline__ float f_up(double d)
{
    float f = (float)d;
    if ((double)f < d) {
        int b = __float_as_int(f);
        b = f > 0 ? b + 1 : (f == 0 ? 1 : b - 1);
        f = __int_as_float(b);
    }
    return f;
}

// Squared distance from a point to a box, minimised over periodic images.
__device__ __forceinline__ double box_d2(const double *b, double x, double y,
                                         double z, int per, const double *L)
{
    const double p[3] = {x, y, z};
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double lo = b[k], hi = b[3 + k];
        double d = p[k] < lo ? lo - p[k] : (p[k] > hi ? p[k] - hi : 0.0);
        if (per) {
            const double a = p[k] + L[k], c = p[k] - L[k];
            const double da = a < lo ? lo - a : (a > hi ? a - hi : 0.0);
            const double dc = c < lo ? lo - c : (c > hi ? c - hi : 0.0);
            d = fmin(d, fmin(da, dc));
        }
        d2 += d * d;
    }
    return d2;
}

// ---------------------------------------------------- structure kernels
__global__ void k_select(int n, const uint32_t *__restrict__ bits, uint32_t mask,
                         int *__restrict__ sel)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sel[i] = ((bits[i] & mask) == mask) ? 1 : 0;
}

__global__ void k_compact(int n, const int *__restrict__ sel,
                          const int64_t *__restrict__ pos, int *__restrict__ gidx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && sel[i]) gidx[pos[i]] = i;
}

__global__ void k_bbox(int n, const int *__restrict__ gidx, const double *__restrict__ r,
                       double *__restrict__ part)
{
    __shared__ double s[6][256];
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int i = gidx[k];
        for (int d = 0; d < 3; ++d) {
            const double v = r[3 * (size_t)i + d];
            lo[d] = fmin(lo[d], v);
            hi[d] = fmax(hi[d], v);
        }
    }
    for (int d = 0; d < 3; ++d) {
        s[d][threadIdx.x] = lo[d];
        s[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < 3; ++d) {
                s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + w]);
                s[3 + d][threadIdx.x] = fmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 6; ++d) part[6 * (size_t)blockIdx.x + d] = s[d][0];
}

// The Morton grid from k_bbox's nbb partial boxes (one wavefront): out[0..2]
// the lower corner, out[3] the scale 65535 / the largest extent (1 for a
// point), so the keys' quantisation never leaves the device.
__global__ void k_bbox_final(int nbb, const double *__restrict__ part, double *__restrict__ out)
{
    const int l = threadIdx.x;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int b = l; b < nbb; b += 64)
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], part[6 * (size_t)b + d]);
            hi[d] = fmax(hi[d], part[6 * (size_t)b + 3 + d]);
        }
    for (int o = 32; o > 0; o >>= 1)
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], o));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], o));
        }
    if (l == 0) {
        double ext = 0;
        for (int d = 0; d < 3; ++d) ext = fmax(ext, hi[d] - lo[d]);
        if (!(ext > 0)) ext = 1.0;
        for (int d = 0; d < 3; ++d) out[d] = lo[d];
        out[3] = 65535.0 / ext;
    }
}

// The one bucket of slots [0, n), n <= 64 (one wavefront): its box, the box
// in FP32 rounded outwards, and its density over the quantisation cube.
__global__ void k_bbox_slots(int n, const double *__restrict__ grid, const double *__restrict__ sx,
                             const double *__restrict__ sy, const double *__restrict__ sz,
                             double *__restrict__ box, float *__restrict__ fbox,
                             double *__restrict__ den)
{
    const int l = threadIdx.x;
    double v[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, u[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    if (l < n) {
        v[0] = u[0] = sx[l];
        v[1] = u[1] = sy[l];
        v[2] = u[2] = sz[l];
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int d = 0; d < 3; ++d) {
            v[d] = fmin(v[d], __shfl_xor(v[d], o));
            u[d] = fmax(u[d], __shfl_xor(u[d], o));
        }
    if (l == 0) {
        for (int d = 0; d < 3; ++d) {
            box[d] = v[d];
            box[3 + d] = u[d];
            fbox[d] = f_down(v[d]);
            fbox[3 + d] = f_up(u[d]);
        }
        const double sc = grid[3];
        den[0] = ldexp((double)n * sc * sc * sc, -48);
    }
}

__device__ __forceinline__ uint64_t spread21(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ void k_morton(int n, const int *__restrict__ gidx, const double *__restrict__ r,
                         const double *__restrict__ grid, uint64_t *__restrict__ key,
                         int *__restrict__ val)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = gidx[k];
    const double lo[3] = {grid[0], grid[1], grid[2]}, sc[3] = {grid[3], grid[3], grid[3]};
    uint64_t q[3];
    for (int d = 0; d < 3; ++d) {
        double t = (r[3 * (size_t)i + d] - lo[d]) * sc[d];
        t = fmin(fmax(t, 0.0), 65535.0);
        q[d] = (uint64_t)t;
    }
    key[k] = spread21(q[0]) | (spread21(q[1]) << 1) | (spread21(q[2]) << 2);
    val[k] = i;
}

__global__ void k_gather_pos(int n, const int *__restrict__ order, const double *__restrict__ r,
                             double *__restrict__ sx, double *__restrict__ sy,
                             double *__restrict__ sz)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int i = order[s];
    sx[s] = r[3 * (size_t)i];
    sy[s] = r[3 * (size_t)i + 1];
    sz[s] = r[3 * (size_t)i + 2];
}

// ------------------------------------------------- radix tree (LBVH)
// Binary radix tree over the sorted Morton keys (Karras 2012): internal
// node i in [0, n-2] covers the slot range rng[2i]..rng[2i+1]; a child code
// c >= 0 is internal node c, c < 0 is the single particle ~c.  Equal keys are
// ordered by slot.  Buckets are the maximal subtrees of at most bkt
// particles (the caller's nBucket, at most BKT): every bucket is one
// Morton-prefix cell, so its box is tight.  A node of more than bkt
// particles covers whole consecutive buckets, so its box is the union of a
// bucket range, read from a sparse table of bucket boxes (FP32, rounded
// outwards); no bottom-up pass is needed.  The kNN serves one bucket's
// queries with one wavefront, two buckets' candidates per load, so a bucket
// holds at most BKT = 32 particles.
constexpr int BKT = 32;

__device__ __forceinline__ int lcp_delta(const uint64_t *__restrict__ k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a != b) return __clzll(a ^ b);
    return 64 + __clz(i ^ j);
}

__global__ void k_lbvh(int n, const uint64_t *__restrict__ key, int *__restrict__ child,
                       int *__restrict__ rng)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (lcp_delta(key, n, i, i + 1) - lcp_delta(key, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = lcp_delta(key, n, i, i - d);
    int lmax = 2;
    while (lcp_delta(key, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (lcp_delta(key, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = lcp_delta(key, n, i, j);
    int sp = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (lcp_delta(key, n, i, i + (sp + t) * d) > dnode) sp += t;
    } while (t > 1);
    const int g = i + sp * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    child[2 * i] = (lo == g) ? ~g : g;
    child[2 * i + 1] = (hi == g + 1) ? ~(g + 1) : g + 1;
    rng[2 * i] = lo;
    rng[2 * i + 1] = hi;
}

__device__ __forceinline__ void child_info(int c, const int *__restrict__ rng, int &first, int &cnt)
{
    if (c >= 0) {
        first = rng[2 * c];
        cnt = rng[2 * c + 1] - first + 1;
    } else {
        first = ~c;
        cnt = 1;
    }
}

// Mark bucket starts: children of nodes larger than bkt that are not.  For
// each, also note the parent's particle count and the length of the key
// prefix its particles share (the parent's Morton cell).
__global__ void k_bucket_mark(int n, int bkt, const uint64_t *__restrict__ key,
                              const int *__restrict__ child, const int *__restrict__ rng,
                              int *__restrict__ flag, int *__restrict__ cnt,
                              int2 *__restrict__ par)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int lo = rng[2 * i], hi = rng[2 * i + 1];
    if (hi - lo + 1 <= bkt) return;
    const int pre = min(lcp_delta(key, n, lo, hi), 64) - 16;   // of the 48 key bits
    for (int k = 0; k < 2; ++k) {
        int f, m;
        child_info(child[2 * i + k], rng, f, m);
        if (m <= bkt) {
            flag[f] = 1;
            cnt[f] = m;
            par[f] = make_int2(hi - lo + 1, pre);
        }
    }
}

// Bucket b: first slot, count, FP64 box, FP32 outward box (sparse table
// level 0), the bucket of each of its slots, and a number density: the
// parent's particle count over the volume of its Morton cell (a key prefix
// of p bits spans 2^(48-p) quantisation cells of volume 1/sc^3).
__global__ void k_bucket_start(int n, const int *__restrict__ flag, const int *__restrict__ cnt,
                               const int2 *__restrict__ par, const double *__restrict__ grid,
                               const int64_t *__restrict__ pos, int *__restrict__ bfirst,
                               int *__restrict__ bcnt, double *__restrict__ bden,
                               int *__restrict__ bof)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int f = flag[k];
    const int b = (int)pos[k] + f - 1;   // the buckets starting at or before k, less one
    bof[k] = b;
    if (!f) return;
    const int2 pp = par[k];
    const double sc = grid[3];
    bden[b] = ldexp((double)pp.x * sc * sc * sc, pp.y - 48);
    bfirst[b] = k;
    bcnt[b] = cnt[k];
}

// Each bucket's box (FP64, and FP32 rounded outwards), 32 lanes per bucket
// (a bucket holds at most 32 slots), the slots' positions read coalesced.
__global__ __launch_bounds__(256) void k_bucket_box(int nb, const int *__restrict__ bfirst,
                                                    const int *__restrict__ bcnt,
                                                    const double *__restrict__ sx,
                                                    const double *__restrict__ sy,
                                                    const double *__restrict__ sz,
                                                    double *__restrict__ bbox,
                                                    float *__restrict__ fbox)
{
    const int b = (blockIdx.x * blockDim.x + threadIdx.x) >> 5, l = threadIdx.x & 31;
    if (b >= nb) return;
    const int f = bfirst[b], m = bcnt[b];
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    if (l < m) {
        lo[0] = hi[0] = sx[f + l];
        lo[1] = hi[1] = sy[f + l];
        lo[2] = hi[2] = sz[f + l];
    }
    for (int o = 16; o > 0; o >>= 1)
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], o));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], o));
        }
    if (l == 0)
        for (int d = 0; d < 3; ++d) {
            bbox[6 * (size_t)b + d] = lo[d];
            bbox[6 * (size_t)b + 3 + d] = hi[d];
            fbox[6 * (size_t)b + d] = f_down(lo[d]);
            fbox[6 * (size_t)b + 3 + d] = f_up(hi[d]);
        }
}

// Sparse table levels lv0 + 1 .. lv0 + nl (nl <= 4) from level lv0: level
// lv0 + j at b is the union of the 2^j level-lv0 boxes at b + m 2^lv0 (the
// same unions the pairwise construction forms, min and max being exact),
// stored where b + 2^(lv0 + j) <= nb.
__global__ void k_sparse_levels(int nb, int lv0, int nl, float *__restrict__ sp)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int w = 1 << lv0;
    if (b + 2 * w > nb) return;
    const float *src = sp + 6 * (size_t)lv0 * nb;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int m = 0; m < (1 << nl); ++m) {
        if (b + (m + 1) * w > nb) break;
        const float *t = src + 6 * ((size_t)b + (size_t)m * w);
        for (int d = 0; d < 3; ++d) {
            lo[d] = fminf(lo[d], t[d]);
            hi[d] = fmaxf(hi[d], t[3 + d]);
        }
        const int j = 31 - __clz(m + 1);   // m + 1 = 2^j: level lv0 + j is complete
        if (m > 0 && (m + 1) == (1 << j) && b + (m + 1) * w <= nb) {
            float *o = sp + 6 * ((size_t)(lv0 + j) * nb + b);
            for (int d = 0; d < 3; ++d) {
                o[d] = lo[d];
                o[3 + d] = hi[d];
            }
        }
    }
}

// Child record of an internal node: the child's particle count, its bucket
// index when it is a bucket (count <= bkt) or its internal-node index
// otherwise, and its box in FP32 rounded outwards, so a box test on the
// record never excludes what the FP64 box would include.
struct __align__(16) CRec {
    int code, cnt;
    float lo[3], hi[3];
};

__global__ void k_child_rec(int n, int nb, int bkt, const int *__restrict__ child,
                            const int *__restrict__ rng, const int *__restrict__ bof,
                            const float *__restrict__ sparse, CRec *__restrict__ rec)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2 * (n - 1)) return;
    const int c = child[e];
    int f, m;
    child_info(c, rng, f, m);
    const int b0 = bof[f], b1 = bof[f + m - 1];
    const int len = b1 - b0 + 1, lv = 31 - __clz(len);
    const float *t0 = sparse + 6 * ((size_t)lv * nb + b0);
    const float *t1 = sparse + 6 * ((size_t)lv * nb + b1 - (1 << lv) + 1);
    CRec r;
    r.code = m > bkt ? c : b0;
    r.cnt = m;
    for (int d = 0; d < 3; ++d) {
        r.lo[d] = fminf(t0[d], t1[d]);
        r.hi[d] = fmaxf(t0[3 + d], t1[3 + d]);
    }
    rec[e] = r;
}

// -------------------------------------------------------- kNN / gather
struct QArgs {
    int n, K, per, mode;               // mode 0 kNN, 1 gather count, 2 gather fill
    int bkt;                           // bucket size of the structure
    double L[3];
    const double *sx, *sy, *sz;
    const CRec *rec;                   // radix tree child records
    const int *bof;                    // slot -> bucket
    const double *bbox;                // bucket boxes
    const int *bfirst, *bcnt;          // bucket first slot, count
    const double *bden;                // bucket number density estimate
    const float *fbox;                 // bucket boxes, FP32 rounded outwards
    double slack;                      // first trial ball = slack * fBall2 hint
    const int *order;                  // slot -> caller (tie key)
    const uint32_t *qflag;             // slot order: nonzero = query
    const double *guess;               // slot order: fBall2 guess (<= 0: none)
    double *fb2;                       // kNN out / gather in (slot order)
    int *nbr;                          // kNN: K per slot; fill: CSR entries
    int *cnt;                          // gather count out
    const int64_t *off;                // gather fill offsets
    unsigned int *err;                 // [0] failures, [1] boundary ties, [10] kNN retries
    const int *qlist;                  // optional: block b serves slot qlist[b]
    const unsigned *nq;                // optional, with qlist: its length, on the device
    const double *grl2;                // optional, kNN retries: the bucket groups' radii^2
    const int *gnb;                    //   and sizes (-1: overflowed)
};

// Squared radius of a ball expected to hold 2K particles at number density
// den (a cold start, with no fBall2 to go by).
__device__ __forceinline__ double cold_r2(double den, int K)
{
    double r3 = (2.0 * K) / den * (3.0 / (4.0 * M_PI));
    if (!(r3 > 0) || !(r3 < DBL_MAX)) r3 = 1e-30;
    return cbrt(r3 * r3);
}

__device__ __forceinline__ bool kv_less(double da, int ca, double db, int cb)
{
    return da < db || (da == db && ca < cb);
}

// Bitonic sort of 256 (d2, caller, slot) triples held 4 per lane
// (element e = lane + 64 k), ascending in (d2, caller).
__device__ void bitonic256(double (&d)[4], int (&c)[4], int (&s)[4], int lane)
{
    for (int size = 2; size <= 256; size <<= 1) {
        for (int j = size >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const int kk = j >> 6;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int k2 = k ^ kk;
                    if (k2 <= k) continue;
                    const int e = lane + 64 * k;
                    const bool asc = (e & size) == 0;
                    const bool sw = asc ? kv_less(d[k2], c[k2], d[k], c[k])
                                        : kv_less(d[k], c[k], d[k2], c[k2]);
                    if (sw) {
                        const double td = d[k]; d[k] = d[k2]; d[k2] = td;
                        const int tc = c[k]; c[k] = c[k2]; c[k2] = tc;
                        const int ts = s[k]; s[k] = s[k2]; s[k2] = ts;
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int e = lane + 64 * k;
                    const double pd = __shfl_xor(d[k], j);
                    const int pc = __shfl_xor(c[k], j);
                    const int ps = __shfl_xor(s[k], j);
                    const bool lower = (lane & j) == 0;
                    const bool asc = (e & size) == 0;
                    const bool take = (lower == asc) ? kv_less(pd, pc, d[k], c[k])
                                                     : kv_less(d[k], c[k], pd, pc);
                    if (take) {
                        d[k] = pd;
                        c[k] = pc;
                        s[k] = ps;
                    }
                }
            }
        }
    }
}

// Bitonic sort of 128 triples held 2 per lane.
__device__ void bitonic128(double (&d)[2], int (&c)[2], int (&s)[2], int lane)
{
    for (int size = 2; size <= 128; size <<= 1) {
        for (int j = size >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const bool asc = (lane & size) == 0;
                const bool sw = asc ? kv_less(d[1], c[1], d[0], c[0]) : kv_less(d[0], c[0], d[1], c[1]);
                if (sw) {
                    const double td = d[0]; d[0] = d[1]; d[1] = td;
                    const int tc = c[0]; c[0] = c[1]; c[1] = tc;
                    const int ts = s[0]; s[0] = s[1]; s[1] = ts;
                }
            } else {
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const int e = lane + 64 * k;
                    const double pd = __shfl_xor(d[k], j);
                    const int pc = __shfl_xor(c[k], j);
                    const int ps = __shfl_xor(s[k], j);
                    const bool lower = (lane & j) == 0;
                    const bool asc = (e & size) == 0;
                    const bool take = (lower == asc) ? kv_less(pd, pc, d[k], c[k])
                                                     : kv_less(d[k], c[k], pd, pc);
                    if (take) {
                        d[k] = pd;
                        c[k] = pc;
                        s[k] = ps;
                    }
                }
            }
        }
    }
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wavefront-wide breadth-first search of the radix tree: appends to
// (bf, bc) the buckets whose box passes test(box) and returns their number
// (their particle total in *tot), or -1 when the frontier or the bucket
// list overflows.  fr holds 2*FC ints; bf and bc may be LDS or global.
template <typename Test>
__device__ int wave_bfs(const QArgs &A, Test test, int *fr, int FC, int *bf, int *bc, int BC,
                        int lane, int *tot)
{
    if (A.n <= A.bkt) {   // the root is the only bucket
        if (lane == 0) {
            bf[0] = 0;
            bc[0] = A.n;
        }
        *tot = A.n;
        wave_sync();
        return 1;
    }
    int nb = 0, nf = 1, cur = 0, np = 0;
    if (lane == 0) fr[0] = 0;
    wave_sync();
    while (nf > 0) {
        int nn = 0;
        for (int base = 0; base < 2 * nf; base += 64) {
            const int k = base + lane;
            bool hit = false, big = false;
            CRec r;
            r.code = r.cnt = 0;
            if (k < 2 * nf) {
                r = A.rec[2 * fr[cur * FC + (k >> 1)] + (k & 1)];
                const double b[6] = {r.lo[0], r.lo[1], r.lo[2], r.hi[0], r.hi[1], r.hi[2]};
                hit = test(b);
                big = r.cnt > A.bkt;
            }
            const unsigned long long mb = __ballot(hit && !big);
            const unsigned long long mi = __ballot(hit && big);
            if (hit && !big) {
                const int p = nb + __popcll(mb & ((1ull << lane) - 1));
                if (p < BC) {
                    bf[p] = r.code;
                    bc[p] = r.cnt;
                }
                np += r.cnt;
            }
            if (hit && big) {
                const int p = nn + __popcll(mi & ((1ull << lane) - 1));
                if (p < FC) fr[(cur ^ 1) * FC + p] = r.code;
            }
            nb += __popcll(mb);
            nn += __popcll(mi);
        }
        wave_sync();
        if (nn > FC || nb > BC) return -1;
        nf = nn;
        cur ^= 1;
    }
    for (int o = 32; o > 0; o >>= 1) np += __shfl_xor(np, o);
    *tot = np;
    return nb;
}

constexpr int QB = 512;   // buckets per single-query search

// LDS of a single-query search.
struct QLds {
    int fr[2 * FCAP];
    int bfst[QB], bcnt[QB];
    double cd[CAP];
    int cc[CAP], cs[CAP];
};

// One query slot q, by one 64-lane wavefront.  mode 0: exact kNN by (d2,
// caller index) with a radius grown or shrunk until the ball holds between
// K and CAP particles; mode 1/2: ball gather at fBall2 (d2 <= fBall2).
__device__ void query_one(const QArgs &A, int q, QLds &S, int lane)
{
    int *fr = S.fr, *bfst = S.bfst, *bcnt = S.bcnt, *cc = S.cc, *cs = S.cs;
    double *cd = S.cd;
    if (!A.qflag[q]) {
        if (A.mode == 1 && lane == 0) A.cnt[q] = 0;
        return;
    }
    const double x = A.sx[q], y = A.sy[q], z = A.sz[q];
    const int K = A.K;
    double R2;
    if (A.mode != 0) {
        R2 = A.fb2[q];
    } else if (A.guess && A.guess[q] > 0) {
        R2 = A.guess[q] * A.slack;
    } else {
        R2 = cold_r2(A.bden[A.bof[q]], K);
    }
    if (A.mode == 0 && A.grl2) {
        // a retry whose bucket group was whole: the group radius held too
        // few (or no settling radius) for it, so start beyond it
        const int b = A.bof[q];
        if (A.gnb[b] >= 0) R2 = fmax(R2, 1.6 * A.grl2[b]);
    }
    // kNN: lo2 < R2 < hi2 brackets the radii that hold K..CAP particles (lo2:
    // the largest tried that held fewer than K; hi2: the smallest that held
    // more than CAP or met too many buckets); once both are known the next
    // radius bisects them, so a thin shell of equidistant neighbours (an
    // isolated particle far from the rest) is still found
    double lo2 = 0.0, hi2 = DBL_MAX;
    auto bisect = [&]() {
        return hi2 > 4.0 * lo2 ? sqrt(lo2) * sqrt(hi2) : 0.5 * (lo2 + hi2);
    };
    for (int attempt = 0; attempt < MAXTRY; ++attempt) {
        const double R2b = R2 * (1.0 + 1e-12);
        int np;
        const int nb = wave_bfs(
            A, [&](const double *b) { return box_d2(b, x, y, z, A.per, A.L) <= R2b; }, fr, FCAP,
            bfst, bcnt, QB, lane, &np);
        if (nb < 0) {
            if (A.mode != 0) break;   // a ball gather cannot change its radius
            hi2 = fmin(hi2, R2);
            R2 = lo2 > 0.0 ? bisect() : R2 * 0.25;
            continue;
        }
        for (int k = lane; k < nb; k += 64) bfst[k] = A.bfirst[bfst[k]];
        wave_sync();
        int total = 0;
        for (int bi = 0; bi < nb; bi += 2) {
            const int b = bi + (lane >> 5), o = lane & 31;
            bool ok = false;
            double d2 = 0.0;
            int slot = 0;
            if (b < nb && o < bcnt[b]) {
                slot = bfst[b] + o;
                const double dx = img_sep(x, A.sx[slot], A.per, A.L[0]);
                const double dy = img_sep(y, A.sy[slot], A.per, A.L[1]);
                const double dz = img_sep(z, A.sz[slot], A.per, A.L[2]);
                d2 = sq_dist(dx, dy, dz);
                ok = d2 <= R2;
            }
            const unsigned long long m = __ballot(ok);
            if (ok) {
                const int p = total + __popcll(m & ((1ull << lane) - 1));
                if (p < CAP) {
                    cd[p] = d2;
                    cs[p] = slot;
                    cc[p] = A.order[slot];
                }
            }
            total += __popcll(m);
        }
        wave_sync();
        if (A.mode == 0) {
            if (total < K) {
                lo2 = fmax(lo2, R2);
                const double f = total > 0 ? cbrt((2.0 * K / total) * (2.0 * K / total)) : 4.0;
                R2 = hi2 < DBL_MAX ? bisect() : R2 * fmin(fmax(f, 1.6), 16.0);
                continue;
            }
            if (total > CAP) {
                hi2 = fmin(hi2, R2);
                const double f = cbrt((1.5 * K / total) * (1.5 * K / total));
                R2 = lo2 > 0.0 ? bisect() : R2 * fmin(fmax(f, 0.05), 0.9);
                continue;
            }
        } else if (total > CAP) {
            break;   // ball gather larger than the buffer: reported below
        }
        double d[4];
        int c[4], s[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = lane + 64 * k;
            if (e < total) {
                d[k] = cd[e];
                c[k] = cc[e];
                s[k] = cs[e];
            } else {
                d[k] = DBL_MAX;
                c[k] = 0x7fffffff;
                s[k] = -1;
            }
        }
        bitonic256(d, c, s, lane);
        if (A.mode == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int e = lane + 64 * k;
                if (e < K) A.nbr[(size_t)q * K + e] = s[k];
            }
            const int eK = K - 1;
            double dK = 0, dK1 = DBL_MAX;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double a = __shfl(d[k], eK & 63);
                if (k == (eK >> 6)) dK = a;
                const double b = __shfl(d[k], K & 63);
                if (k == (K >> 6) && K < CAP) dK1 = b;
            }
            if (lane == 0) {
                A.fb2[q] = dK;
                if (total > K && dK1 == dK) atomicAdd(&A.err[1], 1u);
            }
        } else if (A.mode == 1) {
            if (lane == 0) A.cnt[q] = total;
        } else {
            const int64_t o = A.off[q];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int e = lane + 64 * k;
                if (e < total) A.nbr[o + e] = s[k];
            }
        }
        return;
    }
    if (lane == 0) {
        atomicAdd(&A.err[0], 1u);
        if (A.mode == 0) A.fb2[q] = -1.0;
        if (A.mode == 1) A.cnt[q] = 0;
    }
}

// One 64-lane workgroup per query slot (block b serves slot b, or qlist[b]);
// with nq set, the first *nq entries of qlist shared out over the grid, so
// a retry list whose length only the device knows needs no host round trip.
__global__ __launch_bounds__(64) void k_query(QArgs A)
{
    __shared__ QLds S;
    const int lane = threadIdx.x;
    if (A.nq) {
        const int nq = (int)*A.nq;
        for (int b = blockIdx.x; b < nq; b += gridDim.x) {
            query_one(A, A.qlist[b], S, lane);
            wave_sync();
        }
        return;
    }
    query_one(A, A.qlist ? A.qlist[blockIdx.x] : (int)blockIdx.x, S, lane);
}

// ---------------------------------------------- kNN by bucket groups
// Group of a bucket: every bucket within a group radius RL of the bucket's
// box.  k_knn_bucket, one wavefront per bucket, finds the group (a
// breadth-first descent of the radix tree, whose waits on the child records
// other wavefronts' selections fill) into LDS, then settles the bucket's
// queries against it, holding the group's bucket boxes and ranges in LDS.
// For each query the wavefront
// takes the group's buckets within a trial radius of it and reads their
// positions from the slot-ordered arrays (each bucket is 32 consecutive
// slots at most, one load per half-wavefront); squared distances, computed
// exactly as the reference computes them, stay in registers and ballot
// counts choose a squared radius T whose ball holds between K and CAP
// particles.  The ball is compacted and the K nearest selected from it by
// (d2, caller index).  Every particle within RL of a query is in the group,
// so the selection is exact.  Queries that cannot be settled this way go to
// a retry list served by k_query.
constexpr int KL_CM = 4096;            // particles per group
constexpr int KL_MB = 192;             // buckets per group
constexpr int KL_SB = 32;              // buckets of a query's own ball
constexpr int KL_NS = KL_SB / 2;       // candidates per lane
constexpr int KL_FC = 512;             // descent frontier per wavefront

struct KLArgs {
    QArgs Q;
    int nb;                            // buckets
    const int *bfirst, *bcnt;
    int *gnb;                          // buckets in the group, -1: too large
    double *grl2;                      // group radius^2
    int *retry;
    unsigned *nretry;
};

// Squared gap between two boxes, minimised over periodic images of a.
__device__ __forceinline__ double boxbox_d2(const double *a, const double *b, int per,
                                            const double *L)
{
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        double g = fmax(0.0, fmax(a[k] - b[3 + k], b[k] - a[3 + k]));
        if (per) {
            const double g1 = fmax(0.0, fmax(a[k] + L[k] - b[3 + k], b[k] - a[3 + k] - L[k]));
            const double g2 = fmax(0.0, fmax(a[k] - L[k] - b[3 + k], b[k] - a[3 + k] + L[k]));
            g = fmin(g, fmin(g1, g2));
        }
        d2 += g * g;
    }
    return d2;
}

// One wavefront per bucket: group radius from the fBall2 guesses of the
// bucket's particles (else from the bucket's density) and the group list.
// r-th smallest (0-based) of the values v[k] held by the elements
// lane + 64 k with ok[k] set, by a wavefront quickselect whose pivot is the
// lowest-numbered element still in range; *nlt and *neq count the values
// below and equal to it.  Returns ~0 if the range empties (never for
// 0 <= r < number of elements).
template <int KE>
__device__ uint64_t wave_select(const uint64_t (&v)[KE], const bool (&ok)[KE], int r, int *nlt,
                                int *neq)
{
    uint64_t lo = 0, hi = ~0ull;
    int base = 0;
    for (int it = 0; it < 64 * KE + 1; ++it) {
        uint64_t p = 0;
        bool found = false;
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            const unsigned long long m = __ballot(ok[k] && v[k] >= lo && v[k] <= hi);
            if (!found && m) {
                const int src = __ffsll((long long)m) - 1;
                const uint32_t plo = __builtin_amdgcn_readlane((uint32_t)v[k], src);
                const uint32_t phi = __builtin_amdgcn_readlane((uint32_t)(v[k] >> 32), src);
                p = ((uint64_t)phi << 32) | plo;
                found = true;
            }
        }
        if (!found) break;
        int cl = 0, ce = 0;
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            cl += __popcll(__ballot(ok[k] && v[k] >= lo && v[k] < p));
            ce += __popcll(__ballot(ok[k] && v[k] == p));
        }
        if (r < base + cl) {
            hi = p - 1;
        } else if (r < base + cl + ce) {
            *nlt = base + cl;
            *neq = ce;
            return p;
        } else {
            base += cl + ce;
            lo = p + 1;
        }
    }
    *nlt = *neq = 0;
    return ~0ull;
}

// The K nearest of a settled ball of cnt <= 64 KE particles held in LDS as
// (d2, slot): the K-th smallest d2 by quickselect on its bit pattern
// (non-negative doubles order as their bits), then every particle below it
// and, when equal distances straddle the K-th place, those of the smallest
// caller indices among the equal ones.  The list is written in ball order.
// Returns false (nothing written) if the selection does not come to K.
template <int KE>
__device__ bool knn_select(const QArgs &A, int q, int cnt, const double *bd, const int *bs,
                           double T, int lane)
{
    const int K = A.K;
    uint64_t v[KE];
    bool ok[KE];
    int s[KE];
    // 64 bins of d2 / T (every d2 of the ball is <= T): floor(64 d2 / T) is
    // monotone in d2, so each bin is a contiguous range of keys and the K-th
    // smallest lies in the first bin where the running count reaches K
    const double inv = T > 0 ? 64.0 / T : 0.0;
    int bin[KE];
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int e = lane + 64 * k;
        ok[k] = e < cnt;
        const double d = ok[k] ? bd[e] : 0.0;
        v[k] = ok[k] ? (uint64_t)__double_as_longlong(d) : ~0ull;
        s[k] = ok[k] ? bs[e] : -1;
        bin[k] = ok[k] ? min(63, (int)(d * inv)) : 64;
    }
    // the first bin b whose running count reaches K (bin 63 holds the rest,
    // and cnt >= K), by bisection on ballot counts of bin <= mid
    int blo = 0, bhi = 63;
    while (blo < bhi) {
        const int mid = (blo + bhi) >> 1;
        int c = 0;
#pragma unroll
        for (int k = 0; k < KE; ++k) c += __popcll(__ballot(bin[k] <= mid));
        if (c >= K) bhi = mid;
        else blo = mid + 1;
    }
    const int b = blo;
    int below = 0;                     // keys in the bins under b
#pragma unroll
    for (int k = 0; k < KE; ++k) below += __popcll(__ballot(bin[k] < b));
    bool okb[KE];
#pragma unroll
    for (int k = 0; k < KE; ++k) okb[k] = bin[k] == b;
    int nlt, neq;
    const uint64_t pK = wave_select<KE>(v, okb, K - 1 - below, &nlt, &neq);
    nlt += below;            // equal keys share a bin, so neq is whole
    bool take[KE];
    if (nlt + neq == K) {
#pragma unroll
        for (int k = 0; k < KE; ++k) take[k] = ok[k] && v[k] <= pK;
    } else {   // equal distances straddle the K-th place: caller index decides
        uint64_t c[KE];
        bool eq[KE];
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            eq[k] = ok[k] && v[k] == pK;
            c[k] = eq[k] ? (uint64_t)(uint32_t)A.order[s[k]] : ~0ull;
        }
        int l2, q2;
        const uint64_t cK = wave_select<KE>(c, eq, K - nlt - 1, &l2, &q2);
#pragma unroll
        for (int k = 0; k < KE; ++k) take[k] = ok[k] && (v[k] < pK || (eq[k] && c[k] <= cK));
        if (lane == 0) atomicAdd(&A.err[1], 1u);
    }
    unsigned long long m[KE];
    int tot = 0;
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        m[k] = __ballot(take[k]);
        tot += __popcll(m[k]);
    }
    if (tot != K) return false;   // not reached for a settled ball
    int base = 0;
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        if (take[k]) A.nbr[(size_t)q * K + base + __popcll(m[k] & ((1ull << lane) - 1))] = s[k];
        base += __popcll(m[k]);
    }
    if (lane == 0) A.fb2[q] = __longlong_as_double((long long)pK);
    return true;
}

// Squared distance of a candidate from a query, as the reference
// computes it; IMG = false when no periodic image can apply to the pair
// (then img_sep is a plain difference).
template <bool IMG>
__device__ __forceinline__ double pair_d2(double x, double y, double z, double px, double py,
                                          double pz, const double *L)
{
    if (IMG)
        return sq_dist(img_sep(x, px, 1, L[0]), img_sep(y, py, 1, L[1]), img_sep(z, pz, 1, L[2]));
    return sq_dist(x - px, y - py, z - pz);
}

// LDS state of a kNN group (one wavefront).
struct KLGroup {
    const float (*lbx)[6];             // bucket boxes (FP32, outwards)
    const int *lfst, *lcnt;            // bucket first slot, count
    int nb;
};

// Settles one query over the nsel <= KL_SB buckets listed in sel (first
// slot, count; two buckets per 64 lanes): finds T <= Tsel with K <= |{d2 <= T}| <= CAP and
// compacts that ball into (bd, bs).  Returns 0 when settled (*cnt = ball
// size), 1 when fewer than K lie within Tsel (*cnt = their number), 2 when
// no such T was found.
template <bool IMG>
__device__ int knn_settle(const QArgs &A, const KLGroup &G, const int2 *sel, int nsel,
                          double x, double y, double z, double Tsel, double *bd, int *bs,
                          int lane, int *cnt_out, double *T_out, unsigned &ncand)
{
    const int K = A.K;
    const int ns = (nsel + 1) >> 1;    // slots in use (wave-uniform)
    int ci[KL_NS];
#pragma unroll
    for (int k = 0; k < KL_NS; ++k) {
        ci[k] = -1;
        const int e = 2 * k + (lane >> 5), o = lane & 31;
        if (k < ns && e < nsel) {
            const int2 fc = sel[e];
            if (o < fc.y) ci[k] = fc.x + o;
        }
    }
    double dd[KL_NS];
#pragma unroll
    for (int k = 0; k < KL_NS; ++k) {
        dd[k] = DBL_MAX;
        if (k < ns && ci[k] >= 0) {
            const int c = ci[k];
            dd[k] = pair_d2<IMG>(x, y, z, A.sx[c], A.sy[c], A.sz[c], A.L);
            ++ncand;                   // per lane; summed over the wavefront at the end
        }
    }
    double T = Tsel;
    int cnt = 0;
    bool ok = false;
    for (int it = 0; it < 16; ++it) {
        cnt = 0;
#pragma unroll
        for (int k = 0; k < KL_NS; ++k)
            if (k < ns) cnt += __popcll(__ballot(dd[k] <= T));
        if (cnt >= K && cnt <= CAP) {
            ok = true;
            break;
        }
        if (cnt < K) {
            if (T >= Tsel) {
                *cnt_out = cnt;
                return 1;
            }
            const double f = cnt > 0 ? cbrt((1.5 * K / cnt) * (1.5 * K / cnt)) : 4.0;
            T = fmin(Tsel, T * fmin(fmax(f, 1.2), 16.0));
        } else {
            const double f = cbrt((1.25 * K / cnt) * (1.25 * K / cnt));
            T *= fmin(fmax(f, 0.05), 0.95);
        }
    }
    if (!ok) return 2;
    int base = 0;
#pragma unroll
    for (int k = 0; k < KL_NS; ++k) {
        if (k < ns) {
            const bool in = dd[k] <= T;
            const unsigned long long m = __ballot(in);
            if (in) {
                const int p = base + __popcll(m & ((1ull << lane) - 1));
                bd[p] = dd[k];
                bs[p] = ci[k];
            }
            base += __popcll(m);
        }
    }
    *cnt_out = cnt;
    *T_out = T;
    return 0;
}

// One query: choose the group's buckets within Tsel of it, settle, and
// widen Tsel (up to the group radius) while fewer than K lie within it, or
// narrow it while the ball meets more than KL_SB buckets.  Returns false if
// the query must be retried by k_query.
template <bool IMG>
__device__ bool knn_query(const QArgs &A, const KLGroup &G, int2 *sel, int q, double x,
                          double y, double z, double Tsel, double RL2, double *bd, int *bs,
                          int lane, unsigned &npass, unsigned &ncand)
{
    const int K = A.K;
    for (int pass = 0; pass < 8; ++pass) {
        npass += pass > 0;
        const double Tb = Tsel * (1.0 + 1e-12);
        int nsel = 0;
        for (int kb = 0; kb < G.nb; kb += 64) {
            const int k = kb + lane;
            bool in = false;
            if (k < G.nb) {
                const double b[6] = {G.lbx[k][0], G.lbx[k][1], G.lbx[k][2],
                                     G.lbx[k][3], G.lbx[k][4], G.lbx[k][5]};
                // without images for the group the plain box distance is exact
                in = box_d2(b, x, y, z, IMG ? A.per : 0, A.L) <= Tb;
            }
            const unsigned long long m = __ballot(in);
            const int p = nsel + __popcll(m & ((1ull << lane) - 1));
            if (in && p < KL_SB) sel[p] = make_int2(G.lfst[k], G.lcnt[k]);
            nsel += __popcll(m);
        }
        wave_sync();
        if (nsel > KL_SB) {            // too many buckets: a smaller ball first
            Tsel *= 0.5;
            continue;
        }
        int cnt = 0;
        double Tball = 0;
        const int st = knn_settle<IMG>(A, G, sel, nsel, x, y, z, Tsel, bd, bs, lane, &cnt, &Tball,
                                       ncand);
        if (st == 1 && Tsel < RL2) {   // widen the ball towards the group radius
            const double f = cnt > 0 ? cbrt((1.5 * K / cnt) * (1.5 * K / cnt)) : 4.0;
            Tsel = fmin(RL2, Tsel * fmin(fmax(f, 1.2), 16.0));
            wave_sync();
            continue;
        }
        if (st != 0) {
            if (lane == 0) atomicAdd(&A.err[st == 1 ? 3 : 4], 1u);
            return false;
        }
        wave_sync();
        const bool done = cnt <= 128 ? knn_select<2>(A, q, cnt, bd, bs, Tball, lane)
                                     : knn_select<4>(A, q, cnt, bd, bs, Tball, lane);
        if (!done && lane == 0) atomicAdd(&A.err[4], 1u);
        return done;
    }
    if (lane == 0) atomicAdd(&A.err[4], 1u);
    return false;
}

__device__ __forceinline__ double readlane_d(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// One wavefront per bucket.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_knn_bucket(KLArgs KA)
{
    const QArgs &A = KA.Q;
    __shared__ float lbx[KL_MB][6];
    __shared__ int lfst[KL_MB], lcnt[KL_MB];
    // the descent's frontier, then (once the group is found) each query's
    // buckets (first slot, count) and ball
    constexpr int QB_ = KL_SB * 8 + CAP * 12, FB_ = 2 * KL_FC * 4;
    __shared__ __align__(16) unsigned char ubuf[QB_ > FB_ ? QB_ : FB_];
    int *fr = reinterpret_cast<int *>(ubuf);
    int2 *sel = reinterpret_cast<int2 *>(ubuf);
    double *bd = reinterpret_cast<double *>(ubuf + KL_SB * 8);
    int *bs = reinterpret_cast<int *>(ubuf + KL_SB * 8 + CAP * 8);
    const int lane = threadIdx.x;
    const int g = xcd_block(blockIdx.x, gridDim.x);
    const int s0 = KA.bfirst[g], nq = KA.bcnt[g];
    const int K = A.K;
    // the group radius: from the queries' fBall2 hints, else the bucket's
    // density estimate
    double qg = 0, RL2;
    int nb;
    {
        double gmax = 0.0;
        bool none = false;
        if (lane < nq) {
            qg = A.guess ? A.guess[s0 + lane] : 0.0;
            none = !(qg > 0);
            gmax = none ? 0.0 : qg;
        }
        for (int o = 32; o > 0; o >>= 1) gmax = fmax(gmax, __shfl_xor(gmax, o));
        const unsigned long long nh = __ballot(none);
        const double R2 = nh == 0 ? gmax * A.slack * 1.25   // 1.44 fBall2 at the default slack
                                  : cold_r2(A.bden[g], K);
        const double *lb = A.bbox + 6 * (size_t)g;
        const double bx[6] = {lb[0], lb[1], lb[2], lb[3], lb[4], lb[5]};
        const double RL2b = R2 * (1.0 + 1e-12);
        int tot = 0;
        // the group's buckets (bucket index, count) into lfst, lcnt
        int nbg = wave_bfs(
            A, [&](const double *b) { return boxbox_d2(b, bx, A.per, A.L) <= RL2b; }, fr, KL_FC,
            lfst, lcnt, KL_MB, lane, &tot);
        if (nbg == 0 || tot > KL_CM) nbg = -1;   // none only for a NaN radius
        if (lane == 0) {   // for the retries
            KA.gnb[g] = nbg;
            KA.grl2[g] = R2;
            if (nbg >= 0) atomicMax(&A.err[5], (unsigned)nbg);
        }
        nb = nbg;
        RL2 = R2;
    }
    if (nb < 0) {          // group too large: every query retries
        if (lane == 0) atomicAdd(&A.err[2], 1u);
        if (lane < nq && A.qflag[s0 + lane]) {
            const unsigned p = atomicAdd(KA.nretry, 1u);
            KA.retry[p] = s0 + lane;
        }
        return;
    }
    // the bucket's queries, one per lane, broadcast as each is served
    double qx = 0, qy = 0, qz = 0;
    int qf = 0;
    if (lane < nq) {
        qx = A.sx[s0 + lane];
        qy = A.sy[s0 + lane];
        qz = A.sz[s0 + lane];
        qf = A.qflag[s0 + lane] != 0;
    }
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int k = lane; k < nb; k += 64) {
        const int b = lfst[k];          // bucket index, as the descent left it
        lfst[k] = A.bfirst[b];
        for (int d = 0; d < 3; ++d) {
            const float l = A.fbox[6 * (size_t)b + d], h = A.fbox[6 * (size_t)b + 3 + d];
            lbx[k][d] = l;
            lbx[k][3 + d] = h;
            lo[d] = fminf(lo[d], l);
            hi[d] = fmaxf(hi[d], h);
        }
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int d = 0; d < 3; ++d) {
            lo[d] = fminf(lo[d], __shfl_xor(lo[d], o));
            hi[d] = fmaxf(hi[d], __shfl_xor(hi[d], o));
        }
    // img_sep(p, q) takes an image only when fl(p - q) leaves [-L/2, L/2];
    // rounding is monotone, so no pair of a query p in the bucket box and a
    // group particle q (inside the outward boxes) does when the extreme
    // differences stay inside
    bool img = false;
    if (A.per) {
        const double *qb = A.bbox + 6 * (size_t)g;
        for (int d = 0; d < 3; ++d) {
            const double h = 0.5 * A.L[d];
            if (!(qb[3 + d] - (double)lo[d] <= h) || !(qb[d] - (double)hi[d] >= -h)) img = true;
        }
    }
    wave_sync();
    KLGroup G;
    G.lbx = lbx;
    G.lfst = lfst;
    G.lcnt = lcnt;
    G.nb = nb;
    unsigned npass = 0, ncand = 0;     // search work: passes (uniform), distances (per lane)
    for (int j = 0; j < nq; ++j) {
        if (!__builtin_amdgcn_readlane(qf, j)) continue;
        const int q = s0 + j;
        const double x = readlane_d(qx, j), y = readlane_d(qy, j), z = readlane_d(qz, j),
                     gq = readlane_d(qg, j);
        // first ball: slack * fBall2 from the hint, else 0.8 RL^2 (~1.6K expected)
        const double Tsel = gq > 0 ? fmin(gq * A.slack, RL2) : 0.8 * RL2;
        const bool done =
            img ? knn_query<true>(A, G, sel, q, x, y, z, Tsel, RL2, bd, bs, lane, npass, ncand)
                : knn_query<false>(A, G, sel, q, x, y, z, Tsel, RL2, bd, bs, lane, npass, ncand);
        if (!done && lane == 0) {
            const unsigned p = atomicAdd(KA.nretry, 1u);
            KA.retry[p] = q;
        }
        wave_sync();
    }
    for (int o = 32; o > 0; o >>= 1) ncand += __shfl_xor(ncand, o);
    if (lane == 0) {
        if (npass) atomicAdd(&A.err[6], npass);
        atomicAdd(reinterpret_cast<unsigned long long *>(A.err + 8), (unsigned long long)ncand);
    }
}

// -------------------------------------------------------------- kernels
template <typename T> __device__ __forceinline__ T tsqrt(T x);
template <> __device__ __forceinline__ float tsqrt<float>(float x) { return __builtin_sqrtf(x); }
template <> __device__ __forceinline__ double tsqrt<double>(double x) { return sqrt(x); }
template <typename T> __device__ __forceinline__ T trcp(T x);
template <> __device__ __forceinline__ float trcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double trcp<double>(double x) { return 1.0 / x; }

// M4 kernel W (smoothfcn.c:199-205) and dW/(r/h) (smoothfcn.c:206-213),
// in units where only the 1/pi normalisation is left to the caller.
template <typename T> __device__ __forceinline__ T m4_w(T r2)
{
    T ak = T(2) - tsqrt<T>(r2);
    return r2 < T(1) ? (T(1) - T(0.75) * ak * r2) : T(0.25) * ak * ak * ak;
}
template <typename T> __device__ __forceinline__ T m4_dw(T r2)
{
    const T r = tsqrt<T>(r2);
    if (r2 < T(1)) return T(-3) + T(2.25) * r;
    const T t = T(2) - r;
    return T(-0.75) * t * t * trcp<T>(r);
}

// Per-slot record DenDVDX reads for each neighbour, so that one neighbour
// costs a handful of 16-byte loads from one place instead of a load from each
// of a dozen arrays: the position and fBall2 in FP64 (separations are formed
// in FP64), the pair fields in T, and the flags.
template <typename T> struct DRec;
template <> struct alignas(16) DRec<double> {
    double x, y, z, fb2;
    double m, vx, vy, vz;            // fMass, vPred
    double u;                        // uPred
    uint32_t qf, act;                // smoothed (has a list), TYPE_ACTIVE
};
template <> struct alignas(16) DRec<float> {
    double x, y, z, fb2;
    float vx, vy, vz, m;
    float u;
    uint32_t qf, act, pad;
};
static_assert(sizeof(DRec<double>) == 80 && sizeof(DRec<float>) == 64, "DRec layout");

template <typename T>
__global__ void k_dpack(int n, const int *__restrict__ order, const double *__restrict__ sx,
                        const double *__restrict__ sy, const double *__restrict__ sz,
                        const double *__restrict__ fb2, const uint32_t *__restrict__ qf,
                        const uint32_t *__restrict__ act, const double *__restrict__ m,
                        const double *__restrict__ vp, const double *__restrict__ u,
                        DRec<T> *__restrict__ rec)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int i = order[s];
    DRec<T> r;
    r.x = sx[s];
    r.y = sy[s];
    r.z = sz[s];
    r.fb2 = fb2[s];
    r.m = (T)m[i];
    r.vx = (T)vp[3 * (size_t)i];
    r.vy = (T)vp[3 * (size_t)i + 1];
    r.vz = (T)vp[3 * (size_t)i + 2];
    r.u = (T)u[i];
    r.qf = qf[s];
    r.act = act[s];
    if constexpr (sizeof(T) == 4) r.pad = 0;
    rec[s] = r;
}

struct DArgs {
    int n, op;                  // op: GG_SMX_DENSITY or GG_SMX_DENDVDX
    int per;
    double L[3];
    const int64_t *off;                      // CSR (slot order)
    const int *idx;
    const void *rec;                         // DRec<T> per slot
    double H, a, gamma;
    int bCannonical, iViscosityLimiter, bConstantDiffusion;
    double *rho, *divv, *curl, *dvds, *divvt, *bal, *diff;   // slot order
    uint32_t *nbract;                        // slot order: 32 = neighbour of an active particle
    int *ecnt, *eidx;                        // force-pass extras, ecap per particle
    int ecap;
    unsigned *err;                           // err[7]: an extras list overflowed
};

// Lanes per particle of the SPH pair kernels: 4 (the default), 8 or 16,
// chosen per structure by gg_tune_sph_lanes.
template <int L> struct Lanes {
    static_assert(L == 4 || L == 8 || L == 16, "4, 8 or 16 lanes per particle");
    static constexpr int shift = L == 4 ? 2 : (L == 8 ? 3 : 4);
};

// DenDVDX (smoothfcn.c:3390-3627): L lanes per particle over its list,
// pair arithmetic in T, sums in FP64.
template <typename T, int L>
__global__ __launch_bounds__(256) void k_dendvdx(DArgs A)
{
    const int g = (xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x) >> Lanes<L>::shift;
    const int sub = threadIdx.x & (L - 1);
    if (g >= A.n) return;
    const int i = g;
    const DRec<T> *__restrict__ rec = static_cast<const DRec<T> *>(A.rec);
    // the list's range and first neighbour issued with the particle's record
    // (a particle that is not smoothed has an empty list, or a fixed-K one)
    const int64_t o = A.off[i], e1 = A.off[i + 1];
    int qn = o + sub < e1 ? A.idx[o + sub] : 0;
    const DRec<T> I = rec[i];
    if (!I.qf) return;
    const double fb2 = I.fb2;
    const double ih2d = 4.0 / fb2;
    const T vFac = (T)(A.bCannonical ? 1. / (A.a * A.a) : 1.0);
    const double x = I.x, y = I.y, z = I.z;
    const T vxi = I.vx, vyi = I.vy, vzi = I.vz;
    const bool pact = I.act & 1u;
    double rho = 0, xx = 0, xy = 0, xz = 0, yx = 0, yy = 0, yz = 0, zx = 0, zy = 0, zz = 0;
    double dn = 0, gx = 0, gy = 0, gz = 0;
    unsigned qact = 0;
    auto pair = [&](int q, const DRec<T> &Q) {
        const double dxd = img_sep(x, Q.x, A.per, A.L[0]);
        const double dyd = img_sep(y, Q.y, A.per, A.L[1]);
        const double dzd = img_sep(z, Q.z, A.per, A.L[2]);
        const T dx = (T)dxd, dy = (T)dyd, dz = (T)dzd;
        const T r2 = (T)(sq_dist(dxd, dyd, dzd) * ih2d);
        const T mq = Q.m, vxq = Q.vx, vyq = Q.vy, vzq = Q.vz;
        if (pact) A.nbract[q] = 32u;   // every writer stores the same flag
        qact |= Q.act;
        {   // extras of the force pass: q's own list misses i when i lies
            // outside q's ball, seen from q (or q has no list)
            bool extra = !Q.qf;
            if (!extra) {
                double d2q = sq_dist(dxd, dyd, dzd);
                if (A.per && (fabs(x - Q.x) > 0.5 * A.L[0] || fabs(y - Q.y) > 0.5 * A.L[1] ||
                              fabs(z - Q.z) > 0.5 * A.L[2]))
                    d2q = sq_dist(img_sep(Q.x, x, 1, A.L[0]), img_sep(Q.y, y, 1, A.L[1]),
                                  img_sep(Q.z, z, 1, A.L[2]));
                extra = d2q > Q.fb2;
            }
            if (extra) {
                const int k = atomicAdd(&A.ecnt[q], 1);
                if (k < A.ecap) A.eidx[(size_t)q * A.ecap + k] = i;
                else atomicOr(&A.err[7], 1u);
            }
        }
        rho += (double)(m4_w<T>(r2) * mq);
        if (A.op == GG_SMX_DENDVDX) {
            const T rs1 = m4_dw<T>(r2) * mq;
            const T dvx = (vxq - vxi) * vFac;
            const T dvy = (vyq - vyi) * vFac;
            const T dvz = (vzq - vzi) * vFac;
            const T uq = Q.u;
            xx += (double)(dvx * dx * rs1); xy += (double)(dvx * dy * rs1); xz += (double)(dvx * dz * rs1);
            yx += (double)(dvy * dx * rs1); yy += (double)(dvy * dy * rs1); yz += (double)(dvy * dz * rs1);
            zx += (double)(dvz * dx * rs1); zy += (double)(dvz * dy * rs1); zz += (double)(dvz * dz * rs1);
            dn += (double)((dx * dx + dy * dy + dz * dz) * rs1);
            gx += (double)(uq * dx * rs1); gy += (double)(uq * dy * rs1); gz += (double)(uq * dz * rs1);
        }
    };
    // the next neighbour's record is loaded while this one is used, and the
    // index after it with it, so no step waits on a load issued in it
    DRec<T> Qn;
    if (o + sub < e1) Qn = rec[qn];
    int qnn = o + sub + L < e1 ? A.idx[o + sub + L] : 0;
    for (int64_t e = o + sub; e < e1; e += L) {
        const int q = qn;
        const DRec<T> Q = Qn;
        qn = qnn;
        if (e + L < e1) Qn = rec[qn];
        if (e + 2 * L < e1) qnn = A.idx[e + 2 * L];
        pair(q, Q);
    }
    for (int w = 1; w < L; w <<= 1) {
        rho += __shfl_xor(rho, w);
        qact |= __shfl_xor(qact, w);
        if (A.op == GG_SMX_DENDVDX) {
            xx += __shfl_xor(xx, w); xy += __shfl_xor(xy, w); xz += __shfl_xor(xz, w);
            yx += __shfl_xor(yx, w); yy += __shfl_xor(yy, w); yz += __shfl_xor(yz, w);
            zx += __shfl_xor(zx, w); zy += __shfl_xor(zy, w); zz += __shfl_xor(zz, w);
            dn += __shfl_xor(dn, w);
            gx += __shfl_xor(gx, w); gy += __shfl_xor(gy, w); gz += __shfl_xor(gz, w);
        }
    }
    if (sub != 0) return;
    const double ih = sqrt(ih2d);
    const double fNorm = M_1_PI * ih2d * ih;
    if (qact & 1u) A.nbract[i] = 32u;
    A.rho[i] = rho * fNorm;
    if (A.op != GG_SMX_DENDVDX) return;
    const double trace = xx + yy + zz;
    const double fN1 = (dn != 0 ? 3 / fabs(dn) : 0);
    const double dv = fN1 * trace + 3 * A.H;
    A.divvt[i] = fN1 * trace;
    A.divv[i] = dv;
    const double c0 = fN1 * (zy - yz), c1 = fN1 * (xz - zx), c2 = fN1 * (yx - xy);
    A.curl[3 * (size_t)i] = c0;
    A.curl[3 * (size_t)i + 1] = c1;
    A.curl[3 * (size_t)i + 2] = c2;
    const double Hc = (fN1 != 0 ? A.H / fN1 : 0);
    double gn = gx * gx + gy * gy + gz * gz;
    if (gn > 0) gn = 1 / sqrt(gn);
    gx *= gn; gy *= gn; gz *= gn;
    const double dvdr = (((xx + Hc) * gx + xy * gy + xz * gz) * gx
                         + (yx * gx + (yy + Hc) * gy + yz * gz) * gy
                         + (zx * gx + zy * gy + (zz + Hc) * gz) * gz) * fN1;
    const double ds = (dv < 0 ? 1.5 * (dvdr - (1. / 3.) * dv) : dvdr);
    A.dvds[i] = ds;
    const double cm = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    double bs;
    if (A.iViscosityLimiter == 0) bs = 1;
    else if (A.iViscosityLimiter == 1) bs = (dv != 0.0) ? fabs(dv) / (fabs(dv) + cm) : 0;
    else bs = (ds != 0) ? fabs(ds) / (fabs(ds) + cm) : 0.01;   // ALPHAMIN, pkd.h:41
    A.bal[i] = bs;
    const double ot = (1. / 3.) * trace;
    const double sxx = xx - ot, syy = yy - ot, szz = zz - ot;
    const double sxy = 0.5 * (xy + yx), sxz = 0.5 * (xz + zx), syz = 0.5 * (yz + zy);
    const double S2 = fN1 * fN1 * (sxx * sxx + syy * syy + szz * szz
                                   + 2 * (sxy * sxy + sxz * sxz + syz * syz));
    A.diff[i] = A.bConstantDiffusion ? 1 : 0.25 * fb2 * sqrt(2 * S2);
}

struct DSArgs {
    int n, sym, per;
    double L[3];
    const double *sx, *sy, *sz, *fb2;        // slot order
    const int64_t *off, *roff;               // own lists; lists of the owners holding i
    const int *idx, *ridx;
    const uint32_t *qflag;
    const double *m;                         // slot order
    const double *rho0;                      // slot order: what unsmoothed particles hold
    double *rho;                             // slot order
};

// Density (smoothfcn.c:239-256) when sym == 0: the gather over i's own list,
// bit for bit the density DenDVDX computes.  DensitySym (smoothfcn.c:258-274)
// when sym == 1, in gather form: i sums the half-normalised kernel of its own
// list with h_i and m_q, then over the owners q whose lists hold i (sorted by
// slot, so the order is fixed) the half-normalised kernel with h_q and m_q,
// on top of 0 (smooth-active i, initDensity smoothfcn.c:229-232) or of the
// density i held.  L lanes per particle, as DenDVDX; pair arithmetic in T,
// sums in FP64.
template <typename T, int L>
__global__ __launch_bounds__(256) void k_density(DSArgs A)
{
    const int i = (xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x) >> Lanes<L>::shift;
    const int sub = threadIdx.x & (L - 1);
    if (i >= A.n) return;
    const bool own = A.qflag[i] != 0;
    if (!A.sym && !own) return;
    const double x = A.sx[i], y = A.sy[i], z = A.sz[i];
    double so = 0, sr = 0, ih2 = 0;
    if (own) {
        ih2 = 4.0 / A.fb2[i];
        for (int64_t e = A.off[i] + sub; e < A.off[i + 1]; e += L) {
            const int q = A.idx[e];
            const double d2 = sq_dist(img_sep(x, A.sx[q], A.per, A.L[0]),
                                      img_sep(y, A.sy[q], A.per, A.L[1]),
                                      img_sep(z, A.sz[q], A.per, A.L[2]));
            so += (double)(m4_w<T>((T)(d2 * ih2)) * (T)A.m[q]);
        }
    }
    if (A.sym) {
        for (int64_t e = A.roff[i] + sub; e < A.roff[i + 1]; e += L) {
            const int q = A.ridx[e];
            const double d2 = sq_dist(img_sep(x, A.sx[q], A.per, A.L[0]),
                                      img_sep(y, A.sy[q], A.per, A.L[1]),
                                      img_sep(z, A.sz[q], A.per, A.L[2]));
            const double ih2q = 4.0 / A.fb2[q];
            const double fq = 0.5 * M_1_PI * sqrt(ih2q) * ih2q;
            sr += fq * (double)(m4_w<T>((T)(d2 * ih2q)) * (T)A.m[q]);
        }
    }
    for (int w = 1; w < L; w <<= 1) {
        so += __shfl_xor(so, w);
        sr += __shfl_xor(sr, w);
    }
    if (sub != 0) return;
    const double fN = M_1_PI * ih2 * sqrt(ih2);
    if (!A.sym) A.rho[i] = so * fN;
    else A.rho[i] = (own ? 0.0 : A.rho0[i]) + 0.5 * fN * so + sr;
}

// Pressure of the adiabatic gas (pkdGasPressure, pkd.c:5750-5785), caller
// order, particles with the selection mask only.
__global__ void k_gas_pressure(int n, const uint32_t *__restrict__ bits, uint32_t mask,
                               const double *__restrict__ upred, const double *__restrict__ rho,
                               const double *__restrict__ fb2, const double *__restrict__ udp,
                               const double *__restrict__ uda, const double *__restrict__ udd,
                               double gamma, double dtFacCourant, double dResolveJeans,
                               double *__restrict__ pr2, double *__restrict__ c,
                               double *__restrict__ dt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (bits[i] & mask) != mask) return;
    const double prg = (gamma - 1) * upred[i];
    const double cg = sqrt(gamma * prg);
    const double floorJ = 0.25 * fb2[i] * dResolveJeans * rho[i];
    double pr = prg;
    const double prj = (pr < floorJ ? floorJ - pr : 0);
    pr += prj;
    pr2[i] = pr / rho[i];
    const double ci = sqrt(cg * cg + 2.0 * prj);   // GAMMA_JEANS = 2
    c[i] = ci;
    const double udot = udp[i] + uda[i] + udd[i];  // UDOT_HYDRO, pkd.h:65
    double dti;
    if (udot > 0) dti = dtFacCourant * sqrt(fb2[i] * 0.25 / (4 * (ci * ci + (5. / 3.) * udot * dt[i])));
    else dti = dtFacCourant * sqrt(fb2[i] * 0.25) / (2 * ci);
    if (dti < dt[i]) dt[i] = dti;
}

// Apply the hydro results to the caller-order fields: for smooth-active
// particles initSphPressureTermsParticle (smoothfcn.c:2204-2221) resets
// mumax, dtNew and the uDot terms of ACTIVE ones; a is accumulated.
__global__ void k_hydro_scatter(int n, const int *__restrict__ order,
                                const uint32_t *__restrict__ qflag, const double *__restrict__ res,
                                const uint32_t *__restrict__ bits, double *__restrict__ acc,
                                double *__restrict__ pdv, double *__restrict__ av,
                                double *__restrict__ dif, double *__restrict__ mumax,
                                double *__restrict__ dtnew, int touched_only)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int c = order[s];
    const double *o = res + 8 * (size_t)s;
    const bool active = bits[c] & 1u;
    const bool init = qflag[s] && active;
    double mu = init ? 0.0 : mumax[c];
    double dn = init ? GG_FLT_MAX : dtnew[c];
    mumax[c] = fmax(mu, o[6]);
    dtnew[c] = fmin(dn, o[7]);
    (void)touched_only;
    if (active) {
        acc[3 * (size_t)c] += o[0];
        acc[3 * (size_t)c + 1] += o[1];
        acc[3 * (size_t)c + 2] += o[2];
        pdv[c] = (init ? 0.0 : pdv[c]) + o[3];
        av[c] = (init ? 0.0 : av[c]) + o[4];
        dif[c] = (init ? 0.0 : dif[c]) + o[5];
    }
}

// ------------------------------------------- hydro, own list + extras
// Per-particle record of the force pass, slot order: the position and
// fBall2 in FP64, then in the pair precision velocity, mass, density,
// P/rho^2, sound speed, Balsara switch, h = 2h/2, the kernel normalisation
// 0.5/pi * ih2^2 / h (SphPressureTermsSym, smoothfcn.c:2302-2305), 4/fBall2,
// Courant dt, uPred, diff, activity and whether it has a list.  One record
// per neighbour: a few 16-byte loads from one place.
template <typename T> struct alignas(16) HRec {
    double x, y, z, fb2;
    T vx, vy, vz, m;
    T rho, pr2, c, bal;
    T ph, fn1, ih2, dt;
    T u, diff, act, qf;
};

template <typename T>
__global__ void k_hpack(int n, const int *__restrict__ order, const double *__restrict__ sx,
                        const double *__restrict__ sy, const double *__restrict__ sz,
                        const uint32_t *__restrict__ qflag, const double *__restrict__ fb2,
                        const double *__restrict__ m, const double *__restrict__ vp,
                        const double *__restrict__ u, const double *__restrict__ rho,
                        const double *__restrict__ pr2, const double *__restrict__ c,
                        const double *__restrict__ bal, const double *__restrict__ diff,
                        const double *__restrict__ dt, const uint32_t *__restrict__ bits,
                        HRec<T> *__restrict__ rec)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int i = order[s];
    const double f = fb2[s];
    const double ph = sqrt(0.25 * f), ih2 = 4.0 / f;
    const double fNorm = 0.5 * M_1_PI * ih2 / ph;
    HRec<T> r;
    r.x = sx[s];
    r.y = sy[s];
    r.z = sz[s];
    r.fb2 = f;
    r.vx = (T)vp[3 * (size_t)i];
    r.vy = (T)vp[3 * (size_t)i + 1];
    r.vz = (T)vp[3 * (size_t)i + 2];
    r.m = (T)m[i];
    r.rho = (T)rho[i];
    r.pr2 = (T)pr2[i];
    r.c = (T)c[i];
    r.bal = (T)bal[i];
    r.ph = (T)ph;
    r.fn1 = (T)(fNorm * ih2);
    r.ih2 = (T)ih2;
    r.dt = (T)dt[i];
    r.u = (T)u[i];
    r.diff = (T)diff[i];
    r.act = (T)((bits[i] & 1u) ? 1 : 0);
    r.qf = (T)(qflag[s] ? 1 : 0);
    rec[s] = r;
}

template <typename T> struct HPar {
    T a, H, alpha, beta, dtFacCourant, vFac;
    double dtFacDiffusion, kDiff;
};

struct HAcc {
    double ax = 0, ay = 0, az = 0, pdv = 0, av = 0, dif = 0, mu = 0;
    double dtn = GG_FLT_MAX;
};

// One pair of SphPressureTermsSym (SphPressureTerms.h:97-152) as it affects
// particle "self": role 0, self is p (the list owner, whose smoothing length
// is used); role 1, self is q.  dx = x_p - x_q (minimum image from p).
template <typename T>
__device__ __forceinline__ void hpair(const HRec<T> &P, const HRec<T> &Q, T dx, T dy, T dz,
                                      T d2, int role, const HPar<T> &hp, HAcc &acc)
{
    const bool pa = P.act > T(0), qa = Q.act > T(0);
    if (!pa && !qa) return;
    const bool sa = role ? qa : pa;
    const T r2 = d2 * P.ih2;
    const T rs1 = m4_dw<T>(r2) * P.fn1;
    const T rp = rs1 * P.m, rq = rs1 * Q.m;
    const T dvdotdr = hp.vFac * ((P.vx - Q.vx) * dx + (P.vy - Q.vy) * dy + (P.vz - Q.vz) * dz)
                      + d2 * hp.H;
    const T rs = role ? rp : rq;
    // PRES_PDV(a,b) = a, PRES_ACC(a,b) = a+b (smoothfcn.c:20-23)
    if (sa) acc.pdv += (double)(rs * (role ? Q.pr2 : P.pr2) * dvdotdr);
    T Acc = P.pr2 + Q.pr2;
    T dtv;
    if (dvdotdr >= T(0)) {
        dtv = hp.dtFacCourant * P.ph * trcp<T>(T(2) * (P.c > Q.c ? P.c : Q.c));
    } else {
        const T absmu = -dvdotdr * hp.a * trcp<T>(tsqrt<T>(d2));
        acc.mu = fmax(acc.mu, (double)absmu);
        T visc = hp.alpha * (P.c + Q.c) + hp.beta * T(1.5) * absmu;
        dtv = hp.dtFacCourant * P.ph * trcp<T>(T(0.625) * (P.c + Q.c) + T(0.375) * visc);
        visc = (T(0.5) * (P.bal + Q.bal)) * visc * absmu * trcp<T>(P.rho + Q.rho);
        if (sa) acc.av += (double)(rs * (T(0.5) * visc) * dvdotdr);
        Acc += visc;
    }
    if (sa) {
        Acc *= rs * hp.a;
        const T f = role ? Acc : -Acc;     // a_p -= Acc dx, a_q += Acc dx
        acc.ax += (double)(f * dx);
        acc.ay += (double)(f * dy);
        acc.az += (double)(f * dz);
    }
    if (hp.kDiff > 0) {   // thermal diffusion (SphPressureTerms.h:32-70)
        const double dsum = (double)P.diff + (double)Q.diff;
        if (dsum != 0) {
            const double diffTh = 2 * hp.kDiff * dsum / ((double)P.rho + (double)Q.rho);
            const double dtd = hp.dtFacDiffusion * (double)P.ph * (double)P.ph /
                               (diffTh * (double)P.rho);
            if (dtd < (double)dtv) dtv = (T)dtd;
            const double du = diffTh * ((double)P.u - (double)Q.u);
            if (sa) acc.dif += role ? -du * (double)rp : du * (double)rq;
        }
    }
    // SETDTNEW_PQ (SphPressureTerms.h:97-100)
    acc.dtn = fmin(acc.dtn, (double)dtv);
    acc.dtn = fmin(acc.dtn, 4.0 * (double)(role ? P.dt : Q.dt));
}

struct H2Args {
    int n, per;
    double L[3];
    const int64_t *off, *eoff;         // eoff: CSR extras, or nullptr for
    const int *idx, *eidx;             // ecap-strided lists of ecnt entries
    const int *ecnt;
    int ecap;
    double a, H, alpha, beta, dtFacCourant, dtFacDiffusion, kDiff;
    int bCannonical;
    double *out;     // per slot: ax ay az uDotPdV uDotAV uDotDiff mumax dtNew
};

// Gather form of the symmetric force pass, L lanes per particle i.  Over
// i's own list: the pair as p = i, and, when i lies in the neighbour's ball,
// the pair as q = i.  Over i's extras (particles whose balls hold i but which
// are not in i's list): the pair as q = i.
template <typename T, int L>
__global__ __launch_bounds__(256) void k_hydro2(H2Args A, const HRec<T> *__restrict__ rec)
{
    const int i = (xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x) >> Lanes<L>::shift;
    const int sub = threadIdx.x & (L - 1);
    if (i >= A.n) return;
    HPar<T> hp;
    hp.a = (T)A.a;
    hp.H = (T)A.H;
    hp.alpha = (T)A.alpha;
    hp.beta = (T)A.beta;
    hp.dtFacCourant = (T)A.dtFacCourant;
    hp.vFac = (T)(A.bCannonical ? 1. / (A.a * A.a) : 1.0);
    hp.dtFacDiffusion = A.dtFacDiffusion;
    hp.kDiff = A.kDiff;
    // the own list's range and first index issued with the particle's record
    const int64_t o = A.off[i], e1 = A.off[i + 1];
    int qn = o + sub < e1 ? A.idx[o + sub] : 0;
    const HRec<T> I = rec[i];
    const double xi = I.x, yi = I.y, zi = I.z;
    // the extras' range and first index, loaded ahead of the own list
    const int64_t x0 = A.eoff ? A.eoff[i] : (int64_t)i * A.ecap;
    const int64_t x1 = A.eoff ? A.eoff[i + 1] : x0 + A.ecnt[i];
    int pn = x0 + sub < x1 ? A.eidx[x0 + sub] : 0;
    HAcc acc;
    if (I.qf > T(0)) {
        for (int64_t e = o + sub; e < e1; e += L) {   // the next index in flight
            const int q = qn;
            if (e + L < e1) qn = A.idx[e + L];
            const HRec<T> Q = rec[q];
            const double xq = Q.x, yq = Q.y, zq = Q.z;
            const double dx = img_sep(xi, xq, A.per, A.L[0]);
            const double dy = img_sep(yi, yq, A.per, A.L[1]);
            const double dz = img_sep(zi, zq, A.per, A.L[2]);
            hpair<T>(I, Q, (T)dx, (T)dy, (T)dz, (T)sq_dist(dx, dy, dz), 0, hp, acc);
            if (Q.qf > T(0)) {
                const double ex = img_sep(xq, xi, A.per, A.L[0]);
                const double ey = img_sep(yq, yi, A.per, A.L[1]);
                const double ez = img_sep(zq, zi, A.per, A.L[2]);
                const double d2q = sq_dist(ex, ey, ez);
                if (d2q <= Q.fb2)
                    hpair<T>(Q, I, (T)ex, (T)ey, (T)ez, (T)d2q, 1, hp, acc);
            }
        }
    }
    for (int64_t e = x0 + sub; e < x1; e += L) {
        const int p = pn;
        if (e + L < x1) pn = A.eidx[e + L];
        const HRec<T> P = rec[p];
        const double ex = img_sep(P.x, xi, A.per, A.L[0]);
        const double ey = img_sep(P.y, yi, A.per, A.L[1]);
        const double ez = img_sep(P.z, zi, A.per, A.L[2]);
        hpair<T>(P, I, (T)ex, (T)ey, (T)ez, (T)sq_dist(ex, ey, ez), 1, hp, acc);
    }
    for (int w = 1; w < L; w <<= 1) {
        acc.ax += __shfl_xor(acc.ax, w);
        acc.ay += __shfl_xor(acc.ay, w);
        acc.az += __shfl_xor(acc.az, w);
        acc.pdv += __shfl_xor(acc.pdv, w);
        acc.av += __shfl_xor(acc.av, w);
        acc.dif += __shfl_xor(acc.dif, w);
        acc.mu = fmax(acc.mu, __shfl_xor(acc.mu, w));
        acc.dtn = fmin(acc.dtn, __shfl_xor(acc.dtn, w));
    }
    if (sub == 0) {
        double *o = A.out + 8 * (size_t)i;
        o[0] = acc.ax; o[1] = acc.ay; o[2] = acc.az; o[3] = acc.pdv; o[4] = acc.av;
        o[5] = acc.dif; o[6] = acc.mu; o[7] = acc.dtn;
    }
}

// Extras: q receives p when q is in p's list but p is not in q's ball (or q
// has no list).  16 lanes per list owner p.  mode 0 counts, mode 1 fills at
// eoff + cursor.
__global__ __launch_bounds__(256) void k_extra(int n, int mode, int per, double Lx, double Ly,
                                               double Lz, const uint32_t *__restrict__ qflag,
                                               const int64_t *__restrict__ off,
                                               const int *__restrict__ idx,
                                               const double *__restrict__ sx,
                                               const double *__restrict__ sy,
                                               const double *__restrict__ sz,
                                               const double *__restrict__ fb2,
                                               int *__restrict__ cnt,
                                               const int64_t *__restrict__ eoff,
                                               int *__restrict__ eidx)
{
    const int p = (blockIdx.x * 256 + threadIdx.x) >> 4;
    const int sub = threadIdx.x & 15;
    if (p >= n || !qflag[p]) return;
    const double xp = sx[p], yp = sy[p], zp = sz[p];
    for (int64_t e = off[p] + sub; e < off[p + 1]; e += 16) {
        const int q = idx[e];
        bool extra = !qflag[q];
        if (!extra) {
            const double ex = img_sep(sx[q], xp, per, Lx);
            const double ey = img_sep(sy[q], yp, per, Ly);
            const double ez = img_sep(sz[q], zp, per, Lz);
            extra = sq_dist(ex, ey, ez) > fb2[q];
        }
        if (!extra) continue;
        const int k = atomicAdd(&cnt[q], 1);
        if (mode) eidx[eoff[q] + k] = p;
    }
}

// Bitwise OR of the selection bits over all particles (is any gas flagged?).
__global__ void k_iota(int n, int *__restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// out[0] |= the bits of mask some particle has, out[1] |= those some lacks.
__global__ void k_any_bits(int n, const uint32_t *__restrict__ bits, uint32_t mask,
                           unsigned *__restrict__ out)
{
    __shared__ unsigned s[2];
    if (threadIdx.x < 2) s[threadIdx.x] = 0;
    __syncthreads();
    unsigned v = 0, w = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        v |= bits[i] & mask;
        w |= ~bits[i] & mask;
    }
    if (v) atomicOr(&s[0], v);
    if (w) atomicOr(&s[1], w);
    __syncthreads();
    if (threadIdx.x < 2 && s[threadIdx.x]) atomicOr(&out[threadIdx.x], s[threadIdx.x]);
}

// ------------------------------------------------------- list plumbing
__global__ void k_rev_count(int n, const uint32_t *__restrict__ qflag,
                            const int64_t *__restrict__ off, const int *__restrict__ idx,
                            int *__restrict__ rcnt)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !qflag[p]) return;
    for (int64_t e = off[p]; e < off[p + 1]; ++e) atomicAdd(&rcnt[idx[e]], 1);
}

__global__ void k_rev_fill(int n, const uint32_t *__restrict__ qflag,
                           const int64_t *__restrict__ off, const int *__restrict__ idx,
                           const int64_t *__restrict__ roff, int *__restrict__ rcur,
                           int *__restrict__ ridx)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !qflag[p]) return;
    for (int64_t e = off[p]; e < off[p + 1]; ++e) {
        const int q = idx[e];
        const int k = atomicAdd(&rcur[q], 1);
        ridx[roff[q] + k] = p;
    }
}

// Sort each reverse list so the summation order is deterministic.
// Each particle's extras list (ecap-strided, ecnt entries, in the order the
// density pass's atomics left them) sorted by slot in place, so the force
// pass sums in a fixed order; each block's total of entries goes to
// part[block] (256 threads).
__global__ __launch_bounds__(256) void k_rev_sort_fixed(int n, int ecap,
                                                        const int *__restrict__ ecnt,
                                                        int *__restrict__ eidx,
                                                        int64_t *__restrict__ part)
{
    __shared__ int64_t ws[4];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = q < n ? ecnt[q] : 0;
    int64_t t = m;
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
    if (q >= n || m <= 1) return;
    int *a = eidx + (size_t)q * ecap;
    if (m <= 16) {   // most lists: 16-byte loads, a bitonic network in registers
        int v[16];
        int4 *a4 = reinterpret_cast<int4 *>(a);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int4 t = 4 * c < m ? a4[c] : make_int4(0, 0, 0, 0);
            v[4 * c] = 4 * c < m ? t.x : INT_MAX;
            v[4 * c + 1] = 4 * c + 1 < m ? t.y : INT_MAX;
            v[4 * c + 2] = 4 * c + 2 < m ? t.z : INT_MAX;
            v[4 * c + 3] = 4 * c + 3 < m ? t.w : INT_MAX;
        }
#pragma unroll
        for (int k = 2; k <= 16; k <<= 1)
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int l = i ^ j;
                    if (l > i) {
                        const int lo = min(v[i], v[l]), hi = max(v[i], v[l]);
                        const bool up = (i & k) == 0;
                        v[i] = up ? lo : hi;
                        v[l] = up ? hi : lo;
                    }
                }
        // the padding sorts last: entries past m (within the list's
        // capacity) take INT_MAX, which no reader looks at
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (4 * c < m) a4[c] = make_int4(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]);
        return;
    }
    for (int i = 1; i < m; ++i) {
        const int v = a[i];
        int j = i - 1;
        while (j >= 0 && a[j] > v) {
            a[j + 1] = a[j];
            --j;
        }
        a[j + 1] = v;
    }
}

__global__ void k_rev_sort(int n, const int64_t *__restrict__ roff, int *__restrict__ ridx)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int *a = ridx + roff[q];
    const int m = (int)(roff[q + 1] - roff[q]);
    for (int i = 1; i < m; ++i) {
        const int v = a[i];
        int j = i - 1;
        while (j >= 0 && a[j] > v) {
            a[j + 1] = a[j];
            --j;
        }
        a[j + 1] = v;
    }
}

__global__ void k_fixed_off(int n, int K, int64_t *__restrict__ off)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) off[i] = (int64_t)K * i;
}

template <typename S, typename D>
__global__ void k_gather(int n, const int *__restrict__ order, const S *__restrict__ src,
                         int width, D *__restrict__ dst)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int c = order[s];
    for (int k = 0; k < width; ++k) dst[(size_t)width * s + k] = (D)src[(size_t)width * c + k];
}

__global__ void k_gather_comp(int n, const int *__restrict__ order, const double *__restrict__ src,
                              int width, int comp, double *__restrict__ dst)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    dst[s] = src[(size_t)width * order[s] + comp];
}

__global__ void k_count_diff(int n, const double *__restrict__ a, const double *__restrict__ b,
                             const uint32_t *__restrict__ qa, const uint32_t *__restrict__ qb,
                             unsigned *__restrict__ out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    if (qa[s] != qb[s] || (qa[s] && a[s] != b[s])) atomicAdd(out, 1u);
}

template <typename T>
__global__ void k_scatter(int n, const int *__restrict__ order, const uint32_t *__restrict__ qflag,
                          const T *__restrict__ src, int width, T *__restrict__ dst)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || (qflag && !qflag[s])) return;
    const int c = order[s];
    for (int k = 0; k < width; ++k) dst[(size_t)width * c + k] = src[(size_t)width * s + k];
}

// Up to eight FP64 fields from slot order into the caller's order in one
// launch (the smoothed slots only, when qflag is given).
struct ScatterSet {
    const double *src[8];
    double *dst[8];
    int width[8];
    int nf;
};

__global__ void k_scatter_set(int n, const int *__restrict__ order,
                              const uint32_t *__restrict__ qflag, ScatterSet F)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || (qflag && !qflag[s])) return;
    const int c = order[s];
    for (int f = 0; f < F.nf; ++f) {
        const int w = F.width[f];
        for (int k = 0; k < w; ++k) F.dst[f][(size_t)w * c + k] = F.src[f][(size_t)w * s + k];
    }
}

// Slot-order flags (query, TYPE_ACTIVE) and each block's number of queries
// in qcnt[block] (256 threads).
__global__ __launch_bounds__(256) void k_qflag(int n, const int *__restrict__ order,
                                               const uint32_t *__restrict__ bits, uint32_t mask,
                                               uint32_t *__restrict__ qf, uint32_t *__restrict__ act,
                                               uint32_t *__restrict__ nba, int *__restrict__ qcnt)
{
    __shared__ int wc[4];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    bool q = false;
    if (s < n) {
        const uint32_t b = bits[order[s]];
        q = (b & mask) == mask;
        qf[s] = q ? 1u : 0u;
        act[s] = b & 1u;
        if (nba) nba[s] = 0;
    }
    const int c = __popcll(__ballot(q));
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) qcnt[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ void k_flags_out(int n, const int *__restrict__ order, const uint32_t *__restrict__ qf,
                            const uint32_t *__restrict__ nba, uint32_t *__restrict__ bits)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int c = order[s];
    uint32_t b = bits[c];   // the operators set flags, never clear them
    if (qf[s]) b |= GG_TYPE_SMOOTHDONE;
    if (nba && (nba[s] & 32u)) b |= GG_TYPE_NbrOfACTIVE;
    bits[c] = b;
}


// The FP32 form's j-side record of the density pass, slot order, gathered
// from the store: vPred, fMass (one 16-byte load per pair) and uPred.
// ------------------------------------------------------------ host side
static inline int nblk(int n, int b) { return (n + b - 1) / b; }

static int gather_field(GGPKD pkd, int field, int width, int comp, DevBuf &dst)
{
    SphDev &S = pkd->sph;
    GG_TRY(ensure_field(pkd, field));
    GG_TRY(dst.reserve(sizeof(double) * std::max(S.n, 1)));
    hipLaunchKernelGGL(k_gather_comp, dim3(nblk(S.n, 256)), dim3(256), 0, pkd->stream, S.n,
                       S.order.as<int>(), pkd->f[field].as<double>(), width, comp,
                       dst.as<double>());
    GG_HIP(hipGetLastError());
    return GG_OK;
}

static int scatter_field(GGPKD pkd, const DevBuf &src, int field, int width,
                         const uint32_t *qf)
{
    SphDev &S = pkd->sph;
    GG_TRY(ensure_field(pkd, field));
    hipLaunchKernelGGL((k_scatter<double>), dim3(nblk(S.n, 256)), dim3(256), 0, pkd->stream,
                       S.n, S.order.as<int>(), qf, src.as<double>(), width,
                       pkd->f[field].as<double>());
    GG_HIP(hipGetLastError());
    return GG_OK;
}

// The SPH structure holds TYPE_GAS particles, or every particle when no
// particle carries TYPE_GAS.
// The answer is kept until the type bits are put again or the store changes
// size: no kernel here sets or clears TYPE_GAS.
static int gas_mask(GGPKD pkd, uint32_t *mask)
{
    const int n = pkd->nLocal;
    SphDev &S = pkd->sph;
    if (S.gasMaskValid) {
        *mask = S.gasMask;
        return GG_OK;
    }
    GG_TRY(S.diffcnt.reserve(2 * sizeof(unsigned)));
    GG_HIP(hipMemsetAsync(S.diffcnt.p, 0, 2 * sizeof(unsigned), pkd->stream));
    hipLaunchKernelGGL(k_any_bits, dim3(std::max(1, std::min(nblk(n, 256), 1024))), dim3(256), 0,
                       pkd->stream, n, pkd->f[GG_IACTIVE].as<uint32_t>(), (uint32_t)GG_TYPE_GAS,
                       S.diffcnt.as<unsigned>());
    unsigned any[2] = {0, 0};
    GG_HIP(hipMemcpyAsync(any, S.diffcnt.p, 2 * sizeof(unsigned), hipMemcpyDeviceToHost,
                          pkd->stream));
    GG_HIP(hipStreamSynchronize(pkd->stream));
    *mask = any[0] ? GG_TYPE_GAS : 0;
    S.gasMask = *mask;
    S.allSelected = !any[0] || !any[1];   // no gas (all taken), or all gas
    S.gasMaskValid = 1;
    return GG_OK;
}

// Query flags (slot order) for particles with all bits of qmask.
static int make_flags(GGPKD pkd, uint32_t qmask)
{
    SphDev &S = pkd->sph;
    const int n = S.n;
    GG_TRY(S.qf.reserve(sizeof(uint32_t) * n));
    GG_TRY(S.act.reserve(sizeof(uint32_t) * n));
    GG_TRY(S.nba.reserve(sizeof(uint32_t) * n));
    GG_TRY(S.qcnt.reserve(sizeof(int) * nblk(n, 256)));
    hipLaunchKernelGGL(k_qflag, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n,
                       S.order.as<int>(), pkd->f[GG_IACTIVE].as<uint32_t>(), qmask,
                       S.qf.as<uint32_t>(), S.act.as<uint32_t>(), S.nba.as<uint32_t>(),
                       S.qcnt.as<int>());
    GG_HIP(hipGetLastError());
    return GG_OK;
}

static QArgs query_args(GGPKD pkd, int mode, int K)
{
    SphDev &S = pkd->sph;
    QArgs Q;
    Q.n = S.n;
    Q.bkt = S.nBucket;
    Q.K = K;
    Q.per = S.bPeriodic;
    Q.mode = mode;
    for (int d = 0; d < 3; ++d) Q.L[d] = S.L[d];
    Q.sx = S.sx.as<double>();
    Q.sy = S.sy.as<double>();
    Q.sz = S.sz.as<double>();
    Q.rec = S.crec.as<CRec>();
    Q.bof = S.bof.as<int>();
    Q.bbox = S.bbox.as<double>();
    Q.bfirst = S.bfirst.as<int>();
    Q.bcnt = S.bcnt.as<int>();
    Q.bden = S.bden.as<double>();
    Q.fbox = S.sparse.as<float>();
    Q.slack = S.seedSlack;
    Q.order = S.order.as<int>();
    Q.qflag = S.qf.as<uint32_t>();
    Q.guess = nullptr;
    Q.fb2 = S.fb2s.as<double>();
    Q.nbr = nullptr;
    Q.cnt = nullptr;
    Q.off = nullptr;
    Q.err = S.err.as<unsigned>();
    Q.qlist = nullptr;
    Q.nq = nullptr;
    Q.grl2 = nullptr;
    Q.gnb = nullptr;
    return Q;
}

// Read-back through pinned host memory: rb_queue copies device words behind
// the work queued so far and records an event; rb_wait returns once they
// have arrived, while the work queued after them still runs.
static int rb_queue(GGPKD pkd, const void *src, size_t words, size_t at)
{
    SphDev &S = pkd->sph;
    if (at + words > S.rbCap) {
        const size_t cap = std::max<size_t>(at + words, 1024);
        if (S.rb) GG_HIP(hipHostFree(S.rb));
        S.rb = nullptr;
        S.rbCap = 0;
        GG_HIP(hipHostMalloc((void **)&S.rb, sizeof(uint32_t) * cap, hipHostMallocDefault));
        S.rbCap = cap;
    }
    if (!S.rbEv) GG_HIP(hipEventCreateWithFlags(&S.rbEv, hipEventDisableTiming));
    GG_HIP(hipMemcpyAsync(S.rb + at, src, sizeof(uint32_t) * words, hipMemcpyDeviceToHost,
                          pkd->stream));
    GG_HIP(hipEventRecord(S.rbEv, pkd->stream));
    return GG_OK;
}

static int rb_wait(GGPKD pkd)
{
    GG_HIP(hipEventSynchronize(pkd->sph.rbEv));
    return GG_OK;
}

// The twelve error words and the per-block query counts of make_flags,
// queued (then the caller may queue more work) and read in one wait.
static int queue_err_queries(GGPKD pkd)
{
    SphDev &S = pkd->sph;
    GG_TRY(rb_queue(pkd, S.qcnt.p, nblk(S.n, 256), 12));
    return rb_queue(pkd, S.err.p, 12, 0);
}

static int wait_err_queries(GGPKD pkd, unsigned *e, int64_t *nq)
{
    SphDev &S = pkd->sph;
    GG_TRY(rb_wait(pkd));
    for (int k = 0; k < 12; ++k) e[k] = S.rb[k];
    int64_t t = 0;
    const size_t nbq = nblk(S.n, 256);
    for (size_t k = 0; k < nbq; ++k) t += (int)S.rb[12 + k];
    *nq = t;
    return GG_OK;
}

static int read_err(GGPKD pkd, unsigned *e)
{
    GG_HIP(hipMemcpyAsync(e, pkd->sph.err.p, sizeof(unsigned) * 2, hipMemcpyDeviceToHost,
                          pkd->stream));
    GG_HIP(hipStreamSynchronize(pkd->stream));
    return GG_OK;
}

}  // namespace gg

using namespace gg;

extern "C" int gg_build_local(GGPKD pkd, int nBucket, int bPeriodic, const double *fPeriod)
{
    GG_CHECK(nBucket >= 1 && nBucket <= BKT, GG_EINVAL, "nBucket %d outside 1..%d", nBucket,
             BKT);
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GG_CHECK(pkd->f[GG_R].p, GG_ESTATE, "positions must be put before gg_build_local");
    GG_CHECK(!bPeriodic || fPeriod, GG_EINVAL, "periodic build needs fPeriod");
    if (bPeriodic)
        for (int d = 0; d < 3; ++d)
            GG_CHECK(fPeriod[d] > 0, GG_EINVAL, "fPeriod[%d] must be positive", d);
    GG_TRY(ensure_field(pkd, GG_IACTIVE));
    SphDev &S = pkd->sph;
    hipStream_t st = pkd->stream;
    const int N = pkd->nLocal;
    GG_CHECK(N > 0, GG_EINVAL, "empty particle store");
    timer_mark(pkd, 0);
    uint32_t mask;
    GG_TRY(gas_mask(pkd, &mask));
    GG_TRY(S.sel.reserve(sizeof(int) * N));
    GG_TRY(S.pos.reserve(sizeof(int64_t) * (N + 1)));
    GG_TRY(S.gidx.reserve(sizeof(int) * N));
    int64_t ng = N;
    if (S.allSelected) {   // every particle: no selection to count
        hipLaunchKernelGGL(k_iota, dim3(nblk(N, 256)), dim3(256), 0, st, N, S.gidx.as<int>());
    } else {
        hipLaunchKernelGGL(k_select, dim3(nblk(N, 256)), dim3(256), 0, st, N,
                           pkd->f[GG_IACTIVE].as<uint32_t>(), mask, S.sel.as<int>());
        GG_TRY(exclusive_scan_i64(st, S.sel.as<int>(), S.pos.as<int64_t>(), N, S.sortTmp, &ng));
    }
    const int n = (int)ng;
    GG_CHECK(n > 0, GG_EINVAL, "no particles for the SPH structure");
    if (!S.allSelected)
        hipLaunchKernelGGL(k_compact, dim3(nblk(N, 256)), dim3(256), 0, st, N, S.sel.as<int>(),
                           S.pos.as<int64_t>(), S.gidx.as<int>());
    const int nbb = std::min(nblk(n, 256), 1024);
    GG_TRY(S.part.reserve(sizeof(double) * (6 * (size_t)nbb + 4)));
    hipLaunchKernelGGL(k_bbox, dim3(nbb), dim3(256), 0, st, n, S.gidx.as<int>(),
                       pkd->f[GG_R].as<double>(), S.part.as<double>());
    // the Morton grid (lower corner, scale) stays on the device, after the partials
    double *grid = S.part.as<double>() + 6 * (size_t)nbb;
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, st, nbb, S.part.as<double>(), grid);
    GG_TRY(S.key.reserve(sizeof(uint64_t) * n));
    GG_TRY(S.order.reserve(sizeof(int) * n));
    GG_TRY(S.k2.reserve(sizeof(uint64_t) * n));
    GG_TRY(S.v2.reserve(sizeof(int) * n));
    hipLaunchKernelGGL(k_morton, dim3(nblk(n, 256)), dim3(256), 0, st, n, S.gidx.as<int>(),
                       pkd->f[GG_R].as<double>(), (const double *)grid, S.k2.as<uint64_t>(),
                       S.v2.as<int>());
    size_t tmpb = 0;
    GG_HIP(rocprim::radix_sort_pairs(nullptr, tmpb, S.k2.as<uint64_t>(), S.key.as<uint64_t>(),
                                     S.v2.as<int>(), S.order.as<int>(), n, 0, 48, st));
    GG_TRY(S.sortTmp.reserve(std::max<size_t>(tmpb, 16)));
    GG_HIP(rocprim::radix_sort_pairs(S.sortTmp.p, tmpb, S.k2.as<uint64_t>(), S.key.as<uint64_t>(),
                                     S.v2.as<int>(), S.order.as<int>(), n, 0, 48, st));
    GG_TRY(S.sx.reserve(sizeof(double) * n));
    GG_TRY(S.sy.reserve(sizeof(double) * n));
    GG_TRY(S.sz.reserve(sizeof(double) * n));
    hipLaunchKernelGGL(k_gather_pos, dim3(nblk(n, 256)), dim3(256), 0, st, n, S.order.as<int>(),
                       pkd->f[GG_R].as<double>(), S.sx.as<double>(), S.sy.as<double>(),
                       S.sz.as<double>());
    // radix tree over the sorted keys, its buckets, their boxes and the
    // child records the searches descend
    const int ni = std::max(n - 1, 1);
    GG_TRY(S.child.reserve(sizeof(int) * 2 * ni));
    GG_TRY(S.rng.reserve(sizeof(int) * 2 * ni));
    GG_TRY(S.bflag.reserve(sizeof(int) * n));
    GG_TRY(S.bcode.reserve(sizeof(int) * n));
    GG_TRY(S.bof.reserve(sizeof(int) * n));
    int nb = 1;
    if (n > nBucket) {
        GG_HIP(hipMemsetAsync(S.bflag.p, 0, sizeof(int) * n, st));
        hipLaunchKernelGGL(k_lbvh, dim3(nblk(n - 1, 256)), dim3(256), 0, st, n,
                           S.key.as<uint64_t>(), S.child.as<int>(), S.rng.as<int>());
        GG_TRY(S.bpar.reserve(sizeof(int2) * n));
        hipLaunchKernelGGL(k_bucket_mark, dim3(nblk(n - 1, 256)), dim3(256), 0, st, n, nBucket,
                           S.key.as<uint64_t>(), S.child.as<int>(), S.rng.as<int>(),
                           S.bflag.as<int>(), S.bcode.as<int>(), S.bpar.as<int2>());
        int64_t tb = 0;
        GG_TRY(exclusive_scan_i64(st, S.bflag.as<int>(), S.pos.as<int64_t>(), n, S.sortTmp, &tb));
        nb = (int)tb;
        int nlev = 1;
        while ((2 << (nlev - 1)) <= nb) ++nlev;   // levels 0 .. floor(log2 nb)
        GG_TRY(S.bfirst.reserve(sizeof(int) * nb));
        GG_TRY(S.bcnt.reserve(sizeof(int) * nb));
        GG_TRY(S.bbox.reserve(sizeof(double) * 6 * nb));
        GG_TRY(S.sparse.reserve(sizeof(float) * 6 * (size_t)nb * nlev));
        float *sp = S.sparse.as<float>();
        GG_TRY(S.bden.reserve(sizeof(double) * nb));
        hipLaunchKernelGGL(k_bucket_start, dim3(nblk(n, 256)), dim3(256), 0, st, n,
                           S.bflag.as<int>(), S.bcode.as<int>(), S.bpar.as<int2>(),
                           (const double *)grid, S.pos.as<int64_t>(), S.bfirst.as<int>(),
                           S.bcnt.as<int>(), S.bden.as<double>(), S.bof.as<int>());
        hipLaunchKernelGGL(k_bucket_box, dim3((int)(((int64_t)nb * 32 + 255) / 256)), dim3(256), 0, st, nb,
                           S.bfirst.as<int>(), S.bcnt.as<int>(), S.sx.as<double>(),
                           S.sy.as<double>(), S.sz.as<double>(), S.bbox.as<double>(), sp);
        for (int lv = 0; lv + 1 < nlev; lv += 4)
            hipLaunchKernelGGL(k_sparse_levels, dim3(nblk(nb, 256)), dim3(256), 0, st, nb, lv,
                               std::min(4, nlev - 1 - lv), sp);
        GG_TRY(S.crec.reserve(sizeof(CRec) * 2 * (size_t)(n - 1)));
        hipLaunchKernelGGL(k_child_rec, dim3(nblk(2 * (n - 1), 256)), dim3(256), 0, st, n, nb,
                           nBucket, S.child.as<int>(), S.rng.as<int>(), S.bof.as<int>(), sp,
                           S.crec.as<CRec>());
    } else {   // a single bucket: the whole set; the searches need no tree
        GG_TRY(S.bfirst.reserve(sizeof(int)));
        GG_TRY(S.bcnt.reserve(sizeof(int)));
        GG_TRY(S.bbox.reserve(sizeof(double) * 6));
        GG_TRY(S.sparse.reserve(sizeof(float) * 6));
        GG_TRY(S.bden.reserve(sizeof(double)));
        GG_HIP(hipMemsetAsync(S.bof.p, 0, sizeof(int) * n, st));
        const int one[2] = {0, n};
        GG_HIP(hipMemcpyAsync(S.bfirst.p, &one[0], sizeof(int), hipMemcpyHostToDevice, st));
        GG_HIP(hipMemcpyAsync(S.bcnt.p, &one[1], sizeof(int), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_bbox_slots, dim3(1), dim3(64), 0, st, n, (const double *)grid,
                           S.sx.as<double>(),
                           S.sy.as<double>(), S.sz.as<double>(), S.bbox.as<double>(),
                           S.sparse.as<float>(), S.bden.as<double>());
    }
    GG_HIP(hipGetLastError());
    S.n = n;
    S.nb = nb;
    S.bPeriodic = bPeriodic;
    for (int d = 0; d < 3; ++d) S.L[d] = bPeriodic ? fPeriod[d] : 0.0;
    S.nBucket = nBucket;
    S.built = 1;
    S.listValid = 0;
    timer_mark(pkd, 1);
    const int ph[1] = {2};
    timer_collect(pkd, 0, 1, ph);
    return GG_OK;
}

// The kNN lists' bookkeeping once the operator has run: the search radii
// (-1 marks a failed search), the convergence check, what gg_resmooth may
// reuse, and the number of particles smoothed.
// The device part, queued before the host waits for the error words: the
// radii and flags gg_resmooth compares with (fBall2 itself goes into the
// store with the operator's outputs).  A failed search leaves no list valid.
static int knn_store(GGPKD pkd)
{
    SphDev &S = pkd->sph;
    hipStream_t st = pkd->stream;
    const int n = S.n;
    S.listValid = 0;
    GG_TRY(S.fb2knn.reserve(sizeof(double) * n));
    GG_TRY(S.qfknn.reserve(sizeof(uint32_t) * n));
    GG_HIP(hipMemcpyAsync(S.fb2knn.p, S.fb2s.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    GG_HIP(hipMemcpyAsync(S.qfknn.p, S.qf.p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    return GG_OK;
}

static int knn_finish(GGPKD pkd, const unsigned *e, int64_t nq, int K, int *pnSmoothed)
{
    SphDev &S = pkd->sph;
    S.fails = e[0];
    S.ties = e[1];
    S.retried = e[10];
    for (int k = 0; k < 4; ++k) S.diag[k] = e[2 + k];
    S.passes = e[6];
    uint64_t cand;
    std::memcpy(&cand, &e[8], sizeof(cand));
    S.cand = (int64_t)cand;
    GG_CHECK(e[0] == 0, GG_EUNSUP,
             "kNN search did not converge for %u particles (degenerate positions?)", e[0]);
    S.listK = K;
    S.listValid = (S.ties == 0);
    S.listQueries = (int)nq;
    if (pnSmoothed) *pnSmoothed = (int)nq;
    return GG_OK;
}

// SMX_DENSITY over the fixed-K lists gg_smooth has just found: Density, or
// DensitySym (bSymmetric) through the reverse lists.  Neither marks
// TYPE_NbrOfACTIVE nor records the force pass's extras.
static int density_pass(GGPKD pkd, int K, int bSymmetric, int *pnSmoothed)
{
    SphDev &S = pkd->sph;
    hipStream_t st = pkd->stream;
    const int n = S.n;
    S.exValid = 0;
    GG_TRY(gather_field(pkd, GG_MASS, 1, 0, S.m));
    GG_TRY(S.o_rho.reserve(sizeof(double) * n));
    DSArgs D;
    D.n = n;
    D.sym = bSymmetric;
    D.per = S.bPeriodic;
    for (int d = 0; d < 3; ++d) D.L[d] = S.L[d];
    D.sx = S.sx.as<double>();
    D.sy = S.sy.as<double>();
    D.sz = S.sz.as<double>();
    D.fb2 = S.fb2s.as<double>();
    D.off = S.off.as<int64_t>();
    D.idx = S.nbr.as<int>();
    D.roff = nullptr;
    D.ridx = nullptr;
    D.qflag = S.qf.as<uint32_t>();
    D.m = S.m.as<double>();
    D.rho0 = nullptr;
    D.rho = S.o_rho.as<double>();
    if (bSymmetric) {   // lists of the owners that hold each particle, sorted
        GG_TRY(gather_field(pkd, GG_DENSITY, 1, 0, S.rho));
        GG_TRY(S.rcnt.reserve(sizeof(int) * (n + 1)));
        GG_TRY(S.roff.reserve(sizeof(int64_t) * (n + 1)));
        GG_HIP(hipMemsetAsync(S.rcnt.p, 0, sizeof(int) * (n + 1), st));
        hipLaunchKernelGGL(k_rev_count, dim3(nblk(n, 256)), dim3(256), 0, st, n,
                           S.qf.as<uint32_t>(), S.off.as<int64_t>(), S.nbr.as<int>(),
                           S.rcnt.as<int>());
        int64_t rtot = 0;
        GG_TRY(exclusive_scan_i64(st, S.rcnt.as<int>(), S.roff.as<int64_t>(), n, S.sortTmp, &rtot));
        GG_TRY(S.ridx.reserve(sizeof(int) * std::max<int64_t>(rtot, 1)));
        GG_TRY(S.rcur.reserve(sizeof(int) * (n + 1)));
        GG_HIP(hipMemsetAsync(S.rcur.p, 0, sizeof(int) * (n + 1), st));
        hipLaunchKernelGGL(k_rev_fill, dim3(nblk(n, 256)), dim3(256), 0, st, n,
                           S.qf.as<uint32_t>(), S.off.as<int64_t>(), S.nbr.as<int>(),
                           S.roff.as<int64_t>(), S.rcur.as<int>(), S.ridx.as<int>());
        hipLaunchKernelGGL(k_rev_sort, dim3(nblk(n, 256)), dim3(256), 0, st, n,
                           S.roff.as<int64_t>(), S.ridx.as<int>());
        GG_HIP(hipGetLastError());
        D.roff = S.roff.as<int64_t>();
        D.ridx = S.ridx.as<int>();
        D.rho0 = S.rho.as<double>();
    }
    timer_mark(pkd, 4);
    const int L = S.lpp, grid = (int)(((int64_t)n * L + 255) / 256);
    const bool f32d = pkd->prec == GG_PREC_FP32;
#define GG_DEN(T, LL) hipLaunchKernelGGL((k_density<T, LL>), dim3(grid), dim3(256), 0, st, D)
    if (L == 4) { if (f32d) GG_DEN(float, 4); else GG_DEN(double, 4); }
    else if (L == 16) { if (f32d) GG_DEN(float, 16); else GG_DEN(double, 16); }
    else { if (f32d) GG_DEN(float, 8); else GG_DEN(double, 8); }
#undef GG_DEN
    GG_HIP(hipGetLastError());
    timer_mark(pkd, 5);
    GG_TRY(queue_err_queries(pkd));
    GG_TRY(knn_store(pkd));
    GG_TRY(scatter_field(pkd, S.fb2s, GG_FBALL2, 1, S.qf.as<uint32_t>()));
    // DensitySym writes every particle of the structure (the unsmoothed ones
    // receive their neighbours' halves); Density the smoothed ones
    GG_TRY(scatter_field(pkd, S.o_rho, GG_DENSITY, 1, bSymmetric ? nullptr : S.qf.as<uint32_t>()));
    hipLaunchKernelGGL(k_flags_out, dim3(nblk(n, 256)), dim3(256), 0, st, n, S.order.as<int>(),
                       S.qf.as<uint32_t>(), (const uint32_t *)nullptr,
                       pkd->f[GG_IACTIVE].as<uint32_t>());
    GG_HIP(hipGetLastError());
    timer_mark(pkd, 6);
    unsigned e[12];
    int64_t nq = 0;
    GG_TRY(wait_err_queries(pkd, e, &nq));
    GG_TRY(knn_finish(pkd, e, nq, K, pnSmoothed));
    const int ph[6] = {9, 3, 10, bSymmetric ? 6 : 8, 4, 8};
    timer_collect(pkd, 0, 6, ph);
    return GG_OK;
}

// DenDVDX (smoothfcn.c:3390-3627) over the CSR lists (off, idx) of the
// smooth-active particles: slot-order inputs, the force pass's extras reset,
// then the kernel.  The caller reads the error words.
static int dendvdx_launch(GGPKD pkd, const gg_smf *smf, const int64_t *off, const int *idx)
{
    SphDev &S = pkd->sph;
    hipStream_t st = pkd->stream;
    const int n = S.n;
    const bool f32 = pkd->prec == GG_PREC_FP32;
    for (int f : {GG_MASS, GG_VPRED, GG_UPRED}) GG_TRY(ensure_field(pkd, f));
    GG_TRY(S.drec.reserve((f32 ? sizeof(DRec<float>) : sizeof(DRec<double>)) * (size_t)n));
#define GG_DPACK(T)                                                                             \
    hipLaunchKernelGGL(k_dpack<T>, dim3(nblk(n, 256)), dim3(256), 0, st, n, S.order.as<int>(), \
                       S.sx.as<double>(), S.sy.as<double>(), S.sz.as<double>(),               \
                       S.fb2s.as<double>(), S.qf.as<uint32_t>(), S.act.as<uint32_t>(),        \
                       pkd->f[GG_MASS].as<double>(), pkd->f[GG_VPRED].as<double>(),           \
                       pkd->f[GG_UPRED].as<double>(), S.drec.as<DRec<T>>())
    if (f32) GG_DPACK(float);
    else GG_DPACK(double);
#undef GG_DPACK
    GG_TRY(S.o_rho.reserve(sizeof(double) * n));
    GG_TRY(S.o_divv.reserve(sizeof(double) * n));
    GG_TRY(S.o_curl.reserve(sizeof(double) * 3 * n));
    GG_TRY(S.o_dvds.reserve(sizeof(double) * n));
    GG_TRY(S.o_divvt.reserve(sizeof(double) * n));
    GG_TRY(S.o_bal.reserve(sizeof(double) * n));
    GG_TRY(S.o_diff.reserve(sizeof(double) * n));
    DArgs D;
    D.n = n;
    D.op = GG_SMX_DENDVDX;
    D.per = S.bPeriodic;
    for (int d = 0; d < 3; ++d) D.L[d] = S.L[d];
    D.off = off;
    D.idx = idx;
    D.rec = S.drec.p;
    D.H = smf->H;
    D.a = smf->a;
    D.gamma = smf->gamma;
    D.bCannonical = smf->bCannonical;
    D.iViscosityLimiter = smf->iViscosityLimiter;
    D.bConstantDiffusion = smf->bConstantDiffusion;
    D.rho = S.o_rho.as<double>();
    D.divv = S.o_divv.as<double>();
    D.curl = S.o_curl.as<double>();
    D.dvds = S.o_dvds.as<double>();
    D.divvt = S.o_divvt.as<double>();
    D.bal = S.o_bal.as<double>();
    D.diff = S.o_diff.as<double>();
    D.nbract = S.nba.as<uint32_t>();
    GG_TRY(S.ecnt.reserve(sizeof(int) * (size_t)n));
    GG_TRY(S.eidx.reserve(sizeof(int) * (size_t)n * ECAP));
    GG_HIP(hipMemsetAsync(S.ecnt.p, 0, sizeof(int) * (size_t)n, st));
    D.ecnt = S.ecnt.as<int>();
    D.eidx = S.eidx.as<int>();
    D.ecap = ECAP;
    D.err = S.err.as<unsigned>();
    const int L = S.lpp, grid = (int)(((int64_t)n * L + 255) / 256);
    const bool f32d = pkd->prec == GG_PREC_FP32;
#define GG_DDX(T, LL) hipLaunchKernelGGL((k_dendvdx<T, LL>), dim3(grid), dim3(256), 0, st, D)
    if (L == 4) { if (f32d) GG_DDX(float, 4); else GG_DDX(double, 4); }
    else if (L == 16) { if (f32d) GG_DDX(float, 16); else GG_DDX(double, 16); }
    else { if (f32d) GG_DDX(float, 8); else GG_DDX(double, 8); }
#undef GG_DDX
    GG_HIP(hipGetLastError());
    return GG_OK;
}

// DenDVDX's outputs into the store (smoothed particles only) and the flags.
static int dendvdx_store(GGPKD pkd, bool withFb2)
{
    SphDev &S = pkd->sph;
    hipStream_t st = pkd->stream;
    const int n = S.n;
    const uint32_t *qf = S.qf.as<uint32_t>();
    ScatterSet F;
    F.nf = 0;
    auto add = [&](const DevBuf &src, int field, int width) -> int {
        GG_TRY(ensure_field(pkd, field));
        F.src[F.nf] = src.as<double>();
        F.dst[F.nf] = pkd->f[field].as<double>();
        F.width[F.nf] = width;
        ++F.nf;
        return GG_OK;
    };
    GG_TRY(add(S.o_rho, GG_DENSITY, 1));
    GG_TRY(add(S.o_divv, GG_DIVV, 1));
    GG_TRY(add(S.o_curl, GG_CURLV, 3));
    GG_TRY(add(S.o_dvds, GG_DVDS, 1));
    GG_TRY(add(S.o_divvt, GG_FDIVV_T, 1));
    GG_TRY(add(S.o_bal, GG_BALSARA, 1));
    GG_TRY(add(S.o_diff, GG_DIFF, 1));
    if (withFb2) GG_TRY(add(S.fb2s, GG_FBALL2, 1));
    hipLaunchKernelGGL(k_scatter_set, dim3(nblk(n, 256)), dim3(256), 0, st, n,
                       S.order.as<int>(), qf, F);
    hipLaunchKernelGGL(k_flags_out, dim3(nblk(n, 256)), dim3(256), 0, st, n, S.order.as<int>(),
                       S.qf.as<uint32_t>(), S.nba.as<uint32_t>(), pkd->f[GG_IACTIVE].as<uint32_t>());
    GG_HIP(hipGetLastError());
    return GG_OK;
}

extern "C" int gg_smooth(GGPKD pkd, const gg_smf *smf, int nSmooth, int bPeriodic,
                         int bSymmetric, int iSmoothType, int *pnSmoothed)
{
    GG_CHECK(pkd && smf, GG_EINVAL, "null argument");
    SphDev &S = pkd->sph;
    GG_CHECK(S.built, GG_ESTATE, "gg_smooth before gg_build_local");
    GG_CHECK(bPeriodic == S.bPeriodic, GG_EINVAL, "bPeriodic differs from gg_build_local");
    GG_CHECK(iSmoothType == GG_SMX_DENSITY || iSmoothType == GG_SMX_DENDVDX, GG_EUNSUP,
             "gg_smooth implements SMX_DENSITY and SMX_DENDVDX");
    GG_CHECK(bSymmetric == 0 || bSymmetric == 1, GG_EINVAL, "bSymmetric %d is not 0 or 1",
             bSymmetric);
    GG_CHECK(nSmooth >= 1 && nSmooth <= 128, GG_EINVAL, "nSmooth %d outside 1..128", nSmooth);
    GG_CHECK(nSmooth <= S.n, GG_EINVAL, "nSmooth %d exceeds the %d particles", nSmooth, S.n);
    hipStream_t st = pkd->stream;
    const int n = S.n, K = nSmooth;
    GG_TRY(make_flags(pkd, GG_TYPE_SMOOTHACTIVE));
    GG_TRY(S.nbr.reserve(sizeof(int) * (size_t)K * n));
    GG_TRY(S.fb2s.reserve(sizeof(double) * n));
    GG_TRY(S.err.reserve(sizeof(unsigned) * 12));
    GG_HIP(hipMemsetAsync(S.err.p, 0, sizeof(unsigned) * 12, st));
    const bool hint = S.seedSlack > 0.0;
    if (hint) GG_TRY(gather_field(pkd, GG_FBALL2, 1, 0, S.guess));
    GG_TRY(S.retry.reserve(sizeof(int) * n));
    KLArgs KL;
    KL.Q = query_args(pkd, 0, K);
    KL.Q.guess = hint ? S.guess.as<double>() : nullptr;
    KL.Q.nbr = S.nbr.as<int>();
    KL.nb = S.nb;
    KL.bfirst = S.bfirst.as<int>();
    KL.bcnt = S.bcnt.as<int>();
    GG_TRY(S.grpn.reserve(sizeof(int) * S.nb));
    GG_TRY(S.grpr.reserve(sizeof(double) * S.nb));
    KL.gnb = S.grpn.as<int>();
    KL.grl2 = S.grpr.as<double>();
    KL.retry = S.retry.as<int>();
    KL.nretry = S.err.as<unsigned>() + 10;   // err[10]: retries (zeroed above)
    timer_mark(pkd, 0);
    timer_mark(pkd, 1);   // the group lists are found inside k_knn_bucket
    hipLaunchKernelGGL(k_knn_bucket, dim3(S.nb), dim3(64), 0, st, KL);
    timer_mark(pkd, 2);
    GG_HIP(hipGetLastError());
    {   // queries the leaf groups could not settle; their number stays on the device
        QArgs Q = query_args(pkd, 0, K);
        Q.guess = hint ? S.guess.as<double>() : nullptr;
        Q.nbr = S.nbr.as<int>();
        Q.qlist = S.retry.as<int>();
        Q.nq = KL.nretry;
        Q.grl2 = KL.grl2;
        Q.gnb = KL.gnb;
        hipLaunchKernelGGL(k_query, dim3(std::min(n, 4096)), dim3(64), 0, st, Q);
        GG_HIP(hipGetLastError());
    }
    timer_mark(pkd, 3);
    GG_TRY(S.off.reserve(sizeof(int64_t) * (n + 1)));
    hipLaunchKernelGGL(k_fixed_off, dim3(nblk(n + 1, 256)), dim3(256), 0, st, n, K,
                       S.off.as<int64_t>());
    if (iSmoothType == GG_SMX_DENSITY) return density_pass(pkd, K, bSymmetric, pnSmoothed);
    timer_mark(pkd, 4);
    GG_TRY(dendvdx_launch(pkd, smf, S.off.as<int64_t>(), S.nbr.as<int>()));
    timer_mark(pkd, 5);
    GG_TRY(queue_err_queries(pkd));
    GG_TRY(knn_store(pkd));
    GG_TRY(dendvdx_store(pkd, true));
    GG_HIP(hipGetLastError());
    timer_mark(pkd, 6);
    // the stores run while the host reads the error words
    unsigned e[12];
    int64_t nq = 0;
    GG_TRY(wait_err_queries(pkd, e, &nq));
    S.exValid = (e[7] == 0);   // the extras lists are whole
    GG_TRY(knn_finish(pkd, e, nq, K, pnSmoothed));
    const int ph[6] = {9, 3, 10, 8, 4, 8};
    timer_collect(pkd, 0, 6, ph);
    return GG_OK;
}

extern "C" int gg_gas_pressure(GGPKD pkd, const gg_smf *smf)
{
    GG_CHECK(pkd && smf, GG_EINVAL, "null argument");
    const int n = pkd->nLocal;
    const int fl[] = {GG_UPRED, GG_DENSITY, GG_FBALL2, GG_UDOTPDV, GG_UDOTAV, GG_UDOTDIFF,
                      GG_POVERRHO2, GG_C, GG_DT, GG_IACTIVE};
    for (int f : fl) GG_TRY(ensure_field(pkd, f));
    uint32_t mask;
    GG_TRY(gas_mask(pkd, &mask));
    hipLaunchKernelGGL(k_gas_pressure, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n,
                       pkd->f[GG_IACTIVE].as<uint32_t>(), mask, pkd->f[GG_UPRED].as<double>(),
                       pkd->f[GG_DENSITY].as<double>(), pkd->f[GG_FBALL2].as<double>(),
                       pkd->f[GG_UDOTPDV].as<double>(), pkd->f[GG_UDOTAV].as<double>(),
                       pkd->f[GG_UDOTDIFF].as<double>(), smf->gamma, smf->dtFacCourant,
                       smf->dResolveJeans, pkd->f[GG_POVERRHO2].as<double>(),
                       pkd->f[GG_C].as<double>(), pkd->f[GG_DT].as<double>());
    GG_HIP(hipGetLastError());
    return GG_OK;
}

// SPH bucket-size autotuner: for each candidate nBucket, build the search
// structure and run the DenDVDX pass (kNN + density) twice from the
// caller's fBall2, time the device work of the second run (structure build,
// group lists, kNN, retries, DenDVDX and its stores, from the handle's
// events), keep the fastest, and finish with a build and a pass at it, so
// the structure and every output are those of gg_build_local(best) +
// gg_smooth.
extern "C" int gg_tune_sph(GGPKD pkd, const gg_smf *smf, int nSmooth, int bPeriodic,
                           const double *fPeriod, const int *cand, int nCand, double *ms,
                           int *pBest)
{
    GG_CHECK(pkd && smf && cand && nCand >= 1, GG_EINVAL, "null argument or no candidate");
    for (int k = 0; k < nCand; ++k)
        GG_CHECK(cand[k] >= 1 && cand[k] <= BKT, GG_EINVAL, "candidate nBucket %d outside 1..%d",
                 cand[k], BKT);
    const int n = pkd->nLocal;
    GG_TRY(ensure_field(pkd, GG_FBALL2));
    hipStream_t st = pkd->stream;
    DevBuf save;
    GG_TRY(save.reserve(sizeof(double) * (size_t)std::max(n, 1)));
    const size_t w1 = sizeof(double) * (size_t)n;
    auto restore = [&]() -> int {
        if (n > 0)
            GG_HIP(hipMemcpyAsync(pkd->f[GG_FBALL2].p, save.p, w1, hipMemcpyDeviceToDevice, st));
        return GG_OK;
    };
    auto pass = [&](int nb) -> int {
        int ns = 0;
        GG_TRY(restore());
        GG_TRY(gg_build_local(pkd, nb, bPeriodic, fPeriod));
        return gg_smooth(pkd, smf, nSmooth, bPeriodic, 0, GG_SMX_DENDVDX, &ns);
    };
    int rc = GG_OK;
    if (n > 0 && hipMemcpyAsync(save.p, pkd->f[GG_FBALL2].p, w1, hipMemcpyDeviceToDevice, st) !=
                     hipSuccess)
        rc = GG_EHIP;
    int best = -1;
    double bestMs = 0;
    for (int k = 0; k < nCand && rc == GG_OK; ++k) {
        for (int rep = 0; rep < 2 && rc == GG_OK; ++rep) rc = pass(cand[k]);
        if (rc != GG_OK) break;
        timer_resolve(pkd);
        const double *t = pkd->tm.ms;
        const double tk = t[2] + t[3] + t[4] + t[8] + t[9] + t[10];
        if (ms) ms[k] = tk;
        if (best < 0 || tk < bestMs) {
            best = k;
            bestMs = tk;
        }
    }
    if (rc == GG_OK) rc = pass(cand[best]);
    save.release();
    if (rc != GG_OK) return rc;
    if (pBest) *pBest = cand[best];
    return GG_OK;
}

extern "C" int gg_set_sph_lanes(GGPKD pkd, int lanes)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GG_CHECK(lanes == 4 || lanes == 8 || lanes == 16, GG_EINVAL,
             "lanes per particle %d is not 4, 8 or 16", lanes);
    pkd->sph.lpp = lanes;
    return GG_OK;
}

extern "C" int gg_tune_sph_lanes(GGPKD pkd, const gg_smf *smf, int nSmooth, int bPeriodic,
                                 const int *cand, int nCand, double *ms, int *pBest)
{
    GG_CHECK(pkd && smf && cand && nCand >= 1, GG_EINVAL, "null argument or no candidate");
    for (int k = 0; k < nCand; ++k)
        GG_CHECK(cand[k] == 4 || cand[k] == 8 || cand[k] == 16, GG_EINVAL,
                 "candidate lanes per particle %d is not 4, 8 or 16", cand[k]);
    SphDev &S = pkd->sph;
    GG_CHECK(S.built, GG_ESTATE, "gg_tune_sph_lanes before gg_build_local");
    const int n = pkd->nLocal;
    hipStream_t st = pkd->stream;
    // what the passes accumulate into (or start from): saved once, put back
    // before every pass
    const int keep[] = {GG_FBALL2, GG_A, GG_UDOTPDV, GG_UDOTAV, GG_UDOTDIFF, GG_MUMAX,
                        GG_DTNEW, GG_DT};
    constexpr int NK = sizeof(keep) / sizeof(keep[0]);
    DevBuf save[NK];
    size_t wb[NK];
    int rc = GG_OK;
    for (int k = 0; k < NK && rc == GG_OK; ++k) {
        rc = ensure_field(pkd, keep[k]);
        wb[k] = sizeof(double) * (size_t)field_width(keep[k]) * (size_t)n;
        if (rc == GG_OK) rc = save[k].reserve(std::max<size_t>(wb[k], 8));
        if (rc == GG_OK && n > 0 &&
            hipMemcpyAsync(save[k].p, pkd->f[keep[k]].p, wb[k], hipMemcpyDeviceToDevice, st) !=
                hipSuccess)
            rc = GG_EHIP;
    }
    const int lpp0 = S.lpp;
    auto pass = [&](int L) -> int {
        for (int k = 0; k < NK; ++k)
            if (n > 0)
                GG_HIP(hipMemcpyAsync(pkd->f[keep[k]].p, save[k].p, wb[k], hipMemcpyDeviceToDevice,
                                      st));
        S.lpp = L;
        int ns = 0;
        GG_TRY(gg_smooth(pkd, smf, nSmooth, bPeriodic, 0, GG_SMX_DENDVDX, &ns));
        GG_TRY(gg_gas_pressure(pkd, smf));
        return gg_resmooth(pkd, smf, nSmooth, bPeriodic, 1, GG_SMX_SPHPRESSURETERMS);
    };
    int best = -1;
    double bestMs = 0;
    for (int k = 0; k < nCand && rc == GG_OK; ++k) {
        for (int rep = 0; rep < 2 && rc == GG_OK; ++rep) rc = pass(cand[k]);
        if (rc != GG_OK) break;
        timer_resolve(pkd);
        const double tk = pkd->tm.ms[4] + pkd->tm.ms[7];   // DenDVDX + force pass
        if (ms) ms[k] = tk;
        if (best < 0 || tk < bestMs) {
            best = k;
            bestMs = tk;
        }
    }
    if (rc == GG_OK) rc = pass(cand[best]);
    for (auto &b : save) b.release();
    if (rc != GG_OK) {
        S.lpp = lpp0;
        return rc;
    }
    if (pBest) *pBest = cand[best];
    return GG_OK;
}

extern "C" int gg_resmooth(GGPKD pkd, const gg_smf *smf, int nSmooth, int bPeriodic,
                           int bSymmetric, int iSmoothType)
{
    (void)nSmooth;
    GG_CHECK(pkd && smf, GG_EINVAL, "null argument");
    SphDev &S = pkd->sph;
    GG_CHECK(S.built, GG_ESTATE, "gg_resmooth before gg_build_local");
    GG_CHECK(bPeriodic == S.bPeriodic, GG_EINVAL, "bPeriodic differs from gg_build_local");
    GG_CHECK(iSmoothType == GG_SMX_SPHPRESSURETERMS || iSmoothType == GG_SMX_DENDVDX, GG_EUNSUP,
             "gg_resmooth implements SMX_SPHPRESSURETERMS and SMX_DENDVDX");
    GG_CHECK(bSymmetric == 1 || (iSmoothType == GG_SMX_DENDVDX && bSymmetric == 0), GG_EUNSUP,
             "SphPressureTerms is the symmetric operator (bSymmetric=1)");
    hipStream_t st = pkd->stream;
    const int n = S.n;
    timer_mark(pkd, 0);
    GG_TRY(make_flags(pkd, GG_TYPE_SMOOTHACTIVE));
    GG_TRY(S.fb2s.reserve(sizeof(double) * n));
    GG_TRY(gather_field(pkd, GG_FBALL2, 1, 0, S.fb2s));
    GG_TRY(S.err.reserve(sizeof(unsigned) * 12));
    GG_HIP(hipMemsetAsync(S.err.p, 0, sizeof(unsigned) * 12, st));
    // reuse the kNN lists when they are exactly the ball lists of the
    // stored radii (same queries, same fBall2, no boundary ties)
    bool reuse = false;
    if (S.listValid) {
        GG_TRY(S.diffcnt.reserve(sizeof(unsigned)));
        GG_HIP(hipMemsetAsync(S.diffcnt.p, 0, sizeof(unsigned), st));
        hipLaunchKernelGGL(k_count_diff, dim3(nblk(n, 256)), dim3(256), 0, st, n,
                           S.fb2s.as<double>(), S.fb2knn.as<double>(), S.qf.as<uint32_t>(),
                           S.qfknn.as<uint32_t>(), S.diffcnt.as<unsigned>());
        GG_TRY(rb_queue(pkd, S.diffcnt.p, 1, 0));
    }
    timer_mark(pkd, 1);
    // the force pass's per-particle records in slot order (they do not depend
    // on the lists), packed while the host reads the count
    const bool f32 = pkd->prec == GG_PREC_FP32;
    if (iSmoothType == GG_SMX_SPHPRESSURETERMS) {
        const int in[] = {GG_MASS, GG_VPRED, GG_UPRED, GG_DENSITY, GG_POVERRHO2, GG_C,
                          GG_BALSARA, GG_DIFF, GG_DT};
        for (int f : in) GG_TRY(ensure_field(pkd, f));
        GG_TRY(S.hrec.reserve((f32 ? sizeof(HRec<float>) : sizeof(HRec<double>)) * (size_t)n));
        GG_TRY(S.res.reserve(sizeof(double) * 8 * n));
#define GG_HPACK(T)                                                                           \
    hipLaunchKernelGGL(k_hpack<T>, dim3(nblk(n, 256)), dim3(256), 0, st, n, S.order.as<int>(), \
                       S.sx.as<double>(), S.sy.as<double>(), S.sz.as<double>(),               \
                       S.qf.as<uint32_t>(), S.fb2s.as<double>(), pkd->f[GG_MASS].as<double>(), \
                       pkd->f[GG_VPRED].as<double>(), pkd->f[GG_UPRED].as<double>(),          \
                       pkd->f[GG_DENSITY].as<double>(), pkd->f[GG_POVERRHO2].as<double>(),     \
                       pkd->f[GG_C].as<double>(), pkd->f[GG_BALSARA].as<double>(),             \
                       pkd->f[GG_DIFF].as<double>(), pkd->f[GG_DT].as<double>(),               \
                       pkd->f[GG_IACTIVE].as<uint32_t>(), S.hrec.as<HRec<T>>())
        if (f32) GG_HPACK(float);
        else GG_HPACK(double);
#undef GG_HPACK
        GG_HIP(hipGetLastError());
    }
    timer_mark(pkd, 2);
    if (S.listValid) {
        GG_TRY(rb_wait(pkd));
        reuse = (S.rb[0] == 0);
    }
    const int64_t *off;
    const int *idx;
    if (reuse) {
        off = S.off.as<int64_t>();
        idx = S.nbr.as<int>();
    } else {
        GG_TRY(S.gcnt.reserve(sizeof(int) * (n + 1)));
        GG_TRY(S.goff.reserve(sizeof(int64_t) * (n + 1)));
        QArgs Q = query_args(pkd, 1, 0);
        Q.cnt = S.gcnt.as<int>();
        hipLaunchKernelGGL(k_query, dim3(n), dim3(64), 0, st, Q);
        GG_HIP(hipGetLastError());
        int64_t tot = 0;
        GG_TRY(exclusive_scan_i64(st, S.gcnt.as<int>(), S.goff.as<int64_t>(), n, S.sortTmp, &tot));
        GG_TRY(S.gent.reserve(sizeof(int) * std::max<int64_t>(tot, 1)));
        Q = query_args(pkd, 2, 0);
        Q.off = S.goff.as<int64_t>();
        Q.nbr = S.gent.as<int>();
        hipLaunchKernelGGL(k_query, dim3(n), dim3(64), 0, st, Q);
        GG_HIP(hipGetLastError());
        unsigned e[2];
        GG_TRY(read_err(pkd, e));
        GG_CHECK(e[0] == 0, GG_EUNSUP,
                 "ball gather overflowed its %d-entry buffer for %u particles", CAP, e[0]);
        off = S.goff.as<int64_t>();
        idx = S.gent.as<int>();
    }
    timer_mark(pkd, 3);
    if (iSmoothType == GG_SMX_DENDVDX) {   // DenDVDX over the ball lists
        GG_TRY(dendvdx_launch(pkd, smf, off, idx));
        timer_mark(pkd, 4);
        GG_TRY(rb_queue(pkd, S.err.p, 12, 0));
        GG_TRY(dendvdx_store(pkd, false));
        timer_mark(pkd, 5);
        GG_TRY(rb_wait(pkd));
        // the extras belong to the kNN lists only when those were the lists
        S.exValid = reuse && S.rb[7] == 0;
        const int ph[5] = {5, 8, 5, 4, 8};
        timer_collect(pkd, 0, 5, ph);
        return GG_OK;
    }
    const bool fixedExtras = reuse && S.exValid;
    if (fixedExtras) {   // the density pass found them: sort each list
        GG_TRY(S.nxd.reserve(sizeof(int64_t) * nblk(n, 256)));
        hipLaunchKernelGGL(k_rev_sort_fixed, dim3(nblk(n, 256)), dim3(256), 0, st, n, ECAP,
                           S.ecnt.as<int>(), S.eidx.as<int>(), S.nxd.as<int64_t>());
        S.nExtraDev = nblk(n, 256);   // partial totals, added up when the statistics are read
    } else {
    // extras: particles whose balls hold i but which are not in i's list
    GG_TRY(S.rcnt.reserve(sizeof(int) * (n + 1)));
    GG_TRY(S.roff.reserve(sizeof(int64_t) * (n + 1)));
    GG_HIP(hipMemsetAsync(S.rcnt.p, 0, sizeof(int) * (n + 1), st));
    hipLaunchKernelGGL(k_extra, dim3(nblk(n * 16, 256)), dim3(256), 0, st, n, 0, S.bPeriodic,
                       S.L[0], S.L[1], S.L[2], S.qf.as<uint32_t>(), off, idx, S.sx.as<double>(),
                       S.sy.as<double>(), S.sz.as<double>(), S.fb2s.as<double>(),
                       S.rcnt.as<int>(), (const int64_t *)nullptr, (int *)nullptr);
    int64_t etot = 0;
    GG_TRY(exclusive_scan_i64(st, S.rcnt.as<int>(), S.roff.as<int64_t>(), n, S.sortTmp, &etot));
    S.nExtra = etot;
    S.nExtraDev = 0;
    GG_TRY(S.ridx.reserve(sizeof(int) * std::max<int64_t>(etot, 1)));
    GG_HIP(hipMemsetAsync(S.rcnt.p, 0, sizeof(int) * (n + 1), st));
    hipLaunchKernelGGL(k_extra, dim3(nblk(n * 16, 256)), dim3(256), 0, st, n, 1, S.bPeriodic,
                       S.L[0], S.L[1], S.L[2], S.qf.as<uint32_t>(), off, idx, S.sx.as<double>(),
                       S.sy.as<double>(), S.sz.as<double>(), S.fb2s.as<double>(),
                       S.rcnt.as<int>(), S.roff.as<int64_t>(), S.ridx.as<int>());
    hipLaunchKernelGGL(k_rev_sort, dim3(nblk(n, 256)), dim3(256), 0, st, n, S.roff.as<int64_t>(),
                       S.ridx.as<int>());
    }
    GG_HIP(hipGetLastError());
    timer_mark(pkd, 4);
    H2Args H;
    H.n = n;
    H.per = S.bPeriodic;
    for (int d = 0; d < 3; ++d) H.L[d] = S.L[d];
    H.off = off;
    H.eoff = fixedExtras ? nullptr : S.roff.as<int64_t>();
    H.ecnt = fixedExtras ? S.ecnt.as<int>() : nullptr;
    H.ecap = ECAP;
    H.idx = idx;
    H.eidx = fixedExtras ? S.eidx.as<int>() : S.ridx.as<int>();
    H.a = smf->a;
    H.H = smf->H;
    H.alpha = smf->alpha;
    H.beta = smf->beta;
    H.dtFacCourant = smf->dtFacCourant;
    H.dtFacDiffusion = smf->dtFacDiffusion;
    H.kDiff = smf->dThermalDiffusionCoeff;
    H.bCannonical = smf->bCannonical;
    H.out = S.res.as<double>();
    const int L = S.lpp, grid = (int)(((int64_t)n * L + 255) / 256);
#define GG_H2(T, LL)                                                                          \
    hipLaunchKernelGGL((k_hydro2<T, LL>), dim3(grid), dim3(256), 0, st, H,                  \
                       (const HRec<T> *)S.hrec.as<HRec<T>>())
    if (L == 4) { if (f32) GG_H2(float, 4); else GG_H2(double, 4); }
    else if (L == 16) { if (f32) GG_H2(float, 16); else GG_H2(double, 16); }
    else { if (f32) GG_H2(float, 8); else GG_H2(double, 8); }
#undef GG_H2
    GG_HIP(hipGetLastError());
    timer_mark(pkd, 5);
    const int fl[] = {GG_A, GG_UDOTPDV, GG_UDOTAV, GG_UDOTDIFF, GG_MUMAX, GG_DTNEW};
    for (int f : fl) GG_TRY(ensure_field(pkd, f));
    hipLaunchKernelGGL(k_hydro_scatter, dim3(nblk(n, 256)), dim3(256), 0, st, n,
                       S.order.as<int>(), S.qf.as<uint32_t>(), S.res.as<double>(),
                       pkd->f[GG_IACTIVE].as<uint32_t>(), pkd->f[GG_A].as<double>(),
                       pkd->f[GG_UDOTPDV].as<double>(), pkd->f[GG_UDOTAV].as<double>(),
                       pkd->f[GG_UDOTDIFF].as<double>(), pkd->f[GG_MUMAX].as<double>(),
                       pkd->f[GG_DTNEW].as<double>(), 0);
    hipLaunchKernelGGL(k_flags_out, dim3(nblk(n, 256)), dim3(256), 0, st, n, S.order.as<int>(),
                       S.qf.as<uint32_t>(), (const uint32_t *)nullptr,
                       pkd->f[GG_IACTIVE].as<uint32_t>());
    GG_HIP(hipGetLastError());
    timer_mark(pkd, 6);
    const int ph[6] = {5, 8, 5, 6, 7, 8};
    timer_collect(pkd, 0, 6, ph);
    return GG_OK;
}

extern "C" int gg_smooth_lists(GGPKD pkd, int nSmooth, int *nbr)
{
    GG_CHECK(pkd && nbr, GG_EINVAL, "null argument");
    SphDev &S = pkd->sph;
    GG_CHECK(S.built && S.listK == nSmooth, GG_ESTATE, "no kNN lists with nSmooth %d", nSmooth);
    const int n = S.n, K = nSmooth, N = pkd->nLocal;
    std::vector<int> lst((size_t)n * K), ord(n);
    std::vector<uint32_t> qf(n);
    GG_HIP(hipMemcpy(lst.data(), S.nbr.p, sizeof(int) * (size_t)n * K, hipMemcpyDeviceToHost));
    GG_HIP(hipMemcpy(ord.data(), S.order.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    GG_HIP(hipMemcpy(qf.data(), S.qfknn.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < (size_t)N * K; ++k) nbr[k] = -1;
    for (int s = 0; s < n; ++s) {
        if (!qf[s]) continue;
        for (int k = 0; k < K; ++k) nbr[(size_t)ord[s] * K + k] = ord[lst[(size_t)s * K + k]];
    }
    return GG_OK;
}

namespace gg {
// ---------------------------------------------------------- store reorder
__global__ void k_fill_i32(int n, int v, int *__restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Slots of the SPH structure: their caller indices are not "other".
__global__ void k_mark_slots(int ns, const int *__restrict__ order, int *__restrict__ flag)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ns) flag[order[s]] = 0;
}

__global__ void k_perm_build(int n, int ns, const int *__restrict__ order,
                             const int *__restrict__ flag, const int64_t *__restrict__ pos,
                             int *__restrict__ perm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i < ns) perm[i] = order[i];
    if (flag[i]) perm[ns + (int)pos[i]] = i;   // the others, in their previous order
}

// w interleaved components per particle, in np planes of n particles each
template <typename T>
__global__ void k_permute(int n, int w, int np, const int *__restrict__ perm,
                          const T *__restrict__ src, T *__restrict__ dst)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const size_t o = (size_t)w * perm[k];
    for (int p = 0; p < np; ++p)
        for (int c = 0; c < w; ++c)
            dst[(size_t)p * w * n + (size_t)w * k + c] = src[(size_t)p * w * n + o + c];
}


// The permuted copy goes to a spare buffer of the same capacity, which then
// takes the old one's place: spares are kept per capacity, so after the first
// call no allocation happens.
static DevBuf &spare(GGPKD pkd, size_t bytes)
{
    for (DevBuf &b : pkd->sph.spares)
        if (b.bytes == bytes) return b;
    pkd->sph.spares.emplace_back();
    return pkd->sph.spares.back();
}

template <typename T>
static int permute_buf(GGPKD pkd, DevBuf &buf, int w, int np, const int *perm)
{
    const int n = pkd->nLocal;
    DevBuf &tmp = spare(pkd, buf.bytes);
    GG_TRY(tmp.reserve(buf.bytes));
    hipLaunchKernelGGL((k_permute<T>), dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n, w, np,
                       perm, buf.as<T>(), tmp.as<T>());
    GG_HIP(hipGetLastError());
    std::swap(buf, tmp);
    return GG_OK;
}
}  // namespace gg

// Permutes the particle store into the SPH structure's slot order (the gas
// in Morton order first, then the other particles in their previous order),
// as pkdBuildLocal reorders pStore while it builds (pkd.c:2732-2848), so that
// the gathers and scatters between slot and store order read consecutive
// memory.  perm (optional, host): the previous index of the particle now at
// index k.
extern "C" int gg_reorder(GGPKD pkd, int *perm)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    SphDev &S = pkd->sph;
    GG_CHECK(S.built, GG_ESTATE, "gg_reorder before gg_build_local");
    const int n = pkd->nLocal;
    GG_CHECK(S.n <= n, GG_ESTATE, "SPH structure larger than the store");
    hipStream_t st = pkd->stream;
    if (n == 0) return GG_OK;
    DevBuf &P = S.sel;          // scratch: the permutation
    DevBuf &F = S.diffcnt;      // scratch: flags, then their prefix sums in S.pos
    GG_TRY(P.reserve(sizeof(int) * n));
    GG_TRY(F.reserve(sizeof(int) * n));
    GG_TRY(S.pos.reserve(sizeof(int64_t) * ((size_t)n + 1)));
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n, 256)), dim3(256), 0, st, n, 1, F.as<int>());
    hipLaunchKernelGGL(k_mark_slots, dim3(nblk(S.n, 256)), dim3(256), 0, st, S.n,
                       S.order.as<int>(), F.as<int>());
    GG_TRY(exclusive_scan_i64(st, F.as<int>(), S.pos.as<int64_t>(), n, S.sortTmp, nullptr));
    hipLaunchKernelGGL(k_perm_build, dim3(nblk(n, 256)), dim3(256), 0, st, n, S.n,
                       S.order.as<int>(), F.as<int>(), S.pos.as<int64_t>(), P.as<int>());
    GG_HIP(hipGetLastError());
    const int *pp = P.as<int>();
    int rc = GG_OK;
    for (int f = 0; f < GG_NFIELDS && rc == GG_OK; ++f) {
        if (!pkd->f[f].p) continue;
        rc = field_is32(f) ? permute_buf<uint32_t>(pkd, pkd->f[f], 1, 1, pp)
                             : permute_buf<double>(pkd, pkd->f[f], field_width(f), 1, pp);
    }
    HmlpDev &M = pkd->mlp;
    if (rc == GG_OK && M.nHist == n) {   // features: 4 planes of n floats
        rc = permute_buf<float>(pkd, M.hx, 1, 4, pp);
        if (rc == GG_OK) rc = permute_buf<int>(pkd, M.hv, 1, 1, pp);
    }
    if (rc == GG_OK) {
        hipLaunchKernelGGL(k_iota, dim3(nblk(S.n, 256)), dim3(256), 0, st, S.n, S.order.as<int>());
        if (perm && hipMemcpyAsync(perm, P.p, sizeof(int) * n, hipMemcpyDeviceToHost, st) != hipSuccess)
            rc = GG_EHIP;
    }
    if (hipStreamSynchronize(st) != hipSuccess && rc == GG_OK) rc = GG_EHIP;
    GG_CHECK(rc == GG_OK, rc, "gg_reorder: %s", rc == GG_EHIP ? "HIP error" : "failed");
    // the gravity tree and its lists refer to the old indices
    pkd->grav.built = 0;
    pkd->grav.listsCurrent = 0;
    return GG_OK;
}

extern "C" int gg_set_seed_slack(GGPKD pkd, double slack)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GG_CHECK(slack == 0.0 || (slack >= 1.0 && slack <= 2.0), GG_EINVAL,
             "slack %g: 0 (no hint) or in [1, 2]", slack);
    pkd->sph.seedSlack = slack;
    return GG_OK;
}

extern "C" int gg_smooth_stats(GGPKD pkd, int64_t *s, int n)
{
    GG_CHECK(pkd && s, GG_EINVAL, "null argument");
    SphDev &S = pkd->sph;
    if (S.nExtraDev > 0) {
        std::vector<int64_t> part(S.nExtraDev);
        GG_HIP(hipMemcpyAsync(part.data(), S.nxd.p, sizeof(int64_t) * part.size(),
                              hipMemcpyDeviceToHost, pkd->stream));
        GG_HIP(hipStreamSynchronize(pkd->stream));
        S.nExtra = 0;
        for (int64_t v : part) S.nExtra += v;
        S.nExtraDev = 0;
    }
    const int64_t v[10] = {(int64_t)S.retried, (int64_t)S.ties, S.nExtra, (int64_t)S.listQueries,
                           (int64_t)S.diag[0], (int64_t)S.diag[1], (int64_t)S.diag[2],
                           (int64_t)S.diag[3], (int64_t)S.passes, S.cand};
    for (int k = 0; k < n && k < 10; ++k) s[k] = v[k];
    return GG_OK;
}
