// Gravity tree build on the device: the tree of gasoline's pkdBuildBinary
// (pkd.c:2627-2723), node for node.
//
// Level by level, every cell of a level is split at the midpoint of the
// longest axis of its particles' bounding box (pkd.c:2493-2505) unless it
// holds nBucket particles or fewer or has no extent (pkd.c:2489).  The
// partition reproduces pkdUpperPart's Hoare scan (pkd.c:1109-1157): a Hoare
// scan swaps the k-th particle of the lower part that belongs above with the
// k-th particle, counted from the top, of the upper part that belongs below,
// so each particle's destination follows from prefix counts and the cells
// of a level partition in parallel.  pkdActiveTypeOrder's partitions
// (pkd.c:1304-1355) run the same way over each system.  Cells are numbered
// in preorder afterwards, from subtree sizes.  Mass, softening and centre of
// mass are summed as the reference sums them (buckets over their particles
// in storage order, cells from their children, lower then upper), and
// bounds, Bmax and fOpen2 are exact, so every quantity the walk decides on
// is the reference's to the bit.  The multipole moments (pkdCalcCell,
// pkd.c:2018-2135) are direct sums over each cell's particles in a fixed
// but different order; they enter only the force arithmetic.
//
// Several independent systems build at once: each is a root of its own.
#include "gg_internal.h"
#include "gg_scan.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

namespace gg {
namespace {

int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

constexpr int MOM_CH = 2048;     // particles per moment work item
constexpr int NM = 32;           // per item: raw Q (6), O (10), H (15), Bmax
constexpr int MOM_SMALL = 64;    // cells of at most this many particles: a thread each
constexpr int BOX_CH = 1024;     // particles per bounds work item
constexpr int LMAX = 512;        // cells of at most LMAX particles finish in one workgroup

// Per cell in creation (breadth-first) order.
struct GNode {
    int lo, hi, par, side;       // particle range, parent, 0 lower / 1 upper
    int ch0, ch1, dim, root;     // children (-1), split axis (-1: bucket), root cell
    int s, size, pre, depth;     // split position, subtree size, preorder id, level
    double split;
};

__global__ void k_gt_init(int n, const double *__restrict__ r, double *__restrict__ x,
                          double *__restrict__ y, double *__restrict__ z, int *__restrict__ src)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = r[3 * (size_t)i];
    y[i] = r[3 * (size_t)i + 1];
    z[i] = r[3 * (size_t)i + 2];
    src[i] = i;
}

// Flag of position i of a partitioning cell: 1 = belongs to the lower part.
// mode 0: coordinate below the cell's split; mode 1: type bits & mask set.
// Evaluated inside the prefix-sum pass and again by the rank and swap
// kernels, each time before the position's particle can have moved.
struct FlagFn {
    const int *pnode;
    const GNode *nd;
    const double *x, *y, *z;
    const int *src;
    const uint32_t *bits;
    int a, b, mode;
    uint32_t mask;
    __device__ __forceinline__ int operator()(int64_t i) const
    {
        const int c = pnode[i];
        if (c < a || c >= b) return 0;
        const GNode &g = nd[c];
        if (mode == 1) return (bits[src[i]] & mask) != 0;
        if (g.dim < 0) return 0;
        const double v = g.dim == 0 ? x[i] : (g.dim == 1 ? y[i] : z[i]);
        return v < g.split;
    }
};

// Split position s = lo + number of lower-part particles.  A split that
// leaves one side empty (the reference asserts against it, pkd.c:2512-2516)
// sets *err.
__global__ void k_gt_count(int a, int b, int mode, const int64_t *__restrict__ F,
                           GNode *__restrict__ nd, int *__restrict__ err)
{
    const int c = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b) return;
    GNode &g = nd[c];
    const bool part = mode == 1 || g.dim >= 0;
    g.s = part ? g.lo + (int)(F[g.hi + 1] - F[g.lo]) : g.lo;
    if (mode == 0 && part && (g.s == g.lo || g.s > g.hi)) atomicOr(err, 1);
}

// Positions of the upper-part positions that hold a lower particle, by
// rank counted down from hi: posR[lo + k].
__global__ void k_gt_rank(int n, FlagFn fl, const int64_t *__restrict__ F, int *__restrict__ posR)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = fl.pnode[i];
    if (c < fl.a || c >= fl.b) return;
    const GNode &g = fl.nd[c];
    if (fl.mode == 0 && g.dim < 0) return;
    if (i >= g.s && fl(i)) {
        const int k = (int)(F[g.hi + 1] - F[i + 1]);
        posR[g.lo + k] = i;
    }
}

// The swaps of the Hoare scan: the lower-part position holding the k-th
// upper particle (counted up from lo) <-> posR[lo + k], each pair swapped
// by its lower member.  With descend set, every position of the level also
// moves to its cell's child (lower part: ch0, upper: ch1; -1 in a bucket).
__global__ void k_gt_swap(int n, FlagFn fl, int descend, const int64_t *__restrict__ F,
                          const int *__restrict__ posR, double *__restrict__ x,
                          double *__restrict__ y, double *__restrict__ z, int *__restrict__ src,
                          int *pnode)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = fl.pnode[i];
    if (c < fl.a || c >= fl.b) return;
    const GNode &g = fl.nd[c];
    if (fl.mode == 0 && g.dim < 0) {
        if (descend) pnode[i] = -1;
        return;
    }
    if (i < g.s && !fl(i)) {
        const int k = (i - g.lo) - (int)(F[i] - F[g.lo]);
        const int j = posR[g.lo + k];
        double t = x[i]; x[i] = x[j]; x[j] = t;
        t = y[i]; y[i] = y[j]; y[j] = t;
        t = z[i]; z[i] = z[j]; z[j] = t;
        const int u = src[i]; src[i] = src[j]; src[j] = u;
    }
    if (descend) pnode[i] = i < g.s ? g.ch0 : g.ch1;
}

// Work items of the cells [a, b): ceil(count / ch) each for cells of more
// than minc particles, none for the others.
__global__ void k_gt_items(int a, int b, int ch, int minc, const GNode *__restrict__ nd,
                           int *__restrict__ nit)
{
    const int c = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b) return;
    const int cnt = nd[c].hi - nd[c].lo + 1;
    nit[c - a] = cnt > minc ? (cnt + ch - 1) / ch : 0;
}

// The cell (relative index) of work item k: the last r with ioff[r] <= k
// (cells without items share their offset with the next cell).
__device__ __forceinline__ int item_cell(const int64_t *__restrict__ ioff, int ncell, int64_t k)
{
    int lo = 0, hi = ncell - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ioff[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Bounds of BOX_CH particles of a cell of more than BOX_CH particles (one
// wavefront per work item): part[6k..] = min xyz, max xyz.
__global__ __launch_bounds__(256) void k_gt_box(int a, int b, const int64_t *__restrict__ ioff,
                                                const GNode *__restrict__ nd,
                                                const double *__restrict__ x,
                                                const double *__restrict__ y,
                                                const double *__restrict__ z,
                                                double *__restrict__ part)
{
    const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, ncell = b - a;
    if (k >= ioff[ncell]) return;
    const int r = item_cell(ioff, ncell, k);
    const GNode &g = nd[a + r];
    const int first = g.lo + (int)(k - ioff[r]) * BOX_CH;
    const int last = min(g.hi, first + BOX_CH - 1);
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int i = first + lane; i <= last; i += 64) {
        const double v[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; ++d) {
            lo[d] = v[d] < lo[d] ? v[d] : lo[d];
            hi[d] = v[d] > hi[d] ? v[d] : hi[d];
        }
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int d = 0; d < 3; ++d) {
            const double l = __shfl_xor(lo[d], o), h = __shfl_xor(hi[d], o);
            lo[d] = l < lo[d] ? l : lo[d];
            hi[d] = h > hi[d] ? h : hi[d];
        }
    if (lane == 0)
        for (int d = 0; d < 3; ++d) {
            part[6 * k + d] = lo[d];
            part[6 * k + 3 + d] = hi[d];
        }
}

// Split decision of each cell of the level (pkd.c:2489-2505), one wavefront
// per cell: its bounds from its particles, or from its items' bounds for a
// cell of more than BOX_CH particles.
__global__ __launch_bounds__(256) void k_gt_decide(int a, int b, int nBucket, int depth,
                                                   const int64_t *__restrict__ ioff,
                                                   const double *__restrict__ part,
                                                   const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   const double *__restrict__ z,
                                                   GNode *__restrict__ nd, int *__restrict__ nch)
{
#pragma clang fp contract(off)
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r >= b - a) return;
    GNode &g = nd[a + r];
    double bmin[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, bmax[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    if (g.hi - g.lo + 1 > BOX_CH) {
        for (int64_t it = ioff[r] + lane; it < ioff[r + 1]; it += 64)
            for (int d = 0; d < 3; ++d) {
                const double l = part[6 * it + d], h = part[6 * it + 3 + d];
                bmin[d] = l < bmin[d] ? l : bmin[d];
                bmax[d] = h > bmax[d] ? h : bmax[d];
            }
    } else {
        for (int i = g.lo + lane; i <= g.hi; i += 64) {
            const double v[3] = {x[i], y[i], z[i]};
            for (int d = 0; d < 3; ++d) {
                bmin[d] = v[d] < bmin[d] ? v[d] : bmin[d];
                bmax[d] = v[d] > bmax[d] ? v[d] : bmax[d];
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int d = 0; d < 3; ++d) {
            const double l = __shfl_xor(bmin[d], o), h = __shfl_xor(bmax[d], o);
            bmin[d] = l < bmin[d] ? l : bmin[d];
            bmax[d] = h > bmax[d] ? h : bmax[d];
        }
    if (lane) return;
    bool extent = false;
    nch[r] = 0;
    for (int d = 0; d < 3; ++d) extent |= bmax[d] > bmin[d];
    g.dim = -1;
    g.split = 0.0;
    g.ch0 = g.ch1 = -1;
    g.depth = depth;
    if (g.hi - g.lo + 1 > nBucket && extent) {
        if (g.hi - g.lo + 1 <= LMAX) {   // split by k_gt_local
            g.dim = -2;
            return;
        }
        int d = 0;
        for (int k = 1; k < 3; ++k)
            if (bmax[k] - bmin[k] > bmax[d] - bmin[d]) d = k;
        g.dim = d;
        g.split = 0.5 * (bmin[d] + bmax[d]);
        nch[r] = 2;
    }
}

// Child ids of the level's split cells: next + their offsets.
__global__ void k_gt_kids(int a, int b, int next, const int64_t *__restrict__ choff,
                          GNode *__restrict__ nd)
{
    const int c = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b) return;
    GNode &g = nd[c];
    if (g.dim < 0) return;
    g.ch0 = next + (int)choff[c - a];
    g.ch1 = g.ch0 + 1;
}

// The children's records once the split positions are known.
__global__ void k_gt_children(int a, int b, GNode *__restrict__ nd)
{
    const int c = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b) return;
    const GNode &g = nd[c];
    if (g.dim < 0) return;
    GNode &l = nd[g.ch0], &u = nd[g.ch1];
    l.lo = g.lo;
    l.hi = g.s - 1;
    u.lo = g.s;
    u.hi = g.hi;
    l.par = u.par = c;
    l.side = 0;
    u.side = 1;
    l.root = u.root = g.root;
}

// Subtree sizes and preorder ids, over the nodes of one depth (list).
__global__ void k_gt_size(int m, const int *__restrict__ list, GNode *__restrict__ nd)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    GNode &g = nd[list[k]];
    g.size = 1 + (g.ch0 >= 0 ? nd[g.ch0].size : 0) + (g.ch1 >= 0 ? nd[g.ch1].size : 0);
}

__global__ void k_gt_pre(int m, const int *__restrict__ list, GNode *__restrict__ nd)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const GNode &g = nd[list[k]];
    if (g.ch0 >= 0) nd[g.ch0].pre = g.pre + 1;
    if (g.ch1 >= 0) nd[g.ch1].pre = g.pre + 1 + (g.ch0 >= 0 ? nd[g.ch0].size : 0);
}

// Nodes by depth: a count per depth, then each node to its depth's list.
// Each workgroup counts in LDS
// first, so the global counters see one atomic per workgroup and depth.
constexpr int DMAX = 2048;       // depths a tree may have (level loop + local builds)

__global__ __launch_bounds__(256) void k_gt_depth_count(int nn, int ndep,
                                                        const GNode *__restrict__ nd,
                                                        int *__restrict__ cnt)
{
    __shared__ int h[DMAX];
    for (int d = threadIdx.x; d < ndep; d += 256) h[d] = 0;
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c < nn) atomicAdd(&h[nd[c].depth], 1);
    __syncthreads();
    for (int d = threadIdx.x; d < ndep; d += 256)
        if (h[d]) atomicAdd(&cnt[d], h[d]);
}

__global__ __launch_bounds__(256) void k_gt_depth_list(int nn, int ndep,
                                                       const GNode *__restrict__ nd,
                                                       int *__restrict__ cur,
                                                       int *__restrict__ list)
{
    __shared__ int h[DMAX], b0[DMAX];
    for (int d = threadIdx.x; d < ndep; d += 256) h[d] = 0;
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    const bool ok = c < nn;
    const int dep = ok ? nd[c].depth : 0;
    const int r = ok ? atomicAdd(&h[dep], 1) : 0;
    __syncthreads();
    for (int d = threadIdx.x; d < ndep; d += 256)
        if (h[d]) b0[d] = atomicAdd(&cur[d], h[d]);
    __syncthreads();
    if (ok) list[b0[dep] + r] = c;
}

// Cells marked for the local build, and the scratch node slots each
// reserves for its descendants (2 count - 2 at most).
__global__ void k_gt_lmark(int nn, const GNode *__restrict__ nd, int *__restrict__ flag,
                           int *__restrict__ res)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nn) return;
    const bool l = nd[c].dim == -2;
    flag[c] = l;
    res[c] = l ? 2 * (nd[c].hi - nd[c].lo + 1) - 2 : 0;
}

__global__ void k_gt_lcompact(int nn, const int *__restrict__ flag, const int64_t *__restrict__ pos,
                              int *__restrict__ list)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nn && flag[c]) list[pos[c]] = c;
}

// The subtree of one marked cell (at most LMAX particles), built in LDS by
// one workgroup with the rules of the level loop: per level, each cell's
// bounds and split decision (a wavefront per cell), child ids in order,
// then each split cell's partition by the Hoare pairing of the k-th
// misplaced particles from either end (a wavefront per cell).  The cell is
// local node 0; descendant k >= 1 (in creation order) goes to scratch slot
// lres[root] + k - 1 with its links as local ids, and *used receives the
// node count; k_gt_lcopy then moves them into place.  The particles are
// written back in their new order.
__global__ __launch_bounds__(256) void k_gt_local(const int *__restrict__ lroot,
                                                  const int64_t *__restrict__ lres, int nBucket,
                                                  GNode *__restrict__ nd, GNode *__restrict__ scr,
                                                  int *__restrict__ used,
                                                  double *__restrict__ px, double *__restrict__ py,
                                                  double *__restrict__ pz, int *__restrict__ perm,
                                                  int *__restrict__ err)
{
#pragma clang fp contract(off)
    __shared__ double lx[LMAX], ly[LMAX], lz[LMAX];
    __shared__ int ls[LMAX];
    __shared__ short posR[LMAX];
    __shared__ short qa[2][LMAX];              // cells of the current / next level
    __shared__ short nlo[2 * LMAX], nhi[2 * LMAX];   // cell ranges (local positions)
    __shared__ short coff[LMAX], cs[LMAX];     // child offset (-1: none), split position
    __shared__ signed char cdim[LMAX];
    __shared__ double csplit[LMAX];
    __shared__ int wsum[4];
    __shared__ int nsplit_sh;
    const int rootId = lroot[blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int glo = nd[rootId].lo, cnt = nd[rootId].hi - glo + 1;
    const int troot = nd[rootId].root, rdepth = nd[rootId].depth;
    GNode *const sb = scr + lres[rootId] - 1;   // local node k >= 1 at sb[k]
    for (int i = threadIdx.x; i < cnt; i += 256) {
        lx[i] = px[glo + i];
        ly[i] = py[glo + i];
        lz[i] = pz[glo + i];
        ls[i] = perm[glo + i];
    }
    if (threadIdx.x == 0) {
        nlo[0] = 0;
        nhi[0] = (short)(cnt - 1);
        qa[0][0] = 0;
    }
    __syncthreads();
    int nq = 1, cur = 0, nloc = 1, depth = 0;
    while (nq > 0) {
        // bounds, decision and lower-part count, a wavefront per cell
        for (int t = w; t < nq; t += 4) {
            const int k = qa[cur][t];
            const int lo = nlo[k], hi = nhi[k];
            double bmin[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
            double bmax[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
            for (int i = lo + lane; i <= hi; i += 64) {
                const double v[3] = {lx[i], ly[i], lz[i]};
                for (int d = 0; d < 3; ++d) {
                    bmin[d] = v[d] < bmin[d] ? v[d] : bmin[d];
                    bmax[d] = v[d] > bmax[d] ? v[d] : bmax[d];
                }
            }
            for (int o = 32; o > 0; o >>= 1)
                for (int d = 0; d < 3; ++d) {
                    const double l = __shfl_xor(bmin[d], o), h = __shfl_xor(bmax[d], o);
                    bmin[d] = l < bmin[d] ? l : bmin[d];
                    bmax[d] = h > bmax[d] ? h : bmax[d];
                }
            bool extent = false;
            for (int d = 0; d < 3; ++d) extent |= bmax[d] > bmin[d];
            int dim = -1;
            double split = 0.0;
            if (hi - lo + 1 > nBucket && extent) {
                int d = 0;
                for (int q = 1; q < 3; ++q)
                    if (bmax[q] - bmin[q] > bmax[d] - bmin[d]) d = q;
                dim = d;
                split = 0.5 * (bmin[d] + bmax[d]);
            }
            int s = lo;
            if (dim >= 0) {
                const double *cc = dim == 0 ? lx : (dim == 1 ? ly : lz);
                int nl = 0;
                for (int c0 = lo; c0 <= hi; c0 += 64) {
                    const int i = c0 + lane;
                    nl += __popcll(__ballot(i <= hi && cc[i] < split));
                }
                s = lo + nl;
                if (s == lo || s > hi) {   // one side empty: reported, left a bucket
                    if (lane == 0) atomicOr(err, 1);
                    dim = -1;
                    s = lo;
                }
            }
            if (lane == 0) {
                GNode &g = k == 0 ? nd[rootId] : sb[k];
                g.lo = glo + lo;
                g.hi = glo + hi;
                g.dim = dim;
                g.split = dim >= 0 ? split : 0.0;
                g.ch0 = g.ch1 = -1;
                g.s = glo + s;
                g.root = troot;
                g.depth = rdepth + depth;
                coff[t] = dim >= 0 ? 1 : 0;
                cdim[t] = (signed char)dim;
                csplit[t] = split;
                cs[t] = (short)s;
            }
        }
        __syncthreads();
        {   // child offsets: exclusive sum of the split flags
            int v[4], sum = 0;
            for (int q = 0; q < 4; ++q) {
                const int t = threadIdx.x * 4 + q;
                v[q] = t < nq ? coff[t] : 0;
                sum += v[q];
            }
            int inc = sum;
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(inc, o);
                if (lane >= o) inc += u;
            }
            if (lane == 63) wsum[w] = inc;
            __syncthreads();
            int ex = inc - sum, tot = 0;
            for (int q = 0; q < 4; ++q) {
                if (q < w) ex += wsum[q];
                tot += wsum[q];
            }
            for (int q = 0; q < 4; ++q) {
                const int t = threadIdx.x * 4 + q;
                if (t < nq) coff[t] = v[q] ? (short)ex : (short)-1;
                ex += v[q];
            }
            if (threadIdx.x == 0) nsplit_sh = tot;
        }
        __syncthreads();
        const int nsplit = nsplit_sh;
        // partitions and children, a wavefront per split cell
        for (int t = w; t < nq; t += 4) {
            const int co = coff[t];
            if (co < 0) continue;
            const int k = qa[cur][t];
            const int lo = nlo[k], hi = nhi[k], s = cs[t];
            const double split = csplit[t];
            const double *cc = cdim[t] == 0 ? lx : (cdim[t] == 1 ? ly : lz);
            const int nl = s - lo;
            int run = 0;
            for (int c0 = lo; c0 <= hi; c0 += 64) {   // upper positions holding a lower particle
                const int i = c0 + lane;
                const bool f = i <= hi && cc[i] < split;
                const unsigned long long m = __ballot(f);
                const int incl = run + __popcll(m & (lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1)));
                if (f && i >= s) posR[lo + (nl - incl)] = (short)i;
                run += __popcll(m);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            run = 0;
            for (int c0 = lo; c0 < s; c0 += 64) {     // lower positions holding an upper one
                const int i = c0 + lane;
                const bool in = i < s;
                const bool f = in && cc[i] < split;
                const unsigned long long m = __ballot(f);
                const int excl = run + __popcll(m & ((1ull << lane) - 1));
                if (in && !f) {
                    const int j = posR[lo + (i - lo) - excl];
                    double d = lx[i]; lx[i] = lx[j]; lx[j] = d;
                    d = ly[i]; ly[i] = ly[j]; ly[j] = d;
                    d = lz[i]; lz[i] = lz[j]; lz[j] = d;
                    const int u = ls[i]; ls[i] = ls[j]; ls[j] = u;
                }
                run += __popcll(m);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane == 0) {
                const int a0 = nloc + 2 * co, a1 = a0 + 1;
                nlo[a0] = (short)lo;
                nhi[a0] = (short)(s - 1);
                nlo[a1] = (short)s;
                nhi[a1] = (short)hi;
                qa[cur ^ 1][2 * co] = (short)a0;
                qa[cur ^ 1][2 * co + 1] = (short)a1;
                GNode &g = k == 0 ? nd[rootId] : sb[k];
                g.ch0 = a0;
                g.ch1 = a1;
                sb[a0].par = k;
                sb[a0].side = 0;
                sb[a1].par = k;
                sb[a1].side = 1;
            }
        }
        __syncthreads();
        nloc += 2 * nsplit;
        nq = 2 * nsplit;
        cur ^= 1;
        ++depth;
    }
    for (int i = threadIdx.x; i < cnt; i += 256) {
        px[glo + i] = lx[i];
        py[glo + i] = ly[i];
        pz[glo + i] = lz[i];
        perm[glo + i] = ls[i];
    }
    if (threadIdx.x == 0) used[blockIdx.x] = nloc - 1;
}

// The descendants of local build r from scratch to nodes nng + dbase[r] ...,
// links turned from local ids (0: the marked cell) to node ids.
__global__ __launch_bounds__(256) void k_gt_lcopy(const int *__restrict__ lroot,
                                                  const int64_t *__restrict__ lres,
                                                  const int *__restrict__ used,
                                                  const int64_t *__restrict__ dbase, int nng,
                                                  GNode *__restrict__ nd,
                                                  const GNode *__restrict__ scr)
{
    const int r = blockIdx.x, rootId = lroot[r], u = used[r];
    const int db = nng + (int)dbase[r];
    const GNode *const sb = scr + lres[rootId] - 1;
    auto fin = [&](int k) { return k < 0 ? -1 : (k == 0 ? rootId : db + k - 1); };
    for (int k = 1 + threadIdx.x; k <= u; k += 256) {
        GNode g = sb[k];
        g.ch0 = fin(g.ch0);
        g.ch1 = fin(g.ch1);
        g.par = fin(g.par);
        nd[db + k - 1] = g;
    }
    if (threadIdx.x == 0) {
        GNode &g = nd[rootId];
        g.ch0 = fin(g.ch0);
        g.ch1 = fin(g.ch1);
    }
}

// Storage-order particle fields: masses, softenings, type bits by source.
__global__ void k_gt_gather(int n, const int *__restrict__ src, const double *__restrict__ m,
                            const double *__restrict__ h, const uint32_t *__restrict__ bits,
                            double *__restrict__ pm, double *__restrict__ ph,
                            uint32_t *__restrict__ pact)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = src[i];
    pm[i] = m[s];
    ph[i] = h[s];
    pact[i] = bits[s];
}

// Mass, mass-weighted softening and centre of mass (pkd.c:2510-2576):
// buckets over their particles in storage order, cells from their
// children, lower then upper.  Deepest level first, one depth per launch.
__global__ void k_gt_mass(int m, const int *__restrict__ list, const GNode *__restrict__ nd,
                          const double *__restrict__ x, const double *__restrict__ y,
                          const double *__restrict__ z, const double *__restrict__ pm,
                          const double *__restrict__ ph, double *__restrict__ cm)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int c = list[k];
    const GNode &g = nd[c];
    double fm, M = 0.0, S = 0.0, R[3] = {0.0, 0.0, 0.0};
    if (g.dim >= 0) {
        for (int k = 0; k < 2; ++k) {
            const int ch = k ? g.ch1 : g.ch0;
            if (ch == -1) continue;
            const double *q = cm + 5 * (size_t)ch;
            fm = q[0];
            M += fm;
            S += fm * q[1];
            for (int d = 0; d < 3; ++d) R[d] += fm * q[2 + d];
        }
    } else {
        for (int i = g.lo; i <= g.hi; ++i) {
            fm = pm[i];
            M += fm;
            S += fm * ph[i];
            R[0] += fm * x[i];
            R[1] += fm * y[i];
            R[2] += fm * z[i];
        }
    }
    if (M > 0) {
        S /= M;
        for (int d = 0; d < 3; ++d) R[d] /= M;
    }
    double *o = cm + 5 * (size_t)c;
    o[0] = M;
    o[1] = S;
    for (int d = 0; d < 3; ++d) o[2 + d] = R[d];
}

// Cell of each moment item (a thread per cell writes its items).
__global__ void k_gt_itemcell(int nn, const int64_t *__restrict__ ioff, int *__restrict__ icell)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nn) return;
    for (int64_t k = ioff[c]; k < ioff[c + 1]; ++k) icell[k] = c;
}

// One particle's terms of the moments about (cx, cy, cz): raw Q (6),
// reduced O (10) and H (15), and its distance for Bmax.
template <int ORDER>
__device__ __forceinline__ void mom_accum(double mm, double dx, double dy, double dz, double *Q,
                                          double *O, double *H, double &Bmax)
{
#pragma clang fp contract(off)
    constexpr bool wantO = ORDER >= 3, wantH = ORDER >= 4;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double d1 = sqrt(d2);
    if (d1 > Bmax) Bmax = d1;
    if (wantH) {
        H[0] += mm * (dx * dx * dx * dx - 6.0 / 7.0 * d2 * (dx * dx - 0.1 * d2));
        H[1] += mm * (dx * dy * dy * dy - 3.0 / 7.0 * d2 * dx * dy);
        H[2] += mm * (dx * dx * dx * dy - 3.0 / 7.0 * d2 * dx * dy);
        H[3] += mm * (dy * dy * dy * dy - 6.0 / 7.0 * d2 * (dy * dy - 0.1 * d2));
        H[4] += mm * (dx * dx * dx * dz - 3.0 / 7.0 * d2 * dx * dz);
        H[5] += mm * (dy * dy * dy * dz - 3.0 / 7.0 * d2 * dy * dz);
        H[6] += mm * (dx * dx * dy * dy - 1.0 / 7.0 * d2 * (dx * dx + dy * dy - 0.2 * d2));
        H[7] += mm * (dx * dx * dy * dz - 1.0 / 7.0 * d2 * dy * dz);
        H[8] += mm * (dx * dy * dy * dz - 1.0 / 7.0 * d2 * dx * dz);
        H[9] += mm * (dx * dx * dz * dz - 1.0 / 7.0 * d2 * (dx * dx + dz * dz - 0.2 * d2));
        H[10] += mm * (dx * dy * dz * dz - 1.0 / 7.0 * d2 * dx * dy);
        H[11] += mm * (dx * dz * dz * dz - 3.0 / 7.0 * d2 * dx * dz);
        H[12] += mm * (dy * dy * dz * dz - 1.0 / 7.0 * d2 * (dy * dy + dz * dz - 0.2 * d2));
        H[13] += mm * (dy * dz * dz * dz - 3.0 / 7.0 * d2 * dy * dz);
        H[14] += mm * (dz * dz * dz * dz - 6.0 / 7.0 * d2 * (dz * dz - 0.1 * d2));
    }
    if (wantO) {
        O[0] += mm * (dx * dx * dx - 0.6 * d2 * dx);
        O[1] += mm * (dx * dy * dy - 0.2 * d2 * dx);
        O[2] += mm * (dx * dx * dy - 0.2 * d2 * dy);
        O[3] += mm * (dy * dy * dy - 0.6 * d2 * dy);
        O[4] += mm * (dx * dx * dz - 0.2 * d2 * dz);
        O[5] += mm * (dy * dy * dz - 0.2 * d2 * dz);
        O[6] += mm * dx * dy * dz;
        O[7] += mm * (dx * dz * dz - 0.2 * d2 * dx);
        O[8] += mm * (dy * dz * dz - 0.2 * d2 * dy);
        O[9] += mm * (dz * dz * dz - 0.6 * d2 * dz);
    }
    Q[0] += mm * dx * dx;
    Q[1] += mm * dy * dy;
    Q[2] += mm * dz * dz;
    Q[3] += mm * dx * dy;
    Q[4] += mm * dx * dz;
    Q[5] += mm * dy * dz;
}

// Partial moments of one item (one wavefront): raw Q (6), reduced O (10)
// and H (15) about the cell's centre of mass, and the largest distance.
// Per-lane sums, then a fixed butterfly: the result does not depend on
// scheduling.
template <int ORDER>
__global__ __launch_bounds__(256) void k_gt_moments(int nitem, const int *__restrict__ icell,
                                                    const int64_t *__restrict__ ioff,
                                                    const GNode *__restrict__ nd,
                                                    const double *__restrict__ cm,
                                                    const double *__restrict__ x,
                                                    const double *__restrict__ y,
                                                    const double *__restrict__ z,
                                                    const double *__restrict__ pm,
                                                    double *__restrict__ part)
{
#pragma clang fp contract(off)
    const int it = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (it >= nitem) return;
    const int c = icell[it];
    const GNode &g = nd[c];
    if (g.hi - g.lo + 1 <= MOM_SMALL) return;   // k_gt_moments_small's
    const int first = g.lo + (int)(it - ioff[c]) * MOM_CH;
    const int last = min(g.hi, first + MOM_CH - 1);
    const double cx = cm[5 * (size_t)c + 2], cy = cm[5 * (size_t)c + 3], cz = cm[5 * (size_t)c + 4];
    constexpr bool wantO = ORDER >= 3, wantH = ORDER >= 4;
    double Q[6] = {0, 0, 0, 0, 0, 0}, O[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double H[15] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double Bmax = 0.0;
    for (int i = first + lane; i <= last; i += 64) {
        const double mm = pm[i];
        const double dx = x[i] - cx, dy = y[i] - cy, dz = z[i] - cz;
        mom_accum<ORDER>(mm, dx, dy, dz, Q, O, H, Bmax);
    }
    for (int o = 32; o > 0; o >>= 1) {
        for (int k = 0; k < 6; ++k) Q[k] += __shfl_xor(Q[k], o);
        if (wantO)
            for (int k = 0; k < 10; ++k) O[k] += __shfl_xor(O[k], o);
        if (wantH)
            for (int k = 0; k < 15; ++k) H[k] += __shfl_xor(H[k], o);
        Bmax = fmax(Bmax, __shfl_xor(Bmax, o));
    }
    if (lane == 0) {
        double *p = part + (size_t)NM * it;
        for (int k = 0; k < 6; ++k) p[k] = Q[k];
        for (int k = 0; k < 10; ++k) p[6 + k] = O[k];
        for (int k = 0; k < 15; ++k) p[16 + k] = H[k];
        p[31] = Bmax;
    }
}

// Cells of at most MOM_SMALL particles (most cells: the buckets and the
// levels just above them): one thread each sums its cell's particles in
// order, with no cross-lane reduction.  Each such cell is one item.
template <int ORDER>
__global__ __launch_bounds__(256) void k_gt_moments_small(int nitem, const int *__restrict__ icell,
                                                          const GNode *__restrict__ nd,
                                                          const double *__restrict__ cm,
                                                          const double *__restrict__ x,
                                                          const double *__restrict__ y,
                                                          const double *__restrict__ z,
                                                          const double *__restrict__ pm,
                                                          double *__restrict__ part)
{
#pragma clang fp contract(off)
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nitem) return;
    const int c = icell[it];
    const GNode &g = nd[c];
    if (g.hi - g.lo + 1 > MOM_SMALL) return;
    const double cx = cm[5 * (size_t)c + 2], cy = cm[5 * (size_t)c + 3], cz = cm[5 * (size_t)c + 4];
    double Q[6] = {0, 0, 0, 0, 0, 0}, O[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double H[15] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double Bmax = 0.0;
    for (int i = g.lo; i <= g.hi; ++i) {
        const double mm = pm[i];
        const double dx = x[i] - cx, dy = y[i] - cy, dz = z[i] - cz;
        mom_accum<ORDER>(mm, dx, dy, dz, Q, O, H, Bmax);
    }
    double *p = part + (size_t)NM * it;
    for (int k = 0; k < 6; ++k) p[k] = Q[k];
    for (int k = 0; k < 10; ++k) p[6 + k] = O[k];
    for (int k = 0; k < 15; ++k) p[16 + k] = H[k];
    p[31] = Bmax;
}

// Per cell in preorder: the parity record (iDim, pLower, pUpper, iLower,
// iNext, 0; mass, softening, c.o.m., fOpen2, Bmax, fSplit), the walk record,
// the cell record and O/H: moments summed over the cell's items in order,
// traceless Q as the walk stores it (walk.c:41-48), fOpen2 by pkdCalcOpen
// (pkd.c:2228-2264).
// Cells of more than one moment item (one wavefront per cell): lane j sums
// the items j, j + 64, ..., a fixed butterfly combines the lanes, and the
// total replaces the cell's first item.
__global__ __launch_bounds__(256) void k_gt_combine(int nn, int nv,
                                                    const int64_t *__restrict__ ioff,
                                                    double *__restrict__ part)
{
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (b >= nn) return;
    const int64_t i0 = ioff[b], i1 = ioff[b + 1];
    if (i1 - i0 < 2) return;
    double M[31];
    for (int k = 0; k < 31; ++k) M[k] = 0.0;
    double Bmax = 0.0;
    for (int64_t it = i0 + lane; it < i1; it += 64) {
        const double *p = part + (size_t)NM * it;
        for (int k = 0; k < nv; ++k) M[k] += p[k];
        Bmax = fmax(Bmax, p[31]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        for (int k = 0; k < nv; ++k) M[k] += __shfl_xor(M[k], o);
        Bmax = fmax(Bmax, __shfl_xor(Bmax, o));
    }
    if (lane) return;
    double *p = part + (size_t)NM * i0;
    for (int k = 0; k < nv; ++k) p[k] = M[k];
    p[31] = Bmax;
}

__global__ void k_gt_emit(int nn, int iOrder, int iOpenType, double dCrit,
                          const GNode *__restrict__ nd, const double *__restrict__ cm,
                          const int64_t *__restrict__ ioff, const double *__restrict__ part,
                          int *__restrict__ tint, double *__restrict__ tdbl,
                          WalkNode *__restrict__ wn, double *__restrict__ cell,
                          double *__restrict__ momOH)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nn) return;
    const GNode &g = nd[b];
    const int c = g.pre;
    const double *q = cm + 5 * (size_t)b;
    const double *M = part + (size_t)NM * ioff[b];
    const double Bmax = M[31];
    double dOpen;
    if (iOpenType == 1) {
        dOpen = 2 / sqrt(3.0) * Bmax / dCrit;
        if (dOpen < Bmax) dOpen = Bmax;
    } else {
        dOpen = Bmax;
    }
    const double fOpen2 = g.hi >= g.lo ? dOpen * dOpen : 0.0;
    const GNode &rt = nd[g.root];
    const int end = rt.pre + rt.size;
    const int first = g.ch0 >= 0 ? nd[g.ch0].pre : -1;
    const int second = g.ch1 >= 0 ? nd[g.ch1].pre : -1;
    const int next = c + g.size < end ? c + g.size : -1;
    int *ti = tint + 6 * (size_t)c;   // iLower: the first child is c + 1 in preorder
    ti[0] = g.dim;
    ti[1] = g.lo;
    ti[2] = g.hi;
    ti[3] = g.dim >= 0 ? c + 1 : -1;
    ti[4] = next;
    ti[5] = 0;
    double *td = tdbl + 8 * (size_t)c;
    for (int k = 0; k < 5; ++k) td[k] = q[k];
    td[5] = fOpen2;
    td[6] = Bmax;
    td[7] = g.split;
    WalkNode w;
    w.cx = q[2];
    w.cy = q[3];
    w.cz = q[4];
    w.fOpen2 = fOpen2;
    w.fSoft = q[1];
    w.pLower = g.lo;
    w.pUpper = g.hi;
    w.iLower = first;
    w.iNext = next;
    w.iUpper = second;
    w.pad1 = 0;
    wn[c] = w;
    double *r = cell + (size_t)CR_N * c;
    r[CR_M] = q[0];
    for (int d = 0; d < 3; ++d) r[CR_X + d] = q[2 + d];
    r[CR_H] = q[1];
    for (int k = 0; k < 6; ++k) r[CR_QR + k] = M[k];
    const double tr = M[0] + M[1] + M[2];
    r[CR_QT] = M[0] - tr / 3.0;
    r[CR_QT + 1] = M[1] - tr / 3.0;
    r[CR_QT + 2] = M[2] - tr / 3.0;
    r[CR_QT + 3] = M[3];
    r[CR_QT + 4] = M[4];
    r[CR_QT + 5] = M[5];
    if (iOrder >= 3)
        for (int k = 0; k < 25; ++k) momOH[(size_t)OH_N * c + k] = M[6 + k];
}

// Buckets with an active particle: flag by preorder id; bounds of the
// active particles, largest softening, active count (pkd.c:2916-2937,
// walk.c:319-324); the largest bucket, active or not (the fused walk maps a
// bucket's targets, and an opened bucket's particles, onto one wavefront).
__global__ void k_gt_bucket(int nn, const GNode *__restrict__ nd, const double *__restrict__ x,
                            const double *__restrict__ y, const double *__restrict__ z,
                            const double *__restrict__ ph, const uint32_t *__restrict__ pact,
                            int *__restrict__ flag, double *__restrict__ bb8,
                            int *__restrict__ maxb)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nn) return;
    const GNode &g = nd[b];
    const int c = g.pre;
    if (g.dim >= 0 || g.lo > g.hi) {
        flag[c] = 0;
        return;
    }
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    int na = 0;
    double fsm = 0.0;
    for (int i = g.lo; i <= g.hi; ++i) {
        if (ph[i] > fsm) fsm = ph[i];
        if (!(pact[i] & 1u)) continue;
        ++na;
        const double r[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; ++d) {
            if (r[d] < lo[d]) lo[d] = r[d];
            if (r[d] > hi[d]) hi[d] = r[d];
        }
    }
    flag[c] = na > 0;
    atomicMax(maxb, g.hi - g.lo + 1);
    double *o = bb8 + 8 * (size_t)c;
    for (int d = 0; d < 3; ++d) {
        o[d] = lo[d];
        o[3 + d] = hi[d];
    }
    o[6] = fsm;
    o[7] = (double)na;
}

__global__ void k_gt_bucket_list(int nn, const int *__restrict__ flag,
                                 const int64_t *__restrict__ pos, const double *__restrict__ bb8,
                                 const int *__restrict__ rootOf, int *__restrict__ bucket,
                                 double *__restrict__ bbox, int *__restrict__ broot)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nn || !flag[c]) return;
    const int k = (int)pos[c];
    bucket[k] = c;
    for (int d = 0; d < 8; ++d) bbox[8 * (size_t)k + d] = bb8[8 * (size_t)c + d];
    broot[k] = rootOf[c];
}

__global__ void k_gt_rootof(int nn, const GNode *__restrict__ nd, int *__restrict__ rootOf)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nn) return;
    rootOf[nd[b].pre] = nd[nd[b].root].pre;
}

__global__ void k_gt_pos2root(int n, int nr, const int *__restrict__ rlo,
                              int *__restrict__ pos2root)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nr - 1;   // last root whose first position <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (rlo[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    pos2root[i] = lo;
}

__global__ void k_gt_roots(int nr, const int *__restrict__ rlo, const int *__restrict__ rhi,
                           GNode *__restrict__ nd)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr) return;
    GNode &g = nd[k];
    g.lo = rlo[k];
    g.hi = rhi[k];
    g.par = -1;
    g.side = -1;
    g.root = k;
    g.ch0 = g.ch1 = -1;
    g.dim = 0;      // partitioned by type first
}

}  // namespace

// One partition pass over the cells [a, b) (mode 0: their splits, and every
// position moves to its child; mode 1: type bits & mask, every cell).
static int gt_partition(GGPKD pkd, int a, int b, int mode, uint32_t mask)
{
    GravDev &G = pkd->grav;
    hipStream_t st = pkd->stream;
    const int n = pkd->nLocal;
    GNode *nd = G.gnode.as<GNode>();
    FlagFn fl;
    fl.pnode = G.gpn.as<int>();
    fl.nd = nd;
    fl.x = G.px.as<double>();
    fl.y = G.py.as<double>();
    fl.z = G.pz.as<double>();
    fl.src = G.perm.as<int>();
    fl.bits = pkd->f[GG_IACTIVE].as<uint32_t>();
    fl.a = a;
    fl.b = b;
    fl.mode = mode;
    fl.mask = mask;
    GG_TRY(exclusive_scan(st, fl, G.gF.as<int64_t>(), n, G.scan, nullptr));
    hipLaunchKernelGGL(k_gt_count, dim3(nblk(b - a, 256)), dim3(256), 0, st, a, b, mode,
                       G.gF.as<int64_t>(), nd, G.gerr.as<int>());
    hipLaunchKernelGGL(k_gt_rank, dim3(nblk(n, 256)), dim3(256), 0, st, n, fl,
                       G.gF.as<int64_t>(), G.gposR.as<int>());
    hipLaunchKernelGGL(k_gt_swap, dim3(nblk(n, 256)), dim3(256), 0, st, n, fl, mode == 0 ? 1 : 0,
                       G.gF.as<int64_t>(), G.gposR.as<int>(), G.px.as<double>(),
                       G.py.as<double>(), G.pz.as<double>(), G.perm.as<int>(), G.gpn.as<int>());
    GG_HIP(hipGetLastError());
    return GG_OK;
}

int gtree_build(GGPKD pkd, int nSys, const int *sysOff, int nBucket, int iOpenType, double dCrit,
                int iOrder)
{
    GravDev &G = pkd->grav;
    hipStream_t st = pkd->stream;
    const int n = pkd->nLocal;
    G.built = 0;
    // roots: the non-empty systems
    std::vector<int> rlo;
    for (int s = 0; s < nSys; ++s)
        if (sysOff[s + 1] > sysOff[s]) rlo.push_back(sysOff[s]);
    const int nr = (int)rlo.size();
    for (int k = 0; k < nr; ++k) rlo.push_back(k + 1 < nr ? rlo[k + 1] - 1 : n - 1);
    const size_t n1 = (size_t)std::max(n, 1);
    const size_t cap = 2 * n1 + 2;   // cells at most: every split has two children
    GG_TRY(G.px.reserve(sizeof(double) * n1));
    GG_TRY(G.py.reserve(sizeof(double) * n1));
    GG_TRY(G.pz.reserve(sizeof(double) * n1));
    GG_TRY(G.pm.reserve(sizeof(double) * n1));
    GG_TRY(G.ph.reserve(sizeof(double) * n1));
    GG_TRY(G.pact.reserve(sizeof(uint32_t) * n1));
    GG_TRY(G.perm.reserve(sizeof(int) * n1));
    GG_TRY(G.acc.reserve(sizeof(double) * 5 * n1));
    GG_TRY(G.gpn.reserve(sizeof(int) * n1));
    GG_TRY(G.gF.reserve(sizeof(int64_t) * (n1 + 1)));
    GG_TRY(G.gposR.reserve(sizeof(int) * n1));
    GG_TRY(G.gnode.reserve(sizeof(GNode) * cap));
    GG_TRY(G.gbox.reserve(sizeof(double) * 6 * (2 * (n1 / BOX_CH) + 2)));
    GG_TRY(G.gnit.reserve(sizeof(int) * cap));
    GG_TRY(G.gioff.reserve(sizeof(int64_t) * (cap + 1)));
    GG_TRY(G.gnch.reserve(sizeof(int) * cap));
    GG_TRY(G.gchoff.reserve(sizeof(int64_t) * (cap + 1)));
    GG_TRY(G.gerr.reserve(sizeof(int) * 2));
    GG_TRY(G.groot.reserve(sizeof(int) * 2 * std::max(nr, 1)));
    GG_HIP(hipMemsetAsync(G.gerr.p, 0, sizeof(int) * 2, st));
    GNode *nd = G.gnode.as<GNode>();
    int nn = 0;
    std::vector<std::pair<int, int>> levels, depths;
    if (nr > 0) {
        const int *rl = G.groot.as<int>(), *rh = rl + nr;
        GG_HIP(hipMemcpyAsync(G.groot.p, rlo.data(), sizeof(int) * 2 * nr, hipMemcpyHostToDevice,
                              st));
        hipLaunchKernelGGL(k_gt_roots, dim3(nblk(nr, 256)), dim3(256), 0, st, nr, rl, rh, nd);
        // every position starts in its system's root (the offsets cover the store)
        hipLaunchKernelGGL(k_gt_pos2root, dim3(nblk(n, 256)), dim3(256), 0, st, n, nr, rl,
                           G.gpn.as<int>());
        hipLaunchKernelGGL(k_gt_init, dim3(nblk(n, 256)), dim3(256), 0, st, n,
                           pkd->f[GG_R].as<double>(), G.px.as<double>(), G.py.as<double>(),
                           G.pz.as<double>(), G.perm.as<int>());
        GG_HIP(hipGetLastError());
        // msrBuildTree's active-type orders (master.c:4262), then the tree
        GG_TRY(gt_partition(pkd, 0, nr, 1, GG_TYPE_ACTIVE | GG_TYPE_TREEACTIVE));
        GG_TRY(gt_partition(pkd, 0, nr, 1, GG_TYPE_TREEACTIVE));
        int a = 0, b = nr;
        while (b > a) {
            levels.push_back({a, b});
            // bounds by work items of BOX_CH particles (at most one per cell
            // plus n / BOX_CH), then the split decisions
            hipLaunchKernelGGL(k_gt_items, dim3(nblk(b - a, 256)), dim3(256), 0, st, a, b, BOX_CH,
                               BOX_CH, nd, G.gnit.as<int>());
            GG_TRY(exclusive_scan_i64(st, G.gnit.as<int>(), G.gioff.as<int64_t>(), b - a, G.scan,
                                      nullptr));
            hipLaunchKernelGGL(k_gt_box, dim3(nblk(2 * (int64_t)(n / BOX_CH) + 1, 4)), dim3(256),
                               0, st, a, b, G.gioff.as<int64_t>(), nd, G.px.as<double>(),
                               G.py.as<double>(), G.pz.as<double>(), G.gbox.as<double>());
            hipLaunchKernelGGL(k_gt_decide, dim3(nblk(b - a, 4)), dim3(256), 0, st, a, b, nBucket,
                               (int)levels.size() - 1, G.gioff.as<int64_t>(), G.gbox.as<double>(),
                               G.px.as<double>(),
                               G.py.as<double>(), G.pz.as<double>(), nd, G.gnch.as<int>());
            // child ids (two per split), then the partition, which also moves
            // every position to its child
            int64_t nch = 0;
            GG_TRY(exclusive_scan_i64(st, G.gnch.as<int>(), G.gchoff.as<int64_t>(), b - a,
                                      G.scan, &nch));
            int err = 0;
            GG_HIP(hipMemcpyAsync(&err, G.gerr.p, sizeof(int), hipMemcpyDeviceToHost, st));
            GG_HIP(hipStreamSynchronize(st));
            GG_CHECK(!err, GG_EUNSUP,
                     "a cell split leaves one side empty (positions differ in the last bit only)");
            GG_CHECK((size_t)b + (size_t)nch <= cap, GG_EHIP, "tree cell capacity exceeded");
            if (nch > 0) {
                hipLaunchKernelGGL(k_gt_kids, dim3(nblk(b - a, 256)), dim3(256), 0, st, a, b, b,
                                   G.gchoff.as<int64_t>(), nd);
                GG_TRY(gt_partition(pkd, a, b, 0, 0));
                hipLaunchKernelGGL(k_gt_children, dim3(nblk(b - a, 256)), dim3(256), 0, st, a, b,
                                   nd);
            }
            GG_HIP(hipGetLastError());
            a = b;
            b += (int)nch;
        }
        const int nng = b;
        // cells of at most LMAX particles: their subtrees in one launch, each
        // in node slots reserved after the level loop's cells
        GG_TRY(G.gbflag.reserve(sizeof(int) * (size_t)nng));
        hipLaunchKernelGGL(k_gt_lmark, dim3(nblk(nng, 256)), dim3(256), 0, st, nng, nd,
                           G.gbflag.as<int>(), G.gnit.as<int>());
        int64_t nloc = 0, nres = 0;
        GG_TRY(exclusive_scan_i64(st, G.gbflag.as<int>(), G.gchoff.as<int64_t>(), nng, G.scan,
                                  &nloc));
        GG_TRY(exclusive_scan_i64(st, G.gnit.as<int>(), G.gioff.as<int64_t>(), nng, G.scan, &nres));
        GG_CHECK((size_t)nng + (size_t)nres <= cap, GG_EHIP, "tree cell capacity exceeded");
        nn = nng;
        if (nloc > 0) {
            GG_TRY(G.glroot.reserve(sizeof(int) * (size_t)nloc));
            GG_TRY(G.glused.reserve(sizeof(int) * (size_t)nloc));
            GG_TRY(G.glbase.reserve(sizeof(int64_t) * ((size_t)nloc + 1)));
            GG_TRY(G.gscr.reserve(sizeof(GNode) * (size_t)std::max<int64_t>(nres, 1)));
            hipLaunchKernelGGL(k_gt_lcompact, dim3(nblk(nng, 256)), dim3(256), 0, st, nng,
                               G.gbflag.as<int>(), G.gchoff.as<int64_t>(), G.glroot.as<int>());
            hipLaunchKernelGGL(k_gt_local, dim3((unsigned)nloc), dim3(256), 0, st,
                               G.glroot.as<int>(), G.gioff.as<int64_t>(), nBucket, nd,
                               G.gscr.as<GNode>(), G.glused.as<int>(), G.px.as<double>(),
                               G.py.as<double>(), G.pz.as<double>(), G.perm.as<int>(),
                               G.gerr.as<int>());
            GG_HIP(hipGetLastError());
            int64_t nused = 0;
            GG_TRY(exclusive_scan_i64(st, G.glused.as<int>(), G.glbase.as<int64_t>(), (int)nloc,
                                      G.scan, &nused));
            int err = 0;
            GG_HIP(hipMemcpyAsync(&err, G.gerr.p, sizeof(int), hipMemcpyDeviceToHost, st));
            GG_HIP(hipStreamSynchronize(st));
            GG_CHECK(!err, GG_EUNSUP,
                     "a cell split leaves one side empty (positions differ in the last bit only)");
            nn = nng + (int)nused;
            hipLaunchKernelGGL(k_gt_lcopy, dim3((unsigned)nloc), dim3(256), 0, st,
                               G.glroot.as<int>(), G.gioff.as<int64_t>(), G.glused.as<int>(),
                               G.glbase.as<int64_t>(), nng, nd, G.gscr.as<GNode>());
            GG_HIP(hipGetLastError());
        }
        // nodes by depth (the level loop's levels, then up to LMAX local ones)
        const int ndep = (int)levels.size() + LMAX + 1;
        GG_CHECK(ndep <= DMAX, GG_EUNSUP, "tree deeper than %d levels", DMAX);
        GG_TRY(G.gdcnt.reserve(sizeof(int) * 2 * (size_t)ndep));
        GG_TRY(G.gdlist.reserve(sizeof(int) * (size_t)nn));
        GG_HIP(hipMemsetAsync(G.gdcnt.p, 0, sizeof(int) * ndep, st));
        hipLaunchKernelGGL(k_gt_depth_count, dim3(nblk(nn, 256)), dim3(256), 0, st, nn, ndep, nd,
                           G.gdcnt.as<int>());
        std::vector<int> dcnt((size_t)ndep), doff((size_t)ndep + 1, 0);
        GG_HIP(hipMemcpyAsync(dcnt.data(), G.gdcnt.p, sizeof(int) * ndep, hipMemcpyDeviceToHost,
                              st));
        GG_HIP(hipStreamSynchronize(st));
        int nd_used = 0;
        for (int d = 0; d < ndep; ++d) {
            doff[d + 1] = doff[d] + dcnt[d];
            if (dcnt[d]) nd_used = d + 1;
        }
        GG_HIP(hipMemcpyAsync(G.gdcnt.as<int>() + ndep, doff.data(), sizeof(int) * ndep,
                              hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_gt_depth_list, dim3(nblk(nn, 256)), dim3(256), 0, st, nn, ndep, nd,
                           G.gdcnt.as<int>() + ndep, G.gdlist.as<int>());
        depths.clear();
        for (int d = 0; d < nd_used; ++d) depths.push_back({doff[d], dcnt[d]});
        const int *dl = G.gdlist.as<int>();
        // preorder numbering from subtree sizes
        for (int d = (int)depths.size() - 1; d >= 0; --d)
            if (depths[d].second)
                hipLaunchKernelGGL(k_gt_size, dim3(nblk(depths[d].second, 256)), dim3(256), 0, st,
                                   depths[d].second, dl + depths[d].first, nd);
        std::vector<GNode> roots((size_t)nr);
        GG_HIP(hipMemcpyAsync(roots.data(), nd, sizeof(GNode) * nr, hipMemcpyDeviceToHost, st));
        GG_HIP(hipStreamSynchronize(st));
        int pre = 0;
        for (int k = 0; k < nr; ++k) {
            roots[k].pre = pre;
            pre += roots[k].size;
        }
        GG_HIP(hipMemcpyAsync(nd, roots.data(), sizeof(GNode) * nr, hipMemcpyHostToDevice, st));
        for (size_t d = 0; d < depths.size(); ++d)
            if (depths[d].second)
                hipLaunchKernelGGL(k_gt_pre, dim3(nblk(depths[d].second, 256)), dim3(256), 0, st,
                                   depths[d].second, dl + depths[d].first, nd);
        GG_HIP(hipGetLastError());
        GG_HIP(hipStreamSynchronize(st));
    }
    const bool empty = nn == 0;
    const int nNodes = empty ? 1 : nn;
    // storage-order particle fields
    if (n > 0)
        hipLaunchKernelGGL(k_gt_gather, dim3(nblk(n, 256)), dim3(256), 0, st, n,
                           G.perm.as<int>(), pkd->f[GG_MASS].as<double>(),
                           pkd->f[GG_SOFT].as<double>(), pkd->f[GG_IACTIVE].as<uint32_t>(),
                           G.pm.as<double>(), G.ph.as<double>(), G.pact.as<uint32_t>());
    GG_TRY(G.wnode.reserve(sizeof(WalkNode) * nNodes));
    GG_TRY(G.cell.reserve(sizeof(double) * CR_N * nNodes));
    GG_TRY(G.momOH.reserve(sizeof(double) * OH_N * (iOrder >= 3 ? nNodes : 1)));
    GG_TRY(G.tint.reserve(sizeof(int) * 6 * nNodes));
    GG_TRY(G.tdbl.reserve(sizeof(double) * 8 * nNodes));
    GG_TRY(G.gcm.reserve(sizeof(double) * 5 * nNodes));
    int nb = 0;
    if (empty) {   // an empty store: one empty bucket (pkd.c:2684-2712)
        const int ti[6] = {-1, 0, -1, -1, -1, 0};
        const double td[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        WalkNode w{};
        w.pLower = 0;
        w.pUpper = -1;
        w.iLower = w.iNext = w.iUpper = -1;
        GG_HIP(hipMemcpyAsync(G.tint.p, ti, sizeof ti, hipMemcpyHostToDevice, st));
        GG_HIP(hipMemcpyAsync(G.tdbl.p, td, sizeof td, hipMemcpyHostToDevice, st));
        GG_HIP(hipMemcpyAsync(G.wnode.p, &w, sizeof w, hipMemcpyHostToDevice, st));
        GG_HIP(hipStreamSynchronize(st));
    } else {
        // mass, softening, c.o.m.: deepest level first
        for (int d = (int)depths.size() - 1; d >= 0; --d)
            if (depths[d].second)
                hipLaunchKernelGGL(k_gt_mass, dim3(nblk(depths[d].second, 256)), dim3(256), 0, st,
                                   depths[d].second, G.gdlist.as<int>() + depths[d].first, nd,
                                   G.px.as<double>(), G.py.as<double>(), G.pz.as<double>(),
                                   G.pm.as<double>(), G.ph.as<double>(), G.gcm.as<double>());
        // moments: work items of MOM_CH particles per cell
        hipLaunchKernelGGL(k_gt_items, dim3(nblk(nn, 256)), dim3(256), 0, st, 0, nn, MOM_CH, 0,
                           nd, G.gnit.as<int>());
        int64_t nitem = 0;
        GG_TRY(exclusive_scan_i64(st, G.gnit.as<int>(), G.gioff.as<int64_t>(), nn, G.scan, &nitem));
        GG_TRY(G.gpart.reserve(sizeof(double) * NM * std::max<int64_t>(nitem, 1)));
        GG_TRY(G.gicell.reserve(sizeof(int) * std::max<int64_t>(nitem, 1)));
        hipLaunchKernelGGL(k_gt_itemcell, dim3(nblk(nn, 256)), dim3(256), 0, st, nn,
                           G.gioff.as<int64_t>(), G.gicell.as<int>());
        if (nitem > 0) {
            auto mom = iOrder >= 4 ? k_gt_moments<4> : (iOrder >= 3 ? k_gt_moments<3>
                                                                      : k_gt_moments<2>);
            hipLaunchKernelGGL(mom, dim3(nblk(nitem, 4)), dim3(256), 0, st, (int)nitem,
                               G.gicell.as<int>(), G.gioff.as<int64_t>(), nd, G.gcm.as<double>(),
                               G.px.as<double>(), G.py.as<double>(), G.pz.as<double>(),
                               G.pm.as<double>(), G.gpart.as<double>());
            auto small = iOrder >= 4 ? k_gt_moments_small<4>
                                     : (iOrder >= 3 ? k_gt_moments_small<3> : k_gt_moments_small<2>);
            hipLaunchKernelGGL(small, dim3(nblk(nitem, 256)), dim3(256), 0, st, (int)nitem,
                               G.gicell.as<int>(), nd, G.gcm.as<double>(), G.px.as<double>(),
                               G.py.as<double>(), G.pz.as<double>(), G.pm.as<double>(),
                               G.gpart.as<double>());
        }
        const int nv = iOrder >= 4 ? 31 : (iOrder >= 3 ? 16 : 6);
        hipLaunchKernelGGL(k_gt_combine, dim3(nblk(nn, 4)), dim3(256), 0, st, nn, nv,
                           G.gioff.as<int64_t>(), G.gpart.as<double>());
        hipLaunchKernelGGL(k_gt_emit, dim3(nblk(nn, 256)), dim3(256), 0, st, nn, iOrder,
                           iOpenType, dCrit, nd, G.gcm.as<double>(), G.gioff.as<int64_t>(),
                           G.gpart.as<double>(), G.tint.as<int>(), G.tdbl.as<double>(),
                           G.wnode.as<WalkNode>(), G.cell.as<double>(), G.momOH.as<double>());
        // active buckets in preorder
        GG_TRY(G.gbflag.reserve(sizeof(int) * nn));
        GG_TRY(G.gbb8.reserve(sizeof(double) * 8 * nn));
        GG_TRY(G.grootof.reserve(sizeof(int) * nn));
        hipLaunchKernelGGL(k_gt_bucket, dim3(nblk(nn, 256)), dim3(256), 0, st, nn, nd,
                           G.px.as<double>(), G.py.as<double>(), G.pz.as<double>(),
                           G.ph.as<double>(), G.pact.as<uint32_t>(), G.gbflag.as<int>(),
                           G.gbb8.as<double>(), G.gerr.as<int>() + 1);
        hipLaunchKernelGGL(k_gt_rootof, dim3(nblk(nn, 256)), dim3(256), 0, st, nn, nd,
                           G.grootof.as<int>());
        int64_t nb64 = 0;
        GG_TRY(exclusive_scan_i64(st, G.gbflag.as<int>(), G.gchoff.as<int64_t>(), nn, G.scan,
                                  &nb64));
        nb = (int)nb64;
        int maxb = 0;
        GG_HIP(hipMemcpyAsync(&maxb, G.gerr.as<int>() + 1, sizeof(int), hipMemcpyDeviceToHost,
                              st));
        GG_HIP(hipStreamSynchronize(st));
        G.maxBucket = maxb;     // > 64 only for cells of coincident particles (pkd.c:2489)
        GG_TRY(G.bucket.reserve(sizeof(int) * std::max(nb, 1)));
        GG_TRY(G.broot.reserve(sizeof(int) * std::max(nb, 1)));
        GG_TRY(G.bbox.reserve(sizeof(double) * 8 * std::max(nb, 1)));
        hipLaunchKernelGGL(k_gt_bucket_list, dim3(nblk(nn, 256)), dim3(256), 0, st, nn,
                           G.gbflag.as<int>(), G.gchoff.as<int64_t>(), G.gbb8.as<double>(),
                           G.grootof.as<int>(), G.bucket.as<int>(), G.bbox.as<double>(),
                           G.broot.as<int>());
        GG_HIP(hipGetLastError());
    }
    GG_TRY(G.cnt.reserve(sizeof(int) * 3 * std::max(nb, 1)));
    GG_TRY(G.off.reserve(sizeof(int64_t) * (3 * (size_t)std::max(nb, 1) + 1)));
    GG_HIP(hipStreamSynchronize(st));
    G.nb = nb;
    G.n = n;
    G.nNodes = nNodes;
    G.iRoot = 0;
    G.nSys = nSys;
    G.iOrder = iOrder;
    G.nBucket = nBucket;
    G.built = 1;
    G.listsCurrent = 0;
    ++G.treeId;
    return GG_OK;
}

}  // namespace gg
