// Host-side gravity tree build for libgasoline_amd.
//
// Produces the same tree as gasoline's pkdBuildBinary (pkd.c:2627-2723):
// the same particle partition order, preorder cell numbering, squeezed
// bounds, FP64 centre of mass and mass-weighted softening, multipole
// moments about the centre of mass and OPEN_JOSH opening radii.  Decisions
// the walk makes later compare these FP64 values, so the arithmetic follows
// the reference's operation order and the file is compiled with
// -ffp-contract=off.
//
// The structure differs from the reference: an explicit task stack instead
// of recursion, a compact 32-byte particle record for the partition swaps,
// children combined in one reverse-preorder sweep, moments computed with
// OpenMP over nodes, and threading from subtree sizes.
#include "gg_tree.h"

#include <cmath>
#include <cstring>
#include <cfloat>
#include <utility>

namespace gg {

namespace {

struct Rec {           // what the partition moves
    double r[3];
    int src;
    int pad;
};

// Hoare partition about fSplit on axis d, with the reference's scan and
// swap sequence (pkdUpperPart, pkd.c:1109-1157).  Returns the first index
// of the upper part.
int hoare_split(Rec *p, int d, double fSplit, int i, int j)
{
    if (i > j) return i;
    --i;
    ++j;
    while (++i < j && p[i].r[d] < fSplit) {
    }
    while (i < --j && p[j].r[d] >= fSplit) {
    }
    if (i >= j) return i;
    std::swap(p[i], p[j]);
    for (;;) {
        while (p[++i].r[d] < fSplit) {
        }
        while (p[--j].r[d] >= fSplit) {
        }
        if (i > j) return i;
        std::swap(p[i], p[j]);
    }
}

// Active-type-first Hoare partition (pkdActiveTypeOrder, pkd.c:1304-1355).
void type_first(Rec *p, const uint32_t *bits, int n, uint32_t mask)
{
    auto test = [&](int k) { return (bits[p[k].src] & mask) != 0; };
    int i = 0, j = n - 1;
    if (i > j) return;
    --i;
    ++j;
    while (++i < j && test(i)) {
    }
    while (i < --j && !test(j)) {
    }
    if (i >= j) return;
    std::swap(p[i], p[j]);
    for (;;) {
        while (test(++i)) {
        }
        while (!test(--j)) {
        }
        if (i > j) return;
        std::swap(p[i], p[j]);
    }
}

struct Task {
    int lo, hi, parent, upper;
};

}  // namespace

int build_binary_host(HostGravTree &t, int n, const double *r,
                      const double *m, const double *soft,
                      const uint32_t *iActive, int nBucket, int iOpenType,
                      double dCrit, int iOrder)
{
    const uint32_t TYPE_ACTIVE = 1u, TYPE_TREEACTIVE = 2u;
    std::vector<uint32_t> allbits;
    if (!iActive) {
        allbits.assign((size_t)n, TYPE_ACTIVE | TYPE_TREEACTIVE);
        iActive = allbits.data();
    }
    t = HostGravTree();
    t.n = n;
    t.nBucket = nBucket;
    t.iOrder = iOrder;
    std::vector<Rec> p((size_t)n);
    for (int i = 0; i < n; ++i) {
        p[i].r[0] = r[3 * i];
        p[i].r[1] = r[3 * i + 1];
        p[i].r[2] = r[3 * i + 2];
        p[i].src = i;
        p[i].pad = 0;
    }
    // msrBuildTree (master.c:4262) then pkdBuildBinary (pkd.c:2639)
    type_first(p.data(), iActive, n, TYPE_ACTIVE | TYPE_TREEACTIVE);
    type_first(p.data(), iActive, n, TYPE_TREEACTIVE);

    // --- structure: preorder numbering with an explicit stack
    std::vector<int> lower, upper;
    std::vector<Task> stack;
    if (n > 0) stack.push_back({0, n - 1, -1, 0});
    while (!stack.empty()) {
        Task tk = stack.back();
        stack.pop_back();
        int id = (int)t.iDim.size();
        if (tk.parent >= 0) (tk.upper ? upper : lower)[tk.parent] = id;
        double bmin[3], bmax[3];
        for (int d = 0; d < 3; ++d) bmin[d] = bmax[d] = p[tk.lo].r[d];
        for (int i = tk.lo + 1; i <= tk.hi; ++i)
            for (int d = 0; d < 3; ++d) {
                if (p[i].r[d] < bmin[d]) bmin[d] = p[i].r[d];
                else if (p[i].r[d] > bmax[d]) bmax[d] = p[i].r[d];
            }
        bool extent = false;
        for (int d = 0; d < 3; ++d) extent |= bmax[d] > bmin[d];
        t.pLower.push_back(tk.lo);
        t.pUpper.push_back(tk.hi);
        for (int d = 0; d < 3; ++d) {
            t.bmin.push_back(bmin[d]);
            t.bmax.push_back(bmax[d]);
        }
        lower.push_back(-1);
        upper.push_back(-1);
        if (tk.hi - tk.lo + 1 > nBucket && extent) {
            int d = 0;
            for (int k = 1; k < 3; ++k)
                if (bmax[k] - bmin[k] > bmax[d] - bmin[d]) d = k;
            double fSplit = 0.5 * (bmin[d] + bmax[d]);
            int mid = hoare_split(p.data(), d, fSplit, tk.lo, tk.hi);
            t.iDim.push_back(d);
            t.fSplit.push_back(fSplit);
            if (mid <= tk.hi) stack.push_back({mid, tk.hi, id, 1});
            if (tk.lo <= mid - 1) stack.push_back({tk.lo, mid - 1, id, 0});
        } else {
            t.iDim.push_back(-1);
            t.fSplit.push_back(0.0);
        }
    }
    int nn = (int)t.iDim.size();
    t.iRoot = 0;
    if (nn == 0) {  // empty store: one empty bucket (pkd.c:2684-2712)
        nn = 1;
        t.iDim.push_back(-1);
        t.fSplit.push_back(0.0);
        t.pLower.push_back(0);
        t.pUpper.push_back(-1);
        for (int d = 0; d < 3; ++d) {
            t.bmin.push_back(DBL_MAX);
            t.bmax.push_back(-DBL_MAX);
        }
        lower.push_back(-1);
        upper.push_back(-1);
    }
    t.nNodes = nn;

    // --- particles in storage order
    t.perm.resize(n);
    t.x.resize(n);
    t.y.resize(n);
    t.z.resize(n);
    t.m.resize(n);
    t.h.resize(n);
    t.act.resize(n);
    for (int i = 0; i < n; ++i) {
        int s = p[i].src;
        t.perm[i] = s;
        t.x[i] = p[i].r[0];
        t.y[i] = p[i].r[1];
        t.z[i] = p[i].r[2];
        t.m[i] = m[s];
        t.h[i] = soft[s];
        t.act[i] = iActive[s];
    }

    // --- mass, softening, centre of mass: buckets sum their particles in
    // storage order, cells combine lower then upper (BuildBinary,
    // pkd.c:2510-2576); children have larger preorder ids than parents.
    t.mass.assign(nn, 0.0);
    t.soft.assign(nn, 0.0);
    t.com.assign(3 * (size_t)nn, 0.0);
    for (int c = nn - 1; c >= 0; --c) {
        double fm, M = 0.0, S = 0.0, R[3] = {0.0, 0.0, 0.0};
        if (t.iDim[c] >= 0) {
            for (int k = 0; k < 2; ++k) {
                int ch = k ? upper[c] : lower[c];
                if (ch == -1) continue;
                fm = t.mass[ch];
                M += fm;
                S += fm * t.soft[ch];
                for (int d = 0; d < 3; ++d) R[d] += fm * t.com[3 * ch + d];
            }
        } else {
            for (int i = t.pLower[c]; i <= t.pUpper[c]; ++i) {
                fm = t.m[i];
                M += fm;
                S += fm * t.h[i];
                R[0] += fm * t.x[i];
                R[1] += fm * t.y[i];
                R[2] += fm * t.z[i];
            }
        }
        if (M > 0) {
            S /= M;
            for (int d = 0; d < 3; ++d) R[d] /= M;
        }
        t.mass[c] = M;
        t.soft[c] = S;
        for (int d = 0; d < 3; ++d) t.com[3 * c + d] = R[d];
    }

    // --- multipoles about the centre of mass (pkdCalcCell, pkd.c:2018-2135)
    // and opening radius (pkdCalcOpen, pkd.c:2228-2264).
    t.mom.assign((size_t)MOM_N * nn, 0.0);
    t.bmaxr.assign(nn, 0.0);
    t.fOpen2.assign(nn, 0.0);
    const bool wantO = iOrder >= 3, wantH = iOrder >= 4;
#pragma omp parallel for schedule(dynamic, 64)
    for (int c = 0; c < nn; ++c) {
        double *Q = &t.mom[(size_t)MOM_N * c + MOM_QR];
        double *O = &t.mom[(size_t)MOM_N * c + MOM_O];
        double *H = &t.mom[(size_t)MOM_N * c + MOM_H];
        const double cx = t.com[3 * c], cy = t.com[3 * c + 1], cz = t.com[3 * c + 2];
        double Bmax = 0.0;
        for (int i = t.pLower[c]; i <= t.pUpper[c]; ++i) {
            const double mm = t.m[i];
            const double dx = t.x[i] - cx, dy = t.y[i] - cy, dz = t.z[i] - cz;
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double d1 = std::sqrt(d2);
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
        // traceless quadrupole as the walk stores it (SETILIST, walk.c:41-48)
        double tr = Q[0] + Q[1] + Q[2];
        double *T = &t.mom[(size_t)MOM_N * c + MOM_QT];
        T[0] = Q[0] - tr / 3.0;
        T[1] = Q[1] - tr / 3.0;
        T[2] = Q[2] - tr / 3.0;
        T[3] = Q[3];
        T[4] = Q[4];
        T[5] = Q[5];
        double dOpen;
        if (iOpenType == 1) {
            dOpen = 2 / std::sqrt(3.0) * Bmax / dCrit;
            if (dOpen < Bmax) dOpen = Bmax;
        } else {
            dOpen = Bmax;
        }
        t.bmaxr[c] = Bmax;
        t.fOpen2[c] = dOpen * dOpen;
    }
    if (n == 0) t.fOpen2[0] = 0.0;

    // --- threading (pkdThreadTree, pkd.c:2590-2620): in preorder the first
    // child is c+1 and the next node after a subtree is c + size(c).
    std::vector<int> size(nn, 1);
    for (int c = nn - 1; c >= 0; --c) {
        if (lower[c] >= 0) size[c] += size[lower[c]];
        if (upper[c] >= 0) size[c] += size[upper[c]];
    }
    t.iLower.assign(nn, -1);
    t.iNext.assign(nn, -1);
    for (int c = 0; c < nn; ++c) {
        if (t.iDim[c] >= 0) t.iLower[c] = c + 1;
        int nx = c + size[c];
        t.iNext[c] = nx < nn ? nx : -1;
    }
    return 0;
}

int build_forest_host(HostGravTree &t, std::vector<int> &roots, std::vector<int> &nodeRoot,
                      int nSys, const int *off, const double *r, const double *m,
                      const double *soft, const uint32_t *iActive, int nBucket,
                      int iOpenType, double dCrit, int iOrder)
{
    std::vector<HostGravTree> part((size_t)nSys);
    std::vector<int> rc((size_t)nSys, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < nSys; ++s) {
        const int o = off[s], n = off[s + 1] - off[s];
        rc[s] = build_binary_host(part[s], n, r + 3 * (size_t)o, m + o, soft + o,
                                  iActive ? iActive + o : nullptr, nBucket, iOpenType, dCrit,
                                  iOrder);
    }
    for (int s = 0; s < nSys; ++s)
        if (rc[s]) return rc[s];
    t = HostGravTree();
    t.nBucket = nBucket;
    t.iOrder = iOrder;
    roots.assign((size_t)nSys, -1);
    nodeRoot.clear();
    int nb = 0, pb = 0;
    for (int s = 0; s < nSys; ++s) {
        const HostGravTree &q = part[s];
        roots[s] = q.nNodes ? nb + q.iRoot : -1;
        for (int i = 0; i < q.n; ++i) t.perm.push_back(q.perm[i] + off[s]);
        t.x.insert(t.x.end(), q.x.begin(), q.x.end());
        t.y.insert(t.y.end(), q.y.begin(), q.y.end());
        t.z.insert(t.z.end(), q.z.begin(), q.z.end());
        t.m.insert(t.m.end(), q.m.begin(), q.m.end());
        t.h.insert(t.h.end(), q.h.begin(), q.h.end());
        t.act.insert(t.act.end(), q.act.begin(), q.act.end());
        for (int c = 0; c < q.nNodes; ++c) {
            t.iDim.push_back(q.iDim[c]);
            t.pLower.push_back(q.pLower[c] + pb);
            t.pUpper.push_back(q.pUpper[c] + pb);
            t.iLower.push_back(q.iLower[c] >= 0 ? q.iLower[c] + nb : -1);
            t.iNext.push_back(q.iNext[c] >= 0 ? q.iNext[c] + nb : -1);
            nodeRoot.push_back(roots[s]);
        }
        t.fSplit.insert(t.fSplit.end(), q.fSplit.begin(), q.fSplit.end());
        t.bmin.insert(t.bmin.end(), q.bmin.begin(), q.bmin.end());
        t.bmax.insert(t.bmax.end(), q.bmax.begin(), q.bmax.end());
        t.mass.insert(t.mass.end(), q.mass.begin(), q.mass.end());
        t.soft.insert(t.soft.end(), q.soft.begin(), q.soft.end());
        t.com.insert(t.com.end(), q.com.begin(), q.com.end());
        t.fOpen2.insert(t.fOpen2.end(), q.fOpen2.begin(), q.fOpen2.end());
        t.mom.insert(t.mom.end(), q.mom.begin(), q.mom.end());
        t.bmaxr.insert(t.bmaxr.end(), q.bmaxr.begin(), q.bmaxr.end());
        nb += q.nNodes;
        pb += q.n;
    }
    t.n = pb;
    t.nNodes = nb;
    t.iRoot = nSys ? roots[0] : -1;
    return 0;
}

}  // namespace gg
