// Host-side gravity tree build (pure C++; no HIP headers).
#pragma once

#include <cstdint>
#include <vector>

namespace gg {

// Gravity tree on the host, in the tree's storage order (pkdBuildBinary).
struct HostGravTree {
    int n = 0, nNodes = 0, iRoot = -1, nBucket = 0, iOrder = 2;
    std::vector<int> perm;                       // slot -> caller index
    std::vector<double> x, y, z, m, h;           // storage order
    std::vector<uint32_t> act;                   // iActive bits, storage order
    // per node
    std::vector<int> iDim, pLower, pUpper, iLower, iNext;
    std::vector<double> fSplit, bmin, bmax;      // bmin/bmax: 3 per node
    std::vector<double> mass, soft, com, fOpen2; // com: 3 per node
    std::vector<double> mom;                     // MOM_N per node
    std::vector<double> bmaxr;                   // Bmax per node
};

// Moment record per node: Q raw (6), Q traceless (6), O (10), H (15).
constexpr int MOM_QR = 0, MOM_QT = 6, MOM_O = 12, MOM_H = 22, MOM_N = 37;

int build_binary_host(HostGravTree &t, int n, const double *r,
                      const double *m, const double *soft,
                      const uint32_t *iActive, int nBucket, int iOpenType,
                      double dCrit, int iOrder);

// Independent systems in one numbering: system s's particles are caller
// indices [off[s], off[s+1]); one tree per system, built in parallel, then
// concatenated (node and slot indices offset, each tree's threading ending in
// -1).  roots[s] is system s's root; nodeRoot[c] the root of node c's tree.
int build_forest_host(HostGravTree &t, std::vector<int> &roots, std::vector<int> &nodeRoot,
                      int nSys, const int *off, const double *r, const double *m,
                      const double *soft, const uint32_t *iActive, int nBucket,
                      int iOpenType, double dCrit, int iOrder);

}  // namespace gg
