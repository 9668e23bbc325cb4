// Internal declarations shared by the translation units of libgasoline_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/gasoline_amd.h"

namespace gg {

void set_error(const char *fmt, ...);

#define GG_HIP(call)                                                        \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) {                                             \
            ::gg::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,      \
                            hipGetErrorString(e_));                         \
            return GG_EHIP;                                                 \
        }                                                                   \
    } while (0)

#define GG_CHECK(cond, code, ...)                                           \
    do {                                                                    \
        if (!(cond)) {                                                      \
            ::gg::set_error(__VA_ARGS__);                                   \
            return (code);                                                  \
        }                                                                   \
    } while (0)

#define GG_TRY(expr)                                                        \
    do {                                                                    \
        int r_ = (expr);                                                    \
        if (r_ != GG_OK) return r_;                                         \
    } while (0)

// Device buffer that only grows; contents are not preserved on growth.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int reserve(size_t nbytes);
    void release();
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Per node, what the walk reads for a visited cell (64 bytes).
struct WalkNode {
    double cx, cy, cz, fOpen2, fSoft;
    int pLower, pUpper, iLower, iNext;
    int iUpper, pad1;       // second child (-1 for a bucket)
};

// Per node FP64 record read when a cell is on an interaction list: mass,
// centre of mass, softening, raw and traceless quadrupole.
constexpr int CR_M = 0, CR_X = 1, CR_H = 4, CR_QR = 5, CR_QT = 11, CR_N = 20;
constexpr int OH_N = 25;  // octupole (10) + hexadecapole (15)

// Gravity state on the device.
struct GravDev {
    int built = 0, n = 0, nNodes = 0, iRoot = -1, iOrder = 2, nBucket = 0;
    int nb = 0;            // active buckets
    int maxBucket = 0;     // particles in the largest bucket (> 64: coincident particles)
    int64_t nEnt = 0;      // list entries of the last walk
    DevBuf px, py, pz, pm, ph, pact, perm;    // storage-order particles
    DevBuf wnode;          // WalkNode per node
    DevBuf cell;           // CellRec per node
    DevBuf momOH;          // O,H per node (iOrder > 2)
    DevBuf bucket;         // active bucket ids
    DevBuf bbox;           // per active bucket: 6 bounds + fSoftMax + nActive
    DevBuf cnt, off;       // per active bucket: 3 counts, 3 offsets
    DevBuf ent;            // list entries
    DevBuf acc;            // per slot: ax ay az pot dtg (FP64)
    DevBuf stats;          // reduction scratch
    DevBuf scan;           // scan scratch
    DevBuf werr;           // walk frontier overflow flag
    DevBuf broot;          // per active bucket: the root of its tree
    DevBuf tint, tdbl;     // per node: 6 ints, 8 doubles (gg_tree_info's layout)
    // tree build scratch
    DevBuf gpn, gF, gposR, gnode, gbox, gnch, gchoff, gerr, groot, gcm;
    DevBuf gnit, gioff, gicell, gpart, gbflag, gbb8, grootof, glroot, gdcnt, gdlist, glused, glbase, gscr;
    DevBuf ewt, emom, ewh;          // Ewald: moment partials, root moments, k-space table
    DevBuf ewl;                     // Ewald: real-space terms per slot (for fWeight)
    int64_t ewLoops = 0;            // real-space terms of the last Ewald pass
    int nEwh = 0;                   // its k-space terms
    int nSys = 1;                   // systems (roots) of the tree
    hipGraph_t graph = nullptr;         // captured step (gg_grav_graph)
    hipGraphExec_t gexec = nullptr;
    int graphOrder = 0;
    long treeId = 0, graphTree = -1;
    int serialWalk = 0;    // the last listed walk fell back to one thread per bucket
    int listsCurrent = 0;  // cnt/off/ent hold the lists of the current tree
    int fusedWalk = 0;     // the last gg_grav_all walked inside the interaction kernel
    int sched = 0;         // fused walk: 0 one workgroup per bucket, 1 persistent queue
    DevBuf qorder, qbin, qhist, qnext;   // persistent queue (heaviest first), its scratch
    long qTree = -1;       // tree whose costs ordered qorder (-1: identity order)
    int qGrid = 0, qGridKey = 0;   // workgroups of the persistent grid (and its kernel)
    void release_all()
    {
        DevBuf *all[] = {&px, &py, &pz, &pm, &ph, &pact, &perm, &wnode, &cell, &momOH, &bucket,
                         &bbox, &cnt, &off, &ent, &acc, &stats, &scan, &werr, &broot, &tint,
                         &tdbl, &gpn, &gF, &gposR, &gnode, &gbox, &gnch, &gchoff, &gerr,
                         &groot, &gcm, &gnit, &gioff, &gicell, &gpart, &gbflag, &gbb8, &grootof, &glroot, &gdcnt, &gdlist, &glused, &glbase, &gscr,
                         &ewt, &emom, &ewh, &ewl, &qorder, &qbin, &qhist, &qnext};
        for (DevBuf *b : all) b->release();
    }
};

// SPH state on the device (search structure, lists and per-call scratch).
struct SphDev {
    int built = 0, n = 0, bPeriodic = 0, nb = 0, nBucket = 0;
    int lpp = 4;                               // lanes per particle of DenDVDX and the force pass
    double L[3] = {0, 0, 0};
    int listK = 0, listValid = 0, listQueries = 0;
    unsigned ties = 0, fails = 0, retried = 0, diag[4] = {0, 0, 0, 0}, passes = 0;
    int64_t cand = 0;                          // kNN candidates examined
    double seedSlack = 1.15;                   // first trial ball / fBall2 hint
    int gasMaskValid = 0;                      // gasMask holds whether any particle is TYPE_GAS
    int allSelected = 0;                       // with it: every particle goes into the structure
    uint32_t gasMask = 0;
    int64_t nExtra = 0;
    int nExtraDev = 0;                         // > 0: nExtra still in that many partials in nxd
    DevBuf key, order, k2, v2, sortTmp, sel, pos, gidx, part;
    DevBuf sx, sy, sz;                         // slot order positions
    DevBuf child, rng, sparse;                 // radix tree, bucket-box sparse table
    DevBuf bflag, bcode, bpar, bfirst, bcnt, bbox, bden, bof; // buckets
    DevBuf crec, grpn, grpr;                   // child records, kNN groups (size, radius^2)
    DevBuf nbr, fb2s, fb2knn, qfknn, off;      // kNN lists (slot ids), CSR
    DevBuf gcnt, goff, gent;                   // ball-gather CSR
    DevBuf rcnt, roff, rcur, ridx;             // extras lists
    DevBuf ecnt, eidx;                         // extras found by the density pass
    int exValid = 0;
    DevBuf retry, hrec;                        // kNN retries, force records
    DevBuf qf, act, nba, err, guess, diffcnt, nxd, qcnt;
    DevBuf m, vx, vy, vz, u, rho, pr2, c, bal, dif, dt;   // slot-order inputs
    DevBuf drec;                               // DenDVDX per-slot records
    DevBuf o_rho, o_divv, o_curl, o_dvds, o_divvt, o_bal, o_diff, res;
    std::vector<DevBuf> spares;                // gg_reorder's spare buffers, by capacity
    uint32_t *rb = nullptr;                    // pinned read-back words (error words, counts)
    size_t rbCap = 0;                          // words
    hipEvent_t rbEv = nullptr;                 // recorded behind the read-back copies
    void release_all()
    {
        for (DevBuf &b : spares) b.release();
        spares.clear();
        if (rb) (void)hipHostFree(rb);
        rb = nullptr;
        rbCap = 0;
        if (rbEv) (void)hipEventDestroy(rbEv);
        rbEv = nullptr;
        DevBuf *all[] = {&key, &order, &k2, &v2, &sortTmp, &sel, &pos, &gidx, &part,
                         &sx, &sy, &sz, &child, &rng, &sparse, &bflag, &bcode, &bpar, &bden,
                         &bfirst, &bcnt, &bbox, &bof, &crec, &grpn, &grpr, &nbr, &fb2s, &fb2knn, &qfknn, &off,
                         &gcnt, &goff, &gent, &rcnt, &roff, &rcur, &ridx, &ecnt, &eidx, &qf, &act, &nba,
                         &err, &guess, &diffcnt, &nxd, &qcnt, &retry, &hrec, &m, &vx, &vy, &vz, &u, &rho, &pr2, &c,
                         &bal, &dif, &dt, &drec, &o_rho, &o_divv, &o_curl, &o_dvds, &o_divvt,
                         &o_bal, &o_diff, &res};
        for (DevBuf *b : all) b->release();
    }
};

// h-proposal MLP state on the device.
constexpr int HMLP_H = 32;                                  // hidden width
constexpr int HMLP_NW = HMLP_H * 4 + HMLP_H + HMLP_H * HMLP_H + HMLP_H + HMLP_H + 1;   // 1249
constexpr int HMLP_NP = HMLP_NW + 8;                        // + input mean, scale
static_assert(HMLP_NW == GG_HMLP_NW && HMLP_NP == GG_HMLP_NP, "header constants");
struct HmlpDev {
    int init = 0, normSet = 0, nHist = 0, step = 0;
    DevBuf par, m, v, grad, gpart, lpart;   // parameters, Adam moments, gradients
    DevBuf hx, hh, hv;                      // history: features, log h, valid (caller order)
    DevBuf sel, pos, samp, scan, stat;      // sample list, scratch
    void release_all()
    {
        DevBuf *all[] = {&par, &m, &v, &grad, &gpart, &lpart, &hx, &hh, &hv,
                         &sel, &pos, &samp, &scan, &stat};
        for (DevBuf *b : all) b->release();
    }
};

// Pinned host staging of gg_put_aos / gg_get_aos: two buffers, one being
// packed or unpacked while the other's copies run.
struct AosStage {
    void *buf[2] = {nullptr, nullptr};
    size_t bytes = 0;
    hipEvent_t ev[2];
    int created = 0;
    int pending[2] = {0, 0};
};

// Phase timers from HIP events on the handle's stream.  Each API call marks
// its phases in a frame of its own (a ring of NFR frames), and its elapsed
// times are read only when someone asks for them (gg_timings, the tuners)
// or when its frame comes round again, so no call waits for its own work
// just to time it.  ms[] then holds what the last call of each phase took.
struct Timers {
    static constexpr int NFR = 8, NEV = 16, NPEND = 32;
    hipEvent_t ev[NFR][NEV];
    int created = 0, frame = 0;
    double ms[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    struct Pend {
        int frame, first, n;
        int phase[8];
    };
    Pend pend[NPEND];
    int npend = 0;
};


// Workgroups are dealt to the eight XCDs in turn (workgroup i to XCD i mod 8),
// each XCD with its own L2.  This maps the dispatch index to a logical block
// so that every XCD serves one contiguous range of logical blocks: with the
// work in Morton order the data an XCD reads stays in its own L2.  A
// bijection of [0, nblocks) whatever nblocks is.
__device__ __forceinline__ int xcd_block(int bid, int nblocks)
{
    const int q = nblocks >> 3, r = nblocks & 7, x = bid & 7;
    return x * q + (x < r ? x : r) + (bid >> 3);
}

// The same dealing in chunks of c logical blocks: XCD x serves chunks x,
// x + 8, x + 16, ... (blocks past the last whole round of 8 chunks keep their
// index).  Neighbouring blocks still share an L2 while uneven work (a dense
// core in Morton order) spreads over all eight XCDs.  A bijection of
// [0, nblocks).
__device__ __forceinline__ int xcd_chunk(int bid, int nblocks, int c)
{
    const int whole = nblocks / (8 * c) * (8 * c);
    if (bid >= whole) return bid;
    const int x = bid & 7, j = bid >> 3;
    return ((j / c) * 8 + x) * c + j % c;
}

}  // namespace gg

struct gg_pkd_s {
    int device = 0;
    int nStore = 0, nLocal = 0;
    int prec = GG_PREC_FP32;
    double fPeriod[3] = {0, 0, 0};  // periodic box (gg_set_period); 0: not set
    hipStream_t stream = nullptr;
    gg::DevBuf f[GG_NFIELDS];      // caller-order fields
    gg::GravDev grav;
    gg::SphDev sph;
    gg::HmlpDev mlp;
    gg::Timers tm;
    gg::AosStage aos;
    gg::DevBuf red;                // reductions of the elementwise updates
};

namespace gg {
int field_width(int field);                 // values per particle
int field_is32(int field);                  // 32-bit integer field (iActive, iRung)
int ensure_field(GGPKD pkd, int field);     // allocate (zeroed) if missing
// out[0..n]: exclusive prefix sums, out[n] = total (out holds n+1 entries);
// the total is copied to *total_host (after a stream synchronisation) unless
// total_host is null.
int exclusive_scan_i64(hipStream_t s, const int *in, int64_t *out, int n,
                       DevBuf &scratch, int64_t *total_host);
// Gravity tree of the systems [sysOff[s], sysOff[s+1]) of the store, one
// root per non-empty system (gg_build_binary, gg_build_forest).
int gtree_build(GGPKD pkd, int nSys, const int *sysOff, int nBucket, int iOpenType, double dCrit,
                int iOrder);
void sph_release(GGPKD pkd);
// Ewald correction of the active particles added to the gravity slots
// (gg_ewald.hip); *flop = its real-space flop count.
int ewald_apply(GGPKD pkd, int nReps, int iOrder, double L, double fEwCut, double fEwhCut,
                double *flop);
void timer_mark(GGPKD pkd, int slot);
// The device address of gg_direct_grav_noalloc's fixed scratch array on
// `device` (the current device), looked up once per device (gg_initialize).
int direct_scratch_setup(int device);
void timer_collect(GGPKD pkd, int first_slot, int nphases, const int *phase);
void timer_collect_now(GGPKD pkd, int first_slot, int nphases, const int *phase);
void timer_resolve(GGPKD pkd);
}  // namespace gg
