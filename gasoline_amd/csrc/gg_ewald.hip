// Ewald summation for periodic gravity (pkdGravAll with bEwald,
// pkd.c:2953-2962): the root's moments (pkdCalcRoot pkd.c:4395-4470,
// pkdDistribRoot :4472-4493), the k-space table (pkdEwaldInit
// ewald.c:180-247, built on the host from those moments) and the
// correction of each active particle (pkdBucketEwald ewald.c:15-178), one
// thread per particle in FP64.  The replicas -nReps..nReps themselves are
// walked by the interaction kernel; the correction subtracts their
// multipole part (-erf in the hole) and adds the rest of the lattice.
#include "gg_internal.h"

#include <cmath>
#include <vector>

#pragma clang fp contract(off)

namespace gg {
namespace {

int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

constexpr int RM_T = 256;     // threads of the moment reduction
constexpr int RM_B = 256;     // its blocks (fixed: the summation order is too)

// Complete octupole (10) and hexadecapole (15) about the root's centre of
// mass, in the order xxx xyy xxy yyy xxz yyz xyz xzz yzz zzz, xxxx xyyy xxxy
// yyyy xxxz yyyz xxyy xxyz xyyz xxzz xyzz xzzz yyzz yzzz zzzz; per block.
__global__ __launch_bounds__(RM_T) void k_root_moments(int n, const double *__restrict__ px,
                                                       const double *__restrict__ py,
                                                       const double *__restrict__ pz,
                                                       const double *__restrict__ pm,
                                                       const double *__restrict__ cell,
                                                       double *__restrict__ part)
{
    __shared__ double sm[RM_T];
    double v[25];
    for (int k = 0; k < 25; ++k) v[k] = 0;
    const double cx = cell[CR_X], cy = cell[CR_X + 1], cz = cell[CR_X + 2];
    for (int i = blockIdx.x * RM_T + threadIdx.x; i < n; i += RM_T * RM_B) {
        const double m = pm[i], dx = px[i] - cx, dy = py[i] - cy, dz = pz[i] - cz;
        v[0] += m * (dx * dx * dx);
        v[1] += m * (dx * dy * dy);
        v[2] += m * (dx * dx * dy);
        v[3] += m * (dy * dy * dy);
        v[4] += m * (dx * dx * dz);
        v[5] += m * (dy * dy * dz);
        v[6] += m * dx * dy * dz;
        v[7] += m * (dx * dz * dz);
        v[8] += m * (dy * dz * dz);
        v[9] += m * (dz * dz * dz);
        v[10] += m * (dx * dx * dx * dx);
        v[11] += m * (dx * dy * dy * dy);
        v[12] += m * (dx * dx * dx * dy);
        v[13] += m * (dy * dy * dy * dy);
        v[14] += m * (dx * dx * dx * dz);
        v[15] += m * (dy * dy * dy * dz);
        v[16] += m * (dx * dx * dy * dy);
        v[17] += m * (dx * dx * dy * dz);
        v[18] += m * (dx * dy * dy * dz);
        v[19] += m * (dx * dx * dz * dz);
        v[20] += m * (dx * dy * dz * dz);
        v[21] += m * (dx * dz * dz * dz);
        v[22] += m * (dy * dy * dz * dz);
        v[23] += m * (dy * dz * dz * dz);
        v[24] += m * (dz * dz * dz * dz);
    }
    for (int k = 0; k < 25; ++k) {
        sm[threadIdx.x] = v[k];
        __syncthreads();
        for (int h = RM_T / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) sm[threadIdx.x] += sm[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[25 * (size_t)blockIdx.x + k] = sm[0];
        __syncthreads();
    }
}

__global__ void k_root_moments_sum(const double *__restrict__ part, double *__restrict__ out)
{
    const int k = threadIdx.x;
    if (k >= 25) return;
    double s = 0;
    for (int b = 0; b < RM_B; ++b) s += part[25 * (size_t)b + k];
    out[k] = s;
}

// The root's moments as ilcnRoot holds them.
struct EMom {
    double m, x, y, z;
    double Q[6];     // raw quadrupole xx yy zz xy xz yz
    double O[10], H[15];
};

// MEVAL (meval.h:21-80) of the complete multipole at (dx, dy, dz), falling
// through from ORDER to the monopole.
template <int ORDER>
__device__ __forceinline__ void meval(const EMom &M, const double *gam, double dx, double dy,
                                      double dz, double &ax, double &ay, double &az, double &fPot)
{
    const double xx = M.Q[0], yy = M.Q[1], zz = M.Q[2], xy = M.Q[3], xz = M.Q[4], yz = M.Q[5];
    double Qmirx, Qmiry, Qmirz, Qmir, Qta = 0.0;
    if constexpr (ORDER >= 4) {
        const double *H = M.H;
        const double xxxx = H[0], xyyy = H[1], xxxy = H[2], yyyy = H[3], xxxz = H[4],
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
        const double Qxx = xxxx + xxyy + xxzz, Qxy = xxxy + xyyy + xyzz, Qxz = xxxz + xyyz + xzzz;
        const double Qyy = xxyy + yyyy + yyzz, Qyz = xxyz + yyyz + yzzz, Qzz = xxzz + yyzz + zzzz;
        const double Qtr = (1.0 / 8.0) * (Qxx + Qyy + Qzz);
        const double Qhx = 0.5 * (Qxx * dx + Qxy * dy + Qxz * dz);
        const double Qhy = 0.5 * (Qxy * dx + Qyy * dy + Qyz * dz);
        const double Qhz = 0.5 * (Qxz * dx + Qyz * dy + Qzz * dz);
        const double Qh = 0.5 * (Qhx * dx + Qhy * dy + Qhz * dz);
        fPot -= gam[4] * Qmir - gam[3] * Qh + gam[2] * Qtr;
        Qta += gam[5] * Qmir - gam[4] * Qh + gam[3] * Qtr;
        ax += gam[4] * Qmirx - gam[3] * Qhx;
        ay += gam[4] * Qmiry - gam[3] * Qhy;
        az += gam[4] * Qmirz - gam[3] * Qhz;
    }
    if constexpr (ORDER >= 3) {
        const double *O = M.O;
        const double xxx = O[0], xyy = O[1], xxy = O[2], yyy = O[3], xxz = O[4], yyz = O[5],
                     xyz = O[6], xzz = O[7], yzz = O[8], zzz = O[9];
        Qmirx = (1.0 / 2.0) * (xzz * dz * dz + 2 * xyz * dy * dz + xyy * dy * dy
                + 2 * xxz * dx * dz + 2 * xxy * dx * dy + xxx * dx * dx);
        Qmiry = (1.0 / 2.0) * (yzz * dz * dz + 2 * xyz * dx * dz + xxy * dx * dx
                + 2 * yyz * dy * dz + 2 * xyy * dx * dy + yyy * dy * dy);
        Qmirz = (1.0 / 2.0) * (yyz * dy * dy + 2 * xyz * dx * dy + xxz * dx * dx
                + 2 * yzz * dy * dz + 2 * xzz * dx * dz + zzz * dz * dz);
        Qmir = (1.0 / 3.0) * (Qmirx * dx + Qmiry * dy + Qmirz * dz);
        const double Qx = 0.5 * (xxx + xyy + xzz), Qy = 0.5 * (xxy + yyy + yzz),
                     Qz = 0.5 * (xxz + yyz + zzz);
        const double Qtr = Qx * dx + Qy * dy + Qz * dz;
        fPot -= gam[3] * Qmir - gam[2] * Qtr;
        Qta += gam[4] * Qmir - gam[3] * Qtr;
        ax += gam[3] * Qmirx - gam[2] * Qx;
        ay += gam[3] * Qmiry - gam[2] * Qy;
        az += gam[3] * Qmirz - gam[2] * Qz;
    }
    if constexpr (ORDER >= 2) {
        Qmirx = (1.0 / 1.0) * (xz * dz + xy * dy + xx * dx);
        Qmiry = (1.0 / 1.0) * (yz * dz + xy * dx + yy * dy);
        Qmirz = (1.0 / 1.0) * (yz * dy + xz * dx + zz * dz);
        Qmir = (1.0 / 2.0) * (Qmirx * dx + Qmiry * dy + Qmirz * dz);
        const double Qtr = 0.5 * (xx + yy + zz);
        fPot -= gam[2] * Qmir - gam[1] * Qtr;
        Qta += gam[3] * Qmir - gam[2] * Qtr;
        ax += gam[2] * Qmirx;
        ay += gam[2] * Qmiry;
        az += gam[2] * Qmirz;
    }
    fPot -= gam[0] * M.m;
    Qta += gam[1] * M.m;
    ax -= dx * Qta;
    ay -= dy * Qta;
    az -= dz * Qta;
}

struct EArgs {
    int n, nReps, nEwReps, nEwh;
    double L, alpha, alpha2, k1, ka, fEwCut2;
    const double *px, *py, *pz;
    const uint32_t *pact;
    const EMom *mom;
    const double *ewt;          // nEwh x (hx, hy, hz, hCfac, hSfac)
    double *acc;                // per slot: ax ay az pot dtg (added to)
    unsigned long long *nLoop;
    int *slotLoops;             // per slot: real-space terms evaluated (fWeight)
};

// pkdBucketEwald (ewald.c:15-178) for the active particle of slot i.
template <int ORDER>
__global__ __launch_bounds__(256) void k_ewald(EArgs A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned loops = 0;
    if (i < A.n && (A.pact[i] & 1u)) {
        const EMom &M = *A.mom;
        const double L = A.L, alpha = A.alpha, alpha2 = A.alpha2, ka = A.ka;
        double fPot = M.m * A.k1, ax = 0.0, ay = 0.0, az = 0.0;
        const double dx = A.px[i] - M.x, dy = A.py[i] - M.y, dz = A.pz[i] - M.z;
        const int nE = A.nEwReps, nR = A.nReps;
        for (int ix = -nE; ix <= nE; ++ix) {
            const bool bInHolex = ix >= -nR && ix <= nR;
            const double dxo = dx + ix * L;
            for (int iy = -nE; iy <= nE; ++iy) {
                const bool bInHolexy = bInHolex && iy >= -nR && iy <= nR;
                const double dyo = dy + iy * L;
                for (int iz = -nE; iz <= nE; ++iz) {
                    const bool bInHole = bInHolexy && iz >= -nR && iz <= nR;
                    const double dzo = dz + iz * L;
                    double r2 = dxo * dxo + dyo * dyo + dzo * dzo;
                    if (r2 > A.fEwCut2 && !bInHole) continue;
                    double gam[6], alphan;
                    if (r2 < 3.0e-3 * L * L) {   // series about the origin
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
                        const double r = sqrt(r2), dir = 1 / r, dir2 = dir * dir;
                        double ea = exp(-r2 * alpha2);
                        ea *= ka * dir2;
                        gam[0] = bInHole ? -erf(alpha * r) : erfc(alpha * r);
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
                    meval<ORDER>(M, gam, dxo, dyo, dzo, ax, ay, az, fPot);
                    ++loops;
                }
            }
        }
        for (int k = 0; k < A.nEwh; ++k) {
            const double *e = A.ewt + 5 * (size_t)k;
            const double hdotx = e[0] * dx + e[1] * dy + e[2] * dz;
            const double c = cos(hdotx), s = sin(hdotx);
            fPot += e[3] * c + e[4] * s;
            ax += e[0] * (e[3] * s - e[4] * c);
            ay += e[1] * (e[3] * s - e[4] * c);
            az += e[2] * (e[3] * s - e[4] * c);
        }
        A.slotLoops[i] = (int)loops;
        double *o = A.acc + 5 * (size_t)i;
        o[0] += ax;
        o[1] += ay;
        o[2] += az;
        o[3] += fPot;
    }
    unsigned long long l = loops;
    for (int w = 32; w > 0; w >>= 1) l += __shfl_xor(l, w);
    if ((threadIdx.x & 63) == 0 && l) atomicAdd(A.nLoop, l);
}

// QEVAL (qeval.h:21-64) of the root moments, for the k-space table.
void qeval_host(int iOrder, const EMom &M, const double *gam, double dx, double dy, double dz,
                double &ax, double &ay, double &az, double &pot)
{
    const double xx = M.Q[0], yy = M.Q[1], zz = M.Q[2], xy = M.Q[3], xz = M.Q[4], yz = M.Q[5];
    const double *O = M.O, *H = M.H;
    double Qmirx, Qmiry, Qmirz, Qmir, Qta = 0.0;
    if (iOrder >= 4) {
        const double xxxx = H[0], xyyy = H[1], xxxy = H[2], yyyy = H[3], xxxz = H[4],
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
        pot -= gam[4] * Qmir;
        Qta += gam[5] * Qmir;
        ax += gam[4] * Qmirx;
        ay += gam[4] * Qmiry;
        az += gam[4] * Qmirz;
    }
    if (iOrder >= 3) {
        const double xxx = O[0], xyy = O[1], xxy = O[2], yyy = O[3], xxz = O[4], yyz = O[5],
                     xyz = O[6], xzz = O[7], yzz = O[8], zzz = O[9];
        Qmirx = (1.0 / 2.0) * (xzz * dz * dz + 2 * xyz * dy * dz + xyy * dy * dy
                + 2 * xxz * dx * dz + 2 * xxy * dx * dy + xxx * dx * dx);
        Qmiry = (1.0 / 2.0) * (yzz * dz * dz + 2 * xyz * dx * dz + xxy * dx * dx
                + 2 * yyz * dy * dz + 2 * xyy * dx * dy + yyy * dy * dy);
        Qmirz = (1.0 / 2.0) * (yyz * dy * dy + 2 * xyz * dx * dy + xxz * dx * dx
                + 2 * yzz * dy * dz + 2 * xzz * dx * dz + zzz * dz * dz);
        Qmir = (1.0 / 3.0) * (Qmirx * dx + Qmiry * dy + Qmirz * dz);
        pot -= gam[3] * Qmir;
        Qta += gam[4] * Qmir;
        ax += gam[3] * Qmirx;
        ay += gam[3] * Qmiry;
        az += gam[3] * Qmirz;
    }
    if (iOrder >= 2) {
        Qmirx = (1.0 / 1.0) * (xz * dz + xy * dy + xx * dx);
        Qmiry = (1.0 / 1.0) * (yz * dz + xy * dx + yy * dy);
        Qmirz = (1.0 / 1.0) * (yz * dy + xz * dx + zz * dz);
        Qmir = (1.0 / 2.0) * (Qmirx * dx + Qmiry * dy + Qmirz * dz);
        pot -= gam[2] * Qmir;
        Qta += gam[3] * Qmir;
        ax += gam[2] * Qmirx;
        ay += gam[2] * Qmiry;
        az += gam[2] * Qmirz;
    }
    pot -= gam[0] * M.m;
    Qta += gam[1] * M.m;
    ax -= dx * Qta;
    ay -= dy * Qta;
    az -= dz * Qta;
}

}  // namespace

// The Ewald correction of every active particle of the (single-tree) store,
// added to G.acc; *flop receives pkdBucketEwald's flop count summed.
int ewald_apply(GGPKD pkd, int nReps, int iOrder, double L, double fEwCut, double fEwhCut,
                double *flop)
{
    GravDev &G = pkd->grav;
    hipStream_t st = pkd->stream;
    *flop = 0;
    if (!iOrder) return GG_OK;           // ewald.c:39
    const int n = G.n;
    // the root's moments (its cell record holds mass, c.o.m. and raw Q)
    GG_TRY(G.ewt.reserve(sizeof(double) * 25 * RM_B));
    GG_TRY(G.emom.reserve(sizeof(EMom) + sizeof(double) * 25 + sizeof(unsigned long long)));
    double *rootCell = G.cell.as<double>() + (size_t)CR_N * G.iRoot;
    hipLaunchKernelGGL(k_root_moments, dim3(RM_B), dim3(RM_T), 0, st, n, G.px.as<double>(),
                       G.py.as<double>(), G.pz.as<double>(), G.pm.as<double>(), rootCell,
                       G.ewt.as<double>());
    double *mo = (double *)((char *)G.emom.p + sizeof(EMom));
    hipLaunchKernelGGL(k_root_moments_sum, dim3(1), dim3(64), 0, st, G.ewt.as<double>(), mo);
    GG_HIP(hipGetLastError());
    double rc[CR_N], oh[25];
    GG_HIP(hipMemcpyAsync(rc, rootCell, sizeof rc, hipMemcpyDeviceToHost, st));
    GG_HIP(hipMemcpyAsync(oh, mo, sizeof oh, hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    EMom M;
    M.m = rc[CR_M];
    M.x = rc[CR_X];
    M.y = rc[CR_X + 1];
    M.z = rc[CR_X + 2];
    for (int k = 0; k < 6; ++k) M.Q[k] = rc[CR_QR + k];
    for (int k = 0; k < 10; ++k) M.O[k] = oh[k];
    for (int k = 0; k < 15; ++k) M.H[k] = oh[10 + k];
    // k-space table (pkdEwaldInit)
    const int hReps = (int)std::ceil(fEwhCut);
    const double alpha = 2.0 / L, k4 = M_PI * M_PI / (alpha * alpha * L * L);
    std::vector<double> tab;
    for (int hx = -hReps; hx <= hReps; ++hx)
        for (int hy = -hReps; hy <= hReps; ++hy)
            for (int hz = -hReps; hz <= hReps; ++hz) {
                const int h2 = hx * hx + hy * hy + hz * hz;
                if (h2 == 0 || h2 > fEwhCut * fEwhCut) continue;
                double gam[6], ax = 0, ay = 0, az = 0, mfacc = 0, mfacs = 0;
                gam[0] = std::exp(-k4 * h2) / (M_PI * h2 * L);
                gam[1] = 2 * M_PI / L * gam[0];
                gam[2] = -2 * M_PI / L * gam[1];
                gam[3] = 2 * M_PI / L * gam[2];
                gam[4] = -2 * M_PI / L * gam[3];
                gam[5] = 2 * M_PI / L * gam[4];
                gam[1] = gam[3] = gam[5] = 0.0;
                qeval_host(iOrder, M, gam, hx, hy, hz, ax, ay, az, mfacc);
                gam[0] = std::exp(-k4 * h2) / (M_PI * h2 * L);
                gam[1] = 2 * M_PI / L * gam[0];
                gam[2] = -2 * M_PI / L * gam[1];
                gam[3] = 2 * M_PI / L * gam[2];
                gam[4] = -2 * M_PI / L * gam[3];
                gam[5] = 2 * M_PI / L * gam[4];
                gam[0] = gam[2] = gam[4] = 0.0;
                ax = ay = az = 0;
                qeval_host(iOrder, M, gam, hx, hy, hz, ax, ay, az, mfacs);
                const double e[5] = {2 * M_PI / L * hx, 2 * M_PI / L * hy, 2 * M_PI / L * hz,
                                     mfacc, mfacs};
                tab.insert(tab.end(), e, e + 5);
            }
    const int nEwh = (int)(tab.size() / 5);
    GG_TRY(G.ewh.reserve(sizeof(double) * std::max<size_t>(tab.size(), 1)));
    unsigned long long *nl = (unsigned long long *)((char *)G.emom.p + sizeof(EMom) +
                                                    sizeof(double) * 25);
    GG_HIP(hipMemcpyAsync(G.emom.p, &M, sizeof M, hipMemcpyHostToDevice, st));
    if (nEwh)
        GG_HIP(hipMemcpyAsync(G.ewh.p, tab.data(), sizeof(double) * tab.size(),
                              hipMemcpyHostToDevice, st));
    GG_HIP(hipMemsetAsync(nl, 0, sizeof(unsigned long long), st));
    EArgs A;
    A.n = n;
    A.nReps = nReps;
    const int ce = (int)std::ceil(fEwCut);
    A.nEwReps = ce > nReps ? ce : nReps;
    A.nEwh = nEwh;
    A.L = L;
    A.alpha = alpha;
    A.alpha2 = alpha * alpha;
    A.k1 = M_PI / (A.alpha2 * L * L * L);
    A.ka = 2.0 * alpha / std::sqrt(M_PI);
    A.fEwCut2 = fEwCut * fEwCut * L * L;
    A.px = G.px.as<double>();
    A.py = G.py.as<double>();
    A.pz = G.pz.as<double>();
    A.pact = G.pact.as<uint32_t>();
    A.mom = (const EMom *)G.emom.p;
    A.ewt = G.ewh.as<double>();
    A.acc = G.acc.as<double>();
    A.nLoop = nl;
    GG_TRY(G.ewl.reserve(sizeof(int) * (size_t)std::max(n, 1)));
    A.slotLoops = G.ewl.as<int>();
    const dim3 grid(nblk(n, 256)), block(256);
    switch (iOrder) {
    case 1: hipLaunchKernelGGL(k_ewald<1>, grid, block, 0, st, A); break;
    case 2: hipLaunchKernelGGL(k_ewald<2>, grid, block, 0, st, A); break;
    case 3: hipLaunchKernelGGL(k_ewald<3>, grid, block, 0, st, A); break;
    default: hipLaunchKernelGGL(k_ewald<4>, grid, block, 0, st, A); break;
    }
    GG_HIP(hipGetLastError());
    unsigned long long loops = 0;
    GG_HIP(hipMemcpyAsync(&loops, nl, sizeof loops, hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    static const int nMultiFlop[5] = {10, 10, 48, 151, 343};   // MEVAL_FLOP, meval.h:84
    G.ewLoops = (int64_t)loops;
    G.nEwh = nEwh;
    *flop = (double)loops * (104 + nMultiFlop[iOrder > 4 ? 4 : iOrder]);
    return GG_OK;
}

}  // namespace gg
