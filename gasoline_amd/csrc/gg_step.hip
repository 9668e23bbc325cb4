// The updates on either side of the force path, in FP64 on the device:
// pkdInitAccel (pkd.c:5443-5456), pkdKick (pkd.c:3780-3967, the NOCOOLING
// gas branch) and pkdDrift (pkd.c:3686-3778, without the sliding-patch and
// inflow/outflow options).  Elementwise over the store, caller order.
#include "gg_internal.h"

namespace gg {
namespace {

int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

__global__ void k_init_accel(int n, const uint32_t *__restrict__ bits, double *__restrict__ a,
                             double *__restrict__ pot, double *__restrict__ dtg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !(bits[i] & GG_TYPE_ACTIVE)) return;
    pot[i] = 0;
    dtg[i] = 0;
    for (int j = 0; j < 3; ++j) a[3 * (size_t)i + j] = 0;
}

struct KickArgs {
    int n, bGasU;
    double vf1, vf2, vpf1, vpf2, du, dup;
    const uint32_t *bits;
    const double *a, *udpdv, *udav, *uddif;
    double *v, *vpred, *u, *upred;
};

// Active particles: vPred = v dvPredFacOne + a dvPredFacTwo,
// v = v dvFacOne + a dvFacTwo; gas also uPred = u + uDot duPredDelta,
// u = u + uDot duDelta with uDot = uDotPdV + uDotAV + uDotDiff (pkd.h:65).
__global__ void k_kick(KickArgs K)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K.n) return;
    const uint32_t b = K.bits[i];
    if (!(b & GG_TYPE_ACTIVE)) return;
    for (int j = 0; j < 3; ++j) {
        const size_t e = 3 * (size_t)i + j;
        const double v = K.v[e], a = K.a[e];
        K.vpred[e] = v * K.vpf1 + a * K.vpf2;
        K.v[e] = v * K.vf1 + a * K.vf2;
    }
    if (K.bGasU && (b & GG_TYPE_GAS)) {
        const double ud = K.udpdv[i] + K.udav[i] + K.uddif[i];
        const double u = K.u[i];
        K.upred[i] = u + ud * K.dup;
        K.u[i] = u + ud * K.du;
    }
}

// r += dDelta v, wrapped into [c - L/2, c + L/2) when periodic.
__global__ void k_drift(int n, double dt, int per, double c0, double c1, double c2, double L0,
                        double L1, double L2, const double *__restrict__ v, double *__restrict__ r,
                        unsigned *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double c[3] = {c0, c1, c2}, L[3] = {L0, L1, L2};
    for (int j = 0; j < 3; ++j) {
        const size_t e = 3 * (size_t)i + j;
        double x = r[e] + dt * v[e];
        if (per) {
            if (x >= c[j] + 0.5 * L[j]) x -= L[j];
            if (x < c[j] - 0.5 * L[j]) x += L[j];
            if (!(x >= c[j] - 0.5 * L[j] && x < c[j] + 0.5 * L[j])) atomicAdd(out, 1u);
        }
        r[e] = x;
    }
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_init_accel(GGPKD pkd)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    const int n = pkd->nLocal;
    if (n == 0) return GG_OK;
    const int f[] = {GG_A, GG_POT, GG_DTGRAV, GG_IACTIVE};
    for (int k : f) GG_TRY(ensure_field(pkd, k));
    hipLaunchKernelGGL(k_init_accel, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n,
                       pkd->f[GG_IACTIVE].as<uint32_t>(), pkd->f[GG_A].as<double>(),
                       pkd->f[GG_POT].as<double>(), pkd->f[GG_DTGRAV].as<double>());
    GG_HIP(hipGetLastError());
    GG_HIP(hipStreamSynchronize(pkd->stream));
    return GG_OK;
}

extern "C" int gg_kick(GGPKD pkd, double dvFacOne, double dvFacTwo, double dvPredFacOne,
                       double dvPredFacTwo, double duDelta, double duPredDelta, int bIsothermal)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    const int n = pkd->nLocal;
    if (n == 0) return GG_OK;
    const int f[] = {GG_V, GG_VPRED, GG_A, GG_U, GG_UPRED, GG_UDOTPDV, GG_UDOTAV, GG_UDOTDIFF,
                     GG_IACTIVE};
    for (int k : f) GG_TRY(ensure_field(pkd, k));
    KickArgs K;
    K.n = n;
    K.bGasU = !bIsothermal;
    K.vf1 = dvFacOne;
    K.vf2 = dvFacTwo;
    K.vpf1 = dvPredFacOne;
    K.vpf2 = dvPredFacTwo;
    K.du = duDelta;
    K.dup = duPredDelta;
    K.bits = pkd->f[GG_IACTIVE].as<uint32_t>();
    K.a = pkd->f[GG_A].as<double>();
    K.udpdv = pkd->f[GG_UDOTPDV].as<double>();
    K.udav = pkd->f[GG_UDOTAV].as<double>();
    K.uddif = pkd->f[GG_UDOTDIFF].as<double>();
    K.v = pkd->f[GG_V].as<double>();
    K.vpred = pkd->f[GG_VPRED].as<double>();
    K.u = pkd->f[GG_U].as<double>();
    K.upred = pkd->f[GG_UPRED].as<double>();
    hipLaunchKernelGGL(k_kick, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, K);
    GG_HIP(hipGetLastError());
    GG_HIP(hipStreamSynchronize(pkd->stream));
    return GG_OK;
}

extern "C" int gg_drift(GGPKD pkd, double dDelta, const double *fCenter, int bPeriodic,
                        const double *fPeriod)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GG_CHECK(!bPeriodic || (fCenter && fPeriod), GG_EINVAL, "periodic drift needs fCenter and fPeriod");
    const int n = pkd->nLocal;
    if (n == 0) return GG_OK;
    GG_CHECK(pkd->f[GG_R].p, GG_ESTATE, "positions must be put before gg_drift");
    GG_TRY(ensure_field(pkd, GG_V));
    DevBuf bad;
    GG_TRY(bad.reserve(sizeof(unsigned)));
    const double c[3] = {bPeriodic ? fCenter[0] : 0, bPeriodic ? fCenter[1] : 0,
                         bPeriodic ? fCenter[2] : 0};
    const double L[3] = {bPeriodic ? fPeriod[0] : 0, bPeriodic ? fPeriod[1] : 0,
                         bPeriodic ? fPeriod[2] : 0};
    unsigned nb = 0;
    int rc = GG_OK;
    if (hipMemsetAsync(bad.p, 0, sizeof(unsigned), pkd->stream) != hipSuccess) rc = GG_EHIP;
    if (rc == GG_OK) {
        hipLaunchKernelGGL(k_drift, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n, dDelta,
                           bPeriodic, c[0], c[1], c[2], L[0], L[1], L[2],
                           pkd->f[GG_V].as<double>(), pkd->f[GG_R].as<double>(),
                           bad.as<unsigned>());
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&nb, bad.p, sizeof(unsigned), hipMemcpyDeviceToHost, pkd->stream) !=
                hipSuccess ||
            hipStreamSynchronize(pkd->stream) != hipSuccess)
            rc = GG_EHIP;
    }
    bad.release();
    GG_CHECK(rc == GG_OK, rc, "gg_drift: HIP launch or copy failed");
    pkd->sph.built = 0;    // the trees describe the old positions
    pkd->sph.listValid = 0;
    pkd->grav.built = 0;
    // pkdDrift asserts that every particle ends inside the box (pkd.c:3760)
    GG_CHECK(nb == 0, GG_EINVAL, "%u coordinates left the periodic box in one drift", nb);
    return GG_OK;
}
