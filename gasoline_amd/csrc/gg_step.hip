// The updates on either side of the force path, in FP64 on the device:
// pkdInitAccel (pkd.c:5443-5456), pkdKick (pkd.c:3780-3967, the NOCOOLING
// gas branch), pkdKickVpred (pkd.c:6489-6552), pkdDrift (pkd.c:3686-3778,
// without the sliding-patch and inflow/outflow options) and the timestep
// criteria and rung assignment (pkd.c:4627-4826, 5893-5979).  Elementwise
// over the store, caller order.  No contraction into FMA: every product and
// sum rounds as in the reference's x86-64 build, so the results are the
// reference's bit for bit.
#include "gg_internal.h"

#include <cfloat>
#include <cstring>
#include <initializer_list>
#include <vector>

#pragma clang fp contract(off)

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

// Active particles: v = v dvFacOne + a dvFacTwo; gas also
// vPred = v dvPredFacOne + a dvPredFacTwo, uPred = u + uDot duPredDelta and
// u = u + uDot duDelta with uDot = uDotPdV + uDotAV + uDotDiff (pkd.h:65).
// A GASOLINE build writes vPred of gas only (pkd.c:3805-3812, 3953-3957).
__global__ void k_kick(KickArgs K)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K.n) return;
    const uint32_t b = K.bits[i];
    if (!(b & GG_TYPE_ACTIVE)) return;
    const bool gas = (b & GG_TYPE_GAS) != 0;
    for (int j = 0; j < 3; ++j) {
        const size_t e = 3 * (size_t)i + j;
        const double v = K.v[e], a = K.a[e];
        if (gas) K.vpred[e] = v * K.vpf1 + a * K.vpf2;
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

// pkdKickVpred, NOCOOLING: gas particles only; bad counts uPred < 0
// (the reference asserts uPred >= 0).
__global__ void k_kick_vpred(int n, double vf1, double vf2, double du, int bGasU,
                             const uint32_t *__restrict__ bits, const double *__restrict__ a,
                             const double *__restrict__ udpdv, const double *__restrict__ udav,
                             const double *__restrict__ uddif, double *__restrict__ vpred,
                             double *__restrict__ upred, unsigned *__restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !(bits[i] & GG_TYPE_GAS)) return;
    for (int j = 0; j < 3; ++j) {
        const size_t e = 3 * (size_t)i + j;
        vpred[e] = vpred[e] * vf1 + a[e] * vf2;
    }
    if (bGasU) {
        const double u = upred[i] + (udpdv[i] + udav[i] + uddif[i]) * du;
        upred[i] = u;
        if (!(u >= 0.0)) atomicAdd(bad, 1u);
    }
}

__global__ void k_init_dt(int n, double dDelta, const uint32_t *__restrict__ bits,
                          double *__restrict__ dt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (bits[i] & GG_TYPE_ACTIVE)) dt[i] = dDelta;
}

// pkdGravStep: dt = min(dt, dEta/sqrt(dtGrav)); bad counts dtGrav <= 0.
__global__ void k_grav_step(int n, double dEta, const uint32_t *__restrict__ bits,
                            const double *__restrict__ dtg, double *__restrict__ dt,
                            unsigned *__restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !(bits[i] & GG_TYPE_ACTIVE)) return;
    if (!(dtg[i] > 0.0)) {
        atomicAdd(bad, 1u);
        return;
    }
    const double d = dEta / sqrt(dtg[i]);
    if (d < dt[i]) dt[i] = d;
}

// pkdAccelStep: the acceleration criterion (softening, or h for gas) and
// the sqrt(phi) criterion.
__global__ void k_accel_step(int n, double dEta, double dAccFac, int bEpsAcc, int bSqrtPhi,
                             const uint32_t *__restrict__ bits, const double *__restrict__ a,
                             const double *__restrict__ soft, const double *__restrict__ fb2,
                             const double *__restrict__ pot, double *__restrict__ dt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = bits[i];
    if (!(b & GG_TYPE_ACTIVE)) return;
    double acc = 0;
    for (int j = 0; j < 3; ++j) acc += a[3 * (size_t)i + j] * a[3 * (size_t)i + j];
    acc = sqrt(acc) * dAccFac;
    double dT = DBL_MAX;
    if (bEpsAcc && acc > 0) {
        if (b & GG_TYPE_GAS) dT = dEta * sqrt(sqrt(0.25 * fb2[i]) / acc);
        else dT = dEta * sqrt(soft[i] / acc);
    }
    if (bSqrtPhi && acc > 0) {
        const double dtemp = dEta * 3.5 * sqrt(dAccFac * fabs(pot[i])) / acc;
        if (dtemp < dT) dT = dtemp;
    }
    if (dT < dt[i]) dt[i] = dT;
}

__global__ void k_density_step(int n, double dEta, double dRhoFac,
                               const uint32_t *__restrict__ bits, const double *__restrict__ rho,
                               double *__restrict__ dt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !(bits[i] & GG_TYPE_ACTIVE)) return;
    const double dT = dEta / sqrt(rho[i] * dRhoFac);
    if (dT < dt[i]) dt[i] = dT;
}

// Order-preserving map of a double onto uint64 (for atomic min/max).
__device__ __forceinline__ unsigned long long dkey(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

struct SphStepArgs {
    int n;
    double dtFacCourant, dEtauDot, duMinDt, dDiffCoeff, dEtaDiffusion, dResolveJeans;
    const uint32_t *bits;
    const double *fb2, *c, *u, *udpdv, *udav, *uddif, *rho, *diff;
    double *dtNew, *dt;
    unsigned long long *minKey;      // smallest dt of the gas, as dkey
};

// pkdSphStep, NOCOOLING without THERMALCOND/TWOPHASE/UNONCOOL.
__global__ void k_sph_step(SphStepArgs A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long key = ~0ull;
    if (i < A.n && (A.bits[i] & GG_TYPE_GAS)) {
        if (A.bits[i] & GG_TYPE_ACTIVE) {
            const double ph = sqrt(0.25 * A.fb2[i]);
            double dT = A.dtNew[i];
            A.dtNew[i] = (double)FLT_MAX;
            const double uTotDot = A.udpdv[i] + A.udav[i] + A.uddif[i];
            if (uTotDot > 0) {
                // GAMMA_NONCOOL = 5/3 (pkd.h:228)
                const double dtExtrap =
                    A.dtFacCourant *
                    sqrt(A.fb2[i] * 0.25 / (4 * (A.c[i] * A.c[i] + (5. / 3.) * uTotDot * A.dt[i])));
                if (dtExtrap < dT) dT = dtExtrap;
            }
            if (A.dEtauDot > 0.0 && A.udpdv[i] < 0.0) {
                // Jeans floor (pkd.c:5668-5688) over GAMMA_JEANS - 1 = 1, PONRHOFLOOR 0
                const double floorJ = 0.25 * A.fb2[i] * A.dResolveJeans * A.rho[i];
                double uEff = 0 + floorJ / (2.0 - 1) + A.u[i];
                if (uEff < A.duMinDt) uEff = A.duMinDt;
                const double dTu = A.dEtauDot * uEff / fabs(A.udpdv[i]);
                if (dTu < dT) dT = dTu;
            }
            if (A.diff[i] > 0 && A.dDiffCoeff > 0) {
                const double dTD = A.dEtaDiffusion * ph * ph / (A.dDiffCoeff * A.diff[i]);
                if (dTD < dT) dT = dTD;
            }
            if (dT < A.dt[i]) A.dt[i] = dT;
        }
        key = dkey(A.dt[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long k2 = __shfl_xor(key, o);
        key = k2 < key ? k2 : key;
    }
    if ((threadIdx.x & 63) == 0 && key != ~0ull) atomicMin(A.minKey, key);
}

// pkdOneParticleDtToRung (pkd.c:4704-4726)
__device__ __forceinline__ int one_dt_to_rung(int iRung, double dDelta, double dt)
{
    const double dSteps = dDelta / dt;
    int iSteps = (int)floor(dSteps);
    if (fmod(dDelta, dt) == 0.0) iSteps--;   // an integer boundary goes to the lower rung
    if (iSteps < 0) iSteps = 0;
    int r = iRung;
    while (iSteps) {
        ++r;
        iSteps >>= 1;
    }
    return r;
}

// pkdEmergencyAdjust (pkd.c:3971-4030, GASOLINE, NOCOOLING): a gas particle on a rung
// below iRung whose Courant step dtNew fell under dDeltaThresh moves to the
// rung of dt = dtNew (capped at iMaxRung - 1) and is un-kicked to its
// predicted velocity.  red[0] = particles moved, red[1] = largest rung
// they took, red[2] = ideal largest rung, red[3 + r] = moved particles on
// rung r.
__global__ void k_emergency(int n, int iRung, int iMaxRung, double dDelta, double dDeltaThresh,
                            const uint32_t *__restrict__ bits, const double *__restrict__ dtNew,
                            double *__restrict__ dt, int *__restrict__ rung,
                            const double *__restrict__ vpred, double *__restrict__ v,
                            int *__restrict__ red)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !(bits[i] & GG_TYPE_GAS)) return;
    if (!(rung[i] < iRung && dtNew[i] < dDeltaThresh)) return;
    const double d = dtNew[i];
    dt[i] = d;
    int t = one_dt_to_rung(iRung, dDelta, d);
    atomicMax(&red[2], t + 1);
    if (t >= iMaxRung) t = iMaxRung - 1;
    rung[i] = t;
    atomicAdd(&red[0], 1);
    atomicMax(&red[1], t);
    atomicAdd(&red[3 + t], 1);
    for (int j = 0; j < 3; ++j) v[3 * (size_t)i + j] = vpred[3 * (size_t)i + j];
}

// pkdDtToRung (pkd.c:4729-4809) without the diagnostic print.  red[0]
// counts broken assertions, red[1] = largest rung, red[2] = ideal largest
// rung (k_rung_count adds red[3] = particles on red[1]).
__global__ void k_dt_to_rung(int n, int iRung, double dDelta, int iMaxRung, int bAll,
                             const uint32_t *__restrict__ bits, const double *__restrict__ soft,
                             const double *__restrict__ dt, int *__restrict__ rung,
                             double *__restrict__ dtOld, int *__restrict__ red)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int top = -1, ideal = 0;
    if (i < n) {
        int r = rung[i];
        if (r >= iRung) {
            if (!(bits[i] & GG_TYPE_ACTIVE)) {
                atomicAdd(&red[0], 1);
            } else if (bAll) {
                const double d = dt[i];
                if (!(soft[i] > 0.0) || !(d > 0.0) || !(dDelta / d < 2.1e9)) {
                    atomicAdd(&red[0], 1);
                } else {
                    int t = one_dt_to_rung(iRung, dDelta, d);
                    ideal = t + 1;
                    if (t >= iMaxRung) t = iMaxRung - 1;
                    r = t;
                    dtOld[i] = d;
                }
            } else {
                r = dDelta <= dt[i] ? iRung : iRung + 1;
            }
            rung[i] = r;
        }
        top = r;
    }
    for (int o = 32; o > 0; o >>= 1) {
        top = max(top, __shfl_xor(top, o));
        ideal = max(ideal, __shfl_xor(ideal, o));
    }
    if ((threadIdx.x & 63) == 0) {
        if (top >= 0) atomicMax(&red[1], top);
        if (ideal > 0) atomicMax(&red[2], ideal);
    }
}

// The type-bit updates of pkd.c:5114-5281 in one kernel.  A particle
// matches when its bits pass the test of the operation (and, with a rung
// test, it sits on rung iRung or above it with bGreater); then
//   TB_ACTIVE  (pkdActiveType, ActiveTypeRung, ActiveMaskRung, ActiveExactType):
//              iSetMask set on the matching particles, cleared on the others;
//   TB_SET     (pkdSetType):   iSetMask set on the matching ones;
//   TB_RESET   (pkdResetType): iSetMask cleared on the matching ones.
// Tests: TT_ANY bits & iTestMask nonzero (TYPETest), TT_EXACT (bits &
// iFilterMask) == iTestMask (TYPEFilter), TT_ALL every particle.
// count[0] = matching particles, count[1] = particles whose TYPE_ACTIVE bit
// changed, count[2] = particles whose TYPE_GAS bit changed.
enum { TB_ACTIVE = 0, TB_SET = 1, TB_RESET = 2 };
enum { TT_ANY = 0, TT_EXACT = 1, TT_ALL = 2 };
__global__ void k_type_bits(int n, int op, int test, uint32_t iFilterMask, uint32_t iTestMask,
                            uint32_t iSetMask, int bRung, int iRung, int bGreater,
                            const int *__restrict__ rung, uint32_t *__restrict__ bits,
                            int *__restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool on = false, dAct = false, dGas = false;
    if (i < n) {
        const uint32_t b = bits[i];
        on = test == TT_ALL ? true
             : test == TT_EXACT ? (b & iFilterMask) == iTestMask : (b & iTestMask) != 0;
        if (bRung) {
            const int r = rung[i];
            on = on && (r == iRung || (bGreater && r > iRung));
        }
        uint32_t nb = b;
        if (op == TB_ACTIVE) nb = on ? (b | iSetMask) : (b & ~iSetMask);
        else if (on) nb = op == TB_SET ? (b | iSetMask) : (b & ~iSetMask);
        dAct = ((nb ^ b) & GG_TYPE_ACTIVE) != 0;
        dGas = ((nb ^ b) & GG_TYPE_GAS) != 0;
        if (nb != b) bits[i] = nb;
    }
    const int c0 = __popcll(__ballot(on)), c1 = __popcll(__ballot(dAct)),
              c2 = __popcll(__ballot(dGas));
    if ((threadIdx.x & 63) == 0) {
        if (c0) atomicAdd(count, c0);
        if (c1) atomicAdd(count + 1, c1);
        if (c2) atomicAdd(count + 2, c2);
    }
}

__global__ void k_rung_count(int n, const int *__restrict__ rung, int *__restrict__ red)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = i < n && rung[i] == red[1];
    const int c = __popcll(__ballot(on));
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&red[3], c);
}

}  // namespace
}  // namespace gg

using namespace gg;

namespace {
int need_fields(GGPKD pkd, std::initializer_list<int> f)
{
    for (int k : f) GG_TRY(ensure_field(pkd, k));
    return GG_OK;
}
}  // namespace

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

extern "C" int gg_kick_vpred(GGPKD pkd, double dvFacOne, double dvFacTwo, double duDelta,
                             int bIsothermal)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    const int n = pkd->nLocal;
    if (n == 0) return GG_OK;
    GG_TRY(need_fields(pkd, {GG_VPRED, GG_A, GG_UPRED, GG_UDOTPDV, GG_UDOTAV, GG_UDOTDIFF,
                             GG_IACTIVE}));
    GG_TRY(pkd->red.reserve(sizeof(unsigned)));
    GG_HIP(hipMemsetAsync(pkd->red.p, 0, sizeof(unsigned), pkd->stream));
    hipLaunchKernelGGL(k_kick_vpred, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n, dvFacOne,
                       dvFacTwo, duDelta, !bIsothermal, pkd->f[GG_IACTIVE].as<uint32_t>(),
                       pkd->f[GG_A].as<double>(), pkd->f[GG_UDOTPDV].as<double>(),
                       pkd->f[GG_UDOTAV].as<double>(), pkd->f[GG_UDOTDIFF].as<double>(),
                       pkd->f[GG_VPRED].as<double>(), pkd->f[GG_UPRED].as<double>(),
                       pkd->red.as<unsigned>());
    GG_HIP(hipGetLastError());
    unsigned bad = 0;
    GG_HIP(hipMemcpyAsync(&bad, pkd->red.p, sizeof(unsigned), hipMemcpyDeviceToHost, pkd->stream));
    GG_HIP(hipStreamSynchronize(pkd->stream));
    GG_CHECK(bad == 0, GG_EINVAL, "gg_kick_vpred: %u gas particles with uPred < 0", bad);
    return GG_OK;
}

extern "C" int gg_init_dt(GGPKD pkd, double dDelta)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    const int n = pkd->nLocal;
    if (n == 0) return GG_OK;
    GG_TRY(need_fields(pkd, {GG_DT, GG_IACTIVE}));
    hipLaunchKernelGGL(k_init_dt, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n, dDelta,
                       pkd->f[GG_IACTIVE].as<uint32_t>(), pkd->f[GG_DT].as<double>());
    GG_HIP(hipGetLastError());
    GG_HIP(hipStreamSynchronize(pkd->stream));
    return GG_OK;
}

extern "C" int gg_grav_step(GGPKD pkd, double dEta)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    const int n = pkd->nLocal;
    if (n == 0) return GG_OK;
    GG_TRY(need_fields(pkd, {GG_DT, GG_DTGRAV, GG_IACTIVE}));
    GG_TRY(pkd->red.reserve(sizeof(unsigned)));
    GG_HIP(hipMemsetAsync(pkd->red.p, 0, sizeof(unsigned), pkd->stream));
    hipLaunchKernelGGL(k_grav_step, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n, dEta,
                       pkd->f[GG_IACTIVE].as<uint32_t>(), pkd->f[GG_DTGRAV].as<double>(),
                       pkd->f[GG_DT].as<double>(), pkd->red.as<unsigned>());
    GG_HIP(hipGetLastError());
    unsigned bad = 0;
    GG_HIP(hipMemcpyAsync(&bad, pkd->red.p, sizeof(unsigned), hipMemcpyDeviceToHost, pkd->stream));
    GG_HIP(hipStreamSynchronize(pkd->stream));
    GG_CHECK(bad == 0, GG_EINVAL, "gg_grav_step: %u active particles with dtGrav <= 0", bad);
    return GG_OK;
}

extern "C" int gg_accel_step(GGPKD pkd, double dEta, double dVelFac, double dAccFac,
                             int bDoGravity, int bEpsAcc, int bSqrtPhi, double dhMinOverSoft)
{
    (void)dVelFac;        // the velocity enters no criterion (pkd.c:4657-4664)
    (void)bDoGravity;
    (void)dhMinOverSoft;
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    const int n = pkd->nLocal;
    if (n == 0) return GG_OK;
    GG_TRY(need_fields(pkd, {GG_DT, GG_A, GG_SOFT, GG_FBALL2, GG_POT, GG_IACTIVE}));
    hipLaunchKernelGGL(k_accel_step, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n, dEta,
                       dAccFac, bEpsAcc, bSqrtPhi, pkd->f[GG_IACTIVE].as<uint32_t>(),
                       pkd->f[GG_A].as<double>(), pkd->f[GG_SOFT].as<double>(),
                       pkd->f[GG_FBALL2].as<double>(), pkd->f[GG_POT].as<double>(),
                       pkd->f[GG_DT].as<double>());
    GG_HIP(hipGetLastError());
    GG_HIP(hipStreamSynchronize(pkd->stream));
    return GG_OK;
}

extern "C" int gg_density_step(GGPKD pkd, double dEta, double dRhoFac)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    const int n = pkd->nLocal;
    if (n == 0) return GG_OK;
    GG_TRY(need_fields(pkd, {GG_DT, GG_DENSITY, GG_IACTIVE}));
    hipLaunchKernelGGL(k_density_step, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n, dEta,
                       dRhoFac, pkd->f[GG_IACTIVE].as<uint32_t>(),
                       pkd->f[GG_DENSITY].as<double>(), pkd->f[GG_DT].as<double>());
    GG_HIP(hipGetLastError());
    GG_HIP(hipStreamSynchronize(pkd->stream));
    return GG_OK;
}

extern "C" int gg_sph_step(GGPKD pkd, double dCosmoFac, double dEtaCourant, double dEtauDot,
                           double duMinDt, double dDiffCoeff, double dEtaDiffusion,
                           double dResolveJeans, int bViscosityLimitdt, double *pdtMinGas)
{
    (void)bViscosityLimitdt;   // read by no branch of the NOCOOLING build (pkd.c:5901-5979)
    GG_CHECK(pkd && pdtMinGas, GG_EINVAL, "null argument");
    const int n = pkd->nLocal;
    *pdtMinGas = DBL_MAX;
    if (n == 0) return GG_OK;
    GG_TRY(need_fields(pkd, {GG_DT, GG_DTNEW, GG_FBALL2, GG_C, GG_U, GG_UDOTPDV, GG_UDOTAV,
                             GG_UDOTDIFF, GG_DENSITY, GG_DIFF, GG_IACTIVE}));
    GG_TRY(pkd->red.reserve(sizeof(unsigned long long)));
    GG_HIP(hipMemsetAsync(pkd->red.p, 0xff, sizeof(unsigned long long), pkd->stream));
    SphStepArgs A;
    A.n = n;
    A.dtFacCourant = dEtaCourant * dCosmoFac * 2 / 1.6;   // pkdDtFacCourant, pkd.c:5893-5895
    A.dEtauDot = dEtauDot;
    A.duMinDt = duMinDt;
    A.dDiffCoeff = dDiffCoeff;
    A.dEtaDiffusion = dEtaDiffusion;
    A.dResolveJeans = dResolveJeans;
    A.bits = pkd->f[GG_IACTIVE].as<uint32_t>();
    A.fb2 = pkd->f[GG_FBALL2].as<double>();
    A.c = pkd->f[GG_C].as<double>();
    A.u = pkd->f[GG_U].as<double>();
    A.udpdv = pkd->f[GG_UDOTPDV].as<double>();
    A.udav = pkd->f[GG_UDOTAV].as<double>();
    A.uddif = pkd->f[GG_UDOTDIFF].as<double>();
    A.rho = pkd->f[GG_DENSITY].as<double>();
    A.diff = pkd->f[GG_DIFF].as<double>();
    A.dtNew = pkd->f[GG_DTNEW].as<double>();
    A.dt = pkd->f[GG_DT].as<double>();
    A.minKey = pkd->red.as<unsigned long long>();
    hipLaunchKernelGGL(k_sph_step, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, A);
    GG_HIP(hipGetLastError());
    unsigned long long key = ~0ull;
    GG_HIP(hipMemcpyAsync(&key, pkd->red.p, sizeof(key), hipMemcpyDeviceToHost, pkd->stream));
    GG_HIP(hipStreamSynchronize(pkd->stream));
    if (key != ~0ull) {   // undo the order-preserving map
        const unsigned long long b = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
        std::memcpy(pdtMinGas, &b, sizeof(double));
    }
    return GG_OK;
}

extern "C" int gg_dt_to_rung(GGPKD pkd, int iRung, double dDelta, int iMaxRung, int bAll,
                             int bDiagExceed, int *pnMaxRung, int *piMaxRungIdeal,
                             int *piMaxRungOut)
{
    (void)bDiagExceed;   // selects diagnostic prints only (pkd.c:4748-4789)
    GG_CHECK(pkd && pnMaxRung && piMaxRungIdeal && piMaxRungOut, GG_EINVAL, "null argument");
    const int n = pkd->nLocal;
    *pnMaxRung = *piMaxRungIdeal = *piMaxRungOut = 0;
    if (n == 0) return GG_OK;
    GG_TRY(need_fields(pkd, {GG_IRUNG, GG_DT, GG_DTOLD, GG_SOFT, GG_IACTIVE}));
    GG_TRY(pkd->red.reserve(4 * sizeof(int)));
    GG_HIP(hipMemsetAsync(pkd->red.p, 0, 4 * sizeof(int), pkd->stream));
    hipLaunchKernelGGL(k_dt_to_rung, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n, iRung,
                       dDelta, iMaxRung, bAll, pkd->f[GG_IACTIVE].as<uint32_t>(),
                       pkd->f[GG_SOFT].as<double>(), pkd->f[GG_DT].as<double>(),
                       pkd->f[GG_IRUNG].as<int>(), pkd->f[GG_DTOLD].as<double>(),
                       pkd->red.as<int>());
    hipLaunchKernelGGL(k_rung_count, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n,
                       pkd->f[GG_IRUNG].as<int>(), pkd->red.as<int>());
    GG_HIP(hipGetLastError());
    int red[4];
    GG_HIP(hipMemcpyAsync(red, pkd->red.p, sizeof(red), hipMemcpyDeviceToHost, pkd->stream));
    GG_HIP(hipStreamSynchronize(pkd->stream));
    GG_CHECK(red[0] == 0, GG_EINVAL,
             "gg_dt_to_rung: %d particles on rung >= %d are inactive or have fSoft <= 0, dt <= 0 "
             "or dDelta/dt >= 2.1e9 (rungs are undefined)", red[0], iRung);
    *piMaxRungOut = red[1];
    *pnMaxRung = red[3];
    *piMaxRungIdeal = red[2];
    return GG_OK;
}

// Run k_type_bits and keep the trees that do not depend on what changed:
// the gravity tree holds the active buckets and their active boxes (TYPE_ACTIVE),
// the SPH structure its gas particles (TYPE_GAS); the other bits are read by
// the calls that use them.
static int type_bits(GGPKD pkd, int op, int test, unsigned iFilterMask, unsigned iTestMask,
                     unsigned iSetMask, int bRung, int iRung, int bGreater, int *pn)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    const int n = pkd->nLocal;
    if (pn) *pn = 0;
    if (n == 0) return GG_OK;
    if (bRung) GG_TRY(need_fields(pkd, {GG_IRUNG, GG_IACTIVE}));
    else GG_TRY(need_fields(pkd, {GG_IACTIVE}));
    GG_TRY(pkd->red.reserve(3 * sizeof(int)));
    GG_HIP(hipMemsetAsync(pkd->red.p, 0, 3 * sizeof(int), pkd->stream));
    hipLaunchKernelGGL(k_type_bits, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n, op, test,
                       (uint32_t)iFilterMask, (uint32_t)iTestMask, (uint32_t)iSetMask, bRung,
                       iRung, bGreater, bRung ? pkd->f[GG_IRUNG].as<int>() : (const int *)nullptr,
                       pkd->f[GG_IACTIVE].as<uint32_t>(), pkd->red.as<int>());
    GG_HIP(hipGetLastError());
    int c[3] = {0, 0, 0};
    GG_HIP(hipMemcpyAsync(c, pkd->red.p, sizeof(c), hipMemcpyDeviceToHost, pkd->stream));
    GG_HIP(hipStreamSynchronize(pkd->stream));
    if (c[1]) pkd->grav.built = 0;
    if (c[2]) {
        pkd->sph.gasMaskValid = 0;
        pkd->sph.built = 0;
    }
    if (pn) *pn = c[0];
    return GG_OK;
}

extern "C" int gg_emergency_adjust(GGPKD pkd, int iRung, int iMaxRung, double dDelta,
                                   double dDeltaThresh, int *pnUn, int *piMaxRungIdeal,
                                   int *pnMaxRung, int *piMaxRungOut)
{
    GG_CHECK(pkd && pnUn && piMaxRungIdeal && pnMaxRung && piMaxRungOut, GG_EINVAL,
             "null argument");
    GG_CHECK(dDeltaThresh < dDelta, GG_EINVAL, "dDeltaThresh must be below dDelta");
    GG_CHECK(iMaxRung >= 1 && iMaxRung <= 60, GG_EINVAL, "iMaxRung %d outside 1..60", iMaxRung);
    const int n = pkd->nLocal;
    *pnUn = *piMaxRungIdeal = *pnMaxRung = *piMaxRungOut = 0;
    if (n == 0) return GG_OK;
    GG_TRY(need_fields(pkd, {GG_IRUNG, GG_DT, GG_DTNEW, GG_V, GG_VPRED, GG_IACTIVE}));
    const int nred = 3 + iMaxRung;
    GG_TRY(pkd->red.reserve(sizeof(int) * nred));
    GG_HIP(hipMemsetAsync(pkd->red.p, 0, sizeof(int) * nred, pkd->stream));
    hipLaunchKernelGGL(k_emergency, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n, iRung,
                       iMaxRung, dDelta, dDeltaThresh, pkd->f[GG_IACTIVE].as<uint32_t>(),
                       pkd->f[GG_DTNEW].as<double>(), pkd->f[GG_DT].as<double>(),
                       pkd->f[GG_IRUNG].as<int>(), pkd->f[GG_VPRED].as<double>(),
                       pkd->f[GG_V].as<double>(), pkd->red.as<int>());
    GG_HIP(hipGetLastError());
    std::vector<int> red(nred);
    GG_HIP(hipMemcpyAsync(red.data(), pkd->red.p, sizeof(int) * nred, hipMemcpyDeviceToHost,
                          pkd->stream));
    GG_HIP(hipStreamSynchronize(pkd->stream));
    *pnUn = red[0];
    *piMaxRungOut = red[1];
    *piMaxRungIdeal = red[2];
    *pnMaxRung = red[0] ? red[3 + red[1]] : 0;
    return GG_OK;
}

extern "C" int gg_active_mask_rung(GGPKD pkd, unsigned iSetMask, int iRung, int bGreater,
                                   int *pnActive)
{
    return type_bits(pkd, TB_ACTIVE, TT_ALL, 0, 0, iSetMask, 1, iRung, bGreater, pnActive);
}

extern "C" int gg_active_type(GGPKD pkd, unsigned iTestMask, unsigned iSetMask, int *pnActive)
{
    return type_bits(pkd, TB_ACTIVE, TT_ANY, 0, iTestMask, iSetMask, 0, 0, 0, pnActive);
}

extern "C" int gg_active_type_rung(GGPKD pkd, unsigned iTestMask, unsigned iSetMask, int iRung,
                                   int bGreater, int *pnActive)
{
    return type_bits(pkd, TB_ACTIVE, TT_ANY, 0, iTestMask, iSetMask, 1, iRung, bGreater,
                     pnActive);
}

extern "C" int gg_active_exact_type(GGPKD pkd, unsigned iFilterMask, unsigned iTestMask,
                                    unsigned iSetMask, int *pnActive)
{
    return type_bits(pkd, TB_ACTIVE, TT_EXACT, iFilterMask, iTestMask, iSetMask, 0, 0, 0,
                     pnActive);
}

extern "C" int gg_set_type(GGPKD pkd, unsigned iTestMask, unsigned iSetMask, int *pn)
{
    return type_bits(pkd, TB_SET, TT_ANY, 0, iTestMask, iSetMask, 0, 0, 0, pn);
}

extern "C" int gg_reset_type(GGPKD pkd, unsigned iTestMask, unsigned iSetMask, int *pn)
{
    return type_bits(pkd, TB_RESET, TT_ANY, 0, iTestMask, iSetMask, 0, 0, 0, pn);
}

extern "C" int gg_sph_curr_rung(GGPKD pkd, int iRung, int bGreater, int *piCurrent)
{
    // pkdSphCurrRung (pkd.c:5876-5890): is a gas particle on the rung(s)?
    // (set bits of nothing: only the count of TYPE_GAS particles on them)
    GG_CHECK(piCurrent, GG_EINVAL, "null argument");
    int c = 0;
    GG_TRY(type_bits(pkd, TB_SET, TT_ANY, 0, GG_TYPE_GAS, 0, 1, iRung, bGreater, &c));
    *piCurrent = c > 0;
    return GG_OK;
}
