// Handle management, particle-store I/O, device scan and timing for the
// C ABI of libgasoline_amd (include/gasoline_amd.h).
#include "gg_internal.h"
#include "gg_scan.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>

namespace gg {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int DevBuf::reserve(size_t nbytes)
{
    if (nbytes <= bytes && p) return GG_OK;
    release();
    size_t want = nbytes < 256 ? 256 : nbytes;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        bytes = 0;
        set_error("hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
        return GG_ENOMEM;
    }
    bytes = want;
    return GG_OK;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int field_width(int field)
{
    switch (field) {
    case GG_R:
    case GG_V:
    case GG_VPRED:
    case GG_A:
    case GG_CURLV:
        return 3;
    default:
        return 1;
    }
}

static size_t field_bytes(int field, int n)
{
    if (field == GG_IACTIVE) return sizeof(uint32_t) * (size_t)n;
    return sizeof(double) * (size_t)field_width(field) * (size_t)n;
}

__global__ void k_fill_u32(size_t n, uint32_t v, uint32_t *__restrict__ p)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f64(size_t n, double v, double *__restrict__ p)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Allocate a field on first use.  Defaults: iActive = ACTIVE | TREEACTIVE |
// SMOOTHACTIVE, dt and dtNew = FLT_MAX (floattype.h:9), everything else 0.
int ensure_field(GGPKD pkd, int field)
{
    DevBuf &b = pkd->f[field];
    if (b.p) return GG_OK;
    const int n = pkd->nStore > 0 ? pkd->nStore : 1;
    const size_t nb = field_bytes(field, n);
    GG_TRY(b.reserve(nb));
    const size_t cnt = field == GG_IACTIVE ? (size_t)n : (size_t)field_width(field) * n;
    const unsigned grid = (unsigned)((cnt + 255) / 256);
    if (field == GG_IACTIVE) {
        pkd->sph.gasMaskValid = 0;
        hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, pkd->stream, cnt,
                           GG_TYPE_ACTIVE | GG_TYPE_TREEACTIVE | GG_TYPE_SMOOTHACTIVE,
                           b.as<uint32_t>());
    } else {
        const double v = (field == GG_DT || field == GG_DTNEW) ? (double)3.402823466E+38F : 0.0;
        hipLaunchKernelGGL(k_fill_f64, dim3(grid), dim3(256), 0, pkd->stream, cnt, v,
                           b.as<double>());
    }
    GG_HIP(hipGetLastError());
    return GG_OK;
}

// ------------------------------------------------------------ scan
int exclusive_scan_i64(hipStream_t st, const int *in, int64_t *out, int n, DevBuf &scratch,
                       int64_t *total_host)
{
    return exclusive_scan(st, ScanArray{in}, out, n, scratch, total_host);
}

// ----------------------------------------------------------- timing
void timer_mark(GGPKD pkd, int slot)
{
    if (!pkd->tm.created) {
        for (int k = 0; k < 16; ++k) (void)hipEventCreate(&pkd->tm.ev[k]);
        pkd->tm.created = 1;
    }
    (void)hipEventRecord(pkd->tm.ev[slot], pkd->stream);
}

// Elapsed time between consecutive marks first..first+nphases, stored into
// the phase slots (the stream must have been synchronised).
void timer_collect(GGPKD pkd, int first, int nphases, const int *phase)
{
    for (int k = 0; k < nphases; ++k) pkd->tm.ms[phase[k]] = 0;
    for (int k = 0; k < nphases; ++k) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, pkd->tm.ev[first + k], pkd->tm.ev[first + k + 1]) == hipSuccess)
            pkd->tm.ms[phase[k]] += ms;
    }
}

void sph_release(GGPKD pkd) { pkd->sph.release_all(); }

}  // namespace gg

using namespace gg;

extern "C" const char *gg_last_error(void) { return g_err; }

extern "C" int gg_initialize(GGPKD *ppkd, int device, int nStore)
{
    GG_CHECK(ppkd, GG_EINVAL, "null output pointer");
    GG_CHECK(nStore >= 0, GG_EINVAL, "nStore must be >= 0");
    *ppkd = nullptr;
    int ndev = 0;
    GG_HIP(hipGetDeviceCount(&ndev));
    GG_CHECK(device >= 0 && device < ndev, GG_EINVAL, "device %d not in 0..%d", device, ndev - 1);
    GG_HIP(hipSetDevice(device));
    gg_pkd_s *p = new gg_pkd_s();
    p->device = device;
    p->nStore = nStore;
    p->nLocal = nStore;
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        delete p;
        set_error("hipStreamCreate failed");
        return GG_EHIP;
    }
    *ppkd = p;
    return GG_OK;
}

extern "C" int gg_finish(GGPKD pkd)
{
    if (!pkd) return GG_OK;
    (void)hipSetDevice(pkd->device);
    (void)hipStreamSynchronize(pkd->stream);
    for (auto &b : pkd->f) b.release();
    GravDev &G = pkd->grav;
    G.release_all();
    if (G.gexec) (void)hipGraphExecDestroy(G.gexec);
    if (G.graph) (void)hipGraphDestroy(G.graph);
    sph_release(pkd);
    pkd->mlp.release_all();
    if (pkd->tm.created)
        for (int k = 0; k < 16; ++k) (void)hipEventDestroy(pkd->tm.ev[k]);
    (void)hipStreamDestroy(pkd->stream);
    delete pkd;
    return GG_OK;
}

extern "C" int gg_set_local(GGPKD pkd, int nLocal)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GG_CHECK(nLocal >= 0 && nLocal <= pkd->nStore, GG_EINVAL, "nLocal %d outside 0..%d",
             nLocal, pkd->nStore);
    pkd->nLocal = nLocal;
    pkd->grav.built = 0;
    pkd->sph.built = 0;
    pkd->sph.gasMaskValid = 0;
    return GG_OK;
}

extern "C" int gg_local(GGPKD pkd) { return pkd ? pkd->nLocal : -1; }

extern "C" int gg_set_precision(GGPKD pkd, int prec)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GG_CHECK(prec == GG_PREC_FP32 || prec == GG_PREC_FP64, GG_EINVAL, "unknown precision %d", prec);
    pkd->prec = prec;
    return GG_OK;
}

extern "C" int gg_put(GGPKD pkd, int field, const void *src, int bDevice)
{
    GG_CHECK(pkd && src, GG_EINVAL, "null argument");
    GG_CHECK(field >= 0 && field < GG_NFIELDS, GG_EINVAL, "unknown field %d", field);
    GG_TRY(ensure_field(pkd, field));
    const size_t nb = field_bytes(field, pkd->nLocal);
    if (nb)
        GG_HIP(hipMemcpyAsync(pkd->f[field].p, src, nb,
                              bDevice ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                              pkd->stream));
    GG_HIP(hipStreamSynchronize(pkd->stream));
    if (field == GG_IACTIVE) pkd->sph.gasMaskValid = 0;
    if (field == GG_R || field == GG_MASS || field == GG_SOFT || field == GG_IACTIVE) {
        pkd->grav.built = 0;
        if (field != GG_MASS && field != GG_SOFT) pkd->sph.built = 0;
    }
    return GG_OK;
}

extern "C" int gg_get(GGPKD pkd, int field, void *dst, int bDevice)
{
    GG_CHECK(pkd && dst, GG_EINVAL, "null argument");
    GG_CHECK(field >= 0 && field < GG_NFIELDS, GG_EINVAL, "unknown field %d", field);
    GG_TRY(ensure_field(pkd, field));
    const size_t nb = field_bytes(field, pkd->nLocal);
    if (nb)
        GG_HIP(hipMemcpyAsync(dst, pkd->f[field].p, nb,
                              bDevice ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                              pkd->stream));
    GG_HIP(hipStreamSynchronize(pkd->stream));
    return GG_OK;
}

extern "C" int gg_timings(GGPKD pkd, double *t, int n)
{
    GG_CHECK(pkd && t, GG_EINVAL, "null argument");
    for (int k = 0; k < n && k < 12; ++k) t[k] = pkd->tm.ms[k];
    return GG_OK;
}
