// Handle management, particle-store I/O, device scan and timing for the
// C ABI of libgasoline_amd (include/gasoline_amd.h).
#include "gg_internal.h"
#include "gg_scan.h"

#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <thread>

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

int field_is32(int field) { return field == GG_IACTIVE || field == GG_IRUNG; }

static size_t field_bytes(int field, int n)
{
    if (field_is32(field)) return sizeof(uint32_t) * (size_t)n;
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
// SMOOTHACTIVE, dt and dtNew = FLT_MAX (floattype.h:9), iRung and everything
// else 0.
int ensure_field(GGPKD pkd, int field)
{
    DevBuf &b = pkd->f[field];
    if (b.p) return GG_OK;
    const int n = pkd->nStore > 0 ? pkd->nStore : 1;
    const size_t nb = field_bytes(field, n);
    GG_TRY(b.reserve(nb));
    const size_t cnt = field_is32(field) ? (size_t)n : (size_t)field_width(field) * n;
    const unsigned grid = (unsigned)((cnt + 255) / 256);
    if (field == GG_IACTIVE) {
        pkd->sph.gasMaskValid = 0;
        hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, pkd->stream, cnt,
                           GG_TYPE_ACTIVE | GG_TYPE_TREEACTIVE | GG_TYPE_SMOOTHACTIVE,
                           b.as<uint32_t>());
    } else if (field == GG_IRUNG) {
        hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, pkd->stream, cnt, 0u,
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
    Timers &T = pkd->tm;
    if (!T.created) {
        for (int f = 0; f < Timers::NFR; ++f)
            for (int k = 0; k < Timers::NEV; ++k) (void)hipEventCreate(&T.ev[f][k]);
        T.created = 1;
    }
    if (slot == 0) {   // a new call: the next frame, once its old marks are read
        const int next = (T.frame + 1) % Timers::NFR;
        for (int i = 0; i < T.npend; ++i)
            if (T.pend[i].frame == next) {
                timer_resolve(pkd);
                break;
            }
        T.frame = next;
    }
    (void)hipEventRecord(T.ev[T.frame][slot], pkd->stream);
}

// The phases between consecutive marks first..first+nphases of the current
// call, read later (timer_resolve): each phase slot then holds this call's
// time (phases listed twice add up).
void timer_collect(GGPKD pkd, int first, int nphases, const int *phase)
{
    Timers &T = pkd->tm;
    if (T.npend == Timers::NPEND) timer_resolve(pkd);
    Timers::Pend &p = T.pend[T.npend++];
    p.frame = T.frame;
    p.first = first;
    p.n = nphases < 8 ? nphases : 8;
    for (int k = 0; k < p.n; ++k) p.phase[k] = phase[k];
}

// The same, read at once (the stream has been synchronised), after every
// earlier call's times.
void timer_collect_now(GGPKD pkd, int first, int nphases, const int *phase)
{
    timer_collect(pkd, first, nphases, phase);
    timer_resolve(pkd);
}

// Read every pending call's times, oldest first.
void timer_resolve(GGPKD pkd)
{
    Timers &T = pkd->tm;
    for (int i = 0; i < T.npend; ++i) {
        const Timers::Pend &p = T.pend[i];
        hipEvent_t *ev = T.ev[p.frame];
        (void)hipEventSynchronize(ev[p.first + p.n]);
        for (int k = 0; k < p.n; ++k) T.ms[p.phase[k]] = 0;
        for (int k = 0; k < p.n; ++k) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, ev[p.first + k], ev[p.first + k + 1]) == hipSuccess)
                T.ms[p.phase[k]] += ms;
        }
    }
    T.npend = 0;
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
    GG_TRY(direct_scratch_setup(device));   // the no-alloc direct sum's array on this GPU
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

static void aos_release(GGPKD pkd);

extern "C" int gg_finish(GGPKD pkd)
{
    if (!pkd) return GG_OK;
    (void)hipSetDevice(pkd->device);
    (void)hipStreamSynchronize(pkd->stream);
    for (auto &b : pkd->f) b.release();
    pkd->red.release();
    GravDev &G = pkd->grav;
    G.release_all();
    if (G.gexec) (void)hipGraphExecDestroy(G.gexec);
    if (G.graph) (void)hipGraphDestroy(G.graph);
    sph_release(pkd);
    pkd->mlp.release_all();
    aos_release(pkd);
    if (pkd->tm.created)
        for (int f = 0; f < Timers::NFR; ++f)
            for (int k = 0; k < Timers::NEV; ++k) (void)hipEventDestroy(pkd->tm.ev[f][k]);
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

extern "C" int gg_set_period(GGPKD pkd, const double *fPeriod)
{
    GG_CHECK(pkd && fPeriod, GG_EINVAL, "null argument");
    for (int d = 0; d < 3; ++d)
        GG_CHECK(fPeriod[d] > 0, GG_EINVAL, "fPeriod[%d] = %g is not positive", d, fPeriod[d]);
    for (int d = 0; d < 3; ++d) pkd->fPeriod[d] = fPeriod[d];
    return GG_OK;
}

extern "C" int gg_set_precision(GGPKD pkd, int prec)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GG_CHECK(prec == GG_PREC_FP32 || prec == GG_PREC_FP64, GG_EINVAL, "unknown precision %d", prec);
    pkd->prec = prec;
    return GG_OK;
}

// What a new value of a field invalidates: the trees hold positions, masses,
// softenings and type bits; the SPH structure positions and type bits.
static void put_done(GGPKD pkd, int field)
{
    if (field == GG_IACTIVE) pkd->sph.gasMaskValid = 0;
    if (field == GG_R || field == GG_MASS || field == GG_SOFT || field == GG_IACTIVE) {
        pkd->grav.built = 0;
        if (field != GG_MASS && field != GG_SOFT) pkd->sph.built = 0;
    }
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
    // a host source is free again on return; a device copy stays queued on
    // the handle's stream, ahead of everything the handle does next
    if (!bDevice) GG_HIP(hipStreamSynchronize(pkd->stream));
    put_done(pkd, field);
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

// ------------------------------------------------ array-of-structs I/O
static const size_t AOS_CHUNK = 32u << 20;   // staging bytes per buffer

static int aos_stage(GGPKD pkd, size_t bytes)
{
    AosStage &A = pkd->aos;
    if (!A.created) {
        for (int k = 0; k < 2; ++k) GG_HIP(hipEventCreateWithFlags(&A.ev[k], hipEventDisableTiming));
        A.created = 1;
    }
    if (A.bytes >= bytes) return GG_OK;
    for (int k = 0; k < 2; ++k) {
        if (A.buf[k]) (void)hipHostFree(A.buf[k]);
        A.buf[k] = nullptr;
    }
    A.bytes = 0;
    for (int k = 0; k < 2; ++k) {
        hipError_t e = hipHostMalloc(&A.buf[k], bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            A.buf[k] = nullptr;
            set_error("hipHostMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
            return GG_ENOMEM;
        }
    }
    A.bytes = bytes;
    return GG_OK;
}

static void aos_release(GGPKD pkd)
{
    AosStage &A = pkd->aos;
    for (int k = 0; k < 2; ++k) {
        if (A.buf[k]) (void)hipHostFree(A.buf[k]);
        A.buf[k] = nullptr;
    }
    if (A.created)
        for (int k = 0; k < 2; ++k) (void)hipEventDestroy(A.ev[k]);
    A.created = 0;
    A.bytes = 0;
}

static size_t aos_width(int field)
{
    return field_is32(field) ? sizeof(uint32_t) : sizeof(double) * field_width(field);
}

// Validates the member list; w[k] = bytes of member k, *row = their sum.
static int aos_check(GGPKD pkd, const void *base, size_t stride, const gg_aos_field *f, int nf,
                     std::vector<size_t> &w, size_t *row)
{
    GG_CHECK(pkd && f, GG_EINVAL, "null argument");
    GG_CHECK(base || pkd->nLocal == 0, GG_EINVAL, "null particle array");
    GG_CHECK(nf >= 0 && nf <= GG_NFIELDS, GG_EINVAL, "%d members", nf);
    w.resize(nf);
    *row = 0;
    for (int k = 0; k < nf; ++k) {
        GG_CHECK(f[k].field >= 0 && f[k].field < GG_NFIELDS, GG_EINVAL, "member %d: unknown field %d",
                 k, f[k].field);
        for (int j = 0; j < k; ++j)
            GG_CHECK(f[j].field != f[k].field, GG_EINVAL, "field %d listed twice", f[k].field);
        w[k] = aos_width(f[k].field);
        GG_CHECK(f[k].offset + w[k] <= stride, GG_EINVAL,
                 "member %d (field %d, offset %zu, %zu bytes) outside the %zu-byte particle", k,
                 f[k].field, f[k].offset, w[k], stride);
        *row += w[k];
    }
    return GG_OK;
}

// Rows [r0, r1) between the particle array and the staging buffer, which
// holds one segment per member, seg[k] bytes in; split over host threads.
static void aos_move(char *aos, size_t stride, const gg_aos_field *f, int nf, const size_t *w,
                     const size_t *seg, int r0, int r1, char *stage, bool pack)
{
    const int rows = r1 - r0;
    int nt = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 8u);
    nt = std::max(1, std::min(nt, rows / 16384));
    auto work = [&](int a, int b) {
        for (int r = a; r < b; ++r) {
            char *p = aos + (size_t)r * stride;
            const size_t i = (size_t)(r - r0);
            for (int k = 0; k < nf; ++k) {
                char *s = stage + seg[k] + i * w[k];
                if (pack) std::memcpy(s, p + f[k].offset, w[k]);
                else std::memcpy(p + f[k].offset, s, w[k]);
            }
        }
    };
    if (nt == 1) {
        work(r0, r1);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back(work, r0 + (int)((int64_t)rows * t / nt), r0 + (int)((int64_t)rows * (t + 1) / nt));
    for (auto &x : th) x.join();
}

extern "C" int gg_put_aos(GGPKD pkd, const void *base, size_t stride, const gg_aos_field *f, int nf)
{
    std::vector<size_t> w;
    size_t row = 0;
    GG_TRY(aos_check(pkd, base, stride, f, nf, w, &row));
    const int n = pkd->nLocal;
    if (n == 0 || nf == 0) return GG_OK;
    for (int k = 0; k < nf; ++k) GG_TRY(ensure_field(pkd, f[k].field));
    const int C = (int)std::max<size_t>(1, std::min<size_t>(AOS_CHUNK / row, (size_t)n));
    GG_TRY(aos_stage(pkd, (size_t)C * row));
    AosStage &A = pkd->aos;
    std::vector<size_t> seg(nf);
    for (int k = 0, o = 0; k < nf; o += (int)(w[k] * C), ++k) seg[k] = (size_t)o;
    for (int c = 0, r0 = 0; r0 < n; ++c, r0 += C) {
        const int b = c & 1, r1 = std::min(n, r0 + C);
        if (A.pending[b]) GG_HIP(hipEventSynchronize(A.ev[b]));   // its copies have read it
        A.pending[b] = 0;
        char *stage = (char *)A.buf[b];
        aos_move((char *)base, stride, f, nf, w.data(), seg.data(), r0, r1, stage, true);
        for (int k = 0; k < nf; ++k)
            GG_HIP(hipMemcpyAsync((char *)pkd->f[f[k].field].p + (size_t)r0 * w[k], stage + seg[k],
                                  (size_t)(r1 - r0) * w[k], hipMemcpyHostToDevice, pkd->stream));
        GG_HIP(hipEventRecord(A.ev[b], pkd->stream));
        A.pending[b] = 1;
    }
    GG_HIP(hipStreamSynchronize(pkd->stream));
    A.pending[0] = A.pending[1] = 0;
    for (int k = 0; k < nf; ++k) put_done(pkd, f[k].field);
    return GG_OK;
}

extern "C" int gg_get_aos(GGPKD pkd, void *base, size_t stride, const gg_aos_field *f, int nf)
{
    std::vector<size_t> w;
    size_t row = 0;
    GG_TRY(aos_check(pkd, base, stride, f, nf, w, &row));
    const int n = pkd->nLocal;
    if (n == 0 || nf == 0) return GG_OK;
    for (int k = 0; k < nf; ++k) GG_TRY(ensure_field(pkd, f[k].field));
    const int C = (int)std::max<size_t>(1, std::min<size_t>(AOS_CHUNK / row, (size_t)n));
    GG_TRY(aos_stage(pkd, (size_t)C * row));
    AosStage &A = pkd->aos;
    std::vector<size_t> seg(nf);
    for (int k = 0, o = 0; k < nf; o += (int)(w[k] * C), ++k) seg[k] = (size_t)o;
    const int nc = (n + C - 1) / C;
    auto issue = [&](int c) -> int {
        const int b = c & 1, r0 = c * C, r1 = std::min(n, r0 + C);
        char *stage = (char *)A.buf[b];
        for (int k = 0; k < nf; ++k)
            GG_HIP(hipMemcpyAsync(stage + seg[k], (char *)pkd->f[f[k].field].p + (size_t)r0 * w[k],
                                  (size_t)(r1 - r0) * w[k], hipMemcpyDeviceToHost, pkd->stream));
        GG_HIP(hipEventRecord(A.ev[b], pkd->stream));
        return GG_OK;
    };
    GG_TRY(issue(0));
    for (int c = 0; c < nc; ++c) {
        // chunk c+1 lands in the other buffer, whose chunk c-1 is unpacked
        if (c + 1 < nc) GG_TRY(issue(c + 1));
        GG_HIP(hipEventSynchronize(A.ev[c & 1]));
        const int r0 = c * C, r1 = std::min(n, r0 + C);
        aos_move((char *)base, stride, f, nf, w.data(), seg.data(), r0, r1, (char *)A.buf[c & 1],
                 false);
    }
    GG_HIP(hipStreamSynchronize(pkd->stream));
    return GG_OK;
}

extern "C" int gg_timings(GGPKD pkd, double *t, int n)
{
    GG_CHECK(pkd && t, GG_EINVAL, "null argument");
    timer_resolve(pkd);
    for (int k = 0; k < n && k < 12; ++k) t[k] = pkd->tm.ms[k];
    return GG_OK;
}
