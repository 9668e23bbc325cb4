// h-proposal MLP: a 4-32-32-1 tanh network in FP32 that proposes each gas
// particle's next smoothing length from its previous h, density, velocity
// divergence and the step dt.  It only seeds the kNN search radius (the
// fBall2 hint that gg_smooth reads), so the neighbour sets stay exact
// whatever it predicts.  It has no counterpart in the reference.
//
// Parameter vector (HMLP_NP floats): W1[32][4], b1[32], W2[32][32], b2[32],
// W3[32], b3, then the input mean[4] and scale[4].  The network maps the
// standardised features x^ = (x - mean) / scale of
//   x = (log h, log rho, divv * dt, log dt)
// to y = log h_next - log h, through a1 = tanh(W1 x^ + b1),
// a2 = tanh(W2 a1 + b2), y = W3 . a2 + b3.  Training minimises the mean of
// (y - t)^2 over a minibatch with Adam; gradients are reduced in a fixed
// order, so a training run is reproducible bit for bit.
#include "gg_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace gg {
namespace {

constexpr int H = HMLP_H;
constexpr int OW1 = 0, OB1 = OW1 + 4 * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + H, OMU = HMLP_NW, OSD = HMLP_NW + 4;
static_assert(OB3 + 1 == HMLP_NW, "parameter layout");
constexpr int FB = 64;        // samples per block of the gradient pass
constexpr int NBMAX = 4096;   // gradient blocks per minibatch (minibatch <= FB * NBMAX)

int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

__device__ __forceinline__ uint32_t hash3(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ (c + 0x165667B1u) * 0xC2B2AE3Du;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return h;
}

// Forward pass of one sample with parameters in LDS; a1, a2 out.
__device__ __forceinline__ float mlp_forward(const float *P, const float x[4], float a1[H],
                                             float a2[H], float xh[4])
{
    for (int i = 0; i < 4; ++i) xh[i] = (x[i] - P[OMU + i]) / P[OSD + i];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        float z = P[OB1 + j];
        for (int i = 0; i < 4; ++i) z += P[OW1 + 4 * j + i] * xh[i];
        a1[j] = tanhf(z);
    }
    float y = P[OB3];
#pragma unroll
    for (int k = 0; k < H; ++k) {
        float z = P[OB2 + k];
        for (int j = 0; j < H; ++j) z += P[OW2 + H * k + j] * a1[j];
        a2[k] = tanhf(z);
        y += P[OW3 + k] * a2[k];
    }
    return y;
}

// Output only (no activations kept).
__device__ __forceinline__ float mlp_y(const float *P, const float x[4])
{
    float xh[4], a1[H];
    for (int i = 0; i < 4; ++i) xh[i] = (x[i] - P[OMU + i]) / P[OSD + i];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        float z = P[OB1 + j];
        for (int i = 0; i < 4; ++i) z += P[OW1 + 4 * j + i] * xh[i];
        a1[j] = tanhf(z);
    }
    float y = P[OB3];
    for (int k = 0; k < H; ++k) {
        float z = P[OB2 + k];
#pragma unroll
        for (int j = 0; j < H; ++j) z += P[OW2 + H * k + j] * a1[j];
        y += P[OW3 + k] * tanhf(z);
    }
    return y;
}

__device__ __forceinline__ void load_params(const float *par, float *P)
{
    for (int p = threadIdx.x; p < HMLP_NP; p += blockDim.x) P[p] = par[p];
    __syncthreads();
}

// Features of particle i after a density pass.
__global__ void k_record(int n, const uint32_t *__restrict__ bits, const double *__restrict__ fb2,
                         const double *__restrict__ rho, const double *__restrict__ divv,
                         double dt, float *__restrict__ hx, int *__restrict__ hv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool ok = (bits[i] & GG_TYPE_SMOOTHDONE) && fb2[i] > 0 && rho[i] > 0 &&
                    isfinite(divv[i]) && isfinite(fb2[i]) && isfinite(rho[i]);
    hv[i] = ok;
    if (!ok) return;
    hx[i] = (float)(0.5 * log(fb2[i]) - M_LN2);   // h = sqrt(fBall2) / 2
    hx[(size_t)n + i] = (float)log(rho[i]);
    hx[2 * (size_t)n + i] = (float)(divv[i] * dt);
    hx[3 * (size_t)n + i] = (float)log(dt);
}

// Training samples: history valid and smoothed again since.
__global__ void k_select(int n, const int *__restrict__ hv, const uint32_t *__restrict__ bits,
                         const double *__restrict__ fb2, int *__restrict__ sel)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sel[i] = hv[i] && (bits[i] & GG_TYPE_SMOOTHDONE) && fb2[i] > 0 && isfinite(fb2[i]);
}

__global__ void k_compact(int n, const int *__restrict__ sel, const int64_t *__restrict__ pos,
                          int *__restrict__ samp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && sel[i]) samp[pos[i]] = i;
}

// Per-block sums of each feature and its square over the samples.
__global__ void k_stats(int ns, const int *__restrict__ samp, const float *__restrict__ hx, int n,
                        double *__restrict__ part)
{
    __shared__ double s[8][256];
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = blockIdx.x * 256 + threadIdx.x; k < ns; k += gridDim.x * 256) {
        const int i = samp[k];
        for (int f = 0; f < 4; ++f) {
            const double v = hx[(size_t)f * n + i];
            a[f] += v;
            a[4 + f] += v * v;
        }
    }
    for (int f = 0; f < 8; ++f) s[f][threadIdx.x] = a[f];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int f = 0; f < 8; ++f) s[f][threadIdx.x] += s[f][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int f = 0; f < 8; ++f) part[8 * (size_t)blockIdx.x + f] = s[f][0];
}

struct GArgs {
    const float *par;
    int nb;                     // minibatch size
    int ns;                     // samples available
    const int *samp;            // sample -> particle (nullptr: identity)
    uint32_t seed, iter;        // minibatch draw (seed 0: samples 0..nb-1 in order)
    const float *x;             // features, SoA [4][ld]
    int ld;
    const float *tgt;           // targets (nullptr: from fb2 and x[0])
    const double *fb2;
    float *gpart;               // [blocks][HMLP_NW]
    double *lpart;              // [blocks]
};

// Gradient of the mean squared error over the minibatch, one sample per
// lane, FB samples per block; per-block partial sums in a fixed order.
__global__ __launch_bounds__(FB) void k_grad(GArgs A)
{
    __shared__ float P[HMLP_NP];
    __shared__ float sx[FB][4], s1[FB][H + 1], s2[FB][H + 1], d1[FB][H + 1], d2[FB][H + 1];
    __shared__ float d3[FB];
    __shared__ double sl[FB];
    load_params(A.par, P);
    const int t = threadIdx.x;
    const int k = blockIdx.x * FB + t;
    float a1[H], a2[H], xh[4] = {0, 0, 0, 0};
    float e = 0.f;
    bool ok = k < A.nb;
    if (ok) {
        int s = A.seed ? (int)(hash3(A.seed, A.iter, (uint32_t)k) % (uint32_t)A.ns) : k;
        const int i = A.samp ? A.samp[s] : s;
        float x[4];
        for (int f = 0; f < 4; ++f) x[f] = A.x[(size_t)f * A.ld + i];
        const float y = mlp_forward(P, x, a1, a2, xh);
        const float tv = A.tgt ? A.tgt[i] : (float)(0.5 * log(A.fb2[i]) - M_LN2 - (double)x[0]);
        e = y - tv;
    } else {
        for (int j = 0; j < H; ++j) a1[j] = a2[j] = 0.f;
    }
    const float g3 = ok ? 2.f * e / (float)A.nb : 0.f;
    float g2[H];
#pragma unroll
    for (int q = 0; q < H; ++q) g2[q] = g3 * P[OW3 + q] * (1.f - a2[q] * a2[q]);
#pragma unroll
    for (int j = 0; j < H; ++j) {
        float b = 0.f;
        for (int q = 0; q < H; ++q) b += P[OW2 + H * q + j] * g2[q];
        d1[t][j] = b * (1.f - a1[j] * a1[j]);
        s1[t][j] = a1[j];
        s2[t][j] = a2[j];
        d2[t][j] = g2[j];
    }
    for (int f = 0; f < 4; ++f) sx[t][f] = xh[f];
    d3[t] = g3;
    sl[t] = ok ? (double)e * (double)e : 0.0;
    __syncthreads();
    float *gp = A.gpart + (size_t)blockIdx.x * HMLP_NW;
    for (int p = t; p < HMLP_NW; p += FB) {
        float g = 0.f;
        if (p < OB1) {
            const int j = (p - OW1) >> 2, i = (p - OW1) & 3;
            for (int s = 0; s < FB; ++s) g += d1[s][j] * sx[s][i];
        } else if (p < OW2) {
            for (int s = 0; s < FB; ++s) g += d1[s][p - OB1];
        } else if (p < OB2) {
            const int q = (p - OW2) / H, j = (p - OW2) % H;
            for (int s = 0; s < FB; ++s) g += d2[s][q] * s1[s][j];
        } else if (p < OW3) {
            for (int s = 0; s < FB; ++s) g += d2[s][p - OB2];
        } else if (p < OB3) {
            for (int s = 0; s < FB; ++s) g += d3[s] * s2[s][p - OW3];
        } else {
            for (int s = 0; s < FB; ++s) g += d3[s];
        }
        gp[p] = g;
    }
    if (t == 0) {
        double l = 0;
        for (int s = 0; s < FB; ++s) l += sl[s];
        A.lpart[blockIdx.x] = l;
    }
}

// Sum of the block partials in block order; loss = sum / nb.
__global__ void k_reduce(int nblocks, int nb, const float *__restrict__ gpart,
                         const double *__restrict__ lpart, float *__restrict__ grad,
                         double *__restrict__ loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < HMLP_NW) {
        float g = 0.f;
        for (int b = 0; b < nblocks; ++b) g += gpart[(size_t)b * HMLP_NW + p];
        grad[p] = g;
    }
    if (p == 0) {
        double l = 0;
        for (int b = 0; b < nblocks; ++b) l += lpart[b];
        *loss = l / nb;
    }
}

__global__ void k_adam(float *__restrict__ par, const float *__restrict__ grad,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                       float b2, float c1, float c2, float eps)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= HMLP_NW) return;
    const float g = grad[p];
    const float mp = b1 * m[p] + (1.f - b1) * g;
    const float vp = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mp;
    v[p] = vp;
    par[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// Forward pass over samples: y[k] for x[.][k] (SoA, ld = n).
__global__ void k_eval(int n, const float *__restrict__ par, const float *__restrict__ x,
                       float *__restrict__ y)
{
    __shared__ float P[HMLP_NP];
    load_params(par, P);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float xi[4];
    for (int f = 0; f < 4; ++f) xi[f] = x[(size_t)f * n + k];
    y[k] = mlp_y(P, xi);
}

// fBall2 seed (2 h_next)^2 from the history features, for particles with a
// valid history; others keep their stored fBall2.
__global__ void k_predict(int n, const float *__restrict__ par, const float *__restrict__ hx,
                          const int *__restrict__ hv, double *__restrict__ fb2)
{
    __shared__ float P[HMLP_NP];
    load_params(par, P);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !hv[i]) return;
    float xi[4];
    for (int f = 0; f < 4; ++f) xi[f] = hx[(size_t)f * n + i];
    const float y = mlp_y(P, xi);
    if (!isfinite(y)) return;
    const double h = exp((double)xi[0] + (double)fminf(fmaxf(y, -5.f), 5.f));
    fb2[i] = 4.0 * h * h;
}

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int ensure_init(GGPKD pkd)
{
    GG_CHECK(pkd->mlp.init, GG_ESTATE, "gg_hmlp_init has not been called");
    return GG_OK;
}

// Gradient over a minibatch into mlp.grad; *loss (host) the minibatch loss.
int run_grad(GGPKD pkd, const GArgs &A0, double *loss)
{
    HmlpDev &M = pkd->mlp;
    hipStream_t st = pkd->stream;
    const int nblocks = nblk(A0.nb, FB);
    GG_CHECK(nblocks <= NBMAX, GG_EINVAL, "minibatch %d exceeds %d", A0.nb, FB * NBMAX);
    GG_TRY(M.gpart.reserve(sizeof(float) * HMLP_NW * (size_t)nblocks));
    GG_TRY(M.lpart.reserve(sizeof(double) * (nblocks + 1)));
    GArgs A = A0;
    A.par = M.par.as<float>();
    A.gpart = M.gpart.as<float>();
    A.lpart = M.lpart.as<double>();
    hipLaunchKernelGGL(k_grad, dim3(nblocks), dim3(FB), 0, st, A);
    hipLaunchKernelGGL(k_reduce, dim3(nblk(HMLP_NW, 256)), dim3(256), 0, st, nblocks, A.nb,
                       M.gpart.as<float>(), M.lpart.as<double>(), M.grad.as<float>(),
                       M.lpart.as<double>() + nblocks);
    GG_HIP(hipGetLastError());
    if (loss) {
        GG_HIP(hipMemcpyAsync(loss, M.lpart.as<double>() + nblocks, sizeof(double),
                              hipMemcpyDeviceToHost, st));
        GG_HIP(hipStreamSynchronize(st));
    }
    return GG_OK;
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_hmlp_init(GGPKD pkd, unsigned seed)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    HmlpDev &M = pkd->mlp;
    GG_TRY(M.par.reserve(sizeof(float) * HMLP_NP));
    GG_TRY(M.m.reserve(sizeof(float) * HMLP_NW));
    GG_TRY(M.v.reserve(sizeof(float) * HMLP_NW));
    GG_TRY(M.grad.reserve(sizeof(float) * HMLP_NW));
    std::vector<float> w(HMLP_NP, 0.f);
    uint64_t s = 0x243F6A8885A308D3ull ^ seed;
    auto uni = [&](double a) {
        return (float)(a * (2.0 * (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0) - 1.0));
    };
    const double a1 = std::sqrt(6.0 / (4 + H)), a2 = std::sqrt(6.0 / (H + H));
    for (int p = OW1; p < OB1; ++p) w[p] = uni(a1);
    for (int p = OW2; p < OB2; ++p) w[p] = uni(a2);
    // W3 = b3 = 0: the untrained network proposes h_next = h, the plain seed
    for (int f = 0; f < 4; ++f) w[OSD + f] = 1.f;
    GG_HIP(hipMemcpyAsync(M.par.p, w.data(), sizeof(float) * HMLP_NP, hipMemcpyHostToDevice,
                          pkd->stream));
    GG_HIP(hipMemsetAsync(M.m.p, 0, sizeof(float) * HMLP_NW, pkd->stream));
    GG_HIP(hipMemsetAsync(M.v.p, 0, sizeof(float) * HMLP_NW, pkd->stream));
    GG_HIP(hipStreamSynchronize(pkd->stream));
    M.init = 1;
    M.normSet = 0;
    M.nHist = 0;
    M.step = 0;
    return GG_OK;
}

extern "C" int gg_hmlp_params(GGPKD pkd, float *w, int n, int bSet)
{
    GG_CHECK(pkd && w, GG_EINVAL, "null argument");
    GG_TRY(ensure_init(pkd));
    GG_CHECK(n == HMLP_NP, GG_EINVAL, "the parameter vector holds %d floats, not %d", HMLP_NP, n);
    HmlpDev &M = pkd->mlp;
    if (bSet) {
        GG_HIP(hipMemcpyAsync(M.par.p, w, sizeof(float) * HMLP_NP, hipMemcpyHostToDevice,
                              pkd->stream));
        M.normSet = 1;
    } else {
        GG_HIP(hipMemcpyAsync(w, M.par.p, sizeof(float) * HMLP_NP, hipMemcpyDeviceToHost,
                              pkd->stream));
    }
    GG_HIP(hipStreamSynchronize(pkd->stream));
    return GG_OK;
}

extern "C" int gg_hmlp_record(GGPKD pkd, double dt)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GG_TRY(ensure_init(pkd));
    GG_CHECK(dt > 0 && std::isfinite(dt), GG_EINVAL, "dt must be positive");
    const int n = pkd->nLocal;
    GG_CHECK(n > 0, GG_EINVAL, "empty particle store");
    const int need[] = {GG_FBALL2, GG_DENSITY, GG_DIVV, GG_IACTIVE};
    for (int f : need) GG_CHECK(pkd->f[f].p, GG_ESTATE, "field %d missing: record after gg_smooth", f);
    HmlpDev &M = pkd->mlp;
    GG_TRY(M.hx.reserve(sizeof(float) * 4 * (size_t)n));
    GG_TRY(M.hv.reserve(sizeof(int) * (size_t)n));
    hipLaunchKernelGGL(k_record, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n,
                       pkd->f[GG_IACTIVE].as<uint32_t>(), pkd->f[GG_FBALL2].as<double>(),
                       pkd->f[GG_DENSITY].as<double>(), pkd->f[GG_DIVV].as<double>(), dt,
                       M.hx.as<float>(), M.hv.as<int>());
    GG_HIP(hipGetLastError());
    GG_HIP(hipStreamSynchronize(pkd->stream));
    M.nHist = n;
    return GG_OK;
}

extern "C" int gg_hmlp_train(GGPKD pkd, int nIter, int nBatch, double lr, unsigned seed,
                             double *pLossFirst, double *pLossLast, int *pnSamples)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GG_TRY(ensure_init(pkd));
    HmlpDev &M = pkd->mlp;
    const int n = pkd->nLocal;
    GG_CHECK(M.nHist == n && n > 0, GG_ESTATE,
             "no history for the %d particles: gg_hmlp_record after the previous gg_smooth", n);
    GG_CHECK(nIter >= 0 && nBatch >= 1 && nBatch <= FB * NBMAX, GG_EINVAL,
             "nIter >= 0 and 1 <= nBatch <= %d", FB * NBMAX);
    GG_CHECK(lr > 0, GG_EINVAL, "lr must be positive");
    hipStream_t st = pkd->stream;
    GG_TRY(M.sel.reserve(sizeof(int) * (size_t)n));
    GG_TRY(M.pos.reserve(sizeof(int64_t) * ((size_t)n + 1)));
    GG_TRY(M.samp.reserve(sizeof(int) * (size_t)n));
    hipLaunchKernelGGL(k_select, dim3(nblk(n, 256)), dim3(256), 0, st, n, M.hv.as<int>(),
                       pkd->f[GG_IACTIVE].as<uint32_t>(), pkd->f[GG_FBALL2].as<double>(),
                       M.sel.as<int>());
    int64_t ns64 = 0;
    GG_TRY(exclusive_scan_i64(st, M.sel.as<int>(), M.pos.as<int64_t>(), n, M.scan, &ns64));
    const int ns = (int)ns64;
    if (pnSamples) *pnSamples = ns;
    if (pLossFirst) *pLossFirst = 0;
    if (pLossLast) *pLossLast = 0;
    if (ns == 0 || nIter == 0) return GG_OK;
    hipLaunchKernelGGL(k_compact, dim3(nblk(n, 256)), dim3(256), 0, st, n, M.sel.as<int>(),
                       M.pos.as<int64_t>(), M.samp.as<int>());
    if (!M.normSet) {   // input standardisation from the first training set, then fixed
        const int nbs = std::min(nblk(ns, 256), 256);
        GG_TRY(M.stat.reserve(sizeof(double) * 8 * nbs));
        hipLaunchKernelGGL(k_stats, dim3(nbs), dim3(256), 0, st, ns, M.samp.as<int>(),
                           M.hx.as<float>(), n, M.stat.as<double>());
        std::vector<double> part(8 * (size_t)nbs);
        GG_HIP(hipMemcpyAsync(part.data(), M.stat.p, sizeof(double) * 8 * nbs,
                              hipMemcpyDeviceToHost, st));
        GG_HIP(hipStreamSynchronize(st));
        float ms[8];
        for (int f = 0; f < 4; ++f) {
            double a = 0, b = 0;
            for (int k = 0; k < nbs; ++k) {
                a += part[8 * (size_t)k + f];
                b += part[8 * (size_t)k + 4 + f];
            }
            const double mu = a / ns, var = std::max(b / ns - mu * mu, 0.0);
            ms[f] = (float)mu;
            ms[4 + f] = (float)std::max(std::sqrt(var), 1e-3);
        }
        GG_HIP(hipMemcpyAsync(M.par.as<float>() + OMU, ms, sizeof(float) * 8,
                              hipMemcpyHostToDevice, st));
        GG_HIP(hipStreamSynchronize(st));
        M.normSet = 1;
    }
    GArgs A{};
    A.nb = std::min(nBatch, ns);
    A.ns = ns;
    A.samp = M.samp.as<int>();
    A.seed = seed ? seed : 1u;
    A.x = M.hx.as<float>();
    A.ld = n;
    A.tgt = nullptr;
    A.fb2 = pkd->f[GG_FBALL2].as<double>();
    const float b1 = 0.9f, b2 = 0.999f;
    for (int it = 0; it < nIter; ++it) {
        A.iter = (uint32_t)M.step;   // minibatch k of Adam step s: hash(seed, s, k)
        double l = 0;
        const bool want = (it == 0 && pLossFirst) || (it == nIter - 1 && pLossLast);
        GG_TRY(run_grad(pkd, A, want ? &l : nullptr));
        if (it == 0 && pLossFirst) *pLossFirst = l;
        if (it == nIter - 1 && pLossLast) *pLossLast = l;
        ++M.step;
        const float c1 = (float)(1.0 - std::pow((double)b1, M.step));
        const float c2 = (float)(1.0 - std::pow((double)b2, M.step));
        hipLaunchKernelGGL(k_adam, dim3(nblk(HMLP_NW, 256)), dim3(256), 0, st, M.par.as<float>(),
                           M.grad.as<float>(), M.m.as<float>(), M.v.as<float>(), (float)lr, b1, b2,
                           c1, c2, 1e-8f);
        GG_HIP(hipGetLastError());
    }
    GG_HIP(hipStreamSynchronize(st));
    return GG_OK;
}

extern "C" int gg_hmlp_predict(GGPKD pkd)
{
    GG_CHECK(pkd, GG_EINVAL, "null handle");
    GG_TRY(ensure_init(pkd));
    HmlpDev &M = pkd->mlp;
    const int n = pkd->nLocal;
    GG_CHECK(M.nHist == n && n > 0, GG_ESTATE,
             "no history for the %d particles: gg_hmlp_record after the previous gg_smooth", n);
    GG_TRY(ensure_field(pkd, GG_FBALL2));
    hipLaunchKernelGGL(k_predict, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n,
                       M.par.as<float>(), M.hx.as<float>(), M.hv.as<int>(),
                       pkd->f[GG_FBALL2].as<double>());
    GG_HIP(hipGetLastError());
    GG_HIP(hipStreamSynchronize(pkd->stream));
    return GG_OK;
}

extern "C" int gg_hmlp_eval(GGPKD pkd, const float *x, int n, float *y)
{
    GG_CHECK(pkd && x && y && n > 0, GG_EINVAL, "null argument or n <= 0");
    GG_TRY(ensure_init(pkd));
    HmlpDev &M = pkd->mlp;
    DevBuf bx, by;
    GG_TRY(bx.reserve(sizeof(float) * 4 * (size_t)n));
    GG_TRY(by.reserve(sizeof(float) * (size_t)n));
    int rc = GG_OK;
    if (hipMemcpyAsync(bx.p, x, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice,
                       pkd->stream) != hipSuccess)
        rc = GG_EHIP;
    if (rc == GG_OK) {
        hipLaunchKernelGGL(k_eval, dim3(nblk(n, 256)), dim3(256), 0, pkd->stream, n,
                           M.par.as<float>(), bx.as<float>(), by.as<float>());
        if (hipMemcpyAsync(y, by.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost,
                           pkd->stream) != hipSuccess ||
            hipStreamSynchronize(pkd->stream) != hipSuccess)
            rc = GG_EHIP;
    }
    bx.release();
    by.release();
    GG_CHECK(rc == GG_OK, rc, "gg_hmlp_eval: HIP copy or launch failed");
    return GG_OK;
}

extern "C" int gg_hmlp_grad(GGPKD pkd, const float *x, const float *t, int n, float *g,
                            double *loss)
{
    GG_CHECK(pkd && x && t && g && n > 0, GG_EINVAL, "null argument or n <= 0");
    GG_CHECK(n <= FB * NBMAX, GG_EINVAL, "at most %d samples", FB * NBMAX);
    GG_TRY(ensure_init(pkd));
    HmlpDev &M = pkd->mlp;
    DevBuf bx, bt;
    GG_TRY(bx.reserve(sizeof(float) * 4 * (size_t)n));
    GG_TRY(bt.reserve(sizeof(float) * (size_t)n));
    int rc = GG_OK;
    if (hipMemcpyAsync(bx.p, x, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice,
                       pkd->stream) != hipSuccess ||
        hipMemcpyAsync(bt.p, t, sizeof(float) * (size_t)n, hipMemcpyHostToDevice,
                       pkd->stream) != hipSuccess)
        rc = GG_EHIP;
    if (rc == GG_OK) {
        GArgs A{};
        A.nb = n;
        A.ns = n;
        A.x = bx.as<float>();
        A.ld = n;
        A.tgt = bt.as<float>();
        rc = run_grad(pkd, A, loss);
        if (rc == GG_OK &&
            (hipMemcpyAsync(g, M.grad.p, sizeof(float) * HMLP_NW, hipMemcpyDeviceToHost,
                            pkd->stream) != hipSuccess ||
             hipStreamSynchronize(pkd->stream) != hipSuccess))
            rc = GG_EHIP;
    }
    bx.release();
    bt.release();
    GG_CHECK(rc == GG_OK, rc, "gg_hmlp_grad: HIP copy or launch failed");
    return GG_OK;
}
