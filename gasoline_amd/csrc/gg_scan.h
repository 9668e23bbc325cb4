// Exclusive prefix sums on the device, over any per-element input (an array,
// or flags computed on the fly by the caller's functor).
//
// Three passes: per-block exclusive sums (256 threads x 8 consecutive
// elements, a wavefront shuffle scan and a combine of the 4 wavefront
// totals), an exclusive scan of the block totals by one workgroup, and the
// block offsets added back.
#pragma once

#include "gg_internal.h"

namespace gg {

constexpr int SCAN_T = 256, SCAN_V = 8, SCAN_E = SCAN_T * SCAN_V;

struct ScanArray {
    const int *p;
    __device__ __forceinline__ int operator()(int64_t i) const { return p[i]; }
};

// Exclusive prefix of v over the workgroup (every thread must call);
// *btot = the workgroup's total.
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t *ws, int64_t *btot)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t t = __shfl_up(inc, o);
        if (lane >= o) inc += t;
    }
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    int64_t off = 0, tot = 0;
    for (int k = 0; k < SCAN_T / 64; ++k) {
        if (k < w) off += ws[k];
        tot += ws[k];
    }
    __syncthreads();
    *btot = tot;
    return off + inc - v;
}

// Inputs are read, and outputs written, with consecutive lanes on
// consecutive elements; LDS turns that into 8 consecutive elements per
// thread for the scan itself.
template <class In>
__global__ __launch_bounds__(SCAN_T) void k_scan_block(int n, In in, int64_t *__restrict__ out,
                                                       int64_t *__restrict__ bsum)
{
    __shared__ int64_t ws[SCAN_T / 64];
    __shared__ int64_t buf[SCAN_E];
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_E;
    for (int k = 0; k < SCAN_V; ++k) {
        const int e = k * SCAN_T + threadIdx.x;
        buf[e] = b0 + e < n ? in(b0 + e) : 0;
    }
    __syncthreads();
    int64_t v[SCAN_V];
    int64_t sum = 0;
    for (int k = 0; k < SCAN_V; ++k) {
        v[k] = buf[threadIdx.x * SCAN_V + k];
        sum += v[k];
    }
    int64_t tot;
    int64_t ex = block_excl_scan(sum, ws, &tot);
    for (int k = 0; k < SCAN_V; ++k) {
        buf[threadIdx.x * SCAN_V + k] = ex;
        ex += v[k];
    }
    __syncthreads();
    for (int k = 0; k < SCAN_V; ++k) {
        const int e = k * SCAN_T + threadIdx.x;
        if (b0 + e < n) out[b0 + e] = buf[e];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Exclusive scan of the block totals by one workgroup, carrying across chunks.
static __global__ __launch_bounds__(SCAN_T) void k_scan_sums(int nb, int64_t *__restrict__ bsum,
                                                             int64_t *__restrict__ total)
{
    __shared__ int64_t ws[SCAN_T / 64];
    int64_t carry = 0;
    for (int base = 0; base < nb; base += SCAN_E) {
        const int b0 = base + threadIdx.x * SCAN_V;
        int64_t v[SCAN_V];
        int64_t sum = 0;
        for (int k = 0; k < SCAN_V; ++k) {
            v[k] = b0 + k < nb ? bsum[b0 + k] : 0;
            sum += v[k];
        }
        int64_t tot;
        int64_t ex = carry + block_excl_scan(sum, ws, &tot);
        for (int k = 0; k < SCAN_V; ++k) {
            if (b0 + k < nb) bsum[b0 + k] = ex;
            ex += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

static __global__ void k_scan_add(int n, const int64_t *__restrict__ bsum,
                                  int64_t *__restrict__ out, const int64_t *__restrict__ total)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += bsum[i / SCAN_E];
    if (i == n) out[n] = *total;
}

// out[0..n] = exclusive prefix sums of in(0..n-1); out[n] = the total, also
// copied to *total_host (after a stream synchronisation) unless it is null.
template <class In>
int exclusive_scan(hipStream_t st, In in, int64_t *out, int n, DevBuf &scratch,
                   int64_t *total_host)
{
    const int nb = (n + SCAN_E - 1) / SCAN_E;
    GG_TRY(scratch.reserve(sizeof(int64_t) * ((size_t)nb + 2)));
    int64_t *bsum = scratch.as<int64_t>();
    int64_t *tot = bsum + nb + 1;
    if (nb > 0) {
        hipLaunchKernelGGL(k_scan_block<In>, dim3(nb), dim3(SCAN_T), 0, st, n, in, out, bsum);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_T), 0, st, nb, bsum, tot);
        hipLaunchKernelGGL(k_scan_add, dim3((int)(((int64_t)n + 1 + 255) / 256)), dim3(256), 0,
                           st, n, bsum, out, tot);
        GG_HIP(hipGetLastError());
        if (total_host) {
            GG_HIP(hipMemcpyAsync(total_host, tot, sizeof(int64_t), hipMemcpyDeviceToHost, st));
            GG_HIP(hipStreamSynchronize(st));
        }
    } else {
        GG_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), st));
        if (total_host) *total_host = 0;
    }
    return GG_OK;
}

}  // namespace gg
