// Sharpness metric of the dispersion search: per B-scan, the log of the mean of the brightest
// fraction of intensity pixels I = |A|^2 below the first zmin depth rows.
// Reference: octrecon/estimation/dispersion.py:46-63 (_metric 'top_percentile': I = A^2 in the
// magnitude's float32, cast to float64; k = max(1, int(M * 0.01)); mean of the k largest; log).
//
// Exact top-k by radix selection on the IEEE bit patterns of I (non-negative floats order as
// unsigned integers): four 8-bit passes narrow the k-th largest value, then one pass sums
// everything above it in float64 and adds the ties.  One workgroup per B-scan.
#pragma once
#include "oct_common.hpp"

namespace oct {

constexpr int kSelThreads = 1024;

__global__ __launch_bounds__(kSelThreads) void topfrac_metric_kernel(const float* __restrict__ mag, int nX, int nz,
                                                                     int zmin, long long kk, double* __restrict__ out) {
    const int b = blockIdx.x;
    const float* A = mag + (long long)b * nX * nz;
    const int w = nz - zmin;
    const long long M = (long long)nX * w;
    __shared__ unsigned hist[256];
    __shared__ unsigned s_prefix, s_mask;
    __shared__ long long s_k;
    __shared__ double s_sum[kSelThreads];
    __shared__ unsigned long long s_cnt[kSelThreads];
    if (threadIdx.x == 0) {
        s_prefix = 0u;
        s_mask = 0u;
        s_k = kk;          // rank (1-based, from the top) of the value searched for
    }
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 24 - 8 * pass;
        for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0u;
        __syncthreads();
        const unsigned prefix = s_prefix, mask = s_mask;
        for (long long e = threadIdx.x; e < M; e += blockDim.x) {
            const long long x = e / w, z = zmin + e - x * w;
            const float a = A[x * nz + z];
            const unsigned u = __builtin_bit_cast(unsigned, a * a);
            if ((u & mask) == prefix) atomicAdd(&hist[(u >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            long long k = s_k;
            int d = 255;
            for (; d > 0; --d) {
                if ((long long)hist[d] >= k) break;
                k -= hist[d];
            }
            s_prefix = prefix | ((unsigned)d << shift);
            s_mask = mask | (255u << shift);
            s_k = k;
        }
        __syncthreads();
    }
    const unsigned thr = s_prefix;        // bit pattern of the k-th largest value
    double sum = 0.0;
    unsigned long long cnt = 0;
    for (long long e = threadIdx.x; e < M; e += blockDim.x) {
        const long long x = e / w, z = zmin + e - x * w;
        const float a = A[x * nz + z];
        const float v = a * a;
        if (__builtin_bit_cast(unsigned, v) > thr) {
            sum += (double)v;
            ++cnt;
        }
    }
    s_sum[threadIdx.x] = sum;
    s_cnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = kSelThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + s];
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double tv = (double)__builtin_bit_cast(float, thr);
        const double total = s_sum[0] + (double)(kk - (long long)s_cnt[0]) * tv;
        out[b] = log(total / (double)kk);
    }
}

// Dispersion search windows: out[b, j] = hann_n[j] exp(-i beta_b dk2[j]) (hann_n = Hann/rms/N,
// dk2 = (k - mean k)^2 of the equispaced k grid, octrecon/estimation/dispersion.py:66-107), float64
// phase, stored as float32 pairs, with
//   beta_b = ((head ? head[b * head_stride] * head_scale : betas[b]) + offset) * to_ref.
// Grid (ceil(n / 256), count).
__global__ __launch_bounds__(256) void windows_kernel(const double* __restrict__ betas, const float* __restrict__ head,
                                                      int head_stride, double head_scale, double offset,
                                                      double to_ref, const double* __restrict__ hann_n,
                                                      const double* __restrict__ dk2, int n,
                                                      float* __restrict__ out) {
    const int b = blockIdx.y;
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const double beta = ((head ? (double)head[(long long)b * head_stride] * head_scale : betas[b]) + offset) * to_ref;
    double s, c;
    sincos(beta * dk2[j], &s, &c);
    float2 w;
    w.x = (float)(hann_n[j] * c);
    w.y = (float)(-hann_n[j] * s);
    reinterpret_cast<float2*>(out)[(long long)b * n + j] = w;
}

}  // namespace oct
