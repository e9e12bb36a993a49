// Sharpness metric of the dispersion search: per B-scan, the log of the mean of the brightest
// fraction of intensity pixels I = |A|^2 below the first zmin depth rows.
// Reference: octrecon/estimation/dispersion.py:46-63 (_metric 'top_percentile': I = A^2 in the
// magnitude's float32, cast to float64; k = max(1, int(M * 0.01)); mean of the k largest; log).
//
#pragma once
#include "oct_common.hpp"

namespace oct {

// Exact top-k by radix selection on the IEEE bit patterns of I (non-negative floats order as
// unsigned integers), spread over many workgroups per B-scan: three histogram passes fix the
// bits [20, 31), [10, 20) and [0, 10) of the k-th largest key (each followed by a one-block
// select), then one pass sums everything above it in float64 (per-workgroup partials, reduced
// in a fixed order), and the ties at the threshold are added.  Deterministic: counts are
// integers and every float64 sum runs in a fixed order.
constexpr int kTfThreads = 256;
constexpr int kTfBins = 2048;

struct TopfracState {
    unsigned prefix;      // key bits fixed so far
    int pad;
    long long k_rem;      // rank (from the top, 1-based) of the threshold among keys sharing the prefix
};

// Selection key of element e (x-major over columns z >= zmin): the float32 bits of |A|^2 (the
// top-fraction metric squares in float32) or of |A| itself (order statistics of |A|, whose
// order is that of the float64 squares).  Non-negative floats order as their unsigned bits.
__device__ __forceinline__ unsigned tf_key(const float* A, long long e, int w, int nz, int zmin, bool square) {
    const long long x = e / w, z = zmin + e - x * w;
    const float a = A[x * nz + z];
    return __builtin_bit_cast(unsigned, square ? a * a : a);
}

__global__ __launch_bounds__(kTfThreads) void topfrac_init_kernel(unsigned* __restrict__ hist, TopfracState* st,
                                                                  long long kk, int B) {
    const long long i = (long long)blockIdx.x * kTfThreads + threadIdx.x;
    if (i < (long long)B * kTfBins) hist[i] = 0u;
    if (i < B) {
        st[i].prefix = 0u;
        st[i].k_rem = kk;
    }
}

// pass: bits [lo, lo + nbits) of keys whose bits above lo + nbits equal the prefix.  Grid (S, V):
// selection state b works on B-scan b / per (per > 1: several ranks of one B-scan).
__global__ __launch_bounds__(kTfThreads) void topfrac_hist_kernel(const float* __restrict__ mag, int nX, int nz,
                                                                  int zmin, const TopfracState* __restrict__ st,
                                                                  unsigned* __restrict__ hist, int lo, int nbits,
                                                                  int per, int square) {
    __shared__ unsigned h[kTfBins];
    const int b = blockIdx.y;
    for (int i = threadIdx.x; i < kTfBins; i += kTfThreads) h[i] = 0u;
    __syncthreads();
    const float* A = mag + (long long)(b / per) * nX * nz;
    const int w = nz - zmin;
    const long long M = (long long)nX * w;
    const int hi = lo + nbits;
    const unsigned prefix = st[b].prefix;
    const unsigned dmask = (1u << nbits) - 1u;
    for (long long e = (long long)blockIdx.x * kTfThreads + threadIdx.x; e < M; e += (long long)gridDim.x * kTfThreads) {
        const unsigned u = tf_key(A, e, w, nz, zmin, square != 0);
        if (hi >= 32 || (u >> hi) == (prefix >> hi)) atomicAdd(&h[(u >> lo) & dmask], 1u);
    }
    __syncthreads();
    unsigned* g = hist + (long long)b * kTfBins;
    for (int i = threadIdx.x; i <= (int)dmask; i += kTfThreads)
        if (h[i]) atomicAdd(&g[i], h[i]);
}

// One block per B-scan: the digit d holding the k_rem-th largest key; prefix |= d << lo,
// k_rem -= (count of keys above d); clears the histogram for the next pass.
__global__ __launch_bounds__(kTfThreads) void topfrac_select_kernel(unsigned* __restrict__ hist, TopfracState* st,
                                                                    int lo, int nbits) {
    constexpr int kPer = kTfBins / kTfThreads;     // bins per thread, descending digit order
    __shared__ unsigned long long tot[kTfThreads];
    __shared__ int s_d;
    __shared__ long long s_above;
    const int b = blockIdx.x;
    const int nb = 1 << nbits;
    unsigned* g = hist + (long long)b * kTfBins;
    unsigned c[kPer];
    unsigned long long t = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int d = nb - 1 - (threadIdx.x * kPer + q);      // thread 0 holds the largest digits
        c[q] = d >= 0 ? g[d] : 0u;
        t += c[q];
    }
    tot[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        const long long k = st[b].k_rem;
        long long above = 0;
        int th = 0;
        for (; th < kTfThreads - 1; ++th) {
            if (above + (long long)tot[th] >= k) break;
            above += (long long)tot[th];
        }
        s_d = th;
        s_above = above;
    }
    __syncthreads();
    if ((int)threadIdx.x == s_d) {
        const long long k = st[b].k_rem;
        long long above = s_above;
        int q = 0;
        for (; q < kPer - 1; ++q) {
            if (above + (long long)c[q] >= k) break;
            above += (long long)c[q];
        }
        const int d = nb - 1 - (threadIdx.x * kPer + q);
        st[b].prefix |= (unsigned)d << lo;
        st[b].k_rem = k - above;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += kTfThreads) g[i] = 0u;
}

// Sum in float64 of the keys above the threshold: per-block partial part[b * S + s].  Grid (S, B).
__global__ __launch_bounds__(kTfThreads) void topfrac_sum_kernel(const float* __restrict__ mag, int nX, int nz,
                                                                 int zmin, const TopfracState* __restrict__ st,
                                                                 double* __restrict__ part) {
    __shared__ double red[kTfThreads];
    const int b = blockIdx.y;
    const float* A = mag + (long long)b * nX * nz;
    const int w = nz - zmin;
    const long long M = (long long)nX * w;
    const unsigned thr = st[b].prefix;
    double sum = 0.0;
    for (long long e = (long long)blockIdx.x * kTfThreads + threadIdx.x; e < M; e += (long long)gridDim.x * kTfThreads) {
        const unsigned u = tf_key(A, e, w, nz, zmin, true);
        if (u > thr) sum += (double)__builtin_bit_cast(float, u);
    }
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int s = kTfThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long long)b * gridDim.x + blockIdx.x] = red[0];
}

// out[b] = log((sum of partials in order + k_rem * threshold) / k).  Grid (ceil(B / 256)).
__global__ __launch_bounds__(kTfThreads) void topfrac_final_kernel(const TopfracState* __restrict__ st,
                                                                   const double* __restrict__ part, int S, long long kk,
                                                                   int B, double* __restrict__ out) {
    const int b = blockIdx.x * kTfThreads + threadIdx.x;
    if (b >= B) return;
    double total = 0.0;
    for (int s = 0; s < S; ++s) total += part[(long long)b * S + s];
    total += (double)st[b].k_rem * (double)__builtin_bit_cast(float, st[b].prefix);
    out[b] = log(total / (double)kk);
}

// Order statistics: state v takes its own rank (1-based, from the top) ranks[v].
__global__ __launch_bounds__(kTfThreads) void order_init_kernel(unsigned* __restrict__ hist, TopfracState* st,
                                                                const long long* __restrict__ ranks, int V) {
    const long long i = (long long)blockIdx.x * kTfThreads + threadIdx.x;
    if (i < (long long)V * kTfBins) hist[i] = 0u;
    if (i < V) {
        st[i].prefix = 0u;
        st[i].k_rem = ranks[i];
    }
}

// out[v] = the selected key of state v as a float (|A|^2 in float32, or |A|), in float64.
__global__ __launch_bounds__(kTfThreads) void order_final_kernel(const TopfracState* __restrict__ st, int V,
                                                                 double* __restrict__ out) {
    const int v = blockIdx.x * kTfThreads + threadIdx.x;
    if (v < V) out[v] = (double)__builtin_bit_cast(float, st[v].prefix);
}

// Moment metrics (octrecon/estimation/dispersion.py:58-62), I = |A|^2 squared in float32 and
// widened to float64, over x and z >= zmin.  Pass 1: per-workgroup partial sums of I and I^2,
// part[(b S + s) 2 + {0, 1}].  Grid (S, B).
__global__ __launch_bounds__(kTfThreads) void moments_kernel(const float* __restrict__ mag, int nX, int nz, int zmin,
                                                             double* __restrict__ part) {
    __shared__ double r1[kTfThreads], r2[kTfThreads];
    const int b = blockIdx.y;
    const float* A = mag + (long long)b * nX * nz;
    const int w = nz - zmin;
    const long long M = (long long)nX * w;
    double s1 = 0.0, s2 = 0.0;
    for (long long e = (long long)blockIdx.x * kTfThreads + threadIdx.x; e < M; e += (long long)gridDim.x * kTfThreads) {
        const double I = (double)__builtin_bit_cast(float, tf_key(A, e, w, nz, zmin, true));
        s1 += I;
        s2 += I * I;
    }
    r1[threadIdx.x] = s1;
    r2[threadIdx.x] = s2;
    __syncthreads();
    for (int h = kTfThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            r1[threadIdx.x] += r1[threadIdx.x + h];
            r2[threadIdx.x] += r2[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[((long long)b * gridDim.x + blockIdx.x) * 2] = r1[0];
        part[((long long)b * gridDim.x + blockIdx.x) * 2 + 1] = r2[0];
    }
}

// Pass 2 of the entropy metric: partial sums of p log(p + 1e-30), p = I / sum(I) (the total
// from pass 1's partials, added in a fixed order by every workgroup).  ent[b S + s].  Grid (S, B).
__global__ __launch_bounds__(kTfThreads) void entropy_kernel(const float* __restrict__ mag, int nX, int nz, int zmin,
                                                             const double* __restrict__ part, double* __restrict__ ent) {
    __shared__ double red[kTfThreads];
    const int b = blockIdx.y;
    const int S = gridDim.x;
    double tot = 0.0;
    for (int s = 0; s < S; ++s) tot += part[((long long)b * S + s) * 2];
    const float* A = mag + (long long)b * nX * nz;
    const int w = nz - zmin;
    const long long M = (long long)nX * w;
    double acc = 0.0;
    for (long long e = (long long)blockIdx.x * kTfThreads + threadIdx.x; e < M; e += (long long)gridDim.x * kTfThreads) {
        const double p = (double)__builtin_bit_cast(float, tf_key(A, e, w, nz, zmin, true)) / tot;
        acc += p * log(p + 1e-30);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kTfThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) ent[(long long)b * S + blockIdx.x] = red[0];
}

// kind 1 (kurtosis): sum I^2 / (sum I)^2; kind 2 (entropy): sum of the pass-2 partials.
// Partials added in a fixed order.  Grid (ceil(B / 256)).
__global__ __launch_bounds__(kTfThreads) void moments_final_kernel(const double* __restrict__ part,
                                                                   const double* __restrict__ ent, int S, int B,
                                                                   int kind, double* __restrict__ out) {
    const int b = blockIdx.x * kTfThreads + threadIdx.x;
    if (b >= B) return;
    double s1 = 0.0, s2 = 0.0;
    if (kind == 1) {
        for (int s = 0; s < S; ++s) {
            s1 += part[((long long)b * S + s) * 2];
            s2 += part[((long long)b * S + s) * 2 + 1];
        }
        out[b] = s2 / (s1 * s1);
    } else {
        for (int s = 0; s < S; ++s) s1 += ent[(long long)b * S + s];
        out[b] = s1;
    }
}

// Legacy glass-slide metric (octrecon/estimation/dispersion.py:49-51): per depth z >= zmin the
// mean over x of log(|A| + 1e-12) (the sum of float32 |A| + 1e-12 in float32, then log and mean
// in float64), L[b, z - zmin]; then the maximum over z.  Grid (ceil(w / 256), B): one column per
// thread, so the x loop reads along z across the workgroup.
__global__ __launch_bounds__(kTfThreads) void logmean_kernel(const float* __restrict__ mag, int nX, int nz, int zmin,
                                                             double* __restrict__ L) {
    const int b = blockIdx.y;
    const int w = nz - zmin;
    const int z = blockIdx.x * kTfThreads + threadIdx.x;
    if (z >= w) return;
    const float* A = mag + (long long)b * nX * nz + zmin + z;
    double acc = 0.0;
    for (int x = 0; x < nX; ++x) acc += log((double)(A[(long long)x * nz] + 1e-12f));
    L[(long long)b * w + z] = acc / (double)nX;
}

// out[b] = max_z L[b, z]; one workgroup per B-scan.
__global__ __launch_bounds__(kTfThreads) void rowmax_kernel(const double* __restrict__ L, int w,
                                                            double* __restrict__ out) {
    __shared__ double red[kTfThreads];
    const int b = blockIdx.x;
    double m = -INFINITY;
    for (int z = threadIdx.x; z < w; z += kTfThreads) m = fmax(m, L[(long long)b * w + z]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int h = kTfThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = red[0];
}

// Dispersion search windows: out[b, j] = hann_n[j] exp(-i beta_b dk2[j]) (hann_n = Hann/rms/N,
// dk2 = (k - mean k)^2 of the equispaced k grid, octrecon/estimation/dispersion.py:66-107), float64
// phase, stored as float32 pairs in the pair order of the spectral kernels (element j at float2
// position pair_pos(j) of window b; n = 2048), with
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
    reinterpret_cast<float2*>(out)[(long long)b * n + pair_pos(j)] = w;           // the pair order
}

}  // namespace oct
