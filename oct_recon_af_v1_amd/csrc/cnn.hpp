// Learned beta / focus estimator: BF16 convolutional network kernels (training and inference).
//
// Storage: activations and weights in bfloat16 (round to nearest even), master weights, gradients
// and every accumulation in float32.  Layers:
//   conv 3x3, stride 2, zero padding 1, ReLU      x (N, Ci, H, W) -> y (N, Co, H/2, W/2)
//   fully connected (+ ReLU or linear head)        x (N, K)        -> y (N, O)
// Backward kernels take the gradient of the layer's pre-activation output (dpre) and produce the
// gradient of its input already multiplied by the ReLU mask of the layer below (x > 0), i.e. the
// dpre of the layer below.  Every index below is checked against the shapes on the host.
//
// There is no counterpart in the reference (octrecon/estimation/dispersion.py:154-267 and
// focus.py:191-218 are the searches this network replaces; they remain the labels and the referee).
#pragma once
#include "oct_common.hpp"

namespace oct {

__device__ __forceinline__ float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)0x7FC0u;       // NaN stays NaN
    u += 0x7FFFu + ((u >> 16) & 1u);                                      // round to nearest even
    return (uint16_t)(u >> 16);
}

constexpr int kCnnThreads = 256;
constexpr int kConvCoB = 16;   // output channels per thread / wave (forward)
constexpr int kConvCiB = 16;   // input channels per thread (backward data)

// y[n, co, oy, ox] = relu(b[co] + sum_ci sum_{ky,kx} w[co, ci, ky, kx] x[n, ci, 2oy-1+ky, 2ox-1+kx])
// Block = an 8 x 32 tile of output positions of one image and CB (16, or 8 to double the blocks
// of a small launch) output channels; the input patch of the tile and the CB channels' weights
// are staged in LDS for 8 input channels at a time, and every thread forms its position's CB
// channels from LDS.  The patch rows are read as 33 aligned bf16 pairs (input columns
// 2 ox0 - 2 .. 2 ox0 + 63; W is even, so a pair is wholly inside or outside the image).  Staging
// is unrolled: a thread issues all its loads of a chunk before its first LDS store, so a chunk
// costs one round of load latency.  Summation order per output: ci ascending, then ky, kx.
// Grid (ceil(Wo / 32), ceil(Ho / 8), N Co / CB).
constexpr int kConvTY = 8, kConvTX = 32, kConvCC = 8;
constexpr int kConvPH = 2 * kConvTY + 1, kConvPP = kConvTX + 1, kConvPW = 2 * kConvPP;
constexpr int kConvPairs = kConvCC * kConvPH * kConvPP;                     // [c][py][pair]
constexpr int kConvPairsPT = (kConvPairs + kCnnThreads - 1) / kCnnThreads;  // per thread

template <int CB>
__global__ __launch_bounds__(kCnnThreads) void conv_fwd_kernel(const uint16_t* __restrict__ x,
                                                               const uint16_t* __restrict__ w,
                                                               const float* __restrict__ bias,
                                                               uint16_t* __restrict__ y, int N, int Ci, int H,
                                                               int W, int Co) {
    constexpr int kWts = kConvCC * 9 * CB;                                  // [c][tap][channel]
    constexpr int kWtsPT = (kWts + kCnnThreads - 1) / kCnnThreads;
    __shared__ float patch[kConvCC * kConvPH * kConvPW];
    __shared__ float wl[kWts];
    const int Ho = H >> 1, Wo = W >> 1;
    const int groups = Co / CB;
    const int n = blockIdx.z / groups, co0 = (blockIdx.z - (blockIdx.z / groups) * groups) * CB;
    const int oyb = blockIdx.y * kConvTY, oxb = blockIdx.x * kConvTX;
    const int ty = threadIdx.x / kConvTX, tx = threadIdx.x - (threadIdx.x / kConvTX) * kConvTX;
    const int iyb = 2 * oyb - 1, ixb = 2 * oxb - 2;                      // patch column 0 = input ixb
    // pair e = c (17 33) + py 33 + k holds input columns ixb + 2k, +1 of chunk channel c: its
    // offset in the chunk, or -1 outside the image (the same for every chunk)
    int offs[kConvPairsPT];
#pragma unroll
    for (int i = 0; i < kConvPairsPT; ++i) {
        const int e = threadIdx.x + i * kCnnThreads;
        const int c = e / (kConvPH * kConvPP);
        const int r = e - c * (kConvPH * kConvPP);
        const int py = r / kConvPP, k = r - (r / kConvPP) * kConvPP;
        const int iy = iyb + py, ix = ixb + 2 * k;
        offs[i] = ((unsigned)iy < (unsigned)H && (unsigned)ix < (unsigned)W) ? (c * H + iy) * W + ix : -1;
    }
    float acc[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) acc[j] = bias[co0 + j];
    for (int c0 = 0; c0 < Ci; c0 += kConvCC) {
        const int cc = min(kConvCC, Ci - c0);
        const uint16_t* xc = x + ((long long)n * Ci + c0) * H * W;
        const int cw = cc * 9;
        // all loads of the chunk first; a pair is inside the image and the chunk iff
        // 0 <= offset < cc H W (the others read the chunk's first word and are zeroed at the
        // store); weight e = j (cc 9) + c 9 + tap, the layout of w
        const unsigned lim = (unsigned)(cc * H * W);
        uint32_t pv[kConvPairsPT];
        uint16_t pw[kWtsPT];
#pragma unroll
        for (int i = 0; i < kConvPairsPT; ++i)
            pv[i] = *reinterpret_cast<const uint32_t*>(xc + ((unsigned)offs[i] < lim ? offs[i] : 0));
#pragma unroll
        for (int i = 0; i < kWtsPT; ++i) {
            const int e = threadIdx.x + i * kCnnThreads;
            const int j = e / cw, r = e - (e / cw) * cw;
            pw[i] = w[e < cw * CB ? ((long long)(co0 + j) * Ci + c0) * 9 + r : 0LL];
        }
        if (c0) __syncthreads();                                         // last chunk's reads done
#pragma unroll
        for (int i = 0; i < kConvPairsPT; ++i) {
            const int e = threadIdx.x + i * kCnnThreads;
            if (e < kConvPairs) {
                const uint32_t v = (unsigned)offs[i] < lim ? pv[i] : 0u;
                patch[2 * e] = bf2f((uint16_t)(v & 0xFFFFu));            // little endian: lower column first
                patch[2 * e + 1] = bf2f((uint16_t)(v >> 16));
            }
        }
#pragma unroll
        for (int i = 0; i < kWtsPT; ++i) {
            const int e = threadIdx.x + i * kCnnThreads;
            if (e < cw * CB) {
                const int j = e / cw, r = e - (e / cw) * cw;
                wl[r * CB + j] = bf2f(pw[i]);
            }
        }
        __syncthreads();
        for (int c = 0; c < cc; ++c) {
            float v[9];
#pragma unroll
            for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                for (int kx = 0; kx < 3; ++kx)
                    v[3 * ky + kx] = patch[(c * kConvPH + 2 * ty + ky) * kConvPW + 2 * tx + kx + 1];
#pragma unroll
            for (int k = 0; k < 9; ++k)
#pragma unroll
                for (int j = 0; j < CB; ++j) acc[j] = __builtin_fmaf(wl[(c * 9 + k) * CB + j], v[k], acc[j]);
        }
    }
    const int oy = oyb + ty, ox = oxb + tx;
    if (oy < Ho && ox < Wo) {
#pragma unroll
        for (int j = 0; j < CB; ++j)
            y[((long long)n * Co + co0 + j) * Ho * Wo + (long long)oy * Wo + ox] = f2bf(fmaxf(acc[j], 0.f));
    }
}

// The same convolution on the matrix cores (v_mfma_f32_16x16x32_bf16, float32 accumulation), for
// Co a multiple of 16, at least 8 input channels and outputs at least 16 wide.  As a GEMM, y[co][p] = sum_k W[co][k] X[k][p]
// over k = (tap, ci), the block's TY x TX output positions p of one image (64 PQ, PQ = 1, 2 or 4)
// and its 16 CT output channels.  Input channels are staged 16 at a time:
//   patch: the tile's (2 TY + 1) x (2 TX + 2) input pixels with the 16 channels innermost (48-byte
//          pixels: 16-byte fragment reads, at most 2-way bank conflicts), zeros outside the image
//          and past Ci;
//   wl:    per output channel the 20 groups g = 2 tap + c8 of 8 channels (tap 9 zero), so that
//          one 32-deep product takes the groups 4 m .. 4 m + 3: taps 2 m, 2 m + 1 of 16 channels.
// Lane l of a wave holds, for one 16 x 16 x 32 product, k group 4 m + (l >> 4) of output channel
// row (l & 15) (A) and of position column (l & 15) (B); the result rows 4 (l >> 4) + i, column
// (l & 15).  Wave w owns the PQ position tiles 16 (PQ w .. PQ w + PQ - 1) and the block's CT
// channel tiles.  Bias is added to the finished sum, then ReLU and the bf16 rounding.  Grid
// (ceil(Wo / TX), ceil(Ho / TY), N Co / (16 CT)): blocks never span images and every output's
// sum is the same for any tile shape (k order, then one product per k block), so an image's
// outputs do not depend on the batch or on the launch shape.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
constexpr int kCmCC = 16;                     // input channels per staged chunk
constexpr int kCmCS = 24;                     // bf16 slots per staged pixel (16 channels + 8 pad)
constexpr int kCmGroups = 20;                 // weight groups per output channel and chunk
constexpr int kCmWRow = kCmGroups * 8 + 8;    // bf16 per staged weight row (336 bytes)

template <int TX, int TY, int CT>
__global__ __launch_bounds__(kCnnThreads) void conv_fwd_mfma_kernel(const uint16_t* __restrict__ x,
                                                                    const uint16_t* __restrict__ w,
                                                                    const float* __restrict__ bias,
                                                                    uint16_t* __restrict__ y, int N, int Ci, int H,
                                                                    int W, int Co) {
    constexpr int PQ = TX * TY / 64;                             // position tiles per wave
    static_assert(PQ == 1 || PQ == 2 || PQ == 4, "64, 128 or 256 positions per block");
    constexpr int PH = 2 * TY + 1, PW = 2 * TX + 2;              // pixel column px = input ix - (2 oxb - 2)
    constexpr int kPairs = PH * (PW / 2);
    __shared__ __align__(16) uint16_t patch[PH * PW * kCmCS];
    __shared__ __align__(16) uint16_t wl[16 * CT * kCmWRow];
    const int Ho = H >> 1, Wo = W >> 1;
    const int splits = Co / (16 * CT);
    const int n = blockIdx.z / splits, cob = (blockIdx.z - (blockIdx.z / splits) * splits) * 16 * CT;
    const int oyb = blockIdx.y * TY, oxb = blockIdx.x * TX;
    const int iyb = 2 * oyb - 1, ixb = 2 * oxb - 2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int grp = lane >> 4;
    int boff[PQ];                                                // pixel (2 ty, 2 tx + 1): tap (0, 0)
#pragma unroll
    for (int q = 0; q < PQ; ++q) {
        const int pos = 16 * (PQ * wave + q) + (lane & 15);
        const int ty = pos / TX, tx = pos - (pos / TX) * TX;
        boff[q] = ((2 * ty) * PW + 2 * tx + 1) * kCmCS;
    }
    f32x4_t acc[CT][PQ];
#pragma unroll
    for (int t = 0; t < CT; ++t)
#pragma unroll
        for (int q = 0; q < PQ; ++q) acc[t][q] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    const uint16_t* xn = x + (long long)n * Ci * H * W;
    for (int c0 = 0; c0 < Ci; c0 += kCmCC) {
        if (c0) __syncthreads();                                 // the last chunk's reads are done
        // patch: item = (8-channel group c8, row py, pixel pair k): 8 channels of pixels 2k, 2k+1
        for (int e = tid; e < 2 * kPairs; e += kCnnThreads) {
            const int c8 = e / kPairs, r = e - (e / kPairs) * kPairs;
            const int py = r / (PW / 2), k = r - (r / (PW / 2)) * (PW / 2);
            const int iy = iyb + py, ix = ixb + 2 * k;
            const bool inside = (unsigned)iy < (unsigned)H && (unsigned)ix < (unsigned)W;   // W even: whole pairs
            uint32_t v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int ci = c0 + 8 * c8 + j;
                v[j] = (inside && ci < Ci) ? *reinterpret_cast<const uint32_t*>(xn + ((long long)ci * H + iy) * W + ix) : 0u;
            }
            uint4 lo, hi;                                        // pixel 2k: low halves, 2k + 1: high halves
            lo.x = __builtin_amdgcn_perm(v[1], v[0], 0x05040100u);
            lo.y = __builtin_amdgcn_perm(v[3], v[2], 0x05040100u);
            lo.z = __builtin_amdgcn_perm(v[5], v[4], 0x05040100u);
            lo.w = __builtin_amdgcn_perm(v[7], v[6], 0x05040100u);
            hi.x = __builtin_amdgcn_perm(v[1], v[0], 0x07060302u);
            hi.y = __builtin_amdgcn_perm(v[3], v[2], 0x07060302u);
            hi.z = __builtin_amdgcn_perm(v[5], v[4], 0x07060302u);
            hi.w = __builtin_amdgcn_perm(v[7], v[6], 0x07060302u);
            *reinterpret_cast<uint4*>(&patch[(py * PW + 2 * k) * kCmCS + 8 * c8]) = lo;
            *reinterpret_cast<uint4*>(&patch[(py * PW + 2 * k + 1) * kCmCS + 8 * c8]) = hi;
        }
        // weights: element (co, g, j) = w[co][c0 + 8 (g & 1) + j][g >> 1], zero for tap 9 / past Ci
        for (int e = tid; e < 16 * CT * kCmGroups * 8; e += kCnnThreads) {
            const int co = e / (kCmGroups * 8), r = e - (e / (kCmGroups * 8)) * (kCmGroups * 8);
            const int g = r >> 3, tap = g >> 1, ci = c0 + 8 * (g & 1) + (r & 7);
            wl[co * kCmWRow + r] = (tap < 9 && ci < Ci) ? w[((long long)(cob + co) * Ci + ci) * 9 + tap] : (uint16_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < kCmGroups / 4; ++m) {
            const int gl = 4 * m + grp;
            const int tap = gl >> 1;
            const int tp = tap < 9 ? tap : 0;                    // tap 9 has zero weights: read any pixel
            const int ky = tp / 3, kx = tp - (tp / 3) * 3;
            const int doff = (ky * PW + kx) * kCmCS + 8 * (gl & 1);
            bf16x8_t b[PQ];
#pragma unroll
            for (int q = 0; q < PQ; ++q) b[q] = *reinterpret_cast<const bf16x8_t*>(&patch[boff[q] + doff]);
#pragma unroll
            for (int t = 0; t < CT; ++t) {
                const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(&wl[(16 * t + (lane & 15)) * kCmWRow + 8 * gl]);
#pragma unroll
                for (int q = 0; q < PQ; ++q) acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[q], acc[t][q], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < PQ; ++q) {
        const int pos = 16 * (PQ * wave + q) + (lane & 15);
        const int ty = pos / TX, tx = pos - (pos / TX) * TX;
        const int oy = oyb + ty, ox = oxb + tx;
        if (oy < Ho && ox < Wo) {
#pragma unroll
            for (int t = 0; t < CT; ++t)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int co = cob + 16 * t + 4 * grp + i;
                    y[((long long)n * Co + co) * Ho * Wo + (long long)oy * Wo + ox] = f2bf(fmaxf(acc[t][q][i] + bias[co], 0.f));
                }
        }
    }
}

// The same convolution for layers with few output positions (small batches, deep layers): one
// wave per (P horizontally adjacent output positions, 16 output channels); lanes split the
// reduction over kk = ci 9 + tap (the weights of one output channel are contiguous in kk) and use
// each loaded weight for the P positions.  The 64 lanes' partial sums of the 16 P outputs are then
// combined by full xor butterflies (P = 1) or, for P > 1, by a reduce-scatter butterfly: at distance m = 32, 16, ..., 1 a lane keeps half of its
// values and adds the partner lane's (lane ^ m) copies of them, own value first.  Every output is
// thereby summed over the same pairs in the same order as a full xor butterfly of each value, with
// about 16 P exchanges instead of 96 P; the lane that ends with output idx = q 16 + j writes it.
// Bias last.  Wo % P == 0; grid ceil(N Ho (Wo / P) (Co / 16) / 4) blocks of 4 waves.
template <int P>
__global__ __launch_bounds__(kCnnThreads) void conv_fwd_wave_kernel(const uint16_t* __restrict__ x,
                                                                    const uint16_t* __restrict__ w,
                                                                    const float* __restrict__ bias,
                                                                    uint16_t* __restrict__ y, int N, int Ci, int H,
                                                                    int W, int Co) {
    constexpr int V = 16 * P;                 // outputs per wave: [q][j]
    constexpr int R = 64 / V;                 // lanes that end up holding each output
    const int Ho = H >> 1, Wo = W >> 1, Wg = Wo / P;
    const int groups = Co / 16;
    const long long task = (long long)blockIdx.x * (kCnnThreads / 64) + (threadIdx.x >> 6);
    if (task >= (long long)N * Ho * Wg * groups) return;                 // uniform per wave
    const int lane = threadIdx.x & 63;
    const int g = (int)(task % groups);
    const long long p = task / groups;
    const int ox0 = (int)(p % Wg) * P;
    const long long t = p / Wg;
    const int oy = (int)(t % Ho);
    const int n = (int)(t / Ho);
    const int K = Ci * 9;
    const int co0 = g * 16;
    float a[V];
#pragma unroll
    for (int i = 0; i < V; ++i) a[i] = 0.f;
    const uint16_t* xn = x + (long long)n * Ci * H * W;
    for (int kk = lane; kk < K; kk += 64) {
        const int ci = kk / 9, tap = kk - (kk / 9) * 9;
        const int iy = 2 * oy - 1 + tap / 3, dx = tap % 3 - 1;
        float wv[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) wv[j] = bf2f(w[(long long)(co0 + j) * K + kk]);
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int ix = 2 * (ox0 + q) + dx;
            const float v = (iy >= 0 && ix >= 0) ? bf2f(xn[((long long)ci * H + iy) * W + ix]) : 0.f;
#pragma unroll
            for (int j = 0; j < 16; ++j) a[q * 16 + j] = __builtin_fmaf(wv[j], v, a[q * 16 + j]);
        }
    }
    if constexpr (P == 1) {
        // one position: 16 independent full butterflies pipeline better than the dependent
        // reduce-scatter steps (same pairs, same order, same sums)
#pragma unroll
        for (int j = 0; j < 16; ++j)
            for (int off = 32; off > 0; off >>= 1) a[j] += __shfl_xor(a[j], off);
        if (lane < 16) {
            float v = a[0];
#pragma unroll
            for (int j = 1; j < 16; ++j) v = (lane == j) ? a[j] : v;
            v += bias[co0 + lane];
            y[((long long)n * Co + co0 + lane) * Ho * Wo + (long long)oy * Wo + ox0] = f2bf(fmaxf(v, 0.f));
        }
        return;
    }
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int m = 32 >> s;
        const int cnt = (V >> s) > 1 ? (V >> s) : 1;                     // values held before this step
        if (cnt > 1) {
            const bool upper = (lane & m) != 0;                          // keeps the upper half
#pragma unroll
            for (int i = 0; i < cnt / 2; ++i) {
                const float keep = upper ? a[i + cnt / 2] : a[i];
                const float send = upper ? a[i] : a[i + cnt / 2];
                a[i] = keep + __shfl_xor(send, m);
            }
        } else {
            a[0] += __shfl_xor(a[0], m);
        }
    }
    if ((lane & (R - 1)) == 0) {
        const int idx = lane / R, q = idx / 16, j = idx - (idx / 16) * 16;
        const float v = a[0] + bias[co0 + j];
        y[((long long)n * Co + co0 + j) * Ho * Wo + (long long)oy * Wo + ox0 + q] = f2bf(fmaxf(v, 0.f));
    }
}

// dx[n, ci, iy, ix] = sum_co sum_taps w[co, ci, ky, kx] dpre[n, co, oy, ox] over iy = 2oy-1+ky,
// ix = 2ox-1+kx, times (xmask[n, ci, iy, ix] > 0) when xmask is given.  One thread owns the 2 x 2
// input positions (2a + {0,1}, 2b + {0,1}) for 16 input channels, so every thread uses all 9 taps
// (uniform weight reads).  Grid (ceil(N Ho Wo / 256), Ci / 16).
__global__ __launch_bounds__(kCnnThreads) void conv_bwd_data_kernel(const float* __restrict__ dpre,
                                                                    const uint16_t* __restrict__ w,
                                                                    const uint16_t* __restrict__ xmask,
                                                                    float* __restrict__ dx, int N, int Ci, int H,
                                                                    int W, int Co) {
    const int Ho = H >> 1, Wo = W >> 1;
    const long long P = (long long)N * Ho * Wo;
    const long long p = (long long)blockIdx.x * kCnnThreads + threadIdx.x;
    if (p >= P) return;
    const int ci0 = blockIdx.y * kConvCiB;
    const int b = (int)(p % Wo);
    const long long t = p / Wo;
    const int a = (int)(t % Ho);
    const int n = (int)(t / Ho);
    const bool ra = a + 1 < Ho, rb = b + 1 < Wo;
    float acc[4][kConvCiB];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < kConvCiB; ++j) acc[q][j] = 0.f;
    for (int co = 0; co < Co; ++co) {
        const float* g = dpre + ((long long)n * Co + co) * Ho * Wo;
        const float g00 = g[(long long)a * Wo + b];
        const float g01 = rb ? g[(long long)a * Wo + b + 1] : 0.f;
        const float g10 = ra ? g[(long long)(a + 1) * Wo + b] : 0.f;
        const float g11 = (ra && rb) ? g[(long long)(a + 1) * Wo + b + 1] : 0.f;
        const uint16_t* wp = w + ((long long)co * Ci + ci0) * 9;
#pragma unroll
        for (int j = 0; j < kConvCiB; ++j) {
            float wk[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) wk[k] = bf2f(wp[j * 9 + k]);
            acc[0][j] = __builtin_fmaf(wk[4], g00, acc[0][j]);                         // (2a,   2b)
            acc[1][j] = __builtin_fmaf(wk[5], g00, __builtin_fmaf(wk[3], g01, acc[1][j]));   // (2a,   2b+1)
            acc[2][j] = __builtin_fmaf(wk[7], g00, __builtin_fmaf(wk[1], g10, acc[2][j]));   // (2a+1, 2b)
            acc[3][j] = __builtin_fmaf(
                wk[8], g00, __builtin_fmaf(wk[6], g01, __builtin_fmaf(wk[2], g10, __builtin_fmaf(wk[0], g11, acc[3][j]))));
        }
    }
#pragma unroll
    for (int j = 0; j < kConvCiB; ++j) {
        const long long base = ((long long)n * Ci + ci0 + j) * H * W;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long long o = base + (long long)(2 * a + (q >> 1)) * W + 2 * b + (q & 1);
            float v = acc[q][j];
            if (xmask && !(bf2f(xmask[o]) > 0.f)) v = 0.f;
            dx[o] = v;
        }
    }
}

// Per (co, ci, split): partial sums over a slice of the N Ho Wo output positions of
//   dw[co, ci, ky, kx] = sum dpre[n, co, oy, ox] x[n, ci, 2oy-1+ky, 2ox-1+kx],  db[co] = sum dpre
// into part[((s Co + co) Ci + ci) 10 + k] (k = 9: the bias sum).  Grid (Co Ci, S).
__global__ __launch_bounds__(kCnnThreads) void conv_bwd_weight_kernel(const float* __restrict__ dpre,
                                                                      const uint16_t* __restrict__ x,
                                                                      float* __restrict__ part, int N, int Ci,
                                                                      int H, int W, int Co) {
    const int co = blockIdx.x / Ci, ci = blockIdx.x - (blockIdx.x / Ci) * Ci;
    const int s = blockIdx.y, S = gridDim.y;
    const int Ho = H >> 1, Wo = W >> 1;
    const long long P = (long long)N * Ho * Wo;
    float acc[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) acc[k] = 0.f;
    for (long long p = (long long)s * kCnnThreads + threadIdx.x; p < P; p += (long long)S * kCnnThreads) {
        const int ox = (int)(p % Wo);
        const long long t = p / Wo;
        const int oy = (int)(t % Ho);
        const int n = (int)(t / Ho);
        const float g = dpre[((long long)n * Co + co) * Ho * Wo + (long long)oy * Wo + ox];
        const uint16_t* xp = x + ((long long)n * Ci + ci) * H * W;
        const int iy0 = 2 * oy - 1, ix0 = 2 * ox - 1;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky)
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const int iy = iy0 + ky, ix = ix0 + kx;
                const float v = (iy >= 0 && ix >= 0) ? bf2f(xp[(long long)iy * W + ix]) : 0.f;
                acc[3 * ky + kx] = __builtin_fmaf(g, v, acc[3 * ky + kx]);
            }
        acc[9] += g;
    }
    __shared__ float red[kCnnThreads / 64][10];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        float v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 10) {
        float v = red[0][threadIdx.x];
        for (int wv = 1; wv < kCnnThreads / 64; ++wv) v += red[wv][threadIdx.x];
        part[(((long long)s * Co + co) * Ci + ci) * 10 + threadIdx.x] = v;
    }
}

// dw[(co Ci + ci) 9 + k] = sum_s part[.., k]; db[co] = sum_s part[(s, co, ci = 0), 9].
__global__ __launch_bounds__(kCnnThreads) void conv_bwd_weight_reduce_kernel(const float* __restrict__ part,
                                                                             float* __restrict__ dw,
                                                                             float* __restrict__ db, int Ci, int Co,
                                                                             int S) {
    const long long i = (long long)blockIdx.x * kCnnThreads + threadIdx.x;
    const long long nw = (long long)Co * Ci * 9;
    if (i < nw) {
        const long long cc = i / 9;
        const int k = (int)(i - cc * 9);
        float v = 0.f;
        for (int s = 0; s < S; ++s) v += part[((long long)s * Co * Ci + cc) * 10 + k];
        dw[i] = v;
    } else if (i < nw + Co) {
        const int co = (int)(i - nw);
        float v = 0.f;
        for (int s = 0; s < S; ++s) v += part[((long long)s * Co * Ci + (long long)co * Ci) * 10 + 9];
        db[co] = v;
    }
}

// ---------------------------------------------------------------- backward on the matrix cores
// Both backward convolutions as GEMMs on v_mfma_f32_16x16x32_bf16 with float32 sums.  The float32
// output gradient enters as two bf16 parts, hi = bf16(d) and lo = bf16(d - hi) (together within
// 2^-16 |d| of d), and every product is made for both; x and w are bf16 already, so each product
// is exact and the sums are the only float32 rounding.

__device__ __forceinline__ void split_bf16(float d, uint16_t& hi, uint16_t& lo) {
    hi = f2bf(d);
    lo = f2bf(d - bf2f(hi));                 // d - hi is exact in float32
}

// Weight gradient, step 1: xcol[c][p] (bf16; rows c < ncols_pad, columns p < ppad = P rounded up
// to 32) with row c = 9 ci + 3 ky + kx holding x[n, ci, 2 oy - 1 + ky, 2 ox - 1 + kx] (0 outside
// the image) at p = (n Ho + oy) Wo + ox, row 9 Ci holding 1 (the bias column), and 0 in the other
// rows and past P.  One thread per 8 consecutive p of one row (one 16-byte store).
__global__ __launch_bounds__(kCnnThreads) void conv_im2col_kernel(const uint16_t* __restrict__ x,
                                                                  uint16_t* __restrict__ xcol, int Ci, int H,
                                                                  int W, int P, int ppad, int ncols_pad) {
    const long long e = (long long)blockIdx.x * kCnnThreads + threadIdx.x;
    const int per_row = ppad / 8;
    if (e >= (long long)ncols_pad * per_row) return;
    const int c = (int)(e / per_row);
    const int p0 = 8 * (int)(e - (long long)c * per_row);
    const int Ho = H >> 1, Wo = W >> 1;
    uint16_t v[8];
    if (c < 9 * Ci) {
        const int ci = c / 9, tap = c - 9 * ci, ky = tap / 3, kx = tap - 3 * ky;
        int ox = p0 % Wo, t = p0 / Wo;
        int oy = t % Ho, n = t / Ho;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int iy = 2 * oy - 1 + ky, ix = 2 * ox - 1 + kx;
            const bool in = p0 + j < P && iy >= 0 && iy < H && ix >= 0 && ix < W;
            v[j] = in ? x[(((long long)n * Ci + ci) * H + iy) * W + ix] : (uint16_t)0;
            if (++ox == Wo) {
                ox = 0;
                if (++oy == Ho) {
                    oy = 0;
                    ++n;
                }
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = (c == 9 * Ci && p0 + j < P) ? (uint16_t)0x3F80u : (uint16_t)0;   // bf16 1.0
    }
    uint4 o;
    o.x = v[0] | ((uint32_t)v[1] << 16);
    o.y = v[2] | ((uint32_t)v[3] << 16);
    o.z = v[4] | ((uint32_t)v[5] << 16);
    o.w = v[6] | ((uint32_t)v[7] << 16);
    *reinterpret_cast<uint4*>(xcol + (long long)c * ppad + p0) = o;
}

// Weight gradient, step 2: part[s][co][c] = sum over slice s of the positions of
// dpre[n, co, p] xcol[c][p].  One wave per (16 output channels, 4 column tiles of 16, slice);
// per 32 positions each lane holds 8 consecutive positions of its channel (A: both bf16 parts)
// and of its column (B).  HoWo % 8 == 0: a lane's 8 positions lie in one image (two 16-byte
// loads); otherwise they are gathered one by one.  Slices are whole multiples of 32 positions.
__global__ __launch_bounds__(kCnnThreads) void conv_bwd_weight_mfma_kernel(const float* __restrict__ dpre,
                                                                           const uint16_t* __restrict__ xcol,
                                                                           float* __restrict__ part, int Co, int HoWo,
                                                                           int P, int ppad, int ncols_pad, int S,
                                                                           int slice) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ntiles = ncols_pad / 16, ngroups = (ntiles + 3) / 4;
    const long long task = (long long)blockIdx.x * (kCnnThreads / 64) + wave;
    if (task >= (long long)(Co / 16) * ngroups * S) return;                // whole wave
    const int s = (int)(task % S);
    const int rest = (int)(task / S);
    const int ng = rest % ngroups, mt = rest / ngroups;
    const int g = lane >> 4, l16 = lane & 15;
    const int co = 16 * mt + l16;
    const bool vec = (HoWo % 8) == 0;
    f32x4_t acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    const int k_begin = s * slice;
    const int k_end = min(ppad, k_begin + slice);
    for (int k0 = k_begin; k0 < k_end; k0 += 32) {
        const int p = k0 + 8 * g;
        float d[8];
        if (vec) {
            if (p < P) {
                const int n = p / HoWo, q = p - n * HoWo;
                const float4* src = reinterpret_cast<const float4*>(dpre + ((long long)n * Co + co) * HoWo + q);
                const float4 a0 = src[0], a1 = src[1];
                d[0] = a0.x; d[1] = a0.y; d[2] = a0.z; d[3] = a0.w;
                d[4] = a1.x; d[5] = a1.y; d[6] = a1.z; d[7] = a1.w;
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) d[j] = 0.f;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int pj = p + j;
                const int n = pj / HoWo, q = pj - n * HoWo;
                d[j] = pj < P ? dpre[((long long)n * Co + co) * HoWo + q] : 0.f;
            }
        }
        bf16x8_t ah, al;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint16_t h, l;
            split_bf16(d[j], h, l);
            ah[j] = __builtin_bit_cast(__bf16, h);
            al[j] = __builtin_bit_cast(__bf16, l);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int tile = 4 * ng + t;
            if (tile < ntiles) {                                           // uniform
                const bf16x8_t b = *reinterpret_cast<const bf16x8_t*>(xcol + (long long)(16 * tile + l16) * ppad + p);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, b, acc[t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int tile = 4 * ng + t;
        if (tile < ntiles) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                part[((long long)s * Co + 16 * mt + 4 * g + i) * ncols_pad + 16 * tile + l16] = acc[t][i];
        }
    }
}

// Weight gradient, step 3: dw[co][c] = sum_s part[s][co][c] for c < 9 Ci, db[co] = the bias column.
// Block = 16 columns x 16 slice groups: thread (r, j) sums slices r, r + 16, ... of column
// 16 blockIdx.x + j (reads along c), then the 16 group sums are added in group order through LDS
// (a fixed order: the same bits on every run).  Grid (ceil((9 Ci + 1) / 16), Co).
__global__ __launch_bounds__(kCnnThreads) void conv_bwd_weight_mfma_reduce_kernel(const float* __restrict__ part,
                                                                                  float* __restrict__ dw,
                                                                                  float* __restrict__ db, int Ci,
                                                                                  int Co, int ncols_pad, int S) {
    const int j = threadIdx.x & 15, r = threadIdx.x >> 4;
    const int co = blockIdx.y, c = 16 * blockIdx.x + j;
    const int nc = 9 * Ci + 1;
    float v = 0.f;
    if (c < nc)
        for (int s = r; s < S; s += 16) v += part[((long long)s * Co + co) * ncols_pad + c];
    __shared__ float red[16][17];
    red[r][j] = v;
    __syncthreads();
    if (r == 0 && c < nc) {
        float t = 0.f;
#pragma unroll
        for (int q = 0; q < 16; ++q) t += red[q][j];
        if (c < 9 * Ci)
            dw[(long long)co * 9 * Ci + c] = t;
        else
            db[co] = t;
    }
}

// Data gradient, step 1: dth / dtl [p][co] = the two bf16 parts of dpre[n, co, p] (position-major,
// so that 8 consecutive channels are one 16-byte load).  One thread per (8 channels, position);
// consecutive threads take consecutive positions (coalesced reads of each channel).
__global__ __launch_bounds__(kCnnThreads) void conv_dpre_split_kernel(const float* __restrict__ dpre,
                                                                      uint16_t* __restrict__ dth,
                                                                      uint16_t* __restrict__ dtl, int Co, int HoWo,
                                                                      int P) {
    const long long e = (long long)blockIdx.x * kCnnThreads + threadIdx.x;
    if (e >= (long long)(Co / 8) * P) return;
    const int c8 = (int)(e / P), p = (int)(e - (long long)c8 * P);
    const int n = p / HoWo, q = p - n * HoWo;
    uint16_t h[8], l[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) split_bf16(dpre[((long long)n * Co + 8 * c8 + j) * HoWo + q], h[j], l[j]);
    uint4 oh, ol;
    oh.x = h[0] | ((uint32_t)h[1] << 16); oh.y = h[2] | ((uint32_t)h[3] << 16);
    oh.z = h[4] | ((uint32_t)h[5] << 16); oh.w = h[6] | ((uint32_t)h[7] << 16);
    ol.x = l[0] | ((uint32_t)l[1] << 16); ol.y = l[2] | ((uint32_t)l[3] << 16);
    ol.z = l[4] | ((uint32_t)l[5] << 16); ol.w = l[6] | ((uint32_t)l[7] << 16);
    *reinterpret_cast<uint4*>(dth + (long long)p * Co + 8 * c8) = oh;
    *reinterpret_cast<uint4*>(dtl + (long long)p * Co + 8 * c8) = ol;
}

// Data gradient, step 2: wt[tap][ci][co] = w[co][ci][tap] (bf16).
__global__ __launch_bounds__(kCnnThreads) void conv_w_transpose_kernel(const uint16_t* __restrict__ w,
                                                                       uint16_t* __restrict__ wt, int Ci, int Co) {
    const int e = blockIdx.x * kCnnThreads + threadIdx.x;
    if (e >= 9 * Ci * Co) return;
    const int co = e % Co, r = e / Co;
    const int ci = r % Ci, tap = r / Ci;
    wt[e] = w[((long long)co * Ci + ci) * 9 + tap];
}

// Data gradient, step 3.  Input pixels of one parity class (py, px) = (iy % 2, ix % 2) take the
// taps with ky = 1 (py = 0) or ky in {0, 2} (py = 1), the same for x, and the output position of
// tap ky is oy = a + (ky == 0) for iy = 2 a + py:
//   dx[n, ci, 2 a + py, 2 b + px] = sum_taps sum_co w[co, ci, ky, kx] dpre[n, co, a + dy, b + dx]
// (terms past the last output row or column are 0), times (xmask > 0) when xmask is given.  A GEMM
// per class: M = Ci, K = (taps, co), N = the N Ho Wo positions (a, b).  One wave per (class,
// 16 positions), all Ci / 16 row tiles; K in steps of 32 (a lane's 8 values lie in one tap).
// Grid (ceil(N Ho Wo / 64), 4).
__global__ __launch_bounds__(kCnnThreads) void conv_bwd_data_mfma_kernel(const uint16_t* __restrict__ dth,
                                                                         const uint16_t* __restrict__ dtl,
                                                                         const uint16_t* __restrict__ wt,
                                                                         const uint16_t* __restrict__ xmask,
                                                                         float* __restrict__ dx, int N, int Ci, int H,
                                                                         int W, int Co) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int Ho = H >> 1, Wo = W >> 1;
    const int P = N * Ho * Wo;
    const int tile = blockIdx.x * (kCnnThreads / 64) + wave;
    if (16 * tile >= P) return;                                            // whole wave
    const int py = blockIdx.y >> 1, px = blockIdx.y & 1;
    const int ny = py ? 2 : 1, nx = px ? 2 : 1;
    const int K = ny * nx * Co;
    const int g = lane >> 4, l16 = lane & 15;
    const int pos = 16 * tile + l16;
    const int posc = pos < P ? pos : P - 1;
    const int b = posc % Wo, t0 = posc / Wo;
    const int a = t0 % Ho, n = t0 / Ho;
    f32x4_t acc[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[m] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    const int mtiles = Ci / 16;
    for (int k0 = 0; k0 < K; k0 += 32) {
        const int k = k0 + 8 * g;
        bf16x8_t bh, bl;
        int ky = 0, kx = 0, co = 0;
        bool valid = false;
        if (k < K) {
            const int t = k / Co;
            co = k - t * Co;
            const int ty = t / nx, tx = t - ty * nx;
            ky = py ? 2 * ty : 1;                                          // ty 0 -> ky 0 (shift 1), 1 -> ky 2
            kx = px ? 2 * tx : 1;
            const int aa = a + (ky == 0), bb = b + (kx == 0);
            valid = pos < P && aa < Ho && bb < Wo;
            const long long row = ((long long)n * Ho + (valid ? aa : a)) * Wo + (valid ? bb : b);
            bh = *reinterpret_cast<const bf16x8_t*>(dth + row * Co + co);
            bl = *reinterpret_cast<const bf16x8_t*>(dtl + row * Co + co);
        }
        if (!valid) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                bh[j] = __builtin_bit_cast(__bf16, (uint16_t)0);
                bl[j] = __builtin_bit_cast(__bf16, (uint16_t)0);
            }
        }
        const int tap = 3 * ky + kx;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (m < mtiles) {                                              // uniform
                bf16x8_t av;
                if (k < K) {
                    av = *reinterpret_cast<const bf16x8_t*>(wt + ((long long)tap * Ci + 16 * m + l16) * Co + co);
                } else {
#pragma unroll
                    for (int j = 0; j < 8; ++j) av[j] = __builtin_bit_cast(__bf16, (uint16_t)0);
                }
                acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bh, acc[m], 0, 0, 0);
                acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bl, acc[m], 0, 0, 0);
            }
        }
    }
    if (pos >= P) return;
    const int iy = 2 * a + py, ix = 2 * b + px;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        if (m < mtiles) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int ci = 16 * m + 4 * g + i;
                const long long o = (((long long)n * Ci + ci) * H + iy) * W + ix;
                float v = acc[m][i];
                if (xmask && !(bf2f(xmask[o]) > 0.f)) v = 0.f;
                dx[o] = v;
            }
        }
    }
}

// y[n, o] = b[o] + sum_k w[o, k] x[n, k]: one wave per output (lanes over k, then a wave
// reduction); ReLU + bf16 store, or a float32 store.  Grid ceil(N O / 4) blocks of 4 waves.
__global__ __launch_bounds__(kCnnThreads) void fc_fwd_kernel(const uint16_t* __restrict__ x,
                                                             const uint16_t* __restrict__ w,
                                                             const float* __restrict__ bias,
                                                             uint16_t* __restrict__ ybf, float* __restrict__ yf,
                                                             int N, int K, int O) {
    const long long i = (long long)blockIdx.x * (kCnnThreads / 64) + (threadIdx.x >> 6);
    if (i >= (long long)N * O) return;                    // uniform per wave
    const int lane = threadIdx.x & 63;
    const int n = (int)(i / O), o = (int)(i - (long long)(i / O) * O);
    const uint16_t* xr = x + (long long)n * K;
    const uint16_t* wr = w + (long long)o * K;
    float acc = 0.f;
#pragma unroll 8
    for (int k = lane; k < K; k += 64) acc = __builtin_fmaf(bf2f(wr[k]), bf2f(xr[k]), acc);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) {
        acc += bias[o];
        if (ybf) ybf[i] = f2bf(fmaxf(acc, 0.f));
        else yf[i] = acc;
    }
}

// dx[n, k] = sum_o dpre[n, o] w[o, k] (o ascending), times (xmask[n, k] > 0) when given.
__global__ __launch_bounds__(kCnnThreads) void fc_bwd_data_kernel(const float* __restrict__ dpre,
                                                                  const uint16_t* __restrict__ w,
                                                                  const uint16_t* __restrict__ xmask,
                                                                  float* __restrict__ dx, int N, int K, int O) {
    const long long i = (long long)blockIdx.x * kCnnThreads + threadIdx.x;
    if (i >= (long long)N * K) return;
    const int n = (int)(i / K), k = (int)(i - (long long)(i / K) * K);
    float acc = 0.f;
    for (int o = 0; o < O; ++o) acc = __builtin_fmaf(dpre[(long long)n * O + o], bf2f(w[(long long)o * K + k]), acc);
    if (xmask && !(bf2f(xmask[i]) > 0.f)) acc = 0.f;
    dx[i] = acc;
}

// dw[o, k] = sum_n dpre[n, o] x[n, k];  db[o] = sum_n dpre[n, o]  (n ascending).
__global__ __launch_bounds__(kCnnThreads) void fc_bwd_weight_kernel(const float* __restrict__ dpre,
                                                                    const uint16_t* __restrict__ x,
                                                                    float* __restrict__ dw, float* __restrict__ db,
                                                                    int N, int K, int O) {
    const long long i = (long long)blockIdx.x * kCnnThreads + threadIdx.x;
    const long long nw = (long long)O * K;
    if (i < nw) {
        const int o = (int)(i / K), k = (int)(i - (long long)(i / K) * K);
        float acc = 0.f;
        for (int n = 0; n < N; ++n) acc = __builtin_fmaf(dpre[(long long)n * O + o], bf2f(x[(long long)n * K + k]), acc);
        dw[i] = acc;
    } else if (i < nw + O) {
        const int o = (int)(i - nw);
        float acc = 0.f;
        for (int n = 0; n < N; ++n) acc += dpre[(long long)n * O + o];
        db[o] = acc;
    }
}

// Heteroscedastic Gaussian negative log-likelihood, head (N, 2T) = (mu_0..mu_T-1, s_0..s_T-1),
// s = log variance, targets (N, T), per-target weights wt (T), T <= 4:
//   L = (1/N) sum_n sum_j wt_j 0.5 ((t - mu)^2 exp(-s) + s)
// dhead = dL/dhead; loss[0] = L, loss[1 + j] = (1/N) sum_n (t - mu_j)^2.  One block.
constexpr int kNllMaxT = 4;
__global__ __launch_bounds__(kCnnThreads) void gauss_nll_kernel(const float* __restrict__ head,
                                                                const float* __restrict__ target,
                                                                const float* __restrict__ wt, float* __restrict__ loss,
                                                                float* __restrict__ dhead, int N, int T) {
    __shared__ float red[kNllMaxT + 1][kCnnThreads];
    float acc[kNllMaxT + 1];
#pragma unroll
    for (int j = 0; j <= kNllMaxT; ++j) acc[j] = 0.f;
    const float invN = 1.0f / (float)N;
    for (int n = threadIdx.x; n < N; n += kCnnThreads) {
        for (int j = 0; j < T; ++j) {
            const float mu = head[(long long)n * 2 * T + j];
            const float s = head[(long long)n * 2 * T + T + j];
            const float e = expf(-s);
            const float r = target[(long long)n * T + j] - mu;
            acc[kNllMaxT] += wt[j] * 0.5f * (r * r * e + s);
            acc[j] += r * r;
            dhead[(long long)n * 2 * T + j] = -wt[j] * r * e * invN;
            dhead[(long long)n * 2 * T + T + j] = wt[j] * 0.5f * (1.f - r * r * e) * invN;
        }
    }
#pragma unroll
    for (int j = 0; j <= kNllMaxT; ++j) red[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (int st = kCnnThreads / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
#pragma unroll
            for (int j = 0; j <= kNllMaxT; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        loss[0] = red[kNllMaxT][0] * invN;
        for (int j = 0; j < T; ++j) loss[1 + j] = red[j][0] * invN;
    }
}

// Adam (bias-corrected, decoupled weight decay):  m = b1 m + (1-b1) g;  v = b2 v + (1-b2) g^2;
// p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps) + lr wd p;  pbf = bf16(p).
__global__ __launch_bounds__(kCnnThreads) void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                                           float* __restrict__ m, float* __restrict__ v,
                                                           uint16_t* __restrict__ pbf, long long count, float lr,
                                                           float b1, float b2, float eps, float wd, float c1,
                                                           float c2) {
    const long long i = (long long)blockIdx.x * kCnnThreads + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float pi = p[i];
    pi -= lr * wd * pi;
    pi -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    pbf[i] = f2bf(pi);
}

// Network input from magnitude volumes (C B, nX, rows), channel-major (image v = c B + b), row
// stride `xstride` floats per A-scan:
//   out[b, c, h, w] = bf16(clamp((20 log10(max_{j < zpool} mag[v, x0 + w, z0 + h zpool + j]) - offset)
//                                 * inv_scale, -8, 8))
// with magnitudes below 1e-30 taken as 1e-30.  Block = a 64 (h) x 64 (w) tile of one image:
// lanes run along h to read (rows are contiguous in mag; kVec4: zpool = 4 read as one aligned
// float4), the bf16 results are transposed through LDS, and lanes run along w to write (w is
// contiguous in out).  Grid (ceil(H / 64) ceil(W / 64), C B).
constexpr int kCropT = 64;

template <bool kVec4>
__global__ __launch_bounds__(kCnnThreads) void logmag_crop_kernel(const float* __restrict__ mag, long long xstride,
                                                                  int nX, int x0, int z0, int zpool, int H, int W,
                                                                  float offset, float inv_scale,
                                                                  uint16_t* __restrict__ out, int C, int B) {
    __shared__ uint16_t tile[kCropT][kCropT + 2];                        // [w][h]
    constexpr int kRows = kCropT / (kCnnThreads / 64);                   // 16 per wave
    const int v = blockIdx.y;
    const int c = v / B, b = v - (v / B) * B;
    const int th = (H + kCropT - 1) / kCropT;
    const int h0 = (blockIdx.x % th) * kCropT, w0 = (blockIdx.x / th) * kCropT;
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int h = h0 + lane;
#pragma unroll
    for (int i = 0; i < kRows; ++i) {
        const int wl = q + 4 * i, w = w0 + wl;
        if (h < H && w < W) {
            const float* src = mag + ((long long)v * nX + x0 + w) * xstride + z0 + (long long)h * zpool;
            float m;
            if constexpr (kVec4) {
                const float4 t = *reinterpret_cast<const float4*>(src);
                m = fmaxf(fmaxf(fmaxf(t.x, t.y), t.z), t.w);
            } else {
                m = src[0];
                for (int j = 1; j < zpool; ++j) m = fmaxf(m, src[j]);
            }
            const float db = 20.0f * log10f(fmaxf(m, 1e-30f));
            tile[wl][lane] = f2bf(fminf(fmaxf((db - offset) * inv_scale, -8.f), 8.f));
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kRows; ++i) {
        const int hl = q + 4 * i, w = w0 + lane;
        if (h0 + hl < H && w < W) out[(((long long)b * C + c) * H + h0 + hl) * W + w] = tile[lane][hl];
    }
}

// Exact median over x of mag[b, x, z] for every (b, z): float64 mean of the two middle order
// statistics (the single middle one for odd nX), as numpy.median of the float32 values cast to
// float64.  Block (b, 8 depth rows), the columns staged in LDS, order statistics by bisection on
// the order-preserving integer keys.  Grid (ceil(nz / 8), B), LDS 8 (nX + 1) floats, nX <= 1024.
constexpr int kMedZ = 8;
constexpr int kMedMaxX = 1024;

__device__ __forceinline__ int med_key(float f) {
    const int i = __builtin_bit_cast(int, f);
    return i >= 0 ? i : (i ^ 0x7FFFFFFF);
}

__device__ float wave_select(const float* col, int nX, int k) {
    // smallest key u with #(key <= u) >= k + 1
    long long lo = -2147483648LL, hi = 2147483647LL;
    const int lane = threadIdx.x & 63;
    while (lo < hi) {
        const long long mid = lo + ((hi - lo) >> 1);
        int cnt = 0;
        for (int x = lane; x < nX; x += 64) cnt += (med_key(col[x]) <= mid) ? 1 : 0;
        for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
        if (cnt >= k + 1) hi = mid;
        else lo = mid + 1;
    }
    const int key = (int)lo;
    return __builtin_bit_cast(float, key >= 0 ? key : (key ^ 0x7FFFFFFF));
}

// kOpc: the optical-path correction applied on the fly (octrecon/estimation/focus.py:110-129):
// row z of column x reads row clip(z + shift[f, x], 0, nz - 1) and exists only for
// vlo[f, x] <= z <= vhi[f, x] (f = frame_idx[b]); missing values are left out of the median
// (numpy.nanmedian), NaN when a row has none.  They are staged as NaN, whose keys sort above
// every number, so the order statistics of the valid values are unchanged.
template <bool kOpc>
__global__ __launch_bounds__(kCnnThreads) void depth_median_kernel(const float* __restrict__ mag, long long xstride,
                                                                   int nX, int nz, const int* __restrict__ fidx,
                                                                   const int* __restrict__ shift,
                                                                   const int* __restrict__ vlo,
                                                                   const int* __restrict__ vhi,
                                                                   double* __restrict__ out) {
    extern __shared__ float cols[];   // [kMedZ][nX + 1]
    const int b = blockIdx.y;
    const int z0 = blockIdx.x * kMedZ;
    const int ld = nX + 1;
    const long long orow = kOpc ? (long long)fidx[b] * nX : 0;
    for (int e = threadIdx.x; e < nX * kMedZ; e += kCnnThreads) {
        const int x = e / kMedZ, zz = e - x * kMedZ;
        const int z = z0 + zz;
        float v = 0.f;
        if (z < nz) {
            if constexpr (kOpc) {
                const long long o = orow + x;
                const bool ok = z >= vlo[o] && z <= vhi[o];
                const int src = min(max(z + shift[o], 0), nz - 1);
                v = ok ? mag[((long long)b * nX + x) * xstride + src] : __builtin_nanf("");
            } else {
                v = mag[((long long)b * nX + x) * xstride + z];
            }
        }
        cols[zz * ld + x] = v;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    for (int zz = wave; zz < kMedZ; zz += kCnnThreads / 64) {
        const int z = z0 + zz;
        if (z >= nz) break;
        const float* col = cols + zz * ld;
        int nv = nX;
        if constexpr (kOpc) {
            int c = 0;
            for (int x = lane; x < nX; x += 64) c += (col[x] == col[x]) ? 1 : 0;
            for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
            nv = c;
        }
        double r = __builtin_nan("");
        if (nv > 0) {
            const float v1 = wave_select(col, nX, (nv - 1) / 2);
            const float v2 = (nv & 1) ? v1 : wave_select(col, nX, nv / 2);
            r = 0.5 * ((double)v1 + (double)v2);
        }
        if (lane == 0) out[(long long)b * nz + z] = r;
    }
}

}  // namespace oct
