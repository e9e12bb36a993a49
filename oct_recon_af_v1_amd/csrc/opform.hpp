// Operator ("MFMA") form of the spectral stage.
//
// The map from one background-subtracted spectrum d (N samples) to its cropped complex A-scan
// is linear:  X[z] = sum_j A[z][j] d[j],  A = F_crop diag(win / N) S',
//   S' = S + (1/N) (1 - S 1) 1^T   (banded sinc operator with the mean removal / re-add folded
//   in as a rank-1 term), F the inverse-DFT kernel exp(+2 pi i z i / N).
// Reference path it reproduces: octrecon/core/spectral.py:123-153 (linearise, window, ifft, crop).
//
// A is built on the GPU in float64 (opbuild_*), scaled by 2^sa and split into fp16 hi + lo
// parts; rows are interleaved in blocks of 32 = (16 z real parts, the same 16 z imaginary
// parts).  opgemm_kernel then applies it to batches of A-scans on the matrix cores with
// v_mfma_f32_32x32x16_f16:  acc = Ahi dhi + Ahi dlo + Alo dhi  (fp32 accumulation), where
// d = raw - bg is split on load into dhi = fp16(d), dlo = fp16(d - dhi).  The epilogue forms
// |X| = 2^-sa sqrt(re^2 + im^2) in registers (each lane holds both parts of its bins).
#pragma once
#include "oct_common.hpp"

namespace oct {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ operator build (FP64)
// c[zl] = (1/N) sum_i (win_i / N) e^{2 pi i z i / N} u_i,  u = 1 - S 1.  One block per z.
__global__ __launch_bounds__(256) void opbuild_c_kernel(const double* __restrict__ win, const double* __restrict__ u,
                                                        int N, int z0, double2* __restrict__ cz) {
    const int zl = blockIdx.x;
    const long long z = z0 + zl;
    double re = 0.0, im = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const long long m = (z * i) % N;
        double s, c;
        sincospi(2.0 * (double)m / (double)N, &s, &c);
        const double wr = win[2 * i] / N, wi = win[2 * i + 1] / N;
        // (wr + i wi)(c + i s) u_i
        re += (wr * c - wi * s) * u[i];
        im += (wr * s + wi * c) * u[i];
    }
    __shared__ double sr[256], si[256];
    sr[threadIdx.x] = re;
    si[threadIdx.x] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            si[threadIdx.x] += si[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) cz[zl] = make_double2(sr[0] / N, si[0] / N);
}

// A[zl][j] = c[zl] + sum_t (win_i / N) e^{2 pi i z i / N} S[i][j],  i = tb_start[j] + t,
// written as fp16 hi/lo (scaled by 2^sa) in the interleaved row layout; rows of z >= z0+nzc
// (padding up to a multiple of 16) are zero.  Optionally also the float64 operator (nzc, N, 2).
__global__ __launch_bounds__(256) void opbuild_a_kernel(const int* __restrict__ tb_start, const int* __restrict__ tb_count,
                                                        const double* __restrict__ tb_w, int TT,
                                                        const double* __restrict__ win, const double2* __restrict__ cz,
                                                        int N, int z0, int nzc, int nzc_pad, double scale,
                                                        _Float16* __restrict__ op_hi, _Float16* __restrict__ op_lo,
                                                        double* __restrict__ op64) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int zl = blockIdx.y;
    if (j >= N || zl >= nzc_pad) return;
    double re = 0.0, im = 0.0;
    if (zl < nzc) {
        const long long z = z0 + zl;
        const double2 c = cz[zl];
        re = c.x;
        im = c.y;
        const int i0 = tb_start[j];
        const int cnt = tb_count[j];
        for (int t = 0; t < cnt; ++t) {
            const int i = i0 + t;
            const double w = tb_w[(long long)j * TT + t];
            const long long m = (z * i) % N;
            double s, cc;
            sincospi(2.0 * (double)m / (double)N, &s, &cc);
            const double wr = win[2 * i] / N, wi = win[2 * i + 1] / N;
            re += (wr * cc - wi * s) * w;
            im += (wr * s + wi * cc) * w;
        }
        if (op64) {
            op64[((long long)zl * N + j) * 2] = re;
            op64[((long long)zl * N + j) * 2 + 1] = im;
        }
    }
    const int blk = zl >> 4, zi = zl & 15;
    const long long rr = (long long)(32 * blk + zi) * N + j;
    const long long ri = (long long)(32 * blk + 16 + zi) * N + j;
    const double sre = re * scale, sim = im * scale;
    const _Float16 hr = (_Float16)sre, hi_ = (_Float16)sim;
    op_hi[rr] = hr;
    op_hi[ri] = hi_;
    op_lo[rr] = (_Float16)(sre - (double)hr);
    op_lo[ri] = (_Float16)(sim - (double)hi_);
}

// Step-tiled copy of the operator for opgemm_rf_kernel: for row block rb of kR rows and K step s
// of 32, the kR x 32 halves are contiguous (row-major inside the tile), so a workgroup stages one
// step's rows with 16-byte loads that cover 1 KB per wave instruction (8 whole cache lines) instead
// of 64 bytes from each of 32 rows.  dst[((rb nk + s) kR + rr) 32 + kk] = src[(rb kR + rr) N + 32 s + kk].
// One thread per 16-byte chunk of the row-major source.
__global__ __launch_bounds__(256) void opack_kernel(const _Float16* __restrict__ src_hi, const _Float16* __restrict__ src_lo,
                                                    int r_pad, int N, int kR, _Float16* __restrict__ dst_hi,
                                                    _Float16* __restrict__ dst_lo) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int per_row = N / 8;
    if (e >= (long long)r_pad * per_row) return;
    const int row = (int)(e / per_row);
    const int k = 8 * (int)(e - (long long)row * per_row);
    const int rb = row / kR, rr = row - rb * kR;
    const int s = k >> 5, kk = k & 31;
    const long long d = (((long long)rb * (N / 32) + s) * kR + rr) * 32 + kk;
    *reinterpret_cast<half8_t*>(dst_hi + d) = *reinterpret_cast<const half8_t*>(src_hi + 8 * e);
    *reinterpret_cast<half8_t*>(dst_lo + d) = *reinterpret_cast<const half8_t*>(src_lo + 8 * e);
}

// ------------------------------------------------------------------ GEMM + magnitude
constexpr int kOpCols = 256;       // A-scans per workgroup (4 waves x 64)
constexpr int kOpRows = 128;       // operator rows per workgroup (4 tiles of 32 = 64 z bins)
constexpr int kOpK = 32;           // k per stage
constexpr int kOpPad = 40;         // LDS row stride in halves (80 B)
constexpr int kOpBgN = 2048;       // spectrum length of the background rows staged in LDS

struct OpGemmArgs {
    const void* raw;
    const float* bg;
    const _Float16* op_hi;        // (R_pad, N)
    const _Float16* op_lo;
    const _Float16* opt_hi;       // step-tiled copies (opack_kernel) for opgemm_rf_kernel
    const _Float16* opt_lo;
    float* out;                   // magnitude (M, out_stride), bins [0, nzc)
    long long M;
    int lines, line0, nX;
    int N;
    int nzc;                      // kept bins
    int row_blocks;               // R_pad / kOpRows
    int col_blocks;               // ceil(M / kOpCols)
    long long out_stride;
    float inv_scale;              // 2^-sa
};

template <typename RawT>
__global__ __launch_bounds__(256, 2) void opgemm_kernel(OpGemmArgs a) {
    __shared__ __align__(16) _Float16 dhi[kOpCols * kOpPad];
    __shared__ __align__(16) _Float16 dlo[kOpCols * kOpPad];
    __shared__ __align__(16) _Float16 ahi[kOpRows * kOpPad];
    __shared__ __align__(16) _Float16 alo[kOpRows * kOpPad];

    // XCD-aware tile order: consecutive workgroup ids go to different XCDs, so the row blocks
    // of one A-scan block are given ids that land on the same XCD (shared L2 for its raw data).
    const int nblk = a.row_blocks * a.col_blocks;
    const int id = blockIdx.x;
    int rb, cb;
    {
        const int xcd = id % 8;
        const int slot = id / 8;
        const int per_xcd = (nblk + 7) / 8;
        const int lin = xcd * per_xcd + slot;     // position in the (col, row) order for this XCD
        if (lin >= nblk) return;
        cb = lin / a.row_blocks;
        rb = lin - cb * a.row_blocks;
    }
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const long long col0 = (long long)cb * kOpCols;
    const int row0 = rb * kOpRows;

    f32x16_t acc[4][2];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[r][c][v] = 0.f;

    // this thread's A-scan for the data tile
    const long long p = col0 + tid;
    const bool pvalid = p < a.M;
    const RawT* rrow = nullptr;
    const float* brow = nullptr;
    if (pvalid) {
        const long long b = p / a.nX;
        const long long x = p - b * a.nX;
        rrow = reinterpret_cast<const RawT*>(a.raw) + (b * a.lines + a.line0 + x) * (long long)a.N;
        brow = a.bg + b * a.N;
    }
    // operator rows loaded by this thread: row r = tid / 2, 16 halves at k + 16 * (tid & 1)
    const int orow = tid >> 1;
    const int ohalf = (tid & 1) * 16;
    const _Float16* ohp = a.op_hi + (long long)(row0 + orow) * a.N + ohalf;
    const _Float16* olp = a.op_lo + (long long)(row0 + orow) * a.N + ohalf;

    // Software pipeline: the raw samples and operator rows of step k0 + kOpK are loaded into
    // registers while the matrix cores work on step k0 from LDS; the background row is read
    // (L1-resident, one row per file) when the staged samples are converted.
    constexpr int kRawVec = 16 / sizeof(RawT);                  // samples per 16-byte load
    constexpr int kRawLd = kOpK / kRawVec;                      // 16-byte loads per thread per step
    uint4 rv[kRawLd];
    half8_t oh[2], ol[2];
    auto fetch = [&](int k0) {
        if (pvalid) {
#pragma unroll
            for (int g = 0; g < kRawLd; ++g) rv[g] = *reinterpret_cast<const uint4*>(rrow + k0 + kRawVec * g);
        }
        const half8_t* sh = reinterpret_cast<const half8_t*>(ohp + k0);
        const half8_t* sl = reinterpret_cast<const half8_t*>(olp + k0);
        oh[0] = sh[0];
        oh[1] = sh[1];
        ol[0] = sl[0];
        ol[1] = sl[1];
    };
    fetch(0);
    for (int k0 = 0; k0 < a.N; k0 += kOpK) {
        // ---- stage data: d = raw - bg -> fp16 hi / lo, and the operator rows
        {
            _Float16* hrow = dhi + tid * kOpPad;
            _Float16* lrow = dlo + tid * kOpPad;
            if (pvalid) {
#pragma unroll
                for (int g = 0; g < kOpK / 8; ++g) {
                    float f[8];
                    if constexpr (sizeof(RawT) == 2) {
                        unpack16<RawT>(rv[g], f);
                    } else {
                        float f0[4], f1[4];
                        unpack16<RawT>(rv[2 * g], f0);
                        unpack16<RawT>(rv[2 * g + 1], f1);
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            f[e] = f0[e];
                            f[4 + e] = f1[e];
                        }
                    }
                    const float4 b0 = *reinterpret_cast<const float4*>(brow + k0 + 8 * g);
                    const float4 b1 = *reinterpret_cast<const float4*>(brow + k0 + 8 * g + 4);
                    const float bb[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
                    half8_t h, l;
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const float d = f[e] - bb[e];
                        const _Float16 dh = (_Float16)d;
                        h[e] = dh;
                        l[e] = (_Float16)(d - (float)dh);
                    }
                    *reinterpret_cast<half8_t*>(hrow + 8 * g) = h;
                    *reinterpret_cast<half8_t*>(lrow + 8 * g) = l;
                }
            } else {
                half8_t z = {};
#pragma unroll
                for (int g = 0; g < kOpK / 8; ++g) {
                    *reinterpret_cast<half8_t*>(hrow + 8 * g) = z;
                    *reinterpret_cast<half8_t*>(lrow + 8 * g) = z;
                }
            }
            half8_t* th = reinterpret_cast<half8_t*>(ahi + orow * kOpPad + ohalf);
            half8_t* tl = reinterpret_cast<half8_t*>(alo + orow * kOpPad + ohalf);
            th[0] = oh[0];
            th[1] = oh[1];
            tl[0] = ol[0];
            tl[1] = ol[1];
        }
        __syncthreads();
        if (k0 + kOpK < a.N) fetch(k0 + kOpK);                  // in flight during the products
#pragma unroll
        for (int ks = 0; ks < kOpK; ks += 16) {
            const int kk = ks + 8 * (lane >> 5);
            half8_t bh[2], bl[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int col = wave * 64 + c * 32 + (lane & 31);
                bh[c] = *reinterpret_cast<const half8_t*>(dhi + col * kOpPad + kk);
                bl[c] = *reinterpret_cast<const half8_t*>(dlo + col * kOpPad + kk);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = r * 32 + (lane & 31);
                const half8_t ah = *reinterpret_cast<const half8_t*>(ahi + row * kOpPad + kk);
                const half8_t al = *reinterpret_cast<const half8_t*>(alo + row * kOpPad + kk);
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh[c], acc[r][c], 0, 0, 0);
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl[c], acc[r][c], 0, 0, 0);
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh[c], acc[r][c], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }

    // ---- epilogue: |X| for 8 bins per lane per tile
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const long long pc = col0 + wave * 64 + c * 32 + (lane & 31);
        if (pc >= a.M) continue;
        float* orow_out = a.out + pc * a.out_stride;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const int m = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);         // 0..15
                const int zl = (row0 / 32 + r) * 16 + m;
                if (zl < a.nzc) {
                    const float re = acc[r][c][v], im = acc[r][c][v + 8];
                    orow_out[zl] = __builtin_sqrtf(re * re + im * im) * a.inv_scale;
                }
            }
        }
    }
}

// opgemm_kernel with the data fragments kept in registers.  Each lane loads and converts the
// samples of its own B fragments directly (raw - bg -> fp16 hi / lo): for the wave's two 32-column
// tiles c, lanes 0-31 take samples k..k+7 and lanes 32-63 k+8..k+15 of A-scans 32 c + (lane & 31).
// These are exactly opgemm_kernel's fragments, and every accumulator receives the same products
// in the same order, so the results are the same bits.  Only the operator rows go through LDS,
// double-buffered with one workgroup barrier per K step of 32.  RT row tiles of 32 per workgroup.
//
// Software pipeline, per K step `it` in two halves of 16 (ks = 0, 1):
//   half 0: the products of (it, ks 0), and beside them the conversion of (it, ks 1) and the
//           load of step it + NS's second 16 samples;
//   half 1: the products of (it, ks 1), and beside them the conversion of (it + 1, ks 0) and the
//           load of step it + 1 + NS's first 16 samples.
// Each half is one basic block (clamped addresses instead of branches: lanes past the last
// A-scan load a valid row and their columns are never stored), so the conversion's VALU work
// issues in the shadow of the matrix instructions.  Raw samples are loaded NS steps ahead into a
// ring of NS register slots (NS = 2 for 16-bit samples, 1 for float32).
// kLdsBg (files of at least 256 A-scans, so a workgroup's A-scans lie in at most two files): the
// two background rows are read once into LDS instead of by every lane at every step.
// LDS 2 x RT x 5 KB (+ 16 KB).
template <typename RawT, bool kLdsBg, int RT>
__global__ __launch_bounds__(256, 2) void opgemm_rf_kernel(OpGemmArgs a) {
    constexpr int kR = 32 * RT;                                  // operator rows per workgroup
    __shared__ __align__(16) _Float16 ahi[2][kR * kOpPad];
    __shared__ __align__(16) _Float16 alo[2][kR * kOpPad];
    __shared__ __align__(16) float bgs[kLdsBg ? 2 : 1][kLdsBg ? kOpBgN : 4];   // background rows of the block's files

    const int nblk = a.row_blocks * a.col_blocks;
    int rb, cb;
    {
        const int id = blockIdx.x;
        const int xcd = id % 8;
        const int slot = id / 8;
        const int per_xcd = (nblk + 7) / 8;
        const int lin = xcd * per_xcd + slot;
        if (lin >= nblk) return;                                 // whole workgroup
        cb = lin / a.row_blocks;
        rb = lin - cb * a.row_blocks;
    }
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const long long col0 = (long long)cb * kOpCols;
    const int row0 = rb * kR;
    const int nk = a.N / kOpK;

    // This lane's data fragments: A-scans pc[c] = col0 + 64 wave + 32 c + (lane & 31), c = 0, 1,
    // samples k0 + 16 ks + ko .. + 7 (ko = 8 (lane >> 5)): the B-fragment layout of the matrix
    // cores, loaded directly (a wave's load touches 32 A-scan rows, two lanes per 32-byte segment).
    // A lane past the last A-scan reads the last A-scan's row (in the block's files).
    const int ko = 8 * (lane >> 5);
    const long long fb0 = col0 / a.nX;                           // first file of the block
    const RawT* rrow[2];
    const float* brow[2];
    int fsel[2];                                                 // kLdsBg: file - fb0 (0 or 1)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        long long pc = col0 + wave * 64 + c * 32 + (lane & 31);
        pc = pc < a.M ? pc : a.M - 1;
        const long long b = pc / a.nX;
        const long long x = pc - b * a.nX;
        rrow[c] = reinterpret_cast<const RawT*>(a.raw) + (b * a.lines + a.line0 + x) * (long long)a.N + ko;
        brow[c] = a.bg + b * a.N + ko;
        fsel[c] = (int)(b - fb0);
    }
    if constexpr (kLdsBg) {
        // nX >= kOpCols: the block's A-scans lie in files fb0 and fb0 + 1 (if present)
        const long long nfiles = (a.M + a.nX - 1) / a.nX;
        for (int e = tid; e < 2 * (kOpBgN / 4); e += 256) {
            const int f = e / (kOpBgN / 4), q = e - f * (kOpBgN / 4);
            float4 v = {0.f, 0.f, 0.f, 0.f};
            if (fb0 + f < nfiles) v = *reinterpret_cast<const float4*>(a.bg + (fb0 + f) * a.N + 4 * q);
            *reinterpret_cast<float4*>(&bgs[f][4 * q]) = v;
        }
    }
    // operator staging from the step-tiled copy: a step is 4 kR 16-byte chunks per part (hi, lo);
    // thread tid stages chunks tid and tid + 256 (chunk e: row e / 4, halves 8 (e % 4) ..)
    constexpr int kChunks = 4 * kR;
    constexpr int kPer = (kChunks + 255) / 256;                  // 1 or 2
    const _Float16* ohp = a.opt_hi + (long long)rb * nk * kR * kOpK + 8 * tid;
    const _Float16* olp = a.opt_lo + (long long)rb * nk * kR * kOpK + 8 * tid;
    constexpr int kRawVec = 16 / sizeof(RawT);
    constexpr int kFragLd = 8 / kRawVec;                         // 16-byte loads per fragment
    constexpr int NS = sizeof(RawT) == 2 ? 2 : 1;                // raw register slots (steps ahead)
    uint4 rv[NS][2][2][kFragLd];                                 // [slot][ks][c]
    half8_t oh[2], ol[2];
    half8_t dh[2][2], dl[2][2];                                  // fragments [ks][c]
    // -1 kept opaque: fma(hi, -1, d) is then formed by the mixed-precision FMA (f16 operand, f32
    // addend) instead of a conversion back to f32 and a subtraction; d - hi is exact, so the
    // rounded low part has the same bits either way
    float mone = -1.0f;
    asm volatile("" : "+v"(mone));

    auto fetch_raw = [&](auto slotI, auto jI, int step) {
        constexpr int S = decltype(slotI)::value, J = decltype(jI)::value;
        const int k0 = (step < nk ? step : nk - 1) * kOpK + 16 * J;      // past the end: reload the last step
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int q = 0; q < kFragLd; ++q)
                rv[S][J][c][q] = *reinterpret_cast<const uint4*>(rrow[c] + k0 + kRawVec * q);
    };
    auto fetch_op = [&](int step) {
        const long long o = (long long)(step < nk ? step : nk - 1) * kR * kOpK;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            if (tid + 256 * i < kChunks) {
                oh[i] = *reinterpret_cast<const half8_t*>(ohp + o + 2048 * i);
                ol[i] = *reinterpret_cast<const half8_t*>(olp + o + 2048 * i);
            }
        }
    };
    auto convert = [&](auto slotI, auto jI, int step) {
        constexpr int S = decltype(slotI)::value, J = decltype(jI)::value;
        const int k0 = (step < nk ? step : nk - 1) * kOpK + 16 * J;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            float f[8];
            if constexpr (sizeof(RawT) == 2) {
                unpack16<RawT>(rv[S][J][c][0], f);
            } else {
                float f0[4], f1[4];
                unpack16<RawT>(rv[S][J][c][0], f0);
                unpack16<RawT>(rv[S][J][c][1], f1);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    f[e] = f0[e];
                    f[4 + e] = f1[e];
                }
            }
            float4 b0, b1;
            if constexpr (kLdsBg) {
                b0 = *reinterpret_cast<const float4*>(&bgs[fsel[c]][k0 + ko]);
                b1 = *reinterpret_cast<const float4*>(&bgs[fsel[c]][k0 + ko + 4]);
            } else {
                b0 = *reinterpret_cast<const float4*>(brow[c] + k0);
                b1 = *reinterpret_cast<const float4*>(brow[c] + k0 + 4);
            }
            const float bb[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
            half8_t h, l;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float d = f[e] - bb[e];
                const _Float16 hv = (_Float16)d;
                h[e] = hv;
                l[e] = (_Float16)__builtin_fmaf((float)hv, mone, d);
            }
            dh[J][c] = h;
            dl[J][c] = l;
        }
    };
    auto stage_op = [&](int sb) {
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int e = tid + 256 * i;
            if (e < kChunks) {
                *reinterpret_cast<half8_t*>(&ahi[sb][(e >> 2) * kOpPad + 8 * (e & 3)]) = oh[i];
                *reinterpret_cast<half8_t*>(&alo[sb][(e >> 2) * kOpPad + 8 * (e & 3)]) = ol[i];
            }
        }
    };

    f32x16_t acc[RT][2];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[r][c][v] = 0.f;

    auto products = [&](int sb, auto jI) {
        constexpr int J = decltype(jI)::value;
        const int kk = 16 * J + ko;
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const int row = r * 32 + (lane & 31);
            const half8_t ah = *reinterpret_cast<const half8_t*>(&ahi[sb][row * kOpPad + kk]);
            const half8_t al = *reinterpret_cast<const half8_t*>(&alo[sb][row * kOpPad + kk]);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, dh[J][c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, dl[J][c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, dh[J][c], acc[r][c], 0, 0, 0);
            }
        }
    };

    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    static_for<0, NS>([&](auto s) {                              // steps 0 .. NS-1, both halves
        fetch_raw(s, I0{}, decltype(s)::value);
        fetch_raw(s, I1{}, decltype(s)::value);
    });
    fetch_op(0);
    stage_op(0);
    if constexpr (kLdsBg) __syncthreads();                       // background rows staged
    convert(I0{}, I0{}, 0);
    fetch_raw(I0{}, I0{}, NS);
    fetch_op(1);
    __syncthreads();
    for (int it0 = 0; it0 < nk; it0 += NS) {
        static_for<0, NS>([&](auto sI) {
            constexpr int S = decltype(sI)::value;
            constexpr int S1 = (S + 1) % NS;
            const int it = it0 + S;
            if (it < nk) {
                const int sb = it & 1;
                if (it + 1 < nk) stage_op(sb ^ 1);               // read by nobody in this step
                fetch_op(it + 2);
                products(sb, I0{});
                convert(sI, I1{}, it);
                fetch_raw(sI, I1{}, it + NS);
                products(sb, I1{});
                convert(std::integral_constant<int, S1>{}, I0{}, it + 1);
                fetch_raw(std::integral_constant<int, S1>{}, I0{}, it + 1 + NS);
                __syncthreads();
            }
        });
    }

#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const long long pc = col0 + wave * 64 + c * 32 + (lane & 31);
        if (pc >= a.M) continue;
        float* orow_out = a.out + pc * a.out_stride;
#pragma unroll
        for (int r = 0; r < RT; ++r) {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const int m = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);         // 0..15
                const int zl = (row0 / 32 + r) * 16 + m;
                if (zl < a.nzc) {
                    const float re = acc[r][c][v], im = acc[r][c][v + 8];
                    orow_out[zl] = __builtin_sqrtf(re * re + im * im) * a.inv_scale;
                }
            }
        }
    }
}

// opgemm_rf_kernel with coalesced sample loads (16-bit samples).  The direct fragment loads of
// opgemm_rf_kernel touch 32 A-scan rows per wave instruction (32 bytes of each), and the vector
// L1 cannot keep a 128-byte line until its four parts have been read, so every part is a separate
// L2 request.  Here a wave loads each line (64 samples = two K steps) with two back-to-back
// instructions: instruction (h, j) gives lane l the 16-byte chunk l % 4 of half h of the line of
// A-scan 16 j + l / 4 (16 half lines per instruction, the two halves of the same 16 lines issued
// one after the other).  Each K step 2 m + h the wave writes its four half-h registers into its
// own LDS region [64 A-scans][32 samples + 8 pad] and reads the B fragments back (16 bytes per
// lane: A-scan 32 c + l % 32, samples 16 ks + 8 (l / 32)); the padded 80-byte rows keep the
// fragment reads free of bank conflicts.  The fragments, the conversion and the products are
// opgemm_rf_kernel's, so the results are the same bits.  The region is the wave's alone, so its
// writes and reads need no barrier.
// The step body has no branches and its loads are issued in the order they are consumed
// (operator rows for step it + 2 at the top of step it, the next line pair in the middle of even
// steps), so each wait for loaded data waits for exactly that data and not for the prefetches
// issued after it.
// LDS 2 x RT x 5 KB (operator) + 20 KB (samples) + 16 KB (background rows).
template <typename RawT, bool kLdsBg, int RT>
__global__ __launch_bounds__(256, 2) void opgemm_lt_kernel(OpGemmArgs a) {
    static_assert(sizeof(RawT) == 2, "16-bit samples");
    constexpr int kR = 32 * RT;
    constexpr int kRawPad = 40;                                  // samples per staged A-scan row
    __shared__ __align__(16) _Float16 ahi[2][kR * kOpPad];
    __shared__ __align__(16) _Float16 alo[2][kR * kOpPad];
    __shared__ __align__(16) RawT sraw[4][64 * kRawPad];
    __shared__ __align__(16) float bgs[kLdsBg ? 2 : 1][kLdsBg ? kOpBgN : 4];

    const int nblk = a.row_blocks * a.col_blocks;
    int rb, cb;
    {
        const int id = blockIdx.x;
        const int xcd = id % 8;
        const int slot = id / 8;
        const int per_xcd = (nblk + 7) / 8;
        const int lin = xcd * per_xcd + slot;
        if (lin >= nblk) return;                                 // whole workgroup
        cb = lin / a.row_blocks;
        rb = lin - cb * a.row_blocks;
    }
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const long long col0 = (long long)cb * kOpCols;
    const int nk = a.N / kOpK;                                   // even (N = 2048)
    // A-scan indices and files in 32 bits (the host keeps M below 2^31)
    const unsigned nX = (unsigned)a.nX, last = (unsigned)(a.M - 1);
    const unsigned w0 = (unsigned)col0 + 64u * wave;

    // fragment side: background rows of A-scans w0 + 32 c + lane % 32
    const int ko = 8 * (lane >> 5);
    const unsigned fb0 = (unsigned)col0 / nX;
    const float* brow[2];
    int fsel[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        unsigned pc = w0 + 32u * c + (lane & 31);
        pc = pc < last ? pc : last;
        const unsigned b = pc / nX;
        brow[c] = a.bg + (long long)b * a.N + ko;
        fsel[c] = (int)(b - fb0);
    }
    // load side: raw rows of A-scans w0 + 16 j + lane / 4 (past the last A-scan: the last row)
    const RawT* rrow[4];
    static_for<0, 4>([&](auto jI) {
        constexpr int j = decltype(jI)::value;
        unsigned pc = w0 + 16u * j + (lane >> 2);
        pc = pc < last ? pc : last;
        const unsigned b = pc / nX;
        const unsigned x = pc - b * nX;
        rrow[j] = reinterpret_cast<const RawT*>(a.raw) + ((long long)b * a.lines + a.line0 + x) * (long long)a.N +
                  8 * (lane & 3);
    });
    if constexpr (kLdsBg) {
        const long long nfiles = (a.M + a.nX - 1) / a.nX;
        for (int e = tid; e < 2 * (kOpBgN / 4); e += 256) {
            const int f = e / (kOpBgN / 4), qq = e - f * (kOpBgN / 4);
            float4 v = {0.f, 0.f, 0.f, 0.f};
            if (fb0 + f < nfiles) v = *reinterpret_cast<const float4*>(a.bg + (fb0 + f) * a.N + 4 * qq);
            *reinterpret_cast<float4*>(&bgs[f][4 * qq]) = v;
        }
    }
    // operator staging: chunk e = tid + 256 i of the step tile (row e / 4, halves 8 (e % 4) ..);
    // a thread past the tile's chunks repeats a chunk of the first round (same data, same place)
    constexpr int kChunks = 4 * kR;
    constexpr int kPer = (kChunks + 255) / 256;
    int ech[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int e = tid + 256 * i;
        ech[i] = e < kChunks ? e : (e - 256 * i) % kChunks;
    }
    const _Float16* ohp = a.opt_hi + (long long)rb * nk * kR * kOpK;
    const _Float16* olp = a.opt_lo + (long long)rb * nk * kR * kOpK;
    RawT* const wraw = &sraw[wave][0];
    RawT* const wdst = wraw + (lane >> 2) * kRawPad + 8 * (lane & 3);
    u32x4_t rv[2][4];                                            // [half][j]; a native vector type, indexed
    half8_t oh[kPer], ol[kPer];                                  // with constants only, stays in registers
    half8_t dh[2][2], dl[2][2];
    float mone = -1.0f;
    asm volatile("" : "+v"(mone));

    auto load_pair = [&](int m) {                                // samples 64 m .. 64 m + 63
        const int k0 = 64 * (m < nk / 2 ? m : nk / 2 - 1);
        static_for<0, 4>([&](auto jI) {
            constexpr int j = decltype(jI)::value;
            rv[0][j] = *reinterpret_cast<const u32x4_t*>(rrow[j] + k0);
            rv[1][j] = *reinterpret_cast<const u32x4_t*>(rrow[j] + k0 + 32);
        });
    };
    auto write_half = [&](auto hI) {                             // step 2 m + h's samples into the region
        constexpr int h = decltype(hI)::value;
        static_for<0, 4>([&](auto jI) {
            constexpr int j = decltype(jI)::value;
            *reinterpret_cast<u32x4_t*>(wdst + 16 * j * kRawPad) = rv[h][j];
        });
    };
    // the region is written and read by one wave, whose LDS instructions execute in issue order:
    // the compiler must only keep the order (a hardware wait here would also wait for the loads
    // in flight)
    auto fence = [] { asm volatile("" ::: "memory"); };
    auto fetch_op = [&](int step) {
        const long long o = (long long)(step < nk ? step : nk - 1) * kR * kOpK;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            oh[i] = *reinterpret_cast<const half8_t*>(ohp + o + 8 * ech[i]);
            ol[i] = *reinterpret_cast<const half8_t*>(olp + o + 8 * ech[i]);
        }
    };
    auto stage_op = [&](int sb) {
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int e = ech[i];
            *reinterpret_cast<half8_t*>(&ahi[sb][(e >> 2) * kOpPad + 8 * (e & 3)]) = oh[i];
            *reinterpret_cast<half8_t*>(&alo[sb][(e >> 2) * kOpPad + 8 * (e & 3)]) = ol[i];
        }
    };
    auto convert = [&](auto jI, int step) {                      // fragments (step, ks J) from the region
        constexpr int J = decltype(jI)::value;
        const int k0 = (step < nk ? step : nk - 1) * kOpK + 16 * J;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const uint4 u = *reinterpret_cast<const uint4*>(wraw + (32 * c + (lane & 31)) * kRawPad + 16 * J + ko);
            float f[8];
            unpack16<RawT>(u, f);
            float4 b0, b1;
            if constexpr (kLdsBg) {
                b0 = *reinterpret_cast<const float4*>(&bgs[fsel[c]][k0 + ko]);
                b1 = *reinterpret_cast<const float4*>(&bgs[fsel[c]][k0 + ko + 4]);
            } else {
                b0 = *reinterpret_cast<const float4*>(brow[c] + k0);
                b1 = *reinterpret_cast<const float4*>(brow[c] + k0 + 4);
            }
            const float bb[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
            half8_t h, l;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float d = f[e] - bb[e];
                const _Float16 hv = (_Float16)d;
                h[e] = hv;
                l[e] = (_Float16)__builtin_fmaf((float)hv, mone, d);
            }
            dh[J][c] = h;
            dl[J][c] = l;
        }
    };

    f32x16_t acc[RT][2];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[r][c][v] = 0.f;

    auto products = [&](int sb, auto jI) {
        constexpr int J = decltype(jI)::value;
        const int kk = 16 * J + ko;
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const int row = r * 32 + (lane & 31);
            const half8_t ah = *reinterpret_cast<const half8_t*>(&ahi[sb][row * kOpPad + kk]);
            const half8_t al = *reinterpret_cast<const half8_t*>(&alo[sb][row * kOpPad + kk]);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, dh[J][c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, dl[J][c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, dh[J][c], acc[r][c], 0, 0, 0);
            }
        }
    };

    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    fetch_op(0);
    load_pair(0);
    stage_op(0);
    fetch_op(1);
    write_half(I0{});
    fence();
    if constexpr (kLdsBg) __syncthreads();                       // background rows staged
    convert(I0{}, 0);
    __syncthreads();
    // step it: the region holds step it's samples, dh / dl [0] its first fragments, operator
    // buffer it % 2 its rows, oh / ol step it + 1's rows; rv holds the line pair of step it + 1
    auto step = [&](int it, auto parity) {
        constexpr int P = decltype(parity)::value;               // it % 2
        stage_op(P ^ 1);                                         // step it + 1's rows (unused after the last step)
        fetch_op(it + 2);
        products(P, I0{});
        convert(I1{}, it);
        fence();
        write_half(std::integral_constant<int, P ^ 1>{});        // step it + 1's samples
        if constexpr (P == 0) load_pair(it / 2 + 1);             // the next line pair
        fence();
        products(P, I1{});
        convert(I0{}, it + 1);
        __syncthreads();
    };
    for (int it = 0; it < nk; it += 2) {
        step(it, I0{});
        step(it + 1, I1{});
    }

    const int row0 = rb * kR;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const long long pc = col0 + wave * 64 + c * 32 + (lane & 31);
        if (pc >= a.M) continue;
        float* orow_out = a.out + pc * a.out_stride;
#pragma unroll
        for (int r = 0; r < RT; ++r) {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const int m = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);         // 0..15
                const int zl = (row0 / 32 + r) * 16 + m;
                if (zl < a.nzc) {
                    const float re = acc[r][c][v], im = acc[r][c][v + 8];
                    orow_out[zl] = __builtin_sqrtf(re * re + im * im) * a.inv_scale;
                }
            }
        }
    }
}

}  // namespace oct
