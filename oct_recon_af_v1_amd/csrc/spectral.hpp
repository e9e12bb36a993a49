// Fused spectral kernels: raw spectra -> background subtraction -> mean removal -> banded sinc
// k-linearisation -> complex window (Hann/rms x dispersion, 1/N folded in) -> 2048-point
// inverse FFT -> depth crop -> |.| (or complex, or the linearised interferogram).
// spectral_kernel (8 A-scans per 512-thread workgroup, 2 waves per SIMD) and spectral16_kernel
// (W A-scans per W-wave workgroup, 4 waves per SIMD, the default) compute the same values.
//
// Reference: octrecon/core/spectral.py:136-153 (SpectralProcessor.magnitude),
// octrecon/core/gpu_kernels.py:10-38 (per-sample fused pre-FFT), octrecon/core/cpu_kernels.py:12-41.
//
// Work decomposition (one workgroup = 8 wave64 = 8 A-scans per group, persistent over groups;
// each A-scan owns one 16.5 KB LDS region):
//   A  each wave loads one A-scan (16-byte loads), subtracts the file's background row,
//      reduces the line mean and writes d - m to the first half of its region;
//   B  sinc: wave w computes output samples i = lane + 64 q, q in [4w, 4w+4), of ALL eight
//      A-scans with its taps and band starts held in registers (each operator entry is read
//      once per group of eight A-scans), and writes the equispaced interferogram to the
//      second half of each region;
//   D  each wave transforms its own A-scan: window multiply on load, N = 64 n1 + n2
//      (lane = n2) -> 32-point DFT over n1 in registers, twiddle W_N^(n2 k1) from an LDS
//      table, one padded LDS transpose, 32-point DFT over the even/odd halves of n2 in lane
//      pairs, and a DPP pair exchange that forms only the kept bins k = k1 + 32 k2, k2 < 32
//      (the first N/2 outputs).
// Two workgroup barriers per group (after A and after B).
#pragma once
#include "oct_common.hpp"

namespace oct {

constexpr int kN = 2048;
constexpr int kNZ = kN / 2;
constexpr int kQ = kN / 64;            // 32 values per lane in the transform layout
constexpr int kStride = 66;            // padded transpose row (complex elements)
constexpr int kBufCF = 32 * kStride;   // complex elements per A-scan buffer (>= kN)

enum OutKind { KIND_MAG = 0, KIND_CPLX = 1, KIND_EQ = 2 };

struct SpectralArgs {
    const void* raw;         // (B, lines, N) int16 / uint16 / float32
    const float* bg;         // (B, N) background row per raw file
    const int* band_start;   // (N) first source sample of each output's tap window
    const float* band_w;     // (N, TP) tap weights
    const float* win;        // (N, 2) complex window scaled by 1/N, in the pair order (pair_pos)
    long long win_fstride;   // floats between the windows of consecutive raw files (0: one window)
    long long win_rstride;   // floats between the windows of consecutive repeats of the batch
    int raw_files;           // files in raw / bg: A-scan file index f reads raw file f % raw_files
    const float2* tw;        // (32, 64) twiddles exp(+2 pi i n2 k1 / N), element k1 * 64 + n2 in the pair order
    const cf* c64;           // (65) exp(+2 pi i j / 64), j = 0..64: the pruned second stage's twiddles
    const float4* tw8;       // (8, 64): slot s of lane n2 = (W^(16 n2) for s = 0 else W^(s n2), W^((s + 8) n2))
    void* out;
    long long M;             // A-scans = B * nX
    int lines;               // A-lines per raw file
    int line0;               // first scan A-line of a file
    int nX;                  // scan A-lines per file
    int row0, rows;          // kept depth rows [row0, row0 + rows)
    int k2lo, k2n;           // k2n in {1, 2}: the rows lie in k2 blocks [k2lo, k2lo + k2n); 0: any rows
    long long out_stride;    // elements between consecutive A-scans in out
};

// exp(+2 pi i j / 64) for j = 0..64, built once per plan: the pruned second stage reads its
// twiddles W_32^r (entry 2 r) and W_64^k2 (entry k2) from it instead of evaluating them in double
// precision in every workgroup
__global__ void stage2_table_kernel(cf* c64) {
    const int j = threadIdx.x;
    if (j <= 64) {
        double sn, cs;
        sincospi((double)j / 32.0, &sn, &cs);
        c64[j] = cf{(float)cs, (float)sn};
    }
}

template <typename RawT>
__device__ __forceinline__ void unpack16(const uint4& u, float (&f)[16 / sizeof(RawT)]);

template <>
__device__ __forceinline__ void unpack16<int16_t>(const uint4& u, float (&f)[8]) {
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        f[2 * i] = (float)(int16_t)(w[i] & 0xFFFFu);
        f[2 * i + 1] = (float)(int16_t)(w[i] >> 16);
    }
}
template <>
__device__ __forceinline__ void unpack16<uint16_t>(const uint4& u, float (&f)[8]) {
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        f[2 * i] = (float)(w[i] & 0xFFFFu);
        f[2 * i + 1] = (float)(w[i] >> 16);
    }
}
template <>
__device__ __forceinline__ void unpack16<float>(const uint4& u, float (&f)[4]) {
    f[0] = __builtin_bit_cast(float, u.x);
    f[1] = __builtin_bit_cast(float, u.y);
    f[2] = __builtin_bit_cast(float, u.z);
    f[3] = __builtin_bit_cast(float, u.w);
}

// W = waves per workgroup = A-scans per group.  W = 8: one 148 KB workgroup per CU with the
// twiddle table in LDS; W = 4: two 68 KB workgroups per CU (one overlaps the other's
// barriers and loads), twiddles read from an L1-resident table (pair order).
template <typename RawT, int TP, int KIND, int W>
__global__ __launch_bounds__(64 * W, 8 / W) void spectral_kernel(SpectralArgs a) {
    constexpr int kWaves = W;
    constexpr int kThreads = 64 * W;
    constexpr int kQW = kQ / W;        // sinc outputs per lane per A-scan owned by one wave
    constexpr bool kLdsTw = (W == 8);
    __shared__ __align__(16) cf buf[kWaves][kBufCF];
    __shared__ __align__(16) cf tws[kLdsTw ? kQ : 1][kLdsTw ? 64 : 1];
    __shared__ float lmean[kWaves];
    __shared__ cf tw2[2][kQ];            // pruned stage: W_32^(m k2) for the (at most 2) k2 kept
    __shared__ cf tw64[2];               //               W_64^(k2)

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;

    if constexpr (KIND != KIND_EQ && kLdsTw) {
        for (int e = threadIdx.x; e < kQ * 64; e += kThreads) {
            const int k1 = e >> 6, n2 = e & 63;
            double s, c;
            sincospi((double)(n2 * k1) / (kN / 2), &s, &c);
            tws[k1][n2] = cf{(float)c, (float)s};
        }
        // made visible to all waves by the first barrier of the group loop
    }

    if constexpr (KIND != KIND_EQ) {
        if (a.k2n > 0 && threadIdx.x < 2 * kQ + 2) {
            const int t = threadIdx.x < 2 * kQ ? threadIdx.x / kQ : threadIdx.x - 2 * kQ;
            const int k2 = a.k2lo + t;                  // <= 64
            if (threadIdx.x < 2 * kQ) {
                const int m = threadIdx.x % kQ;
                tw2[t][m] = a.c64[2 * ((m * k2) % kQ)];     // W_32^(m k2)
            } else {
                tw64[t] = a.c64[k2];                        // W_64^k2
            }
        }
        // made visible to all waves by the first barrier of the group loop
    }

    constexpr int VEC = 16 / sizeof(RawT);
    constexpr int NLD = kN / (64 * VEC);
    const long long ngroups = (a.M + kWaves - 1) / kWaves;

    for (long long g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const long long p0 = g * kWaves;
        const int nvalid = (int)min((long long)kWaves, a.M - p0);
        const long long p = p0 + wave;
        float* D = reinterpret_cast<float*>(&buf[wave][0]);
        // A-scan p = (rep raw_files + b) nX + x; the host keeps M below 2^31, so 32-bit division
        const unsigned fv = (unsigned)p / (unsigned)a.nX;          // virtual file
        const unsigned rep = fv / (unsigned)a.raw_files;
        const unsigned b = fv - rep * (unsigned)a.raw_files;
        const unsigned x = (unsigned)p - fv * (unsigned)a.nX;

        // ---- A: load, background, mean, d - m -> LDS
        if (wave < nvalid) {
            const RawT* r = reinterpret_cast<const RawT*>(a.raw) +
                            (((long long)b * a.lines + a.line0 + x) * (long long)kN);
            const float* bgr = a.bg + (long long)b * kN;
            float d[NLD][VEC];
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < NLD; ++k) {
                const int j0 = VEC * lane + 64 * VEC * k;
                uint4 u = *reinterpret_cast<const uint4*>(r + j0);
                float f[VEC];
                unpack16<RawT>(u, f);
#pragma unroll
                for (int e4 = 0; e4 < VEC / 4; ++e4) {
                    float4 bv = *reinterpret_cast<const float4*>(bgr + j0 + 4 * e4);
                    d[k][4 * e4] = f[4 * e4] - bv.x;
                    d[k][4 * e4 + 1] = f[4 * e4 + 1] - bv.y;
                    d[k][4 * e4 + 2] = f[4 * e4 + 2] - bv.z;
                    d[k][4 * e4 + 3] = f[4 * e4 + 3] - bv.w;
                }
#pragma unroll
                for (int e = 0; e < VEC; ++e) s += d[k][e];
            }
            s = wave_sum_xor(s, (lane ^ 32) << 2);
            const float m = s * (1.0f / kN);
#pragma unroll
            for (int k = 0; k < NLD; ++k) {
                const int j0 = VEC * lane + 64 * VEC * k;
#pragma unroll
                for (int e4 = 0; e4 < VEC / 4; ++e4) {
                    float4 v{d[k][4 * e4] - m, d[k][4 * e4 + 1] - m, d[k][4 * e4 + 2] - m, d[k][4 * e4 + 3] - m};
                    *reinterpret_cast<float4*>(D + j0 + 4 * e4) = v;
                }
            }
            if (lane == 0) lmean[wave] = m;
        }
        __syncthreads();

        // ---- B: banded sinc for all A-scans of the group, this wave's output samples, in
        //      chunks of 4 samples per lane whose taps and band starts sit in registers (each
        //      table entry is read once per group); the interferogram goes to the second half
        //      of each A-scan's LDS region
#pragma unroll 1
        for (int qc = 0; qc < kQW; qc += 4) {
            int bs[4];
            float wt[4][TP];
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) {
                const int i = lane + 64 * (wave * kQW + qc + qq);
                bs[qq] = a.band_start[i];
                const float4* wp = reinterpret_cast<const float4*>(a.band_w + (long long)i * TP);
#pragma unroll
                for (int t4 = 0; t4 < TP / 4; ++t4) {
                    const float4 v = wp[t4];
                    wt[qq][4 * t4] = v.x;
                    wt[qq][4 * t4 + 1] = v.y;
                    wt[qq][4 * t4 + 2] = v.z;
                    wt[qq][4 * t4 + 3] = v.w;
                }
            }
#pragma unroll 1
            for (int aa = 0; aa < nvalid; ++aa) {
                const float* Da = reinterpret_cast<const float*>(&buf[aa][0]);
                float* Ea = reinterpret_cast<float*>(&buf[aa][0]) + kN;
                const float ma = lmean[aa];
#pragma unroll
                for (int qq = 0; qq < 4; ++qq) {
                    const int i = lane + 64 * (wave * kQW + qc + qq);
                    const float* src = Da + bs[qq];
                    float acc = 0.f;
#pragma unroll
                    for (int t = 0; t < TP; ++t) acc = __builtin_fmaf(wt[qq][t], src[t], acc);
                    const float v = acc + ma;
                    if constexpr (KIND == KIND_EQ) {
                        reinterpret_cast<float*>(a.out)[(p0 + aa) * a.out_stride + i] = v;
                    } else {
                        Ea[i] = v;
                    }
                }
            }
        }
        __syncthreads();
        if constexpr (KIND == KIND_EQ) continue;

        // ---- D: 2048-point inverse transform of this wave's A-scan
        if (wave < nvalid) {
            cf* B = &buf[wave][0];
            const float* E = reinterpret_cast<const float*>(B) + kN;
            const float2* wg = reinterpret_cast<const float2*>(a.win + (long long)rep * a.win_rstride +
                                                               (long long)b * a.win_fstride);
            cf v[kQ];
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const float e = E[lane + 64 * q];
                const float2 w = wg[pair_pos(lane + 64 * q)];      // window in the pair order
                v[q] = cf{e * w.x, e * w.y};
            }
            wave_lds_sync();
            dft_inv_dif<kQ>(v);                          // Y[k1] in v[bitrev(k1)]
            static_for<1, kQ>([&](auto k1) {
                constexpr int r = bitrev(decltype(k1)::value, 5);
                if constexpr (kLdsTw) {
                    v[r] = cmul(v[r], tws[decltype(k1)::value][lane]);
                } else {
                    const float2 t = a.tw[pair_pos(decltype(k1)::value * 64 + lane)];   // pair order
                    v[r] = cmul(v[r], cf{t.x, t.y});
                }
            });
            wave_lds_sync();
            static_for<0, kQ>([&](auto k1) {
                B[decltype(k1)::value * kStride + lane] = v[bitrev(decltype(k1)::value, 5)];
            });
            wave_lds_sync();
            const int kk1 = lane >> 1;
            const int h = lane & 1;
#pragma unroll
            for (int m = 0; m < kQ; ++m) v[m] = B[kk1 * kStride + 2 * m + h];
            if (a.k2n > 0) {
                // pruned stage: only the k2 blocks holding the kept rows, each
                // E_h[k2] = sum_m Z[2 m + h] W_32^(m k2) directly (32 MACs with tabled twiddles)
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    if (t < a.k2n) {
                        float ex = 0.f, ey = 0.f;
#pragma unroll
                        for (int m = 0; m < kQ; ++m) {
                            const cf w = tw2[t][m];
                            ex = __builtin_fmaf(v[m].x, w.x, __builtin_fmaf(-v[m].y, w.y, ex));
                            ey = __builtin_fmaf(v[m].x, w.y, __builtin_fmaf(v[m].y, w.x, ey));
                        }
                        cf e{ex, ey};
                        if (h) e = cmul(e, tw64[t]);
                        const cf o{e.x + swap_pair(e.x), e.y + swap_pair(e.y)};
                        const int k = kk1 + 32 * (a.k2lo + t);
                        if (h == t && k >= a.row0 && k < a.row0 + a.rows) {
                            if constexpr (KIND == KIND_MAG) {
                                reinterpret_cast<float*>(a.out)[p * a.out_stride + (k - a.row0)] =
                                    __builtin_amdgcn_sqrtf(o.x * o.x + o.y * o.y);
                            } else {
                                reinterpret_cast<cf*>(a.out)[p * a.out_stride + (k - a.row0)] = o;
                            }
                        }
                    }
                }
                continue;
            }
            dft_inv_dif<kQ>(v);                          // E_h[k2] in v[bitrev(k2)]
            if (h) {
                static_for<1, kQ>([&](auto k2) {
                    constexpr int r = bitrev(decltype(k2)::value, 5);
                    v[r] = twiddle<decltype(k2)::value, 64>(v[r]);
                });
            }
            // X[k1 + 32 k2] = F0[k2] + F1[k2]; lane h keeps k2 = 16 h + j, j < 16
            static_for<0, 16>([&](auto jj) {
                constexpr int j = decltype(jj)::value;
                constexpr int rlo = bitrev(j, 5);
                constexpr int rhi = bitrev(16 + j, 5);
                const cf send = h ? v[rlo] : v[rhi];
                const cf keep = h ? v[rhi] : v[rlo];
                const cf o{keep.x + swap_pair(send.x), keep.y + swap_pair(send.y)};
                const int k = kk1 + 32 * (16 * h + j);
                if (k >= a.row0 && k < a.row0 + a.rows) {
                    if constexpr (KIND == KIND_MAG) {
                        reinterpret_cast<float*>(a.out)[p * a.out_stride + (k - a.row0)] =
                            __builtin_amdgcn_sqrtf(o.x * o.x + o.y * o.y);
                    } else {
                        reinterpret_cast<cf*>(a.out)[p * a.out_stride + (k - a.row0)] = o;
                    }
                }
            });
        }
    }
}

// The same path at 4 waves per SIMD (spectral_kernel runs 2): W A-scans per group and one W-wave
// workgroup, 16 / W workgroups per CU.  Each A-scan owns one 8.25 KB region that holds in turn
//   d - m (phase A);
//   the equispaced interferogram, written in place once every wave has read every region
//   (phase B keeps its outputs in registers across a barrier);
//   the transform's transpose, real parts then imaginary parts (32 x 66 floats each).
// Phase B: wave w computes output samples i = lane + 64 (w 32/W + qq), qq < 32/W, of all W A-scans
// (each tap table entry read once per W A-scans).  Phase D is that of spectral_kernel.  Three
// workgroup barriers per group.  The launch bounds cap the kernel at 128 VGPRs.
constexpr int kRegionF = 32 * kStride;   // 2112 floats >= N

// STAGE2 (compiled apart: the union does not fit in 128 VGPRs):
//   0  the full second stage (any rows);
//   1  at most 32 + kStage2Extra kept rows within two 32-bin blocks (a.k2n > 0): the first 32
//      rows hold at most one bin per k1, so each lane pair forms the direct 32-term sum of that
//      bin's block only; each further row (bin k = row0 + 32 + e) is one 64-term sum over the
//      lanes, X[k] = sum_n2 Z[n2][k mod 32] W_64^(n2 (k / 32)), reduced across the wave;
//   2  more rows within two blocks: the direct sums of both blocks.
constexpr int kStage2Extra = 2;
// W = 16: one workgroup per CU, twiddles in LDS; W = 8 / 4 / 2: two / four / eight workgroups per
// CU (their loads and barriers overlap the others' arithmetic), twiddles from eight 16-byte loads
// of the L1-resident table a.tw8.
template <typename RawT, int TP, int KIND, int STAGE2, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4))) void spectral16_kernel(SpectralArgs a) {
    constexpr bool PRUNED = STAGE2 > 0;
    constexpr int kWaves = W;
    constexpr int kThreads = 64 * kWaves;
    constexpr int kQW = kQ / kWaves;         // output columns per lane per A-scan
    constexpr bool kLdsTw = (W == 16);
    __shared__ __align__(16) float buf[kWaves][kRegionF];
    __shared__ __align__(16) cf tws[kLdsTw ? kQ : 1][kLdsTw ? 64 : 1];
    __shared__ float lmean[kWaves];
    __shared__ cf tw2[2][kQ];
    __shared__ cf tw64[2];
    __shared__ cf w64t[STAGE2 == 1 ? 64 : 1];   // W_64^j, the extra rows' second-stage twiddles

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    if constexpr (KIND != KIND_EQ) {
        if constexpr (kLdsTw) {
            for (int e = threadIdx.x; e < kQ * 64; e += kThreads) {
                const int k1 = e >> 6, n2 = e & 63;
                double sn, cs;
                sincospi((double)(n2 * k1) / (kN / 2), &sn, &cs);
                tws[k1][n2] = cf{(float)cs, (float)sn};
            }
        }
        if (PRUNED && threadIdx.x < 2 * kQ + 2) {
            const int t = threadIdx.x < 2 * kQ ? threadIdx.x / kQ : threadIdx.x - 2 * kQ;
            const int k2 = a.k2lo + t;                  // <= 64
            if (threadIdx.x < 2 * kQ) {
                const int m = threadIdx.x % kQ;
                tw2[t][m] = a.c64[2 * ((m * k2) % kQ)];     // W_32^(m k2)
            } else {
                tw64[t] = a.c64[k2];                        // W_64^k2
            }
        }
        if constexpr (STAGE2 == 1) {
            if (threadIdx.x < 64) w64t[threadIdx.x] = a.c64[threadIdx.x];
        }
        // made visible to all waves by the first barrier of the group loop
    }

    constexpr int VEC = 16 / sizeof(RawT);
    constexpr int NLD = kN / (64 * VEC);
    constexpr int kEx = STAGE2 == 1 ? kStage2Extra : 1;     // extra rows read (STAGE2 == 1)
    const long long ngroups = (a.M + kWaves - 1) / kWaves;

    for (long long g = blockIdx.x; g < ngroups; g += gridDim.x) {
        // lane-derived addresses and row tests are rebuilt in every group: the empty asm makes
        // the lane index opaque, so the compiler cannot hoist dozens of them out of the loop
        // (under the 128-VGPR cap they would be spilled)
        int lane = threadIdx.x & 63;
        asm volatile("" : "+v"(lane));
        int row0 = a.row0, rows = a.rows;
        asm volatile("" : "+s"(row0), "+s"(rows));
        const long long p0 = g * kWaves;
        const int nvalid = (int)min((long long)kWaves, a.M - p0);
        const long long p = p0 + wave;
        float* R = &buf[wave][0];
        const unsigned fv = (unsigned)p / (unsigned)a.nX;          // virtual file
        const unsigned rep = fv / (unsigned)a.raw_files;
        const unsigned b = fv - rep * (unsigned)a.raw_files;
        const unsigned x = (unsigned)p - fv * (unsigned)a.nX;

        // ---- A: load, background, mean, d - m -> this A-scan's region
        if (wave < nvalid) {
            const RawT* r = reinterpret_cast<const RawT*>(a.raw) +
                            (((long long)b * a.lines + a.line0 + x) * (long long)kN);
            const float* bgr = a.bg + (long long)b * kN;
            float d[NLD][VEC];
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < NLD; ++k) {
                const int j0 = VEC * lane + 64 * VEC * k;
                uint4 u = *reinterpret_cast<const uint4*>(r + j0);
                float f[VEC];
                unpack16<RawT>(u, f);
#pragma unroll
                for (int e4 = 0; e4 < VEC / 4; ++e4) {
                    float4 bv = *reinterpret_cast<const float4*>(bgr + j0 + 4 * e4);
                    d[k][4 * e4] = f[4 * e4] - bv.x;
                    d[k][4 * e4 + 1] = f[4 * e4 + 1] - bv.y;
                    d[k][4 * e4 + 2] = f[4 * e4 + 2] - bv.z;
                    d[k][4 * e4 + 3] = f[4 * e4 + 3] - bv.w;
                }
#pragma unroll
                for (int e = 0; e < VEC; ++e) s += d[k][e];
            }
            s = wave_sum_xor(s, (lane ^ 32) << 2);
            const float m = s * (1.0f / kN);
#pragma unroll
            for (int k = 0; k < NLD; ++k) {
                const int j0 = VEC * lane + 64 * VEC * k;
#pragma unroll
                for (int e4 = 0; e4 < VEC / 4; ++e4) {
                    float4 v{d[k][4 * e4] - m, d[k][4 * e4 + 1] - m, d[k][4 * e4 + 2] - m, d[k][4 * e4 + 3] - m};
                    *reinterpret_cast<float4*>(R + j0 + 4 * e4) = v;
                }
            }
            if (lane == 0) lmean[wave] = m;
        }
        __syncthreads();

        // ---- B: banded sinc of this wave's output columns for all A-scans of the group, kept
        //      in registers until every wave has read every region; the columns' taps are read
        //      once per group (L1 / L2)
        // (regions of A-scans past the end of the batch hold stale values: computed, never stored)
        // columns in passes of at most 4, whose taps are held in registers
        constexpr int kCP = kQW < 4 ? kQW : 4;
        float res[kWaves][kQW];
#pragma unroll
        for (int cp = 0; cp < kQW / kCP; ++cp) {
            int bs[kCP];
            float wt[kCP][TP];
#pragma unroll
            for (int qq = 0; qq < kCP; ++qq) {
                const int i = lane + 64 * (wave * kQW + cp * kCP + qq);
                bs[qq] = a.band_start[i];
                const float4* wp = reinterpret_cast<const float4*>(a.band_w + (long long)i * TP);
#pragma unroll
                for (int t4 = 0; t4 < TP / 4; ++t4) {
                    const float4 v = wp[t4];
                    wt[qq][4 * t4] = v.x;
                    wt[qq][4 * t4 + 1] = v.y;
                    wt[qq][4 * t4 + 2] = v.z;
                    wt[qq][4 * t4 + 3] = v.w;
                }
            }
            unsigned bsb[kCP];                   // byte offsets of the band starts in a region
#pragma unroll
            for (int qq = 0; qq < kCP; ++qq) bsb[qq] = 4u * (unsigned)bs[qq];
#pragma unroll
            for (int aa = 0; aa < kWaves; ++aa) {
                const float ma = lmean[aa];
#pragma unroll
                for (int qq = 0; qq < kCP; ++qq) {
                    // one opaque byte address per (A-scan, column): the taps then use the LDS
                    // instructions' immediate offsets (left visible, the compiler forms one
                    // address per pair of taps, since a region's offset exceeds that field)
                    unsigned ba = bsb[qq] + (unsigned)(aa * kRegionF * 4);
                    asm volatile("" : "+v"(ba));
                    const float* src = reinterpret_cast<const float*>(reinterpret_cast<const char*>(&buf[0][0]) + ba);
                    float acc = 0.f;
#pragma unroll
                    for (int t = 0; t < TP; ++t) acc = __builtin_fmaf(wt[qq][t], src[t], acc);
                    res[aa][cp * kCP + qq] = acc + ma;
                    // materialise the sum here: otherwise the compiler sinks the taps' arithmetic
                    // past the barrier into the conditional stores and keeps every loaded tap live
                    asm volatile("" : "+v"(res[aa][cp * kCP + qq]));
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int aa = 0; aa < kWaves; ++aa) {
            if (aa < nvalid) {
#pragma unroll
                for (int qq = 0; qq < kQW; ++qq) {
                    const int i = lane + 64 * (wave * kQW + qq);
                    if constexpr (KIND == KIND_EQ) {
                        reinterpret_cast<float*>(a.out)[(p0 + aa) * a.out_stride + i] = res[aa][qq];
                    } else {
                        buf[aa][i] = res[aa][qq];
                    }
                }
            }
        }
        if constexpr (KIND == KIND_EQ) continue;      // the next group's first barrier orders the reuse
        __syncthreads();

        // ---- D: 2048-point inverse transform of this wave's A-scan
        if (wave < nvalid) {
            // the window in the pair order: one 16-byte load gives a lane its samples q and q + 16
            const float4* wg = reinterpret_cast<const float4*>(a.win + (long long)rep * a.win_rstride +
                                                               (long long)b * a.win_fstride);
            // the 32-point DFT over q by decimation in time: bit-reversed positions, so its first
            // radix-2 stage pairs samples q and q + 16, with the window multiply fused in
            cf v[kQ];
            static_for<0, kQ / 2>([&](auto iI) {
                constexpr int i = decltype(iI)::value;
                constexpr int q = bitrev(2 * i, 5);              // < 16; position 2 i + 1 holds q + 16
                const float e0 = R[lane + 64 * q], e1 = R[lane + 64 * (q + 16)];
                const float4 w01 = wg[q * 64 + lane];
                const float2 w0{w01.x, w01.y}, w1{w01.z, w01.w};
                const float px = e1 * w1.x, py = e1 * w1.y;
                v[2 * i] = cf{__builtin_fmaf(e0, w0.x, px), __builtin_fmaf(e0, w0.y, py)};
                v[2 * i + 1] = cf{__builtin_fmaf(e0, w0.x, -px), __builtin_fmaf(e0, w0.y, -py)};
            });
            wave_lds_sync();
            dft_inv_dit<kQ, 1>(v);                       // Y[k1] in v[k1]
            if constexpr (kLdsTw) {
                static_for<1, kQ>([&](auto k1) { v[decltype(k1)::value] = cmul(v[decltype(k1)::value], tws[decltype(k1)::value][lane]); });
            } else {
                // W^(n2 (j + 16)) = W^(16 n2) W^(n2 j): eight 16-byte loads bring W^(16 n2) and
                // W^(n2 j), j = 1..15; the upper half takes W^(16 n2) first
                const float4 t0 = a.tw8[lane];
                const cf u{t0.x, t0.y};
                v[16] = cmul(v[16], u);
                v[8] = cmul(v[8], cf{t0.z, t0.w});
                v[24] = cmul(cmul(v[24], u), cf{t0.z, t0.w});
                static_for<1, kQ / 4>([&](auto sS) {
                    constexpr int sl = decltype(sS)::value;
                    const float4 t = a.tw8[sl * 64 + lane];
                    const cf ta{t.x, t.y}, tb{t.z, t.w};
                    v[sl] = cmul(v[sl], ta);
                    v[sl + 16] = cmul(cmul(v[sl + 16], u), ta);
                    v[sl + 8] = cmul(v[sl + 8], tb);
                    v[sl + 24] = cmul(cmul(v[sl + 24], u), tb);
                });
            }
            // transpose through the region: real parts, then imaginary parts
            const int kk1 = lane >> 1;
            const int h = lane & 1;
            // (scheduling barriers keep each pass's stores and loads together: mixing them with
            // the transform's arithmetic would hold both halves and the arithmetic's values live)
            float vr[kQ];
            float exr[kEx], exi[kEx];                    // Z[lane][k1] of the extra rows (STAGE2 == 1)
            __builtin_amdgcn_sched_barrier(0);
            static_for<0, kQ>([&](auto k1) { R[decltype(k1)::value * kStride + lane] = v[decltype(k1)::value].x; });
            wave_lds_sync();
#pragma unroll
            for (int m = 0; m < kQ; ++m) vr[m] = R[kk1 * kStride + 2 * m + h];
            if constexpr (STAGE2 == 1) {
#pragma unroll
                for (int e = 0; e < kEx; ++e) exr[e] = R[((row0 + e) & 31) * kStride + lane];
            }
            wave_lds_sync();
            static_for<0, kQ>([&](auto k1) { R[decltype(k1)::value * kStride + lane] = v[decltype(k1)::value].y; });
            wave_lds_sync();
#pragma unroll
            for (int m = 0; m < kQ; ++m) v[m] = cf{vr[m], R[kk1 * kStride + 2 * m + h]};
            if constexpr (STAGE2 == 1) {
#pragma unroll
                for (int e = 0; e < kEx; ++e) exi[e] = R[((row0 + e) & 31) * kStride + lane];
            }
            __builtin_amdgcn_sched_barrier(0);
            // this A-scan's output row, indexed by the depth bin k (k - row0 in [0, rows) kept)
            const long long obase = p * a.out_stride - row0;
            if constexpr (STAGE2 == 1) {
                // the block of this lane pair's bin: k2lo unless k1 + 32 k2lo lies before row0
                const int t = (a.k2n == 2 && kk1 + 32 * a.k2lo < row0) ? 1 : 0;
                float ex = 0.f, ey = 0.f;
#pragma unroll
                for (int m = 0; m < kQ; ++m) {
                    const cf w = tw2[t][m];
                    ex = __builtin_fmaf(v[m].x, w.x, __builtin_fmaf(-v[m].y, w.y, ex));
                    ey = __builtin_fmaf(v[m].x, w.y, __builtin_fmaf(v[m].y, w.x, ey));
                }
                cf e{ex, ey};
                if (h) e = cmul(e, tw64[t]);
                const cf o{e.x + swap_pair(e.x), e.y + swap_pair(e.y)};
                const int k = kk1 + 32 * (a.k2lo + t);
                if (h == 0 && (unsigned)(k - row0) < (unsigned)rows) {
                    if constexpr (KIND == KIND_MAG) {
                        reinterpret_cast<float*>(a.out)[obase + k] = __builtin_amdgcn_sqrtf(o.x * o.x + o.y * o.y);
                    } else {
                        reinterpret_cast<cf*>(a.out)[obase + k] = o;
                    }
                }
                // rows row0 + 32 + e: bin k1 = (row0 + e) mod 32, k2 = (row0 + 32 + e) / 32
#pragma unroll
                for (int e = 0; e < kEx; ++e) {
                    if (e < rows - 32) {                 // uniform
                        const int ke = row0 + 32 + e;
                        const cf w = w64t[(lane * (ke >> 5)) & 63];
                        float xr = __builtin_fmaf(exr[e], w.x, -exi[e] * w.y);
                        float xi = __builtin_fmaf(exr[e], w.y, exi[e] * w.x);
                        xr = wave_sum_xor(xr, (lane ^ 32) << 2);
                        xi = wave_sum_xor(xi, (lane ^ 32) << 2);
                        if (lane == 0) {
                            if constexpr (KIND == KIND_MAG) {
                                reinterpret_cast<float*>(a.out)[obase + ke] = __builtin_amdgcn_sqrtf(xr * xr + xi * xi);
                            } else {
                                reinterpret_cast<cf*>(a.out)[obase + ke] = cf{xr, xi};
                            }
                        }
                    }
                }
            } else if constexpr (STAGE2 == 2) {
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    if (t < a.k2n) {
                        float ex = 0.f, ey = 0.f;
#pragma unroll
                        for (int m = 0; m < kQ; ++m) {
                            const cf w = tw2[t][m];
                            ex = __builtin_fmaf(v[m].x, w.x, __builtin_fmaf(-v[m].y, w.y, ex));
                            ey = __builtin_fmaf(v[m].x, w.y, __builtin_fmaf(v[m].y, w.x, ey));
                        }
                        cf e{ex, ey};
                        if (h) e = cmul(e, tw64[t]);
                        const cf o{e.x + swap_pair(e.x), e.y + swap_pair(e.y)};
                        const int k = kk1 + 32 * (a.k2lo + t);
                        if (h == t && (unsigned)(k - row0) < (unsigned)rows) {
                            if constexpr (KIND == KIND_MAG) {
                                reinterpret_cast<float*>(a.out)[obase + k] = __builtin_amdgcn_sqrtf(o.x * o.x + o.y * o.y);
                            } else {
                                reinterpret_cast<cf*>(a.out)[obase + k] = o;
                            }
                        }
                    }
                }
            } else {
                dft_inv_dif<kQ>(v);                      // E_h[k2] in v[bitrev(k2)]
                if (h) {
                    static_for<1, kQ>([&](auto k2) {
                        constexpr int r = bitrev(decltype(k2)::value, 5);
                        v[r] = twiddle<decltype(k2)::value, 64>(v[r]);
                    });
                }
                static_for<0, 16>([&](auto jj) {
                    constexpr int j = decltype(jj)::value;
                    constexpr int rlo = bitrev(j, 5);
                    constexpr int rhi = bitrev(16 + j, 5);
                    const cf send = h ? v[rlo] : v[rhi];
                    const cf keep = h ? v[rhi] : v[rlo];
                    const cf o{keep.x + swap_pair(send.x), keep.y + swap_pair(send.y)};
                    const int k = kk1 + 32 * (16 * h + j);
                    if ((unsigned)(k - row0) < (unsigned)rows) {
                        if constexpr (KIND == KIND_MAG) {
                            reinterpret_cast<float*>(a.out)[obase + k] = __builtin_amdgcn_sqrtf(o.x * o.x + o.y * o.y);
                        } else {
                            reinterpret_cast<cf*>(a.out)[obase + k] = o;
                        }
                    }
                });
            }
        }
        // no closing barrier: phase D touches only this wave's region, which only this wave
        // rewrites (next phase A); other waves read it only after the next group's first barrier
    }
}

// A-scan binning of the scan lines (Thorlabs AScanBinning; octrecon/core/spectral.py:20-46, the
// legacy filter2(ones(1,B)/B) + decimation): kept line c averages scan lines
// s0 + c B + t - left, t < B (zero outside), s0 = max(1, B / 2) - 1, left = (B + 1) / 2 - 1, as
// the float32 sum of the float32 products line * (1/B) in t order (no contraction: numpy rounds
// each product).  The first `keep` lines (background lines) are copied as float32.
// out (B, keep + n_kept, N) float32; one thread per output sample.
#pragma clang fp contract(off)
template <typename RawT>
__global__ __launch_bounds__(256) void bin_lines_kernel(const RawT* __restrict__ raw, float* __restrict__ out,
                                                        int files, int lines, int keep, int skip, int n_scan,
                                                        int binning, int n_kept, int N) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int out_lines = keep + n_kept;
    if (tid >= (long long)files * out_lines * N) return;
    const int j = (int)(tid % N);
    const long long t2 = tid / N;
    const int l = (int)(t2 % out_lines);
    const int b = (int)(t2 / out_lines);
    const RawT* r = raw + (long long)b * lines * N + j;
    float v;
    if (l < keep) {
        v = (float)r[(long long)l * N];
    } else {
        const float h = 1.0f / (float)binning;
        const int s0 = max(1, binning / 2) - 1, left = (binning + 1) / 2 - 1;
        const int first = s0 + (l - keep) * binning - left;
        v = 0.f;
        for (int t = 0; t < binning; ++t) {
            const int sl = first + t;
            const float x = (sl >= 0 && sl < n_scan) ? (float)r[(long long)(skip + sl) * N] : 0.f;
            const float term = x * h;
            v = (t == 0) ? term : v + term;
        }
    }
    out[tid] = v;
}

// Repeat averaging (octrecon/core/spectral.py:168-177): mag (n_frames nb, size_x na, rows) ->
// out (n_frames, size_x, rows) = mean over the A-scan repeats of the means over the B-scan
// repeats; float32 sums in index order, each divided by its count.
__global__ __launch_bounds__(256) void average_repeats_kernel(const float* __restrict__ mag, float* __restrict__ out,
                                                              int n_frames, int nb, int size_x, int na, int rows,
                                                              long long in_stride, long long out_stride) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long long)n_frames * size_x * rows) return;
    const int z = (int)(tid % rows);
    const long long t2 = tid / rows;
    const int x = (int)(t2 % size_x);
    const int f = (int)(t2 / size_x);
    float acc_a = 0.f;
    for (int ia = 0; ia < na; ++ia) {
        float acc_b = 0.f;
        for (int ib = 0; ib < nb; ++ib) {
            const float v = mag[((long long)(f * nb + ib) * size_x * na + (long long)x * na + ia) * in_stride + z];
            acc_b = (ib == 0) ? v : acc_b + v;
        }
        const float mb = (nb > 1) ? acc_b / (float)nb : acc_b;
        acc_a = (ia == 0) ? mb : acc_a + mb;
    }
    out[((long long)f * size_x + x) * out_stride + z] = (na > 1) ? acc_a / (float)na : acc_a;
}
#pragma clang fp contract(fast)

// Background row per raw file.
//   mode 0 ("lines"): mean of the first `apod` A-lines of the file (float64 sum, stored float32)
//   mode 1 ("frame_mean"): mean over all A-lines of each group of `group` files, float32
//   accumulation in line order, then the mean over the group's files.
// Reference: octrecon/core/cpu_kernels.py:20-25; octrecon/core/spectral.py:54-66.
template <typename RawT>
__global__ void background_kernel(const RawT* __restrict__ raw, float* __restrict__ bg, int B, int lines,
                                  int N, int apod, int mode, int group) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long long)B * N) return;
    const int b = (int)(tid / N);
    const int j = (int)(tid - (long long)b * N);
    if (mode == 0) {
        const RawT* r = raw + (long long)b * lines * N + j;
        double s = 0.0;
        for (int l = 0; l < apod; ++l) s += (double)r[(long long)l * N];
        bg[tid] = (float)(s / apod);
    } else {
        const int g0 = (b / group) * group;
        float acc = 0.f;
        for (int f = 0; f < group; ++f) {
            const RawT* r = raw + (long long)(g0 + f) * lines * N + j;
            float s = 0.f;
            for (int l = 0; l < lines; ++l) s += (float)r[(long long)l * N];
            const float fm = s / (float)lines;
            acc = (f == 0) ? fm : acc + fm;
        }
        bg[tid] = acc / (float)group;
    }
}

// Mode 0 of background_kernel with 16-byte loads: one thread per 16 / sizeof(RawT) consecutive
// samples of a file, the line loop unrolled so that several lines' loads are in flight.  Same
// float64 sums in line order, so the same values.  Needs N % (16 / sizeof(RawT)) == 0 and 16-byte
// aligned raw and bg.
template <typename RawT>
__global__ __launch_bounds__(256) void background_lines_kernel(const RawT* __restrict__ raw, float* __restrict__ bg,
                                                               int B, int lines, int N, int apod) {
    constexpr int VEC = 16 / sizeof(RawT);
    const int nv = N / VEC;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)B * nv) return;
    const int b = (int)(t / nv);
    const int j0 = (int)(t - (long long)b * nv) * VEC;
    const RawT* r = raw + (long long)b * lines * N + j0;
    double s[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) s[e] = 0.0;
#pragma unroll 5
    for (int l = 0; l < apod; ++l) {
        float f[VEC];
        unpack16<RawT>(*reinterpret_cast<const uint4*>(r + (long long)l * N), f);
#pragma unroll
        for (int e = 0; e < VEC; ++e) s[e] += (double)f[e];
    }
    float o[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) o[e] = (float)(s[e] / apod);
    float* dst = bg + (long long)b * N + j0;
#pragma unroll
    for (int e = 0; e < VEC; e += 4) *reinterpret_cast<float4*>(dst + e) = float4{o[e], o[e + 1], o[e + 2], o[e + 3]};
}

}  // namespace oct
