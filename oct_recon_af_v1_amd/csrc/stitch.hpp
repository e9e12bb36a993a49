// Tile stitching, finalisation, volume clim and legacy quantisation kernels.
//
// Reference: octrecon/core/stitching.py:122-140 (TileStitcher.contribute), octrecon/pipeline.py
// :282-284 (accumulation), :296-300 (finalise), :376-388 (clim), octrecon/io/tiff_writer.py
// :26-45, 62-65 (quantisation).
#pragma once
#include "oct_common.hpp"

namespace oct {

// One tile's contribution, output-owned: thread (b, zo, c) forms
//   num = sum_x Wx[c,x] sum_r Wz[zo,r] fz[r] valid(b,x,r) mag[b, x, clip(r + shift)]
//   den = sum_x Wx[c,x] sum_r Wz[zo,r] fz[r] valid(b,x,r)
// from the two non-zero taps of each interpolation row, and adds (or stores) them into the
// accumulators; consecutive threads own consecutive output columns.
struct StitchArgs {
    const float* mag;        // magnitude rows [mrow0, mrow0 + mrows) of each A-scan
    long long mstride;       // floats between A-scans in mag
    int mrow0, mrows;
    int nX;                  // A-scans per frame in the tile
    int nZ;                  // rows of the tile depth grid
    const int* frame_idx;    // (B) tile-local y frame of each batch entry
    const int* shift;        // (nY, nX) optical-path row shift, or null
    const int* vlo;          // (nY, nX) first valid row (row + P/dz >= 0)
    const int* vhi;          // (nY, nX) last valid row  (row + P/dz <= nZ - 1)
    const float* fz;         // (nZ) focus weight per tile row
    const int* zr;           // (nOutZ, 2) tile rows of the two z taps
    const float* zw;         // (nOutZ, 2) their weights
    const int* xr;           // (C, 2) tile columns of the two x taps
    const float* xw;         // (C, 2) their weights
    float* num;              // accumulators: element (b, zo, c) at
    float* den;              //   [(b + out_b0) * nOutZ * ldx + zo * ldx + out_c0 + c]
    int B, nOutZ, C;
    long long ldx;
    int out_b0, out_c0;
    int accumulate;
};

__global__ __launch_bounds__(256) void stitch_kernel(StitchArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int zo = blockIdx.y;
    const int b = blockIdx.z;
    if (c >= a.C) return;
    const int fb = a.frame_idx[b];
    float num = 0.f, den = 0.f;
#pragma unroll
    for (int tx = 0; tx < 2; ++tx) {
        const float wx = a.xw[2 * c + tx];
        if (wx == 0.f) continue;
        const int x = a.xr[2 * c + tx];
        int sh = 0, lo = 0, hi = a.nZ - 1;
        if (a.shift) {
            const long long o = (long long)fb * a.nX + x;
            sh = a.shift[o];
            lo = a.vlo[o];
            hi = a.vhi[o];
        }
        const float* mrow = a.mag + ((long long)b * a.nX + x) * a.mstride - a.mrow0;
        float sa = 0.f, sf = 0.f;
#pragma unroll
        for (int tz = 0; tz < 2; ++tz) {
            const float wz = a.zw[2 * zo + tz];
            if (wz == 0.f) continue;
            const int r = a.zr[2 * zo + tz];
            const float f = (r >= lo && r <= hi) ? a.fz[r] : 0.f;
            const int src = min(max(r + sh, 0), a.nZ - 1);
            const float av = mrow[src] * f;
            sa = __builtin_fmaf(av, wz, sa);
            sf = __builtin_fmaf(f, wz, sf);
        }
        num = __builtin_fmaf(sa, wx, num);
        den = __builtin_fmaf(sf, wx, den);
    }
    const long long o = ((long long)(b + a.out_b0) * a.nOutZ + zo) * a.ldx + a.out_c0 + c;
    if (a.accumulate) {
        a.num[o] += num;
        a.den[o] += den;
    } else {
        a.num[o] = num;
        a.den[o] = den;
    }
}

// Order-preserving map of a float to an int key (for min/max); NaN never reaches it.
__device__ __forceinline__ int fkey(float f);

// ---------------------------------------------------------------- fused y-row stitch + finalise
// All tiles of one y-tile row in one pass, output-owned: thread (f, zo, c) accumulates
//   num = sum_tiles sum_x Wx sum_r Wz fz valid mag,  den = the same without mag
// in registers over every tile whose columns cover c, then finalises in place:
// den < thr -> NaN, else 20 log10(num / den), and keeps per-block finite min/max partials
// (same partial layout as finalize_kernel, so clim_planes_kernel reduces a volume of rows).
// Equivalent to the per-tile contributions + accumulation + finalise of
// octrecon/pipeline.py:261-299, without num/den round trips through memory.
constexpr int kRowMaxTiles = 40;   // keeps the by-value argument block under 4 KB
struct RowTile {
    const float* mag;       // (nF, nX, mstride) rows [mrow0, mrow0 + mstride) of the tile's frames
    long long mstride;
    int mrow0;
    int c0, C;              // output columns [c0, c0 + C)
    const int* zr;          // (nOutZ, 2)
    const float* zw;
    const int* xr;          // (C, 2)
    const float* xw;
    const float* fz;        // (nZ)
};
struct RowArgs {
    RowTile t[kRowMaxTiles];
    const RowTile* tt;      // the tile table in device memory (the kTable variant: any count)
    int ntiles;
    int nX, nZ;
    const int* frame_idx;   // (nF) tile-local y frame of each output plane (for the optical path)
    const int* shift;       // (nY, nX) or null
    const int* vlo;
    const int* vhi;
    int nOutZ, nOutX;
    float thr;
    float* db;              // plane p = plane0 + f: db[(p * nOutZ + zo) * nOutX + c]
    int plane0;
    int* partial;           // (planes, blocks_per_plane, 2)
    int blocks_per_plane;
    int nplanes;            // planes of this call
};

// Block tile: kRowZ output rows x kRowC output columns of one plane.  Lane l of a wave owns row
// zo = z0 + l % kRowZ and a column pair position l / kRowZ, so the magnitude reads of a wave hit
// one or two consecutive A-scans (rows are contiguous in the tile's magnitude), and the z-weight
// reads are contiguous.  The finalised values go through LDS so that the dB stores run along x.
constexpr int kRowZ = 32, kRowC = 64, kRowWaves = 4;
constexpr int kRowLanesC = 64 / kRowZ;                        // columns in flight per wave
constexpr int kRowPasses = kRowC / (kRowWaves * kRowLanesC);  // column passes per wave

// kShift: optical-path shift present (kernels without it carry no per-x lookups);
// kTable: tile descriptors read from the device table a.tt instead of the argument block.
// A workgroup makes its block of kRowPlanes consecutive planes: the tap loads of a tile (depth
// rows and weights, focus factors, column taps) are made once for all of them, and the planes'
// magnitude loads are independent, so each dependent load round has kRowPlanes times as many
// loads in flight.  Per voxel the summation order is unchanged.
constexpr int kRowPlanes = 4;

template <bool kShift, bool kTable>
__global__ __launch_bounds__(64 * kRowWaves) void stitch_row_kernel(RowArgs a) {
    constexpr int P = kRowPlanes;
    __shared__ float tile[kRowZ][kRowC + 1];
    __shared__ int smin[kRowWaves], smax[kRowWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs, so block id i runs
    // on XCD i % 8.  Each XCD takes a contiguous range of the work list, ordered depth block
    // fastest, so the depth blocks of one column range share that XCD's L2.
    const int gx = (a.nOutX + kRowC - 1) / kRowC, gz = (a.nOutZ + kRowZ - 1) / kRowZ;
    const int ngroups = (a.nplanes + P - 1) / P;
    const long long total = (long long)gx * gz * ngroups;
    const long long per_xcd = (long long)gridDim.x / 8;
    const long long w = (long long)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (w >= total) return;                                // uniform per workgroup
    const int zb = (int)(w % gz);
    const int xb = (int)((w / gz) % gx);
    const int f0 = (int)(w / ((long long)gz * gx)) * P;
    const int np = min(P, a.nplanes - f0);                 // planes of this block (uniform)
    const int zo = zb * kRowZ + (lane % kRowZ);
    const int cb0 = xb * kRowC;
    const bool zok = zo < a.nOutZ;
    int fp[P];                                             // plane -> frame (clamped past the last)
#pragma unroll
    for (int p = 0; p < P; ++p) fp[p] = f0 + min(p, np - 1);
    float num[kRowPasses][P], den[kRowPasses][P];
#pragma unroll
    for (int q = 0; q < kRowPasses; ++q)
#pragma unroll
        for (int p = 0; p < P; ++p) num[q][p] = den[q][p] = 0.f;
    // Tiles outer (in order), columns inner: per voxel the same summation order as the
    // per-tile path (tile, then x tap, then z tap).  The body is branch-free so that the loads
    // of all columns issue together: taps outside a tile or of zero weight read an in-range
    // element and contribute an exact +0 through their zero weight.
    const int zc = zok ? zo : 0;
    for (int ti = 0; ti < a.ntiles; ++ti) {
        const RowTile& t = kTable ? a.tt[ti] : a.t[ti];
        if (cb0 + kRowC <= t.c0 || cb0 >= t.c0 + t.C) continue;       // uniform per block
        const int2 rr = reinterpret_cast<const int2*>(t.zr)[zc];
        const float2 ww = reinterpret_cast<const float2*>(t.zw)[zc];
        const int r0 = rr.x, r1 = rr.y;
        const float wz0 = zok ? ww.x : 0.f, wz1 = zok ? ww.y : 0.f;
        const float f0w = t.fz[min(max(r0, 0), a.nZ - 1)], f1w = t.fz[min(max(r1, 0), a.nZ - 1)];
        const int slo = max(t.mrow0, 0), shi = min(t.mrow0 + (int)t.mstride, a.nZ) - 1;   // computed rows
#pragma unroll
        for (int q = 0; q < kRowPasses; ++q) {
            const int c = cb0 + (q * kRowWaves + wave) * kRowLanesC + lane / kRowZ;
            const int cl = c - t.c0;
            const bool in = c < a.nOutX && cl >= 0 && cl < t.C;
            const int cc = min(max(cl, 0), t.C - 1);
            const int2 xr = reinterpret_cast<const int2*>(t.xr)[cc];
            const float2 xw = reinterpret_cast<const float2*>(t.xw)[cc];
#pragma unroll
            for (int tx = 0; tx < 2; ++tx) {
                const float wx = in ? (tx ? xw.y : xw.x) : 0.f;
                const int x = tx ? xr.y : xr.x;
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    int sh = 0, lo = 0, hi = a.nZ - 1;
                    if constexpr (kShift) {
                        const long long o = (long long)a.frame_idx[fp[p]] * a.nX + x;
                        sh = a.shift[o];
                        lo = a.vlo[o];
                        hi = a.vhi[o];
                    }
                    const float* mrow = t.mag + ((long long)fp[p] * a.nX + x) * t.mstride - t.mrow0;
                    const float fw0 = (r0 >= lo && r0 <= hi) ? f0w : 0.f;
                    const float fw1 = (r1 >= lo && r1 <= hi) ? f1w : 0.f;
                    const int s0 = min(max(min(max(r0 + sh, 0), a.nZ - 1), slo), shi);
                    const int s1 = min(max(min(max(r1 + sh, 0), a.nZ - 1), slo), shi);
                    float sa = __builtin_fmaf(mrow[s0] * fw0, wz0, 0.f);
                    float sf = __builtin_fmaf(fw0, wz0, 0.f);
                    sa = __builtin_fmaf(mrow[s1] * fw1, wz1, sa);
                    sf = __builtin_fmaf(fw1, wz1, sf);
                    num[q][p] = __builtin_fmaf(sa, wx, num[q][p]);
                    den[q][p] = __builtin_fmaf(sf, wx, den[q][p]);
                }
            }
        }
    }
    // finalise the block's planes one at a time through the LDS tile (coalesced stores)
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if (p >= np) break;                                // uniform per workgroup
        int kmin = 0x7FFFFFFF, kmax = (int)0x80000000;
#pragma unroll
        for (int q = 0; q < kRowPasses; ++q) {
            const int cl0 = (q * kRowWaves + wave) * kRowLanesC + lane / kRowZ;
            if (zok && cb0 + cl0 < a.nOutX) {
                const float v = (den[q][p] >= a.thr) ? 20.0f * log10f(num[q][p] / den[q][p]) : __builtin_nanf("");
                tile[lane % kRowZ][cl0] = v;
                if (__builtin_isfinite(v)) {
                    const int kk = fkey(v);
                    kmin = min(kmin, kk);
                    kmax = max(kmax, kk);
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            kmin = min(kmin, __shfl_xor(kmin, off));
            kmax = max(kmax, __shfl_xor(kmax, off));
        }
        if (lane == 0) {
            smin[wave] = kmin;
            smax[wave] = kmax;
        }
        __syncthreads();
        const int f = f0 + p;
        for (int e = threadIdx.x; e < kRowZ * kRowC; e += 64 * kRowWaves) {
            const int zr = e / kRowC, cc = e - zr * kRowC;
            const int z = zb * kRowZ + zr, c = cb0 + cc;
            if (z < a.nOutZ && c < a.nOutX) a.db[((long long)(a.plane0 + f) * a.nOutZ + z) * a.nOutX + c] = tile[zr][cc];
        }
        if (threadIdx.x == 0) {
            int mn = smin[0], mx = smax[0];
            for (int w2 = 1; w2 < kRowWaves; ++w2) {
                mn = min(mn, smin[w2]);
                mx = max(mx, smax[w2]);
            }
            const long long o = ((long long)(a.plane0 + f) * a.blocks_per_plane + (long long)zb * gx + xb) * 2;
            a.partial[o] = mn;
            a.partial[o + 1] = mx;
        }
        __syncthreads();                                   // tile and smin reused by the next plane
    }
}
__device__ __forceinline__ int fkey(float f) {
    int i = __builtin_bit_cast(int, f);
    return i >= 0 ? i : (i ^ 0x7FFFFFFF);
}
__device__ __forceinline__ float fkey_inv(int k) {
    return __builtin_bit_cast(float, k >= 0 ? k : (k ^ 0x7FFFFFFF));
}

constexpr int kFinThreads = 256;

// den < thr -> NaN; db = 20 log10(num / den); per-(plane, block) finite min/max partials.
// Grid: (blocks_per_plane, nPlanes).
__global__ __launch_bounds__(kFinThreads) void finalize_kernel(const float* __restrict__ num,
                                                               const float* den,
                                                               float* __restrict__ db, long long plane_elems,
                                                               float thr, int* __restrict__ partial) {
    const int plane = blockIdx.y;
    const long long base = (long long)plane * plane_elems;
    int kmin = 0x7FFFFFFF, kmax = (int)0x80000000;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < plane_elems;
         e += (long long)gridDim.x * blockDim.x) {
        float v;
        if (den) {
            const float d = den[base + e];
            if (!(d >= thr)) {          // d < thr, or d is NaN
                v = __builtin_nanf("");
            } else {
                v = 20.0f * log10f(num[base + e] / d);
            }
        } else {
            v = 20.0f * log10f(num[base + e]);
        }
        db[base + e] = v;
        if (__builtin_isfinite(v)) {
            const int k = fkey(v);
            kmin = min(kmin, k);
            kmax = max(kmax, k);
        }
    }
    __shared__ int smin[kFinThreads], smax[kFinThreads];
    smin[threadIdx.x] = kmin;
    smax[threadIdx.x] = kmax;
    __syncthreads();
    for (int s = kFinThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = min(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const long long o = ((long long)plane * gridDim.x + blockIdx.x) * 2;
        partial[o] = smin[0];
        partial[o + 1] = smax[0];
    }
}

// Per-plane clim (NaN when the plane has no finite value) and the global clim
// (nanmin / nanmax over planes), one block.
// Plane and global clims from the per-(plane, block) finite min / max keys: one wave per plane
// (its lanes take the blocks in turn, then a wave reduction), 4 planes per workgroup.  The global
// clim follows from the plane clims (clim_global_kernel).
__global__ __launch_bounds__(256) void clim_planes_kernel(const int* __restrict__ partial, int nblk, int nplanes,
                                                          float* __restrict__ plane_clim) {
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= nplanes) return;                                   // whole wave
    int kmin = 0x7FFFFFFF, kmax = (int)0x80000000;
    const int2* pp = reinterpret_cast<const int2*>(partial) + (long long)p * nblk;
    for (int i = lane; i < nblk; i += 64) {
        const int2 v = pp[i];
        kmin = min(kmin, v.x);
        kmax = max(kmax, v.y);
    }
    for (int off = 32; off > 0; off >>= 1) {
        kmin = min(kmin, __shfl_xor(kmin, off));
        kmax = max(kmax, __shfl_xor(kmax, off));
    }
    if (lane == 0) {
        const bool any = kmin <= kmax;
        plane_clim[2 * p] = any ? fkey_inv(kmin) : __builtin_nanf("");
        plane_clim[2 * p + 1] = any ? fkey_inv(kmax) : __builtin_nanf("");
    }
}
// clim = the finite min / max over the plane clims (a plane without finite values holds NaN)
__global__ __launch_bounds__(kFinThreads) void clim_global_kernel(const float* __restrict__ plane_clim, int nplanes,
                                                                  float* __restrict__ clim) {
    int gmin = 0x7FFFFFFF, gmax = (int)0x80000000;
    for (int p = threadIdx.x; p < nplanes; p += blockDim.x) {
        const float lo = plane_clim[2 * p], hi = plane_clim[2 * p + 1];
        if (!__builtin_isnan(lo)) {
            gmin = min(gmin, fkey(lo));
            gmax = max(gmax, fkey(hi));
        }
    }
    __shared__ int smin[kFinThreads], smax[kFinThreads];
    smin[threadIdx.x] = gmin;
    smax[threadIdx.x] = gmax;
    __syncthreads();
    for (int s = kFinThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = min(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const bool any = smin[0] <= smax[0];
        clim[0] = any ? fkey_inv(smin[0]) : __builtin_nanf("");
        clim[1] = any ? fkey_inv(smax[0]) : __builtin_nanf("");
    }
}


// ---------------------------------------------------------------- legacy quantisation
// All in float64 with no contraction, operation for operation as the host formulas:
//   bits = min(uint16(round_half_away((db - c1) / (c2 - c1) * 65534)) + 1, 65535), NaN -> 0
//   dequantised = (bits - 1) * (c2 - c1) / 65534 + c1, bits == 0 -> NaN
#pragma clang fp contract(off)
__device__ __forceinline__ unsigned db_to_bits_1(double db, double c1, double c2) {
    if (__builtin_isnan(db)) return 0u;
    const double v = (db - c1) / (c2 - c1) * 65534.0;
    double r;
    if (__builtin_isnan(v)) {
        r = 0.0;
    } else {
        r = __builtin_floor(__builtin_fabs(v) + 0.5);
        if (v < 0) r = -r;
        if (v == 0) r = 0.0;
    }
    r = r < 0.0 ? 0.0 : (r > 65535.0 ? 65535.0 : r);
    unsigned u = (unsigned)r + 1u;
    return u > 65535u ? 65535u : u;
}
__device__ __forceinline__ double bits_to_db_1(unsigned bits, double c1, double c2) {
    if (bits == 0u) return __builtin_nan("");
    return ((double)bits - 1.0) * (c2 - c1) / 65534.0 + c1;
}
#pragma clang fp contract(fast)

__global__ __launch_bounds__(256) void quantize_kernel(const float* __restrict__ db, uint16_t* __restrict__ bits,
                                                       long long plane_elems, int nplanes,
                                                       const float* __restrict__ clim,
                                                       const float* __restrict__ plane_clim, int legacy_double) {
    const long long n = plane_elems * nplanes;
    double c1 = (double)clim[0], c2 = (double)clim[1];
    if (c1 > c2) { const double t = c1; c1 = c2; c2 = t; }
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        double v = (double)db[e];
        if (legacy_double) {
            const int p = (int)(e / plane_elems);
            double p1 = (double)plane_clim[2 * p], p2 = (double)plane_clim[2 * p + 1];
            if (p1 > p2) { const double t = p1; p1 = p2; p2 = t; }
            v = bits_to_db_1(db_to_bits_1(v, p1, p2), p1, p2);
        }
        bits[e] = (uint16_t)db_to_bits_1(v, c1, c2);
    }
}

// The same per element, one plane per grid row (no per-element plane division) and four
// elements per thread (16-byte loads, 8-byte stores).  plane_elems % 4 == 0; db and bits aligned
// to 16 and 8 bytes.  Grid (ceil(plane_elems / 1024), nplanes).
__global__ __launch_bounds__(256) void quantize4_kernel(const float* __restrict__ db, uint16_t* __restrict__ bits,
                                                        long long plane_elems, const float* __restrict__ clim,
                                                        const float* __restrict__ plane_clim, int legacy_double) {
    const long long q = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (q >= plane_elems) return;
    const int p = blockIdx.y;
    double c1 = (double)clim[0], c2 = (double)clim[1];
    if (c1 > c2) { const double t = c1; c1 = c2; c2 = t; }
    double p1 = 0.0, p2 = 0.0;
    if (legacy_double) {
        p1 = (double)plane_clim[2 * p];
        p2 = (double)plane_clim[2 * p + 1];
        if (p1 > p2) { const double t = p1; p1 = p2; p2 = t; }
    }
    const long long e = (long long)p * plane_elems + q;
    const float4 v4 = *reinterpret_cast<const float4*>(db + e);
    const float vin[4] = {v4.x, v4.y, v4.z, v4.w};
    unsigned o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double v = (double)vin[i];
        if (legacy_double) v = bits_to_db_1(db_to_bits_1(v, p1, p2), p1, p2);
        o[i] = db_to_bits_1(v, c1, c2);
    }
    *reinterpret_cast<uint2*>(bits + e) = uint2{o[0] | (o[1] << 16), o[2] | (o[3] << 16)};
}

}  // namespace oct
