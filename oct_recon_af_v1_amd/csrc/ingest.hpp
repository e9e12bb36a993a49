// Raw-file ingest on the device.
#pragma once
#include "oct_common.hpp"

namespace oct {

// Thorlabs SRR payload -> the Thorlabs B-scan file layout (octrecon/io/srr.py:107-124): each
// file's uint16 samples (file_lines, n) keep their low 12 bits (the legacy rem(data, 4096)), and
// output file f holds the background A-lines [apod0, apod1) followed by the scan A-lines
// [scan0, scan1) of input file f, as int16.  One thread per 8 samples (16-byte loads/stores).
__global__ __launch_bounds__(256) void srr_unpack_kernel(const uint16_t* __restrict__ in, int16_t* __restrict__ out,
                                                         int files, int file_lines, int n, int apod0, int apod1,
                                                         int scan0, int scan1) {
    const int a = apod1 - apod0, L = a + (scan1 - scan0);
    const int n8 = n / 8;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long long)files * L * n8) return;
    const int j8 = (int)(tid % n8);
    const long long t2 = tid / n8;
    const int l = (int)(t2 % L);
    const int f = (int)(t2 / L);
    const int src = l < a ? apod0 + l : scan0 + (l - a);
    uint4 v = reinterpret_cast<const uint4*>(in + ((long long)f * file_lines + src) * n)[j8];
    v.x &= 0x0FFF0FFFu;
    v.y &= 0x0FFF0FFFu;
    v.z &= 0x0FFF0FFFu;
    v.w &= 0x0FFF0FFFu;
    reinterpret_cast<uint4*>(out + ((long long)f * L + l) * n)[j8] = v;
}

}  // namespace oct

namespace oct {

// ---------------------------------------------------------------- PackBits (TIFF compression 32773)
// Each row of uint16 codes (little-endian bytes) is encoded on its own as runs of at most 128
// bytes: the row's bytes are cut into chunks of 128 (the last one shorter); a chunk whose bytes
// are all equal (2 bytes or more) becomes a replicate run (header 1 - n, the byte), any other
// chunk a literal run (header n - 1, the n bytes).  Every decoder of the TIFF 6.0 PackBits
// scheme restores the row; the split is simpler than a byte-by-byte run search and loses
// nothing on the noisy rows of a volume, while uniform stretches (NaN planes: code 0) still
// shrink 64-fold.  packbits_size_kernel: bytes of each chunk; packbits_write_kernel: the runs at
// the exclusive prefix sum of those sizes.
constexpr int kPbChunk = 128;

__device__ __forceinline__ bool pb_uniform(const uint8_t* p, int n) {
    const uint8_t b0 = p[0];
    bool same = n >= 2;
    for (int i = 1; i < n && same; ++i) same = (p[i] == b0);
    return same;
}

__global__ __launch_bounds__(256) void packbits_size_kernel(const uint16_t* __restrict__ codes, long long rows,
                                                            int row_elems, int* __restrict__ chunk_bytes) {
    const int rb = 2 * row_elems;
    const int nch = (rb + kPbChunk - 1) / kPbChunk;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= rows * nch) return;
    const long long r = tid / nch;
    const int c = (int)(tid - r * nch);
    const uint8_t* p = reinterpret_cast<const uint8_t*>(codes + r * row_elems) + c * kPbChunk;
    const int n = min(kPbChunk, rb - c * kPbChunk);
    chunk_bytes[tid] = pb_uniform(p, n) ? 2 : n + 1;
}

__global__ __launch_bounds__(256) void packbits_write_kernel(const uint16_t* __restrict__ codes, long long rows,
                                                             int row_elems, const long long* __restrict__ offsets,
                                                             uint8_t* __restrict__ out) {
    const int rb = 2 * row_elems;
    const int nch = (rb + kPbChunk - 1) / kPbChunk;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= rows * nch) return;
    const long long r = tid / nch;
    const int c = (int)(tid - r * nch);
    const uint8_t* p = reinterpret_cast<const uint8_t*>(codes + r * row_elems) + c * kPbChunk;
    const int n = min(kPbChunk, rb - c * kPbChunk);
    uint8_t* o = out + offsets[tid];
    if (pb_uniform(p, n)) {
        o[0] = (uint8_t)(1 - n);            // -(n - 1) as a signed byte
        o[1] = p[0];
    } else {
        o[0] = (uint8_t)(n - 1);
        for (int i = 0; i < n; ++i) o[1 + i] = p[i];
    }
}

}  // namespace oct
