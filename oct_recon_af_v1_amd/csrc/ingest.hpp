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
