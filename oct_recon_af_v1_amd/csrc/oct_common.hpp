// Shared device helpers for the OCT reconstruction kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace oct {

struct cf {
    float x, y;
};

__device__ __forceinline__ cf cadd(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cf cmul(cf a, cf b) {
    return {__builtin_fmaf(a.x, b.x, -a.y * b.y), __builtin_fmaf(a.x, b.y, a.y * b.x)};
}

// ---------------------------------------------------------------- compile-time trig
// Taylor series in double on [-pi/4, pi/4] after octant reduction; exact enough that the
// float twiddle constants are the correctly rounded values of cos/sin(2 pi k / n).
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ct_sin_small(double x) {
    double term = x, sum = x;
    for (int i = 1; i < 16; ++i) {
        term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double ct_cos_small(double x) {
    double term = 1.0, sum = 1.0;
    for (int i = 1; i < 16; ++i) {
        term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
        sum += term;
    }
    return sum;
}
// cos and sin of 2*pi*k/n for integers k, n (n a power of two >= 8)
constexpr double ct_cos2pi(long k, long n) {
    k %= n;
    if (k < 0) k += n;
    // reduce to octant: angle = 2 pi k / n
    long o = (8 * k) / n;          // octant 0..7
    double r = 2.0 * kPi * (double)(8 * k - o * n) / (8.0 * (double)n);   // in [0, pi/4)
    switch (o) {
        case 0: return ct_cos_small(r);
        case 1: return ct_sin_small(kPi / 4 - r);
        case 2: return -ct_sin_small(r);
        case 3: return -ct_cos_small(kPi / 4 - r);
        case 4: return -ct_cos_small(r);
        case 5: return -ct_sin_small(kPi / 4 - r);
        case 6: return ct_sin_small(r);
        default: return ct_cos_small(kPi / 4 - r);
    }
}
constexpr double ct_sin2pi(long k, long n) { return ct_cos2pi(k - n / 4, n); }

// multiply by exp(+2 pi i K / NN) with K, NN compile-time
template <long K, long NN>
__device__ __forceinline__ cf twiddle(cf a) {
    constexpr long k = ((K % NN) + NN) % NN;
    if constexpr (k == 0) {
        return a;
    } else if constexpr (4 * k == NN) {
        return {-a.y, a.x};
    } else if constexpr (2 * k == NN) {
        return {-a.x, -a.y};
    } else if constexpr (4 * k == 3 * NN) {
        return {a.y, -a.x};
    } else if constexpr (8 * k == NN || 8 * k == 3 * NN || 8 * k == 5 * NN || 8 * k == 7 * NN) {
        // odd multiples of pi/4: c = +-sqrt(1/2), s = +-c -> one add and one multiply each
        constexpr float c = (float)ct_cos2pi(k, NN);
        constexpr float sc = (ct_sin2pi(k, NN) > 0) == (ct_cos2pi(k, NN) > 0) ? 1.0f : -1.0f;
        return {c * (a.x - sc * a.y), c * (sc * a.x + a.y)};
    } else {
        constexpr float c = (float)ct_cos2pi(k, NN);
        constexpr float s = (float)ct_sin2pi(k, NN);
        return {__builtin_fmaf(a.x, c, -a.y * s), __builtin_fmaf(a.x, s, a.y * c)};
    }
}

// ---------------------------------------------------------------- pair order of 2048-point tables
// The transform's wave owns elements i = lane + 64 q (q < 32) of a 2048-point complex table (the
// window, the twiddles).  In the pair order a lane's elements q and q + 16 are adjacent, so one
// 16-byte load gives it both: element i sits at float2 position
//   2 ((q mod 16) 64 + lane) + q / 16,  q = i / 64, lane = i mod 64.
__host__ __device__ constexpr int pair_pos(int i) { return 2 * (((i >> 6) & 15) * 64 + (i & 63)) + (i >> 10); }

// ---------------------------------------------------------------- compile-time loops
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

constexpr int ct_log2(int n) { return n <= 1 ? 0 : 1 + ct_log2(n / 2); }
constexpr int bitrev(int v, int bits) {
    int r = 0;
    for (int i = 0; i < bits; ++i) r |= ((v >> i) & 1) << (bits - 1 - i);
    return r;
}

// In-register radix-2 decimation-in-frequency DFT of size NP with kernel exp(+2 pi i/NP)
// (unnormalised inverse DFT).  Input natural order; output Y[k] in v[bitrev(k)].
template <int NP>
__device__ __forceinline__ void dft_inv_dif(cf (&v)[NP]) {
    constexpr int L = ct_log2(NP);
    static_for<0, L>([&](auto s) {
        constexpr int len = NP >> decltype(s)::value;
        constexpr int half = len / 2;
        static_for<0, NP / len>([&](auto blk) {
            static_for<0, half>([&](auto j) {
                constexpr int a = decltype(blk)::value * len + decltype(j)::value;
                constexpr int b = a + half;
                cf x = v[a], y = v[b];
                v[a] = cadd(x, y);
                v[b] = twiddle<decltype(j)::value, len>(csub(x, y));
            });
        });
    });
}

// Radix-2 decimation-in-time butterfly with the compile-time twiddle W = exp(+2 pi i J / LEN),
// J < LEN / 2:  x <- x + W y,  y <- x - W y.  A non-trivial twiddle takes 6 instructions instead
// of 8: x + W y is formed with fused multiply-adds and x - W y as 2 x - (x + W y).
template <long J, long LEN>
__device__ __forceinline__ void bfly_dit(cf& x, cf& y) {
    constexpr long k = J % LEN;
    if constexpr (k == 0) {
        const cf a{x.x + y.x, x.y + y.y}, b{x.x - y.x, x.y - y.y};
        x = a;
        y = b;
    } else if constexpr (4 * k == LEN) {                        // W = i
        const cf a{x.x - y.y, x.y + y.x}, b{x.x + y.y, x.y - y.x};
        x = a;
        y = b;
    } else if constexpr (8 * k == LEN || 8 * k == 3 * LEN) {    // odd multiple of pi / 4
        constexpr float c = (float)ct_cos2pi(k, LEN);
        constexpr float sc = (ct_sin2pi(k, LEN) > 0) == (ct_cos2pi(k, LEN) > 0) ? 1.0f : -1.0f;
        const float t1 = y.x - sc * y.y, t2 = sc * y.x + y.y;   // W y = c (t1, t2)
        const cf a{__builtin_fmaf(c, t1, x.x), __builtin_fmaf(c, t2, x.y)};
        y = cf{__builtin_fmaf(2.0f, x.x, -a.x), __builtin_fmaf(2.0f, x.y, -a.y)};
        x = a;
    } else {
        constexpr float c = (float)ct_cos2pi(k, LEN);
        constexpr float s = (float)ct_sin2pi(k, LEN);
        const cf a{__builtin_fmaf(c, y.x, __builtin_fmaf(-s, y.y, x.x)), __builtin_fmaf(c, y.y, __builtin_fmaf(s, y.x, x.y))};
        y = cf{__builtin_fmaf(2.0f, x.x, -a.x), __builtin_fmaf(2.0f, x.y, -a.y)};
        x = a;
    }
}

// Radix-2 decimation-in-time DFT of size NP with kernel exp(+2 pi i / NP), stages S0 .. log2 NP - 1
// (the caller has done the stages before S0).  Input in bit-reversed order (v[i] = x[bitrev(i)]),
// output natural (v[k] = Y[k]).
template <int NP, int S0>
__device__ __forceinline__ void dft_inv_dit(cf (&v)[NP]) {
    constexpr int L = ct_log2(NP);
    static_for<S0, L>([&](auto s) {
        constexpr int len = 2 << decltype(s)::value;
        constexpr int half = len / 2;
        static_for<0, NP / len>([&](auto blk) {
            static_for<0, half>([&](auto j) {
                constexpr int a = decltype(blk)::value * len + decltype(j)::value;
                bfly_dit<decltype(j)::value, len>(v[a], v[a + half]);
            });
        });
    });
}

// Wave-local LDS ordering point: orders this wave's LDS accesses before/after, without a
// workgroup barrier (all lanes of a wave execute in lockstep).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// value of the neighbouring lane (lane ^ 1) through a DPP quad permutation [1,0,3,2]
__device__ __forceinline__ float swap_pair(float v) {
    int i = __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false);
    return __builtin_bit_cast(float, i);
}

// Cross-lane moves without LDS addresses: DPP (row_ror / quad_perm controls) and ds_swizzle
// (bit-mask mode, xor within 32-lane halves).
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int XOR>
__device__ __forceinline__ float swizzle_xor_f(float v) {
    static_assert(XOR > 0 && XOR < 32, "ds_swizzle bit-mask mode works within 32 lanes");
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), 0x1F | (XOR << 10)));
}

// Sum over the wave by the xor butterfly s += s[lane ^ m], m = 32, 16, ..., 1: every lane ends
// with the same bits as the __shfl_xor loop (same pairs in the same order; addition commutes),
// without its index arithmetic.  x32 = (lane ^ 32) * 4, the ds_bpermute address of the first step.
__device__ __forceinline__ float wave_sum_xor(float s, int x32) {
    s += __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(x32, __builtin_bit_cast(int, s)));
    s += swizzle_xor_f<16>(s);
    s += dpp_f<0x128>(s);          // row_ror:8 = lane ^ 8 within a row of 16
    s += swizzle_xor_f<4>(s);
    s += dpp_f<0x4E>(s);           // quad_perm [2,3,0,1] = lane ^ 2
    s += dpp_f<0xB1>(s);           // quad_perm [1,0,3,2] = lane ^ 1
    return s;
}

}  // namespace oct
