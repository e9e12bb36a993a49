// liboct_hip: C ABI over the gfx950 OCT reconstruction kernels (declarations: include/oct_hip.h).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>

#include "../../include/oct_hip.h"
#include "spectral.hpp"
#include "stitch.hpp"
#include "opform.hpp"
#include "estimate.hpp"
#include "cnn.hpp"
#include "ingest.hpp"

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <mutex>
#include <vector>

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return fail(OCT_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

struct oct_plan {
    int device;
    int n;
    int taps;
    int cus;
    int* band_start;
    float* band_w;
    float* win;   // window / N, in the pair order (oct::pair_pos)
    oct::cf* c64;   // exp(+2 pi i j / 64), j = 0..64 (oct::stage2_table_kernel)
    const float* win_override;   // caller's device window / N (dispersion search), or null
    long long win_fstride;       // floats between per-file override windows (0: one window)
    long long win_rstride;       // floats between the override windows of consecutive batch repeats
    float2* tw;   // (32, 64) transform twiddles W^(n2 k1), element k1 * 64 + n2 in the pair order
    int waves;    // workgroup shape of the fused kernel (see oct_plan_create)
};

namespace {

template <typename RawT, int TP, int W>
int launch_spectral_w(const oct_plan* p, const oct::SpectralArgs& a, int kind, hipStream_t s) {
    const long long ngroups = (a.M + W - 1) / W;
    const int per_cu = 8 / W;
    const int grid = (int)std::min<long long>(ngroups, (long long)p->cus * per_cu);
    if (grid <= 0) return OCT_OK;
    switch (kind) {
        case OCT_OUT_MAG:
            hipLaunchKernelGGL((oct::spectral_kernel<RawT, TP, oct::KIND_MAG, W>), dim3(grid), dim3(64 * W), 0, s, a);
            break;
        case OCT_OUT_COMPLEX:
            hipLaunchKernelGGL((oct::spectral_kernel<RawT, TP, oct::KIND_CPLX, W>), dim3(grid), dim3(64 * W), 0, s, a);
            break;
        default:
            hipLaunchKernelGGL((oct::spectral_kernel<RawT, TP, oct::KIND_EQ, W>), dim3(grid), dim3(64 * W), 0, s, a);
            break;
    }
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

template <typename RawT, int TP, int W>
int launch_spectral16(const oct_plan* p, const oct::SpectralArgs& a, int kind, hipStream_t s) {
    // sixteen rounds of resident workgroups: workgroups that finish early take the remaining
    // groups' share, so a launch's tail is shorter (1.5-2.5 % per tile against one persistent
    // round; a workgroup per group is no better than four rounds)
    const long long ngroups = (a.M + W - 1) / W;
    const int grid = (int)std::min<long long>(ngroups, (long long)p->cus * (16 / W) * 16);
    if (grid <= 0) return OCT_OK;
    const dim3 threads(64 * W);
    const int st2 = a.k2n == 0 ? 0 : (a.rows <= 32 + oct::kStage2Extra ? 1 : 2);
    switch (kind * 3 + st2) {
        case OCT_OUT_MAG * 3 + 0:
            hipLaunchKernelGGL((oct::spectral16_kernel<RawT, TP, oct::KIND_MAG, 0, W>), dim3(grid), threads, 0, s, a);
            break;
        case OCT_OUT_MAG * 3 + 1:
            hipLaunchKernelGGL((oct::spectral16_kernel<RawT, TP, oct::KIND_MAG, 1, W>), dim3(grid), threads, 0, s, a);
            break;
        case OCT_OUT_MAG * 3 + 2:
            hipLaunchKernelGGL((oct::spectral16_kernel<RawT, TP, oct::KIND_MAG, 2, W>), dim3(grid), threads, 0, s, a);
            break;
        case OCT_OUT_COMPLEX * 3 + 0:
            hipLaunchKernelGGL((oct::spectral16_kernel<RawT, TP, oct::KIND_CPLX, 0, W>), dim3(grid), threads, 0, s, a);
            break;
        case OCT_OUT_COMPLEX * 3 + 1:
            hipLaunchKernelGGL((oct::spectral16_kernel<RawT, TP, oct::KIND_CPLX, 1, W>), dim3(grid), threads, 0, s, a);
            break;
        case OCT_OUT_COMPLEX * 3 + 2:
            hipLaunchKernelGGL((oct::spectral16_kernel<RawT, TP, oct::KIND_CPLX, 2, W>), dim3(grid), threads, 0, s, a);
            break;
        default:
            hipLaunchKernelGGL((oct::spectral16_kernel<RawT, TP, oct::KIND_EQ, 0, W>), dim3(grid), threads, 0, s, a);
            break;
    }
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

template <typename RawT>
int launch_spectral(const oct_plan* p, const oct::SpectralArgs& a, int kind, hipStream_t s) {
    if (p->waves == 16) {
        if (p->taps == 12) return launch_spectral16<RawT, 12, 16>(p, a, kind, s);
        return launch_spectral16<RawT, 16, 16>(p, a, kind, s);
    }
    if (p->waves == 28) {            // two 8-wave workgroups per CU
        if (p->taps == 12) return launch_spectral16<RawT, 12, 8>(p, a, kind, s);
        return launch_spectral16<RawT, 16, 8>(p, a, kind, s);
    }
    if (p->waves == 82) {            // eight 2-wave workgroups per CU
        if (p->taps == 12) return launch_spectral16<RawT, 12, 2>(p, a, kind, s);
        return launch_spectral16<RawT, 16, 2>(p, a, kind, s);
    }
    if (p->waves == 44) {            // four 4-wave workgroups per CU
        if (p->taps == 12) return launch_spectral16<RawT, 12, 4>(p, a, kind, s);
        return launch_spectral16<RawT, 16, 4>(p, a, kind, s);
    }
    if (p->waves == 4) {
        if (p->taps == 12) return launch_spectral_w<RawT, 12, 4>(p, a, kind, s);
        return launch_spectral_w<RawT, 16, 4>(p, a, kind, s);
    }
    if (p->taps == 12) return launch_spectral_w<RawT, 12, 8>(p, a, kind, s);
    return launch_spectral_w<RawT, 16, 8>(p, a, kind, s);
}

template <typename RawT, int RT>
void launch_opgemm_rf_rt(const oct::OpGemmArgs& a, int grid, hipStream_t s) {
    // background rows through LDS when a workgroup's A-scans lie in at most two files
    if (a.nX >= oct::kOpCols && a.N == oct::kOpBgN)
        hipLaunchKernelGGL((oct::opgemm_rf_kernel<RawT, true, RT>), dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((oct::opgemm_rf_kernel<RawT, false, RT>), dim3(grid), dim3(256), 0, s, a);
}

template <typename RawT, int RT>
void launch_opgemm_lt_rt(const oct::OpGemmArgs& a, int grid, hipStream_t s) {
    if (a.nX >= oct::kOpCols && a.N == oct::kOpBgN)
        hipLaunchKernelGGL((oct::opgemm_lt_kernel<RawT, true, RT>), dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((oct::opgemm_lt_kernel<RawT, false, RT>), dim3(grid), dim3(256), 0, s, a);
}

// 16-bit samples: the coalesced-load kernel
template <typename RawT>
void launch_opgemm_lt(const oct::OpGemmArgs& a, int rt, int grid, hipStream_t s) {
    switch (rt) {
        case 1: launch_opgemm_lt_rt<RawT, 1>(a, grid, s); break;
        case 2: launch_opgemm_lt_rt<RawT, 2>(a, grid, s); break;
        case 3: launch_opgemm_lt_rt<RawT, 3>(a, grid, s); break;
        default: launch_opgemm_lt_rt<RawT, 4>(a, grid, s); break;
    }
}

template <typename RawT>
void launch_opgemm_rf(const oct::OpGemmArgs& a, int rt, int grid, hipStream_t s) {
    switch (rt) {
        case 1: launch_opgemm_rf_rt<RawT, 1>(a, grid, s); break;
        case 2: launch_opgemm_rf_rt<RawT, 2>(a, grid, s); break;
        case 3: launch_opgemm_rf_rt<RawT, 3>(a, grid, s); break;
        default: launch_opgemm_rf_rt<RawT, 4>(a, grid, s); break;
    }
}

template <typename RawT>
int launch_background(const void* raw, float* bg, int B, int lines, int N, int apod, int mode, int group,
                      hipStream_t s) {
    const int threads = 256;
    constexpr int VEC = 16 / sizeof(RawT);
    if (mode == 0 && N % VEC == 0 && (reinterpret_cast<uintptr_t>(bg) & 15) == 0) {
        const long long n = (long long)B * (N / VEC);
        hipLaunchKernelGGL(oct::background_lines_kernel<RawT>, dim3((unsigned)((n + threads - 1) / threads)),
                           dim3(threads), 0, s, reinterpret_cast<const RawT*>(raw), bg, B, lines, N, apod);
    } else {
        const long long n = (long long)B * N;
        hipLaunchKernelGGL(oct::background_kernel<RawT>, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads),
                           0, s, reinterpret_cast<const RawT*>(raw), bg, B, lines, N, apod, mode, group);
    }
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

template <int TX, int TY>
void launch_conv_mfma(int ct, const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y, int n,
                             int ci, int h, int wd, int co, hipStream_t s) {
    const int Ho = h / 2, Wo = wd / 2;
    const dim3 grid((unsigned)((Wo + TX - 1) / TX), (unsigned)((Ho + TY - 1) / TY), (unsigned)(n * (co / (16 * ct))));
    const dim3 threads(oct::kCnnThreads);
    if (ct == 4)
        hipLaunchKernelGGL((oct::conv_fwd_mfma_kernel<TX, TY, 4>), grid, threads, 0, s, x, w, bias, y, n, ci, h, wd, co);
    else if (ct == 2)
        hipLaunchKernelGGL((oct::conv_fwd_mfma_kernel<TX, TY, 2>), grid, threads, 0, s, x, w, bias, y, n, ci, h, wd, co);
    else
        hipLaunchKernelGGL((oct::conv_fwd_mfma_kernel<TX, TY, 1>), grid, threads, 0, s, x, w, bias, y, n, ci, h, wd, co);
}

}  // namespace

extern "C" {

const char* oct_last_error(void) { return g_err; }

int oct_plan_create(oct_plan** plan, int device, int n_lambda, int taps, const int32_t* band_start,
                    const float* band_w, const float* window) {
    if (!plan || !band_start || !band_w || !window) return fail(OCT_EINVAL, "null argument");
    *plan = nullptr;
    if (n_lambda != oct::kN) return fail(OCT_EUNSUPPORTED, "n_lambda %d unsupported (2048 only)", n_lambda);
    if (taps != 12 && taps != 16) return fail(OCT_EUNSUPPORTED, "tap window %d unsupported (12 or 16)", taps);
    for (int i = 0; i < n_lambda; ++i)
        if (band_start[i] < 0 || band_start[i] > n_lambda - taps)
            return fail(OCT_EINVAL, "band_start[%d]=%d outside [0, %d]", i, band_start[i], n_lambda - taps);
    int prev = 0;
    HIP_TRY(hipGetDevice(&prev));
    HIP_TRY(hipSetDevice(device));
    oct_plan* p = new oct_plan();
    p->device = device;
    p->n = n_lambda;
    p->taps = taps;
    int cus = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    p->cus = (e == hipSuccess && cus > 0) ? cus : 256;
    // workgroup shape of the fused kernel (OCT_HIP_SPECTRAL_WAVES, for measurements): 44 = four
    // 4-wave workgroups per CU (default, fastest), 28 = two of 8, 82 = eight of 2, 16 = one of 16
    // (in-place regions); 8 and 4 = the earlier two-region forms
    const char* ev = std::getenv("OCT_HIP_SPECTRAL_WAVES");
    const int wv = ev ? std::atoi(ev) : 44;
    p->waves = (wv == 4 || wv == 8 || wv == 16 || wv == 28 || wv == 82) ? wv : 44;
    std::vector<float> twh(2 * 32 * 64);
    for (int k1 = 0; k1 < 32; ++k1)
        for (int n2 = 0; n2 < 64; ++n2) {
            const double ang = 2.0 * M_PI * (double)(n2 * k1) / (double)n_lambda;
            const int pp = oct::pair_pos(k1 * 64 + n2);
            twh[2 * pp] = (float)std::cos(ang);
            twh[2 * pp + 1] = (float)std::sin(ang);
        }
    // eight slots per lane: (W^(16 n2) in slot 0, else W^(s n2); W^((s + 8) n2)) as float4
    const size_t tw8_off = twh.size();
    twh.resize(tw8_off + 8 * 64 * 4);
    for (int sl = 0; sl < 8; ++sl)
        for (int n2 = 0; n2 < 64; ++n2) {
            const int ka = sl == 0 ? 16 : sl, kb = sl + 8;
            const double aa = 2.0 * M_PI * (double)(n2 * ka) / (double)n_lambda;
            const double ab = 2.0 * M_PI * (double)(n2 * kb) / (double)n_lambda;
            float* e = &twh[tw8_off + (size_t)(sl * 64 + n2) * 4];
            e[0] = (float)std::cos(aa);
            e[1] = (float)std::sin(aa);
            e[2] = (float)std::cos(ab);
            e[3] = (float)std::sin(ab);
        }
    float* win = new float[2 * n_lambda];
    for (int i = 0; i < n_lambda; ++i) {                // / N is exact (2^-11); stored in the pair order
        const int pp = oct::pair_pos(i);
        win[2 * pp] = window[2 * i] * (1.0f / (float)n_lambda);
        win[2 * pp + 1] = window[2 * i + 1] * (1.0f / (float)n_lambda);
    }
    int rc = OCT_OK;
    if (hipMalloc(&p->band_start, sizeof(int) * n_lambda) != hipSuccess ||
        hipMalloc(&p->band_w, sizeof(float) * n_lambda * taps) != hipSuccess ||
        hipMalloc(&p->win, sizeof(float) * 2 * n_lambda) != hipSuccess ||
        hipMalloc(&p->tw, sizeof(float) * twh.size()) != hipSuccess ||
        hipMemcpy(p->tw, twh.data(), sizeof(float) * twh.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->band_start, band_start, sizeof(int) * n_lambda, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->band_w, band_w, sizeof(float) * n_lambda * taps, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->win, win, sizeof(float) * 2 * n_lambda, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(OCT_EHIP, "plan upload failed: %s", hipGetErrorString(hipGetLastError()));
    }
    if (rc == OCT_OK) {
        hipError_t e2 = hipMalloc(&p->c64, sizeof(oct::cf) * 65);
        if (e2 == hipSuccess) {
            hipLaunchKernelGGL(oct::stage2_table_kernel, dim3(1), dim3(128), 0, nullptr, p->c64);
            e2 = hipGetLastError();
        }
        if (e2 == hipSuccess) e2 = hipStreamSynchronize(nullptr);
        if (e2 != hipSuccess) rc = fail(OCT_EHIP, "plan table build failed: %s", hipGetErrorString(e2));
    }
    delete[] win;
    hipSetDevice(prev);
    if (rc != OCT_OK) {
        oct_plan_destroy(p);
        return rc;
    }
    *plan = p;
    return OCT_OK;
}

int oct_plan_destroy(oct_plan* p) {
    if (!p) return OCT_OK;
    int prev = 0;
    hipGetDevice(&prev);
    hipSetDevice(p->device);
    if (p->band_start) hipFree(p->band_start);
    if (p->band_w) hipFree(p->band_w);
    if (p->win) hipFree(p->win);
    if (p->tw) hipFree(p->tw);
    if (p->c64) hipFree(p->c64);
    hipSetDevice(prev);
    delete p;
    return OCT_OK;
}

size_t oct_spectral_workspace_bytes(const oct_plan* p, int n_files) {
    if (!p || n_files <= 0) return 0;
    return sizeof(float) * (size_t)n_files * (size_t)p->n;
}

static int check_raw(const void* raw, int raw_dtype, int n_files, int lines, int apod_mode, int apod_size,
                     int apod_group, int* line0, int* nX) {
    if (!raw) return fail(OCT_EINVAL, "null raw pointer");
    if (n_files < 0 || lines <= 0) return fail(OCT_EINVAL, "bad raw shape (%d, %d)", n_files, lines);
    if (reinterpret_cast<uintptr_t>(raw) % 16) return fail(OCT_EINVAL, "raw pointer must be 16-byte aligned");
    if (raw_dtype < OCT_RAW_I16 || raw_dtype > OCT_RAW_F32) return fail(OCT_EINVAL, "raw dtype %d", raw_dtype);
    if (apod_mode == OCT_APOD_LINES) {
        if (apod_size <= 0 || apod_size >= lines)
            return fail(OCT_EINVAL, "apod_size %d must be in [1, %d)", apod_size, lines);
        *line0 = apod_size;
        *nX = lines - apod_size;
    } else if (apod_mode == OCT_APOD_FRAME_MEAN) {
        if (apod_group <= 0 || (n_files % apod_group))
            return fail(OCT_EINVAL, "batch of %d files is not a whole number of %d-file frames", n_files, apod_group);
        *line0 = 0;
        *nX = lines;
    } else {
        return fail(OCT_EINVAL, "apod mode %d", apod_mode);
    }
    return OCT_OK;
}

int oct_background(const oct_plan* p, const void* raw, int raw_dtype, int n_files, int lines, int apod_mode,
                   int apod_size, int apod_group, float* bg, void* stream) {
    if (n_files == 0) return OCT_OK;
    if (!p || !bg) return fail(OCT_EINVAL, "null argument");
    int line0, nX;
    int rc = check_raw(raw, raw_dtype, n_files, lines, apod_mode, apod_size, apod_group, &line0, &nX);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int mode = apod_mode == OCT_APOD_LINES ? 0 : 1;
    switch (raw_dtype) {
        case OCT_RAW_I16: return launch_background<int16_t>(raw, bg, n_files, lines, p->n, apod_size, mode, apod_group, s);
        case OCT_RAW_U16: return launch_background<uint16_t>(raw, bg, n_files, lines, p->n, apod_size, mode, apod_group, s);
        default: return launch_background<float>(raw, bg, n_files, lines, p->n, apod_size, mode, apod_group, s);
    }
}

static int spectral_impl(const oct_plan* p, const void* raw, int raw_dtype, int n_files, int lines, int apod_mode,
                         int apod_size, int n_repeat, int x0, int n_x, int out_kind, int row0, int rows, void* out,
                         long long out_stride, const float* bg, void* stream);

int oct_binned_lines(int n_scan_lines, int binning) {
    if (n_scan_lines <= 0 || binning < 1) return 0;
    const int s0 = std::max(1, binning / 2) - 1;
    return s0 >= n_scan_lines ? 0 : (n_scan_lines - s0 + binning - 1) / binning;
}

int oct_bin_lines(const void* raw, int raw_dtype, int n_files, int lines, int keep_lines, int skip_lines,
                  int binning, float* out, void* stream) {
    if (n_files <= 0) return OCT_OK;
    if (!raw || !out) return fail(OCT_EINVAL, "null argument");
    if (binning < 1 || keep_lines < 0 || skip_lines < keep_lines || skip_lines >= lines)
        return fail(OCT_EINVAL, "binning %d, keep %d, skip %d of %d lines", binning, keep_lines, skip_lines, lines);
    if (raw_dtype < OCT_RAW_I16 || raw_dtype > OCT_RAW_F32) return fail(OCT_EINVAL, "raw dtype %d", raw_dtype);
    const int n_scan = lines - skip_lines;
    const int n_kept = oct_binned_lines(n_scan, binning);
    const long long tot = (long long)n_files * (keep_lines + n_kept) * (long long)oct::kN;
    if ((tot + 255) / 256 > 0x7FFFFFFFLL) return fail(OCT_EINVAL, "batch too large");
    const unsigned grid = (unsigned)((tot + 255) / 256);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    switch (raw_dtype) {
        case OCT_RAW_I16:
            hipLaunchKernelGGL(oct::bin_lines_kernel<int16_t>, dim3(grid), dim3(256), 0, s, static_cast<const int16_t*>(raw),
                               out, n_files, lines, keep_lines, skip_lines, n_scan, binning, n_kept, oct::kN);
            break;
        case OCT_RAW_U16:
            hipLaunchKernelGGL(oct::bin_lines_kernel<uint16_t>, dim3(grid), dim3(256), 0, s, static_cast<const uint16_t*>(raw),
                               out, n_files, lines, keep_lines, skip_lines, n_scan, binning, n_kept, oct::kN);
            break;
        default:
            hipLaunchKernelGGL(oct::bin_lines_kernel<float>, dim3(grid), dim3(256), 0, s, static_cast<const float*>(raw),
                               out, n_files, lines, keep_lines, skip_lines, n_scan, binning, n_kept, oct::kN);
            break;
    }
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_srr_unpack(const uint16_t* in, int n_files, int file_lines, int n_lambda, int apod_start, int apod_end,
                   int scan_start, int scan_end, int16_t* out, void* stream) {
    if (n_files <= 0) return OCT_OK;
    if (!in || !out) return fail(OCT_EINVAL, "null argument");
    if (n_lambda <= 0 || (n_lambda % 8)) return fail(OCT_EINVAL, "n_lambda %d must be a positive multiple of 8", n_lambda);
    if (apod_start < 0 || apod_end <= apod_start || scan_start < 0 || scan_end <= scan_start ||
        apod_end > file_lines || scan_end > file_lines)
        return fail(OCT_EINVAL, "regions [%d, %d) / [%d, %d) outside %d A-lines", apod_start, apod_end, scan_start,
                    scan_end, file_lines);
    if (((uintptr_t)in | (uintptr_t)out) & 15) return fail(OCT_EINVAL, "buffers must be 16-byte aligned");
    const long long L = (long long)(apod_end - apod_start) + (scan_end - scan_start);
    const long long tot = (long long)n_files * L * (n_lambda / 8);
    if ((tot + 255) / 256 > 0x7FFFFFFFLL) return fail(OCT_EINVAL, "batch too large");
    hipLaunchKernelGGL(oct::srr_unpack_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), in, out, n_files, file_lines, n_lambda, apod_start,
                       apod_end, scan_start, scan_end);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_packbits_chunks(int row_elems) { return row_elems > 0 ? (2 * row_elems + oct::kPbChunk - 1) / oct::kPbChunk : 0; }

int oct_packbits_sizes(const uint16_t* codes, long long rows, int row_elems, int32_t* chunk_bytes, void* stream) {
    if (rows <= 0 || row_elems <= 0) return OCT_OK;
    if (!codes || !chunk_bytes) return fail(OCT_EINVAL, "null argument");
    const long long tot = rows * oct_packbits_chunks(row_elems);
    if ((tot + 255) / 256 > 0x7FFFFFFFLL) return fail(OCT_EINVAL, "too many rows");
    hipLaunchKernelGGL(oct::packbits_size_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), codes, rows, row_elems, chunk_bytes);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_packbits_write(const uint16_t* codes, long long rows, int row_elems, const int64_t* chunk_offsets,
                       uint8_t* out, void* stream) {
    if (rows <= 0 || row_elems <= 0) return OCT_OK;
    if (!codes || !chunk_offsets || !out) return fail(OCT_EINVAL, "null argument");
    const long long tot = rows * oct_packbits_chunks(row_elems);
    if ((tot + 255) / 256 > 0x7FFFFFFFLL) return fail(OCT_EINVAL, "too many rows");
    hipLaunchKernelGGL(oct::packbits_write_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), codes, rows, row_elems,
                       reinterpret_cast<const long long*>(chunk_offsets), out);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_average_repeats(const float* mag, float* out, int n_frames, int bscan_avg, int size_x, int ascan_avg, int rows,
                        long long in_stride, long long out_stride, void* stream) {
    if (n_frames <= 0 || size_x <= 0 || rows <= 0) return OCT_OK;
    if (!mag || !out) return fail(OCT_EINVAL, "null argument");
    if (bscan_avg < 1 || ascan_avg < 1 || in_stride < rows || out_stride < rows)
        return fail(OCT_EINVAL, "bad repeat shape (%d, %d, strides %lld %lld, rows %d)", bscan_avg, ascan_avg, in_stride,
                    out_stride, rows);
    const long long tot = (long long)n_frames * size_x * rows;
    if ((tot + 255) / 256 > 0x7FFFFFFFLL) return fail(OCT_EINVAL, "batch too large");
    hipLaunchKernelGGL(oct::average_repeats_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), mag, out, n_frames, bscan_avg, size_x, ascan_avg, rows,
                       in_stride, out_stride);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_spectral(const oct_plan* p, const void* raw, int raw_dtype, int n_files, int lines, int apod_mode,
                 int apod_size, int out_kind, int row0, int rows, void* out, long long out_stride,
                 const float* bg, void* stream) {
    return spectral_impl(p, raw, raw_dtype, n_files, lines, apod_mode, apod_size, 1, 0, -1, out_kind, row0, rows, out,
                         out_stride, bg, stream);
}

int oct_spectral_repeat(const oct_plan* p, const void* raw, int raw_dtype, int n_files, int lines, int apod_mode,
                        int apod_size, int n_repeat, int x0, int n_x, int out_kind, int row0, int rows, void* out,
                        long long out_stride, const float* bg, void* stream) {
    if (n_repeat < 1) return fail(OCT_EINVAL, "n_repeat %d", n_repeat);
    return spectral_impl(p, raw, raw_dtype, n_files, lines, apod_mode, apod_size, n_repeat, x0, n_x, out_kind, row0,
                         rows, out, out_stride, bg, stream);
}

static int spectral_impl(const oct_plan* p, const void* raw, int raw_dtype, int n_files, int lines, int apod_mode,
                         int apod_size, int n_repeat, int x0, int n_x, int out_kind, int row0, int rows, void* out,
                         long long out_stride, const float* bg, void* stream) {
    if (n_files == 0) return OCT_OK;
    if (!p || !out || !bg) return fail(OCT_EINVAL, "null argument");
    int line0, nX;
    int rc = check_raw(raw, raw_dtype, n_files, lines, apod_mode, apod_size, 1, &line0, &nX);
    if (rc) return rc;
    if (n_x >= 0) {                         // a window of the scan lines
        if (x0 < 0 || n_x < 1 || x0 + n_x > nX) return fail(OCT_EINVAL, "scan lines [%d, %d) outside [0, %d)", x0, x0 + n_x, nX);
        line0 += x0;
        nX = n_x;
    }
    if (out_kind < OCT_OUT_MAG || out_kind > OCT_OUT_LINEARISED) return fail(OCT_EINVAL, "out kind %d", out_kind);
    const int nz = p->n / 2;
    if (out_kind != OCT_OUT_LINEARISED) {
        if (row0 < 0 || rows <= 0 || row0 + rows > nz)
            return fail(OCT_EINVAL, "rows [%d, %d) outside [0, %d)", row0, row0 + rows, nz);
        if (out_stride < rows) return fail(OCT_EINVAL, "out_stride %lld < rows %d", out_stride, rows);
    } else {
        row0 = 0;
        rows = p->n;
        if (out_stride < p->n) return fail(OCT_EINVAL, "out_stride %lld < N", out_stride);
    }
    oct::SpectralArgs a;
    a.raw = raw;
    a.bg = bg;
    a.band_start = p->band_start;
    a.band_w = p->band_w;
    a.win = p->win_override ? p->win_override : p->win;
    a.win_fstride = p->win_override ? p->win_fstride : 0;
    a.win_rstride = p->win_override ? p->win_rstride : 0;
    a.raw_files = n_files;
    a.tw = p->tw;
    a.tw8 = reinterpret_cast<const float4*>(reinterpret_cast<const float*>(p->tw) + 2 * 32 * 64);
    a.c64 = p->c64;
    a.out = out;
    a.M = (long long)n_repeat * n_files * nX;
    if (a.M >= (1LL << 31)) return fail(OCT_EINVAL, "%lld A-scans in one launch (at most 2^31 - 1)", a.M);
    a.lines = lines;
    a.line0 = line0;
    a.nX = nX;
    a.row0 = row0;
    a.rows = rows;
    {
        // rows inside at most two blocks of 32 bins k = k1 + 32 k2: the kernel forms only those
        // k2 of its second transform stage (oct::spectral_kernel, pruned stage)
        const int k2lo = row0 / 32, k2hi = (row0 + rows - 1) / 32;
        a.k2lo = k2lo;
        a.k2n = (out_kind != OCT_OUT_LINEARISED && k2hi - k2lo < 2) ? k2hi - k2lo + 1 : 0;
    }
    a.out_stride = out_stride;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    switch (raw_dtype) {
        case OCT_RAW_I16: return launch_spectral<int16_t>(p, a, out_kind, s);
        case OCT_RAW_U16: return launch_spectral<uint16_t>(p, a, out_kind, s);
        default: return launch_spectral<float>(p, a, out_kind, s);
    }
}

int oct_stitch(const oct_stitch_desc* d, void* stream) {
    if (!d || !d->mag || !d->frame_idx || !d->fz || !d->z_rows || !d->z_w || !d->x_cols || !d->x_w || !d->num ||
        !d->den)
        return fail(OCT_EINVAL, "null argument");
    if (d->shift && (!d->valid_lo || !d->valid_hi)) return fail(OCT_EINVAL, "shift without valid bounds");
    if (d->B <= 0 || d->n_out_z <= 0 || d->C <= 0) return OCT_OK;
    if (d->n_x <= 0 || d->n_z <= 0 || d->ldx < d->out_c0 + d->C) return fail(OCT_EINVAL, "bad stitch shape");
    if (d->mag_row0 != 0 || d->mag_rows != d->n_z)
        if (d->mag_row0 < 0 || d->mag_rows <= 0) return fail(OCT_EINVAL, "bad magnitude row window");
    oct::StitchArgs a;
    a.mag = d->mag;
    a.mstride = d->mag_stride;
    a.mrow0 = d->mag_row0;
    a.mrows = d->mag_rows;
    a.nX = d->n_x;
    a.nZ = d->n_z;
    a.frame_idx = d->frame_idx;
    a.shift = d->shift;
    a.vlo = d->valid_lo;
    a.vhi = d->valid_hi;
    a.fz = d->fz;
    a.zr = d->z_rows;
    a.zw = d->z_w;
    a.xr = d->x_cols;
    a.xw = d->x_w;
    a.num = d->num;
    a.den = d->den;
    a.B = d->B;
    a.nOutZ = d->n_out_z;
    a.C = d->C;
    a.ldx = d->ldx;
    a.out_b0 = d->out_b0;
    a.out_c0 = d->out_c0;
    a.accumulate = d->accumulate;
    dim3 grid((d->C + 255) / 256, d->n_out_z, d->B);
    hipLaunchKernelGGL(oct::stitch_kernel, grid, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), a);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_stitch_row_blocks(int n_out_z, int n_out_x) {
    if (n_out_z <= 0 || n_out_x <= 0) return 0;
    const int gx = (n_out_x + oct::kRowC - 1) / oct::kRowC;
    const int gz = (n_out_z + oct::kRowZ - 1) / oct::kRowZ;
    return gx * gz;
}

static_assert(sizeof(oct::RowTile) == sizeof(oct_row_tile) && offsetof(oct::RowTile, fz) == offsetof(oct_row_tile, fz) &&
                  offsetof(oct::RowTile, mstride) == offsetof(oct_row_tile, mag_stride) &&
                  offsetof(oct::RowTile, zr) == offsetof(oct_row_tile, z_rows),
              "the device tile table is read with the oct_row_tile layout");

static int stitch_row_impl(const oct_row_tile* tiles, const oct_row_tile* dev_tiles, int n_tiles, int n_frames, int n_x,
                           int n_z, const int32_t* frame_idx, const int32_t* shift, const int32_t* valid_lo,
                           const int32_t* valid_hi, int n_out_z, int n_out_x, float thr, float* db, int plane0,
                           int* partial, void* stream) {
    if (n_frames <= 0 || n_out_z <= 0 || n_out_x <= 0) return OCT_OK;
    if (!tiles || !frame_idx || !db || !partial) return fail(OCT_EINVAL, "null argument");
    if (n_tiles < 0 || (!dev_tiles && n_tiles > oct::kRowMaxTiles))
        return fail(OCT_EINVAL, "%d tiles (at most %d per row in the argument block)", n_tiles, oct::kRowMaxTiles);
    if (dev_tiles && ((uintptr_t)dev_tiles & 7)) return fail(OCT_EINVAL, "tile table must be 8-byte aligned");
    if (shift && (!valid_lo || !valid_hi)) return fail(OCT_EINVAL, "shift without valid bounds");
    if (n_x <= 0 || n_z <= 0 || plane0 < 0) return fail(OCT_EINVAL, "bad row shape");
    oct::RowArgs a;
    a.tt = reinterpret_cast<const oct::RowTile*>(dev_tiles);
    for (int i = 0; i < n_tiles; ++i) {
        const oct_row_tile& t = tiles[i];
        if (!t.mag || !t.z_rows || !t.z_w || !t.x_cols || !t.x_w || !t.fz) return fail(OCT_EINVAL, "tile %d: null pointer", i);
        // tap pairs are read as 8-byte vectors
        if (((uintptr_t)t.z_rows | (uintptr_t)t.z_w | (uintptr_t)t.x_cols | (uintptr_t)t.x_w) & 7)
            return fail(OCT_EINVAL, "tile %d: tap arrays must be 8-byte aligned", i);
        if (t.c0 < 0 || t.C <= 0 || t.c0 + t.C > n_out_x) return fail(OCT_EINVAL, "tile %d: columns outside the output", i);
        if (i >= oct::kRowMaxTiles) continue;             // the kernel reads the device table
        a.t[i].mag = t.mag;
        a.t[i].mstride = t.mag_stride;
        a.t[i].mrow0 = t.mag_row0;
        a.t[i].c0 = t.c0;
        a.t[i].C = t.C;
        a.t[i].zr = t.z_rows;
        a.t[i].zw = t.z_w;
        a.t[i].xr = t.x_cols;
        a.t[i].xw = t.x_w;
        a.t[i].fz = t.fz;
    }
    a.ntiles = n_tiles;
    a.nX = n_x;
    a.nZ = n_z;
    a.frame_idx = frame_idx;
    a.shift = shift;
    a.vlo = valid_lo;
    a.vhi = valid_hi;
    a.nOutZ = n_out_z;
    a.nOutX = n_out_x;
    a.thr = thr;
    a.db = db;
    a.plane0 = plane0;
    a.partial = partial;
    a.blocks_per_plane = oct_stitch_row_blocks(n_out_z, n_out_x);
    if (n_frames > 65535) return fail(OCT_EINVAL, "%d planes in one call (at most 65535)", n_frames);
    a.nplanes = n_frames;
    const long long total = (long long)a.blocks_per_plane * ((n_frames + oct::kRowPlanes - 1) / oct::kRowPlanes);
    if (total > (1LL << 30)) return fail(OCT_EINVAL, "too many planes in one call");
    dim3 grid((unsigned)((total + 7) / 8 * 8));            // whole multiple of the 8 XCDs
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const dim3 blk(64 * oct::kRowWaves);
    if (dev_tiles) {
        if (shift) hipLaunchKernelGGL((oct::stitch_row_kernel<true, true>), grid, blk, 0, s, a);
        else hipLaunchKernelGGL((oct::stitch_row_kernel<false, true>), grid, blk, 0, s, a);
    } else {
        if (shift) hipLaunchKernelGGL((oct::stitch_row_kernel<true, false>), grid, blk, 0, s, a);
        else hipLaunchKernelGGL((oct::stitch_row_kernel<false, false>), grid, blk, 0, s, a);
    }
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_stitch_row(const oct_row_tile* tiles, int n_tiles, int n_frames, int n_x, int n_z, const int32_t* frame_idx,
                   const int32_t* shift, const int32_t* valid_lo, const int32_t* valid_hi, int n_out_z, int n_out_x,
                   float thr, float* db, int plane0, int* partial, void* stream) {
    return stitch_row_impl(tiles, nullptr, n_tiles, n_frames, n_x, n_z, frame_idx, shift, valid_lo, valid_hi, n_out_z,
                           n_out_x, thr, db, plane0, partial, stream);
}

int oct_stitch_row_table(const oct_row_tile* tiles, const oct_row_tile* dev_tiles, int n_tiles, int n_frames, int n_x,
                         int n_z, const int32_t* frame_idx, const int32_t* shift, const int32_t* valid_lo,
                         const int32_t* valid_hi, int n_out_z, int n_out_x, float thr, float* db, int plane0,
                         int* partial, void* stream) {
    if (!dev_tiles) return fail(OCT_EINVAL, "null argument");
    return stitch_row_impl(tiles, dev_tiles, n_tiles, n_frames, n_x, n_z, frame_idx, shift, valid_lo, valid_hi, n_out_z,
                           n_out_x, thr, db, plane0, partial, stream);
}

// plane clims (a wave per plane), then the global clim from them
static void launch_clim(const int* partial, int nblk, int n_planes, float* plane_clim, float* clim, hipStream_t s) {
    hipLaunchKernelGGL(oct::clim_planes_kernel, dim3((unsigned)((n_planes + 3) / 4)), dim3(256), 0, s, partial, nblk,
                       n_planes, plane_clim);
    hipLaunchKernelGGL(oct::clim_global_kernel, dim3(1), dim3(oct::kFinThreads), 0, s, plane_clim, n_planes, clim);
}

int oct_volume_clim(const int* partial, int n_planes, int blocks_per_plane, float* plane_clim, float* clim,
                    void* stream) {
    if (!partial || !plane_clim || !clim) return fail(OCT_EINVAL, "null argument");
    if (n_planes <= 0 || blocks_per_plane <= 0) return fail(OCT_EINVAL, "empty volume");
    if (reinterpret_cast<uintptr_t>(partial) & 7) return fail(OCT_EINVAL, "partials must be 8-byte aligned");
    launch_clim(partial, blocks_per_plane, n_planes, plane_clim, clim, reinterpret_cast<hipStream_t>(stream));
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

static int finalize_blocks(long long plane_elems) {
    long long b = (plane_elems + oct::kFinThreads * 8 - 1) / (oct::kFinThreads * 8);
    return (int)std::max<long long>(1, std::min<long long>(b, 256));
}

size_t oct_finalize_workspace_bytes(int n_planes, long long plane_elems) {
    if (n_planes <= 0 || plane_elems <= 0) return 0;
    return sizeof(int) * 2 * (size_t)n_planes * (size_t)finalize_blocks(plane_elems);
}

int oct_finalize(const float* num, const float* den, float* db, int n_planes, long long plane_elems, float thr,
                 float* plane_clim, float* clim, void* workspace, size_t workspace_bytes, void* stream) {
    if (!num || !db || !plane_clim || !clim) return fail(OCT_EINVAL, "null argument");
    if (n_planes <= 0 || plane_elems <= 0) return fail(OCT_EINVAL, "empty volume");
    const size_t need = oct_finalize_workspace_bytes(n_planes, plane_elems);
    if (!workspace || workspace_bytes < need)
        return fail(OCT_EINVAL, "workspace of %zu bytes needed, %zu given", need, workspace_bytes);
    if (reinterpret_cast<uintptr_t>(workspace) & 7) return fail(OCT_EINVAL, "workspace must be 8-byte aligned");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int nb = finalize_blocks(plane_elems);
    int* partial = reinterpret_cast<int*>(workspace);
    hipLaunchKernelGGL(oct::finalize_kernel, dim3(nb, n_planes), dim3(oct::kFinThreads), 0, s, num, den, db,
                       plane_elems, thr, partial);
    HIP_TRY(hipGetLastError());
    launch_clim(partial, nb, n_planes, plane_clim, clim, s);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_quantize(const float* db, uint16_t* bits, int n_planes, long long plane_elems, const float* clim,
                 const float* plane_clim, int legacy_double, void* stream) {
    if (n_planes <= 0 || plane_elems <= 0) return OCT_OK;
    if (!db || !bits || !clim || (legacy_double && !plane_clim)) return fail(OCT_EINVAL, "null argument");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const long long gx = (plane_elems + 1023) / 1024;
    if (plane_elems % 4 == 0 && n_planes <= 65535 && gx <= 0x7FFFFFFFLL && (reinterpret_cast<uintptr_t>(db) & 15) == 0 &&
        (reinterpret_cast<uintptr_t>(bits) & 7) == 0) {
        hipLaunchKernelGGL(oct::quantize4_kernel, dim3((unsigned)gx, (unsigned)n_planes), dim3(256), 0, s, db, bits,
                           plane_elems, clim, plane_clim, legacy_double);
    } else {
        const long long n = (long long)n_planes * plane_elems;
        const int blocks = (int)std::min<long long>((n + 255) / 256, 1 << 16);
        hipLaunchKernelGGL(oct::quantize_kernel, dim3(blocks), dim3(256), 0, s, db, bits, plane_elems, n_planes, clim,
                           plane_clim, legacy_double);
    }
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

// ---------------------------------------------------------------- dispersion search support
int oct_plan_use_window(oct_plan* p, const float* dev_window) {
    if (!p) return fail(OCT_EINVAL, "null plan");
    if (reinterpret_cast<uintptr_t>(dev_window) % 16) return fail(OCT_EINVAL, "window must be 16-byte aligned");
    p->win_override = dev_window;
    p->win_fstride = 0;
    p->win_rstride = 0;
    return OCT_OK;
}

int oct_plan_use_windows(oct_plan* p, const float* dev_windows, long long file_stride, long long repeat_stride) {
    if (!p) return fail(OCT_EINVAL, "null plan");
    for (long long st : {file_stride, repeat_stride})
        if (st < 0 || (st > 0 && st < 2LL * p->n) || st % 4)
            return fail(OCT_EINVAL, "window stride %lld (0 or a multiple of 4 floats >= 2N)", st);
    if (reinterpret_cast<uintptr_t>(dev_windows) % 16) return fail(OCT_EINVAL, "windows must be 16-byte aligned");
    p->win_override = dev_windows;
    p->win_fstride = dev_windows ? file_stride : 0;
    p->win_rstride = dev_windows ? repeat_stride : 0;
    return OCT_OK;
}

static int topfrac_splits(int n_bscans, long long M) {
    long long s = (1024 + n_bscans - 1) / n_bscans;
    s = std::min<long long>(s, std::max<long long>(1, M / 4096));
    return (int)std::max<long long>(1, std::min<long long>(s, 4096));
}

size_t oct_topfrac_metric_workspace_bytes(int n_bscans, int n_x, int n_z, int z_min) {
    if (n_bscans <= 0 || n_x <= 0 || z_min < 0 || z_min >= n_z) return 0;
    const long long M = (long long)n_x * (n_z - z_min);
    const size_t S = (size_t)topfrac_splits(n_bscans, M);
    return (size_t)n_bscans * (sizeof(unsigned) * oct::kTfBins + sizeof(oct::TopfracState) + sizeof(double) * S);
}

int oct_topfrac_metric(const float* mag, int n_bscans, int n_x, int n_z, int z_min, double frac, double* out,
                       void* workspace, size_t workspace_bytes, void* stream) {
    if (n_bscans <= 0) return OCT_OK;
    if (!mag || !out || !workspace) return fail(OCT_EINVAL, "null argument");
    if (n_x <= 0 || z_min < 0 || z_min >= n_z) return fail(OCT_EINVAL, "bad metric shape (%d, %d, zmin %d)", n_x, n_z, z_min);
    if (!(frac > 0.0 && frac <= 1.0)) return fail(OCT_EINVAL, "fraction %g outside (0, 1]", frac);
    if (n_bscans > 65535) return fail(OCT_EINVAL, "at most 65535 B-scans per call");
    if (workspace_bytes < oct_topfrac_metric_workspace_bytes(n_bscans, n_x, n_z, z_min))
        return fail(OCT_EINVAL, "workspace %zu bytes, need %zu", workspace_bytes,
                    oct_topfrac_metric_workspace_bytes(n_bscans, n_x, n_z, z_min));
    if (reinterpret_cast<uintptr_t>(workspace) % 16) return fail(OCT_EINVAL, "workspace must be 16-byte aligned");
    const long long M = (long long)n_x * (n_z - z_min);
    long long k = (long long)(M * frac);
    if (k < 1) k = 1;
    const int S = topfrac_splits(n_bscans, M);
    unsigned* hist = static_cast<unsigned*>(workspace);
    auto* st = reinterpret_cast<oct::TopfracState*>(hist + (size_t)n_bscans * oct::kTfBins);
    auto* part = reinterpret_cast<double*>(st + n_bscans);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const long long ninit = (long long)n_bscans * oct::kTfBins;
    hipLaunchKernelGGL(oct::topfrac_init_kernel, dim3((unsigned)((ninit + oct::kTfThreads - 1) / oct::kTfThreads)),
                       dim3(oct::kTfThreads), 0, s, hist, st, k, n_bscans);
    const int lo[3] = {20, 10, 0}, nb[3] = {11, 10, 10};
    for (int p = 0; p < 3; ++p) {
        hipLaunchKernelGGL(oct::topfrac_hist_kernel, dim3(S, n_bscans), dim3(oct::kTfThreads), 0, s, mag, n_x, n_z, z_min,
                           st, hist, lo[p], nb[p], 1, 1);
        hipLaunchKernelGGL(oct::topfrac_select_kernel, dim3(n_bscans), dim3(oct::kTfThreads), 0, s, hist, st, lo[p], nb[p]);
    }
    hipLaunchKernelGGL(oct::topfrac_sum_kernel, dim3(S, n_bscans), dim3(oct::kTfThreads), 0, s, mag, n_x, n_z, z_min, st,
                       part);
    hipLaunchKernelGGL(oct::topfrac_final_kernel, dim3((n_bscans + oct::kTfThreads - 1) / oct::kTfThreads),
                       dim3(oct::kTfThreads), 0, s, st, part, S, k, n_bscans, out);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- no-allocation metric
// Scratch of oct_topfrac_metric_noalloc: the histograms, selection states and float64 partials
// of the metric, in a module-scope device array (every GPU the library runs on holds its own
// copy until the process ends).  Host code never passes `scratch` itself to a kernel: its name
// is a host-side stand-in, so the device address is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_dev.
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on that device.
int scratch_of_current_device(void** ptr) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(OCT_EINVAL, "device %d outside [0, %d)", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(OCT_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *ptr = g_scratch_dev[dev];
    return OCT_OK;
}
}  // namespace

extern "C" {

int oct_topfrac_metric_noalloc(const float* mag, int n_bscans, int n_x, int n_z, int z_min, double frac, double* out,
                               void* stream) {
    if (n_bscans <= 0) return OCT_OK;
    const size_t need = oct_topfrac_metric_workspace_bytes(n_bscans, n_x, n_z, z_min);
    if (need > kNoAllocScratchBytes)
        return fail(OCT_EINVAL, "metric of %d B-scans x %d x %d needs %zu bytes of scratch, more than the %zu bytes "
                    "of the no-allocation entry point (use oct_topfrac_metric with a workspace)", n_bscans, n_x, n_z,
                    need, kNoAllocScratchBytes);
    void* ws = nullptr;
    const int rc = scratch_of_current_device(&ws);
    if (rc != OCT_OK) return rc;
    // the same kernels as oct_topfrac_metric; its first kernel clears the histograms and states
    return oct_topfrac_metric(mag, n_bscans, n_x, n_z, z_min, frac, out, ws, kNoAllocScratchBytes, stream);
}

size_t oct_order_stats_workspace_bytes(int n_bscans, int n_ranks, int n_x, int n_z, int z_min) {
    if (n_bscans <= 0 || n_ranks <= 0 || n_x <= 0 || z_min < 0 || z_min >= n_z) return 0;
    return (size_t)n_bscans * (size_t)n_ranks * (sizeof(unsigned) * oct::kTfBins + sizeof(oct::TopfracState));
}

int oct_order_stats(const float* mag, int n_bscans, int n_x, int n_z, int z_min, int square, const long long* ranks,
                    int n_ranks, double* out, void* workspace, size_t workspace_bytes, void* stream) {
    if (n_bscans <= 0 || n_ranks <= 0) return OCT_OK;
    if (!mag || !ranks || !out || !workspace) return fail(OCT_EINVAL, "null argument");
    if (n_x <= 0 || z_min < 0 || z_min >= n_z) return fail(OCT_EINVAL, "bad shape (%d, %d, zmin %d)", n_x, n_z, z_min);
    const long long V = (long long)n_bscans * n_ranks;
    if (V > 65535) return fail(OCT_EINVAL, "at most 65535 (B-scan, rank) pairs per call");
    const size_t need = oct_order_stats_workspace_bytes(n_bscans, n_ranks, n_x, n_z, z_min);
    if (workspace_bytes < need) return fail(OCT_EINVAL, "workspace %zu bytes, need %zu", workspace_bytes, need);
    if (reinterpret_cast<uintptr_t>(workspace) % 16) return fail(OCT_EINVAL, "workspace must be 16-byte aligned");
    const long long M = (long long)n_x * (n_z - z_min);
    const int S = topfrac_splits((int)V, M);
    unsigned* hist = static_cast<unsigned*>(workspace);
    auto* st = reinterpret_cast<oct::TopfracState*>(hist + (size_t)V * oct::kTfBins);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const long long ninit = V * oct::kTfBins;
    hipLaunchKernelGGL(oct::order_init_kernel, dim3((unsigned)((ninit + oct::kTfThreads - 1) / oct::kTfThreads)),
                       dim3(oct::kTfThreads), 0, s, hist, st, ranks, (int)V);
    const int lo[3] = {20, 10, 0}, nb[3] = {11, 10, 10};
    for (int p = 0; p < 3; ++p) {
        hipLaunchKernelGGL(oct::topfrac_hist_kernel, dim3(S, (unsigned)V), dim3(oct::kTfThreads), 0, s, mag, n_x, n_z,
                           z_min, st, hist, lo[p], nb[p], n_ranks, square ? 1 : 0);
        hipLaunchKernelGGL(oct::topfrac_select_kernel, dim3((unsigned)V), dim3(oct::kTfThreads), 0, s, hist, st, lo[p],
                           nb[p]);
    }
    hipLaunchKernelGGL(oct::order_final_kernel, dim3((unsigned)((V + oct::kTfThreads - 1) / oct::kTfThreads)),
                       dim3(oct::kTfThreads), 0, s, st, (int)V, out);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

size_t oct_sharpness_metric_workspace_bytes(int metric, int n_bscans, int n_x, int n_z, int z_min) {
    if (n_bscans <= 0 || n_x <= 0 || z_min < 0 || z_min >= n_z) return 0;
    const long long M = (long long)n_x * (n_z - z_min);
    switch (metric) {
        case OCT_METRIC_TOP_PERCENTILE: return oct_topfrac_metric_workspace_bytes(n_bscans, n_x, n_z, z_min);
        case OCT_METRIC_KURTOSIS:
        case OCT_METRIC_ENTROPY: return sizeof(double) * 3 * (size_t)n_bscans * (size_t)topfrac_splits(n_bscans, M);
        case OCT_METRIC_LEGACY_PEAK: return sizeof(double) * (size_t)n_bscans * (size_t)(n_z - z_min);
        default: return 0;
    }
}

int oct_sharpness_metric(int metric, const float* mag, int n_bscans, int n_x, int n_z, int z_min, double* out,
                         void* workspace, size_t workspace_bytes, void* stream) {
    if (metric < OCT_METRIC_TOP_PERCENTILE || metric > OCT_METRIC_LEGACY_PEAK)
        return fail(OCT_EINVAL, "unknown metric %d", metric);
    if (metric == OCT_METRIC_TOP_PERCENTILE)
        return oct_topfrac_metric(mag, n_bscans, n_x, n_z, z_min, 0.01, out, workspace, workspace_bytes, stream);
    if (n_bscans <= 0) return OCT_OK;
    if (!mag || !out || !workspace) return fail(OCT_EINVAL, "null argument");
    if (n_x <= 0 || z_min < 0 || z_min >= n_z) return fail(OCT_EINVAL, "bad metric shape (%d, %d, zmin %d)", n_x, n_z, z_min);
    if (n_bscans > 65535) return fail(OCT_EINVAL, "at most 65535 B-scans per call");
    const size_t need = oct_sharpness_metric_workspace_bytes(metric, n_bscans, n_x, n_z, z_min);
    if (workspace_bytes < need) return fail(OCT_EINVAL, "workspace %zu bytes, need %zu", workspace_bytes, need);
    if (reinterpret_cast<uintptr_t>(workspace) % 16) return fail(OCT_EINVAL, "workspace must be 16-byte aligned");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    double* ws = static_cast<double*>(workspace);
    const int T = oct::kTfThreads;
    if (metric == OCT_METRIC_LEGACY_PEAK) {
        const int w = n_z - z_min;
        hipLaunchKernelGGL(oct::logmean_kernel, dim3((w + T - 1) / T, n_bscans), dim3(T), 0, s, mag, n_x, n_z, z_min, ws);
        hipLaunchKernelGGL(oct::rowmax_kernel, dim3(n_bscans), dim3(T), 0, s, ws, w, out);
    } else {
        const long long M = (long long)n_x * (n_z - z_min);
        const int S = topfrac_splits(n_bscans, M);
        double* part = ws;
        double* ent = ws + 2 * (size_t)n_bscans * S;
        hipLaunchKernelGGL(oct::moments_kernel, dim3(S, n_bscans), dim3(T), 0, s, mag, n_x, n_z, z_min, part);
        if (metric == OCT_METRIC_ENTROPY)
            hipLaunchKernelGGL(oct::entropy_kernel, dim3(S, n_bscans), dim3(T), 0, s, mag, n_x, n_z, z_min, part, ent);
        hipLaunchKernelGGL(oct::moments_final_kernel, dim3((n_bscans + T - 1) / T), dim3(T), 0, s, part, ent, S,
                           n_bscans, metric == OCT_METRIC_KURTOSIS ? 1 : 2, out);
    }
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_dispersion_windows(const double* betas, const float* head, int head_stride, double head_scale, double offset,
                           double to_ref, int count, const double* hann_over_n, const double* dk2, int n, float* out,
                           void* stream) {
    if (count <= 0) return OCT_OK;
    if ((!betas && !head) || !hann_over_n || !dk2 || !out) return fail(OCT_EINVAL, "null argument");
    if (n != oct::kN || count > 65535 || (head && head_stride <= 0))
        return fail(OCT_EINVAL, "bad window shape (%d x %d; n = %d only)", count, n, oct::kN);
    if (reinterpret_cast<uintptr_t>(out) % 16) return fail(OCT_EINVAL, "out must be 16-byte aligned");
    hipLaunchKernelGGL(oct::windows_kernel, dim3((n + 255) / 256, count), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       head ? nullptr : betas, head, head_stride, head_scale, offset, to_ref, hann_over_n, dk2, n, out);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

// ---------------------------------------------------------------- operator (MFMA) form
struct oct_opplan {
    int device;
    int n;
    int z0, nzc, r_pad, nzc_pad, rt;   // rt: operator row tiles of 32 per GEMM workgroup
    int sa;
    int tt;                       // transposed-band width
    double max_win;               // largest |window| of the creation window (sets sa)
    _Float16* op_hi;
    _Float16* op_lo;
    _Float16* opt_hi;             // step-tiled copies for opgemm_rf_kernel (opack_kernel)
    _Float16* opt_lo;
    // kept for in-place rebuilds with another window (oct_opplan_set_window)
    int* d_start;
    int* d_count;
    double* d_w;
    double* d_win;                // the creation window (N, 2)
    double* d_u;
    double2* d_c;                 // (nzc) rank-1 term, rebuilt with each window
};

// the rf kernel's step-tiled operator copy, rebuilt after every opbuild_a_kernel
static void launch_opack(const oct_opplan* p, hipStream_t s) {
    const long long chunks = (long long)p->r_pad * (p->n / 8);
    hipLaunchKernelGGL(oct::opack_kernel, dim3((unsigned)((chunks + 255) / 256)), dim3(256), 0, s, p->op_hi, p->op_lo,
                       p->r_pad, p->n, 32 * p->rt, p->opt_hi, p->opt_lo);
}

int oct_opplan_rows(int nzc, int* row_tiles) {
    if (nzc < 1) return fail(OCT_EINVAL, "%d bins", nzc);
    // rows: 16 bins' real then imaginary parts per tile of 32 rows.  The GEMM takes rt = 1-4 row
    // tiles per workgroup: all of them up to 4 tiles, else 4, or 3 where that pads fewer tiles
    const int tiles = 2 * ((nzc + 15) / 16 * 16) / 32;
    const int rt = tiles <= 4 ? tiles : ((tiles + 2) / 3 * 3 < (tiles + 3) / 4 * 4 ? 3 : 4);
    if (row_tiles) *row_tiles = rt;
    return (tiles + rt - 1) / rt * rt * 32;
}

int oct_opplan_create(oct_opplan** out, int device, int n_lambda, int taps, const int32_t* band_start,
                      const double* band_w, const double* window, int z0, int nzc, int* scale_log2,
                      double* op64_host) {
    if (!out || !band_start || !band_w || !window) return fail(OCT_EINVAL, "null argument");
    *out = nullptr;
    const int N = n_lambda;
    if (N != oct::kN) return fail(OCT_EUNSUPPORTED, "n_lambda %d unsupported (2048 only)", N);
    if (taps <= 0 || taps > 64) return fail(OCT_EINVAL, "taps %d", taps);
    if (z0 < 0 || nzc <= 0 || z0 + nzc > N / 2) return fail(OCT_EINVAL, "bins [%d, %d) outside [0, %d)", z0, z0 + nzc, N / 2);
    // transposed band: for each source j the output rows i that use it, with their weights
    std::vector<std::vector<std::pair<int, double>>> cols(N);
    std::vector<double> u(N, 1.0);
    for (int i = 0; i < N; ++i) {
        const int s0 = band_start[i];
        if (s0 < 0 || s0 + taps > N) return fail(OCT_EINVAL, "band_start[%d]=%d", i, s0);
        for (int t = 0; t < taps; ++t) {
            const double w = band_w[(size_t)i * taps + t];
            u[i] -= w;
            if (w != 0.0) cols[s0 + t].push_back({i, w});
        }
    }
    int TT = 1;
    for (auto& c : cols) TT = std::max<int>(TT, (int)c.size());
    std::vector<int> tb_start(N), tb_count(N);
    std::vector<double> tb_w((size_t)N * TT, 0.0);
    std::vector<int> tb_idx((size_t)N * TT, 0);
    for (int j = 0; j < N; ++j) {
        tb_count[j] = (int)cols[j].size();
        tb_start[j] = cols[j].empty() ? 0 : cols[j][0].first;
        for (int t = 0; t < tb_count[j]; ++t) {
            if (cols[j][t].first != tb_start[j] + t) return fail(OCT_EINVAL, "sinc band is not contiguous in column %d", j);
            tb_w[(size_t)j * TT + t] = cols[j][t].second;
        }
    }
    // scale: keep |A| * 2^sa below 2^14 (fp16 headroom for the hi part)
    double amax_win = 0.0;
    for (int i = 0; i < N; ++i) amax_win = std::max(amax_win, std::hypot(window[2 * i], window[2 * i + 1]));
    double colsum = 0.0;
    for (int j = 0; j < N; ++j) {
        double s = 0.0;
        for (auto& e : cols[j]) s += std::fabs(e.second);
        colsum = std::max(colsum, s);
    }
    double usum = 0.0;
    for (int i = 0; i < N; ++i) usum += std::fabs(u[i]);
    const double bound = amax_win / N * (colsum + usum / N);
    const int sa = (int)std::floor(std::log2(16384.0 / bound));
    const int nzc_pad = (nzc + 15) / 16 * 16;
    int rt = 0;
    const int r_pad = oct_opplan_rows(nzc, &rt);

    int prev = 0;
    HIP_TRY(hipGetDevice(&prev));
    HIP_TRY(hipSetDevice(device));
    oct_opplan* p = new oct_opplan();
    p->device = device;
    p->n = N;
    p->z0 = z0;
    p->nzc = nzc;
    p->r_pad = r_pad;
    p->rt = rt;
    p->nzc_pad = nzc_pad;
    p->sa = sa;
    p->tt = TT;
    p->max_win = amax_win;
    int* d_start = nullptr;
    int* d_count = nullptr;
    double *d_w = nullptr, *d_win = nullptr, *d_u = nullptr, *d_op64 = nullptr;
    double2* d_c = nullptr;
    int rc = OCT_OK;
    const size_t opbytes = sizeof(_Float16) * (size_t)r_pad * N;
    if (hipMalloc(&p->op_hi, opbytes) != hipSuccess || hipMalloc(&p->op_lo, opbytes) != hipSuccess ||
        hipMalloc(&p->opt_hi, opbytes) != hipSuccess || hipMalloc(&p->opt_lo, opbytes) != hipSuccess ||
        hipMemset(p->op_hi, 0, opbytes) != hipSuccess || hipMemset(p->op_lo, 0, opbytes) != hipSuccess ||
        hipMalloc(&d_start, sizeof(int) * N) != hipSuccess || hipMalloc(&d_count, sizeof(int) * N) != hipSuccess ||
        hipMalloc(&d_w, sizeof(double) * N * TT) != hipSuccess || hipMalloc(&d_win, sizeof(double) * 2 * N) != hipSuccess ||
        hipMalloc(&d_u, sizeof(double) * N) != hipSuccess || hipMalloc(&d_c, sizeof(double2) * nzc) != hipSuccess ||
        (op64_host && hipMalloc(&d_op64, sizeof(double) * 2 * (size_t)nzc * N) != hipSuccess) ||
        hipMemcpy(d_start, tb_start.data(), sizeof(int) * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_count, tb_count.data(), sizeof(int) * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_w, tb_w.data(), sizeof(double) * N * TT, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_win, window, sizeof(double) * 2 * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_u, u.data(), sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(OCT_EHIP, "operator upload failed: %s", hipGetErrorString(hipGetLastError()));
    }
    if (rc == OCT_OK) {
        hipLaunchKernelGGL(oct::opbuild_c_kernel, dim3(nzc), dim3(256), 0, 0, d_win, d_u, N, z0, d_c);
        hipLaunchKernelGGL(oct::opbuild_a_kernel, dim3((N + 255) / 256, nzc_pad), dim3(256), 0, 0, d_start, d_count,
                           d_w, TT, d_win, d_c, N, z0, nzc, nzc_pad, std::ldexp(1.0, sa), p->op_hi, p->op_lo, d_op64);
        launch_opack(p, 0);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess && op64_host)
            e = hipMemcpy(op64_host, d_op64, sizeof(double) * 2 * (size_t)nzc * N, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(OCT_EHIP, "operator build failed: %s", hipGetErrorString(e));
    }
    p->d_start = d_start;
    p->d_count = d_count;
    p->d_w = d_w;
    p->d_win = d_win;
    p->d_u = d_u;
    p->d_c = d_c;
    if (d_op64) hipFree(d_op64);
    hipSetDevice(prev);
    if (rc != OCT_OK) {
        oct_opplan_destroy(p);
        return rc;
    }
    if (scale_log2) *scale_log2 = sa;
    *out = p;
    return OCT_OK;
}

int oct_opplan_destroy(oct_opplan* p) {
    if (!p) return OCT_OK;
    int prev = 0;
    hipGetDevice(&prev);
    hipSetDevice(p->device);
    if (p->op_hi) hipFree(p->op_hi);
    if (p->op_lo) hipFree(p->op_lo);
    if (p->opt_hi) hipFree(p->opt_hi);
    if (p->opt_lo) hipFree(p->opt_lo);
    if (p->d_start) hipFree(p->d_start);
    if (p->d_count) hipFree(p->d_count);
    if (p->d_w) hipFree(p->d_w);
    if (p->d_win) hipFree(p->d_win);
    if (p->d_u) hipFree(p->d_u);
    if (p->d_c) hipFree(p->d_c);
    hipSetDevice(prev);
    delete p;
    return OCT_OK;
}

int oct_opplan_scale_bound(const oct_opplan* p, double* max_window_abs) {
    if (!p || !max_window_abs) return fail(OCT_EINVAL, "null argument");
    *max_window_abs = p->max_win;
    return OCT_OK;
}

int oct_opplan_set_window(oct_opplan* p, const double* window, void* stream) {
    if (!p) return fail(OCT_EINVAL, "null argument");
    const double* win = window ? window : p->d_win;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(oct::opbuild_c_kernel, dim3(p->nzc), dim3(256), 0, s, win, p->d_u, p->n, p->z0, p->d_c);
    hipLaunchKernelGGL(oct::opbuild_a_kernel, dim3((p->n + 255) / 256, p->nzc_pad), dim3(256), 0, s, p->d_start,
                       p->d_count, p->d_w, p->tt, win, p->d_c, p->n, p->z0, p->nzc, p->nzc_pad, std::ldexp(1.0, p->sa),
                       p->op_hi, p->op_lo, (double*)nullptr);
    launch_opack(p, s);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_opspectral(const oct_opplan* p, const void* raw, int raw_dtype, int n_files, int lines, int apod_mode,
                   int apod_size, float* out, long long out_stride, const float* bg, void* stream) {
    if (n_files == 0) return OCT_OK;
    if (!p || !out || !bg) return fail(OCT_EINVAL, "null argument");
    int line0, nX;
    int rc = check_raw(raw, raw_dtype, n_files, lines, apod_mode, apod_size, 1, &line0, &nX);
    if (rc) return rc;
    if (out_stride < p->nzc) return fail(OCT_EINVAL, "out_stride %lld < bins %d", out_stride, p->nzc);
    oct::OpGemmArgs a;
    a.raw = raw;
    a.bg = bg;
    a.op_hi = p->op_hi;
    a.op_lo = p->op_lo;
    a.opt_hi = p->opt_hi;
    a.opt_lo = p->opt_lo;
    a.out = out;
    a.M = (long long)n_files * nX;
    a.lines = lines;
    a.line0 = line0;
    a.nX = nX;
    a.N = p->n;
    a.nzc = p->nzc;
    const int rt = p->rt;                                     // row tiles of 32 per workgroup
    a.row_blocks = p->r_pad / (32 * rt);
    a.col_blocks = (int)((a.M + oct::kOpCols - 1) / oct::kOpCols);
    a.out_stride = out_stride;
    a.inv_scale = (float)std::ldexp(1.0, -p->sa);
    const long long nblk = (long long)a.row_blocks * a.col_blocks;
    if (nblk > (1LL << 30) || a.M >= (1LL << 31)) return fail(OCT_EINVAL, "batch too large");
    const int grid = (int)((nblk + 7) / 8 * 8);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // OCT_OPGEMM=classic selects the kernel that stages the data fragments through LDS, =rf the
    // one that loads them directly into registers (same results), for comparisons
    const char* sel = std::getenv("OCT_OPGEMM");
    if (sel && std::strcmp(sel, "classic") == 0 && p->r_pad % oct::kOpRows == 0) {
        switch (raw_dtype) {
            case OCT_RAW_I16: hipLaunchKernelGGL(oct::opgemm_kernel<int16_t>, dim3(grid), dim3(256), 0, s, a); break;
            case OCT_RAW_U16: hipLaunchKernelGGL(oct::opgemm_kernel<uint16_t>, dim3(grid), dim3(256), 0, s, a); break;
            default: hipLaunchKernelGGL(oct::opgemm_kernel<float>, dim3(grid), dim3(256), 0, s, a); break;
        }
    } else if (sel && std::strcmp(sel, "rf") == 0) {
        switch (raw_dtype) {
            case OCT_RAW_I16: launch_opgemm_rf<int16_t>(a, rt, grid, s); break;
            case OCT_RAW_U16: launch_opgemm_rf<uint16_t>(a, rt, grid, s); break;
            default: launch_opgemm_rf<float>(a, rt, grid, s); break;
        }
    } else {
        switch (raw_dtype) {
            case OCT_RAW_I16: launch_opgemm_lt<int16_t>(a, rt, grid, s); break;
            case OCT_RAW_U16: launch_opgemm_lt<uint16_t>(a, rt, grid, s); break;
            default: launch_opgemm_rf<float>(a, rt, grid, s); break;
        }
    }
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}


// ---------------------------------------------------------------- learned estimator
static constexpr long long kConvWaveMin = 2048;

static int conv_shape_ok(int n, int ci, int h, int w, int co) {
    return n > 0 && ci > 0 && co > 0 && h >= 2 && w >= 2 && (h % 2) == 0 && (w % 2) == 0;
}

static int conv_fwd(const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y, int n, int ci, int h, int wd,
                    int co, bool matrix_cores, void* stream) {
    if (!x || !w || !bias || !y) return fail(OCT_EINVAL, "null argument");
    if (!conv_shape_ok(n, ci, h, wd, co)) return fail(OCT_EINVAL, "bad conv shape (%d, %d, %d, %d) -> %d", n, ci, h, wd, co);
    if (co % oct::kConvCoB) return fail(OCT_EINVAL, "output channels %d not a multiple of %d", co, oct::kConvCoB);
    const long long P = (long long)n * (h / 2) * (wd / 2);
    if ((P * (co / 16) + 3) / 4 > 0x7FFFFFFFLL) return fail(OCT_EINVAL, "batch too large");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int Ho = h / 2, Wo = wd / 2;
    if (matrix_cores && co % 16 == 0 && ci >= 8 && Wo >= 16) {
        // matrix cores: 8 x 32 output positions x all channels per block; 2 x 32 when that would
        // leave fewer than 512 blocks; 4 x 16 positions x 16 channels below 32 columns (fewer
        // input channels than 8: the LDS-tiled kernel below, whose chunks are not padded)
        if ((long long)n * co / 16 > 65535) return fail(OCT_EINVAL, "batch too large for one launch");
        const long long big = (long long)((Wo + 31) / 32) * ((Ho + 7) / 8) * n;
        const int ct = co % 64 == 0 ? 4 : co % 32 == 0 ? 2 : 1;
        if (Wo >= 32 && big >= 512)
            launch_conv_mfma<32, 8>(ct, x, w, bias, y, n, ci, h, wd, co, s);
        else if (Wo >= 32)
            launch_conv_mfma<32, 2>(ct, x, w, bias, y, n, ci, h, wd, co, s);
        else
            launch_conv_mfma<16, 4>(1, x, w, bias, y, n, ci, h, wd, co, s);
    } else if (Wo >= 2 * oct::kConvTX) {
        // LDS-tiled: 8 x 32 output positions x 16 channels per block, 8 channels when that
        // would leave fewer than 2 blocks per CU
        const unsigned gx = (Wo + oct::kConvTX - 1) / oct::kConvTX, gy = (Ho + oct::kConvTY - 1) / oct::kConvTY;
        const int cb = ((long long)gx * gy * n * (co / 16) < 512) ? 8 : 16;
        const long long gz = (long long)n * (co / cb);
        if (gz > 65535) return fail(OCT_EINVAL, "batch too large for one launch");
        if ((long long)oct::kConvCC * h * wd > 0x7FFFFFFFLL) return fail(OCT_EINVAL, "image too large");
        if (cb == 8)
            hipLaunchKernelGGL(oct::conv_fwd_kernel<8>, dim3(gx, gy, (unsigned)gz), dim3(oct::kCnnThreads), 0, s, x, w,
                               bias, y, n, ci, h, wd, co);
        else
            hipLaunchKernelGGL(oct::conv_fwd_kernel<16>, dim3(gx, gy, (unsigned)gz), dim3(oct::kCnnThreads), 0, s, x, w,
                               bias, y, n, ci, h, wd, co);
    } else {
        // narrow layers: one wave per (P adjacent positions, 16 channels), lanes across the
        // reduction; P = 4 while that leaves >= 2048 waves (8 per CU), P = 2 only where Wo is
        // not a multiple of 4 and >= 4096 waves remain
        const long long waves = P * (co / 16);
        const int pp = (Wo % 4 == 0 && waves / 4 >= kConvWaveMin) ? 4 : (Wo % 2 == 0 && waves / 2 >= 2 * kConvWaveMin) ? 2 : 1;
        const dim3 grid((unsigned)((waves / pp + 3) / 4));
        if (pp == 4)
            hipLaunchKernelGGL(oct::conv_fwd_wave_kernel<4>, grid, dim3(oct::kCnnThreads), 0, s, x, w, bias, y, n, ci, h, wd, co);
        else if (pp == 2)
            hipLaunchKernelGGL(oct::conv_fwd_wave_kernel<2>, grid, dim3(oct::kCnnThreads), 0, s, x, w, bias, y, n, ci, h, wd, co);
        else
            hipLaunchKernelGGL(oct::conv_fwd_wave_kernel<1>, grid, dim3(oct::kCnnThreads), 0, s, x, w, bias, y, n, ci, h, wd, co);
    }
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_cnn_conv_fwd(const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y, int n, int ci, int h,
                     int wd, int co, void* stream) {
    return conv_fwd(x, w, bias, y, n, ci, h, wd, co, false, stream);
}

int oct_cnn_conv_fwd_mfma(const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y, int n, int ci, int h,
                          int wd, int co, void* stream) {
    return conv_fwd(x, w, bias, y, n, ci, h, wd, co, true, stream);
}

int oct_cnn_conv_bwd_data(const float* dpre, const uint16_t* w, const uint16_t* xmask, float* dx, int n, int ci, int h,
                          int wd, int co, void* stream) {
    if (!dpre || !w || !dx) return fail(OCT_EINVAL, "null argument");
    if (!conv_shape_ok(n, ci, h, wd, co)) return fail(OCT_EINVAL, "bad conv shape (%d, %d, %d, %d) -> %d", n, ci, h, wd, co);
    if (ci % oct::kConvCiB) return fail(OCT_EINVAL, "input channels %d not a multiple of %d", ci, oct::kConvCiB);
    const long long P = (long long)n * (h / 2) * (wd / 2);
    const long long g = (P + oct::kCnnThreads - 1) / oct::kCnnThreads;
    if (g > 0x7FFFFFFFLL) return fail(OCT_EINVAL, "batch too large");
    hipLaunchKernelGGL(oct::conv_bwd_data_kernel, dim3((unsigned)g, ci / oct::kConvCiB), dim3(oct::kCnnThreads), 0,
                       reinterpret_cast<hipStream_t>(stream), dpre, w, xmask, dx, n, ci, h, wd, co);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

static int conv_splits(int n, int ci, int h, int wd, int co) {
    const long long P = (long long)n * (h / 2) * (wd / 2);
    long long s = (2048 + (long long)co * ci - 1) / ((long long)co * ci);
    s = std::min<long long>(s, std::max<long long>(1, P / 1024));
    return (int)std::max<long long>(1, std::min<long long>(s, 1024));
}

size_t oct_cnn_conv_bwd_weight_workspace_bytes(int n, int ci, int h, int wd, int co) {
    if (!conv_shape_ok(n, ci, h, wd, co)) return 0;
    return sizeof(float) * 10 * (size_t)conv_splits(n, ci, h, wd, co) * (size_t)co * (size_t)ci;
}

int oct_cnn_conv_bwd_weight(const float* dpre, const uint16_t* x, float* dw, float* db, int n, int ci, int h, int wd,
                            int co, void* workspace, size_t workspace_bytes, void* stream) {
    if (!dpre || !x || !dw || !db || !workspace) return fail(OCT_EINVAL, "null argument");
    if (!conv_shape_ok(n, ci, h, wd, co)) return fail(OCT_EINVAL, "bad conv shape (%d, %d, %d, %d) -> %d", n, ci, h, wd, co);
    if ((long long)co * ci > 0x7FFFFFFFLL) return fail(OCT_EINVAL, "too many channels");
    const int S = conv_splits(n, ci, h, wd, co);
    if (workspace_bytes < oct_cnn_conv_bwd_weight_workspace_bytes(n, ci, h, wd, co))
        return fail(OCT_EINVAL, "workspace %zu bytes, need %zu", workspace_bytes,
                    oct_cnn_conv_bwd_weight_workspace_bytes(n, ci, h, wd, co));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    float* part = static_cast<float*>(workspace);
    hipLaunchKernelGGL(oct::conv_bwd_weight_kernel, dim3(co * ci, S), dim3(oct::kCnnThreads), 0, s, dpre, x, part, n,
                       ci, h, wd, co);
    HIP_TRY(hipGetLastError());
    const long long tot = (long long)co * ci * 9 + co;
    hipLaunchKernelGGL(oct::conv_bwd_weight_reduce_kernel, dim3((unsigned)((tot + oct::kCnnThreads - 1) / oct::kCnnThreads)),
                       dim3(oct::kCnnThreads), 0, s, part, dw, db, ci, co, S);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

// ---- backward convolutions on the matrix cores
static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

struct BwdWeightMfmaPlan {
    int P, ppad, ncols_pad, S, slice;
    size_t xcol_bytes, part_bytes;
};

static bool bwd_weight_mfma_plan(int n, int ci, int h, int wd, int co, BwdWeightMfmaPlan* q) {
    if (!conv_shape_ok(n, ci, h, wd, co) || co % 16 || co > 65535) return false;
    const long long P = (long long)n * (h / 2) * (wd / 2);
    const long long ncols = 9LL * ci + 1;
    if (P >= (1LL << 30) || ncols > (1 << 20)) return false;
    q->P = (int)P;
    q->ppad = (int)((P + 31) / 32 * 32);
    q->ncols_pad = (int)((ncols + 15) / 16 * 16);
    // slices of whole 32-position steps, about 2048 waves in all, at least 256 positions each
    const long long tasks = (long long)(co / 16) * ((q->ncols_pad / 16 + 3) / 4);
    long long s = std::max<long long>(1, (2048 + tasks - 1) / tasks);
    long long slice = (q->ppad + s - 1) / s;
    slice = std::max<long long>(256, (slice + 31) / 32 * 32);
    q->slice = (int)std::min<long long>(slice, q->ppad);
    q->S = (int)((q->ppad + q->slice - 1) / q->slice);
    q->xcol_bytes = align256(sizeof(uint16_t) * (size_t)q->ncols_pad * (size_t)q->ppad);
    q->part_bytes = align256(sizeof(float) * (size_t)q->S * (size_t)co * (size_t)q->ncols_pad);
    return true;
}

size_t oct_cnn_conv_bwd_weight_mfma_workspace_bytes(int n, int ci, int h, int wd, int co) {
    BwdWeightMfmaPlan q;
    if (!bwd_weight_mfma_plan(n, ci, h, wd, co, &q)) return 0;
    return q.xcol_bytes + q.part_bytes;
}

int oct_cnn_conv_bwd_weight_mfma(const float* dpre, const uint16_t* x, float* dw, float* db, int n, int ci, int h,
                                 int wd, int co, void* workspace, size_t workspace_bytes, void* stream) {
    if (!dpre || !x || !dw || !db || !workspace) return fail(OCT_EINVAL, "null argument");
    BwdWeightMfmaPlan q;
    if (!bwd_weight_mfma_plan(n, ci, h, wd, co, &q))
        return fail(OCT_EINVAL, "bad conv shape (%d, %d, %d, %d) -> %d (output channels a multiple of 16)", n, ci, h,
                    wd, co);
    if (workspace_bytes < q.xcol_bytes + q.part_bytes)
        return fail(OCT_EINVAL, "workspace %zu bytes, need %zu", workspace_bytes, q.xcol_bytes + q.part_bytes);
    if (reinterpret_cast<uintptr_t>(workspace) % 16 || reinterpret_cast<uintptr_t>(dpre) % 16)
        return fail(OCT_EINVAL, "workspace and dpre must be 16-byte aligned");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    uint16_t* xcol = static_cast<uint16_t*>(workspace);
    float* part = reinterpret_cast<float*>(static_cast<char*>(workspace) + q.xcol_bytes);
    const long long items = (long long)q.ncols_pad * (q.ppad / 8);
    hipLaunchKernelGGL(oct::conv_im2col_kernel, dim3((unsigned)((items + oct::kCnnThreads - 1) / oct::kCnnThreads)),
                       dim3(oct::kCnnThreads), 0, s, x, xcol, ci, h, wd, q.P, q.ppad, q.ncols_pad);
    HIP_TRY(hipGetLastError());
    const long long waves = (long long)(co / 16) * ((q.ncols_pad / 16 + 3) / 4) * q.S;
    hipLaunchKernelGGL(oct::conv_bwd_weight_mfma_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(oct::kCnnThreads), 0,
                       s, dpre, xcol, part, co, (h / 2) * (wd / 2), q.P, q.ppad, q.ncols_pad, q.S, q.slice);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(oct::conv_bwd_weight_mfma_reduce_kernel, dim3((unsigned)((9LL * ci + 1 + 15) / 16), (unsigned)co),
                       dim3(oct::kCnnThreads), 0, s, part, dw, db, ci, co, q.ncols_pad, q.S);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

static bool bwd_data_mfma_sizes(int n, int ci, int h, int wd, int co, size_t* dt_bytes, size_t* wt_bytes) {
    if (!conv_shape_ok(n, ci, h, wd, co) || co % 8 || ci % 16 || ci > 64) return false;
    const long long P = (long long)n * (h / 2) * (wd / 2);
    if (P >= (1LL << 30) || (long long)n * ci * h * wd >= (1LL << 40)) return false;
    *dt_bytes = align256(sizeof(uint16_t) * (size_t)P * (size_t)co);
    *wt_bytes = align256(sizeof(uint16_t) * 9 * (size_t)ci * (size_t)co);
    return true;
}

size_t oct_cnn_conv_bwd_data_mfma_workspace_bytes(int n, int ci, int h, int wd, int co) {
    size_t dt = 0, wt = 0;
    if (!bwd_data_mfma_sizes(n, ci, h, wd, co, &dt, &wt)) return 0;
    return 2 * dt + wt;
}

int oct_cnn_conv_bwd_data_mfma(const float* dpre, const uint16_t* w, const uint16_t* xmask, float* dx, int n, int ci,
                               int h, int wd, int co, void* workspace, size_t workspace_bytes, void* stream) {
    if (!dpre || !w || !dx || !workspace) return fail(OCT_EINVAL, "null argument");
    size_t dt = 0, wt = 0;
    if (!bwd_data_mfma_sizes(n, ci, h, wd, co, &dt, &wt))
        return fail(OCT_EINVAL, "bad conv shape (%d, %d, %d, %d) -> %d (input channels a multiple of 16 up to 64, "
                    "output channels a multiple of 8)", n, ci, h, wd, co);
    if (workspace_bytes < 2 * dt + wt) return fail(OCT_EINVAL, "workspace %zu bytes, need %zu", workspace_bytes, 2 * dt + wt);
    if (reinterpret_cast<uintptr_t>(workspace) % 16) return fail(OCT_EINVAL, "workspace must be 16-byte aligned");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    uint16_t* dth = static_cast<uint16_t*>(workspace);
    uint16_t* dtl = reinterpret_cast<uint16_t*>(static_cast<char*>(workspace) + dt);
    uint16_t* wtb = reinterpret_cast<uint16_t*>(static_cast<char*>(workspace) + 2 * dt);
    const int P = n * (h / 2) * (wd / 2);
    const long long items = (long long)(co / 8) * P;
    hipLaunchKernelGGL(oct::conv_dpre_split_kernel, dim3((unsigned)((items + oct::kCnnThreads - 1) / oct::kCnnThreads)),
                       dim3(oct::kCnnThreads), 0, s, dpre, dth, dtl, co, (h / 2) * (wd / 2), P);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(oct::conv_w_transpose_kernel, dim3((unsigned)((9 * ci * co + oct::kCnnThreads - 1) / oct::kCnnThreads)),
                       dim3(oct::kCnnThreads), 0, s, w, wtb, ci, co);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(oct::conv_bwd_data_mfma_kernel, dim3((unsigned)((P + 63) / 64), 4), dim3(oct::kCnnThreads), 0, s,
                       dth, dtl, wtb, xmask, dx, n, ci, h, wd, co);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_cnn_fc_fwd(const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y_relu_bf16, float* y_linear,
                   int n, int k, int o, void* stream) {
    if (!x || !w || !bias) return fail(OCT_EINVAL, "null argument");
    if ((y_relu_bf16 != nullptr) == (y_linear != nullptr)) return fail(OCT_EINVAL, "exactly one output (ReLU bf16 or linear f32)");
    if (n <= 0 || k <= 0 || o <= 0) return fail(OCT_EINVAL, "bad fc shape (%d, %d) -> %d", n, k, o);
    const long long tot = (long long)n * o;
    const long long wpb = oct::kCnnThreads / 64;                  // outputs (waves) per block
    if ((tot + wpb - 1) / wpb > 0x7FFFFFFFLL) return fail(OCT_EINVAL, "batch too large");
    hipLaunchKernelGGL(oct::fc_fwd_kernel, dim3((unsigned)((tot + wpb - 1) / wpb)), dim3(oct::kCnnThreads), 0,
                       reinterpret_cast<hipStream_t>(stream), x, w, bias, y_relu_bf16, y_linear, n, k, o);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_cnn_fc_bwd_data(const float* dpre, const uint16_t* w, const uint16_t* xmask, float* dx, int n, int k, int o,
                        void* stream) {
    if (!dpre || !w || !dx) return fail(OCT_EINVAL, "null argument");
    if (n <= 0 || k <= 0 || o <= 0) return fail(OCT_EINVAL, "bad fc shape (%d, %d) -> %d", n, k, o);
    const long long tot = (long long)n * k;
    hipLaunchKernelGGL(oct::fc_bwd_data_kernel, dim3((unsigned)((tot + oct::kCnnThreads - 1) / oct::kCnnThreads)),
                       dim3(oct::kCnnThreads), 0, reinterpret_cast<hipStream_t>(stream), dpre, w, xmask, dx, n, k, o);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_cnn_fc_bwd_weight(const float* dpre, const uint16_t* x, float* dw, float* db, int n, int k, int o,
                          void* stream) {
    if (!dpre || !x || !dw || !db) return fail(OCT_EINVAL, "null argument");
    if (n <= 0 || k <= 0 || o <= 0) return fail(OCT_EINVAL, "bad fc shape (%d, %d) -> %d", n, k, o);
    const long long tot = (long long)o * k + o;
    hipLaunchKernelGGL(oct::fc_bwd_weight_kernel, dim3((unsigned)((tot + oct::kCnnThreads - 1) / oct::kCnnThreads)),
                       dim3(oct::kCnnThreads), 0, reinterpret_cast<hipStream_t>(stream), dpre, x, dw, db, n, k, o);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_cnn_gauss_nll(const float* head, const float* target, const float* weights, float* loss, float* dhead, int n,
                      int n_targets, void* stream) {
    if (!head || !target || !weights || !loss || !dhead) return fail(OCT_EINVAL, "null argument");
    if (n <= 0 || n_targets <= 0 || n_targets > oct::kNllMaxT) return fail(OCT_EINVAL, "bad loss shape (%d, %d)", n, n_targets);
    hipLaunchKernelGGL(oct::gauss_nll_kernel, dim3(1), dim3(oct::kCnnThreads), 0, reinterpret_cast<hipStream_t>(stream),
                       head, target, weights, loss, dhead, n, n_targets);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_cnn_adam(float* params, const float* grads, float* m, float* v, uint16_t* params_bf16, long long count,
                 float lr, float beta1, float beta2, float eps, float weight_decay, int step, void* stream) {
    if (!params || !grads || !m || !v || !params_bf16) return fail(OCT_EINVAL, "null argument");
    if (count <= 0 || step < 1) return fail(OCT_EINVAL, "bad count %lld / step %d", count, step);
    const float c1 = (float)(1.0 - std::pow((double)beta1, step));
    const float c2 = (float)(1.0 - std::pow((double)beta2, step));
    hipLaunchKernelGGL(oct::adam_kernel, dim3((unsigned)((count + oct::kCnnThreads - 1) / oct::kCnnThreads)),
                       dim3(oct::kCnnThreads), 0, reinterpret_cast<hipStream_t>(stream), params, grads, m, v, params_bf16,
                       count, lr, beta1, beta2, eps, weight_decay, c1, c2);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_logmag_crop(const float* mag, int n_bscans, int n_x, long long x_stride, int n_rows, int x0, int z0, int zpool,
                    int height, int width, float offset_db, float scale_db, uint16_t* out, int channels, void* stream) {
    if (n_bscans <= 0) return OCT_OK;
    if (!mag || !out) return fail(OCT_EINVAL, "null argument");
    if (height <= 0 || width <= 0 || zpool <= 0 || x0 < 0 || z0 < 0 || x0 + width > n_x ||
        z0 + (long long)height * zpool > n_rows || n_rows > x_stride)
        return fail(OCT_EINVAL, "crop x [%d, %d) z [%d, %lld) outside (%d, %d)", x0, x0 + width, z0,
                    z0 + (long long)height * zpool, n_x, n_rows);
    if (channels < 1) return fail(OCT_EINVAL, "%d channels", channels);
    if (!(scale_db > 0.f)) return fail(OCT_EINVAL, "scale must be positive");
    if ((long long)n_bscans * channels > 65535) return fail(OCT_EINVAL, "at most 65535 images per call");
    const long long tiles = (long long)((height + oct::kCropT - 1) / oct::kCropT) * ((width + oct::kCropT - 1) / oct::kCropT);
    if (tiles > 0x7FFFFFFFLL) return fail(OCT_EINVAL, "crop too large");
    const dim3 grid((unsigned)tiles, (unsigned)(n_bscans * channels));
    const bool vec4 = zpool == 4 && z0 % 4 == 0 && x_stride % 4 == 0 && (reinterpret_cast<uintptr_t>(mag) & 15) == 0;
    if (vec4)
        hipLaunchKernelGGL(oct::logmag_crop_kernel<true>, grid, dim3(oct::kCnnThreads), 0,
                           reinterpret_cast<hipStream_t>(stream), mag, x_stride, n_x, x0, z0, zpool, height, width,
                           offset_db, 1.0f / scale_db, out, channels, n_bscans);
    else
        hipLaunchKernelGGL(oct::logmag_crop_kernel<false>, grid, dim3(oct::kCnnThreads), 0,
                           reinterpret_cast<hipStream_t>(stream), mag, x_stride, n_x, x0, z0, zpool, height, width,
                           offset_db, 1.0f / scale_db, out, channels, n_bscans);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_depth_median(const float* mag, int n_bscans, int n_x, long long x_stride, int n_z, double* out, void* stream) {
    if (n_bscans <= 0) return OCT_OK;
    if (!mag || !out) return fail(OCT_EINVAL, "null argument");
    if (n_x <= 0 || n_x > oct::kMedMaxX || n_z <= 0 || n_z > x_stride)
        return fail(OCT_EINVAL, "median shape (%d x %d, stride %lld; at most %d A-scans)", n_x, n_z, x_stride, oct::kMedMaxX);
    if (n_bscans > 65535) return fail(OCT_EINVAL, "at most 65535 B-scans per call");
    const size_t lds = sizeof(float) * oct::kMedZ * (size_t)(n_x + 1);
    hipLaunchKernelGGL(oct::depth_median_kernel<false>, dim3((n_z + oct::kMedZ - 1) / oct::kMedZ, n_bscans),
                       dim3(oct::kCnnThreads), lds, reinterpret_cast<hipStream_t>(stream), mag, x_stride, n_x, n_z,
                       (const int*)nullptr, (const int*)nullptr, (const int*)nullptr, (const int*)nullptr, out);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

int oct_depth_median_opc(const float* mag, int n_bscans, int n_x, long long x_stride, int n_z, const int32_t* frame_idx,
                         const int32_t* shift, const int32_t* valid_lo, const int32_t* valid_hi, double* out,
                         void* stream) {
    if (n_bscans <= 0) return OCT_OK;
    if (!mag || !out || !frame_idx || !shift || !valid_lo || !valid_hi) return fail(OCT_EINVAL, "null argument");
    if (n_x <= 0 || n_x > oct::kMedMaxX || n_z <= 0 || n_z > x_stride)
        return fail(OCT_EINVAL, "median shape (%d x %d, stride %lld; at most %d A-scans)", n_x, n_z, x_stride, oct::kMedMaxX);
    if (n_bscans > 65535) return fail(OCT_EINVAL, "at most 65535 B-scans per call");
    const size_t lds = sizeof(float) * oct::kMedZ * (size_t)(n_x + 1);
    hipLaunchKernelGGL(oct::depth_median_kernel<true>, dim3((n_z + oct::kMedZ - 1) / oct::kMedZ, n_bscans),
                       dim3(oct::kCnnThreads), lds, reinterpret_cast<hipStream_t>(stream), mag, x_stride, n_x, n_z,
                       frame_idx, shift, valid_lo, valid_hi, out);
    HIP_TRY(hipGetLastError());
    return OCT_OK;
}

}  // extern "C"
