/*
 * ORACLE - test infrastructure only.  Nothing in the product path links or calls this file;
 * only tests, the smoke check and the CPU-baseline leg of the benchmark load it, as the
 * checker.
 *
 * C restatement of the reference's CPU hot loop, octrecon/core/cpu_kernels.py:12-41
 * (Numba `fused_pre_fft`), keeping its arithmetic types:
 *   - background  apod[b,j] = (sum_a raw[b,a,j] in float64) / apod_size, stored float32 (:20-25)
 *   - d[j] = raw - apod[b,j] in float32; line mean m accumulated in float64 (:30-36)
 *   - d[j] -= m, stored float32 (:37-38)
 *   - acc = sum_t float32(w[t,i] * d[idx[t,i]]) accumulated in float64 (:39-42)
 *   - out = (acc + m) * window[i] in complex128, stored complex64 (:43)
 * Parallel over A-lines with OpenMP, as the reference's prange over B*nX (:26).
 * Build with -ffp-contract=off so that no multiply-add is fused across those roundings.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define DEFINE_FUSED(NAME, RAW_T)                                                           \
    int NAME(const RAW_T *raw, int B, int interf, int N, int apod_size,                     \
             const int64_t *idx, const float *w, int T, const float *win /* (N,2) */,       \
             float *out /* (B, nX, N, 2) */)                                                \
    {                                                                                      \
        if (apod_size <= 0 || apod_size >= interf || N <= 0 || T <= 0) return -1;          \
        int nX = interf - apod_size;                                                       \
        float *apod = (float *)malloc(sizeof(float) * (size_t)B * N);                     \
        if (!apod) return -2;                                                              \
        _Pragma("omp parallel for schedule(static)")                                       \
        for (int b = 0; b < B; ++b) {                                                      \
            for (int j = 0; j < N; ++j) {                                                  \
                double s = 0.0;                                                            \
                for (int a = 0; a < apod_size; ++a)                                        \
                    s += (double)raw[((size_t)b * interf + a) * N + j];                    \
                apod[(size_t)b * N + j] = (float)(s / apod_size);                          \
            }                                                                              \
        }                                                                                  \
        int fail = 0;                                                                      \
        _Pragma("omp parallel")                                                            \
        {                                                                                  \
            float *d = (float *)malloc(sizeof(float) * (size_t)N);                         \
            if (!d) {                                                                      \
                _Pragma("omp atomic write") fail = 1;                                      \
            } else {                                                                       \
                _Pragma("omp for schedule(static)")                                        \
                for (long p = 0; p < (long)B * nX; ++p) {                                  \
                    int b = (int)(p / nX);                                                 \
                    int x = (int)(p - (long)b * nX);                                       \
                    const RAW_T *r = raw + ((size_t)b * interf + apod_size + x) * N;       \
                    const float *ap = apod + (size_t)b * N;                                \
                    double m = 0.0;                                                        \
                    for (int j = 0; j < N; ++j) {                                          \
                        d[j] = (float)r[j] - ap[j];                                        \
                        m += (double)d[j];                                                 \
                    }                                                                      \
                    m = m / N;                                                             \
                    for (int j = 0; j < N; ++j) d[j] = (float)((double)d[j] - m);          \
                    float *o = out + (size_t)p * N * 2;                                    \
                    for (int i = 0; i < N; ++i) {                                          \
                        double acc = 0.0;                                                  \
                        for (int t = 0; t < T; ++t) {                                      \
                            float prod = w[(size_t)t * N + i] * d[idx[(size_t)t * N + i]]; \
                            acc += (double)prod;                                           \
                        }                                                                  \
                        double v = acc + m;                                                \
                        o[2 * i] = (float)(v * (double)win[2 * i]);                        \
                        o[2 * i + 1] = (float)(v * (double)win[2 * i + 1]);                \
                    }                                                                      \
                }                                                                          \
                free(d);                                                                   \
            }                                                                              \
        }                                                                                  \
        free(apod);                                                                        \
        return fail ? -2 : 0;                                                              \
    }

DEFINE_FUSED(oracle_fused_pre_fft_i16, int16_t)
DEFINE_FUSED(oracle_fused_pre_fft_u16, uint16_t)
DEFINE_FUSED(oracle_fused_pre_fft_f32, float)

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
