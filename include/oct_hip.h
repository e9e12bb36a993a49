/*
 * C ABI of the MI355X (gfx950) OCT reconstruction library, liboct_hip.so.
 *
 * Plain pointers and sizes only.  Every device pointer is a HIP device allocation on the
 * plan's device; `stream` is a hipStream_t (NULL = default stream).  Calls enqueue work and
 * return without synchronising the host, like the reference's backend calls between its
 * synchronize() points (octrecon/pipeline.py:276, 281, 285).  Every function returns 0 on
 * success or a negative status; oct_last_error() gives the message of the calling thread's
 * last failure.  A call on an empty batch (zero files, B-scans, rows, planes or frames) returns
 * 0 before looking at its pointers (an empty tensor's may be NULL) and launches nothing.
 *
 * Reference interfaces these replace (atesfet/oct_recon_AF_v1):
 *   oct_plan_create / oct_plan_destroy
 *       SpectralProcessor.__init__ uploading sinc_idx/sinc_w/window to the device
 *       (octrecon/core/spectral.py:91-110; octrecon/core/torch_backend.py:69-85)
 *   oct_background
 *       background rows of the pre-FFT stage (octrecon/core/cpu_kernels.py:20-25,
 *       octrecon/core/gpu_kernels.py:57; frame_mean: octrecon/core/spectral.py:54-66)
 *   oct_spectral (kind OCT_OUT_MAG)
 *       SpectralProcessor.magnitude (octrecon/core/spectral.py:136-153), i.e. the fused pre-FFT
 *       kernel (octrecon/core/gpu_kernels.py:54-68) + cupy.fft.ifft + abs(.[..., :N/2])
 *   oct_spectral (kind OCT_OUT_COMPLEX)   SpectralProcessor.complex_scan (spectral.py:155-158)
 *   oct_spectral (kind OCT_OUT_LINEARISED) SpectralProcessor.linearise (spectral.py:118-134)
 *   oct_stitch
 *       TileStitcher.contribute + the accumulation into num/den
 *       (octrecon/core/stitching.py:122-140; octrecon/pipeline.py:282-284)
 *   oct_stitch_row + oct_volume_clim
 *       the whole y-tile row loop of Reconstructor.process_row fused: every tile's contribute,
 *       the num/den accumulation and the finalise (octrecon/pipeline.py:261-299)
 *   oct_finalize
 *       finalise of a y-tile row + plane/global clim (octrecon/pipeline.py:296-300, 376-388;
 *       octrecon/io/tiff_writer.py:48-50)
 *   oct_quantize
 *       db_to_bits with optional legacy double quantisation (octrecon/io/tiff_writer.py:26-45, 62-65)
 */
#ifndef OCT_HIP_H
#define OCT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OCT_OK = 0, OCT_EINVAL = -1, OCT_EHIP = -2, OCT_EUNSUPPORTED = -3 };
enum { OCT_RAW_I16 = 0, OCT_RAW_U16 = 1, OCT_RAW_F32 = 2 };
enum { OCT_APOD_LINES = 0, OCT_APOD_FRAME_MEAN = 1 };
enum { OCT_OUT_MAG = 0, OCT_OUT_COMPLEX = 1, OCT_OUT_LINEARISED = 2 };

typedef struct oct_plan oct_plan;

/* Upload the spectral operators (host arrays) to device `device` and build the transform's twiddle
 * tables there (synchronous: call outside any graph capture).
 *   n_lambda     spectral pixels N (2048)
 *   taps         tap window width T of band_w (12 or 16)
 *   band_start   (N) int32: output i uses source samples band_start[i] .. band_start[i]+T-1
 *                (0 <= band_start[i] <= N-T)
 *   band_w       (N, T) float32 weights (0 outside the sinc band)
 *   window       (N, 2) float32 complex window Hann/rms * exp(-i beta (k - k0)^2), unscaled */
int oct_plan_create(oct_plan** plan, int device, int n_lambda, int taps, const int32_t* band_start,
                    const float* band_w, const float* window);
int oct_plan_destroy(oct_plan* plan);

/* Device workspace (bytes) for the background rows of `n_files` raw files. */
size_t oct_spectral_workspace_bytes(const oct_plan* plan, int n_files);

/* Background row per raw file into `bg` (n_files, N) float32:
 *   apod_mode OCT_APOD_LINES: mean of the first apod_size A-lines of the file;
 *   OCT_APOD_FRAME_MEAN: mean of all A-lines of each group of apod_group files. */
int oct_background(const oct_plan* plan, const void* raw, int raw_dtype, int n_files, int lines, int apod_mode,
                   int apod_size, int apod_group, float* bg, void* stream);

/* raw (n_files, lines, N) of raw_dtype -> out, with background rows `bg` from oct_background.
 *   apod_mode OCT_APOD_LINES: the first apod_size A-lines of each file are skipped, the other
 *             lines - apod_size are scan A-lines; OCT_APOD_FRAME_MEAN: every line is a scan A-line.
 *   out_kind OCT_OUT_MAG: float32 |ifft(eq * window)| rows [row0, row0+rows) of N/2;
 *            OCT_OUT_COMPLEX: complex64 (float pairs) of the same rows;
 *            OCT_OUT_LINEARISED: float32 equispaced interferogram, N samples (row0/rows ignored).
 *   out_stride: elements (float, or complex for OCT_OUT_COMPLEX) between consecutive A-scans.
 *   raw must be 16-byte aligned. */
int oct_spectral(const oct_plan* plan, const void* raw, int raw_dtype, int n_files, int lines,
                 int apod_mode, int apod_size, int out_kind, int row0, int rows, void* out,
                 long long out_stride, const float* bg, void* stream);

/* One tile's stitching contribution (see oct_stitch_desc). */
typedef struct oct_stitch_desc {
    const float* mag;         /* magnitude rows [mag_row0, mag_row0 + mag_rows) per A-scan */
    long long mag_stride;     /* floats between A-scans */
    int mag_row0, mag_rows;
    int n_x;                  /* A-scans per frame */
    int n_z;                  /* rows of the tile depth grid (N/2) */
    const int32_t* frame_idx; /* (B) tile-local y frame of each batch entry */
    const int32_t* shift;     /* (nY, n_x) optical-path row shift, or NULL */
    const int32_t* valid_lo;  /* (nY, n_x) first valid row; NULL when shift is NULL */
    const int32_t* valid_hi;  /* (nY, n_x) last valid row */
    const float* fz;          /* (n_z) focus weight per tile row */
    const int32_t* z_rows;    /* (n_out_z, 2) tile rows of the two z interpolation taps */
    const float* z_w;         /* (n_out_z, 2) */
    const int32_t* x_cols;    /* (C, 2) tile columns of the two x interpolation taps */
    const float* x_w;         /* (C, 2) */
    float* num;               /* element (b, zo, c) at [(b+out_b0)*n_out_z*ldx + zo*ldx + out_c0 + c] */
    float* den;
    int B, n_out_z, C;
    long long ldx;
    int out_b0, out_c0;
    int accumulate;           /* 1: +=, 0: = */
} oct_stitch_desc;

int oct_stitch(const oct_stitch_desc* d, void* stream);

/* ---- fused y-row stitching + finalise: every tile of one y-tile row in one output-owned
 * pass (num/den accumulate in registers, never in memory), written as dB planes
 * db[(plane0 + f) * n_out_z * n_out_x + zo * n_out_x + c] with the same per-block finite
 * min/max partials as oct_finalize (partial: (planes, oct_stitch_row_blocks(), 2) ints).
 * Equivalent to the per-tile oct_stitch accumulation followed by oct_finalize
 * (octrecon/pipeline.py:261-299).  Each tile's magnitudes hold all n_frames frames. */
typedef struct oct_row_tile {
    const float* mag;         /* (n_frames, n_x, mag_stride): rows [mag_row0, mag_row0 + mag_stride) */
    long long mag_stride;
    int mag_row0;
    int c0, C;                /* output columns [c0, c0 + C) */
    const int32_t* z_rows;    /* (n_out_z, 2) */
    const float* z_w;
    const int32_t* x_cols;    /* (C, 2) */
    const float* x_w;
    const float* fz;          /* (n_z) */
} oct_row_tile;
int oct_stitch_row_blocks(int n_out_z, int n_out_x);
int oct_stitch_row(const oct_row_tile* tiles, int n_tiles, int n_frames, int n_x, int n_z, const int32_t* frame_idx,
                   const int32_t* shift, const int32_t* valid_lo, const int32_t* valid_hi, int n_out_z, int n_out_x,
                   float thr, float* db, int plane0, int* partial, void* stream);
/* The same for rows of any number of tiles (oct_stitch_row takes at most 40 in its argument
 * block): dev_tiles is a device copy of the host table `tiles` (which is validated), read by the
 * kernel. */
int oct_stitch_row_table(const oct_row_tile* tiles, const oct_row_tile* dev_tiles, int n_tiles, int n_frames, int n_x,
                         int n_z, const int32_t* frame_idx, const int32_t* shift, const int32_t* valid_lo,
                         const int32_t* valid_hi, int n_out_z, int n_out_x, float thr, float* db, int plane0,
                         int* partial, void* stream);
/* plane clims (n_planes, 2) and the global clim (2) from row partials (8-byte aligned) */
int oct_volume_clim(const int* partial, int n_planes, int blocks_per_plane, float* plane_clim, float* clim,
                    void* stream);

/* num/den (n_planes, plane_elems) -> db (same layout) with den < thr -> NaN, 20 log10(num/den);
 * with den == NULL, db = 20 log10(num) (a magnitude volume without stitching);
 * plane_clim (n_planes, 2) and clim (2) = finite min / max (NaN if none). */
size_t oct_finalize_workspace_bytes(int n_planes, long long plane_elems);
int oct_finalize(const float* num, const float* den, float* db, int n_planes, long long plane_elems,
                 float thr, float* plane_clim, float* clim, void* workspace, size_t workspace_bytes,
                 void* stream);

/* db (n_planes, plane_elems) -> uint16 codes with the device clim (2 floats); legacy_double
 * first quantises each plane with plane_clim (n_planes, 2) and dequantises it. */
int oct_quantize(const float* db, uint16_t* bits, int n_planes, long long plane_elems, const float* clim,
                 const float* plane_clim, int legacy_double, void* stream);

/* ---- A-scan binning and repeat averaging (octrecon/core/spectral.py:20-46 bin_ascans /
 * binned_columns; :168-177 average_repeats)
 * oct_binned_lines: scan lines kept by AScanBinning of n_scan_lines lines.
 * oct_bin_lines: raw (n_files, lines, 2048) -> out (n_files, keep_lines + kept, 2048) float32: the
 * first keep_lines lines copied, then the binned scan lines [skip_lines, lines) (filter2 with
 * ones(1,B)/B, zero padded, then every B-th line from max(1, B/2) - 1), float32 products summed
 * in order.
 * oct_average_repeats: mag (n_frames bscan_avg, size_x ascan_avg, rows) -> out (n_frames,
 * size_x, rows): mean over the B-scan repeats, then over the A-scan repeats (float32). */
/* ---- raw-file ingest: Thorlabs SRR payload (n_files, file_lines, n_lambda) uint16 -> Thorlabs file
 * layout (n_files, (apod_end - apod_start) + (scan_end - scan_start), n_lambda) int16: the low 12
 * bits of each sample (legacy rem 4096), background A-lines [apod_start, apod_end) first, then the
 * scan A-lines [scan_start, scan_end) (octrecon/io/srr.py:107-124, SRRTileReader.read_file /
 * read_bscans).  n_lambda a multiple of 8, buffers 16-byte aligned. */
int oct_srr_unpack(const uint16_t* in, int n_files, int file_lines, int n_lambda, int apod_start, int apod_end,
                   int scan_start, int scan_end, int16_t* out, void* stream);
/* ---- PackBits (TIFF compression 32773) of uint16 code rows for the legacy BigTIFF output
 * (octrecon/io/tiff_writer.py:53-71, which has tifffile compress each page): every row is one
 * strip, encoded as runs over consecutive 128-byte chunks of its little-endian bytes (a uniform
 * chunk -> replicate run, any other -> literal run).  oct_packbits_chunks: chunks per row;
 * oct_packbits_sizes: bytes of every chunk (rows x chunks int32); oct_packbits_write: the runs at
 * the exclusive prefix sum chunk_offsets (int64, same shape) of those sizes. */
int oct_packbits_chunks(int row_elems);
int oct_packbits_sizes(const uint16_t* codes, long long rows, int row_elems, int32_t* chunk_bytes, void* stream);
int oct_packbits_write(const uint16_t* codes, long long rows, int row_elems, const int64_t* chunk_offsets,
                       uint8_t* out, void* stream);
int oct_binned_lines(int n_scan_lines, int binning);
int oct_bin_lines(const void* raw, int raw_dtype, int n_files, int lines, int keep_lines, int skip_lines,
                  int binning, float* out, void* stream);
int oct_average_repeats(const float* mag, float* out, int n_frames, int bscan_avg, int size_x, int ascan_avg, int rows,
                        long long in_stride, long long out_stride, void* stream);

/* ---- dispersion sharpness search (octrecon/estimation/dispersion.py:46-107, 154-239)
 * One search evaluation is the spectral path with the candidate's window
 * hann/rms * exp(-i beta (k - k0)^2).  oct_plan_use_window makes later oct_spectral launches
 * with this plan read `dev_window` ((N, 2) float32 device array, 16-byte aligned, already
 * multiplied by 1/N, in the pair order below) instead of the plan's own window; NULL restores it.
 * The pointer is captured at each launch, so a batch of candidates can be queued without host
 * synchronisation.
 * Pair order (N = 2048): element i = l + 64 q (l < 64, q < 32) of a device window is stored at
 * complex position 2 ((q mod 16) 64 + l) + q / 16, so that the transform's lane l reads its
 * elements q and q + 16 with one 16-byte load (oct_dispersion_windows writes this order;
 * spectral.pair_order converts a natural-order window).  oct_plan_create takes its host window
 * in natural order.
 * oct_topfrac_metric: out[b] = log(mean of the k = max(1, int(M*frac)) largest |A|^2 over all
 * x and z >= z_min), M = n_x (n_z - z_min); mag (n_bscans, n_x, n_z) float32, out float64;
 * exact (radix selection), deterministic, workspace of oct_topfrac_metric_workspace_bytes. */
int oct_plan_use_window(oct_plan* plan, const float* dev_window);
/* Per-file candidates: file b of later oct_spectral launches uses the window at
 * dev_windows + b * file_stride floats (file_stride 0: one window for every file, else a multiple
 * of 4 and >= 2N; dev_windows 16-byte aligned; each window in the pair order).
 * Used by the batched per-B-scan search (each B-scan its own bracket), which the reference runs
 * one B-scan at a time (octrecon/estimation/dispersion.py:226-239). */
int oct_plan_use_windows(oct_plan* plan, const float* dev_windows, long long file_stride, long long repeat_stride);
/* oct_spectral over the batch n_repeat times in one launch: virtual file r n_files + b reads raw
 * file b and the window dev_windows + r repeat_stride + b file_stride (oct_plan_use_windows), and
 * writes out rows of A-scan (r n_files + b) n_x + x; only scan lines [x0, x0 + n_x) of every file
 * (n_x < 0: all); bg as computed by oct_background for the batch.  Used to reconstruct one batch
 * at several candidate dispersions at once. */
int oct_spectral_repeat(const oct_plan* plan, const void* raw, int raw_dtype, int n_files, int lines, int apod_mode,
                        int apod_size, int n_repeat, int x0, int n_x, int out_kind, int row0, int rows, void* out,
                        long long out_stride, const float* bg, void* stream);
size_t oct_topfrac_metric_workspace_bytes(int n_bscans, int n_x, int n_z, int z_min);
int oct_topfrac_metric(const float* mag, int n_bscans, int n_x, int n_z, int z_min, double frac, double* out,
                       void* workspace, size_t workspace_bytes, void* stream);
/* ---- the other sharpness metrics of the search (octrecon/estimation/dispersion.py:46-63 _metric)
 * over I = |A|^2 (squared in float32, widened to float64) for all x and z >= z_min of each B-scan:
 *   OCT_METRIC_TOP_PERCENTILE  log mean of the brightest 1 % (oct_topfrac_metric with frac 0.01)
 *   OCT_METRIC_KURTOSIS        sum(I^2) / sum(I)^2
 *   OCT_METRIC_ENTROPY         sum(p log(p + 1e-30)), p = I / sum(I) (the negative entropy)
 *   OCT_METRIC_LEGACY_PEAK     max over z of the mean over x of log(|A| + 1e-12)
 * out (n_bscans) float64; float64 sums with per-workgroup partials added in a fixed order, so
 * results are deterministic; workspace of oct_sharpness_metric_workspace_bytes. */
enum { OCT_METRIC_TOP_PERCENTILE = 0, OCT_METRIC_KURTOSIS = 1, OCT_METRIC_ENTROPY = 2, OCT_METRIC_LEGACY_PEAK = 3 };
size_t oct_sharpness_metric_workspace_bytes(int metric, int n_bscans, int n_x, int n_z, int z_min);
int oct_sharpness_metric(int metric, const float* mag, int n_bscans, int n_x, int n_z, int z_min, double* out,
                         void* workspace, size_t workspace_bytes, void* stream);
/* Exact order statistics per B-scan (the percentile and median of the B-scan signal score,
 * octrecon/estimation/dispersion.py:145-150): out[b n_ranks + r] = the ranks[b n_ranks + r]-th
 * largest (1-based, 1 <= rank <= n_x (n_z - z_min)) of the values over x and z >= z_min of B-scan
 * b, in float64: |A|^2 squared in float32 when square != 0, else |A| (whose order is that of
 * the float64 squares).  ranks: device int64 (n_bscans, n_ranks).  Radix selection as
 * oct_topfrac_metric; workspace of oct_order_stats_workspace_bytes. */
size_t oct_order_stats_workspace_bytes(int n_bscans, int n_ranks, int n_x, int n_z, int z_min);
int oct_order_stats(const float* mag, int n_bscans, int n_x, int n_z, int z_min, int square, const long long* ranks,
                    int n_ranks, double* out, void* workspace, size_t workspace_bytes, void* stream);
/* No counterpart in the reference: oct_topfrac_metric for callers that cannot allocate device
 * memory (code under HIP graph capture, tight search loops, harnesses that forbid hipMalloc).
 * The same kernels and results as oct_topfrac_metric, with the workspace taken from a 4 MiB
 * device array that the library holds on each GPU until the process ends (enough for about 500
 * B-scans per call); a call whose oct_topfrac_metric_workspace_bytes exceeds 4 MiB fails with
 * OCT_EINVAL.  It allocates nothing and uses the device current at the call (the stream must
 * belong to it).  All calls on one GPU share that array: calls that may overlap on one GPU
 * (different streams, or threads) race and must be serialised by the caller. */
int oct_topfrac_metric_noalloc(const float* mag, int n_bscans, int n_x, int n_z, int z_min, double frac, double* out,
                               void* stream);
/* Candidate windows out (count, n, 2) float32 (n = 2048; out 16-byte aligned), each in the pair
 * order of oct_plan_use_window: element j = hann_over_n[j] exp(-i beta_b dk2[j]) in float64, with
 * beta_b = ((head ? head[b * head_stride] * head_scale : betas[b]) + offset) * to_ref (device
 * arrays; `head` lets a network's output set the candidates without a host round trip). */
int oct_dispersion_windows(const double* betas, const float* head, int head_stride, double head_scale, double offset,
                           double to_ref, int count, const double* hann_over_n, const double* dk2, int n, float* out,
                           void* stream);

/* ---- operator (MFMA) form: the whole linear spectral map as one complex operator
 * A = F_crop diag(window / N) S' (N/2-point crop [z0, z0+nzc), S' = banded sinc + rank-1 mean
 * term), built on the device in float64, stored as fp16 hi/lo scaled by 2^scale_log2, and
 * applied as a split-precision (3-product) fp16 GEMM with fp32 accumulation on MFMA.
 * band_w here is float64 (N, taps).  op64_host (optional, (nzc, N, 2) doubles) receives the
 * float64 operator.  Replaces, for magnitudes, the same reference calls as oct_spectral. */
typedef struct oct_opplan oct_opplan;
int oct_opplan_create(oct_opplan** plan, int device, int n_lambda, int taps, const int32_t* band_start,
                      const double* band_w, const double* window, int z0, int nzc, int* scale_log2,
                      double* op64_host);
int oct_opplan_destroy(oct_opplan* plan);
/* Rebuild the operator in place, on `stream`, for another window (N, 2) float64 device array
 * (NULL: the creation window): the same float64 build and fp16 hi/lo split, with the scale
 * 2^scale_log2 fixed at creation, so |window| must not exceed oct_opplan_scale_bound's value
 * (a dispersion change alters only the phase).  Launches stay in stream order with the plan's
 * oct_opspectral calls; calls on one plan must use one stream.  Used to reconstruct each tile
 * with its own dispersion (the reference applies one dispersion per volume,
 * octrecon/pipeline.py:155-190). */
int oct_opplan_set_window(oct_opplan* plan, const double* window, void* stream);
int oct_opplan_scale_bound(const oct_opplan* plan, double* max_window_abs);
/* Rows of the fp16 operator oct_opplan_create allocates and the GEMM applies for nzc bins (32-row
 * tiles of 16 bins' real and imaginary parts, grouped 1-4 per GEMM workgroup; the grouping into
 * *row_tiles when not NULL).  Host arithmetic only: no device is touched.  Returns the row count,
 * or OCT_EINVAL for nzc < 1.  No counterpart in the reference (a sizing query of this library). */
int oct_opplan_rows(int nzc, int* row_tiles);
/* magnitudes |A (raw - bg)| of bins [z0, z0+nzc) -> out (M, out_stride) float32 */
int oct_opspectral(const oct_opplan* plan, const void* raw, int raw_dtype, int n_files, int lines, int apod_mode,
                   int apod_size, float* out, long long out_stride, const float* bg, void* stream);

/* ---- focus detector input (octrecon/estimation/focus.py:209-218, 221-279: the depth profile
 * np.median over x of |A| per B-scan, float32 magnitudes compared in float64)
 * out[b, z] = median over x of mag[b, x, z] (float64 mean of the two middle values for even n_x),
 * mag rows z in [0, n_z) of A-scans spaced x_stride floats; n_x <= 1024. */
int oct_depth_median(const float* mag, int n_bscans, int n_x, long long x_stride, int n_z, double* out, void* stream);
/* The same after the optical-path correction (octrecon/estimation/focus.py:110-129): B-scan b is
 * tile frame frame_idx[b] (device int32); row z of column x reads row clip(z + shift[f, x]) and
 * exists only for valid_lo[f, x] <= z <= valid_hi[f, x] ((n_frames, n_x) int32, as the stitcher's
 * optical-path tables); the median is over the rows that exist (numpy.nanmedian), NaN if none. */
int oct_depth_median_opc(const float* mag, int n_bscans, int n_x, long long x_stride, int n_z, const int32_t* frame_idx,
                         const int32_t* shift, const int32_t* valid_lo, const int32_t* valid_hi, double* out,
                         void* stream);

/* ---- learned beta / focus estimator (no counterpart in the reference: it replaces the searches
 * of octrecon/estimation/dispersion.py:154-267 and focus.py:191-218, which stay the labels, the
 * referee and the fallback).  bfloat16 activations and weights (uint16_t bit patterns, round to
 * nearest even), float32 master weights, gradients and accumulation.
 * Network input: out[b, c, h, w] (n_bscans, channels, height, width) bf16 from magnitude images
 * v = c n_bscans + b of mag (channels n_bscans, n_x, rows):
 *   clamp((20 log10(max_{j<zpool} mag[v, x0 + w, z0 + h zpool + j]) - offset_db) / scale_db, -8, 8). */
int oct_logmag_crop(const float* mag, int n_bscans, int n_x, long long x_stride, int n_rows, int x0, int z0, int zpool,
                    int height, int width, float offset_db, float scale_db, uint16_t* out, int channels, void* stream);
/* conv 3x3 / stride 2 / pad 1 + ReLU: x (n, ci, h, w) -> y (n, co, h/2, w/2); w (co, ci, 3, 3); co % 16 == 0 */
int oct_cnn_conv_fwd(const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y, int n, int ci, int h,
                     int wd, int co, void* stream);
/* The same layer on the matrix cores (bf16 products, float32 sums) where ci >= 8 and w/2 >= 16,
 * the FP32-VALU kernels of oct_cnn_conv_fwd elsewhere: the inference forward.  Same arguments and
 * errors; outputs within one bf16 ulp of oct_cnn_conv_fwd's (the sums' order differs). */
int oct_cnn_conv_fwd_mfma(const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y, int n, int ci, int h,
                          int wd, int co, void* stream);
/* dx = conv^T(dpre) (n, ci, h, w) float32, times (xmask > 0) when xmask is given; ci % 16 == 0 */
int oct_cnn_conv_bwd_data(const float* dpre, const uint16_t* w, const uint16_t* xmask, float* dx, int n, int ci, int h,
                          int wd, int co, void* stream);
size_t oct_cnn_conv_bwd_weight_workspace_bytes(int n, int ci, int h, int wd, int co);
/* dw (co, ci, 3, 3) and db (co) float32 from dpre (n, co, h/2, w/2) and x (n, ci, h, w) */
int oct_cnn_conv_bwd_weight(const float* dpre, const uint16_t* x, float* dw, float* db, int n, int ci, int h, int wd,
                            int co, void* workspace, size_t workspace_bytes, void* stream);
/* The two backward convolutions on the matrix cores (bf16 products, float32 sums; dpre enters as
 * two bf16 parts that hold it to 2^-16): the training backward.  Same arguments and results as
 * oct_cnn_conv_bwd_data / oct_cnn_conv_bwd_weight to float32 rounding, with a workspace (16-byte
 * aligned, size from the *_workspace_bytes query, 0 = unsupported shape).
 * bwd_weight: co % 16 == 0; im2col of x, a split GEMM over the positions, a reduction.
 * bwd_data: ci % 16 == 0 and ci <= 64, co % 8 == 0; one GEMM per input-pixel parity class. */
size_t oct_cnn_conv_bwd_weight_mfma_workspace_bytes(int n, int ci, int h, int wd, int co);
int oct_cnn_conv_bwd_weight_mfma(const float* dpre, const uint16_t* x, float* dw, float* db, int n, int ci, int h,
                                 int wd, int co, void* workspace, size_t workspace_bytes, void* stream);
size_t oct_cnn_conv_bwd_data_mfma_workspace_bytes(int n, int ci, int h, int wd, int co);
int oct_cnn_conv_bwd_data_mfma(const float* dpre, const uint16_t* w, const uint16_t* xmask, float* dx, int n, int ci,
                               int h, int wd, int co, void* workspace, size_t workspace_bytes, void* stream);
/* fully connected y = x w^T + b: exactly one of y_relu_bf16 (ReLU, bf16) or y_linear (float32) */
int oct_cnn_fc_fwd(const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y_relu_bf16, float* y_linear,
                   int n, int k, int o, void* stream);
int oct_cnn_fc_bwd_data(const float* dpre, const uint16_t* w, const uint16_t* xmask, float* dx, int n, int k, int o,
                        void* stream);
int oct_cnn_fc_bwd_weight(const float* dpre, const uint16_t* x, float* dw, float* db, int n, int k, int o,
                          void* stream);
/* heteroscedastic Gaussian NLL over head (n, 2T) = (means, log variances), targets (n, T), T <= 4;
 * loss[0] = mean loss, loss[1 + j] = mean squared error of target j; dhead (n, 2T) */
int oct_cnn_gauss_nll(const float* head, const float* target, const float* weights, float* loss, float* dhead, int n,
                      int n_targets, void* stream);
/* Adam with bias correction and decoupled weight decay; params_bf16 receives the rounded weights */
int oct_cnn_adam(float* params, const float* grads, float* m, float* v, uint16_t* params_bf16, long long count,
                 float lr, float beta1, float beta2, float eps, float weight_decay, int step, void* stream);

const char* oct_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
