/*
 * avgrid.h -- C ABI of libavgrid.so, the MI355X (gfx950) standard gridder.
 *
 * Every pointer argument is a DEVICE pointer unless stated otherwise; sizes
 * are int64_t/int32_t; `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  Every function returns AV_OK (0) or a negative status; the
 * message of the last failure on the calling thread is av_last_error().
 * No function allocates device memory except av_plan_create.
 *
 * Reference paths are relative to src/astroviper/ of casangi/astroviper
 * 0.0.34; each entry point names the reference function it replaces.
 */
#ifndef AVGRID_H
#define AVGRID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AV_OK 0
#define AV_ERR_INVALID (-1)     /* bad argument, shape or size */
#define AV_ERR_HIP (-2)         /* HIP runtime error */
#define AV_ERR_NOMEM (-3)       /* device allocation failed */
#define AV_ERR_UNSUPPORTED (-4) /* valid but not implemented combination */

/* element types */
#define AV_F32 1 /* float32 weights / complex64 visibilities and grids */
#define AV_F64 2 /* float64 weights / complex128 visibilities and grids */

/* av_grid mode bits */
#define AV_GRID_PSF 1         /* grid the weights instead of vis*weight */
#define AV_GRID_ACCUMULATE 2  /* add into grid/sumwt instead of overwriting */
#define AV_GRID_COMPENSATED 4 /* FP32 only: compensated per-tile sums */
#define AV_GRID_ADJOINT 8     /* weight each sample by 1/norm (adjoint of
                                 av_degrid); weights/flags are ignored */
#define AV_GRID_SUBTILE 16    /* FP32 16 x 16 tiles: run the sub-tile
                                 scheduled 4x4x1 matrix-core kernel instead of
                                 the dense 16x16x4 one (same results) */
#define AV_GRID_F32_PRODUCTS 32 /* FP32 16 x 16 tiles: form every product in
                                   FP32 (16x16x4 FP32 matrix-core kernel)
                                   instead of from split FP16 operands */
/* av_degrid mode bits */
#define AV_DEGRID_INCREMENTAL 1 /* add into vis_out instead of overwriting */
#define AV_DEGRID_F32_PRODUCTS 2 /* FP32 16 x 16 tiles: FP32 products (16x16x4
                                    kernel) instead of split FP16 operands */

/* av_plan_create geometry */
#define AV_GEOM_REFERENCE 0 /* grid uses standard_grid_jit rounding and
                               degrid uses dgrid_numba rounding */
#define AV_GEOM_DEGRID 1    /* both use dgrid_numba rounding: av_grid with
                               AV_GRID_ADJOINT is then the exact adjoint of
                               av_degrid */
#define AV_PLAN_ROW_CHANNELS 16 /* or'ed into geometry: every row is one sample
                                   at its own channel (row_chan), so the plan
                                   holds n_rows samples, not n_rows * n_chan */
#define AV_GEOM_CONJUGATE 32    /* or'ed into AV_GEOM_REFERENCE: every
                                   grid-valid sample is placed at its conjugate
                                   centre int(-u + N/2 + 0.5) with its own
                                   kernel offsets, cells outside the grid
                                   dropped, and gets no degrid record: the
                                   second scatter of do_imaging_weight
                                   (standard_grid.py:330-340, 389-430) */

/* ABI version: bumped whenever a signature changes; the bindings refuse a
 * library whose version differs from the one they were written for. */
#define AV_ABI_VERSION 12

const char *av_last_error(void);
int av_version(void);

/* Device properties used for launch shapes (number of CUs, LDS per CU). */
int av_device_info(int32_t *num_cus, int32_t *lds_bytes_per_cu,
                   int32_t *xcds);

/* prolate_spheroidal_function, core/imaging/imaging_utils/
 * gcf_prolate_spheroidal.py:102-158: grdsf and (1-u^2)*grdsf at n points. */
int av_ps_function(const double *u, int64_t n, double *grdsf, double *kern,
                   void *stream);

/* create_prolate_spheroidal_kernel_1D, gcf_prolate_spheroidal.py:87-99:
 * writes oversampling*(support/2+1) float64 values. */
int av_ps_kernel_1d(int32_t oversampling, int32_t support, double *out,
                    void *stream);

/* One axis of the gridding-correction taper kernel_image of
 * create_prolate_spheroidal_kernel, gcf_prolate_spheroidal.py:72-79:
 * out[i] = grdsf(|2*(i - n/2)/n|), i < n.  The taper is the outer product
 * of the u and v axes. */
int av_ps_correction_1d(int32_t n, double *out, void *stream);

/* The full taper kernel_image of create_prolate_spheroidal_kernel,
 * gcf_prolate_spheroidal.py:72-79: out[i][j] = grdsf(|2*(i - n_u/2)/n_u|) *
 * grdsf(|2*(j - n_v/2)/n_v|), float64 [n_u][n_v]. */
int av_ps_taper(int32_t n_u, int32_t n_v, double *out, void *stream);

/* Outer products of create_prolate_spheroidal_kernel's 4-D kernel,
 * gcf_prolate_spheroidal.py:65-67: k4[x][y][i][j] = k1[x][i]*k1[y][j] for
 * x, y < nfill, zero for nfill <= x,y < nrow.  k1 is [nrow][support]. */
int av_ps_outer(const double *k1, int32_t nrow, int32_t nfill,
                int32_t support, double *k4, void *stream);

/* ---------------------------------------------------------------- plans --
 * A plan holds the binning of one set of uvw coordinates into
 * (uv tile, image channel) bins.  It replaces the per-sample coordinate
 * arithmetic of standard_grid_jit (standard_grid.py:318-352) and
 * dgrid_numba (standard_degrid.py:435-462), computed once on the GPU in FP64
 * and shared by av_grid and av_degrid.
 *
 * n_rows      rows (time x baseline) of uvw
 * n_chan      visibility channels (entries of freq); with row_chan each row
 *             is one sample at one of them
 * n_img_chan  grid channels; n_u, n_v grid axes 2 and 3
 * cell_u/v    pixelincr in radians (signed)
 * support     kernel support in [2, 63]; av_grid and av_degrid take plans of
 *             support <= 16, wider ones bin for av_mosaic_grid
 * tile_u/v    tile shape in pixels (0 = default 32), in [support, 128]
 * geometry    AV_GEOM_REFERENCE or AV_GEOM_DEGRID, optionally or'ed with
 *             AV_PLAN_ROW_CHANNELS (then av_plan_set_uvw needs row_chan);
 *             AV_GEOM_REFERENCE may also be or'ed with AV_GEOM_CONJUGATE
 * unit_cap    records per work unit (0 = default 8192)
 * The plan is device-resident; create it once per (shape, tiling). */
typedef struct av_plan av_plan;

int av_plan_create(int64_t n_rows, int32_t n_chan, int32_t n_img_chan,
                   int32_t n_u, int32_t n_v, double cell_u, double cell_v,
                   int32_t support, int32_t oversampling, int32_t tile_u,
                   int32_t tile_v, int32_t geometry, int32_t unit_cap,
                   av_plan **out);
int av_plan_destroy(av_plan *plan);

/* Bin the samples of uvw [n_rows][3] float64 at frequencies freq [n_chan]
 * (or, with row_chan [n_rows] int32 non-NULL, row r has one sample at
 * frequency freq[row_chan[r]] on image channel chan_map[row_chan[r]]).
 * chan_map [n_chan] int32 maps visibility to image channel (NULL =
 * identity when n_img_chan == n_chan, all zero when n_img_chan == 1).
 * Asynchronous on `stream`; no host synchronisation. */
int av_plan_set_uvw(av_plan *plan, const double *uvw, const double *freq,
                    const int32_t *chan_map, const int32_t *row_chan,
                    void *stream);

/* Replace the degridder's pixel transform pos = scale*uvw*freq/c + offset
 * (standard_degrid.py:435-440).  The default, set by av_plan_create, is
 * scale = -cell*N and offset = N/2 per axis as degrid_spheroid_ms4 passes
 * (standard_degrid.py:572-573); dgrid_numba itself accepts any values.
 * Takes effect at the next av_plan_set_uvw. */
int av_plan_set_degrid_transform(av_plan *plan, double scale_u, double scale_v,
                                 double offset_u, double offset_v);

/* Rows are time-major with n_baseline baselines per time step, as the
 * reference's [time][baseline] loops visit them (standard_grid.py:310-319,
 * standard_degrid.py:416-417); 1 (the default) = no known order.  n_baseline
 * must divide n_rows.  The binning then writes each chunk's records baseline
 * by baseline, which makes its record stores contiguous; the records of a
 * bin are the same, in another order within a chunk.  Takes effect at the
 * next av_plan_set_uvw. */
int av_plan_set_row_order(av_plan *plan, int32_t n_baseline);

/* Counts of the last av_plan_set_uvw (synchronises `stream`):
 * out[0] samples, out[1] grid-valid, out[2] degrid-valid, out[3] records,
 * out[4] bins, out[5] grid units, out[6] degrid units, out[7] zero-list. */
int av_plan_stats(av_plan *plan, int64_t out[8], void *stream);

/* standard_grid_jit, core/imaging/imaging_utils/standard_grid.py:246-432
 * (with the flag masking of grid2image_spheroid_ms4, :664-668).
 * vis [rows][n_chan][n_pol] complex (ignored with AV_GRID_PSF),
 * weight [rows][n_chan][n_pol] real, flag [rows][n_chan][n_pol] uint8 or
 * NULL (a flagged sample counts as weight 0), cgk the float64 kernel table
 * of av_ps_kernel_1d, grid [n_img_chan][n_pol][n_u][n_v] complex,
 * sumwt [n_img_chan][n_pol] float64.  dtype AV_F32 or AV_F64 selects
 * complex64/float32 or complex128/float64 for vis, weight and grid. */
int av_grid(av_plan *plan, const void *vis, const void *weight,
            const uint8_t *flag, int32_t n_pol, const double *cgk,
            void *grid, double *sumwt, int32_t dtype, int32_t mode,
            void *stream);

/* dgrid_numba, core/imaging/imaging_utils/standard_degrid.py:388-485 with
 * zero dphase.  grid [n_img_chan][n_grid_pol][n_u][n_v] complex,
 * pol_map [n_pol] int32 grid polarisation of each visibility polarisation
 * (NULL = p % n_grid_pol), flag [rows][n_chan][n_pol] uint8 or NULL,
 * vis_out [rows][n_chan][n_pol] complex.  Flagged or off-grid samples, the
 * polarisations whose pol_map entry lies outside [0, n_grid_pol) and
 * samples whose kernel-weight sum is zero (a caller's table that vanishes
 * over the footprint) are skipped as dgrid_numba skips them
 * (standard_degrid.py:419-423,482): left unchanged with
 * AV_DEGRID_INCREMENTAL and set to 0 without it. */
int av_degrid(av_plan *plan, const void *grid, int32_t n_grid_pol,
              const int32_t *pol_map, const uint8_t *flag, int32_t n_pol,
              const double *cgk, void *vis_out, int32_t dtype, int32_t mode,
              void *stream);

/* av_plan_create + av_plan_set_uvw + av_grid in one call that allocates no
 * device memory (no counterpart in the reference; for callers that cannot
 * allocate: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The plan's arrays live in a static 32 MiB device
 * array of the library, one per GPU, held until the process ends; calls
 * whose plan needs more are refused with AV_ERR_INVALID (use
 * av_plan_create).  Arguments as av_plan_create (reference geometry, default
 * unit cap), av_plan_set_uvw (chan_map, no row_chan) and av_grid.  Every
 * call on a GPU reuses that GPU's array: calls that overlap on one GPU (other
 * streams or threads) race. */
int av_grid_noalloc(const double *uvw, int64_t n_rows, const double *freq,
                    int32_t n_chan, const int32_t *chan_map, int32_t n_img_chan,
                    int32_t n_u, int32_t n_v, double cell_u, double cell_v,
                    int32_t support, int32_t oversampling, int32_t tile_u,
                    int32_t tile_v, const void *vis, const void *weight,
                    const uint8_t *flag, int32_t n_pol, const double *cgk,
                    void *grid, double *sumwt, int32_t dtype, int32_t mode,
                    void *stream);

/* mosaic_grid_jit, core/imaging/imaging_utils/mosaic_grid.py:306-486: the
 * mosaic / A-projection gridder.  Every sample of the plan (created with
 * AV_GEOM_REFERENCE, support >= max(weight_support), so that each footprint
 * lies in the tiles the plan records, and tiles at least that wide) whose
 * centre passes the reference's bounds test (centre +- max(weight_support)
 * inside the grid, mosaic_grid.py:379-383, applied by the call) is
 * convolved with the complex convolution function cfpg[cfb][cfc][cfp]
 * [cu_n][cv_n] (cfpg holds ncb x ncc x ncp of them) (conv_kernel * phase_gradient[0], mosaic_grid.py:359) of
 * cfb = cf_baseline_map[row % n_baseline], cfc = cf_chan_map[chan],
 * cfp = cf_pol_map[pol], over weight_support[cfb][cfc][cfp][2] taps,
 * oversampled by ov_u, ov_v.  vis [rows][n_chan][n_pol] complex (ignored
 * with AV_GRID_PSF), weight [rows][n_chan][n_pol] (imaging weights), grid
 * [n_img_chan][n_pol][n_u][n_v] complex, sumwt [n_img_chan][n_pol] FP64:
 * weight * Re(norm) with AV_GRID_PSF, weight * Re(norm^2) without.  mode:
 * AV_GRID_PSF, AV_GRID_ACCUMULATE.  cf_*_map, weight_support int32 on the
 * device; cfpg complex64 (AV_F32) or complex128 (AV_F64).  A sample
 * polarisation whose convolution function lies outside cfpg is skipped.  The
 * call allocates a small stream-ordered table of kernel sums (ncb * ncc *
 * ncp * (ov_u + 3) * (ov_v + 3) * 16 bytes) and frees it on the stream. */
int av_mosaic_grid(av_plan *plan, const void *vis, const void *weight, int32_t n_pol,
                   int32_t n_baseline, const int32_t *cf_baseline_map,
                   const int32_t *cf_chan_map, const int32_t *cf_pol_map,
                   const void *cfpg, int32_t ncb, int32_t ncc, int32_t ncp,
                   int32_t cu_n, int32_t cv_n, const int32_t *weight_support, int32_t ov_u,
                   int32_t ov_v, void *grid, double *sumwt, int32_t dtype,
                   int32_t mode, void *stream);

/* ------------------------------------------------- plain GPU baseline --
 * One thread per sample of ONE visibility channel, global atomics, no
 * binning: the straightforward GPU port of standard_grid_jit /
 * dgrid_numba used as the comparison baseline and as a second GPU check.
 * Arguments as av_grid/av_degrid plus the channel and its frequency. */
int av_grid_plain(const double *uvw, int64_t n_rows, int32_t n_chan,
                  int32_t chan, double freq, int32_t img_chan,
                  const void *vis, const void *weight, const uint8_t *flag,
                  int32_t n_pol, const double *cgk, int32_t support,
                  int32_t oversampling, double cell_u, double cell_v,
                  int32_t n_u, int32_t n_v, void *grid, double *sumwt,
                  int32_t dtype, int32_t mode, void *stream);
int av_degrid_plain(const double *uvw, int64_t n_rows, int32_t n_chan,
                    int32_t chan, double freq, int32_t img_chan,
                    const void *grid, int32_t n_grid_pol,
                    const uint8_t *flag, int32_t n_pol, const double *cgk,
                    int32_t support, int32_t oversampling, double cell_u,
                    double cell_v, int32_t n_u, int32_t n_v, void *vis_out,
                    int32_t dtype, int32_t mode, void *stream);

/* ---------------------------------------------- FFT pre/post passes --
 * The FFT itself is a library call (rocFFT through the host runtime).
 * av_ifftshift: out = ifftshift(in) over the last two axes of n_planes
 * complex planes [n_u][n_v] (np.fft.ifftshift, standard_grid.py:701),
 * converting from dtype to out_dtype: an FP32 grid can be transformed in
 * FP64, because the taper correction divides by down to grdsf(1)^2 = 1.6e-5
 * at the image edge, which magnifies FFT rounding.
 * av_normalise: the post-FFT step of grid2image_spheroid_ms4,
 * standard_grid.py:696-708: out[p] = real(fftshift(Y[p])) / taper * n_u *
 * n_v / sumwt[p], with taper the outer product of av_ps_correction_1d on
 * both axes; out is float64 (out_dtype AV_F64) or float32 (AV_F32). */
int av_ifftshift(const void *in, void *out, int64_t n_planes, int32_t n_u,
                 int32_t n_v, int32_t dtype, int32_t out_dtype, void *stream);
/* General shift + taper + scale pass used by the measurement operator
 * (image -> taper -> FFT -> degrid, and its adjoint; no reference
 * counterpart, imager.py:88-105 is the untapered forward path):
 * out[k] = scale * in[(k + shift) mod n] per axis, divided by the taper at
 * the source index (taper 1) or destination index (taper 2), 0 = no taper;
 * in/out real (flag 1) or complex (0) of dtype AV_F32/AV_F64; the real part
 * is taken when a complex input goes to a real output. */
int av_shift_taper(const void *in, void *out, int64_t n_planes, int32_t n_u,
                   int32_t n_v, int32_t shift_u, int32_t shift_v,
                   const double *taper_u, const double *taper_v, int32_t taper,
                   double scale, int32_t in_dtype, int32_t in_real,
                   int32_t out_dtype, int32_t out_real, void *stream);
int av_normalise(const void *fft_planes, const double *sumwt,
                 const double *taper_u, const double *taper_v,
                 int64_t n_planes, int32_t n_u, int32_t n_v, int32_t dtype,
                 void *out, int32_t out_dtype, void *stream);
/* The same image through a complex-to-real FFT, for even n_u and n_v
 * (standard_grid.py:693-708 keeps only the real part of the inverse FFT).
 * av_ifftshift_hermitian: out[p][ku][kv], kv in [0, n_v/2], complex FP64,
 * = (-1)^(ku+kv) * (X[ku][kv] + conj(X[-ku][-kv])) / 2 with X =
 * ifftshift(in[p]) (in complex of dtype); its 2-D complex-to-real inverse
 * FFT (backward normalisation) is real(fftshift(ifft2(X))).
 * av_normalise_real: out[p] = image[p] / taper * n_u * n_v / sumwt[p] for
 * that real FP64 image (already shifted); out float64 or float32. */
int av_ifftshift_hermitian(const void *in, void *out, int64_t n_planes,
                           int32_t n_u, int32_t n_v, int32_t dtype, void *stream);
int av_normalise_real(const double *image, const double *sumwt,
                      const double *taper_u, const double *taper_v,
                      int64_t n_planes, int32_t n_u, int32_t n_v, void *out,
                      int32_t out_dtype, void *stream);
/* fft_lm_to_uv of a real model image (fft.py:5-18, as
 * predict_model_visibilities applies it, imager.py:96-99) through a
 * real-to-complex FFT: out[p] [n_u][n_v] complex of dtype is the whole
 * spectrum F, from its half [n_u][n_v/2 + 1] (F[ku][kv] = conj(F[-ku][n_v -
 * kv]) past the half), times (-1)^(ku+kv) when sign != 0 (axes multiples of
 * 4: then fftshift(fft2(ifftshift(x))) = S * fft2(S * x), S = (-1)^(i+j)). */
int av_hermitian_expand(const void *half, void *out, int64_t n_planes,
                        int32_t n_u, int32_t n_v, int32_t sign, int32_t dtype,
                        void *stream);

/* ------------------------------------------------------- Hogbom CLEAN --
 * core/imaging/deconvolvers/hogbom (the _hogbom_ext extension).  Images are
 * [ny][nx] row-major of dtype AV_F32 or AV_F64; mask NULL = no mask (else a
 * pixel is used where mask > 0.5). */

/* Bytes of the device work buffer av_hogbom_clean needs (host output). */
int av_hogbom_work_bytes(int64_t *bytes);

/* maximg, src/hclean.cpp:51-81: out[0] = min, out[1] = max of the masked
 * image, starting from +1e20 / -1e20; out is 2 device values of dtype. */
int av_hogbom_maximg(const void *image, const void *mask, int32_t nx,
                     int32_t ny, int32_t dtype, void *out, void *stream);

/* clean, src/hclean.cpp:106-210, in place: model += components, residual -=
 * gain * peak * psf (psf centred at (nx/2, ny/2)) over iterations start ..
 * niter - 1 inside the clean box [xbeg, xend) x [ybeg, yend); siter is the
 * reference's start iteration (the threshold is thres * 2^((iter -
 * siter)/cspeedup) when cspeedup > 0).  (px0, py0) is the previous peak
 * (0, 0 for a fresh run; a box whose pixels are all 0 keeps it).  With
 * report_cycle > 0 the run also stops, before subtracting, at the first
 * iteration other than skip_report that is siter or has iter % report_cycle
 * == 1 -- where the reference calls its progress and stop callbacks -- so
 * the caller can run them and resume from there.  One cooperative launch;
 * work is a device buffer of av_hogbom_work_bytes() bytes whose header
 * receives {uint32 barrier, uint32 error, int32 iter, px, py, reason,
 * double peak}: reason 0 = threshold reached or niter done, 1 = report
 * point, -1 = a workgroup barrier timed out (the results are then invalid).
 * Calls sharing a work buffer must not overlap. */
int av_hogbom_clean(void *model, void *residual, const void *psf,
                    const void *mask, int32_t nx, int32_t ny, int32_t xbeg,
                    int32_t xend, int32_t ybeg, int32_t yend, int32_t niter,
                    int32_t siter, int32_t start, int32_t report_cycle,
                    int32_t skip_report, double gain, double thres,
                    double cspeedup, int32_t px0, int32_t py0, int32_t dtype,
                    void *work, void *stream);

/* ---------------------------------------------------- imaging weights --
 * core/imaging/imaging_weighting/grid_imaging_weights.py.  uvw [n_rows][3]
 * float64, freq [n_chan], cell in radians; a sample's pixel is int(x + 0.5)
 * of its grid position x, truncated toward zero.  Rows are time-major with
 * n_baseline baselines per time step (n_baseline divides n_rows; 1 serves
 * any row order): samples are taken baseline by baseline, so that
 * neighbouring threads hold consecutive times of one baseline. */

/* grid_imaging_weights_jit (:105-206): for every sample whose pixel lies
 * inside the grid and whose weight (weight[(row * n_chan + chan) *
 * weight_stride], NaN = skip) is not NaN, add the weight at the pixel and at
 * the conjugate pixel (skipped where that lies outside the grid; the
 * reference writes one past the end there) of grid [n_chan][n_u][n_v]
 * float64, and 2 * weight to sum_weight [n_chan].  Accumulates. */
int av_grid_imaging_weights(const double *uvw, int64_t n_rows,
                            int32_t n_baseline, const double *freq, int32_t n_chan,
                            const double *weight, int32_t weight_stride,
                            int32_t n_u, int32_t n_v, double cell_u,
                            double cell_v, double *grid, double *sum_weight,
                            void *stream);

/* degrid_imaging_weights_jit (:274-443), one polarisation: imaging_weight
 * [n_rows][n_chan] = weight [n_rows][n_chan] where the pixel is inside the
 * grid (0 elsewhere), divided by briggs[0][chan] * g + briggs[1][chan]
 * (briggs [2][n_chan]) when the weight and the grid value g are finite and
 * non-zero.  *nan_flag (device int32, set by the caller to 0) becomes
 * non-zero when a Briggs factor used was NaN (the reference raises). */
int av_degrid_imaging_weights(const double *grid, const double *briggs,
                              const double *uvw, int64_t n_rows,
                              int32_t n_baseline, const double *freq,
                              int32_t n_chan,
                              const double *weight, int32_t n_u, int32_t n_v,
                              double cell_u, double cell_v,
                              double *imaging_weight, int32_t *nan_flag,
                              void *stream);

/* --------------------------------------- mosaic convolution functions --
 * The per-pixel parts of make_gridding_convolution_function
 * (core/imaging/make_gridding_convolution_function.py:24-193). */

/* airy_disk_rorder (ipower, casa = 0) and casa_airy_disk_rorder (casa = 1),
 * core/imaging/imaging_utils/make_pb_symmetric.py:148-280, as
 * make_baseline_patterns (make_gridding_convolution_function.py:196-229)
 * evaluates them: out [n_ant][n_freq][n0][n1] float64, the pattern of
 * antenna type a (dish[a], blockage[a] in metres) at wavenumber kf[f] =
 * 2 pi freq / c over pixels ((i - c0) cell0, (j - c1) cell1) in radians,
 * raised to ipower (1: voltage, 2: power); 1 at the centre pixel. */
int av_gcf_aperture_patterns(const double *kf, int32_t n_freq, const double *dish,
                             const double *blockage, int32_t n_ant, int32_t n0,
                             int32_t n1, int32_t c0, int32_t c1, double cell0,
                             double cell1, int32_t ipower, int32_t casa,
                             double *out, void *stream);

/* make_phase_gradient (make_gridding_convolution_function.py:426-485) from
 * its per-field phase steps pix [n_field][2] (radians per pixel): out
 * [n_field][nx][ny] complex128 = exp(1j ((x - cx) pix0 + (y - cy) pix1)). */
int av_gcf_phase_gradient(const double *pix, int32_t n_field, int32_t nx,
                          int32_t ny, int32_t cx, int32_t cy, double *out,
                          void *stream);

#ifdef __cplusplus
}
#endif

#endif /* AVGRID_H */
