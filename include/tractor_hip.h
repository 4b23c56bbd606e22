/*
 * tractor_hip.h -- C ABI of libtractor_hip.so, the MI355X (gfx950) render-and-fit
 * library behind tractor_amd's Python API.
 *
 * Conventions (all entry points):
 *   - every pointer argument that names an array is a DEVICE pointer owned by the
 *     caller (hipMalloc / torch tensors' data_ptr), unless documented as host;
 *   - `stream` is a hipStream_t passed as void* (0 = null stream); work is
 *     enqueued on it and the call returns without synchronising;
 *   - return 0 on success, a negative TRH_E* code on error; trh_last_error()
 *     returns the calling thread's message for the last error.  Nothing aborts.
 */
#ifndef TRACTOR_HIP_H
#define TRACTOR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRH_OK            0
#define TRH_E_ARG        -1   /* invalid argument / shape                   */
#define TRH_E_HIP        -2   /* a HIP runtime call failed                   */
#define TRH_E_NOMEM      -3   /* device allocation failed                    */
#define TRH_E_UNSUPPORTED -4  /* size or mode outside what is implemented    */

#define TRH_MAXK 16           /* max MoG components in one FFT render job    */

/* Epilogue modes of a render job. */
#define TRH_STORE   0         /* out[...] = scale * model                    */
#define TRH_ADD     1         /* out[...] += scale * model (atomic)          */
#define TRH_CHI2    2         /* chi2[slot] = sum(((data - scale*model)*inverr)^2), FP64 */

/*
 * One FFT-path render (ProfileGalaxy._realGetUnitFluxModelPatch FFT branch,
 * tractor/galaxy.py:336-508): the zero-mean MoG's analytic Fourier transform
 * (tractor/mp_fourier.i:431-478) on the sz x (sz/2+1) half grid, times the
 * PSF spectrum (tractor/psf.py:249-280), inverse real FFT (galaxy.py:463),
 * Lanczos-3 shift by (mux, muy) with truncated edge taps
 * (mp_fourier.i:267-429), then the integer shift / crop into the output
 * window:  out[oy][ox] = G[oy - sy][ox - sx]  (zero outside G)
 * (galaxy.py:476-508).
 */
typedef struct trh_fft_job {
    int64_t out_offset;      /* element offset of the window in out (or data/inverr for TRH_CHI2) */
    int32_t out_stride;      /* row stride of the window, elements          */
    int32_t mw, mh;          /* window width, height                         */
    int32_t sx, sy;          /* integer shift                                */
    int32_t psf;             /* index of the PSF spectrum in psf_spec        */
    int32_t K;               /* number of MoG components (<= TRH_MAXK)       */
    int32_t mode;            /* TRH_STORE / TRH_ADD / TRH_CHI2               */
    int32_t slot;            /* chi2 output slot (TRH_CHI2)                  */
    int32_t pad0;
    double  mux, muy;        /* sub-pixel shift, |mu| <= 0.5                 */
    float   scale;           /* counts (1 for a unit-flux patch)             */
    float   pad1;
    float   amp[TRH_MAXK];
    float   var[TRH_MAXK][3];/* (var00, var01, var11) in pixels^2           */
} trh_fft_job;

/*
 * One pixel-space MoG evaluation (c_gauss_2d_grid, tractor/mix.i:452-524, or
 * c_gauss_2d_masked, tractor/gauss_masked.c:1-159 with eval_all_dxy_f,
 * tractor/mix.i:85-142).  Component k is centred at
 * (cxi[k] + cxf[k], cyi[k] + cyf[k]) in the window's pixel coordinates
 * (integer + fractional parts, so large image coordinates keep FP32 precision).
 */
typedef struct trh_mog_job {
    int64_t out_offset;      /* element offset of the window in out          */
    int64_t mask_offset;     /* byte offset of a uint8 mask window, or -1    */
    int32_t out_stride;      /* row stride of out / mask / derivs            */
    int32_t x0, y0;          /* window origin in image pixels                */
    int32_t w, h;            /* window size                                  */
    int32_t comp0, K;        /* components [comp0, comp0+K) of the comp arrays */
    int32_t cutoff_mode;     /* 0: skip Mahalanobis d>=100 (grid); 1: skip exponent < -30 (masked) */
    int32_t mode;            /* TRH_STORE / TRH_ADD                          */
    int32_t want_derivs;     /* write d/dmean_x, d/dmean_y into dx_out, dy_out */
    float   scale;
    float   pad;
} trh_mog_job;

/*
 * One adaptive ring-growth MoG evaluation (c_gauss_2d_approx3,
 * tractor/approx3.c:1-465, called by MixtureOfGaussians.evaluate_grid_approx3,
 * tractor/mixture_profiles.py:345-411).  FP64 throughout.  The box is
 * [x0, x0+w) x [y0, y0+h); components [comp0, comp0+K) are evaluated at
 * (pixel - (fx, fy)); growth starts at (xc, yc) (moved to the brightest edge
 * pixel when outside the box); pixels never reached are written as 0.
 */
#define TRH_APPROX_MAXK   64      /* max components per approx3 job           */
#define TRH_APPROX_MAX_WH 8192    /* max w + h of an approx3 job              */
typedef struct trh_mog_approx_job {
    int64_t out_offset;      /* element offset of the window in out / derivs  */
    int32_t out_stride;      /* row stride (elements)                          */
    int32_t x0, y0, w, h;    /* box                                            */
    int32_t comp0, K;        /* components of the amp/mean/var arrays          */
    int32_t xc, yc;          /* start pixel (image coordinates)                */
    int32_t minradius;       /* rings <= minradius always grow                 */
    int32_t want_derivs;     /* write d/dmean_x, d/dmean_y                     */
    int32_t pad;
    double  fx, fy;          /* offset of the component means                  */
    double  minval;          /* growth stops below this value                  */
} trh_mog_approx_job;

/*
 * Library handle: per-device twiddle tables and scratch.  The scratch and
 * the row slots of the sz >= 256 FFT renders belong to the handle: the entry
 * points that use them (trh_render_fft_batch at sz >= 256,
 * trh_lanczos_shift_batch, trh_solve_normal_f64, trh_fp_imstats) order a
 * call on one stream after the previous call on another, so they may be
 * mixed across streams; a stream being captured into a graph is not
 * ordered, so a graph that uses them must be the handle's only user while it
 * runs.  Fitters (trh_fitter_*) keep row slots of their own.
 */
typedef struct trh_handle trh_handle;

const char *trh_last_error(void);
int trh_version(void);

/* sizeof(trh_fft_job), sizeof(trh_mog_job), sizeof(trh_mog_approx_job) as
 * compiled (host only, no GPU call): lets bindings check their struct layouts. */
int trh_struct_sizes(int *fft_job, int *mog_job, int *mog_approx_job);

/* Device -> host copy of `bytes` bytes, synchronous on `stream` (binding
 * plumbing for reading library-owned device buffers; no reference counterpart). */
int trh_copy_to_host(void *dst, const void *src, uint64_t bytes, void *stream);

/* Create / destroy a handle on `device` (no reference counterpart: replaces the
 * module-level global scratch of tractor/psf.py:60-62). */
int trh_create(trh_handle **h, int device);
int trh_destroy(trh_handle *h);

/*
 * PSF spectra (PixelizedPSF.getFourierTransform, tractor/psf.py:249-280, with
 * _padInImage 214-247): n float32 stamps [n][ph][pw] (ph, pw odd) are
 * padded/cropped into sz x sz and transformed to half spectra stored column
 * by column: out[n][sz/2+1][sz] (complex64 as float pairs; element [i][j] is
 * frequency (v = i/sz, w = j/sz)).  sz in {32,64,128,256,512}.
 * Writes the centre (cx, cy) of the padded stamp to *cx, *cy (host ints).
 */
int trh_psf_spectrum_batch(trh_handle *h, const float *stamps, int n, int ph, int pw,
                           int sz, float *out, int *cx, int *cy, void *stream);

/*
 * Batched FFT-path renders (replaces mp_fourier.gaussian_fourier_transform_zero_mean
 * + np.fft.irfftn + mp_fourier.lanczos_shift_3f, tractor/galaxy.py:459-472,
 * called once per source from galaxy.py:256).  All jobs use FFT size sz.
 * jobs: device array of njobs trh_fft_job.  psf_spec: [npsf][sz/2+1][sz]
 * complex64 (the trh_psf_spectrum_batch layout).  out: float32 (TRH_STORE / TRH_ADD).  data, inverr: float32
 * (TRH_CHI2).  chi2: double slots (TRH_CHI2).  Unused pointers may be NULL.
 * With out NULL and chi2, data, inverr given, every job is a TRH_CHI2 job
 * (its mode field is not read) and a chi^2-specialised kernel runs.
 */
int trh_render_fft_batch(trh_handle *h, int sz, const trh_fft_job *jobs, int njobs,
                         const float *psf_spec, float *out, const float *data,
                         const float *inverr, double *chi2, void *stream);

/*
 * Launch shape of the FFT render kernel for FFT size sz (32..512): threads
 * per block (multiple of 64, <= 1024; one wavefront per spectral column or
 * row pair) and the cap on G rows kept in LDS per pass (`rows`); 0 keeps the
 * default.  Rows are reduced automatically to fit LDS.  Tuning knob of the
 * render schedule (no reference counterpart).
 */
int trh_set_fft_config(trh_handle *h, int sz, int threads, int rows);

/*
 * Batched pixel-space MoG evaluation (replaces tractor.mix.c_gauss_2d_grid,
 * tractor/mix.i:452-524 and c_gauss_2d_masked, tractor/gauss_masked.c:1-159).
 * Component arrays (float32 except the integer centres): amp[c], cxi[c],
 * cyi[c] (int32), cxf[c], cyf[c], var[c][3] = (v00, v01, v11).
 * mask: uint8 base pointer (may be NULL if no job has a mask).  max_pixels:
 * the largest w*h of any job (shapes the grid); njobs <= 65535 per call.
 */
int trh_render_mog_batch(trh_handle *h, const trh_mog_job *jobs, int njobs, int max_pixels,
                         const float *amp, const int32_t *cxi, const int32_t *cyi,
                         const float *cxf, const float *cyf, const float *var,
                         const uint8_t *mask, float *out, float *dx_out, float *dy_out,
                         void *stream);

/*
 * Batched ring-growth MoG evaluation (replaces tractor.mix.c_gauss_2d_approx3,
 * tractor/approx3.c:1-465; one SWIG call per evaluate_grid_approx3,
 * tractor/mixture_profiles.py:382-386).  Device arrays: jobs[njobs];
 * amp[c], mean[c][2], var[c][2][2] (float64); out (and dx_out, dy_out when a
 * job wants derivatives) float64; bbox[njobs][4] = (sx0, sx1, sy0, sy1), the
 * half-open box-relative bounds of the evaluated pixels (all -1 for a job
 * whose w, h or K are out of range).  max_wh >= every job's w + h, and
 * <= TRH_APPROX_MAX_WH (sizes the LDS mark buffers).
 */
int trh_render_mog_approx3_batch(trh_handle *h, const trh_mog_approx_job *jobs, int njobs,
                                 int max_wh, const double *amp, const double *mean,
                                 const double *var, double *out, double *dx_out, double *dy_out,
                                 int32_t *bbox, void *stream);

/*
 * Lanczos-3 shift of float32 images (lanczos_shift_image, tractor/psf.py:26-58
 * -> lanczos_shift_3f, tractor/mp_fourier.i:405-429), batched: n images of
 * h x w at in + i*h*w, shifted by (dx[i], dy[i]) (device arrays of double)
 * into out.  Replaces one SWIG call per point source (tractor/psf.py:180-187).
 */
int trh_lanczos_shift_batch(trh_handle *h, const float *in, float *out, int n, int height,
                            int width, const double *dx, const double *dy, void *stream);

/*
 * trh_lanczos_shift_batch without any allocation (no reference counterpart:
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The x-pass
 * intermediate lives in a static device array of TRH_NOALLOC_SCRATCH_FLOATS
 * floats (32 MiB) per GPU, held until the process ends, so n * height *
 * width must not exceed it (TRH_E_ARG otherwise).  The handle's device must
 * be current.  Every call on one GPU shares that array: calls that overlap
 * in time on one GPU (other streams, other handles) race.  Same results as
 * trh_lanczos_shift_batch (the same kernels).
 */
#define TRH_NOALLOC_SCRATCH_FLOATS (8 * 1024 * 1024)
int trh_lanczos_shift_batch_noalloc(trh_handle *h, const float *in, float *out, int n,
                                    int height, int width, const double *dx, const double *dy,
                                    void *stream);

/*
 * Batched pixelized-PSF point-source renders (PixelizedPSF.getPointSourcePatch
 * with a modelMask, tractor/psf.py:137-190; one SWIG lanczos_shift_3f call per
 * source in the reference).  Job fields as trh_fft_job with: (sx, sy) = window
 * coordinates of the stamp's pixel (0,0); (mux, muy) = the sub-pixel shift
 * px - round(px); psf selects the stamp at stamp + psf * ph * pw; K, amp,
 * var unused.  stamp: float32 [nstamp][ph][pw]; windows of any size (the
 * LDS holds the stamp and pw + 6 filtered columns).
 * With out NULL and chi2, data, inverr given, every job is a TRH_CHI2 job.
 */
int trh_render_point_batch(trh_handle *h, const trh_fft_job *jobs, int njobs,
                           const float *stamp, int ph, int pw, float *out, const float *data,
                           const float *inverr, double *chi2, void *stream);

/*
 * Normal equations of a dense least-squares problem (the update direction of
 * LsqrOptimizer.getUpdateDirection, tractor/lsqr_optimizer.py:328-644, for
 * the object API).  Ab: column-major float32 [ncols+1][nrows], column ncols
 * is b.  out: double [(ncols+1)^2] = Ab^T Ab (FP32 products, FP64 sums).
 * ncols <= 512.
 */
int trh_normal_eq_dense(trh_handle *h, const float *Ab, int nrows, int ncols, double *out,
                        void *stream);

/*
 * FP64 Cholesky solve of the column-scaled, damped normal equations from
 * trh_normal_eq_dense: (D^-1 N D^-1 + damp^2 I) y = D^-1 g, x = D^-1 y
 * (D = sqrt(diag N) if scale_columns, else I; zero columns give x = 0).  This
 * is the least-squares solution scipy's LSQR converges to
 * (tractor/lsqr_optimizer.py:584-591, damp as LSQR's).  *info (host, may be
 * NULL; synchronises) = 0 or the failing pivot + 1.  n <= 512.
 */
int trh_solve_normal_f64(trh_handle *h, const double *Nb, int n, double damp, int scale_columns,
                         double *x, int *info, void *stream);

/*
 * Batched forced photometry (Optimizer.forced_photometry,
 * tractor/optimize.py:16-145, with LsqrOptimizer's unscaled least-squares
 * step, tractor/lsqr_optimizer.py:11-251, on one image of many overlapping
 * sources).  Unit-flux patches u_s live in one float32 buffer: patch s is
 * win[s] = (x0, y0, w, h) (inside the image, row-major) at poff[s].
 *   psf_combine:   out[s] = sum_{q < nq} weights[s][q] * basis[idx[s][q]]
 *                  (1 <= nq <= 16), arrays of `len` floats: a spatially
 *                  varying PSF as a weighted sum of basis images, for stamps
 *                  and, by linearity, their spectra -- bilinear interpolation
 *                  of a PSF grid (nq = 4) or PSFEx's polynomial over its
 *                  eigen-PSFs (PixelizedPsfEx, tractor/psfex.py:326-346,
 *                  481-520); n <= 65535.
 *   psfex_terms:   weights[q][ii] = dx^j dy^k, idx[q][ii] = ii for source
 *                  order[q] (order may be NULL: identity; src rows x, y,
 *                  ... at stride 6), dx = (x - x0) / xscale, dy likewise,
 *                  ii = j + (degree + 1) k - k (k - 1) / 2 over j + k <=
 *                  degree (PsfEx.polynomials, tractor/psfex.py:294-325).
 *   fp_keys:       sort key of each source (src rows: x, y, kind, log r_e,
 *                  ee1, ee2): its tile cell (tile x tile pixels of the
 *                  clipped position, row-major) * 64 + the Morton index of
 *                  its 8 x 8 sub-cell.
 *   fp_plan:       for q < n, source order[q]'s unit-flux patch window
 *                  (win[q] = x0, y0, w, h inside the image), render bucket
 *                  (0 point, 1..5 FFT size 32..512, 6 empty window, 7 FFT
 *                  size out of range), cell stile[q] = key >> 6, area,
 *                  reach extent and TRH_STORE render job (out_offset and psf
 *                  left 0): the no-modelMask patches of
 *                  PixelizedPSF.getPointSourcePatch (tractor/psf.py:137-207)
 *                  and of the galaxies' FFT path (HoggGalaxy.
 *                  _getUnitFluxPatchSize, tractor/galaxy.py:678-691;
 *                  galaxy.py:336-508).  mix: two rows (exp, dev) of
 *                  1 + 2 TRH_MAXK doubles: K, amplitudes, variances.
 *   fp_jobs_bucketed: dst[j] = src[border[j]] with out_offset = poff[q]
 *                  and psf = j - bstart[bucket[q]] (q = border[j]).
 *   fp_overlaps:   the sparsity pattern of A (the sources whose patches share
 *                  pixels: the columns LsqrOptimizer stacks into its sparse
 *                  matrix, tractor/lsqr_optimizer.py:386-431).  Sources are
 *                  numbered in tile order: stile[s] = the cell (ntx x nty
 *                  cells) holding s's position, tile_first[c] .. tile_end[c]
 *                  = the sources of cell c; two windows can only meet when
 *                  their cells are within `reach` in x and in y.  With
 *                  counts: counts[s] = row length (diagonal included); else
 *                  the columns of row s, sorted, at col[row_ptr[s]..],
 *                  diag[s] = the position of (s, s) and, if erow is given,
 *                  erow[k] = the row of entry k.
 *   fp_pair_dots_csr: A_st = sum w_s u_t for every entry of that pattern
 *                  (t >= s computed, mirrored to (t, s)), w_s = u_s inverr^2
 *                  the weighted patches written by fp_rhs (FP32 products,
 *                  FP64 sums).
 *   fp_pair_dots:  A_st = sum u_s u_t inverr^2 over each overlapping pair
 *                  (s, t) = pairs[p] (FP32 products, FP64 sum), written to
 *                  the CSR values at pos[p][0] and pos[p][1].
 *   fp_rhs:        b_s = sum u_s (data - mod0) inverr^2 (mod0 may be NULL) and,
 *                  if wpatch is given, the weighted patches w_s = u_s inverr^2
 *                  (same layout as the patches).
 *   fp_block_inv:  the inverse of each of A's diagonal blocks (FP64,
 *                  Gauss-Jordan in LDS), one block per tile of consecutive
 *                  rows [tile_start[T], tile_start[T+1]) of at most max_n rows
 *                  (max_n^2 + max_n doubles and max_n + 1 ints must fit in
 *                  LDS), written dense and symmetric at blocks +
 *                  block_off[T].  info[T] = k if pivot k is negative, -1 if
 *                  the tile has more than max_n rows.
 *   fp_precond:    z = blockdiag(A_TT)^-1 r from those inverses (one dense
 *                  product per tile; max_n >= every tile size).
 *   fp_spmv:       y = A x (CSR, FP64).
 *   fp_pcg:        A x = b by FP64 conjugate gradients preconditioned with
 *                  those blocks, every scalar kept on the device.  it0 == 0
 *                  starts from x = 0; then `iters` iterations k = it0 ..
 *                  it0+iters-1 are enqueued (call again with it0 advanced
 *                  to continue).  Stops when |r|^2 <= tol^2 |b|^2:
 *                  status[0] = 0 running, 1 converged, 2 breakdown
 *                  (p.Ap <= 0); status[1] = iterations taken; launches after
 *                  the stop return at once.  work: 4 n + 5 ntile + 2 doubles,
 *                  kept between calls; its last two hold |r|^2 and |b|^2.
 *   fp_fitstats:   Optimizer._get_fitstats (tractor/optimize.py:233-376) for
 *                  one image, one unit-flux model per source: out[10 s ..] =
 *                  prochi2, pronpix, promasked, the fracflux ratio
 *                  sum(|m/c - u| |u|) / sum(u^2), sum(|u|) / |c|, proflux,
 *                  npix, sum |u| over the window, 1 if counted, and the
 *                  inverse variance sum (u scale inverr)^2 (_get_iv,
 *                  optimize.py:378-423); sums over u != 0 and inverr > 0,
 *                  zeros for counts 0 or no such pixel.  model: the image
 *                  model without sky; chi = (data - model - sky) inverr;
 *                  counts[s] = flux x photocal scale.  Over several images
 *                  the columns add, then profracflux = sum(ratio) /
 *                  max(1, sum(column 4)) and fracin = sum(column 7) /
 *                  max(1, sum(column 8)).
 *   fp_imstats:    out[0] = sum chi^2 over the image, out[1] = number of
 *                  pixels with inverr > 0 (optimize.py:244-250).
 *   fp_add_patches: img += sum_s flux[s] u_s (model image; Tractor.
 *                  getModelImage, tractor/engine.py:424-445, for these
 *                  patches).
 * (No reference counterpart for fp_block_inv, fp_precond, fp_spmv and
 * fp_pcg: the reference calls scipy's LSQR, lsqr_optimizer.py:586.)
 */
int trh_psf_combine(trh_handle *h, const float *basis, int nbasis, int64_t len,
                    const float *weights, const int32_t *idx, int nq, float *out, int n,
                    void *stream);
int trh_psfex_terms(trh_handle *h, int n, const int64_t *order, const double *src, int degree,
                    double x0, double y0, double xscale, double yscale, float *weights,
                    int32_t *idx, void *stream);
int trh_fp_keys(trh_handle *h, int n, const double *src, int img_h, int img_w, int tile,
                int32_t *key, void *stream);
int trh_fp_plan(trh_handle *h, int n, const int64_t *order, const double *src, const int32_t *key,
                int img_h, int img_w, int ph, int pw, const double *mix, int32_t *win,
                int32_t *bucket, int32_t *stile, int64_t *area, double *ext, trh_fft_job *jobs,
                void *stream);
int trh_fp_jobs_bucketed(trh_handle *h, int n, const int64_t *border, const int32_t *bucket,
                         const int64_t *bstart, const int64_t *poff, const trh_fft_job *src,
                         trh_fft_job *dst, void *stream);
int trh_fp_overlaps(trh_handle *h, int n, const int32_t *win, const int32_t *stile, int ntx,
                    int nty, const int32_t *tile_first, const int32_t *tile_end, int reach,
                    int32_t *counts, const int64_t *row_ptr, int32_t *col, int64_t *diag,
                    int32_t *erow, void *stream);
int trh_fp_pair_dots_csr(trh_handle *h, int n, const int64_t *row_ptr, const int32_t *col,
                         const int32_t *erow, const int32_t *win, const int64_t *poff,
                         const float *patches, const float *wpatch, double *vals, void *stream);
int trh_fp_pair_dots(trh_handle *h, const int32_t *pairs, const int64_t *pos, int npairs,
                     const int32_t *win, const int64_t *poff, const float *patches,
                     const float *inverr, int img_w, double *vals, void *stream);
int trh_fp_rhs(trh_handle *h, int n, const int32_t *win, const int64_t *poff,
               const float *patches, const float *data, const float *mod0, const float *inverr,
               int img_w, double *b, float *wpatch, void *stream);
int trh_fp_block_inv(trh_handle *h, const int64_t *row_ptr, const int32_t *col,
                     const double *vals, const int32_t *tile_start, const int64_t *block_off,
                     int ntile, int max_n, double *blocks, int32_t *info, void *stream);
int trh_fp_precond(trh_handle *h, const int32_t *tile_start, const int64_t *block_off, int ntile,
                   int max_n, const double *blocks, const double *r, double *z, void *stream);
int trh_fp_spmv(trh_handle *h, int n, const int64_t *row_ptr, const int32_t *col,
                const double *vals, const double *x, double *y, void *stream);
int trh_fp_pcg(trh_handle *h, int n, const int64_t *row_ptr, const int32_t *col,
               const double *vals, const int32_t *tile_start, const int64_t *block_off, int ntile,
               int max_n, const double *blocks, const double *b, double tol, int it0, int iters,
               double *x, double *work, int32_t *status, void *stream);
int trh_fp_fitstats(trh_handle *h, int n, const int32_t *win, const int64_t *poff,
                    const float *patches, const double *counts, const float *model,
                    const float *data, const float *inverr, int img_w, double scale, double sky,
                    double *out, void *stream);
int trh_fp_imstats(trh_handle *h, int64_t npix, const float *model, const float *data,
                   const float *inverr, double sky, double *out, void *stream);
int trh_fp_add_patches(trh_handle *h, int n, const int32_t *win, const int64_t *poff,
                       const float *patches, const float *flux, int img_w, float *img,
                       void *stream);

/*
 * EM fit of a K-component (K <= 8) mixture of Gaussians to each of n PSF
 * stamps (float64, ny x nx, pixel (ix, iy) at (x0 + ix, y0 + iy)): the
 * Gaussian approximation of HybridPixelizedPSF(gauss=None) and
 * GaussianMixturePSF.fromStamp (tractor/psf.py:537-597, 600-625), with
 * em_fit_2d_reg (tractor/emfit.i:274): `steps` EM steps, terms with
 * Mahalanobis distance^2 >= 100 dropped, inverse-Wishart prior alpha I on
 * the variances, amp = the summed responsibilities.  amp [n][K], mean
 * [n][K][2] and var [n][K][2][2] (FP64) hold the start and receive the fit;
 * status[s] = -1 if a variance lost positive definiteness (the fit stops
 * there), else 0.  One stamp's nx ny doubles must fit in LDS.
 */
int trh_em_fit_psf(trh_handle *h, int n, const double *stamps, int ny, int nx, int x0, int y0,
                   int K, double alpha, int steps, double *amp, double *mean, double *var,
                   int32_t *status, void *stream);

/*
 * Surrogate CNN (no reference counterpart: the learned initialiser that
 * proposes a source's type and starting parameters for the batched fit).
 * Activations are NHWC bfloat16 stored as uint16; convolutions are 3x3,
 * stride 2, pad 1, 32 output channels, ReLU, computed with BF16 MFMA and FP32
 * accumulation.  H and W are even.
 *   conv_fwd:        x [B][H][W][cin] (cin 1 or 32), wf [32][K] (K = 16 for
 *                    cin 1, taps 9..15 zero; else 288 = tap*32 + ci), bias[32]
 *                    -> y [B][H/2][W/2][32] = relu(conv + bias)
 *   conv_bwd_data:   dy [B][H/2][W/2][32] (gradient at the layer's
 *                    pre-activation), wt [ci][tap][co], x = the layer's input
 *                    [B][H][W][32] -> dx = (W^T dy) masked by x > 0 (the
 *                    previous layer's pre-activation gradient)
 *   conv_bwd_weight: dw [32][K'] += (K' = 9 for cin 1, 288 otherwise; FP32),
 *                    db[32] +=
 *   head:            global average pool of y4 [B][2][2][32], MLP 32 -> 64
 *                    (ReLU) -> 9 = 6 parameters + 3 type logits; params
 *                    W1[64][32], b1[64], W2[9][64], b2[9].  With targets
 *                    [B][6] and labels [B]: loss (mean 0.5 sum rw_i e_i^2,
 *                    parameters 3..5 only for labels > 0, + tw * cross
 *                    entropy) added to *loss, dy4 written, head gradients
 *                    added to dparams; scratch holds B * 169 floats.
 *   adam:            p -= lr_t m / (sqrt(v) + eps) with m, v updated from
 *                    g + wd p.
 *   pack:            FP32 master conv weights [32][9][cin] -> wf (and wt).
 *   cutouts:         32x32 windows of patches (patch s of side P[s] at
 *                    pix_off[s]) centred on pixel (P/2, P/2), zero outside,
 *                    divided by the window maximum (maxv[s]) -> bf16.
 */
int trh_cnn_conv_fwd(trh_handle *h, const uint16_t *x, const uint16_t *wf, const float *bias,
                     uint16_t *y, int B, int H, int W, int cin, void *stream);
int trh_cnn_conv_bwd_data(trh_handle *h, const uint16_t *dy, const uint16_t *wt, const uint16_t *x,
                          uint16_t *dx, int B, int H, int W, void *stream);
int trh_cnn_conv_bwd_weight(trh_handle *h, const uint16_t *dy, const uint16_t *x, float *dw,
                            float *db, int B, int H, int W, int cin, void *stream);
int trh_cnn_head(trh_handle *h, const uint16_t *y4, const float *params, int B,
                 const float *targets, const int32_t *labels, const float *rw, float tw,
                 float *out, float *loss, uint16_t *dy4, float *scratch, float *dparams,
                 void *stream);
int trh_cnn_adam(trh_handle *h, float *p, const float *g, float *m, float *v, int n, float lr_t,
                 float b1, float b2, float eps, float wd, void *stream);
int trh_cnn_pack(trh_handle *h, const float *w, int cin, uint16_t *wf, uint16_t *wt, void *stream);
int trh_cnn_cutouts(trh_handle *h, const float *data, const int64_t *pix_off, const int32_t *P,
                    int n, uint16_t *out, float *maxv, void *stream);

/* ------------------------------------------------------------------------
 * Batched single-source fitting (config workloads: thousands of sources, one
 * modelMask patch each).  Replaces GpuOptimizer's per-source
 * gpu_one_source_update / gpu_one_source_try_updates
 * (tractor/gpu_optimizer.py:184-811) and LsqrOptimizer.getUpdateDirection +
 * ConstrainedOptimizer.tryUpdates (tractor/lsqr_optimizer.py:328-644,
 * tractor/constrained_optimizer.py:36-159) for one-source problems.
 * ------------------------------------------------------------------------ */
typedef struct trh_fitter trh_fitter;

/*
 * Create a fitter for S sources.  Host arrays:
 *   kind[S]      0 PointSource, 1 ExpGalaxy, 2 DevGalaxy, 3 SersicGalaxy
 *                (tractor/sersic.py:362-487; needs trh_fitter_set_sersic)
 *   P[S]         square modelMask size (patch is (0,0,P,P) in patch pixels)
 *   psf_stamp    float32 [ph][pw] PixelizedPSF stamp shared by all sources
 *   cdi[4]       position-derivative matrix (ducks.py:413-423): row-major 2x2
 *   row_stride[S] row stride of each source's window in data / inverr (>= P;
 *                NULL: P, the patches back to back)
 * Device arrays (caller-owned, must outlive the fitter):
 *   data, inverr float32; source s reads the P x P window at pix_off[s]
 *                (host int64 array) with row stride row_stride[s], so
 *                sources may share one image (overlapping windows, the
 *                modelMask of each source in a tile, tractor/engine.py:
 *                341-370) or own contiguous patches
 */
int trh_fitter_create(trh_handle *h, trh_fitter **f, int S, const int32_t *kind,
                      const int32_t *P, const int64_t *pix_off, const int32_t *row_stride,
                      const float *data, const float *inverr, const float *psf_stamp, int ph,
                      int pw, const double *cdi);
int trh_fitter_destroy(trh_fitter *f);

/* Copy parameters in/out: host double [S][TRH_FIT_NPARAM] = (x, y, flux,
 * logre, ee1, ee2, sersic index); point sources use the first 3, exp and dev
 * galaxies the first 6. */
#define TRH_FIT_NPARAM 7
int trh_fitter_set_params(trh_fitter *f, const double *params, void *stream);
int trh_fitter_get_params(trh_fitter *f, double *params, void *stream);

/*
 * The Sersic mixture for kind-3 sources (SersicMixture, tractor/sersic.py:
 * 25-342: amplitudes and log-variances interpolated in the index), as
 * piecewise cubics in a flat double array of ntab values:
 *   R; then per index range [lo, hi), in increasing lo:
 *     lo, hi, K, m, breakpoints b[0..m], then for each component k the m
 *     amplitude pieces and the m log-variance pieces, each 4 coefficients
 *     (highest power first) in powers of (n - b[i]), valid on [b[i], b[i+1]);
 *   then the core fraction: m, b[0..m], m pieces of 4 coefficients.
 * Indices outside [lo of the first range, hi of the last] clamp to it; a
 * zero-variance core carries the core fraction above n = 1.  At most 8
 * ranges, in increasing lo, each starting at or before the previous one's
 * end (no gaps) and at most two covering any index; a table that breaks
 * these rules, or whose length does not match its counts, is rejected
 * (TRH_E_ARG).  tractor_amd.sersic.fitter_table() builds it.
 */
int trh_fitter_set_sersic(trh_fitter *f, const double *tab, int ntab);

/*
 * One LM iteration over all still-active sources (Jacobian renders,
 * J^T W J in FP32 with FP64 reduction, FP64 damped Cholesky with damping
 * `lambda` on the column-scaled normal equations (lambda=0: Gauss-Newton, the
 * reference's direction), batched line search over the reference's alphas
 * with its acceptance rule).  Sources whose step satisfies the optimize_loop
 * stopping rule (constrained_optimizer.py:12-34) with `dchisq`, or that
 * have taken max_iter iterations (optimize_loop's `steps`), become
 * inactive.  Enqueued on `stream`.
 */
int trh_fitter_iterate(trh_fitter *f, double lambda, double dchisq, int max_iter,
                       void *stream);

/*
 * `n_iter` LM iterations replayed from a hipGraph that captures the
 * launches of four iterations when n_iter is a multiple of four, else of
 * one (each captured on first use, and again when lambda, dchisq or
 * max_iter change), on a library-owned stream ordered after and before
 * `stream`.  Same results as calling trh_fitter_iterate n_iter times; not
 * available while per-launch profiling is on.  No reference counterpart
 * (launch scheduling).
 */
int trh_fitter_iterate_graph(trh_fitter *f, double lambda, double dchisq, int max_iter,
                             int n_iter, void *stream);

/*
 * Branched replays for the last iterations of a fit (no reference
 * counterpart: launch scheduling).  Once the last trh_fitter_status read
 * counted fewer than `below_active` active sources, trh_fitter_iterate_graph
 * replays a second captured iteration in which, in each render phase, the
 * sz = 128 launch runs on the fit's stream while the other launches (the
 * other FFT sizes, then the point sources) run on one library-owned side
 * stream between a fork and a join, so that launches that are
 * latency-bound when few renders remain overlap.  Results are identical
 * either way (every render job writes its own window or chi^2 slot).
 * trh_fitter_set_params goes back to the first graph.  0 (the default):
 * never branch.
 */
int trh_fitter_set_branching(trh_fitter *f, int below_active);

/*
 * The number of active sources without synchronising the fit (no reference
 * counterpart: the check between graph replays of BatchFitter.fit).
 * trh_fitter_count_active enqueues a count on `stream`, ordered after the
 * fitter's replays so far; its result reaches pinned host memory on a
 * library-owned stream while later replays run (an earlier count that was
 * not read is dropped).  trh_fitter_count_read returns it: with `wait`
 * after waiting for it, without `wait` -1 if it has not arrived yet.  The
 * count it returns is also what trh_fitter_set_branching's threshold is
 * compared with.
 */
int trh_fitter_count_active(trh_fitter *f, void *stream);
int trh_fitter_count_read(trh_fitter *f, int wait, int32_t *n_active);

/*
 * Hybrid PSF (HybridPixelizedPSF, tractor/psf.py:537-625, as the GPU plugin
 * requires, tractor/gpu_optimizer.py:827): every galaxy render splits its
 * sheared components by the FFT box size sz (ProfileGalaxy, tractor/
 * galaxy.py:412-455; the same ramp as gpu_optimizer.py:961-1011): component
 * k with pixel variance trace vv_k goes to the FFT when (sz/2)^2 >
 * inner_nsigma^2 vv_k and to pixel space when (sz/2)^2 < outer_nsigma^2
 * vv_k; where any component is in both, both amplitudes are ramped by
 * ns_k = (sz/2) / sqrt(vv_k).  The pixel-space part is convolved with the
 * PSF's mixture (K components: amp[K], mean[K][2] in pixels, var[K][3] =
 * var00, var01, var11 in pixels^2) and evaluated on the source's window at
 * the galaxy's position (c_gauss_2d_grid, tractor/mix.i:452-524, skipping
 * Mahalanobis distances >= 100), then added to the FFT render (galaxy.py:
 * 510-516).  Point sources keep the pixelized stamp (the hybrid PSF's
 * getPointSourcePatch).  K = 0 turns the split off.  K <= 4; reference
 * defaults inner 3, outer 4.  Allocates its buffers on the first call
 * (13 P^2 floats per galaxy for the line search's pixel-space renders);
 * the next trh_fitter_iterate_graph captures a new graph.
 */
int trh_fitter_set_hybrid(trh_fitter *f, int K, const double *amp, const double *mean,
                          const double *var, double inner_nsigma, double outer_nsigma);

/*
 * Chained restarts (no counterpart in the reference's optimizers: the
 * fallback of a surrogate-initialised fit, "re-fit from the standard start
 * when the proposal converges worse", run inside the same batch so that the
 * re-fits overlap the rest of the fit).  Host arrays of n pairs: source
 * second[i] waits, stopped, until source first[i] has stopped; it then
 * starts from the parameters set for it (with its own iteration count) if
 * first[i]'s chi^2 is above ref_chisq[i] * (1 + rel_tol) or not a number,
 * and otherwise never runs.  A source may be in one pair only, in one role.
 * The check runs on the device after every iteration (one launch, part of
 * the captured graph).  trh_fitter_set_params puts every pair back to
 * pending; n = 0 removes them.  Ordered after the fitter's graph replays on
 * `stream`, and synchronises it.
 */
int trh_fitter_set_chain(trh_fitter *f, int n, const int32_t *first, const int32_t *second,
                         const double *ref_chisq, double rel_tol, void *stream);

/*
 * Per pair of trh_fitter_set_chain (host int32 [n]; synchronises `stream`):
 * 0 first source still running, 1 second source started, 2 not needed.
 */
int trh_fitter_chain_state(trh_fitter *f, int32_t *state, void *stream);

/*
 * Render every source's model (counts x unit-flux patch at the current
 * parameters, as getModelPatch, tractor/pointsource.py:36-60) into `out`,
 * laid out like `data` (window s at pix_off[s], row stride row_stride[s]):
 * stored, or with `accumulate` added (atomically, so overlapping windows
 * sum: Tractor.getModelImage, tractor/engine.py:424-445).  A source whose
 * model needs an FFT size outside 32..512 is not written and gets
 * TRH_FIT_F_MODEL_SIZE.  Enqueued on `stream`.
 */
int trh_fitter_render_models(trh_fitter *f, float *out, int accumulate, void *stream);

/*
 * Per-source status, host arrays (NULL to skip): active[S] (int32),
 * iterations[S] (int32), chisq[S] (double, at the current parameters as of the
 * last iteration), dlnp[S] (double, last accepted improvement).
 * Synchronises `stream`.
 */
int trh_fitter_status(trh_fitter *f, int32_t *active, int32_t *iterations, double *chisq,
                      double *dlnp, int32_t *n_active, void *stream);

/*
 * Per-source flags (host int32 [S]; synchronises `stream`), reset by
 * trh_fitter_set_params:
 *   TRH_FIT_F_JAC_SIZE    a Jacobian render needed an FFT size outside
 *                         32..512: the source stopped at its parameters then
 *   TRH_FIT_F_LS_SIZE     a line-search trial did: it counted as a
 *                         non-finite log-probability (the walk over the
 *                         steps ends there, tryUpdates' rule)
 *   TRH_FIT_F_MAXITER     stopped at max_iter without meeting the stopping rule
 *   TRH_FIT_F_LIMITED     the last iteration was step-limited (its stop test
 *                         used dchisq_limited = 1e-6)
 *   TRH_FIT_F_MODEL_SIZE  trh_fitter_render_models could not render it
 *   TRH_FIT_F_HYBRID      with a hybrid PSF, a render's pixel-space part had
 *                         more than 64 components (wide galaxy components x
 *                         PSF components): as TRH_FIT_F_JAC_SIZE /
 *                         TRH_FIT_F_LS_SIZE / TRH_FIT_F_MODEL_SIZE
 */
#define TRH_FIT_F_JAC_SIZE    1
#define TRH_FIT_F_LS_SIZE     2
#define TRH_FIT_F_MAXITER     4
#define TRH_FIT_F_LIMITED     8
#define TRH_FIT_F_MODEL_SIZE 16
#define TRH_FIT_F_HYBRID     32
int trh_fitter_flags(trh_fitter *f, int32_t *flags, void *stream);

/*
 * The fitter's stages as standalone batched calls (device arrays unless
 * noted), for callers that drive their own LM loop.
 *
 * trh_jacobian_normal_eq_batch: the normal equations of the single-source
 *   least-squares step of GpuOptimizer.gpu_one_source_update
 *   (tractor/gpu_optimizer.py:670-792: columns from the finite-difference
 *   renders, weighted by inverr, and b = (data - counts R0) inverr), for n
 *   sources.  Source s: kind[s] as in trh_fitter_create, its P x P window
 *   of data / inverr at pix_off[s] with row stride row_stride[s], and its
 *   render planes at renders + render_off[s], P*P floats each: plane 0 the
 *   unit-flux render, planes 1..3 the renders with the three shape
 *   parameters each moved by shape_step (galaxies), plane 4 the render with
 *   the Sersic index moved by index_step[s] (kind 3).  counts[s] = flux.
 *   Columns: x, y (central differences of plane 0 through cdi, a host
 *   row-major 2x2, NULL = identity, times counts), flux (plane 0), the
 *   shape and index forward differences times counts / step.  neq[s][40]:
 *   the packed upper triangle of the NC x NC matrix J^T W J (NC = 6, or 7
 *   for kind 3; point sources fill the x, y, flux entries), then the NC
 *   entries of J^T W r, chi^2 at [36] and the chi^2 of a zero model at [37]
 *   (point sources); FP32 products, FP64 sums.  With index_step NULL the
 *   kind-3 sources are skipped.
 * trh_solve_damped_cholesky_f64_batch: the column-scaled, damped FP64
 *   Cholesky solve of each source's system in neq (as written above;
 *   lambda is added to the scaled diagonal, raised tenfold from 1e-12 when
 *   the system is not positive definite; all-zero columns give X = 0,
 *   gpu_optimizer.py:752-792): X[s][8] (x, y, flux, shape..., index).
 *   With trial given (and params [s][8], nsteps, steplim), also the
 *   line-search trials of ConstrainedOptimizer.getParameterSteps
 *   (constrained_optimizer.py:84-159): trial[s][13][8], their count and
 *   whether a maximum step limited them.
 * trh_line_search_chi2_batch: chi2[s][a] = the chi^2 of source s at trial
 *   parameters trial[s][a][8] over its window, for a < nsteps[s] (13 slots
 *   per source), rendered with the fitter's sources, PSF and data
 *   (gpu_one_source_try_updates, gpu_optimizer.py:341-425, as chi^2 rather
 *   than its change); a trial that needs an FFT size outside 32..512 gets
 *   +inf and TRH_FIT_F_LS_SIZE.  Uses the fitter's job buffers: not
 *   concurrent with its iterations.
 */
int trh_jacobian_normal_eq_batch(trh_handle *h, int n, const int32_t *kind, const int32_t *P,
                                 const int32_t *row_stride, const int64_t *pix_off,
                                 const float *data, const float *inverr,
                                 const int64_t *render_off, const float *renders,
                                 const double *counts, const double *index_step,
                                 const double *cdi, double shape_step, double *neq,
                                 void *stream);
int trh_solve_damped_cholesky_f64_batch(trh_handle *h, int n, const int32_t *kind,
                                        const double *neq, double lambda, const double *params,
                                        double *X, double *trial, int32_t *nsteps,
                                        int32_t *steplim, void *stream);
int trh_line_search_chi2_batch(trh_fitter *f, const double *trial, const int32_t *nsteps,
                               double *chi2, void *stream);

/* Device pointer to the fitter's Jacobian renders (float32, per source
 * 1 + n_shape (+1 Sersic index) planes of P*P at render_off[s]) and normal
 * equations (double [S][40]: the packed upper triangle of the NC x NC matrix,
 * NC = 6 or 7 for a Sersic galaxy, then NC gradient entries; chi^2 at 36).
 * For testing. */
int trh_fitter_debug(trh_fitter *f, const float **renders, const int64_t **render_off,
                     const double **normal_eq, const double **step_x);

/*
 * Per-launch profiling.  trh_fitter_profile(f, 1) resets the counters and
 * makes every later trh_fitter_iterate record HIP events around each launch
 * slot (and synchronise its stream at the end).  trh_fitter_profile_read
 * fills out[15][4] = (total ms, launches, jobs, algorithmic bytes) per slot:
 * 0-4 Jacobian FFT renders at sz 32/64/128/256/512 (bytes: renders written +
 * PSF spectrum read), 5 Jacobian point renders, 6 normal equations (renders +
 * data + inverr read), 7 solve, 8-12 line-search FFT renders (data + inverr
 * read), 13 line-search point renders, 14 job preparation.
 */
int trh_fitter_profile(trh_fitter *f, int enable);
int trh_fitter_profile_read(trh_fitter *f, double *out);

/* Kernel timing of the last trh_fitter_iterate (milliseconds, by HIP events
 * on the fitter's stream): [0] jacobian renders, [1] normal equations,
 * [2] solve + steps, [3] line-search renders, [4] accept, [5] whole iteration. */
int trh_fitter_timing(trh_fitter *f, float *ms6);

#ifdef __cplusplus
}
#endif

#endif /* TRACTOR_HIP_H */
